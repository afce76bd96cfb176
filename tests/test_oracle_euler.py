"""Pins the oracle's 4-dof (Euler 2D) restatement -- ebemvNDOF4,
ebe2mvNDOF4 with BlockDiag2/NOBlockDiag2, the node block-diagonal setup and
apply, Block_scaling_EBE -- CPU only.

* against the reference's own routines: the fixtures tests/golden/eu2d_*
  were written by the reference compiled from its sources
  (tests/golden/make_golden_eu.py); bitwise;
* against the mathematics: the products against a dense assembly of the
  element matrices, the ebe2 forms against the full product and the
  block-preconditioned product, the node blocks against a dense
  block-diagonal extraction, the inverse blocks against numpy's inverse
  (every entry but the one the reference writes differently, which is
  checked against the closed form of that difference), and the block
  scaling against numpy products of the boundary-condition-modified blocks.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("eu2d_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _dense(d):
    neq, nel = int(d["neq"]), int(d["nel"])
    K = np.zeros((neq + 1, neq + 1))
    A = d["A"].reshape(nel, 12, 12)
    for e in range(nel):
        g = d["lm"][12 * e:12 * e + 12]
        K[np.ix_(g, g)] += A[e]
    return K[:neq, :neq]


def _dense_blockdiag(d):
    """the same-node 4x4 blocks of K (node dofs of the same node only)"""
    neq = int(d["neq"])
    K = _dense(d)
    Dm = np.zeros_like(K)
    ids = d["ids"].reshape(-1, 4)
    for I in range(len(ids)):
        g = [x for x in ids[I] if x >= 0]
        Dm[np.ix_(g, g)] = K[np.ix_(g, g)]
    return K, Dm


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_products_golden(o, path):
    d = load_golden(path)
    neq, nel, nn = int(d["neq"]), int(d["nel"]), int(d["nnodes"])
    Q = o.ebemv(neq, nel, 12, d["A"], d["lm"], d["P"].copy())
    assert np.array_equal(Q, d["Q"])
    q2 = o.ebe2mv_ndof4(neq, nel, d["A"], d["lm"], nn, d["Id"], d["P"].copy(), invBD=d["invBD"])
    assert np.array_equal(q2, d["Q2_bd"])
    q3 = o.ebe2mv_ndof4(neq, nel, d["A"], d["lm"], nn, d["Id"], d["P"].copy(), BD=d["BD"])
    assert np.array_equal(q3, d["Q2_not"])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_blockdiag_golden(o, path):
    d = load_golden(path)
    neq, nel, nn = int(d["neq"]), int(d["nel"]), int(d["nnodes"])
    Id = np.where(d["ids"] < 0, neq, d["ids"]).astype(np.int32)
    assert np.array_equal(Id, d["Id"])
    BD, iBD, Fo = o.blockdiag_setup_ebe(neq, nel, nn, d["vertex"], d["A"], Id, d["F"])
    assert np.array_equal(BD, d["BD"]) and np.array_equal(iBD, d["invBD"])
    assert np.array_equal(Fo, d["F_bd"])
    assert np.array_equal(o.blockdiag_apply(neq, nn, Id, iBD, d["P"]), d["z_bd"])
    A2, F2, _, _ = o.block_scaling_ebe(neq, nel, nn, d["vertex"], d["ids"], Id, d["A"], d["F"])
    assert np.array_equal(A2, d["A_sc"]) and np.array_equal(F2, d["F_sc"])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_products_dense(path):
    """ebemvNDOF4 = K P; ebe2 NOBlockDiag2 = K P; ebe2 BlockDiag2 =
    P + invBD (K - Dblk) P, against dense assemblies"""
    d = load_golden(path)
    neq = int(d["neq"])
    K, Dm = _dense_blockdiag(d)
    P = d["P"][:neq]
    KP = K @ P
    scale = np.abs(KP).max()
    assert np.abs(d["Q"][:neq] - KP).max() <= 1e-13 * scale
    assert np.abs(d["Q2_not"][:neq] - KP).max() <= 1e-13 * scale
    # BlockDiag2 applies the reference's inverse blocks to the off-block product
    iBD = d["invBD"].reshape(-1, 4, 4)
    off = (K - Dm) @ P
    want = P.copy()
    ids = d["ids"].reshape(-1, 4)
    offf = np.zeros(neq + 1)
    offf[:neq] = off
    for I in range(len(ids)):
        g = np.where(ids[I] < 0, neq, ids[I])
        for r in range(4):
            if g[r] < neq:
                want[g[r]] = P[g[r]] + iBD[I, r] @ offf[g]
    assert np.abs(d["Q2_bd"][:neq] - want).max() <= 1e-13 * np.abs(want).max()
    assert d["Q"][neq] == 0 and d["Q2_bd"][neq] == 0 and d["Q2_not"][neq] == 0


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_blocks_and_inverse(path):
    """node blocks = sums of the element same-node blocks; inverse blocks =
    numpy's inverse except entry 14 (row 3, column 2), which the reference
    writes with c*j*m in place of the cofactor's c*f*m: the difference is
    exactly c*m*(j - f)/det"""
    d = load_golden(path)
    nel, nn = int(d["nel"]), int(d["nnodes"])
    A = d["A"].reshape(nel, 12, 12)
    BDw = np.zeros((nn, 4, 4))
    for e in range(nel):
        for a in range(3):
            BDw[d["vertex"][3 * e + a]] += A[e, 4 * a:4 * a + 4, 4 * a:4 * a + 4]
    BD = d["BD"].reshape(nn, 4, 4)
    assert np.abs(BD - BDw).max() <= 1e-14 * np.abs(BDw).max()
    iBD = d["invBD"].reshape(nn, 4, 4)
    true = np.linalg.inv(BD)
    mask = np.ones((4, 4), bool)
    mask[3, 2] = False
    assert np.abs((iBD - true)[:, mask]).max() <= 1e-12 * np.abs(true).max()
    B = BD.reshape(nn, 16)
    a, b, c, e, f, g, i, j, m, n, o_ = (B[:, q] for q in (0, 1, 2, 4, 5, 6, 8, 9, 12, 13, 14))
    det = np.linalg.det(BD)
    diff = c * m * (j - f) / det
    assert np.abs(iBD[:, 3, 2] - true[:, 3, 2] - diff).max() <= 1e-12 * np.abs(true).max()
    # the difference is real on these inputs (the defect is exercised)
    assert np.abs(diff).max() > 1e-6


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_block_scaling_math(path):
    """scaled element blocks = invBD_v (A_vw with the prescribed rows and
    columns removed, 1 on the diagonal of a prescribed dof's own block)"""
    d = load_golden(path)
    nel, nn = int(d["nel"]), int(d["nnodes"])
    A = d["A"].reshape(nel, 12, 12)
    A_sc = d["A_sc"].reshape(nel, 12, 12)
    iBD = d["invBD"].reshape(nn, 4, 4)
    pres = (d["ids"] < 0).reshape(nn, 4)
    err = 0.0
    for e in range(nel):
        V = d["vertex"][3 * e:3 * e + 3]
        for v in range(3):
            for w in range(3):
                M = A[e, 4 * v:4 * v + 4, 4 * w:4 * w + 4].copy()
                M[pres[V[v]], :] = 0.0
                M[:, pres[V[w]]] = 0.0
                if v == w:
                    for q in np.where(pres[V[v]])[0]:
                        M[q, q] = 1.0
                want = iBD[V[v]] @ M
                err = max(err, np.abs(A_sc[e, 4 * v:4 * v + 4, 4 * w:4 * w + 4] - want).max())
    assert err <= 1e-13 * np.abs(A_sc).max()
    # F is the block-preconditioned right-hand side
    assert np.array_equal(d["F_sc"], d["F_bd"])


@pytest.mark.skipif(not O.ReferenceEuler.available(), reason="reference build absent")
def test_oracle_vs_reference_random(o):
    """fresh seeded inputs through the compiled reference and the oracle"""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mk", os.path.join(os.path.dirname(__file__), "golden", "make_golden_eu.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    R = O.ReferenceEuler()
    for n, seed in ((4, 11), (9, 12)):
        c = mk.make_case(n, seed)
        nn, nel, neq = c["nnodes"], c["nel"], c["neq"]
        R.setup(nn, nel, neq, c["vertex"], c["ids"], c["lm"], c["A"], "BlockDiag")
        Fr, BDr, iBDr, Idr = R.blockdiag_setup(c["F"])
        BDo, iBDo, Fo = o.blockdiag_setup_ebe(neq, nel, nn, c["vertex"], c["A"], Idr, c["F"])
        assert np.array_equal(BDr, BDo) and np.array_equal(iBDr, iBDo)
        assert np.array_equal(Fr, Fo)
        assert np.array_equal(R.ebe2mv(c["P"]),
                              o.ebe2mv_ndof4(neq, nel, c["A"], c["lm"], nn, Idr, c["P"].copy(),
                                             invBD=iBDr))


# ---- JacobiDOF4_precond_EBE(_setup): the element Jacobi of 12-dof elements --
JAC4 = golden_files("jac4_*.npz")


def _jac4_case(path):
    j = load_golden(path)
    d = load_golden(os.path.join(os.path.dirname(path), str(j["inputs"]) + ".npz"))
    return j, d


@pytest.mark.parametrize("path", JAC4, ids=os.path.basename)
def test_jacobi_dof4_golden(o, path):
    """setup (invDe, preconditioned F) and the apply, bitwise against the
    reference's JacobiDOF4 routines"""
    j, d = _jac4_case(path)
    neq, nel = int(d["neq"]), int(d["nel"])
    invDe, Fj = o.jacobi_ebe_setup_k(neq, nel, 12, d["A"], d["lm"], d["F"])
    assert np.array_equal(invDe, j["invDe"])
    assert np.array_equal(Fj[:neq], j["F_jac"][:neq]) and Fj[neq] == 0.0
    z = o.jacobi_ebe_apply_k(neq, nel, 12, invDe, d["lm"], d["P"])
    assert np.array_equal(z[:neq], j["z_jac"][:neq])


@pytest.mark.parametrize("path", JAC4, ids=os.path.basename)
def test_jacobi_dof4_is_a_diagonal(o, path):
    """independently of the reference: the operator is diagonal, entry i the
    sum over the elements of i of 1/A^e_ii (from a dense walk over lm)"""
    j, d = _jac4_case(path)
    neq, nel = int(d["neq"]), int(d["nel"])
    A = d["A"].reshape(nel, 12, 12)
    diag = np.zeros(neq + 1)
    for e in range(nel):
        for a in range(12):
            diag[d["lm"][12 * e + a]] += 1.0 / A[e, a, a]
    invDe, _ = o.jacobi_ebe_setup_k(neq, nel, 12, d["A"], d["lm"], d["F"])
    z = o.jacobi_ebe_apply_k(neq, nel, 12, invDe, d["lm"], d["P"])
    ref = d["P"][:neq] * diag[:neq]
    assert np.max(np.abs(z[:neq] - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("path", golden_files("ss2d_n*_ebe.npz")[:2], ids=os.path.basename)
def test_jacobi_k_generic_equals_the_pinned_3dof_form(o, path):
    """the k-generic restatement at k = 3 gives the bits of the 3-dof form
    pinned against Jacobi_precond_EBE(_setup).c"""
    d = load_golden(path)
    neq, nel = int(d["neq"]), len(d["lm"]) // 3
    i3, F3 = o.jacobi_ebe_setup(neq, nel, d["A_u0"], d["lm"], d["F_u0"])
    assert np.array_equal(i3, d["invDe"]) and np.array_equal(F3[:neq], d["F_jac"][:neq])
    ik, Fk = o.jacobi_ebe_setup_k(neq, nel, 3, d["A_u0"], d["lm"], d["F_u0"])
    assert np.array_equal(i3, ik) and np.array_equal(F3, Fk)
    P = np.random.default_rng(4).standard_normal(neq + 1)
    assert np.array_equal(o.jacobi_ebe_apply(neq, nel, i3, d["lm"], P),
                          o.jacobi_ebe_apply_k(neq, nel, 3, ik, d["lm"], P))
