"""Pins the oracle's 3D steady-transport restatement (Kuhn mesh, Build_SUPG,
csr_Initialization for k=4, ILU(0)) -- CPU only.

* against the reference's own routines: the fixtures in tests/golden/ss3d_*
  were written by the reference compiled from its sources
  (tests/golden/make_golden3d.py), on the oracle's Kuhn mesh and on the
  reference's shipped mesh DIFCONV_125_384.dat; bitwise;
* against the mathematics: tetrahedron volumes and conformity, Galerkin
  diffusion/convection and SUPG element matrices and loads from shape
  function gradients computed by a matrix inverse, the manufactured DIFCONV
  source against -A lap u + beta.grad u, the ILU(0) identity (LU)_ij = A_ij
  on the pattern, and exact LU on a matrix with no fill-in.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, load_golden

CASES = tuple(os.path.join(GOLDEN, f) for f in ("ss3d_kuhn4.npz", "ss3d_difconv_125_384.npz"))


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


@pytest.mark.parametrize("name", CASES, ids=os.path.basename)
def test_build_supg3d_golden(o, name):
    d = load_golden(name)
    neq = int(d["neq"])
    ke, F = o.build_supg3d(d["x"], d["y"], d["z"], d["utype"], d["vertex"], d["lm"], neq,
                           O.prob3d(O.P3_DIFCONV, A=1.0))
    assert np.array_equal(ke, d["ke"])
    assert np.array_equal(F, d["F"])


@pytest.mark.parametrize("name", CASES, ids=os.path.basename)
def test_lm_and_csr_golden(o, name):
    d = load_golden(name)
    neq, nel = int(d["neq"]), len(d["vertex"]) // 4
    n2, idv = o.number_nodes(d["utype"])
    assert n2 == neq
    assert np.array_equal(o.fill_lm(neq, nel, d["vertex"], d["utype"], idv, nnoel=4), d["lm"])
    nnz, IA, JA, cbe = o.csr_init_k(neq, nel, 4, d["lm"])
    assert np.array_equal(IA, d["IA"]) and np.array_equal(JA, d["JA"])
    # brute force: the pattern is every (row, col) pair of an element, sorted
    pairs = {(a, b) for e in range(nel) for a in d["lm"][4 * e:4 * e + 4]
             for b in d["lm"][4 * e:4 * e + 4] if a < neq and b < neq}
    assert nnz == len(pairs)
    for i in range(neq):
        assert list(JA[IA[i]:IA[i + 1]]) == sorted(b for a, b in pairs if a == i)
    lm = d["lm"].reshape(-1, 4)
    for e in range(nel):
        for a in range(4):
            for b in range(4):
                s = cbe[16 * e + 4 * a + b]
                if lm[e, a] < neq and lm[e, b] < neq:
                    assert IA[lm[e, a]] <= s < IA[lm[e, a] + 1] and JA[s] == lm[e, b]
                else:
                    assert s == nnz
    assert np.array_equal(o.csr_assembly_k(nel, 4, nnz, d["ke"], cbe), d["AA"])


@pytest.mark.parametrize("name", CASES, ids=os.path.basename)
def test_ilu0_golden(o, name):
    d = load_golden(name)
    neq = int(d["neq"])
    IA, JA = d["IA"], d["JA"]
    rc, LU, D = o.ilu0_factor(neq, d["AA"], JA, IA)
    assert rc == 0
    off = np.ones(len(JA), bool)
    for i in range(neq):
        off[IA[i]:IA[i + 1]] &= JA[IA[i]:IA[i + 1]] != i
    assert np.array_equal(LU[off], d["LU"][off])
    assert np.array_equal(D[:neq], d["D"])
    assert np.array_equal(o.ilu0_apply(neq, LU, JA, IA, D, d["P"])[:neq], d["z_ilu"][:neq])
    assert np.array_equal(o.ilu0_apply(neq, LU, JA, IA, D, d["F"])[:neq], d["F_ilu"][:neq])


def test_ebemv_k4_golden(o):
    d = load_golden(CASES[0])
    neq, nel = int(d["neq"]), len(d["vertex"]) // 4
    Q = o.ebemv(neq, nel, 4, d["ke"], d["lm"], d["P"].copy())
    assert np.array_equal(Q, d["Q"])
    A = np.zeros((neq, neq))
    for i in range(neq):
        A[i, d["JA"][d["IA"][i]:d["IA"][i + 1]]] = d["AA"][d["IA"][i]:d["IA"][i + 1]]
    assert np.allclose(Q[:neq], A @ d["P"][:neq], rtol=0, atol=1e-14 * np.abs(Q).max())


def _grads(X):
    """gradients of the four P1 shape functions and the volume of a tet"""
    M = np.hstack([np.ones((4, 1)), X])
    C = np.linalg.inv(M)           # phi_i = C[0,i] + C[1:,i].x
    return C[1:, :].T, abs(np.linalg.det(M)) / 6.0


def test_mesh3d_geometry(o):
    n = 3
    x, y, z, t, v = o.mesh3d(n)
    assert len(x) == (n + 1) ** 3 and len(v) == 4 * 6 * n ** 3
    assert t.sum() == (n - 1) ** 3
    P = np.stack([x, y, z], 1)
    V = v.reshape(-1, 4)
    det = np.linalg.det(np.stack([P[V[:, i]] - P[V[:, 0]] for i in (1, 2, 3)], 1))
    assert np.all(det > 0) and abs(det.sum() / 6 - 1.0) < 1e-12
    faces = {}
    for e in range(len(V)):
        for skip in range(4):
            f = tuple(sorted(np.delete(V[e], skip)))
            faces[f] = faces.get(f, 0) + 1
    for f, cnt in faces.items():
        c = P[list(f)]
        on_boundary = any(np.all(c[:, a] == c[0, a]) and c[0, a] in (0.0, 1.0) for a in range(3))
        assert cnt == (1 if on_boundary else 2)


@pytest.mark.parametrize("kappa,expect_supg", [(5.0, False), (1e-3, True)])
def test_supg3d_closed_forms(o, kappa, expect_supg):
    n = 2
    x, y, z, t, v = o.mesh3d(n)
    neq, idv = o.number_nodes(t)
    nel = len(v) // 4
    lm = o.fill_lm(neq, nel, v, t, idv, nnoel=4)
    beta = np.array([1.0, 0.5, 0.25])
    f = 2.0
    ke, F = o.build_supg3d(x, y, z, t, v, lm, neq,
                           O.prob3d(O.P3_CONST, kappa=kappa, bx=beta[0], by=beta[1], bz=beta[2],
                                    f=f))
    Fx = np.zeros(neq + 1)
    P = np.stack([x, y, z], 1)
    nb = np.linalg.norm(beta)
    any_supg = False
    for e in range(nel):
        G, Vol = _grads(P[v[4 * e:4 * e + 4]])
        h = Vol ** (1.0 / 3.0)
        pe = nb * h / (2.0 * np.sqrt(3.0) * kappa)
        tau = h * max(0.0, 1.0 - 1.0 / pe) / (2.0 * nb)
        any_supg |= tau > 0
        bg = G @ beta
        K = Vol * kappa * G @ G.T + Vol / 4.0 * np.outer(np.ones(4), bg) + tau * Vol * np.outer(bg, bg)
        assert np.allclose(ke[16 * e:16 * e + 16].reshape(4, 4), K, rtol=1e-12,
                           atol=1e-14 * np.abs(K).max())
        Fe = Vol * f / 4.0 + tau * Vol * f * bg
        for r in range(4):
            Fx[lm[4 * e + r]] += Fe[r]
    assert any_supg == expect_supg
    assert np.allclose(F[:neq], Fx[:neq], rtol=1e-12, atol=1e-15)


def test_difconv_source_is_the_pde(o):
    A = 0.7
    pb = O.prob3d(O.P3_DIFCONV, A=A)
    pi = 3.141592654
    rng = np.random.default_rng(5)
    for _ in range(20):
        X, Y, Z = rng.uniform(0, 1, 3)
        s, c = np.sin(pi * np.array([X, Y, Z])), np.cos(pi * np.array([X, Y, Z]))
        lap = -3 * pi * pi * s[0] * s[1] * s[2]
        grad = pi * np.array([c[0] * s[1] * s[2], s[0] * c[1] * s[2], s[0] * s[1] * c[2]])
        expect = -A * lap + grad.sum()
        assert abs(o.lib.orc_prob3d_f(pb, X, Y, Z) - expect) <= 1e-12 * (1 + abs(expect))


def test_ilu0_identity_and_exactness(o):
    d = load_golden(CASES[0])
    neq = int(d["neq"])
    IA, JA, AA = d["IA"], d["JA"], d["AA"]
    rc, LU, D = o.ilu0_factor(neq, AA, JA, IA)
    L = np.eye(neq)
    U = np.zeros((neq, neq))
    for i in range(neq):
        for q in range(IA[i], IA[i + 1]):
            j = JA[q]
            if j < i:
                L[i, j] = LU[q]
            elif j > i:
                U[i, j] = LU[q]
        U[i, i] = 1.0 / D[i]
    A = np.zeros((neq, neq))
    for i in range(neq):
        A[i, JA[IA[i]:IA[i + 1]]] = AA[IA[i]:IA[i + 1]]
    pat = A != 0
    pat |= np.eye(neq, dtype=bool)
    prod = L @ U
    assert np.allclose(prod[pat], A[pat], rtol=0, atol=1e-13 * np.abs(A).max())
    p = d["P"]
    zz = o.ilu0_apply(neq, LU, JA, IA, D, p)
    assert np.allclose(prod @ zz[:neq], p[:neq], atol=1e-12)
    # tridiagonal: no fill-in, ILU(0) is the exact LU
    m = 40
    rng = np.random.default_rng(2)
    T = np.diag(rng.uniform(3, 4, m)) + np.diag(rng.uniform(-1, 1, m - 1), 1) + \
        np.diag(rng.uniform(-1, 1, m - 1), -1)
    IAt = np.zeros(m + 1, np.int32)
    JAt, AAt = [], []
    for i in range(m):
        for j in range(max(0, i - 1), min(m, i + 2)):
            JAt.append(j)
            AAt.append(T[i, j])
        IAt[i + 1] = len(JAt)
    JAt = np.array(JAt, np.int32)
    AAt = np.array(AAt + [0.0])
    rc, LUt, Dt = o.ilu0_factor(m, AAt, JAt, IAt)
    b = rng.standard_normal(m + 1)
    assert np.allclose(o.ilu0_apply(m, LUt, JAt, IAt, Dt, b)[:m], np.linalg.solve(T, b[:m]),
                       atol=1e-12)


def test_pgmres_with_ilu0(o):
    d = load_golden(CASES[1])
    neq = int(d["neq"])
    IA, JA, AA = d["IA"], d["JA"], d["AA"]
    rc, LU, D = o.ilu0_factor(neq, AA, JA, IA)
    B = o.ilu0_apply(neq, LU, JA, IA, D, d["F"])
    X, steps, cycles, rho, _ = o.pgmres(neq, 50, 100, 1e-10, B, "CSR", (AA, JA, IA),
                                        ilu=(LU, JA, IA, D))
    A = np.zeros((neq, neq))
    for i in range(neq):
        A[i, JA[IA[i]:IA[i + 1]]] = AA[IA[i]:IA[i + 1]]
    Xd = np.linalg.solve(A, d["F"][:neq])
    assert np.linalg.norm(X[:neq] - Xd) <= 1e-8 * np.linalg.norm(Xd)
    assert steps < neq
