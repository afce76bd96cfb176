"""4-dof (Euler 2D) element operators on the GPU against the reference's
fixtures (tests/golden/eu2d_*, written by the reference's own ebemvNDOF4,
ebe2mvNDOF4, BlockDiag_precond_EBE_setup, BlockDiag_precond and
Block_scaling_EBE) and the oracle.

Tolerances: the node blocks, their inverses, the block apply and the block
scaling are per node / per element in the reference's operation order with
no FMA contraction: bitwise.  The element product: the node rows (where
built) are bitwise (tests/test_gpu_node_rows.py); the blocked kernel adds a
dof's element contributions in the blocked (breadth-first) element order and
the halo partials in block order, not in mesh element order: 1e-14 relative
(max norm), the tolerance kept here for both; the ebe2 forms inherit it.  FP32 products: 2e-6 relative.
GMRES: the solution within 1e-8 relative of a dense solve of the same
(block-preconditioned) system."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu
CASES = golden_files("eu2d_*.npz")


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _setup(ctx, d, epb=0):
    neq = int(d["neq"])
    op = F.Operator.ebe(ctx, neq, d["lm"], 12, epb)
    op.set_values(d["A"])
    pc = F.Preconditioner.blockdiag(op, int(d["nnodes"]), d["vertex"], d["ids"])
    return op, pc


def _free_nodes(d):
    return (d["ids"].reshape(-1, 4) >= 0).any(axis=1)


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
@pytest.mark.parametrize("epb", [0, 64])
def test_ebemv_ndof4(ctx, path, epb):
    d = load_golden(path)
    neq = int(d["neq"])
    op, _ = _setup(ctx, d, epb)
    Q = op.matvec(d["P"])
    assert _rel(Q[:neq], d["Q"][:neq]) <= 1e-14 and Q[neq] == 0.0
    Q32 = op.matvec(d["P"], np.float32)
    assert _rel(Q32[:neq].astype(np.float64), d["Q"][:neq]) <= 2e-6


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_blockdiag_setup_apply(ctx, path):
    d = load_golden(path)
    neq, nn = int(d["neq"]), int(d["nnodes"])
    op, pc = _setup(ctx, d)
    Fd = ctx.to_device(d["F"])
    pc.setup(Fd)
    assert np.array_equal(Fd.get(), d["F_bd"])
    BD, iBD = pc.blocks()
    act = np.repeat(_free_nodes(d), 16)
    assert np.array_equal(BD[act], d["BD"][act])
    assert np.array_equal(iBD[act], d["invBD"][act])
    assert np.all(BD[~act] == 0) and np.all(iBD[~act] == 0)
    p, z = ctx.to_device(d["P"]), ctx.empty(neq + 1)
    pc.apply(p, z)
    assert np.array_equal(z.get(), d["z_bd"])
    pc.apply(p, p)   # in place (BlockDiag_precond_EBE_setup.c:83 calls it so)
    assert np.array_equal(p.get(), d["z_bd"])
    p32, z32 = ctx.to_device(d["P"], np.float32), ctx.empty(neq + 1, np.float32)
    pc.apply(p32, z32)
    assert _rel(z32.get()[:neq].astype(np.float64), d["z_bd"][:neq]) <= 2e-6
    assert nn == len(d["ids"]) // 4


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ebe2_forms(ctx, path):
    d = load_golden(path)
    neq = int(d["neq"])
    op, pc = _setup(ctx, d)
    pc.setup(None)
    e2 = F.Operator.ebe2(op, pc, F.EBE2_BLOCKDIAG)
    Q = e2.matvec(d["P"])
    assert _rel(Q[:neq], d["Q2_bd"][:neq]) <= 1e-14 and Q[neq] == 0.0
    Q32 = e2.matvec(d["P"], np.float32)
    assert _rel(Q32[:neq].astype(np.float64), d["Q2_bd"][:neq]) <= 2e-6
    n2 = F.Operator.ebe2(op, pc, F.EBE2_NOBLOCKDIAG)
    Q = n2.matvec(d["P"])
    assert _rel(Q[:neq], d["Q2_not"][:neq]) <= 1e-14 and Q[neq] == 0.0
    nel, nn = int(d["nel"]), int(d["nnodes"])
    assert e2.algorithmic_bytes() == pytest.approx(
        nel * (96 * 8 + 48) + 2 * neq * 8 + nn * (16 * 8 + 16) + 3 * neq * 8)


def test_ebe2_needs_setup(ctx):
    d = load_golden(CASES[0])
    op, pc = _setup(ctx, d)
    e2 = F.Operator.ebe2(op, pc, F.EBE2_BLOCKDIAG)
    with pytest.raises(F.FemGpuError):
        e2.matvec(d["P"])


def test_blockdiag_rejects_bad_input(ctx):
    d = load_golden(CASES[0])
    neq = int(d["neq"])
    op = F.Operator.ebe(ctx, neq, d["lm"], 12)
    bad = d["ids"].copy()
    bad[bad >= 0] = 0   # every free dof on equation 0
    with pytest.raises(F.FemGpuError):
        F.Preconditioner.blockdiag(op, int(d["nnodes"]), d["vertex"], bad)
    op3 = F.Operator.ebe(ctx, 4, np.array([0, 1, 2, 1, 2, 3], np.int32), 3)
    with pytest.raises(F.FemGpuError):
        F.Preconditioner.blockdiag(op3, 4, np.array([0, 1, 2, 1, 2, 3], np.int32),
                                   np.arange(16, dtype=np.int32))


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_block_scaling(ctx, path):
    d = load_golden(path)
    nel = int(d["nel"])
    op, pc = _setup(ctx, d)
    Fd = ctx.to_device(d["F"])
    pc.block_scaling(Fd)
    assert np.array_equal(Fd.get(), d["F_sc"])
    A = op.get_values().reshape(nel, 12, 12)
    Aw = d["A_sc"].reshape(nel, 12, 12)
    kept = (d["lm"].reshape(nel, 12) != int(d["neq"])).any(axis=1)
    assert not kept[0]   # element 0 has no free dof: not stored
    assert np.all(A[~kept] == 0)
    # block rows of nodes with a free dof: bitwise; rows of nodes without
    # one (their blocks are left 0, the rows only feed the sink slot) read 0
    free = _free_nodes(d)[d["vertex"].reshape(nel, 3)]          # [nel, 3]
    rows = np.repeat(free, 4, axis=1) & kept[:, None]           # [nel, 12]
    assert np.array_equal(A[rows], Aw[rows])
    assert np.all(A[~np.repeat(free, 4, axis=1)] == 0)
    assert rows.sum() > 0.9 * kept.sum() * 12
    # the FP32 copy follows the scaled values
    Q32 = op.matvec(d["P"], np.float32)
    Qs = O.Oracle().ebemv(int(d["neq"]), nel, 12, d["A_sc"], d["lm"], d["P"].copy())
    assert _rel(Q32[:int(d["neq"])].astype(np.float64), Qs[:int(d["neq"])]) <= 2e-6


def _dense_ebe2(d):
    """I + invBD (K - Dblk) on the free dofs, column by column through the oracle"""
    o = O.Oracle()
    neq, nel, nn = int(d["neq"]), int(d["nel"]), int(d["nnodes"])
    M = np.zeros((neq, neq))
    for j in range(neq):
        e = np.zeros(neq + 1)
        e[j] = 1.0
        M[:, j] = o.ebe2mv_ndof4(neq, nel, d["A"], d["lm"], nn, d["Id"], e, invBD=d["invBD"])[:neq]
    return M


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
@pytest.mark.parametrize("precision", [F.GMRES_FP64, F.GMRES_FP32_IR])
def test_gmres_ebe2_blockdiag(ctx, path, precision):
    """the EBE2 + BlockDiag scheme of the Euler driver (setPreconditioner.c:
    setup BlockDiag_precond_EBE_setup, precond NO_precond, mv ebe2mvNDOF4):
    GMRES on the block-preconditioned system"""
    d = load_golden(path)
    neq = int(d["neq"])
    op, pc = _setup(ctx, d)
    B = ctx.to_device(d["F"])
    pc.setup(B)
    e2 = F.Operator.ebe2(op, pc, F.EBE2_BLOCKDIAG)
    X = ctx.empty(neq + 1)
    info, _, _ = F.gmres(ctx, e2, None, B, X, kmax=30, lmax=200, tol=1e-10, precision=precision)
    assert info.status == F.FG_OK
    x_ref = np.linalg.solve(_dense_ebe2(d), d["F_bd"][:neq])
    assert _rel(X.get()[:neq], x_ref) <= 1e-8


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_gmres_blockdiag_precond(ctx, path):
    """EBE product + BlockDiag as GMRES's left preconditioner"""
    d = load_golden(path)
    neq = int(d["neq"])
    op, pc = _setup(ctx, d)
    B = ctx.to_device(d["F"])
    pc.setup(B)
    X = ctx.empty(neq + 1)
    info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=30, lmax=200, tol=1e-10)
    assert info.status == F.FG_OK
    K = np.zeros((neq + 1, neq + 1))
    A = d["A"].reshape(-1, 12, 12)
    for e in range(len(A)):
        g = d["lm"][12 * e:12 * e + 12]
        K[np.ix_(g, g)] += A[e]
    x_ref = np.linalg.solve(K[:neq, :neq], d["F"][:neq])
    assert _rel(X.get()[:neq], x_ref) <= 1e-8


# ---- JacobiDOF4_precond_EBE(_setup) on the 12-dof element operator ---------
JAC4 = golden_files("jac4_*.npz")


@pytest.mark.parametrize("path", JAC4, ids=os.path.basename)
def test_jacobi_dof4_bitwise(ctx, path):
    """FG_PREC_JACOBI on the k = 12 operator is JacobiDOF4_precond_EBE: the
    setup's preconditioned F and the apply (out of place and in place) equal
    the reference's bits (per incidence 1/A_jj, sums in element order)"""
    j = load_golden(path)
    d = load_golden(os.path.join(os.path.dirname(path), str(j["inputs"]) + ".npz"))
    neq = int(d["neq"])
    op = F.Operator.ebe(ctx, neq, d["lm"], 12)
    op.set_values(d["A"])
    pc = F.Preconditioner(op, F.PREC_JACOBI)
    Fd = ctx.to_device(d["F"])
    pc.setup(Fd)
    Fh = Fd.get()
    assert np.array_equal(Fh[:neq], j["F_jac"][:neq]) and Fh[neq] == 0.0
    p, z = ctx.to_device(d["P"]), ctx.empty(neq + 1)
    pc.apply(p, z)
    assert np.array_equal(z.get()[:neq], j["z_jac"][:neq])
    pc.apply(p, p)                       # in place, as pgmres.c:104 calls it
    assert np.array_equal(p.get()[:neq], j["z_jac"][:neq])
    # FP32 apply: within FP32 rounding
    p32, z32 = ctx.to_device(d["P"].astype(np.float32)), ctx.empty(neq + 1, np.float32)
    pc.apply(p32, z32)
    assert _rel(z32.get()[:neq].astype(np.float64), j["z_jac"][:neq]) <= 2e-6
    # and a Jacobi-preconditioned GMRES against a dense solve of the same system
    B = ctx.to_device(d["F"])
    pc.setup(B)
    X = ctx.empty(neq + 1)
    info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=30, lmax=200, tol=1e-12)
    assert info.status == F.FG_OK
    K = np.zeros((neq + 1, neq + 1))
    A = d["A"].reshape(-1, 12, 12)
    for e in range(A.shape[0]):
        g = d["lm"][12 * e:12 * e + 12]
        K[np.ix_(g, g)] += A[e]
    xs = np.linalg.solve(K[:neq, :neq], d["F"][:neq])
    assert _rel(X.get()[:neq], xs) <= 1e-8
    pc.close()
    op.close()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_mfma_rows_fixtures(ctx, path):
    """the 12x12 row sums on the matrix cores (fg_op_set_row_kernel): the
    reference's ebemvNDOF4 / ebe2mvNDOF4 outputs to the same tolerances"""
    d = load_golden(path)
    neq = int(d["neq"])
    op, pc = _setup(ctx, d)
    op.set_product_kernel(F.PRODUCT_BLOCKED)   # the matrix cores serve the blocked kernel
    op.set_row_kernel(F.ROWS_MFMA)
    Q = op.matvec(d["P"])
    assert _rel(Q[:neq], d["Q"][:neq]) <= 1e-14 and Q[neq] == 0.0
    Q32 = op.matvec(d["P"], np.float32)
    assert _rel(Q32[:neq].astype(np.float64), d["Q"][:neq]) <= 2e-6
    pc.setup(None)
    for form, key in ((F.EBE2_BLOCKDIAG, "Q2_bd"), (F.EBE2_NOBLOCKDIAG, "Q2_not")):
        e2 = F.Operator.ebe2(op, pc, form)
        Q = e2.matvec(d["P"])
        assert _rel(Q[:neq], d[key][:neq]) <= 1e-14 and Q[neq] == 0.0
        Q32 = e2.matvec(d["P"], np.float32)
        assert _rel(Q32[:neq].astype(np.float64), d[key][:neq]) <= 2e-6
        e2.close()


def test_mfma_rows_large_mesh(ctx):
    """96^2 cells (18,432 elements of 12 dofs, 73 blocks, ragged last tiles),
    a quarter of the nodes Dirichlet in one or more components, random
    values: matrix-core rows against the vector rows and the oracle"""
    rng = np.random.default_rng(44)
    n = 96
    x, y, t, v = F.structured_mesh2d(n)
    nn = (n + 1) ** 2
    ids = np.arange(4 * nn).reshape(nn, 4)
    ids[rng.random((nn, 4)) < 0.25] = -1
    free = ids >= 0
    ids[free] = np.arange(int(free.sum()))
    neq = int(free.sum())
    ids[~free] = neq
    tri = np.asarray(v, np.int64).reshape(-1, 3)
    lm = ids[tri].reshape(-1).astype(np.int32)
    nel = tri.shape[0]
    A = rng.standard_normal(144 * nel)
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0
    Qo = O.Oracle().ebemv(neq, nel, 12, A, lm, P.copy())
    op = F.Operator.ebe(ctx, neq, lm, 12)
    op.set_values(A)
    op.set_product_kernel(F.PRODUCT_BLOCKED)
    Qv = op.matvec(P)
    op.set_row_kernel(F.ROWS_MFMA)
    Qm = op.matvec(P)
    assert _rel(Qm[:neq], Qo[:neq]) <= 1e-14 and _rel(Qv[:neq], Qo[:neq]) <= 1e-14
    assert Qm[neq] == 0.0
    Qm32 = op.matvec(P, np.float32)
    assert _rel(Qm32[:neq].astype(np.float64), Qo[:neq]) <= 2e-6
    op.set_row_kernel(F.ROWS_VALU)
    assert np.array_equal(op.matvec(P), Qv)
    op.close()


def test_mfma_rows_rejected_where_not_built(ctx):
    d = load_golden(CASES[0])
    neq = int(d["neq"])
    op = F.Operator.ebe(ctx, neq, d["lm"], 12, 512)
    with pytest.raises(F.FemGpuError):
        op.set_row_kernel(F.ROWS_MFMA)
    op.set_row_kernel(F.ROWS_VALU)
    op.close()
    x, y, t, v = F.structured_mesh2d(4)
    tri = F.Operator.ebe(ctx, 25, np.asarray(v, np.int32), 3)
    with pytest.raises(F.FemGpuError):
        tri.set_row_kernel(F.ROWS_MFMA)
    tri.close()
