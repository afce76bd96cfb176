"""Launch schedules of the blocked EBE/EDE product: the persistent pipelined
grid (fg_op_set_pipeline) must give bit-identical results to one block per
workgroup, on a mesh large enough that every workgroup walks several blocks
(512^2 cells: 524,288 elements, 2,048 blocks of 256 elements), and both must
match the oracle's ebemv/edemv (ebemv.c:4-43, edemv.c:4-39)."""
import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    ctx = F.Context(0)
    n = 512
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    ke = ctx.empty(9 * nel)
    Fv = ctx.empty(neq + 1)
    Fv.zero()
    pb = F.Supg2dProblem(F.P_ROTATING, 1e-5, 0.0, 1.0, 0.0, 0.0, 1.3, F.H_2SQRTAREA)
    mesh.build_supg(pb, ke, Fv)
    P = np.random.default_rng(11).standard_normal(neq + 1)
    P[neq] = 0.0
    yield ctx, neq, nel, mesh.lm(), ke, P
    ctx.close()


@pytest.mark.parametrize("epb", [256, 512, 1024])
def test_ebe_pipeline_identical(system, epb):
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3, epb, product=F.PRODUCT_BLOCKED)
    op.set_values(ke)
    ref64 = op.matvec(P)
    ref32 = op.matvec(P.astype(np.float32), np.float32)
    for wpc in (1, 2, 4):
        op.set_pipeline(wpc)
        assert np.array_equal(op.matvec(P), ref64)
        assert np.array_equal(op.matvec(P.astype(np.float32), np.float32), ref32)
    Qo = O.Oracle().ebemv(neq, nel, 3, ke.get(), lm, P.copy())
    assert np.max(np.abs(ref64[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    op.close()


def _ebemv_f32(neq, lm, ke, P):
    """ebemv.c:24-40 in float32: each element row summed left to right, the
    sums added to Q in element order (np.add.at applies them in index order)"""
    A = np.asarray(ke, np.float32).reshape(-1, 3, 3)
    L = lm.reshape(-1, 3)
    p = np.asarray(P, np.float32).copy()
    p[neq] = 0
    pe = p[L]
    s = (A[:, :, 0] * pe[:, None, 0] + A[:, :, 1] * pe[:, None, 1]) + A[:, :, 2] * pe[:, None, 2]
    Q = np.zeros(neq + 1, np.float32)
    np.add.at(Q, L.ravel(), s.ravel())
    Q[neq] = 0
    return Q


def test_ebe_sliced_rows_bitwise(system):
    """the default element product (sliced rows) is the reference's ebemv bit
    for bit, in FP64 (oracle) and FP32 (float32 restatement)"""
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3)
    op.set_values(ke)
    assert op.product_sliced()
    keh = ke.get()
    Qo = O.Oracle().ebemv(neq, nel, 3, keh, lm, P.copy())
    assert np.array_equal(op.matvec(P), Qo)
    P32 = P.astype(np.float32)
    assert np.array_equal(op.matvec(P32, np.float32), _ebemv_f32(neq, lm, keh, P32))
    op.close()


def test_ebe_product_kernel_per_precision(system):
    """the kernel choice is per precision: blocked FP64 products beside
    sliced-row FP32 products on one operator"""
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3)
    op.set_values(ke)
    op.set_product_kernel(F.PRODUCT_BLOCKED, F.PRODUCT_FP64)
    assert not op.product_sliced(False) and op.product_sliced(True)
    keh = ke.get()
    Qo = O.Oracle().ebemv(neq, nel, 3, keh, lm, P.copy())
    assert np.max(np.abs(op.matvec(P)[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    P32 = P.astype(np.float32)
    assert np.array_equal(op.matvec(P32, np.float32), _ebemv_f32(neq, lm, keh, P32))
    op.set_product_kernel(F.PRODUCT_AUTO, F.PRODUCT_FP64)
    op.set_product_kernel(F.PRODUCT_BLOCKED, F.PRODUCT_FP32)
    assert op.product_sliced(False) and not op.product_sliced(True)
    assert np.array_equal(op.matvec(P), Qo)
    with pytest.raises(F.FemGpuError):
        op.set_product_kernel(F.PRODUCT_AUTO, 0)
    op.close()


def test_ebe_sliced_rows_follow_value_changes(system):
    """new values and a Left/LeftRight scaling reach the row copy before the
    next product"""
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3)
    op.set_values(ke)
    op.matvec(P)
    vals = np.random.default_rng(4).uniform(0.5, 1.5, 9 * nel)
    op.set_values(vals)
    assert np.array_equal(op.matvec(P), O.Oracle().ebemv(neq, nel, 3, vals, lm, P.copy()))
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.scaling(F.SCALING_LEFTRIGHT, ctx.to_device(np.ones(neq + 1)))
    scaled = op.get_values()
    assert not np.array_equal(scaled, vals)
    assert np.array_equal(op.matvec(P), O.Oracle().ebemv(neq, nel, 3, scaled, lm, P.copy()))
    pc.close()
    op.close()


def test_ebe_without_sliced_rows(system):
    """a numbering whose column offsets exceed the row fields keeps the
    blocked product: sliced rows absent and refused, products still right"""
    ctx, neq, nel, lm, ke, P = system
    perm = np.random.default_rng(2).permutation(neq)
    relabel = np.append(perm, neq).astype(np.int32)
    lmx = relabel[lm]
    Px = np.empty_like(P)
    Px[relabel[:neq]] = P[:neq]
    Px[neq] = 0.0
    op = F.Operator.ebe(ctx, neq, lmx, 3)
    op.set_values(ke)
    assert not op.product_sliced()
    with pytest.raises(F.FemGpuError):
        op.set_product_kernel(F.PRODUCT_SLICED)
    Qo = O.Oracle().ebemv(neq, nel, 3, ke.get(), lmx, Px.copy())
    assert np.max(np.abs(op.matvec(Px)[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    op.close()


def test_ede_pipeline_identical(system):
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    op = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(op, order, ebe, ke)
    ref = op.matvec(P)
    op.set_pipeline(3)
    assert np.array_equal(op.matvec(P), ref)
    op.close()


@pytest.mark.parametrize("epb", [256, 1024])
def test_ebe_mesh_ordering(system, epb):
    """blocks of consecutive elements (FG_ORDER_MESH) against breadth-first
    blocks and the oracle, in the mesh's own element order and in a shuffled
    one (the caller's ordering is kept, so blocks are scattered)"""
    ctx, neq, nel, lm, ke, P = system
    keh = ke.get().reshape(nel, 9)
    Qo = O.Oracle().ebemv(neq, nel, 3, keh.ravel(), lm, P.copy())
    perm = np.random.default_rng(5).permutation(nel)
    for lmx, kex in ((lm, keh), (lm.reshape(nel, 3)[perm].ravel(), keh[perm])):
        for order in (F.ORDER_BFS, F.ORDER_MESH):
            op = F.Operator.ebe(ctx, neq, lmx, 3, epb, ordering=order,
                                product=F.PRODUCT_BLOCKED)
            op.set_values(np.ascontiguousarray(kex).ravel())
            for wpc in (0, 2):
                op.set_pipeline(wpc)
                Q = op.matvec(P)
                assert np.max(np.abs(Q[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
                assert Q[neq] == 0.0
                Q32 = op.matvec(P.astype(np.float32), np.float32)
                assert np.max(np.abs(Q32[:neq] - Qo[:neq])) <= 1e-5 * np.max(np.abs(Qo[:neq]))
            op.close()


def test_ede_mesh_ordering(system):
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    ref_op = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(ref_op, order, ebe, ke)
    ref = ref_op.matvec(P)
    vals = ref_op.get_values()
    op = F.Operator.ede(ctx, neq, lm2, 512, ordering=F.ORDER_MESH)
    op.set_values(vals)
    Q = op.matvec(P)
    assert np.max(np.abs(Q[:neq] - ref[:neq])) <= 1e-14 * np.max(np.abs(ref[:neq]))
    op.close()
    ref_op.close()


def test_autotuner_choice_is_correct(system):
    """tune_ebe over block sizes, both orderings and two schedules returns
    the fastest timed candidate, and that operator's product is the oracle's"""
    from fem_codes_amd import autotune
    ctx, neq, nel, lm, ke, P = system
    best, times, op = autotune.tune_ebe(ctx, neq, lm, 3, ke, reps=3, candidates=(256, 512),
                                        pipelines=(0, 2), orderings=autotune.ORDERINGS)
    assert best == min(times, key=times.get)
    assert {k[1] for k in times} == set(autotune.ORDERINGS)
    Qo = O.Oracle().ebemv(neq, nel, 3, ke.get(), lm, P.copy())
    Q = op.matvec(P)
    assert np.max(np.abs(Q[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    op.close()


def test_restart_tuner_on_device(system):
    """tune_restart over real FP32-IR solves of a constant-velocity system
    (256^2 cells) with the Diag preconditioner: every candidate that
    converges reaches the tolerance, and the chosen one is the fastest"""
    from fem_codes_amd import autotune
    ctx = system[0]
    x, y, t, v = F.structured_mesh2d(256)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    ke = ctx.empty(9 * nel)
    B = ctx.empty(neq + 1)
    B.zero()
    pb = F.Supg2dProblem(F.P_CONST, 1e-3, 0.0, 1.0, 0.6, 0.8, 1.0, F.H_2SQRTAREA)
    mesh.build_supg(pb, ke, B)
    op = F.Operator.ebe(ctx, neq, mesh.lm(), 3)
    op.set_values(ke)
    mesh.close()
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.setup(B)
    res = {}

    def solve(km):
        X = ctx.empty(neq + 1)
        info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=km, lmax=200, tol=1e-8,
                             precision=F.GMRES_FP32_IR)
        res[km] = info
        return info.ms * 1e-3, info.status == F.FG_OK

    km, times = autotune.tune_restart(solve, restarts=(20, 40))
    assert times[km] == min(v for v in times.values() if v is not None)
    for k, info in res.items():
        if info.status == F.FG_OK:
            assert info.true_res <= 1.01 * info.eps
    pc.close()
    op.close()


def test_ede_rows_bitwise(system):
    """the half-edge row product (default) is the oracle's edemv bit for bit
    in FP64 (edemv.c:23-36 term grouping, edge order per equation); the
    blocked kernel agrees to 1e-14; FP32 to 1e-6"""
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    op = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(op, order, ebe, ke)
    vals = op.get_values()
    Qo = O.Oracle().edemv(neq, nedge, vals, lm2, P.copy())
    Q = op.matvec(P)
    assert np.array_equal(Q, Qo)
    Q32 = op.matvec(P.astype(np.float32), np.float32)
    assert np.max(np.abs(Q32[:neq] - Qo[:neq])) <= 1e-6 * np.max(np.abs(Qo[:neq]))
    op.set_edge_kernel(F.EDGE_BLOCKED)
    Qb = op.matvec(P)
    assert np.max(np.abs(Qb[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    Qb32 = op.matvec(P.astype(np.float32), np.float32)
    op.set_edge_kernel(F.EDGE_ROWS)
    assert np.array_equal(op.matvec(P), Qo)
    Qr32 = op.matvec(P.astype(np.float32), np.float32)
    assert np.max(np.abs(Qr32[:neq] - Qo[:neq])) <= 1e-6 * np.max(np.abs(Qo[:neq]))
    # the default (AUTO): rows in both precisions
    op.set_edge_kernel(F.EDGE_AUTO)
    assert np.array_equal(op.matvec(P), Qo)
    assert np.array_equal(op.matvec(P.astype(np.float32), np.float32), Qr32)
    assert np.max(np.abs(Qb32[:neq] - Qo[:neq])) <= 1e-6 * np.max(np.abs(Qo[:neq]))
    op.close()


def test_ede_rows_follow_value_changes(system):
    """new values (set_values) and a Left/LeftRight scaling of the stored
    values reach the row copy before the next product"""
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    op = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(op, order, ebe, ke)
    op.matvec(P)                                   # rows in use
    rng = np.random.default_rng(9)
    vals = rng.uniform(0.5, 1.5, 4 * nedge)   # positive diagonals: LeftRight takes sqrt
    op.set_values(vals)
    assert np.array_equal(op.matvec(P), O.Oracle().edemv(neq, nedge, vals, lm2, P.copy()))
    pc = F.Preconditioner(op, F.PREC_DIAG)
    Fd = ctx.to_device(np.ones(neq + 1))
    pc.scaling(F.SCALING_LEFTRIGHT, Fd)
    scaled = op.get_values()
    assert not np.array_equal(scaled, vals)
    assert np.array_equal(op.matvec(P), O.Oracle().edemv(neq, nedge, scaled, lm2, P.copy()))
    pc.close()
    op.close()


def test_ede_edge_kernel_rejected_elsewhere(system):
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3)
    with pytest.raises(F.FemGpuError):
        op.set_edge_kernel(F.EDGE_BLOCKED)
    op.close()


def test_autotuner_edges_includes_rows(system):
    """tune_ebe over edges times the half-edge row kernel beside the blocked
    schedules, and the operator it returns computes edemv"""
    from fem_codes_amd import autotune
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    src = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(src, order, ebe, ke)
    vals = src.get_values()
    src.close()
    best, times, op = autotune.tune_ebe(ctx, neq, lm2, 2, vals, reps=3, edges=True,
                                        candidates=(512, 1024), pipelines=(0,))
    assert (0, F.ORDER_BFS, autotune.ROWS) in times
    assert best == min(times, key=times.get)
    Qo = O.Oracle().edemv(neq, nedge, vals, lm2, P.copy())
    Q = op.matvec(P)
    assert np.max(np.abs(Q[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    op.close()
