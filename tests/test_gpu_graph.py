"""Graph replay of the GMRES restart cycle (FG_ENGINE_GRAPH) against the
kernel-per-operation engine (FG_ENGINE_KERNELS), which runs the same kernels
with the same arguments launched one by one: the same step counts, residual
histories and solutions bit for bit, FP64 (MGS) and FP32-IR (MGS and CGS),
for element, edge and CSR operators.  The captured cycle is kept in the
context: replays of one capture across solves, with the program's own eager
work (other products, setups, host copies) between them, must each match an
eager solve; creating, destroying or reconfiguring an operator must force a
new capture (never a replay of pointers that are gone)."""
import numpy as np
import pytest

import fem_codes_amd as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys2d():
    ctx = F.Context(0)
    n = 96                      # 9,025 equations: above the one-workgroup engine's range
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    ke = ctx.empty(9 * nel)
    Fv = ctx.empty(neq + 1)
    Fv.zero()
    pb = F.Supg2dProblem(F.P_CONST, 1e-3, 0.0, 1.0, 0.6, 0.8, 1.0, F.H_2SQRTAREA)
    mesh.build_supg(pb, ke, Fv)
    yield ctx, mesh, neq, nel, ke, Fv.get()
    ctx.close()


def _ops(ctx, mesh, neq, ke):
    lm = mesh.lm()
    ebe = F.Operator.ebe(ctx, neq, lm, 3)
    ebe.set_values(ke)
    nnz, IA, JA, cbe = F.csr_pattern(neq, lm)
    csr = F.Operator.csr(ctx, neq, IA, JA)
    F.assemble_csr(csr, cbe, ke)
    nedge, order, eb, lm2 = F.ede_pattern(neq, lm)
    ede = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(ede, order, eb, ke)
    return {"EBE": ebe, "CSR": csr, "EDE": ede}


def _solve(ctx, op, Fh, neq, engine, precision, orth, kmax=20):
    pc = F.Preconditioner(op, F.PREC_DIAG)
    B = ctx.to_device(Fh)
    pc.setup(B)
    X = ctx.empty(neq + 1)
    info, rho, e0 = F.gmres(ctx, op, pc, B, X, kmax=kmax, lmax=500, tol=1e-8,
                            precision=precision, engine=engine, orth=orth, hist_cap=4000,
                            e0_cap=501)
    pc.close()
    return info, rho, e0, X.get()[:neq]    # the sink slot X[neq] is not written


CASES = [(F.GMRES_FP64, F.ORTH_MGS), (F.GMRES_FP32_IR, F.ORTH_MGS),
         (F.GMRES_FP32_IR, F.ORTH_CGS)]


@pytest.mark.parametrize("scheme", ["EBE", "CSR", "EDE"])
@pytest.mark.parametrize("precision,orth", CASES)
def test_graph_equals_kernels(sys2d, scheme, precision, orth):
    ctx, mesh, neq, nel, ke, Fh = sys2d
    ops = _ops(ctx, mesh, neq, ke)
    op = ops[scheme]
    ik, rk, ek, Xk = _solve(ctx, op, Fh, neq, F.ENGINE_KERNELS, precision, orth)
    assert ik.status == F.FG_OK and ik.cycles >= 2
    # a capture, then replays of it in later solves, with eager work between
    for rep in range(3):
        ig, rg, eg, Xg = _solve(ctx, op, Fh, neq, F.ENGINE_GRAPH, precision, orth)
        assert ig.status == F.FG_OK
        assert (ig.steps, ig.cycles) == (ik.steps, ik.cycles), rep
        assert np.array_equal(Xg, Xk), rep
        assert np.array_equal(rg[: ig.steps], rk[: ik.steps]), rep
        assert np.array_equal(eg[: ig.cycles + 1], ek[: ik.cycles + 1]), rep
        # eager work on the same stream between replays
        other = ops["CSR" if scheme != "CSR" else "EBE"]
        other.matvec(np.ones(neq + 1))
    for o in ops.values():
        o.close()


def test_graph_recaptures_after_reconfiguration(sys2d):
    """a knob change (the product kernel) and an operator destroyed and
    replaced at run time force a new capture; every solve matches the
    kernel engine under the same configuration"""
    ctx, mesh, neq, nel, ke, Fh = sys2d
    lm = mesh.lm()
    prec = F.GMRES_FP32_IR
    for trial in range(3):
        op = F.Operator.ebe(ctx, neq, lm, 3)
        op.set_values(ke)
        for kind in (F.PRODUCT_AUTO, F.PRODUCT_BLOCKED, F.PRODUCT_AUTO):
            op.set_product_kernel(kind)
            ik, _, _, Xk = _solve(ctx, op, Fh, neq, F.ENGINE_KERNELS, prec, F.ORTH_CGS)
            ig, _, _, Xg = _solve(ctx, op, Fh, neq, F.ENGINE_GRAPH, prec, F.ORTH_CGS)
            assert ig.steps == ik.steps and np.array_equal(Xg, Xk), (trial, kind)
        op.close()


def test_graph_time_loop_equals_kernels(sys2d):
    """the transient time loop (one GMRES solve per correction) with the
    cycle replayed from one capture gives the kernel engine's solution"""
    ctx, mesh, neq, nel, ke, Fh = sys2d
    lm = mesh.lm()
    x, y, t, v = F.structured_mesh2d(96)
    nn = len(x)
    # rotation about the centre of the unit square, a cone-shaped start
    pb = F.Supg2dProblem(F.P_ROTATE_ABOUT, 1e-6, 0.0, 0.0, 0.5, 0.5, 1.0, F.H_2SQRTAREA)
    up = ctx.to_device(np.zeros(nn))
    _, idv = F.number_nodes(t)
    free = t > 0
    cone = np.zeros(neq + 1)
    cone[idv[free]] = np.exp(-50.0 * ((x[free] - 0.5) ** 2 + (y[free] - 0.75) ** 2))
    out = {}
    for engine in (F.ENGINE_KERNELS, F.ENGINE_GRAPH):
        op = F.Operator.ebe(ctx, neq, lm, 3)
        pc = F.Preconditioner(op, F.PREC_DIAG)
        u = ctx.to_device(cone)
        info, hist = F.predictor_old_transport(
            mesh, pb, op, pc, up, u, alpha=0.5, dt=1e-3, final_time=4.5e-3, nl_tol=1e-3,
            nl_maxiter=10, stop_by_norm=True, steady=False, max_steps=4, kmax=20, lmax=200,
            tol=1e-8, precision=F.GMRES_FP32_IR, engine=engine, orth=F.ORTH_CGS)
        out[engine] = (info.steps, info.corrections, int(info.krylov_steps), u.get()[:neq])
        pc.close()
        op.close()
    k, g = out[F.ENGINE_KERNELS], out[F.ENGINE_GRAPH]
    assert k[:3] == g[:3]
    assert np.array_equal(k[3], g[3])
