"""The one-workgroup GMRES engine (systems of at most 4096 equations) against
the reference's outputs, the oracle and the kernel-per-operation engine.

Reference comparisons use the tolerances of test_gpu_parity.py.  Between the
two GPU engines (different reduction trees) the FP64 step counts must agree
within 2 and X within 1e-10 relative, as for the fused engine."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
from conftest import golden_files, load_golden
from test_gpu_parity import _check_solve, _op, _rel, _rho_hists, _scheme_spread

pytestmark = pytest.mark.gpu
FILES = golden_files()


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_small_fp64_matches_reference(ctx, path, orc):
    d = load_golden(path)
    neq = int(d["neq"])
    op = _op(ctx, d)
    B, X = ctx.to_device(d["F_u0"]), ctx.empty(neq + 1)
    info, rho, _ = F.gmres(ctx, op, None, B, X, kmax=30, lmax=500, tol=1e-8, hist_cap=8192,
                           engine=F.ENGINE_SMALL)
    _check_solve(info, X.get(), int(d["it_not"]), d["X_not"], neq, _scheme_spread(path, "X_not"))
    rho_o, spread = _rho_hists(orc, path)
    n1 = min(30, len(rho_o))
    err = np.abs(rho[:n1] - rho_o[:n1])
    assert np.all(err <= 1e-12 * rho_o[:n1] + 1e-14 * rho_o[0] + 4.0 * spread[:n1])
    pc = F.Preconditioner(op, F.PREC_DIAG)
    Bd = ctx.to_device(d["F_u0"])
    pc.setup(Bd)
    info, _, _ = F.gmres(ctx, op, pc, Bd, X, kmax=30, lmax=500, tol=1e-8, engine=F.ENGINE_SMALL)
    _check_solve(info, X.get(), int(d["it_diag"]), d["X_diag"], neq,
                 _scheme_spread(path, "X_diag"))
    assert info.true_res <= 1.01 * info.eps


@pytest.mark.parametrize("path", golden_files("ss2d_*_ebe.npz"), ids=os.path.basename)
def test_small_jacobi_vs_kernels(ctx, path):
    """element Jacobi (Jacobi_precond_EBE.c): both engines reach the same
    outcome, steps within 2, X within 1e-10 relative (or both stagnate)."""
    d = load_golden(path)
    neq = int(d["neq"])
    op = _op(ctx, d)
    pc = F.Preconditioner(op, F.PREC_JACOBI)
    Bj = ctx.to_device(d["F_u0"])
    pc.setup(Bj)
    out = {}
    for eng in (F.ENGINE_KERNELS, F.ENGINE_SMALL):
        X = ctx.empty(neq + 1)
        info, _, e0 = F.gmres(ctx, op, pc, Bj, X, kmax=30, lmax=200, tol=1e-8, engine=eng)
        out[eng] = (info.status, info.steps, X.get()[:neq], e0)
    sk, nk, Xk, ek = out[F.ENGINE_KERNELS]
    ss, ns, Xs, es = out[F.ENGINE_SMALL]
    assert ss == sk
    if sk == F.FG_OK:
        assert abs(nk - ns) <= 2
        assert _rel(Xs, Xk) <= 1e-10
        # restart norms: the first is ||B|| exactly in both (X = 0)
        assert es[0] == pytest.approx(ek[0], rel=1e-14)


def _system(ctx, n, scheme, problem=F.P_CONST):
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    pb = F.Supg2dProblem(problem, 1e-4, 0.0, 1.0, 0.6, 0.8, 1.0, F.H_2SQRTAREA)
    ke, F0 = ctx.empty(9 * nel), ctx.empty(neq + 1)
    F0.zero()
    mesh.build_supg(pb, ke, F0)
    lm = mesh.lm()
    if scheme == "EBE":
        op = F.Operator.ebe(ctx, neq, lm, 3)
        op.set_values(ke)
    elif scheme == "CSR":
        _, IA, JA, cbe = F.csr_pattern(neq, lm)
        op = F.Operator.csr(ctx, neq, IA, JA)
        F.assemble_csr(op, cbe, ke)
    else:
        _, order, ebe, lm2 = F.ede_pattern(neq, lm)
        op = F.Operator.ede(ctx, neq, lm2)
        F.assemble_ede(op, order, ebe, ke)
    return op, F0, neq


@pytest.mark.parametrize("scheme", ["CSR", "EBE", "EDE"])
def test_small_vs_kernels_64(ctx, scheme):
    """the 64x64 reference case (3,969 equations), Diag, GMRES(30) to 1e-8:
    same outcome as the kernel-per-operation engine; AUTO picks this engine
    (bit-identical results, the kernel is deterministic)."""
    op, F0, neq = _system(ctx, 64, scheme)
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.setup(F0)
    res = {}
    for eng in (F.ENGINE_KERNELS, F.ENGINE_SMALL, F.ENGINE_AUTO):
        X = ctx.empty(neq + 1)
        info, rho, e0 = F.gmres(ctx, op, pc, F0, X, kmax=30, lmax=500, tol=1e-8, hist_cap=4096,
                                engine=eng)
        assert info.status == F.FG_OK
        assert info.true_res <= 1.01 * info.eps
        res[eng] = (info, X.get()[:neq], rho, e0)
    ik, Xk, rk, ek = res[F.ENGINE_KERNELS]
    iS, Xs, rs, es = res[F.ENGINE_SMALL]
    assert abs(ik.steps - iS.steps) <= 2
    assert _rel(Xs, Xk) <= 1e-10
    n = min(ik.steps, iS.steps, 30)
    assert np.all(np.abs(rs[:n] - rk[:n]) <= 1e-12 * rk[:n] + 1e-14 * rk[0])
    assert np.all(np.abs(es[:iS.cycles] - ek[:iS.cycles]) <= 1e-10 * ek[:iS.cycles])
    iA, XA, _, _ = res[F.ENGINE_AUTO]
    assert iA.steps == iS.steps and np.array_equal(XA, Xs)
    print(f"\n64x64 {scheme}: {iS.steps} steps, kernels {ik.ms:.2f} ms, "
          f"one workgroup {iS.ms:.2f} ms")


def test_small_zero_rhs(ctx):
    op, F0, neq = _system(ctx, 16, "EBE")
    B = ctx.to_device(np.zeros(neq + 1))
    X = ctx.to_device(np.full(neq + 1, 7.0))
    info, _, _ = F.gmres(ctx, op, None, B, X, kmax=30, lmax=10, tol=1e-8, engine=F.ENGINE_SMALL)
    assert info.status == F.FG_OK and info.steps == 0 and info.rho == 0.0
    assert np.all(X.get()[:neq] == 0.0)


def test_small_nonfinite_rhs_is_breakdown(ctx):
    """a NaN right-hand side: both engines report a breakdown"""
    op, F0, neq = _system(ctx, 16, "CSR")
    b = np.ones(neq + 1)
    b[3] = np.nan
    for eng in (F.ENGINE_SMALL, F.ENGINE_KERNELS):
        X = ctx.empty(neq + 1)
        with pytest.raises(F.FemGpuError, match="status 3"):
            F.gmres(ctx, op, None, ctx.to_device(b), X, kmax=10, lmax=3, tol=1e-8, engine=eng)


def test_small_rejects_what_it_does_not_take(ctx):
    op, F0, neq = _system(ctx, 16, "EBE")
    X = ctx.empty(neq + 1)
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.gmres(ctx, op, None, F0, X, precision=F.GMRES_FP32_IR, engine=F.ENGINE_SMALL)
    big, Fb, nb = _system(ctx, 72, "EBE")   # 5,041 equations
    assert nb > 4096
    Xb = ctx.empty(nb + 1)
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.gmres(ctx, big, None, Fb, Xb, engine=F.ENGINE_SMALL)
    # AUTO falls back to the kernel-per-operation engine
    info, _, _ = F.gmres(ctx, big, None, Fb, Xb, kmax=30, lmax=5, tol=1e-8,
                         engine=F.ENGINE_AUTO)
    assert info.steps > 0
