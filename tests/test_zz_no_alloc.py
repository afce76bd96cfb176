"""fg_gmres_small_noalloc (the allocation-free variant of the one-workgroup
FP64 GMRES, workspace in a fixed per-GPU device array) against the entry
point it varies, fg_gmres with FG_ENGINE_SMALL, through the C ABI.

Bar: the same kernel on the same inputs, so the solution, the step and cycle
counts and the residual history are equal bit for bit; checked on two calls
in a row, the second with a smaller system, so that the second call runs on
the array the first one left behind.  The first system (the 64x64 SUPG
transport case with CSR storage and the Diag preconditioner) is also checked
against the oracle's pgmres: same step count, X within 1e-12.  Sizes beyond
the array and solves the engine does not take are rejected with FG_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _system(ctx, n, scheme, prec, theta):
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    pb = F.Supg2dProblem(F.P_CONST, 1e-4, 0.0, 1.0, float(np.cos(theta)), float(np.sin(theta)),
                         1.0, F.H_2SQRTAREA)
    ke, B = ctx.empty(9 * nel), ctx.empty(neq + 1)
    B.zero()
    mesh.build_supg(pb, ke, B)
    lm = mesh.lm()
    if scheme == "CSR":
        nnz, IA, JA, cbe = F.csr_pattern(neq, lm)
        op = F.Operator.csr(ctx, neq, IA, JA)
        F.assemble_csr(op, cbe, ke)
    else:
        op = F.Operator.ebe(ctx, neq, lm, 3)
        op.set_values(ke)
    pc = F.Preconditioner(op, prec)
    pc.setup(B)
    return dict(op=op, pc=pc, B=B, neq=neq, x=x, y=y, t=t, v=v, pb=pb, lm=lm, nel=nel)


def _opts(kmax=30, lmax=100, tol=1e-8, hist_cap=0, hist=None, e0=None):
    return F.GmresOpts(kmax, lmax, tol, F.GMRES_FP64, 0.0, hist_cap,
                       hist.ctypes.data if hist is not None else None,
                       e0.ctypes.data if e0 is not None else None, F.ENGINE_SMALL, F.ORTH_MGS)


def _solve(ctx, s, fn, kmax=30, lmax=100):
    X = ctx.empty(s["neq"] + 1)
    hist = np.zeros(4096)
    e0 = np.zeros(lmax + 1)
    o = _opts(kmax, lmax, hist_cap=len(hist), hist=hist, e0=e0)
    info = F.GmresInfo()
    rc = fn(ctx.h, s["op"].h, s["pc"].h, s["B"].ptr, X.ptr, C.byref(o), C.byref(info))
    return rc, info, X.get(), hist[:info.steps].copy(), e0[:info.cycles + 1].copy()


def test_noalloc_matches_small_engine_on_two_calls(ctx):
    big = _system(ctx, 64, "CSR", F.PREC_DIAG, 0.7)     # 3,969 equations
    small = _system(ctx, 16, "EBE", F.PREC_JACOBI, 2.1)  # 225 equations
    print(f"\nnoalloc: first call CSR+Diag neq={big['neq']}, second EBE+Jacobi neq={small['neq']}",
          flush=True)
    ref = [_solve(ctx, s, F.lib.fg_gmres) for s in (big, small)]
    got = [_solve(ctx, s, F.lib.fg_gmres_small_noalloc) for s in (big, small)]
    for (rc0, i0, X0, h0, e00), (rc1, i1, X1, h1, e01) in zip(ref, got):
        assert rc0 == F.FG_OK and rc1 == F.FG_OK
        assert (i1.steps, i1.cycles, i1.status) == (i0.steps, i0.cycles, i0.status)
        assert i1.rho == i0.rho and i1.true_res == i0.true_res
        assert np.array_equal(X1, X0)
        assert np.array_equal(h1, h0) and np.array_equal(e01, e00)
    # the first system against the oracle's pgmres
    o = O.Oracle()
    _, idv = o.number_nodes(big["t"])
    nel = big["nel"]
    keo, Fo = o.build_supg2d(big["x"], big["y"], big["t"], idv, big["v"], big["neq"],
                             np.zeros(big["neq"] + 1),
                             O.prob2d(0, 1e-4, 0.0, 1.0, float(np.cos(0.7)), float(np.sin(0.7)),
                                      1.0, 0))
    lmo = o.fill_lm(big["neq"], nel, big["v"], big["t"], idv)
    nnz, IA, JA, cbe = o.csr_init(big["neq"], nel, lmo)
    AA = o.csr_assembly(nel, nnz, keo, cbe)
    D, iD, Fs = o.diag_setup_csr(big["neq"], AA, JA, IA, Fo)
    Xo, steps, _, _, _ = o.pgmres(big["neq"], 30, 100, 1e-8, Fs, "CSR", (AA, JA, IA), invDiag=iD)
    rc, info, X, _, _ = got[0]
    assert info.steps == steps
    n = big["neq"]
    assert np.max(np.abs(X[:n] - Xo[:n])) <= 1e-12 * np.max(np.abs(Xo[:n]))
    # a third call, the big system again after the small one
    again = _solve(ctx, big, F.lib.fg_gmres_small_noalloc)
    assert np.array_equal(again[2], ref[0][2]) and again[1].steps == ref[0][1].steps


def test_noalloc_rejects_what_does_not_fit(ctx):
    s = _system(ctx, 16, "EBE", F.PREC_DIAG, 0.3)
    X = ctx.empty(s["neq"] + 1)
    info = F.GmresInfo()
    # a residual history of 2M doubles does not fit the 8 MiB array
    hist = np.zeros(2 << 20)
    o = _opts(hist_cap=len(hist), hist=hist)
    rc = F.lib.fg_gmres_small_noalloc(ctx.h, s["op"].h, s["pc"].h, s["B"].ptr, X.ptr,
                                      C.byref(o), C.byref(info))
    assert rc == F.FG_ERR_ARG
    assert "scratch" in F.lib.fg_last_error().decode()
    # FP32-IR is not the one-workgroup engine's
    o = _opts()
    o.precision = F.GMRES_FP32_IR
    rc = F.lib.fg_gmres_small_noalloc(ctx.h, s["op"].h, s["pc"].h, s["B"].ptr, X.ptr,
                                      C.byref(o), C.byref(info))
    assert rc == F.FG_ERR_ARG
    # a system above 4096 equations
    big = _system(ctx, 72, "EBE", F.PREC_DIAG, 0.3)   # 5,041 equations
    Xb = ctx.empty(big["neq"] + 1)
    rc = F.lib.fg_gmres_small_noalloc(ctx.h, big["op"].h, big["pc"].h, big["B"].ptr, Xb.ptr,
                                      C.byref(_opts()), C.byref(info))
    assert rc == F.FG_ERR_ARG
    # and a valid call still works after the rejections
    rc, info, X1, _, _ = _solve(ctx, s, F.lib.fg_gmres_small_noalloc)
    rc0, info0, X0, _, _ = _solve(ctx, s, F.lib.fg_gmres)
    assert rc == F.FG_OK and np.array_equal(X1, X0)
