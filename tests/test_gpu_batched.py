"""Batched GMRES (fg_gmres_batched) on the GPU: the six 2D SUPG systems of
the golden fixtures (8x8 and 16x16 meshes; constant, rotating and VALIDACAO
problems; EBE storage) solved together as one block-diagonal operator, each
checked against the reference's own solution and step count with the
criteria of the single-system tests (same step count and X within
max(1e-12, 4 x the reference's inter-scheme spread) for well-conditioned
solves; 25% in steps and 1e-8 in X for the stagnating rotating-flow ones).
Graph replay must give the same bits as direct launches."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
from conftest import golden_files, load_golden
from test_gpu_parity import _check_solve, _rel, _scheme_spread

pytestmark = pytest.mark.gpu
EBE = golden_files("ss2d_*_ebe.npz")


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _batch(ctx, paths, rhs="F_u0", empty_at=None):
    ds = [load_golden(p) for p in paths]
    parts = [(int(d["neq"]), d["lm"]) for d in ds]
    vals = [d["A_u0"] for d in ds]
    if empty_at is not None:   # a problem with no free equation (every dof prescribed)
        parts.insert(empty_at, (0, np.zeros(6, np.int32)))
        vals.insert(empty_at, np.ones(18))
    lm, seg = F.concat_lm(parts)
    n = int(seg[-1])
    op = F.Operator.ebe(ctx, n, lm, 3)
    op.set_values(np.concatenate(vals))
    B = np.zeros(n + 1)
    k = 0
    for i, (neq_p, _) in enumerate(parts):
        if empty_at is not None and i == empty_at:
            continue
        B[seg[i]:seg[i + 1]] = ds[k][rhs][:neq_p]
        k += 1
    return ds, op, seg, B


def _split(X, seg, i):
    return X[seg[i]:seg[i + 1]]


def test_concat_lm():
    lm, seg = F.concat_lm([(2, np.array([0, 1, 2])), (3, np.array([3, 0, 2]))])
    assert seg.tolist() == [0, 2, 5]
    assert lm.tolist() == [0, 1, 5, 5, 2, 4]


@pytest.mark.parametrize("graph", [False, True])
def test_batched_fp64_not(ctx, graph):
    ds, op, seg, B = _batch(ctx, EBE)
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    rc, infos = F.gmres_batched(ctx, op, None, seg, Bd, X, kmax=30, lmax=500, tol=1e-8,
                                graph=graph)
    assert rc == F.FG_OK
    Xh = X.get()
    for i, (d, path) in enumerate(zip(ds, EBE)):
        neq = int(d["neq"])
        assert infos[i].status == F.FG_OK
        _check_solve(infos[i], _split(Xh, seg, i), int(d["it_not"]), d["X_not"], neq,
                     _scheme_spread(path, "X_not"))
        assert infos[i].true_res <= infos[i].eps * (1 + 1e-6)


def test_batched_fp64_diag(ctx):
    ds, op, seg, B = _batch(ctx, EBE)
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.setup(Bd)
    Bh = Bd.get()
    for i, d in enumerate(ds):   # Diag setup is per dof: each problem's F_diag
        assert np.array_equal(_split(Bh, seg, i), d["F_diag"][:int(d["neq"])])
    rc, infos = F.gmres_batched(ctx, op, pc, seg, Bd, X, kmax=30, lmax=500, tol=1e-8)
    assert rc == F.FG_OK
    Xh = X.get()
    for i, (d, path) in enumerate(zip(ds, EBE)):
        _check_solve(infos[i], _split(Xh, seg, i), int(d["it_diag"]), d["X_diag"],
                     int(d["neq"]), _scheme_spread(path, "X_diag"))


@pytest.mark.parametrize("precision,orth", [(F.GMRES_FP64, F.ORTH_MGS),
                                            (F.GMRES_FP32_IR, F.ORTH_MGS),
                                            (F.GMRES_FP32_IR, F.ORTH_CGS)])
@pytest.mark.parametrize("prec", [F.PREC_DIAG, F.PREC_JACOBI])
def test_batched_graph_bitwise(ctx, precision, orth, prec):
    """graph replay of the cycle = direct launches, bit for bit, on two solves
    in a row (the graph is rebuilt per call; each call replays it every cycle)"""
    ds, op, seg, B = _batch(ctx, [p for p in EBE if "rot" not in p])
    pc = F.Preconditioner(op, prec)
    Bd = ctx.to_device(B)
    pc.setup(Bd)
    out = []
    for graph in (False, True, True):
        X = ctx.empty(len(B))
        rc, infos = F.gmres_batched(ctx, op, pc, seg, Bd, X, kmax=20, lmax=200, tol=1e-8,
                                    precision=precision, graph=graph, orth=orth)
        assert rc == F.FG_OK
        out.append((X.get(), [(f.steps, f.cycles) for f in infos]))
    for X, st in out[1:]:
        assert np.array_equal(X, out[0][0]) and st == out[0][1]


@pytest.mark.parametrize("orth", [F.ORTH_MGS, F.ORTH_CGS])
def test_batched_fp32_ir(ctx, orth):
    """FP32-IR (inner cycle MGS or CGS): every problem reaches tol with its
    FP64 residual, and X is within 1e-6 of the reference's Diag-preconditioned
    solution"""
    paths = [p for p in EBE if "rot" not in p]
    ds, op, seg, B = _batch(ctx, paths)
    pc = F.Preconditioner(op, F.PREC_DIAG)
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    pc.setup(Bd)
    rc, infos = F.gmres_batched(ctx, op, pc, seg, Bd, X, kmax=30, lmax=200, tol=1e-8,
                                precision=F.GMRES_FP32_IR, graph=True, orth=orth)
    assert rc == F.FG_OK
    Xh = X.get()
    for i, d in enumerate(ds):
        neq = int(d["neq"])
        assert infos[i].true_res <= infos[i].eps
        assert _rel(_split(Xh, seg, i), d["X_diag"][:neq]) <= 1e-6


def test_batched_uneven_limits(ctx):
    """a problem that runs out of cycles reports NOCONV while the others
    converge; an empty problem is solved at once"""
    paths = [p for p in EBE if "n16_rot" in p or "n8_const" in p]
    ds, op, seg, B = _batch(ctx, paths, empty_at=1)
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    rc, infos = F.gmres_batched(ctx, op, None, seg, Bd, X, kmax=30, lmax=3, tol=1e-8)
    assert rc == F.FG_ERR_NOCONV
    names = [os.path.basename(p) for p in paths]
    k = 0
    for i in range(len(seg) - 1):
        if i == 1:
            assert infos[i].status == F.FG_OK and infos[i].steps == 0
            continue
        if "rot" in names[k]:
            assert infos[i].status == F.FG_ERR_NOCONV and infos[i].cycles == 3
            assert infos[i].steps == 90
        else:
            assert infos[i].status == F.FG_OK
            d = ds[k]
            assert _rel(_split(X.get(), seg, i), d["X_not"][:int(d["neq"])]) <= 1e-10
        k += 1


def test_batched_cgs_needs_fp32_ir(ctx):
    """CGS is offered for the FP32 inner cycle only (the FP64 drop-in keeps
    the reference's MGS)"""
    ds, op, seg, B = _batch(ctx, EBE[:2])
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    with pytest.raises(F.FemGpuError):
        F.gmres_batched(ctx, op, None, seg, Bd, X, precision=F.GMRES_FP64, orth=F.ORTH_CGS)


@pytest.mark.parametrize("precision", [F.GMRES_FP64, F.GMRES_FP32_IR])
def test_batched_nonfinite_rhs_is_breakdown(ctx, precision):
    """a NaN in one problem's right-hand side ends that problem as a
    breakdown (never as converged); the other problems solve as usual"""
    ds, op, seg, B = _batch(ctx, EBE)
    B[seg[1] + 3] = np.nan
    Bd, X = ctx.to_device(B), ctx.empty(len(B))
    rc, infos = F.gmres_batched(ctx, op, None, seg, Bd, X, kmax=30, lmax=500, tol=1e-8,
                                precision=precision)
    assert rc == F.FG_ERR_BREAKDOWN
    assert infos[1].status == F.FG_ERR_BREAKDOWN
    for i in range(len(ds)):
        if i != 1:
            assert infos[i].status == F.FG_OK
            assert infos[i].true_res <= infos[i].eps * (1 + 1e-6)
