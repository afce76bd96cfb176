"""The device time integrator (fg_predictor_old_transport: Predictor_Old.c
with Criteria.c over the transient 2D SUPG build) against the reference's
loop (tests/golden/pred_*, written by the reference's Predictor_Old with a
direct dense solve) and the oracle.

The device solves each correction with GMRES to 1e-13 relative, the
reference fixtures with a direct solve: the step counts, final times and
corrections per step must be identical, u within 1e-9 relative after the
whole run (the solves differ by ~1e-13 per correction)."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu
CASES = golden_files("pred_*.npz")


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _setup(ctx, d, pc_kind=None):
    mesh = F.Mesh2D(ctx, d["x"], d["y"], d["type"], d["vertex"])
    assert np.array_equal(mesh.lm(), d["lm"])
    neq = mesh.neq
    op = F.Operator.ebe(ctx, neq, mesh.lm(), 3)
    pr = d["params"]
    kind = pc_kind if pc_kind is not None else (F.PREC_DIAG if int(pr[8]) else None)
    pc = F.Preconditioner(op, kind) if kind is not None else None
    p = d["prob"]
    pb = F.Supg2dProblem(int(p[0]), p[1], p[2], p[3], p[4], p[5], p[6], int(p[7]))
    up = ctx.to_device(np.full(len(d["x"]), float(d["ud"])))
    return mesh, op, pc, pb, up, neq


def _params(d):
    pr = d["params"]
    return dict(alpha=pr[0], dt=pr[1], final_time=pr[2], nl_tol=pr[3], nl_maxiter=int(pr[4]),
                stop_by_norm=bool(pr[5]), steady=bool(pr[6]), time_tol=pr[7])


@pytest.mark.parametrize("engine", [F.ENGINE_AUTO, F.ENGINE_KERNELS])
@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_predictor_matches_reference(ctx, path, engine):
    d = load_golden(path)
    mesh, op, pc, pb, up, neq = _setup(ctx, d)
    u = ctx.to_device(d["u0"])
    info, hist = F.predictor_old_transport(mesh, pb, op, pc, up, u, tol=1e-13, kmax=30,
                                           lmax=2000, engine=engine, max_steps=400,
                                           **_params(d))
    assert info.status == F.FG_OK and info.unconverged == 0
    assert info.steps == int(d["nsteps"]) and info.t == float(d["t_final"])
    assert np.array_equal(hist, d["hist"])
    assert info.corrections == int(d["hist"].sum())
    ug = u.get()
    assert np.max(np.abs(ug[:neq] - d["u"][:neq])) <= 1e-9 * np.max(np.abs(d["u"][:neq]))
    print(f"\n{os.path.basename(path)} engine {engine}: {info.steps} steps, "
          f"{info.corrections} corrections, {info.krylov_steps} Arnoldi steps, {info.ms:.1f} ms")


def test_predictor_preconditioner_does_not_change_the_answer(ctx):
    """element Jacobi instead of Diag: the same corrections, u within 1e-9"""
    d = load_golden(CASES[1])
    mesh, op, pc, pb, up, neq = _setup(ctx, d, pc_kind=F.PREC_JACOBI)
    u = ctx.to_device(d["u0"])
    info, hist = F.predictor_old_transport(mesh, pb, op, pc, up, u, tol=1e-13, lmax=2000,
                                           max_steps=400, **_params(d))
    assert info.status == F.FG_OK and np.array_equal(hist, d["hist"])
    assert np.max(np.abs(u.get()[:neq] - d["u"][:neq])) <= 1e-9 * np.max(np.abs(d["u"][:neq]))


def test_predictor_max_steps(ctx):
    d = load_golden(CASES[0])
    mesh, op, pc, pb, up, neq = _setup(ctx, d)
    u = ctx.to_device(d["u0"])
    prm = _params(d)
    prm["final_time"] = 1e9
    info, hist = F.predictor_old_transport(mesh, pb, op, pc, up, u, max_steps=3, **prm)
    assert info.status == F.FG_ERR_NOCONV and info.steps == 3 and len(hist) == 3


def test_predictor_rejects_a_foreign_operator(ctx):
    d = load_golden(CASES[0])
    mesh, op, pc, pb, up, neq = _setup(ctx, d)
    other = F.Operator.ebe(ctx, neq, np.full(3 * (mesh.nel - 1), neq, np.int32), 3)
    u = ctx.to_device(d["u0"])
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.predictor_old_transport(mesh, pb, other, None, up, u, **_params(d))


# ---- the Euler driver (03): Predictor_Old over the Euler build -------------
EU_CASES = golden_files("predeu_*.npz")


def _setup_eu(ctx, d):
    neq, nn = int(d["neq"]), int(d["nnodes"])
    m = F.Euler2D(ctx, d["x"], d["y"], d["vertex"], d["lm"], neq)
    op = F.Operator.ebe(ctx, neq, d["lm"], 12)
    blocks = F.Preconditioner.blockdiag(op, nn, d["vertex"], d["ids"])
    pr = d["params"]
    ep = F.EulerParams(1.4, 3.0, int(pr[9]), pr[0], pr[1], 1e-10)
    presc = ctx.to_device(np.tile(d["presc"], nn))
    rot = ctx.to_device((d["v1type"] == -1).astype(np.int32))
    theta = ctx.to_device(d["theta"])
    return m, op, blocks, ep, presc, rot, theta, neq


def _params_eu(d):
    pr = d["params"]
    return dict(alpha=pr[0], dt=pr[1], final_time=pr[2], nl_tol=pr[3], nl_maxiter=int(pr[4]),
                stop_by_norm=bool(pr[5]), steady=bool(pr[6]), time_tol=pr[7])


@pytest.mark.parametrize("path", EU_CASES, ids=os.path.basename)
def test_predictor_euler_matches_reference(ctx, path):
    """the reference's scaling (NO_scaling or Block_scaling_EBE, as the file
    says), and the block-diagonal preconditioner inside GMRES instead: the
    same steps and corrections, u within 1e-9"""
    d = load_golden(path)
    m, op, blocks, ep, presc, rot, theta, neq = _setup_eu(ctx, d)
    ref_mode = F.PRED_EU_BLOCK_SCALING if int(d["params"][8]) else F.PRED_EU_NOT
    for mode in (ref_mode, F.PRED_EU_BLOCKDIAG):
        u = ctx.to_device(d["u0"])
        info, hist = F.predictor_old_euler(m, ep, op, blocks if mode else None, mode, presc,
                                           rot, theta, u, tol=1e-13, lmax=2000, max_steps=200,
                                           **_params_eu(d))
        assert info.status == F.FG_OK and info.unconverged == 0
        assert info.steps == int(d["nsteps"]) and info.t == float(d["t_final"])
        assert np.array_equal(hist, d["hist"])
        ug = u.get()
        assert np.max(np.abs(ug[:neq] - d["u"][:neq])) <= 1e-9 * np.max(np.abs(d["u"][:neq]))
        print(f"\n{os.path.basename(path)} mode {mode}: {info.steps} steps, "
              f"{info.corrections} corrections, {info.krylov_steps} Arnoldi steps")


def test_predictor_euler_rejects_missing_blocks(ctx):
    d = load_golden(EU_CASES[0])
    m, op, blocks, ep, presc, rot, theta, neq = _setup_eu(ctx, d)
    u = ctx.to_device(d["u0"])
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.predictor_old_euler(m, ep, op, None, F.PRED_EU_BLOCK_SCALING, presc, rot, theta, u,
                              **_params_eu(d))


SCH_CASES = golden_files("predsch_*.npz")


@pytest.mark.parametrize("path", SCH_CASES, ids=os.path.basename)
def test_predictor_euler_other_schemes(ctx, path):
    """the driver's EBE2 scheme (ebe2mvNDOF4 with the block setup on R) and
    its edge schemes EDE (edemvNDOF4; NO_precond, and BlockDiag, which does
    not change the answer) and EDE2 (ede2mvNDOF4) against the reference's
    runs of each (tests/golden/predsch_*): the same steps, corrections per
    step and final time, u within 1e-9"""
    d = load_golden(path)
    sch = int(d["scheme"])
    m, op, blocks, ep, presc, rot, theta, neq = _setup_eu(ctx, d)
    nn = int(d["nnodes"])
    g = F.Ede4(ctx, nn, d["vertex"], d["ids"], neq)
    _, _, lm2, en = g.arrays()
    op8 = F.Operator.ebe(ctx, neq, lm2, 8)
    bd8 = F.Preconditioner.blockdiag(op8, nn, en, d["ids"])
    runs = {1: [("EBE2", F.Operator.ebe2(op, blocks, F.EBE2_BLOCKDIAG), blocks, F.PRED_EU_TWO,
                 None)],
            2: [("EDE", op8, None, F.PRED_EU_NOT, g),
                ("EDE+BlockDiag", op8, bd8, F.PRED_EU_BLOCKDIAG, g)],
            3: [("EDE2", F.Operator.ebe2(op8, bd8, F.EBE2_BLOCKDIAG), bd8, F.PRED_EU_TWO, g)]}
    for name, A, pc, mode, edges in runs[sch]:
        u = ctx.to_device(d["u0"])
        info, hist = F.predictor_old_euler(m, ep, A, pc, mode, presc, rot, theta, u, tol=1e-13,
                                           lmax=2000, max_steps=200, edges=edges,
                                           **_params_eu(d))
        assert info.status == F.FG_OK and info.unconverged == 0, name
        assert info.steps == int(d["nsteps"]) and info.t == float(d["t_final"]), name
        assert np.array_equal(hist, d["hist"]), name
        ug = u.get()
        assert np.max(np.abs(ug[:neq] - d["u"][:neq])) <= 1e-9 * np.max(np.abs(d["u"][:neq])), name
        print(f"\n{os.path.basename(path)} {name}: {info.steps} steps, "
              f"{info.corrections} corrections, {info.krylov_steps} Arnoldi steps")


def test_predictor_euler_edge_scheme_rejects_bad_input(ctx):
    d = load_golden(EU_CASES[0])
    m, op, blocks, ep, presc, rot, theta, neq = _setup_eu(ctx, d)
    nn = int(d["nnodes"])
    g = F.Ede4(ctx, nn, d["vertex"], d["ids"], neq)
    _, _, lm2, en = g.arrays()
    op8 = F.Operator.ebe(ctx, neq, lm2, 8)
    bd8 = F.Preconditioner.blockdiag(op8, nn, en, d["ids"])
    u = ctx.to_device(d["u0"])
    for A, pc, mode in ((op8, bd8, F.PRED_EU_BLOCK_SCALING),   # EBE only
                        (op, blocks, F.PRED_EU_BLOCKDIAG),      # the element operator
                        (op8, bd8, F.PRED_EU_TWO)):             # TWO needs the ede2 operator
        with pytest.raises(F.FemGpuError, match="status 2"):
            F.predictor_old_euler(m, ep, A, pc, mode, presc, rot, theta, u, edges=g,
                                  **_params_eu(d))
