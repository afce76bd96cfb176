"""Pins the oracle's time integrator (Predictor_Old.c with Criteria.c over the
transient 2D SUPG build of 02_Transport_Equation) -- CPU only.

* against the reference's own loop: tests/golden/pred_* were written by the
  reference's Predictor_Old, StopBy* criteria, ddot/daxpy, precond setups
  and transient build compiled from their sources, with a direct dense solve
  as the 6-argument solver (tests/golden/make_golden_pred.py); bitwise: the
  final u, the corrections of every step, the step count and final time;
* against the mathematics of the scheme, independently of the oracle's loop
  and of its dense solver: with one correction per step the method is
  u_pred = u_n + (1 - alpha) dt a_n, a_{n+1} = Da = A^{-1} R(u_pred, 0),
  u_{n+1} = u_pred + alpha dt Da, with A and R from the (separately pinned)
  transient build and Da from numpy's solve; checked over two steps
  (alpha = 1: backward Euler; alpha = 1/2: the predictor term of step 2).
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("pred_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _prob(p):
    return O.OrcProb2d(int(p[0]), p[1], p[2], p[3], p[4], p[5], p[6], int(p[7]))


def _run(o, d, **over):
    pr = d["params"]
    args = dict(alpha=pr[0], dt=pr[1], final_time=pr[2], nl_tol=pr[3], nl_maxiter=int(pr[4]),
                stop_by_norm=int(pr[5]), steady=int(pr[6]), time_tol=pr[7], diag=int(pr[8]),
                cap=400)
    args.update(over)
    return o.predictor_old_tr(d["x"], d["y"], d["type"], d["idv"], d["vertex"], d["lm"],
                              int(d["neq"]), _prob(d["prob"]), float(d["ud"]), d["u0"], **args)


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_predictor_golden(o, path):
    d = load_golden(path)
    u, hist, ns, tf, rc = _run(o, d)
    assert rc == 0
    assert ns == int(d["nsteps"]) and tf == float(d["t_final"])
    assert np.array_equal(hist, d["hist"])
    assert np.array_equal(u, d["u"])


def _assemble(A, lm, neq):
    M = np.zeros((neq, neq))
    Ae = A.reshape(-1, 3, 3)
    L = lm.reshape(-1, 3)
    for e in range(len(L)):
        for a in range(3):
            if L[e, a] == neq:
                continue
            for b in range(3):
                if L[e, b] != neq:
                    M[L[e, a], L[e, b]] += Ae[e, a, b]
    return M


@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_predictor_one_correction_two_steps(o, alpha):
    d = load_golden(CASES[0])
    neq = int(d["neq"])
    dt = 0.05
    u, hist, ns, tf, rc = _run(o, d, alpha=alpha, dt=dt, final_time=2 * dt, nl_maxiter=1,
                               stop_by_norm=0, steady=0, diag=0)
    assert rc == 0 and ns == 2 and list(hist) == [1, 1]
    pb = _prob(d["prob"])
    t, idv, ud = d["type"], d["idv"], float(d["ud"])
    un = d["u0"][:neq].copy()
    an = np.zeros(neq)
    for _ in range(2):
        upred = un + (1.0 - alpha) * dt * an
        uf = np.append(upred, 0.0)
        U = np.where(t > 0, uf[np.maximum(idv, 0)], ud)
        A, R = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], neq, U,
                                        np.zeros_like(U), pb, alpha, dt)
        Da = np.linalg.solve(_assemble(A, d["lm"], neq), R[:neq])
        an = Da
        un = upred + alpha * dt * Da
    assert np.max(np.abs(u[:neq] - un)) <= 1e-12 * np.max(np.abs(un))


@pytest.mark.skipif(not O.ReferencePredictor.available(), reason="reference build absent")
def test_predictor_vs_reference_live(o):
    """a case the fixtures do not hold: StopByIterations with NO_precond and
    the steady-state test, against the reference run here and now"""
    d = load_golden(CASES[1])
    R = O.ReferencePredictor()
    p = d["prob"]
    R.setup(d["x"], d["y"], d["type"], d["idv"], d["vertex"], d["lm"], int(d["neq"]),
            int(p[0]), p[1], p[2], p[3], p[4], p[5], p[6], float(d["ud"]))
    args = dict(alpha=0.75, dt=0.2, final_time=20.0, nl_tol=0.0, nl_maxiter=2, stop_by_norm=0,
                steady=1, time_tol=0.05, diag=0, cap=200)
    ur, hr, nr, tr, rcr = R.run(d["u0"], **args)
    uo, ho, no, to, rco = _run(o, d, **args)
    assert rcr == rco == 0 and nr == no and tr == to
    assert np.array_equal(hr, ho) and np.array_equal(ur, uo)


# ---- the same loop over the Euler build (03_Euler_Equations) ---------------
EU_CASES = golden_files("predeu_*.npz")


def _run_eu(o, d, **over):
    pr = d["params"]
    args = dict(alpha=pr[0], dt=pr[1], final_time=pr[2], nl_tol=pr[3], nl_maxiter=int(pr[4]),
                stop_by_norm=int(pr[5]), steady=int(pr[6]), time_tol=pr[7],
                block_scaling=int(pr[8]), cap=200)
    args.update(over)
    rot = (d["v1type"] == -1).astype(np.int32)
    return o.predictor_old_eu(d["x"], d["y"], d["vertex"], d["lm"], int(d["neq"]), d["ids"],
                              d["presc"], rot, d["theta"], 1.4, 3.0, int(pr[9]), 1e-10,
                              d["u0"], **args)


@pytest.mark.parametrize("path", EU_CASES, ids=os.path.basename)
def test_predictor_euler_golden(o, path):
    """bitwise against the reference's Predictor_Old over its Euler build
    (tests/golden/make_golden_pred_eu.py), with and without Block_scaling_EBE"""
    d = load_golden(path)
    u, hist, ns, tf, rc = _run_eu(o, d)
    assert rc == 0
    assert ns == int(d["nsteps"]) and tf == float(d["t_final"])
    assert np.array_equal(hist, d["hist"])
    assert np.array_equal(u, d["u"])


def test_predictor_euler_block_scaling_keeps_the_solution(o):
    """Block_scaling_EBE left-multiplies each node's rows by its inverse block
    (and fixes the prescribed dofs): the scaled system has the same solution,
    so the runs with and without it agree to rounding"""
    d0 = load_golden([p for p in EU_CASES if p.endswith("n6_dimless.npz")][0])
    d1 = load_golden([p for p in EU_CASES if p.endswith("n6_dimless_scaled.npz")][0])
    assert np.array_equal(d0["hist"], d1["hist"])
    neq = int(d0["neq"])
    assert np.max(np.abs(d0["u"][:neq] - d1["u"][:neq])) <= 1e-12 * np.max(np.abs(d0["u"]))


# ---- the Euler driver's other schemes: EBE2, EDE, EDE2 ----------------------
SCH_CASES = golden_files("predsch_*.npz")


@pytest.mark.parametrize("path", SCH_CASES, ids=os.path.basename)
def test_predictor_euler_schemes_golden(o, path):
    """bitwise against the reference's Predictor_Old over its EBE2 / EDE /
    EDE2 schemes (tests/golden/make_golden_pred_eus.py)"""
    d = load_golden(path)
    u, hist, ns, tf, rc = _run_eu(o, d, scheme=int(d["scheme"]))
    assert rc == 0
    assert ns == int(d["nsteps"]) and tf == float(d["t_final"])
    assert np.array_equal(hist, d["hist"])
    assert np.array_equal(u, d["u"])


def test_predictor_euler_edge_scheme_is_the_element_system(o):
    """the edge matrices sum to the element matrices: EDE reproduces the EBE
    run to rounding; EBE2 and EDE2 solve the same (block-preconditioned)
    system as each other, and it is not the EBE system where a node has both
    free and prescribed dofs (the whole node block is inverted)"""
    ebe = load_golden([p for p in EU_CASES if p.endswith("predeu_n6_dimless.npz")][0])
    get = {os.path.basename(p): load_golden(p) for p in SCH_CASES}
    neq = int(ebe["neq"])
    ede, e2, d2 = (get[f"predsch_n6_dimless_{s}.npz"] for s in ("ede", "ebe2", "ede2"))
    scale = np.max(np.abs(ebe["u"][:neq]))
    assert np.max(np.abs(ede["u"][:neq] - ebe["u"][:neq])) <= 1e-12 * scale
    assert np.max(np.abs(e2["u"][:neq] - d2["u"][:neq])) <= 1e-12 * scale
    mixed = ((ebe["ids"].reshape(-1, 4) < 0).any(1) & (ebe["ids"].reshape(-1, 4) >= 0).any(1))
    assert mixed.any()
    assert np.max(np.abs(e2["u"][:neq] - ebe["u"][:neq])) > 1e-8 * scale
