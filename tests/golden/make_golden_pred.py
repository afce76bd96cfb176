"""Write the golden fixtures of the time integrator (Predictor_Old.c with
Criteria.c) over the transient 2D SUPG build of 02_Transport_Equation, from
the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference's
Predictor_Old, StopByNorm / StopByIterations, StopBySteadyState / StopByTime,
ddot, daxpy, NO_precond_setup / Diag_precond_EBE_setup, NO_scaling and the
transient build, compiled from their own sources
(oracle/_ref/libref_pred2d.so, built by ``make -C oracle``), with a direct
dense solve as the 6-argument solver the reference leaves unimplemented, and
stores inputs and outputs as small .npz files next to this script.  Data only.

    python tests/golden/make_golden_pred.py

Cases (seeded): structured n x n meshes of the unit square (constant
velocity) or of [0,10]^2 (rotation about (5,5)); random initial solution;
constant prescribed boundary value; alpha = 0.5; the multicorrection stopped
by the norm test or by the iteration count; the time loop stopped at the
final time or at steady state.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

# name, n, kind, seed, kappa, diag, stop_by_norm, steady, nl_tol, nl_maxiter, time_tol, dt, T
# (every correction loop ends well before nl_maxiter or runs exactly nl_maxiter
# iterations: a loop that stagnates at the cap makes the count of each step
# hinge on the last bits of the solve)
CASES = [
    ("pred_n8_const_norm_time", 8, 0, 1, 1e-3, 0, 1, 0, 1e-3, 20, 0.0, 0.05, 0.5),
    ("pred_n12_cone_diag_steady", 12, 3, 2, 0.1, 1, 1, 1, 1e-4, 30, 0.02, 0.1, 100.0),
    ("pred_n10_const_diag_iter", 10, 0, 3, 1e-3, 1, 0, 0, 0.0, 3, 0.0, 0.1, 1.0),
]


def make_case(n, kind, seed, kappa=1e-3):
    o = O.Oracle()
    rng = np.random.default_rng(seed)
    x, y, t, v = o.mesh2d(n)
    if kind == 3:
        x, y = 10.0 * x, 10.0 * y
    neq, idv = o.number_nodes(t)
    nel = len(v) // 3
    lm = o.fill_lm(neq, nel, v, t, idv)
    bx, by = (5.0, 5.0) if kind == 3 else (float(np.cos(1.1)), float(np.sin(1.1)))
    prob = np.array([kind, kappa, 0.1, 0.5, bx, by, 1.0, 0.0])
    u0 = 0.1 * rng.standard_normal(neq + 1)
    u0[neq] = 0.0
    return dict(x=x, y=y, type=t, idv=idv, vertex=v, lm=lm, neq=neq, prob=prob, ud=0.2, u0=u0)


def main():
    R = O.ReferencePredictor()
    for name, n, kind, seed, kappa, diag, norm, steady, nl_tol, maxit, time_tol, dt, T in CASES:
        c = make_case(n, kind, seed, kappa)
        p = c["prob"]
        R.setup(c["x"], c["y"], c["type"], c["idv"], c["vertex"], c["lm"], c["neq"], kind,
                p[1], p[2], p[3], p[4], p[5], p[6], c["ud"])
        params = dict(alpha=0.5, dt=dt, final_time=T, nl_tol=nl_tol, nl_maxiter=maxit,
                      stop_by_norm=norm, steady=steady, time_tol=time_tol, diag=diag)
        u, hist, nsteps, tf, rc = R.run(c["u0"], cap=400, **params)
        assert rc == 0
        out = dict(c, u=u, hist=hist, nsteps=nsteps, t_final=tf,
                   params=np.array([0.5, dt, T, nl_tol, maxit, norm, steady, time_tol, diag]))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", c["neq"], "steps", nsteps, "corrections", int(hist.sum()))


if __name__ == "__main__":
    main()
