"""Write the golden fixtures of the time integrator (Predictor_Old.c with
Criteria.c) over the Euler 2D SUPG build of 03_Euler_Equations, from the
reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference's
Predictor_Old, StopBy* criteria, ddot/daxpy, NO_precond_setup, NO_scaling /
Block_scaling_EBE and the Euler build (Build_M_K_F_SUPG.c with Delta_CAU,
Ax_Ay_calculations, rotation, eval_U_dU), compiled from their own sources
(oracle/_ref/libref_eu2d.so, built by ``make -C oracle``), with a direct
dense solve as the 6-argument solver the reference leaves unimplemented, and
stores inputs and outputs as small .npz files next to this script.  Data only.

    python tests/golden/make_golden_pred_eu.py

Inputs: the meshes, dofs, states, prescribed values and slip walls of
make_golden_eub.make_case (seeded); gamma 1.4, Mach 3, CAU tolerance 1e-10.
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

# name, n, seed, dimless, block_scaling, alpha, dt, nl_tol, nl_maxiter, stop_by_norm,
# steady, T
CASES = [
    ("predeu_n6_dimless", 6, 21, 1, 0, 1.0, 0.01, 1e-3, 10, 1, 0, 0.05),
    ("predeu_n6_dimless_scaled", 6, 21, 1, 1, 1.0, 0.01, 1e-3, 10, 1, 0, 0.05),
    ("predeu_n9_dimensional_iter", 9, 22, 0, 0, 1.0, 0.01, 0.0, 3, 0, 0, 0.04),
]


def load_eub():
    spec = importlib.util.spec_from_file_location("mk_eub",
                                                  os.path.join(HERE, "make_golden_eub.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    return mk


def main():
    mk = load_eub()
    R = O.ReferenceEuler()
    for name, n, seed, dimless, bs, alpha, dt, nl_tol, maxit, norm, steady, T in CASES:
        c = mk.make_case(n, seed)
        R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"],
                np.zeros(144 * c["nel"]), "NOT")
        R.build_setup(c["x"], c["y"], c["v1type"], 1.4, 3.0, dimless, alpha, dt, 1e-10,
                      mk.PRESC, mk.TH)
        u, hist, ns, tf, rc = R.pred_run(c["u"], alpha, dt, T, nl_tol, maxit, norm, steady,
                                         1e-3, 1e-10, bs, cap=200)
        assert rc == 0
        keep = ("x", "y", "vertex", "ids", "lm", "neq", "nnodes", "nel", "v1type", "theta")
        out = {k: c[k] for k in keep}
        out.update(u0=c["u"], u=u, hist=hist, nsteps=ns, t_final=tf, presc=mk.PRESC,
                   params=np.array([alpha, dt, T, nl_tol, maxit, norm, steady, 1e-3, bs,
                                    dimless]))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", c["neq"], "steps", ns, "corrections", int(hist.sum()))


if __name__ == "__main__":
    main()
