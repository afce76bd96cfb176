"""Write the golden fixtures of the time integrator over the Euler driver's
other matrix-vector schemes (setMatrixVectorProductType.c and
setPreconditioner.c of 03_Euler_Equations): EBE2 (ebe2mvNDOF4 with
BlockDiag_precond_EBE_setup), EDE (ede_assembly, edemvNDOF4, NO_precond) and
EDE2 (ede_assembly, ede2mvNDOF4 with BlockDiag_precond_EDE_setup), from the
reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference's
Predictor_Old, criteria, Euler build, ede_Initialization and those products
and setups, compiled from their own sources (oracle/_ref/libref_eu2d.so,
built by ``make -C oracle``), with a direct dense solve of the matrix that
the scheme's own product gives column by column as the 6-argument solver the
reference leaves unimplemented.  Stores small .npz files next to this
script.  Data only.

    python tests/golden/make_golden_pred_eus.py

The EBE2 and EDE2 systems are not the EBE system where a node has both free
and prescribed dofs: the node block is inverted whole, prescribed rows and
columns included, and applied to the free dofs only.
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

SCHEMES = {1: "ebe2", 2: "ede", 3: "ede2"}
# n, seed, dimless, alpha, dt, nl_tol, nl_maxiter, stop_by_norm, steady, T, schemes
CASES = [
    ("n6_dimless", 6, 21, 1, 1.0, 0.01, 1e-3, 10, 1, 0, 0.05, (1, 2, 3)),
    ("n9_dimensional_iter", 9, 22, 0, 1.0, 0.01, 0.0, 3, 0, 0, 0.04, (1, 3)),
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
    for name, n, seed, dimless, alpha, dt, nl_tol, maxit, norm, steady, T, schemes in CASES:
        c = mk.make_case(n, seed)
        for sch in schemes:
            R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"],
                    np.zeros(144 * c["nel"]), "BlockDiag" if sch in (1, 3) else "NOT")
            R.build_setup(c["x"], c["y"], c["v1type"], 1.4, 3.0, dimless, alpha, dt, 1e-10,
                          mk.PRESC, mk.TH)
            if sch >= 2:
                R.ede_setup()
            u, hist, ns, tf, rc = R.pred_run(c["u"], alpha, dt, T, nl_tol, maxit, norm, steady,
                                             1e-3, 1e-10, 0, cap=200, scheme=sch)
            assert rc == 0
            keep = ("x", "y", "vertex", "ids", "lm", "neq", "nnodes", "nel", "v1type", "theta")
            out = {k: c[k] for k in keep}
            out.update(u0=c["u"], u=u, hist=hist, nsteps=ns, t_final=tf, presc=mk.PRESC,
                       scheme=sch,
                       params=np.array([alpha, dt, T, nl_tol, maxit, norm, steady, 1e-3, 0,
                                        dimless]))
            path = os.path.join(HERE, f"predsch_{name}_{SCHEMES[sch]}.npz")
            np.savez_compressed(path, **out)
            print("wrote", path, "neq", c["neq"], "steps", ns, "corrections", int(hist.sum()))


if __name__ == "__main__":
    main()
