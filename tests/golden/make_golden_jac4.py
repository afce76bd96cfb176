"""Write the golden fixtures of the Euler driver's element Jacobi
(JacobiDOF4_precond_EBE_setup.c / JacobiDOF4_precond_EBE.c, what
03_Euler_Equations/2D/01_CommonFiles/setPreconditioner.c:89-92 selects for
"Jacobi" on the EBE scheme) from the reference itself.

Runs in a container that holds the FEM_CODES tree: the reference routines are
compiled from their own sources into oracle/_ref/libref_eu2d.so by
``make -C oracle``; the inputs are those of make_golden_eu.py (the same
seeded meshes, ids, lm, 12x12 element matrices, P and F), which the
eu2d_n6_seed2 / eu2d_n12_seed5 fixtures already hold, so only the outputs are
stored: invDe [nel*12], F after the setup (left-preconditioned in place), and
z = Jacobi(P) called out of place, as the setup calls it.  Data only; the
reference code does not travel.

    python tests/golden/make_golden_jac4.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from make_golden_eu import make_case  # noqa: E402

CASES = [("jac4_n6_seed2", 6, 2), ("jac4_n12_seed5", 12, 5)]
INPUTS = {"jac4_n6_seed2": "eu2d_n6_seed2", "jac4_n12_seed5": "eu2d_n12_seed5"}


def main():
    R = O.ReferenceEuler()
    for name, n, seed in CASES:
        c = make_case(n, seed)
        R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"], c["A"],
                "Jacobi")
        invDe, F_jac = R.jacobi4_setup(c["F"])
        z_jac = R.jacobi4_apply(c["P"])
        with np.load(os.path.join(HERE, INPUTS[name] + ".npz")) as d:
            assert np.array_equal(d["A"], c["A"]) and np.array_equal(d["F"], c["F"])
        out = dict(inputs=INPUTS[name], neq=c["neq"], nel=c["nel"], invDe=invDe, F_jac=F_jac,
                   z_jac=z_jac)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", c["neq"], "nel", c["nel"])


if __name__ == "__main__":
    main()
