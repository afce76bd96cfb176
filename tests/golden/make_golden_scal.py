"""Write the golden fixtures of the diagonal scalings
(Solvers_and_Preconditioners/scaling.c: Left_scaling_EBE/EDE/CSR,
LeftRight_scaling_EBE/EDE/CSR, Left_unscaling) from the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives those routines
and the SS-2D structures compiled from their own sources
(oracle/_ref/libref_ss2d.so, built by ``make -C oracle``) on the matrices
and F of the SS-2D fixtures (tests/golden/ss2d_n8_rot_*, the build at a
seeded nonzero u), and stores the scaled matrix, F, Diag, invDiag and an
unscaled seeded vector as small .npz files next to this script.  Data only.

    python tests/golden/make_golden_scal.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402


def main():
    for scheme in ("ebe", "ede", "csr"):
        d = dict(np.load(os.path.join(HERE, f"ss2d_n8_rot_{scheme}.npz")))
        r = O.Reference()
        neq = r.setup(d["x"], d["y"], d["type"], d["vertex"], scheme.upper())
        assert neq == int(d["neq"])
        rng = np.random.default_rng(77)
        xv = rng.standard_normal(neq + 1)
        for mode in ("Left", "LeftRight"):
            A, F, D, iD = r.scaling(mode, d["A_u1"], d["F_u1"])
            x = r.unscaling(xv)
            out = {k: d[k] for k in ("x", "y", "type", "vertex", "neq", "lm") if k in d}
            for k in ("IA", "JA", "order", "edge_by_elem", "lm2"):
                if k in d:
                    out[k] = d[k]
            out.update(scheme=scheme, mode=mode, A_in=d["A_u1"], F_in=d["F_u1"], A=A, F=F,
                       Diag=D, invDiag=iD, xv=xv, x_unscaled=x)
            path = os.path.join(HERE, f"scal_n8_rot_{scheme}_{mode.lower()}.npz")
            np.savez_compressed(path, **out)
            print("wrote", path)


if __name__ == "__main__":
    main()
