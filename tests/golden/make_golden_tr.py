"""Write the golden fixtures of the transient 2D SUPG element build
(02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c) from the
reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference
routine compiled from its own sources with that driver's CAU shock capture,
h_shock_2sqrtArea, eval_U_dU, ebe_assembly and setzeros
(oracle/_ref/libref_tr2d.so, built by ``make -C oracle``) and stores inputs
and outputs as small .npz files next to this script.  Data only.

    python tests/golden/make_golden_tr.py

Cases (seeded): the oracle's structured n x n triangle mesh of the unit
square (constant velocity) or of [0,10]^2 (rotation about (5,5), the CONE
problem's velocity); random u, du on the free nodes, a constant prescribed
value on the boundary, alpha = 0.5, dt = 0.01 (and 0.05).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CASES = [("tr2d_n8_const", 8, 1, 0, 0.5, 0.01), ("tr2d_n12_cone", 12, 2, 3, 0.5, 0.05)]


def make_case(n, seed, kind):
    o = O.Oracle()
    rng = np.random.default_rng(seed)
    x, y, t, v = o.mesh2d(n)
    if kind == 3:
        x, y = 10.0 * x, 10.0 * y
    neq, idv = o.number_nodes(t)
    nel = len(v) // 3
    lm = o.fill_lm(neq, nel, v, t, idv)
    kappa = float(10 ** rng.uniform(-5, -2))
    gamma, f = float(rng.uniform(0, 1)), float(rng.uniform(-1, 1))
    bx, by = (5.0, 5.0) if kind == 3 else (float(np.cos(1.1)), float(np.sin(1.1)))
    ud = float(rng.uniform(-1, 1))
    u = rng.standard_normal(neq + 1)
    du = rng.standard_normal(neq + 1)
    u[neq] = du[neq] = 0.0
    U = np.where(t > 0, u[np.maximum(idv, 0)], ud)
    dU = np.where(t > 0, du[np.maximum(idv, 0)], 0.0)
    return dict(x=x, y=y, type=t, idv=idv, vertex=v, lm=lm, neq=neq, u=u, du=du, U=U, dU=dU,
                prob=np.array([kind, kappa, gamma, f, bx, by, 1.0, 0.0]), ud=ud)


def main():
    R = O.ReferenceTransient()
    for name, n, seed, kind, alpha, dt in CASES:
        c = make_case(n, seed, kind)
        p = c["prob"]
        R.setup(c["x"], c["y"], c["type"], c["idv"], c["vertex"], c["lm"], c["neq"], kind,
                p[1], p[2], p[3], p[4], p[5], p[6], c["ud"], alpha, dt)
        A, Rv = R.build(c["u"], c["du"])
        out = dict(c, alpha=alpha, dt=dt, A=A, R=Rv)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", c["neq"])


if __name__ == "__main__":
    main()
