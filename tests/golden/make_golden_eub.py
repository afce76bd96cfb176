"""Write the golden fixtures of the Euler 2D SUPG element build
(03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c, CAU operator) from
the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference
routine compiled from its own sources with that driver's Delta_CAU.c,
Ax_Ay_calculations.c, rotation.c, eval_U_dU.c, ebe_assembly.c and setzeros.c
(oracle/_ref/libref_eu2d.so, built by ``make -C oracle``) and stores inputs
and outputs as small .npz files next to this script.  Data only.

    python tests/golden/make_golden_eub.py

Inputs (seeded, synthetic): the mesh, node dofs and lm of
make_golden_eu.make_case; a nodal state rho ~ 1 +- 5%, v1 ~ 1 +- 10%,
v2 ~ +-0.1, E = |v|^2/2 + 2 + U[0, 0.1] in conserved variables; small
random time derivatives; prescribed values (1, 1.1, 0.05, 2.7); 15% of the
nodes flagged as slip walls (v1Type = -1) with theta = 0.3 + 0.5 x - 0.2 y;
gamma 1.4, Mach 3, alpha 1, dt 0.01, CAU tolerance 1e-10; one case with the
dimensionless Jacobians, one with the dimensional ones.
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CASES = [("eub2d_n6_dimless", 6, 21, 1), ("eub2d_n9_dimensional", 9, 22, 0)]
PRESC = np.array([1.0, 1.1, 0.05, 2.7])
TH = np.array([0.3, 0.5, 0.2])


def make_case(n, seed):
    spec = importlib.util.spec_from_file_location("mk_eu", os.path.join(HERE, "make_golden_eu.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    c = mk.make_case(n, seed)
    o = O.Oracle()
    x, y, _, _ = o.mesh2d(n)
    rng = np.random.default_rng(seed + 1000)
    nn, neq, ids = c["nnodes"], c["neq"], c["ids"]
    rho = 1 + 0.05 * rng.standard_normal(nn)
    v1 = 1 + 0.1 * rng.standard_normal(nn)
    v2 = 0.1 * rng.standard_normal(nn)
    E = 0.5 * (v1 ** 2 + v2 ** 2) + 2 + 0.1 * rng.random(nn)
    Un = np.stack([rho, rho * v1, rho * v2, rho * E], 1).reshape(-1)
    dUn = 0.01 * rng.standard_normal(4 * nn)
    pres = ids < 0
    u, du = np.zeros(neq + 1), np.zeros(neq + 1)
    u[ids[~pres]] = Un[~pres]
    du[ids[~pres]] = dUn[~pres]
    U = np.where(pres, np.tile(PRESC, nn), Un)
    dU = np.where(pres, 0.0, dUn)
    v1type = np.where(rng.random(nn) < 0.15, -1, 1).astype(np.int32)
    theta = TH[0] + TH[1] * x - TH[2] * y
    return dict(c, x=x, y=y, u=u, du=du, U=U, dU=dU, v1type=v1type, theta=theta)


def main():
    R = O.ReferenceEuler()
    for name, n, seed, dimless in CASES:
        c = make_case(n, seed)
        R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"],
                np.zeros(144 * c["nel"]), "NOT")
        R.build_setup(c["x"], c["y"], c["v1type"], 1.4, 3.0, dimless, 1.0, 0.01, 1e-10, PRESC, TH)
        A, Rv = R.build(c["u"], c["du"])
        out = {k: v for k, v in c.items() if k not in ("A", "P", "F")}
        out.update(dimless=dimless, A_build=A, R_build=Rv)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", c["neq"], "nel", c["nel"])


if __name__ == "__main__":
    main()
