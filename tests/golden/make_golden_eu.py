"""Write the golden fixtures of the 4-dof (Euler 2D) element products and
node block-diagonal routines from the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference
routines compiled from their own sources (oracle/_ref/libref_eu2d.so, built
by ``make -C oracle``: ebemvNDOF4, ebe2mvNDOF4 with BlockDiag2_precond /
NOBlockDiag2_precond, BlockDiag_precond_EBE_setup, BlockDiag_precond,
Block_scaling_EBE) and stores inputs and outputs as small .npz files next to
this script.  The fixtures are data only; the reference code does not travel.

    python tests/golden/make_golden_eu.py

Inputs (seeded, synthetic: the Euler element assembly is not on this path):
the oracle's structured n x n triangle mesh, 4 dofs per node numbered as
Fill_ID.c:3-40 does (node by node, dof by dof, prescribed dofs -1), lm as
Fill_LM.c:3-24 (DOF fastest per node, prescribed -> neq), 12x12 element
matrices = 3.5 I + 0.1 N(0,1).  About 10% of the dofs are prescribed at
random, plus every dof of the three nodes of element 0, so that element has
no free dof and its nodes have none either.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CASES = [("eu2d_n6_seed2", 6, 2), ("eu2d_n12_seed5", 12, 5)]


def make_case(n, seed):
    o = O.Oracle()
    rng = np.random.default_rng(seed)
    x, y, _, v = o.mesh2d(n)
    nn, nel = len(x), len(v) // 3
    pres = rng.random(4 * nn) < 0.1
    for a in range(3):
        pres[4 * v[a]:4 * v[a] + 4] = True
    ids = np.full(4 * nn, -1, np.int32)
    neq = int((~pres).sum())
    ids[~pres] = np.arange(neq, dtype=np.int32)
    lm = np.zeros(12 * nel, np.int32)
    for e in range(nel):
        for a in range(3):
            for d in range(4):
                g = ids[4 * v[3 * e + a] + d]
                lm[12 * e + 4 * a + d] = neq if g < 0 else g
    A = (rng.standard_normal((nel, 12, 12)) * 0.1 + 3.5 * np.eye(12)).reshape(-1).copy()
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0
    F = rng.standard_normal(neq + 1)
    F[neq] = 0.0
    return dict(nnodes=nn, nel=nel, neq=neq, vertex=v.astype(np.int32), ids=ids, lm=lm,
                A=A, P=P, F=F)


def main():
    R = O.ReferenceEuler()
    for name, n, seed in CASES:
        c = make_case(n, seed)
        nn, nel, neq = c["nnodes"], c["nel"], c["neq"]
        R.setup(nn, nel, neq, c["vertex"], c["ids"], c["lm"], c["A"], "BlockDiag")
        out = dict(c)
        out["Q"] = R.ebemv(c["P"])
        F_bd, BD, iBD, Id = R.blockdiag_setup(c["F"])
        out.update(F_bd=F_bd, BD=BD, invBD=iBD, Id=Id)
        out["z_bd"] = R.blockdiag_apply(c["P"])
        out["Q2_bd"] = R.ebe2mv(c["P"])
        R.setup(nn, nel, neq, c["vertex"], c["ids"], c["lm"], c["A"], "NOT")
        R.blockdiag_setup(c["F"])
        out["Q2_not"] = R.ebe2mv(c["P"])
        R.setup(nn, nel, neq, c["vertex"], c["ids"], c["lm"], c["A"], "BlockDiag")
        A_sc, F_sc = R.block_scaling(c["F"])
        out.update(A_sc=A_sc, F_sc=F_sc)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "neq", neq, "nel", nel)


if __name__ == "__main__":
    main()
