"""Write the golden fixtures of the Euler driver's edge-by-edge scheme
(03_Euler_Equations/2D/01_CommonFiles/ede_Initialization.c, ede_assembly.c;
00_CommonFiles EDE/edemvNDOF4.c, EDE/ede2mvNDOF4.c with BlockDiag2_precond,
BlockDiag_precond_EDE_setup.c) from the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives those routines
compiled from their own sources (oracle/_ref/libref_eu2d.so, built by
``make -C oracle``) and stores inputs and outputs as small .npz files next to
this script.  Data only.

    python tests/golden/make_golden_ede4.py

Inputs (seeded, synthetic): the oracle's structured n x n triangle mesh with
the nodes relabelled by a random permutation and each element's vertices in
a random one of their 6 orders (so every branch of the vertex sort is taken);
4 dofs per node numbered node by node (Fill_ID.c), about 10% prescribed (-1)
plus every dof of element 0's nodes; 12x12 element matrices 3.5 I +
0.1 N(0,1); random P and F (entry neq = 0).
"""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CASES = [("ede4_n5_seed3", 5, 3), ("ede4_n11_seed8", 11, 8)]
PERMS = np.array(list(itertools.permutations(range(3))), np.int32)


def make_case(n, seed):
    o = O.Oracle()
    rng = np.random.default_rng(seed)
    x, y, _, v = o.mesh2d(n)
    nn, nel = len(x), len(v) // 3
    relabel = rng.permutation(nn).astype(np.int32)
    v = relabel[v].reshape(nel, 3)
    v = np.take_along_axis(v, PERMS[rng.integers(0, 6, nel)], axis=1).reshape(-1).copy()
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
    A12 = (rng.standard_normal((nel, 12, 12)) * 0.1 + 3.5 * np.eye(12)).reshape(-1).copy()
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0
    F = rng.standard_normal(neq + 1)
    F[neq] = 0.0
    return dict(nnodes=nn, nel=nel, neq=neq, vertex=v.astype(np.int32), ids=ids, lm=lm,
                A12=A12, P=P, F=F)


def main():
    R = O.ReferenceEuler()
    for name, n, seed in CASES:
        c = make_case(n, seed)
        R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"], c["A12"],
                "BlockDiag")
        order, ebe, lm2 = R.ede_setup()
        Aede = R.ede_assemble(c["A12"])
        Q = R.edemv(c["P"])
        F_bd, BD, iBD, Id = R.ede_blockdiag_setup(c["F"])
        Q2 = R.ede2mv(c["P"])
        out = dict(c, order=order, edge_by_elem=ebe, lm2=lm2, nedge=len(lm2) // 8, Aede=Aede,
                   Q=Q, F_bd=F_bd, BD=BD, invBD=iBD, Id=Id, Q2=Q2)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, "nel", c["nel"], "nedge", len(lm2) // 8, "neq", c["neq"])


if __name__ == "__main__":
    main()
