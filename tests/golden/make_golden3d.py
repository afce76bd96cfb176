"""Write the golden fixtures of the 3D steady-transport solve path from the
reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference
routines compiled from their own sources (oracle/_ref/libref_ss3d.so, built by
``make -C oracle``: 3D Build_SUPG with the DIFCONV problem, CSR ILU(0)) and
stores inputs and outputs as small .npz files next to this script.  The
fixtures are data only; the reference code does not travel with them.

    python tests/golden/make_golden3d.py

Cases:
  * ss3d_kuhn4: the unit cube cut into 4^3 cells of 6 tetrahedra (the
    oracle's Kuhn mesh generator; 125 nodes, 384 elements, 27 equations);
  * ss3d_difconv_125_384: the reference's own mesh
    02_Mesh/DIFCONV-DIFREA-DIFCONVREA/DIFCONV_125_384.dat (FREEFEM format:
    node type 0 = unknown, elements 1-based), as Preprocess.c:136-199 reads it.
Each file holds: mesh, lm, the 4x4 element matrices and F of Build_SUPG
(DIFCONV, ConstApli = 1), the CSR structure (csr_Initialization with
reordering NOT) and matrix, Q = A P for a seeded P, the ILU(0) factors of
ILUp_precond_setup ("ILU0"), F after that setup (left-preconditioned), and
z = ILUp_precond(P).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

REF_MESH = ("/root/reference/01_SS_Transport_Equation/3D/02_Mesh/"
            "DIFCONV-DIFREA-DIFCONVREA/DIFCONV_125_384.dat")


def read_freefem(path):
    """3D Preprocess.c:136-199 for OriginMesh FREEFEM"""
    tok = open(path).read().split()
    pos = 0
    nn = int(tok[pos]); pos += 1
    xyz = np.zeros((nn, 3))
    ut = np.zeros(nn, np.int32)
    for i in range(nn):
        xyz[i] = [float(tok[pos]), float(tok[pos + 1]), float(tok[pos + 2])]
        ut[i] = 1 if int(tok[pos + 3]) == 0 else 0
        pos += 4
    nel = int(tok[pos]); pos += 1
    v = np.zeros(4 * nel, np.int32)
    for e in range(nel):
        v[4 * e:4 * e + 4] = [int(tok[pos + j]) - 1 for j in range(4)]
        pos += 5
    return xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), ut, v


def one_case(name, x, y, z, ut, v):
    o, r = O.Oracle(), O.Reference3D()
    neq = r.setup(x, y, z, ut, v, 1.0)
    nel = len(v) // 4
    lm = r.lm()
    ke, F = r.build()
    nnz, IA, JA, cbe = o.csr_init_k(neq, nel, 4, lm)
    AA = o.csr_assembly_k(nel, 4, nnz, ke, cbe)
    rng = np.random.default_rng(3000 + nel)
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0
    Q = o.ebemv(neq, nel, 4, ke, lm, P.copy())
    LU, D, Filu = r.ilu0_setup(neq, AA[:nnz], JA, IA, F)
    zilu = r.ilu0_apply(P)
    out = dict(x=x, y=y, z=z, utype=ut, vertex=v, neq=neq, lm=lm, ke=ke, F=F, IA=IA, JA=JA,
               csr_by_elem=cbe, AA=AA, P=P, Q=Q, LU=LU, D=D[:neq], F_ilu=Filu, z_ilu=zilu)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "neq", neq, "nel", nel, "nnz", nnz)


def rcm_case(name, n, seed):
    """Reordering/symrcm.c on the CSR pattern of a Kuhn mesh whose nodes were
    numbered by a seeded random permutation: the pattern before, P, and the
    pattern after MATRIX_ROW/COL_permutation"""
    import ctypes as C
    o = O.Oracle()
    x, y, z, ut, v = o.mesh3d(n)
    nn = len(x)
    perm = np.random.default_rng(seed).permutation(nn)
    inv = np.empty(nn, np.int64)
    inv[perm] = np.arange(nn)
    ut = ut[perm]
    v = inv[v].astype(np.int32)
    neq, idv = o.number_nodes(ut)
    nel = len(v) // 4
    lm = o.fill_lm(neq, nel, v, ut, idv, nnoel=4)
    nnz, IA, JA, _ = o.csr_init_k(neq, nel, 4, lm)
    ref = C.CDLL(O.Reference.PATH)
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    ref.ref_symrcm.argtypes = [C.c_int, C.c_int, ip, ip, ip, ip]
    IA2, JA2 = IA.copy(), JA.copy()
    P = np.zeros(neq, np.int32)
    pT = np.zeros(nnz, np.int32)
    ref.ref_symrcm(neq, nnz, IA2, JA2, P, pT)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), neq=neq, lm=lm, IA=IA, JA=JA, P=P,
                        IA_rcm=IA2, JA_rcm=JA2, pT=pT)
    print(name, "neq", neq, "nnz", nnz)


def main():
    if not (O.Reference3D.available() and O.Reference.available()):
        sys.exit("oracle/_ref/libref_ss3d.so missing: run make -C oracle where the "
                 "FEM_CODES tree is present")
    o = O.Oracle()
    x, y, z, ut, v = o.mesh3d(4)
    one_case("ss3d_kuhn4", x, y, z, ut, v)
    one_case("ss3d_difconv_125_384", *read_freefem(REF_MESH))
    rcm_case("ss3d_rcm_n6_seed3", 6, 3)


if __name__ == "__main__":
    main()
