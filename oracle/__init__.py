"""Python bindings of the oracle (CPU restatement of the FEM_CODES solve path).

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module.  The GPU product
(``fem_codes_amd``) never does.

``Oracle`` wraps ``oracle/liboracle.so`` (order-faithful restatement, built
``-O2 -ffp-contract=off``).  ``Reference`` wraps ``oracle/_ref/libref_ss2d.so``,
the reference's own SS-2D routines compiled from the FEM_CODES tree, when that
library has been built (only where the tree is present).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i, _d, _vp = C.c_int, C.c_double, C.c_void_p

P_CONST, P_ROTATING, P_VALIDACAO = 0, 1, 2
H_2SQRTAREA, H_SQRTAREA = 0, 1


class OrcProb2d(C.Structure):
    _fields_ = [("kind", C.c_int), ("kappa", C.c_double), ("gamma", C.c_double),
                ("f", C.c_double), ("bx", C.c_double), ("by", C.c_double),
                ("scale", C.c_double), ("h_shock", C.c_int)]


P3_DIFCONV, P3_CONST = 0, 1


class OrcProb3d(C.Structure):
    _fields_ = [("kind", C.c_int), ("A", C.c_double), ("kappa", C.c_double),
                ("bx", C.c_double), ("by", C.c_double), ("bz", C.c_double),
                ("sigma", C.c_double), ("f", C.c_double)]


def prob3d(kind=P3_DIFCONV, A=1.0, kappa=0.0, bx=0.0, by=0.0, bz=0.0, sigma=0.0, f=0.0):
    return OrcProb3d(kind, A, kappa, bx, by, bz, sigma, f)


class OrcIluCtx(C.Structure):
    _fields_ = [("neq", C.c_int), ("LU", C.c_void_p), ("JA", C.c_void_p), ("IA", C.c_void_p),
                ("D", C.c_void_p)]


OP_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


class OrcOps(C.Structure):
    _fields_ = [("matvec", C.c_void_p), ("matvec_ctx", C.c_void_p),
                ("precond", C.c_void_p), ("precond_ctx", C.c_void_p),
                ("precondR", C.c_void_p), ("precondR_ctx", C.c_void_p)]


class OrcCsrCtx(C.Structure):
    _fields_ = [("neq", C.c_int), ("AA", C.c_void_p), ("JA", C.c_void_p), ("IA", C.c_void_p)]


class OrcEbeCtx(C.Structure):
    _fields_ = [("neq", C.c_int), ("nel", C.c_int), ("k", C.c_int),
                ("A", C.c_void_p), ("lm", C.c_void_p)]


class OrcEdeCtx(C.Structure):
    _fields_ = [("neq", C.c_int), ("nedge", C.c_int), ("A", C.c_void_p), ("lm2", C.c_void_p)]


class OrcDiagCtx(C.Structure):
    _fields_ = [("neq", C.c_int), ("invDiag", C.c_void_p)]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref when the FEM_CODES tree is present)."""
    out = subprocess.run(["make", "-s", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Thin ctypes layer over liboracle.so; all arrays are numpy."""

    def __init__(self, fast: bool = False):
        name = "liboracle_fast.so" if fast else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        L = self.lib = C.CDLL(path)
        L.orc_ddot.restype = _d
        L.orc_ddot.argtypes = [_i, _dp, _dp]
        L.orc_csrmv.argtypes = [_i, _dp, _ip, _ip, _dp, _dp]
        L.orc_ebemv.argtypes = [_i, _i, _i, _dp, _ip, _dp, _dp]
        L.orc_edemv.argtypes = [_i, _i, _dp, _ip, _dp, _dp]
        L.orc_diag_apply.argtypes = [_i, _dp, _dp, _dp]
        L.orc_diag_setup_ebe.argtypes = [_i, _i, _i, _dp, _ip, _dp, _dp, _dp]
        L.orc_diag_setup_ede.argtypes = [_i, _i, _dp, _ip, _dp, _dp, _dp]
        L.orc_diag_setup_csr.argtypes = [_i, _dp, _ip, _ip, _dp, _dp, _dp]
        L.orc_jacobi_ebe_setup.argtypes = [_i, _i, _dp, _ip, _dp, _dp]
        L.orc_jacobi_ebe_apply.argtypes = [_i, _i, _dp, _ip, _dp, _dp]
        L.orc_jacobi_ebe_setup_k.argtypes = [_i, _i, _i, _dp, _ip, _dp, _dp]
        L.orc_jacobi_ebe_apply_k.argtypes = [_i, _i, _i, _dp, _ip, _dp, _dp]
        L.orc_pgmres.restype = _i
        L.orc_pgmres.argtypes = [_i, _i, _i, _d, _dp, _dp, C.POINTER(OrcOps),
                                 _vp, _i, _vp, C.POINTER(C.c_int)]
        L.orc_mesh2d_structured.argtypes = [_i, _dp, _dp, _ip, _ip]
        L.orc_number_nodes.restype = _i
        L.orc_number_nodes.argtypes = [_i, _ip, _ip]
        L.orc_fill_lm.argtypes = [_i, _i, _i, _ip, _ip, _ip, _ip]
        L.orc_csr_init.restype = _i
        L.orc_csr_init.argtypes = [_i, _i, _ip, _vp, _vp, _vp]
        L.orc_ede_init.restype = _i
        L.orc_ede_init.argtypes = [_i, _i, _ip, _vp, _vp, _vp]
        L.orc_build_supg2d.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _dp,
                                       C.POINTER(OrcProb2d), _dp, _dp]
        L.orc_ede_assembly.argtypes = [_i, _i, _dp, _ip, _ip, _dp]
        L.orc_csr_assembly.argtypes = [_i, _i, _dp, _ip, _dp]
        for nm in ("orc_op_csr", "orc_op_ebe", "orc_op_ede", "orc_op_diag", "orc_op_ilu0"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_csr_init_k.restype = _i
        L.orc_csr_init_k.argtypes = [_i, _i, _i, _ip, _vp, _vp, _vp]
        L.orc_csr_assembly_k.argtypes = [_i, _i, _i, _dp, _ip, _dp]
        L.orc_ilu0_factor.restype = _i
        L.orc_ilu0_factor.argtypes = [_i, _dp, _ip, _ip, _dp, _dp]
        L.orc_ilu0_apply.argtypes = [_i, _dp, _ip, _ip, _dp, _dp, _dp]
        L.orc_ilu0_apply_f32.argtypes = [_i, _fp, _ip, _ip, _fp, _fp, _fp]
        L.orc_mesh3d_structured.argtypes = [_i, _dp, _dp, _dp, _ip, _ip]
        L.orc_prob3d_f.restype = _d
        L.orc_prob3d_f.argtypes = [C.POINTER(OrcProb3d), _d, _d, _d]
        L.orc_build_supg3d.argtypes = [_i, _i, _dp, _dp, _dp, _ip, _ip, _ip,
                                       C.POINTER(OrcProb3d), _dp, _dp]
        L.orc_symrcm.argtypes = [_i, _ip, _ip, _ip]
        L.orc_ebe2mv_ndof4.argtypes = [_i, _i, _dp, _ip, _i, _ip, _vp, _vp, _dp, _dp]
        L.orc_matrix_inverse4.argtypes = [_dp, _dp]
        L.orc_blockdiag_apply.argtypes = [_i, _i, _ip, _dp, _dp, _dp]
        L.orc_blockdiag_setup_ebe.argtypes = [_i, _i, _i, _ip, _dp, _ip, _dp, _dp, _vp]
        L.orc_block_scaling_ebe.argtypes = [_i, _i, _i, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        L.orc_ebe2mv_items.argtypes = [_i, _i, _i, _dp, _ip, _i, _ip, _vp, _vp, _dp, _dp]
        L.orc_blockdiag_setup_items.argtypes = [_i, _i, _i, _i, _ip, _dp, _ip, _dp, _dp, _vp]
        L.orc_scaling_items.argtypes = [_i, _i, _i, _dp, _ip, _dp, _dp, _i, _dp]
        L.orc_scaling_csr.argtypes = [_i, _dp, _ip, _ip, _dp, _dp, _i, _dp]
        L.orc_left_unscaling.argtypes = [_i, _dp, _dp]
        L.orc_ede4_init.restype = _i
        L.orc_ede4_init.argtypes = [_i, _i, _i, _ip, _ip, _ip, _ip, _vp]
        L.orc_ede4_edge_nodes.argtypes = [_i, _ip, _ip, _ip, _ip]
        L.orc_ede4_assembly.argtypes = [_i, _i, _dp, _ip, _ip, _dp]

    def read_params(self, path, kinds):
        """values of a parameter file read with the drivers' fscanf formats
        ('d' as %.17g text)"""
        L = self.lib
        L.orc_read_params.restype = _i
        L.orc_read_params.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, _i]
        buf = C.create_string_buffer(1 << 16)
        n = L.orc_read_params(os.fsencode(path), kinds.encode(), buf, len(buf))
        return [v.decode() for v in buf.raw.split(b"\0")[:max(n, 0)]]

    def read_mesh(self, path, ndim, nint, nvert, ecols):
        L = self.lib
        L.orc_read_mesh.restype = _i
        L.orc_read_mesh.argtypes = [C.c_char_p, _i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i),
                                    _vp, _vp, _vp, _vp]
        nn, ne = _i(), _i()
        p = os.fsencode(path)
        assert L.orc_read_mesh(p, ndim, nint, nvert, ecols, C.byref(nn), C.byref(ne), None, None,
                               None, None) == 0
        xyz = np.zeros((nn.value, ndim))
        ni = np.zeros((nn.value, nint), np.int32)
        v = np.zeros(ne.value * nvert, np.int32)
        ei = np.zeros((ne.value, ecols), np.int32)
        assert L.orc_read_mesh(p, ndim, nint, nvert, ecols, C.byref(nn), C.byref(ne),
                               xyz.ctypes.data, ni.ctypes.data, v.ctypes.data,
                               ei.ctypes.data) == 0
        return xyz, ni, v, ei

    def scaling_items(self, neq, k, A, lm, Diag, invDiag, leftright, F):
        A2, F2 = np.ascontiguousarray(A, np.float64).copy(), np.ascontiguousarray(F, np.float64).copy()
        lm = np.ascontiguousarray(lm, np.int32)
        self.lib.orc_scaling_items(neq, lm.size // k, k, A2, lm, Diag, invDiag, int(leftright), F2)
        return A2, F2

    def scaling_csr(self, neq, AA, JA, IA, Diag, invDiag, leftright, F):
        A2, F2 = np.ascontiguousarray(AA, np.float64).copy(), np.ascontiguousarray(F, np.float64).copy()
        self.lib.orc_scaling_csr(neq, A2, JA, IA, Diag, invDiag, int(leftright), F2)
        return A2, F2

    def left_unscaling(self, neq, invDiag, x):
        x = np.ascontiguousarray(x, np.float64).copy()
        self.lib.orc_left_unscaling(neq, invDiag, x)
        return x

    def ede4_init(self, nnodes, vertex, ids, neq):
        """ede_Initialization.c of the Euler driver: (order, edge_by_elem,
        lm2, edge_nodes)"""
        v = np.ascontiguousarray(vertex, np.int32)
        ids = np.ascontiguousarray(ids, np.int32)
        nel = v.size // 3
        o, e = np.zeros(3 * nel, np.int32), np.zeros(3 * nel, np.int32)
        ne = self.lib.orc_ede4_init(nnodes, nel, neq, v, ids, o, e, None)
        lm2 = np.zeros(8 * ne, np.int32)
        self.lib.orc_ede4_init(nnodes, nel, neq, v, ids, o, e, _ptr(lm2))
        en = np.zeros(2 * ne, np.int32)
        self.lib.orc_ede4_edge_nodes(nel, v, o, e, en)
        return o, e, lm2, en

    def ede4_assembly(self, nedge, A12, order, edge_by_elem):
        Aede = np.zeros(64 * nedge)
        self.lib.orc_ede4_assembly(len(order) // 3, nedge, np.ascontiguousarray(A12, np.float64),
                                   order, edge_by_elem, Aede)
        return Aede

    def ebe2mv_items(self, neq, k, A, lm, nnodes, Id, P, invBD=None, BD=None):
        Q = np.zeros(neq + 1)
        self.lib.orc_ebe2mv_items(neq, len(lm) // k, k, A, lm, nnodes, Id,
                                  _ptr(invBD) if invBD is not None else None,
                                  _ptr(BD) if BD is not None else None,
                                  np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q

    def blockdiag_setup_items(self, neq, nv, nnodes, nodes, A, Id, F=None):
        BD, iBD = np.zeros(16 * nnodes), np.zeros(16 * nnodes)
        Fo = None if F is None else np.ascontiguousarray(F, np.float64).copy()
        nodes = np.ascontiguousarray(nodes, np.int32)
        self.lib.orc_blockdiag_setup_items(neq, nodes.size // nv, nv, nnodes, nodes,
                                           np.ascontiguousarray(A, np.float64), Id, BD, iBD,
                                           _ptr(Fo) if Fo is not None else None)
        return BD, iBD, Fo

    def block_scaling_ebe(self, neq, nel, nnodes, vertex, ids, Id, A, F):
        A2 = np.ascontiguousarray(A, np.float64).copy()
        F2 = np.ascontiguousarray(F, np.float64).copy()
        BD, iBD = np.zeros(16 * nnodes), np.zeros(16 * nnodes)
        self.lib.orc_block_scaling_ebe(neq, nel, nnodes, vertex, ids, Id, A2, BD, iBD, F2)
        return A2, F2, BD, iBD

    def ebe2mv_ndof4(self, neq, nel, A, lm, nnodes, Id, P, invBD=None, BD=None):
        Q = np.zeros(neq + 1)
        self.lib.orc_ebe2mv_ndof4(neq, nel, A, lm, nnodes, Id,
                                  _ptr(invBD) if invBD is not None else None,
                                  _ptr(BD) if BD is not None else None, P, Q)
        return Q

    def blockdiag_setup_ebe(self, neq, nel, nnodes, vertex, A, Id, F=None):
        BD, iBD = np.zeros(16 * nnodes), np.zeros(16 * nnodes)
        Fo = None if F is None else np.ascontiguousarray(F, np.float64).copy()
        self.lib.orc_blockdiag_setup_ebe(neq, nel, nnodes, vertex, A, Id, BD, iBD,
                                         _ptr(Fo) if Fo is not None else None)
        return BD, iBD, Fo

    def blockdiag_apply(self, neq, nnodes, Id, invBD, p):
        z = np.zeros(neq + 1)
        self.lib.orc_blockdiag_apply(neq, nnodes, Id, invBD, np.ascontiguousarray(p), z)
        return z

    def symrcm(self, neq, IA, JA):
        P = np.zeros(max(neq, 1), np.int32)
        self.lib.orc_symrcm(neq, IA, JA, P)
        return P[:neq]

    # ---- 3D / CSR(k) / ILU(0) ----
    def mesh3d(self, n):
        nn, nel = (n + 1) ** 3, 6 * n ** 3
        x, y, z = np.zeros(nn), np.zeros(nn), np.zeros(nn)
        t, v = np.zeros(nn, np.int32), np.zeros(4 * nel, np.int32)
        self.lib.orc_mesh3d_structured(n, x, y, z, t, v)
        return x, y, z, t, v

    def csr_init_k(self, neq, nel, k, lm):
        nnz = self.lib.orc_csr_init_k(neq, nel, k, lm, None, None, None)
        IA = np.zeros(neq + 1, np.int32)
        JA = np.zeros(nnz, np.int32)
        cbe = np.zeros(k * k * nel, np.int32)
        self.lib.orc_csr_init_k(neq, nel, k, lm, _ptr(IA), _ptr(JA), _ptr(cbe))
        return nnz, IA, JA, cbe

    def csr_assembly_k(self, nel, k, nnz, ke, cbe):
        AA = np.zeros(nnz + 1)
        self.lib.orc_csr_assembly_k(nel, k, nnz, ke, cbe, AA)
        return AA

    def ilu0_factor(self, neq, AA, JA, IA):
        LU = np.zeros(len(JA))
        D = np.zeros(max(neq, 1))
        rc = self.lib.orc_ilu0_factor(neq, np.ascontiguousarray(AA[:len(JA)]), JA, IA, LU, D)
        return rc, LU, D

    def ilu0_apply(self, neq, LU, JA, IA, D, p):
        z = np.zeros(neq + 1)
        self.lib.orc_ilu0_apply(neq, LU, JA, IA, D, np.ascontiguousarray(p), z)
        return z

    def ilu0_apply_f32(self, neq, LU, JA, IA, D, p):
        z = np.zeros(neq + 1, np.float32)
        self.lib.orc_ilu0_apply_f32(neq, LU.astype(np.float32), JA, IA, D.astype(np.float32),
                                    np.ascontiguousarray(p, np.float32), z)
        return z

    def build_supg3d(self, x, y, z, utype, vertex, lm, neq, pb):
        nel = len(vertex) // 4
        ke = np.zeros(16 * nel)
        F = np.zeros(neq + 1)
        self.lib.orc_build_supg3d(nel, neq, x, y, z, utype, vertex, lm, C.byref(pb), ke, F)
        return ke, F

    # ---- BLAS / matvec ----
    def ddot(self, x, y):
        return self.lib.orc_ddot(len(x), x, y)

    def csrmv(self, neq, AA, JA, IA, P):
        Q = np.zeros(neq + 1)
        self.lib.orc_csrmv(neq, AA, JA, IA, P, Q)
        return Q

    def ebemv(self, neq, nel, k, A, lm, P):
        Q = np.zeros(neq + 1)
        self.lib.orc_ebemv(neq, nel, k, A, lm, P, Q)
        return Q

    def edemv(self, neq, nedge, A, lm2, P):
        Q = np.zeros(neq + 1)
        self.lib.orc_edemv(neq, nedge, A, lm2, P, Q)
        return Q

    # ---- preconditioner setups: return (Diag, invDiag, F_scaled) ----
    def diag_setup_ebe(self, neq, nel, k, A, lm, F):
        D, iD, F2 = np.zeros(neq + 1), np.zeros(neq + 1), F.copy()
        self.lib.orc_diag_setup_ebe(neq, nel, k, A, lm, D, iD, F2)
        return D, iD, F2

    def diag_setup_ede(self, neq, nedge, A, lm2, F):
        D, iD, F2 = np.zeros(neq + 1), np.zeros(neq + 1), F.copy()
        self.lib.orc_diag_setup_ede(neq, nedge, A, lm2, D, iD, F2)
        return D, iD, F2

    def diag_setup_csr(self, neq, AA, JA, IA, F):
        D, iD, F2 = np.zeros(neq + 1), np.zeros(neq + 1), F.copy()
        self.lib.orc_diag_setup_csr(neq, AA, JA, IA, D, iD, F2)
        return D, iD, F2

    def jacobi_ebe_setup(self, neq, nel, A, lm, F):
        invDe, F2 = np.zeros(3 * nel), F.copy()
        self.lib.orc_jacobi_ebe_setup(neq, nel, A, lm, invDe, F2)
        return invDe, F2

    def jacobi_ebe_apply(self, neq, nel, invDe, lm, p):
        z = np.zeros(neq + 1)
        self.lib.orc_jacobi_ebe_apply(neq, nel, invDe, lm, p, z)
        return z

    def jacobi_ebe_setup_k(self, neq, nel, k, A, lm, F):
        """JacobiDOF4_precond_EBE_setup (k = 12) or any k: (invDe, F preconditioned)"""
        invDe, F2 = np.zeros(k * nel), np.ascontiguousarray(F, np.float64).copy()
        self.lib.orc_jacobi_ebe_setup_k(neq, nel, k, np.ascontiguousarray(A, np.float64),
                                        np.ascontiguousarray(lm, np.int32), invDe, F2)
        return invDe, F2

    def jacobi_ebe_apply_k(self, neq, nel, k, invDe, lm, p):
        z = np.zeros(neq + 1)
        self.lib.orc_jacobi_ebe_apply_k(neq, nel, k, invDe, np.ascontiguousarray(lm, np.int32),
                                        np.ascontiguousarray(p, np.float64), z)
        return z

    # ---- GMRES over a stored operator ----
    def pgmres(self, neq, kmax, lmax, tol, B, scheme, mats, invDiag=None, hist_cap=100000,
               ilu=None):
        """scheme: 'CSR' (mats=(AA,JA,IA)), 'EBE' (mats=(A,lm,k,nel)),
        'EDE' (mats=(A,lm2,nedge)).  invDiag: Diag left preconditioner or None;
        ilu: (LU, JA, IA, D) ILU(0) left preconditioner or None.
        Returns (X, steps, cycles, rho_hist, e0_hist)."""
        L = self.lib
        keep = []
        if scheme == "CSR":
            AA, JA, IA = mats
            ctx = OrcCsrCtx(neq, _ptr(AA), _ptr(JA), _ptr(IA)); fn = L.orc_op_csr
        elif scheme == "EBE":
            A, lm, k, nel = mats
            ctx = OrcEbeCtx(neq, nel, k, _ptr(A), _ptr(lm)); fn = L.orc_op_ebe
        else:
            A, lm2, nedge = mats
            ctx = OrcEdeCtx(neq, nedge, _ptr(A), _ptr(lm2)); fn = L.orc_op_ede
        keep.append(ctx)
        ops = OrcOps()
        ops.matvec = C.cast(fn, C.c_void_p)
        ops.matvec_ctx = C.cast(C.pointer(ctx), C.c_void_p)
        if invDiag is not None:
            dctx = OrcDiagCtx(neq, _ptr(invDiag)); keep.append(dctx)
            ops.precond = C.cast(L.orc_op_diag, C.c_void_p)
            ops.precond_ctx = C.cast(C.pointer(dctx), C.c_void_p)
        if ilu is not None:
            LU, JAi, IAi, Di = ilu
            ictx = OrcIluCtx(neq, _ptr(LU), _ptr(JAi), _ptr(IAi), _ptr(Di)); keep.append(ictx)
            ops.precond = C.cast(L.orc_op_ilu0, C.c_void_p)
            ops.precond_ctx = C.cast(C.pointer(ictx), C.c_void_p)
        X = np.zeros(neq + 1)
        rho = np.zeros(hist_cap)
        e0 = np.zeros(lmax + 1)
        ncyc = C.c_int(0)
        steps = L.orc_pgmres(neq, kmax, lmax, tol, np.ascontiguousarray(B), X, C.byref(ops),
                             _ptr(rho), hist_cap, _ptr(e0), C.byref(ncyc))
        return X, steps, ncyc.value, rho[:min(steps, hist_cap)].copy(), e0[:ncyc.value].copy()

    # ---- mesh / structures / assembly ----
    def mesh2d(self, n):
        nn, nel = (n + 1) ** 2, 2 * n * n
        x, y = np.zeros(nn), np.zeros(nn)
        t, v = np.zeros(nn, np.int32), np.zeros(3 * nel, np.int32)
        self.lib.orc_mesh2d_structured(n, x, y, t, v)
        return x, y, t, v

    def number_nodes(self, type_):
        idv = np.zeros(len(type_), np.int32)
        neq = self.lib.orc_number_nodes(len(type_), type_, idv)
        return neq, idv

    def fill_lm(self, neq, nel, vertex, type_, idv, nnoel=3):
        lm = np.zeros(nnoel * nel, np.int32)
        self.lib.orc_fill_lm(neq, nel, nnoel, vertex, type_, idv, lm)
        return lm

    def csr_init(self, neq, nel, lm):
        nnz = self.lib.orc_csr_init(neq, nel, lm, None, None, None)
        IA, JA = np.zeros(neq + 1, np.int32), np.zeros(nnz, np.int32)
        cbe = np.zeros(9 * nel, np.int32)
        self.lib.orc_csr_init(neq, nel, lm, _ptr(IA), _ptr(JA), _ptr(cbe))
        return nnz, IA, JA, cbe

    def ede_init(self, neq, nel, lm):
        nedge = self.lib.orc_ede_init(neq, nel, lm, None, None, None)
        order, ebe = np.zeros(3 * nel, np.int32), np.zeros(3 * nel, np.int32)
        lm2 = np.zeros(2 * nedge, np.int32)
        self.lib.orc_ede_init(neq, nel, lm, _ptr(order), _ptr(ebe), _ptr(lm2))
        return nedge, order, ebe, lm2

    def build_supg2d(self, x, y, type_, idv, vertex, neq, u, prob: OrcProb2d):
        nel = len(vertex) // 3
        ke, F = np.zeros(9 * nel), np.zeros(neq + 1)
        self.lib.orc_build_supg2d(len(x), nel, neq, x, y, type_, idv, vertex, u,
                                  C.byref(prob), ke, F)
        return ke, F

    def build_supg2d_transient(self, x, y, vertex, lm, neq, U, dU, prob: OrcProb2d, alpha, dt):
        """Build_M_K_R_SUPG: (A[nel*9] = me + alpha dt ke, R[neq+1])"""
        nel = len(vertex) // 3
        A, R = np.zeros(9 * nel), np.zeros(neq + 1)
        L = self.lib
        L.orc_build_supg2d_transient.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _dp, _dp,
                                                 C.POINTER(OrcProb2d), _d, _d, _dp, _dp]
        L.orc_build_supg2d_transient(len(x), nel, neq, x, y, vertex, lm,
                                     np.ascontiguousarray(U, np.float64),
                                     np.ascontiguousarray(dU, np.float64), C.byref(prob),
                                     alpha, dt, A, R)
        return A, R

    def build_euler2d(self, x, y, vertex, lm, neq, U, dU, rot, theta, gamma, mach, dimless,
                      alpha, dt, tol):
        """Build_M_K_F_SUPG (CAU): (A[nel*144] = Me + alpha dt Ke, R[neq+1])"""
        nel = len(vertex) // 3
        A, R = np.zeros(144 * nel), np.zeros(neq + 1)
        L = self.lib
        L.orc_build_euler2d.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _dp, _dp, _vp, _vp, _d,
                                        _d, _i, _d, _d, _d, _dp, _dp]
        rot_a = None if rot is None else np.ascontiguousarray(rot, np.int32)
        th_a = None if theta is None else np.ascontiguousarray(theta, np.float64)
        L.orc_build_euler2d(len(x), nel, neq, x, y, vertex, lm,
                            np.ascontiguousarray(U, np.float64),
                            np.ascontiguousarray(dU, np.float64),
                            _ptr(rot_a) if rot_a is not None else None,
                            _ptr(th_a) if th_a is not None else None, gamma, mach, int(dimless),
                            alpha, dt, tol, A, R)
        return A, R

    def predictor_old_tr(self, x, y, type_, idv, vertex, lm, neq, prob: OrcProb2d, ud, u0,
                         alpha, dt, final_time, nl_tol, nl_maxiter, stop_by_norm, steady,
                         time_tol, diag, cap=0, hist_cap=4096):
        """Predictor_Old over the transient build: (u, corrections per step,
        steps, t_final, status)"""
        L = self.lib
        L.orc_predictor_old_tr.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _ip,
                                           C.POINTER(OrcProb2d), _d, C.POINTER(OrcPredParams),
                                           _i, _dp, _dp, _ip, _i, C.POINTER(_i), C.POINTER(_d)]
        nel = len(vertex) // 3
        pp = OrcPredParams(alpha, dt, final_time, nl_tol, time_tol, nl_maxiter,
                           int(stop_by_norm), int(steady), cap)
        u = np.zeros(neq + 1)
        hist = np.zeros(hist_cap, np.int32)
        ns, tf = _i(0), _d(0.0)
        rc = L.orc_predictor_old_tr(len(x), nel, neq, x, y, type_, idv, vertex, lm,
                                    C.byref(prob), ud, C.byref(pp), int(diag),
                                    np.ascontiguousarray(u0, np.float64), u, hist, hist_cap,
                                    C.byref(ns), C.byref(tf))
        return u, hist[:min(ns.value, hist_cap)].copy(), ns.value, tf.value, rc

    def predictor_old_eu(self, x, y, vertex, lm, neq, ids, presc, rot, theta, gamma, mach,
                         dimless, stab_tol, u0, alpha, dt, final_time, nl_tol, nl_maxiter,
                         stop_by_norm, steady, time_tol, block_scaling, cap=0, hist_cap=4096,
                         scheme=0):
        """Predictor_Old over the Euler build: (u, corrections per step, steps,
        t_final, status); scheme 0 EBE, 1 EBE2, 2 EDE, 3 EDE2"""
        L = self.lib
        L.orc_predictor_old_eu_scheme.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _dp, _vp,
                                                  _vp, _d, _d, _i, _d, C.POINTER(OrcPredParams),
                                                  _i, _i, _dp, _dp, _ip, _i, C.POINTER(_i),
                                                  C.POINTER(_d)]
        nel = len(vertex) // 3
        pp = OrcPredParams(alpha, dt, final_time, nl_tol, time_tol, nl_maxiter,
                           int(stop_by_norm), int(steady), cap)
        rot_a = None if rot is None else np.ascontiguousarray(rot, np.int32)
        th_a = None if theta is None else np.ascontiguousarray(theta, np.float64)
        u = np.zeros(neq + 1)
        hist = np.zeros(hist_cap, np.int32)
        ns, tf = _i(0), _d(0.0)
        rc = L.orc_predictor_old_eu_scheme(len(x), nel, neq, x, y, vertex, lm,
                                           np.ascontiguousarray(ids, np.int32),
                                           np.ascontiguousarray(presc, np.float64),
                                           _ptr(rot_a) if rot_a is not None else None,
                                           _ptr(th_a) if th_a is not None else None, gamma, mach,
                                           int(dimless), stab_tol, C.byref(pp),
                                           int(block_scaling), int(scheme),
                                           np.ascontiguousarray(u0, np.float64), u, hist,
                                           hist_cap, C.byref(ns), C.byref(tf))
        return u, hist[:min(ns.value, hist_cap)].copy(), ns.value, tf.value, rc

    def ede_assembly(self, nel, nedge, ke, order, ebe):
        A = np.zeros(4 * (nedge + 1))
        self.lib.orc_ede_assembly(nel, nedge, ke, order, ebe, A)
        return A

    def csr_assembly(self, nel, nnz, ke, cbe):
        AA = np.zeros(nnz + 1)
        self.lib.orc_csr_assembly(nel, nnz, ke, cbe, AA)
        return AA


class OrcPredParams(C.Structure):
    """orc_pred_params: Predictor_Old / Criteria.c parameters"""
    _fields_ = [("alpha", C.c_double), ("dt", C.c_double), ("final_time", C.c_double),
                ("nl_tol", C.c_double), ("time_tol", C.c_double), ("nl_maxiter", C.c_int),
                ("stop_by_norm", C.c_int), ("steady", C.c_int), ("cap", C.c_int)]


def prob2d(kind=P_CONST, kappa=1e-4, gamma=0.0, f=1.0, bx=1.0, by=0.0, scale=1.0,
           h_shock=H_2SQRTAREA) -> OrcProb2d:
    return OrcProb2d(kind, kappa, gamma, f, bx, by, scale, h_shock)


class Reference:
    """The reference SS-2D routines compiled from their sources (oracle/_ref).

    Available only where oracle/_ref/libref_ss2d.so was built."""

    PATH = os.path.join(HERE, "_ref", "libref_ss2d.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self, fast: bool = False):
        path = self.PATH.replace(".so", "_fast.so") if fast else self.PATH
        L = self.lib = C.CDLL(path)
        L.ref_setup.restype = _i
        L.ref_setup.argtypes = [_i, _i, _dp, _dp, _ip, _ip, C.c_char_p, C.c_char_p]
        L.ref_sizes.argtypes = [C.POINTER(C.c_int)] * 3
        L.ref_get_lm.argtypes = [_ip]
        L.ref_get_ids.argtypes = [_ip]
        L.ref_get_csr.argtypes = [_ip, _ip, _ip]
        L.ref_get_ede.argtypes = [_ip, _ip, _ip]
        L.ref_set_problem.argtypes = [_i, _d, _d, _d, _d, _d, _d, _i]
        L.ref_build.argtypes = [_dp]
        L.ref_get_F.argtypes = [_dp]
        L.ref_set_F.argtypes = [_dp]
        L.ref_get_matrix.argtypes = [_dp]
        L.ref_set_matrix.argtypes = [_dp]
        L.ref_matvec.argtypes = [_dp, _dp]
        L.ref_precond_setup.argtypes = [C.c_char_p, _dp]
        L.ref_get_invdiag.argtypes = [_dp, _dp]
        L.ref_get_invde.argtypes = [_dp]
        L.ref_precond_apply.argtypes = [_dp, _dp]
        L.ref_pgmres.restype = _i
        L.ref_pgmres.argtypes = [_i, _i, _d, _dp, _dp]
        L.ref_ddot.restype = _d
        L.ref_ddot.argtypes = [_i, _dp, _dp]
        self.nel = 0

    def setup(self, x, y, type_, vertex, scheme="EBE", reorder="NOT"):
        self.nel = len(vertex) // 3
        neq = self.lib.ref_setup(len(x), self.nel, x, y, type_, vertex,
                                 scheme.encode(), reorder.encode())
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.lib.ref_sizes(C.byref(a), C.byref(b), C.byref(c))
        self.neq, self.nnz, self.nedge = a.value, b.value, c.value
        self.scheme = scheme
        return neq

    def lm(self):
        out = np.zeros(3 * self.nel, np.int32); self.lib.ref_get_lm(out); return out

    def ids(self, nnodes):
        out = np.zeros(nnodes, np.int32); self.lib.ref_get_ids(out); return out

    def csr(self):
        IA, JA = np.zeros(self.neq + 1, np.int32), np.zeros(self.nnz, np.int32)
        cbe = np.zeros(9 * self.nel, np.int32)
        self.lib.ref_get_csr(IA, JA, cbe)
        return IA, JA, cbe

    def ede(self):
        o, ebe = np.zeros(3 * self.nel, np.int32), np.zeros(3 * self.nel, np.int32)
        lm2 = np.zeros(2 * self.nedge, np.int32)
        self.lib.ref_get_ede(o, ebe, lm2)
        return o, ebe, lm2

    def set_problem(self, p: OrcProb2d):
        self.lib.ref_set_problem(p.kind, p.kappa, p.gamma, p.f, p.bx, p.by, p.scale, p.h_shock)

    def matrix_size(self):
        return {"EBE": 9 * self.nel, "EDE": 4 * (self.nedge + 1), "CSR": self.nnz + 1}[self.scheme]

    def build(self, u):
        self.lib.ref_build(np.ascontiguousarray(u, np.float64))
        A, F = np.zeros(self.matrix_size()), np.zeros(self.neq + 1)
        self.lib.ref_get_matrix(A); self.lib.ref_get_F(F)
        return A, F

    def set_matrix(self, A):
        self.lib.ref_set_matrix(np.ascontiguousarray(A, np.float64))

    def matvec(self, P):
        P = P.copy(); Q = np.zeros(self.neq + 1)
        self.lib.ref_matvec(P, Q)
        return Q

    def precond_setup(self, name, F):
        F2 = F.copy(); self.lib.ref_precond_setup(name.encode(), F2)
        return F2

    def invdiag(self):
        D, iD = np.zeros(self.neq + 1), np.zeros(self.neq + 1)
        self.lib.ref_get_invdiag(D, iD)
        return D, iD

    def invde(self):
        out = np.zeros(3 * self.nel); self.lib.ref_get_invde(out); return out

    def scaling(self, name, A, F):
        """scaling.c Left / LeftRight on A, F: (A, F, Diag, invDiag) after"""
        self.set_matrix(A)
        self.lib.ref_set_F(np.ascontiguousarray(F, np.float64))
        self.lib.ref_scaling.argtypes = [C.c_char_p]
        self.lib.ref_scaling(name.encode())
        A2, F2 = np.zeros(self.matrix_size()), np.zeros(self.neq + 1)
        self.lib.ref_get_matrix(A2); self.lib.ref_get_F(F2)
        D, iD = self.invdiag()
        return A2, F2, D, iD

    def unscaling(self, x):
        x = np.ascontiguousarray(x, np.float64).copy()
        self.lib.ref_unscaling.argtypes = [_dp]
        self.lib.ref_unscaling(x)
        return x

    def precond_apply(self, p):
        z = np.zeros(self.neq + 1); self.lib.ref_precond_apply(p.copy(), z); return z

    def pgmres(self, kmax, lmax, tol, B):
        X = np.zeros(self.neq + 1)
        it = self.lib.ref_pgmres(kmax, lmax, tol, np.ascontiguousarray(B), X)
        return X, it

    def ddot(self, x, y):
        return self.lib.ref_ddot(len(x), x.copy(), y.copy())


class Reference3D:
    """The reference 3D steady-transport routines (Build_SUPG with the DIFCONV
    problem, CSR ILU(0)) compiled from their sources (oracle/_ref).

    Available only where oracle/_ref/libref_ss3d.so was built."""

    PATH = os.path.join(HERE, "_ref", "libref_ss3d.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self, fast: bool = False):
        path = self.PATH.replace(".so", "_fast.so") if fast else self.PATH
        L = self.lib = C.CDLL(path)
        L.ref3d_setup.restype = _i
        L.ref3d_setup.argtypes = [_i, _i, _dp, _dp, _dp, _ip, _ip, _d]
        L.ref3d_lm.argtypes = [_ip]
        L.ref3d_build.argtypes = [_dp, _dp]
        L.ref3d_ilu0_setup.argtypes = [_i, _i, _dp, _ip, _ip, _dp, _dp, _dp]
        L.ref3d_ilu0_apply.argtypes = [_dp, _dp]
        L.ref3d_ebemv_setup.argtypes = [_i, _i, _ip, _dp]
        L.ref3d_ebemv.argtypes = [_dp, _dp]

    def setup(self, x, y, z, utype, vertex, cst=1.0):
        self.nel = len(vertex) // 4
        self.neq = self.lib.ref3d_setup(len(x), self.nel, x, y, z, utype, vertex, cst)
        return self.neq

    def lm(self):
        lm = np.zeros(4 * self.nel, np.int32)
        self.lib.ref3d_lm(lm)
        return lm

    def build(self):
        ke = np.zeros(16 * self.nel)
        F = np.zeros(self.neq + 1)
        self.lib.ref3d_build(ke, F)
        return ke, F

    def ilu0_setup(self, neq, AA, JA, IA, F):
        nnz = len(JA)
        LU = np.zeros(nnz)
        D = np.zeros(max(neq, 1))
        Fp = np.ascontiguousarray(F, np.float64).copy()
        self._keep = (np.ascontiguousarray(AA, np.float64).copy(), JA.copy(), IA.copy())
        self.neq_ilu = neq
        self.lib.ref3d_ilu0_setup(neq, nnz, self._keep[0], self._keep[1], self._keep[2], Fp,
                                  LU, D)
        return LU, D, Fp

    def ilu0_apply(self, p):
        z = np.zeros(self.neq_ilu + 1)
        self.lib.ref3d_ilu0_apply(np.ascontiguousarray(p, np.float64).copy(), z)
        return z

    def ebemv_setup(self, neq, lm, A):
        """the reference ebemv3Dndof1 on these element matrices (nel*16)"""
        self.neq_ebe = neq
        self.lib.ref3d_ebemv_setup(neq, len(lm) // 4, np.ascontiguousarray(lm, np.int32),
                                   np.ascontiguousarray(A, np.float64))

    def ebemv(self, P):
        Q = np.zeros(self.neq_ebe + 1)
        self.lib.ref3d_ebemv(np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q


class ReferenceEuler:
    """The reference's 4-dof (Euler 2D) element products and node block-
    diagonal routines compiled from their sources (oracle/_ref).  `path`
    selects another build of the same harness (the drop-in build links the
    GPU adapter instead of the reference's products and setups and leaves
    out the entry points that read host block arrays)."""

    PATH = os.path.join(HERE, "_ref", "libref_eu2d.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self, fast: bool = False, path: str = None):
        if path is None:
            path = self.PATH.replace(".so", "_fast.so") if fast else self.PATH
        L = self.lib = C.CDLL(path)
        for name, args in (("refeu_setup", [_i, _i, _i, _ip, _ip, _ip, _dp, C.c_char_p]),
                           ("refeu_ebemv", [_dp, _dp]), ("refeu_ebe2mv", [_dp, _dp]),
                           ("refeu_blockdiag_setup", [_dp, _dp, _dp, _ip]),
                           ("refeu_blockdiag_apply", [_dp, _dp]),
                           ("refeu_block_scaling", [_dp, _dp]),
                           ("refeu_jacobi4_setup", [_dp, _dp]),
                           ("refeu_jacobi4_apply", [_dp, _dp])):
            if hasattr(L, name):
                getattr(L, name).argtypes = args

    def setup(self, nnodes, nel, neq, vertex, ids, lm, A, precond="BlockDiag"):
        self.nnodes, self.nel, self.neq = nnodes, nel, neq
        self.lib.refeu_setup(nnodes, nel, neq, vertex, ids, lm, np.ascontiguousarray(A),
                             precond.encode())

    def ebemv(self, P):
        Q = np.zeros(self.neq + 1)
        self.lib.refeu_ebemv(np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q

    def ebe2mv(self, P):
        Q = np.zeros(self.neq + 1)
        self.lib.refeu_ebe2mv(np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q

    def jacobi4_setup(self, F):
        """JacobiDOF4_precond_EBE_setup: (invDe [nel*12], F preconditioned)"""
        Fp = np.ascontiguousarray(F, np.float64).copy()
        invDe = np.zeros(12 * self.nel)
        self.lib.refeu_jacobi4_setup(Fp, invDe)
        return invDe, Fp

    def jacobi4_apply(self, p):
        """JacobiDOF4_precond_EBE out of place, after jacobi4_setup"""
        z = np.zeros(self.neq + 1)
        self.lib.refeu_jacobi4_apply(np.ascontiguousarray(p, np.float64).copy(), z)
        return z

    def blockdiag_setup(self, F):
        Fp = np.ascontiguousarray(F, np.float64).copy()
        BD, iBD = np.zeros(16 * self.nnodes), np.zeros(16 * self.nnodes)
        Id = np.zeros(4 * self.nnodes, np.int32)
        self.lib.refeu_blockdiag_setup(Fp, BD, iBD, Id)
        return Fp, BD, iBD, Id

    def blockdiag_apply(self, p):
        z = np.zeros(self.neq + 1)
        self.lib.refeu_blockdiag_apply(np.ascontiguousarray(p, np.float64).copy(), z)
        return z

    # ---- the edge-by-edge scheme (after setup) ----
    def ede_setup(self):
        """ede_Initialization: (order, edge_by_elem, lm2)"""
        L = self.lib
        L.refeu_ede_setup.restype = _i
        ne = self.nedge = L.refeu_ede_setup()
        o, e = np.zeros(3 * self.nel, np.int32), np.zeros(3 * self.nel, np.int32)
        lm2 = np.zeros(8 * ne, np.int32)
        L.refeu_ede_get.argtypes = [_ip, _ip, _ip]
        L.refeu_ede_get(o, e, lm2)
        return o, e, lm2

    def ede_assemble(self, A12):
        """setzeros + ede_assembly per element: the edge matrices [nedge*64]"""
        L = self.lib
        L.refeu_ede_assemble.argtypes = [_dp, _dp]
        Aede = np.zeros(64 * self.nedge)
        L.refeu_ede_assemble(np.ascontiguousarray(A12, np.float64), Aede)
        return Aede

    def ede_set_matrix(self, Aede):
        self.lib.refeu_ede_set_matrix.argtypes = [_dp]
        self.lib.refeu_ede_set_matrix(np.ascontiguousarray(Aede, np.float64))

    def edemv(self, P):
        self.lib.refeu_edemv.argtypes = [_dp, _dp]
        Q = np.zeros(self.neq + 1)
        self.lib.refeu_edemv(np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q

    def ede_blockdiag_setup(self, F):
        self.lib.refeu_ede_blockdiag_setup.argtypes = [_dp, _dp, _dp, _ip]
        Fp = np.ascontiguousarray(F, np.float64).copy()
        BD, iBD = np.zeros(16 * self.nnodes), np.zeros(16 * self.nnodes)
        Id = np.zeros(4 * self.nnodes, np.int32)
        self.lib.refeu_ede_blockdiag_setup(Fp, BD, iBD, Id)
        return Fp, BD, iBD, Id

    def ede2mv(self, P):
        self.lib.refeu_ede2mv.argtypes = [_dp, _dp]
        Q = np.zeros(self.neq + 1)
        self.lib.refeu_ede2mv(np.ascontiguousarray(P, np.float64).copy(), Q)
        return Q

    def pred_run(self, u0, alpha, dt, final_time, nl_tol, nl_maxiter, stop_by_norm, steady,
                 time_tol, stab_tol, block_scaling, cap=0, hist_cap=4096, scheme=0):
        """the reference's Predictor_Old over its Euler build (after setup and
        build_setup): (u, corrections per step, steps, t_final, status).
        scheme: 0 EBE, 1 EBE2 (BlockDiag), 2 EDE (NOT), 3 EDE2 (BlockDiag);
        2 and 3 after ede_setup"""
        L = self.lib
        L.refeu_pred_scheme.argtypes = [_i]
        L.refeu_pred_scheme(int(scheme))
        L.refeu_pred_run.argtypes = [_dp, _d, _d, _d, _d, _i, _i, _i, _d, _d, _i, _i, _dp, _ip,
                                     _i, C.POINTER(_i), C.POINTER(_d)]
        u = np.zeros(self.neq + 1)
        hist = np.zeros(hist_cap, np.int32)
        ns, tf = _i(0), _d(0.0)
        rc = L.refeu_pred_run(np.ascontiguousarray(u0, np.float64), alpha, dt, final_time,
                              nl_tol, nl_maxiter, int(stop_by_norm), int(steady), time_tol,
                              stab_tol, int(block_scaling), cap, u, hist, hist_cap, C.byref(ns),
                              C.byref(tf))
        return u, hist[:min(ns.value, hist_cap)].copy(), ns.value, tf.value, rc

    def block_scaling(self, F):
        Fp = np.ascontiguousarray(F, np.float64).copy()
        A = np.zeros(144 * self.nel)
        self.lib.refeu_block_scaling(Fp, A)
        return A, Fp

    def build_setup(self, x, y, v1type, gamma, mach, dimless, alpha, dt, tol, presc, th):
        L = self.lib
        L.refeu_build_setup.argtypes = [_dp, _dp, _ip, _d, _d, _i, _d, _d, _d, _dp, _dp]
        L.refeu_build_setup(np.ascontiguousarray(x, np.float64),
                            np.ascontiguousarray(y, np.float64),
                            np.ascontiguousarray(v1type, np.int32), gamma, mach, int(dimless),
                            alpha, dt, tol, np.ascontiguousarray(presc, np.float64),
                            np.ascontiguousarray(th, np.float64))

    def build(self, u, du):
        L = self.lib
        L.refeu_build.argtypes = [_dp, _dp, _dp, _dp]
        A, R = np.zeros(144 * self.nel), np.zeros(self.neq + 1)
        L.refeu_build(np.ascontiguousarray(u, np.float64), np.ascontiguousarray(du, np.float64),
                      A, R)
        return A, R


class ReferenceTransient:
    """The reference's transient 2D SUPG element build (Build_M_K_R_SUPG.c
    of 02_Transport_Equation) compiled from its sources (oracle/_ref)."""

    PATH = os.path.join(HERE, "_ref", "libref_tr2d.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self):
        L = self.lib = C.CDLL(self.PATH)
        L.reftr_setup.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _ip, _i, _d, _d, _d, _d,
                                  _d, _d, _d, _d, _d]
        L.reftr_build.argtypes = [_dp, _dp, _dp, _dp]

    def setup(self, x, y, type_, idv, vertex, lm, neq, kind, kappa, gamma, f, bx, by, scale,
              ud, alpha, dt):
        self.neq, self.nel = neq, len(vertex) // 3
        self.lib.reftr_setup(len(x), self.nel, neq, x, y, type_, idv, vertex, lm, kind, kappa,
                             gamma, f, bx, by, scale, ud, alpha, dt)

    def build(self, u, du):
        A, R = np.zeros(9 * self.nel), np.zeros(self.neq + 1)
        self.lib.reftr_build(np.ascontiguousarray(u, np.float64),
                             np.ascontiguousarray(du, np.float64), A, R)
        return A, R


class ReferencePredictor:
    """The reference's time integrator (Predictor_Old.c, Criteria.c) over the
    transient build of 02_Transport_Equation, compiled from its sources
    (oracle/_ref), with a direct dense solver as the 6-argument solver."""

    PATH = os.path.join(HERE, "_ref", "libref_pred2d.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self, path: str = None):
        L = self.lib = C.CDLL(path or self.PATH)
        L.refpred_setup.argtypes = [_i, _i, _i, _dp, _dp, _ip, _ip, _ip, _ip, _i, _d, _d, _d,
                                    _d, _d, _d, _d]
        L.refpred_run.argtypes = [_dp, _d, _d, _d, _d, _i, _i, _i, _d, _d, _i, _i, _dp, _ip,
                                  _i, C.POINTER(_i), C.POINTER(_d)]

    def setup(self, x, y, type_, idv, vertex, lm, neq, kind, kappa, gamma, f, bx, by, scale,
              ud):
        self.neq = neq
        self.lib.refpred_setup(len(x), len(vertex) // 3, neq, x, y, type_, idv, vertex, lm,
                               kind, kappa, gamma, f, bx, by, scale, ud)

    def run(self, u0, alpha, dt, final_time, nl_tol, nl_maxiter, stop_by_norm, steady,
            time_tol, diag, cap=0, hist_cap=4096):
        u = np.zeros(self.neq + 1)
        hist = np.zeros(hist_cap, np.int32)
        ns, tf = _i(0), _d(0.0)
        rc = self.lib.refpred_run(np.ascontiguousarray(u0, np.float64), alpha, dt, final_time,
                                  nl_tol, nl_maxiter, int(stop_by_norm), int(steady), time_tol,
                                  0.0, int(diag), cap, u, hist, hist_cap, C.byref(ns),
                                  C.byref(tf))
        return u, hist[:min(ns.value, hist_cap)].copy(), ns.value, tf.value, rc
