"""fem_codes_amd -- MI355X-native solve path of FEM_CODES.

The product is ``libfemgpu.so`` (HIP kernels for gfx950 behind the C ABI in
``include/femgpu.h``).  This module is the Python host layer over that ABI:
ctypes bindings, device arrays and the operator / preconditioner / GMRES
objects that mirror the reference's plugin table

    FemFunctions->ProductMatrixVector = ebemv | edemv | csrmv
    FemFunctions->precond_setup / precond = Diag_precond{_EBE,_EDE,_CSR}_setup / Diag_precond
    FemOtherFunctions->solver = pgmres

(01_SS_Transport_Equation/2D/01_CommonFiles/setMatrixVectorProductType.c:6-17,
setPreconditioner.c:36-55, setSolver.c:8).  There is no CPU fallback: every
call goes to the GPU library, and a missing library raises at import.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemgpu.so")

FG_OK, FG_ERR_HIP, FG_ERR_ARG, FG_ERR_BREAKDOWN, FG_ERR_NOCONV, FG_ERR_NOMEM = range(6)
PREC_NONE, PREC_DIAG, PREC_JACOBI, PREC_ILU0, PREC_BLOCKDIAG = 0, 1, 2, 3, 4
ORDER_BFS, ORDER_MESH = 0, 1
ROWS_VALU, ROWS_MFMA = 0, 1
EDGE_ROWS, EDGE_BLOCKED, EDGE_AUTO = 0, 1, 2
PRODUCT_SLICED, PRODUCT_BLOCKED, PRODUCT_AUTO = 0, 1, 2
PRODUCT_FP64, PRODUCT_FP32 = 1, 2
EBE2_BLOCKDIAG, EBE2_NOBLOCKDIAG = 0, 1
P3_DIFCONV, P3_CONST = 0, 1
GMRES_FP64, GMRES_FP32_IR = 0, 1


def build(verbose: bool = False) -> None:
    """Compile libfemgpu.so for gfx950 (hipcc cross-compiles without a GPU)."""
    out = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), "-j8"],
                         capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("libfemgpu build failed:\n" + out.stdout + out.stderr)
    if verbose:
        print(out.stdout, out.stderr)


class FemGpuError(RuntimeError):
    pass


class GmresOpts(C.Structure):
    _fields_ = [("kmax", C.c_int), ("lmax", C.c_int), ("tol", C.c_double),
                ("precision", C.c_int), ("inner_tol", C.c_double), ("hist_cap", C.c_int),
                ("rho_hist", C.c_void_p), ("e0_hist", C.c_void_p), ("engine", C.c_int),
                ("orth", C.c_int)]


ENGINE_AUTO, ENGINE_KERNELS, ENGINE_FUSED, ENGINE_SMALL, ENGINE_GRAPH = 0, 1, 2, 3, 4
ORTH_MGS, ORTH_CGS = 0, 1


class GmresInfo(C.Structure):
    _fields_ = [("steps", C.c_int), ("cycles", C.c_int), ("rho", C.c_double),
                ("eps", C.c_double), ("true_res", C.c_double), ("status", C.c_int),
                ("ms", C.c_double)]


class PredictorParams(C.Structure):
    """fg_predictor_params: Predictor_Old / Criteria.c parameters"""
    _fields_ = [("alpha", C.c_double), ("dt", C.c_double), ("final_time", C.c_double),
                ("nl_tol", C.c_double), ("nl_maxiter", C.c_int), ("stop_by_norm", C.c_int),
                ("steady", C.c_int), ("time_tol", C.c_double), ("max_steps", C.c_int)]


class PredictorInfo(C.Structure):
    """fg_predictor_info"""
    _fields_ = [("steps", C.c_int), ("corrections", C.c_int), ("krylov_steps", C.c_longlong),
                ("unconverged", C.c_int), ("t", C.c_double), ("ms", C.c_double),
                ("status", C.c_int)]


_vp, _i, _d = C.c_void_p, C.c_int, C.c_double
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

# name -> (restype, argtypes); every symbol declared in include/femgpu.h
SIGNATURES = {
    "fg_last_error": (C.c_char_p, []),
    "fg_version": (C.c_char_p, []),
    "fg_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "fg_ctx_destroy": (_i, [_vp]),
    "fg_ctx_sync": (_i, [_vp]),
    "fg_ctx_trim": (_i, [_vp]),
    "fg_malloc": (_i, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "fg_free": (_i, [_vp, _vp]),
    "fg_memcpy_h2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "fg_memcpy_d2h": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "fg_memcpy_d2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "fg_memset": (_i, [_vp, _vp, _i, C.c_size_t]),
    "fg_cast_f64_to_f32": (_i, [_vp, _i, _vp, _vp]),
    "fg_cast_f32_to_f64": (_i, [_vp, _i, _vp, _vp]),
    "fg_op_ebe_create": (_i, [_vp, _i, _i, _i, _ip, _i, C.POINTER(_vp)]),
    "fg_op_ede_create": (_i, [_vp, _i, _i, _ip, _i, C.POINTER(_vp)]),
    "fg_op_ebe_create_ex": (_i, [_vp, _i, _i, _i, _ip, _i, _i, _i, C.POINTER(_vp)]),
    "fg_op_csr_create": (_i, [_vp, _i, _i, _ip, _ip, C.POINTER(_vp)]),
    "fg_op_destroy": (_i, [_vp]),
    "fg_op_set_values_host": (_i, [_vp, _dp]),
    "fg_op_set_values_dev": (_i, [_vp, _vp]),
    "fg_op_apply_f64": (_i, [_vp, _vp, _vp]),
    "fg_op_apply_f32": (_i, [_vp, _vp, _vp]),
    "fg_op_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                        C.POINTER(C.c_longlong)]),
    "fg_op_algorithmic_bytes": (_d, [_vp, _i]),
    "fg_op_time_apply": (_i, [_vp, _i, _vp, _vp, _i, _i, C.POINTER(_d)]),
    "fg_op_set_pipeline": (_i, [_vp, _i]),
    "fg_op_set_row_kernel": (_i, [_vp, _i]),
    "fg_op_set_edge_kernel": (_i, [_vp, _i]),
    "fg_op_set_product_kernel": (_i, [_vp, _i, _i]),
    "fg_op_product_kernel": (_i, [_vp, _i, C.POINTER(_i)]),
    "fg_prec_create": (_i, [_vp, _i, C.POINTER(_vp)]),
    "fg_prec_setup": (_i, [_vp, _vp]),
    "fg_prec_apply_f64": (_i, [_vp, _vp, _vp]),
    "fg_prec_apply_f32": (_i, [_vp, _vp, _vp]),
    "fg_prec_get_diag": (_i, [_vp, _vp, _vp]),
    "fg_prec_set_invdiag": (_i, [_vp, _dp]),
    "fg_prec_destroy": (_i, [_vp]),
    "fg_gmres": (_i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(GmresOpts), C.POINTER(GmresInfo)]),
    "fg_gmres_small_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(GmresOpts),
                                    C.POINTER(GmresInfo)]),
    "fg_gmres_batched": (_i, [_vp, _vp, _vp, _i, _ip, _vp, _vp, C.POINTER(GmresOpts), _i,
                              _vp]),
    "fg_number_nodes": (_i, [_i, _ip, _ip]),
    "fg_fill_lm": (_i, [_i, _i, _i, _ip, _ip, _ip, _ip]),
    "fg_csr_pattern": (_i, [_i, _i, _ip, _vp, _vp, _vp, C.POINTER(C.c_longlong)]),
    "fg_ede_pattern": (_i, [_i, _i, _ip, _vp, _vp, _vp, C.POINTER(_i)]),
    "fg_mesh2d_create": (_i, [_vp, _i, _i, _dp, _dp, _ip, _ip, C.POINTER(_vp)]),
    "fg_mesh2d_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fg_mesh2d_get_lm": (_i, [_vp, _ip]),
    "fg_mesh2d_destroy": (_i, [_vp]),
    "fg_supg2d_build": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fg_supg2d_transient_build": (_i, [_vp, _vp, _d, _d, _vp, _vp, _vp, _vp]),
    "fg_predictor_old_transport": (_i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(PredictorParams),
                                        C.POINTER(GmresOpts), _vp, _vp, _i,
                                        C.POINTER(PredictorInfo)]),
    "fg_param_file_read": (_i, [C.c_char_p, C.c_char_p, C.c_char_p, _i, C.POINTER(_i)]),
    "fg_mesh_file_read": (_i, [C.c_char_p, _i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i), _vp,
                               _vp, _vp, _vp]),
    "fg_scaling_diag": (_i, [_vp, _i, _vp]),
    "fg_unscaling_left": (_i, [_vp, _vp]),
    "fg_ede4_create": (_i, [_vp, _i, _i, _i, _ip, _ip, C.POINTER(_vp)]),
    "fg_ede4_info": (_i, [_vp, C.POINTER(_i)]),
    "fg_ede4_get": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fg_ede4_assemble": (_i, [_vp, _vp, _vp]),
    "fg_ede4_destroy": (_i, [_vp]),
    "fg_predictor_old_euler": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                    C.POINTER(PredictorParams), C.POINTER(GmresOpts), _vp, _vp,
                                    _i, C.POINTER(PredictorInfo)]),
    "fg_predictor_old_euler_ede": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                        C.POINTER(PredictorParams), C.POINTER(GmresOpts), _vp,
                                        _vp, _i, C.POINTER(PredictorInfo)]),
    "fg_op_assemble_csr": (_i, [_vp, _i, _ip, _vp]),
    "fg_op_assemble_ede": (_i, [_vp, _i, _ip, _ip, _vp]),
    "fg_csr_pattern_k": (_i, [_i, _i, _i, _ip, _vp, _vp, _vp, C.POINTER(C.c_longlong)]),
    "fg_op_assemble_csr_k": (_i, [_vp, _i, _i, _ip, _vp]),
    "fg_mesh3d_create": (_i, [_vp, _i, _i, _dp, _dp, _dp, _ip, _ip, C.POINTER(_vp)]),
    "fg_mesh3d_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fg_mesh3d_get_lm": (_i, [_vp, _ip]),
    "fg_mesh3d_destroy": (_i, [_vp]),
    "fg_supg3d_build": (_i, [_vp, _vp, _vp, _vp]),
    "fg_prec_ilu_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fg_prec_ilu_get": (_i, [_vp, _vp, _vp]),
    "fg_prec_ilu_schedule": (_i, [_vp, _i]),
    "fg_prec_ilu_sweeps": (_i, [_vp, _i]),
    "fg_rcm": (_i, [_i, _ip, _ip, _ip]),
    "fg_csr_bandwidth": (_i, [_i, _ip, _ip, C.POINTER(_i)]),
    "fg_prec_blockdiag_create": (_i, [_vp, _i, _ip, _ip, C.POINTER(_vp)]),
    "fg_prec_blockdiag_get": (_i, [_vp, _vp, _vp]),
    "fg_block_scaling_ebe": (_i, [_vp, _vp]),
    "fg_op_ebe2_create": (_i, [_vp, _vp, _i, C.POINTER(_vp)]),
    "fg_op_get_values_host": (_i, [_vp, _dp]),
    "fg_euler2d_create": (_i, [_vp, _i, _i, _i, _dp, _dp, _ip, _ip, C.POINTER(_vp)]),
    "fg_euler2d_destroy": (_i, [_vp]),
    "fg_euler2d_build": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}


class EulerParams(C.Structure):
    """fg_euler2d_params"""
    _fields_ = [("gamma", C.c_double), ("mach", C.c_double), ("dimensionless", C.c_int),
                ("alpha", C.c_double), ("dt", C.c_double), ("tolerance", C.c_double)]


class Supg2dProblem(C.Structure):
    """fg_supg2d_problem: the FemFunctions callbacks of one SS-2D problem."""
    _fields_ = [("kind", C.c_int), ("kappa", C.c_double), ("gamma", C.c_double),
                ("f", C.c_double), ("bx", C.c_double), ("by", C.c_double),
                ("scale", C.c_double), ("h_shock", C.c_int)]


P_CONST, P_ROTATING, P_VALIDACAO, P_ROTATE_ABOUT = 0, 1, 2, 3
H_2SQRTAREA, H_SQRTAREA = 0, 1


def _load():
    if not os.path.exists(LIB_PATH):
        raise FemGpuError(f"{LIB_PATH} is missing: run fem_codes_amd.build() "
                          "(or __graft_entry__.build()) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc: int, what: str, ok=(FG_OK,)) -> int:
    if rc not in ok:
        raise FemGpuError(f"{what}: status {rc}: {lib.fg_last_error().decode()}")
    return rc


# ---------------------------------------------------------------------------
class Context:
    """A device, its stream and reduction scratch (fg_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib.fg_ctx_create(device, C.byref(h)), "fg_ctx_create")
        self.h = h

    def sync(self):
        _check(lib.fg_ctx_sync(self.h), "fg_ctx_sync")

    def trim(self):
        """free the solver workspace kept between solves (fg_ctx_trim)"""
        _check(lib.fg_ctx_trim(self.h), "fg_ctx_trim")

    def close(self):
        if self.h:
            lib.fg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # device arrays
    def empty(self, n: int, dtype=np.float64) -> "DeviceArray":
        return DeviceArray(self, n, dtype)

    def to_device(self, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = DeviceArray(self, a.size, a.dtype)
        d.upload(a)
        return d


class DeviceArray:
    """A flat device buffer owned by a Context (fg_malloc/fg_free)."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = _vp()
        _check(lib.fg_malloc(ctx.h, max(1, self.n * self.dtype.itemsize), C.byref(p)),
               "fg_malloc")
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(lib.fg_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes), "h2d")

    def get(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(lib.fg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def zero(self):
        _check(lib.fg_memset(self.ctx.h, self.ptr, 0, self.nbytes), "memset")

    def free(self):
        if self.ptr and self.ctx.h:
            lib.fg_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
class Operator:
    """A stored matrix on the device: EBE, EDE or CSR (fg_op)."""

    def __init__(self, ctx: Context, h, kind: str):
        self.ctx, self.h, self.kind = ctx, h, kind
        neq, nit, k, nnz = _i(), _i(), _i(), C.c_longlong()
        lib.fg_op_info(h, C.byref(neq), C.byref(nit), C.byref(k), C.byref(nnz))
        self.neq, self.nitems, self.k, self.nnz = neq.value, nit.value, k.value, nnz.value

    @classmethod
    def ebe(cls, ctx, neq, lm, k, elems_per_block=0, ordering=None, product=None):
        """ebemv family: lm (nel*k) with sink = neq.  ordering None or
        ORDER_BFS (blocks grown breadth-first) or ORDER_MESH (runs of
        consecutive elements in lm's order); product None (PRODUCT_AUTO),
        PRODUCT_SLICED or PRODUCT_BLOCKED."""
        lm = np.ascontiguousarray(lm, np.int32)
        h = _vp()
        if ordering is None:
            _check(lib.fg_op_ebe_create(ctx.h, neq, lm.size // k, k, lm, elems_per_block,
                                        C.byref(h)), "fg_op_ebe_create")
        else:
            _check(lib.fg_op_ebe_create_ex(ctx.h, neq, lm.size // k, k, lm, elems_per_block,
                                           int(ordering), 0, C.byref(h)), "fg_op_ebe_create_ex")
        op = cls(ctx, h, "EBE")
        if product is not None:
            op.set_product_kernel(product)
        return op

    @classmethod
    def ede(cls, ctx, neq, lm2, edges_per_block=0, ordering=None):
        """edemv: lm2 (nedge*2) with sink = neq."""
        lm2 = np.ascontiguousarray(lm2, np.int32)
        h = _vp()
        if ordering is None:
            _check(lib.fg_op_ede_create(ctx.h, neq, lm2.size // 2, lm2, edges_per_block,
                                        C.byref(h)), "fg_op_ede_create")
        else:
            _check(lib.fg_op_ebe_create_ex(ctx.h, neq, lm2.size // 2, 2, lm2, edges_per_block,
                                           int(ordering), 1, C.byref(h)), "fg_op_ebe_create_ex")
        return cls(ctx, h, "EDE")

    @classmethod
    def csr(cls, ctx, neq, IA, JA):
        """csrmv: IA (neq+1), JA (nnz)."""
        IA = np.ascontiguousarray(IA, np.int32)
        JA = np.ascontiguousarray(JA, np.int32)
        h = _vp()
        _check(lib.fg_op_csr_create(ctx.h, neq, JA.size, IA, JA, C.byref(h)),
               "fg_op_csr_create")
        return cls(ctx, h, "CSR")

    @classmethod
    def ebe2(cls, base: "Operator", blockdiag: "Preconditioner", form=EBE2_BLOCKDIAG):
        """ebe2mvNDOF4: off-block product of a k = 12 operator followed by
        BlockDiag2_precond (form EBE2_BLOCKDIAG) or NOBlockDiag2_precond
        (EBE2_NOBLOCKDIAG) with the node blocks of `blockdiag`"""
        h = _vp()
        _check(lib.fg_op_ebe2_create(base.h, blockdiag.h, int(form), C.byref(h)),
               "fg_op_ebe2_create")
        op = cls(base.ctx, h, "EBE2")
        op._keep = (base, blockdiag)
        return op

    def get_values(self) -> np.ndarray:
        """element matrices (nitems*k*k, row-major per element) back to host"""
        A = np.zeros(max(self.nitems * self.k * self.k, 1))
        _check(lib.fg_op_get_values_host(self.h, A), "fg_op_get_values_host")
        return A[:self.nitems * self.k * self.k]

    def set_values(self, vals):
        if isinstance(vals, DeviceArray):
            _check(lib.fg_op_set_values_dev(self.h, vals.ptr), "fg_op_set_values_dev")
        else:
            _check(lib.fg_op_set_values_host(self.h, np.ascontiguousarray(vals, np.float64)),
                   "fg_op_set_values_host")

    def apply(self, P: DeviceArray, Q: DeviceArray):
        fn = lib.fg_op_apply_f32 if P.dtype == np.float32 else lib.fg_op_apply_f64
        _check(fn(self.h, P.ptr, Q.ptr), "fg_op_apply")

    def matvec(self, P: np.ndarray, dtype=np.float64) -> np.ndarray:
        """host convenience: Q = A P (P has neq+1 entries)."""
        dP = self.ctx.to_device(np.asarray(P), dtype)
        dQ = self.ctx.empty(self.neq + 1, dtype)
        self.apply(dP, dQ)
        return dQ.get()

    def algorithmic_bytes(self, fp32=False) -> float:
        return lib.fg_op_algorithmic_bytes(self.h, int(fp32))

    def time_apply(self, P: DeviceArray, Q: DeviceArray, reps=20, main_only=False) -> float:
        """mean ms per apply, HIP events on the library stream"""
        ms = _d()
        _check(lib.fg_op_time_apply(self.h, int(P.dtype == np.float32), P.ptr, Q.ptr, reps,
                                    int(main_only), C.byref(ms)), "fg_op_time_apply")
        return ms.value

    def set_pipeline(self, wg_per_cu: int):
        """EBE/EDE schedule: 0 = a block per workgroup, > 0 = persistent grid of
        wg_per_cu workgroups per CU with next-block prefetch"""
        _check(lib.fg_op_set_pipeline(self.h, int(wg_per_cu)), "fg_op_set_pipeline")
        self.pipeline = int(wg_per_cu)

    def set_row_kernel(self, kind):
        """ROWS_VALU or ROWS_MFMA (k = 12 element operators)"""
        _check(lib.fg_op_set_row_kernel(self.h, int(kind)), "fg_op_set_row_kernel")

    def set_edge_kernel(self, kind):
        """EDGE_AUTO (default: rows in both precisions), EDGE_ROWS or
        EDGE_BLOCKED (EDE operators)"""
        _check(lib.fg_op_set_edge_kernel(self.h, int(kind)), "fg_op_set_edge_kernel")

    def set_product_kernel(self, kind, precisions=PRODUCT_FP64 | PRODUCT_FP32):
        """PRODUCT_AUTO (default: sliced rows where the operator has them),
        PRODUCT_SLICED or PRODUCT_BLOCKED (EBE and EDE operators), for the
        FP64 products, the FP32 products or both"""
        _check(lib.fg_op_set_product_kernel(self.h, int(precisions), int(kind)),
               "fg_op_set_product_kernel")

    def product_sliced(self, fp32=False) -> bool:
        """whether the product in use for that precision is a sliced-row kernel"""
        return self.product_form(fp32) != 0

    def product_form(self, fp32=False) -> int:
        """0: a blocked kernel; 1: sliced rows; 2: the k = 4 sliced rows with
        per-equation neighbour tables"""
        v = _i()
        _check(lib.fg_op_product_kernel(self.h, int(bool(fp32)), C.byref(v)),
               "fg_op_product_kernel")
        return int(v.value)

    def close(self):
        if self.h and self.ctx.h:
            lib.fg_op_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Preconditioner:
    """fg_prec: NONE, DIAG (global diagonal) or JACOBI (element-local)."""

    def __init__(self, op: Operator, kind: int, _h=None):
        h = _vp() if _h is None else _h
        if _h is None:
            _check(lib.fg_prec_create(op.h, kind, C.byref(h)), "fg_prec_create")
        self.op, self.h, self.kind = op, h, kind

    @classmethod
    def blockdiag(cls, op: Operator, nnodes: int, vertex, ids):
        """node block-diagonal (BlockDiag_precond*) of a k = 12 EBE operator
        (vertex: nel*3) or of a k = 8 edge operator (vertex: the edge nodes,
        nedge*2); ids (nnodes*4, < 0 = prescribed)"""
        h = _vp()
        _check(lib.fg_prec_blockdiag_create(op.h, int(nnodes),
                                            np.ascontiguousarray(vertex, np.int32),
                                            np.ascontiguousarray(ids, np.int32), C.byref(h)),
               "fg_prec_blockdiag_create")
        pc = cls(op, PREC_BLOCKDIAG, _h=h)
        pc.nnodes = int(nnodes)
        return pc

    def blocks(self):
        """block-diagonal: (BD, invBD), each nnodes*16 row-major node blocks"""
        BD, iBD = np.zeros(16 * self.nnodes), np.zeros(16 * self.nnodes)
        _check(lib.fg_prec_blockdiag_get(self.h, BD.ctypes.data, iBD.ctypes.data),
               "fg_prec_blockdiag_get")
        return BD, iBD

    def block_scaling(self, F: DeviceArray | None):
        """Block_scaling_EBE: setup (F preconditioned) and the operator's
        element matrices scaled in place"""
        _check(lib.fg_block_scaling_ebe(self.h, F.ptr if F is not None else None),
               "fg_block_scaling_ebe")

    def setup(self, F: DeviceArray | None):
        _check(lib.fg_prec_setup(self.h, F.ptr if F is not None else None), "fg_prec_setup")

    def scaling(self, mode: int, F: DeviceArray):
        """Left_scaling_* (SCALING_LEFT) or LeftRight_scaling_* (SCALING_LEFTRIGHT)
        of this Diag preconditioner's operator; F scaled in place"""
        _check(lib.fg_scaling_diag(self.h, int(mode), F.ptr), "fg_scaling_diag")

    def unscaling_left(self, x: DeviceArray):
        """Left_unscaling: x[i] *= sqrt(invDiag[i])"""
        _check(lib.fg_unscaling_left(self.h, x.ptr), "fg_unscaling_left")

    def apply(self, p: DeviceArray, z: DeviceArray):
        fn = lib.fg_prec_apply_f32 if p.dtype == np.float32 else lib.fg_prec_apply_f64
        _check(fn(self.h, p.ptr, z.ptr), "fg_prec_apply")

    def diag(self):
        n = self.op.neq + 1
        D, iD = np.zeros(n), np.zeros(n)
        _check(lib.fg_prec_get_diag(self.h, D.ctypes.data, iD.ctypes.data), "get_diag")
        return D, iD

    def ilu_levels(self):
        """ILU(0): (lower levels, upper levels, kernel segments)"""
        a, b, c = _i(), _i(), _i()
        _check(lib.fg_prec_ilu_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "ilu_info")
        return a.value, b.value, c.value

    def ilu_schedule(self, wide_rows: int):
        """ILU(0): levels wider than wide_rows get a launch each"""
        _check(lib.fg_prec_ilu_schedule(self.h, int(wide_rows)), "ilu_schedule")

    def ilu_sweeps(self, sweeps: int):
        """ILU(0): 0 = exact triangular solves, s > 0 = s Jacobi sweeps each
        (set before setup)"""
        _check(lib.fg_prec_ilu_sweeps(self.h, int(sweeps)), "ilu_sweeps")

    def ilu_factors(self):
        """ILU(0): (LU[nnz] in the operator's CSR slots, D[neq] inverted pivots)"""
        LU = np.zeros(max(int(self.op.nnz), 1))
        D = np.zeros(max(self.op.neq, 1))
        _check(lib.fg_prec_ilu_get(self.h, LU.ctypes.data, D.ctypes.data), "ilu_get")
        return LU[:int(self.op.nnz)], D[:self.op.neq]

    def close(self):
        if self.h and self.op.h and self.op.ctx.h:
            lib.fg_prec_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def concat_lm(parts):
    """Block-diagonal layout of independent problems for gmres_batched.

    parts: sequence of (neq_p, lm_p) with lm_p's sink = neq_p.  Returns
    (lm, seg): problem p's equations are shifted by seg[p]; every sink maps
    to the shared sink slot seg[-1] = sum of neq_p."""
    seg = np.zeros(len(parts) + 1, np.int32)
    for p, (neq_p, _) in enumerate(parts):
        seg[p + 1] = seg[p] + int(neq_p)
    total = int(seg[-1])
    out = []
    for p, (neq_p, lm_p) in enumerate(parts):
        lm_p = np.asarray(lm_p, np.int64)
        out.append(np.where(lm_p == neq_p, total, lm_p + int(seg[p])).astype(np.int32))
    return np.concatenate(out) if out else np.zeros(0, np.int32), seg


def gmres_batched(ctx: Context, op: Operator, prec: Preconditioner | None, seg,
                  B: DeviceArray, X: DeviceArray, kmax=30, lmax=1000, tol=1e-8,
                  precision=GMRES_FP64, inner_tol=0.0, graph=False, orth=ORTH_MGS):
    """fg_gmres_batched: independent problems of a block-diagonal operator
    (problem p = equations [seg[p], seg[p+1])) solved together.  orth:
    ORTH_MGS, or ORTH_CGS for the FP32-IR inner cycle.
    Returns (status, [GmresInfo per problem])."""
    seg = np.ascontiguousarray(seg, np.int32)
    np_ = len(seg) - 1
    infos = (GmresInfo * np_)()
    o = GmresOpts(kmax, lmax, tol, precision, inner_tol, 0, None, None, 0, int(orth))
    rc = lib.fg_gmres_batched(ctx.h, op.h, prec.h if prec is not None else None, np_, seg,
                              B.ptr, X.ptr, C.byref(o), int(bool(graph)), C.cast(infos, _vp))
    _check(rc, "fg_gmres_batched", ok=(FG_OK, FG_ERR_NOCONV, FG_ERR_BREAKDOWN))
    return rc, list(infos)


def gmres(ctx: Context, op: Operator, prec: Preconditioner | None, B: DeviceArray,
          X: DeviceArray, kmax=30, lmax=1000, tol=1e-8, precision=GMRES_FP64,
          inner_tol=0.0, hist_cap=0, e0_cap=0, engine=ENGINE_KERNELS, orth=ORTH_MGS):
    """pgmres (Solvers_and_Preconditioners/pgmres.c:4-211) on the device.

    engine: ENGINE_KERNELS (one kernel per vector operation), ENGINE_FUSED
    (the whole solve in one persistent kernel), ENGINE_SMALL (the whole FP64
    solve in one workgroup, systems of at most 4096 equations),
    ENGINE_GRAPH (KERNELS with each restart cycle replayed from one captured
    hipGraph, kept in the context for later solves) or ENGINE_AUTO (SMALL
    where it applies, else KERNELS).
    orth: ORTH_MGS or ORTH_CGS for the FP32-IR inner cycle.
    Returns (info, rho_hist, e0_hist)."""
    rho = np.zeros(max(1, hist_cap))
    e0 = np.zeros(max(1, e0_cap, lmax + 1))   # the library writes up to lmax+1 entries
    o = GmresOpts(kmax, lmax, tol, precision, inner_tol, hist_cap,
                  rho.ctypes.data if hist_cap else None,
                  e0.ctypes.data if e0_cap else None, engine, orth)
    info = GmresInfo()
    rc = lib.fg_gmres(ctx.h, op.h, prec.h if prec is not None else None, B.ptr, X.ptr,
                      C.byref(o), C.byref(info))
    _check(rc, "fg_gmres", ok=(FG_OK, FG_ERR_NOCONV))
    return info, rho[:min(info.steps, hist_cap)].copy(), e0[:info.cycles + 1].copy()


def predictor_old_transport(mesh: "Mesh2D", prob: "Supg2dProblem", op: Operator,
                            prec: Preconditioner | None, upresc: DeviceArray, u: DeviceArray,
                            alpha, dt, final_time, nl_tol, nl_maxiter, stop_by_norm=True,
                            steady=False, time_tol=0.0, max_steps=0, kmax=30, lmax=1000,
                            tol=1e-8, precision=GMRES_FP64, engine=ENGINE_AUTO,
                            orth=ORTH_MGS, hist_cap=4096):
    """Predictor_Old (Time_Integration/Predictor_Old.c:4-145) over the
    transient build, on the device; u (neq+1) holds the initial solution and
    receives the result.  Returns (info, corrections per step)."""
    pp = PredictorParams(alpha, dt, final_time, nl_tol, nl_maxiter, int(stop_by_norm),
                         int(steady), time_tol, max_steps)
    o = GmresOpts(kmax, lmax, tol, precision, 0.0, 0, None, None, engine, orth)
    hist = np.zeros(max(1, hist_cap), np.int32)
    info = PredictorInfo()
    rc = lib.fg_predictor_old_transport(mesh.h, C.byref(prob), op.h,
                                        prec.h if prec is not None else None, upresc.ptr,
                                        C.byref(pp), C.byref(o), u.ptr, hist.ctypes.data,
                                        hist_cap, C.byref(info))
    _check(rc, "fg_predictor_old_transport", ok=(FG_OK, FG_ERR_NOCONV))
    return info, hist[:min(info.steps, hist_cap)].copy()


class Ede4:
    """fg_ede4: the edge structure of the 4-dof (Euler) driver
    (ede_Initialization.c and ede_assembly.c of 03_Euler_Equations)."""

    def __init__(self, ctx: Context, nnodes, vertex, ids, neq):
        vertex = np.ascontiguousarray(vertex, np.int32)
        ids = np.ascontiguousarray(ids, np.int32)
        h = _vp()
        _check(lib.fg_ede4_create(ctx.h, int(nnodes), vertex.size // 3, int(neq), vertex, ids,
                                  C.byref(h)), "fg_ede4_create")
        self.ctx, self.h, self.nel, self.neq = ctx, h, vertex.size // 3, int(neq)
        n = _i()
        lib.fg_ede4_info(h, C.byref(n))
        self.nedge = n.value

    def arrays(self):
        """(order[nel*3], edge_by_elem[nel*3], lm2[nedge*8], edge_nodes[nedge*2])"""
        o, e = np.zeros(3 * self.nel, np.int32), np.zeros(3 * self.nel, np.int32)
        l2, en = np.zeros(8 * self.nedge, np.int32), np.zeros(2 * self.nedge, np.int32)
        _check(lib.fg_ede4_get(self.h, o.ctypes.data, e.ctypes.data, l2.ctypes.data,
                               en.ctypes.data), "fg_ede4_get")
        return o, e, l2, en

    def assemble(self, A12: DeviceArray, Aede: DeviceArray):
        _check(lib.fg_ede4_assemble(self.h, A12.ptr, Aede.ptr), "fg_ede4_assemble")

    def close(self):
        if self.h and self.ctx.h:
            lib.fg_ede4_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PRED_EU_NOT, PRED_EU_BLOCK_SCALING, PRED_EU_BLOCKDIAG, PRED_EU_TWO = 0, 1, 2, 3
SCALING_LEFT, SCALING_LEFTRIGHT = 1, 2


def predictor_old_euler(mesh: "Euler2D", ep: "EulerParams", op: Operator,
                        blocks: Preconditioner | None, mode: int, presc: DeviceArray,
                        rot: DeviceArray | None, theta: DeviceArray | None, u: DeviceArray,
                        alpha, dt, final_time, nl_tol, nl_maxiter, stop_by_norm=True,
                        steady=False, time_tol=0.0, max_steps=0, kmax=30, lmax=1000, tol=1e-8,
                        precision=GMRES_FP64, engine=ENGINE_AUTO, orth=ORTH_MGS,
                        hist_cap=4096, edges: "Ede4 | None" = None):
    """Predictor_Old over the Euler build, on the device (mode: PRED_EU_NOT,
    PRED_EU_BLOCK_SCALING, PRED_EU_BLOCKDIAG, PRED_EU_TWO with op the ebe2
    operator); with `edges` the edge-by-edge scheme (op: the k = 8 operator
    on the edges' lm2, or its ede2 operator).  u (neq+1) holds the initial
    solution and receives the result.  Returns (info, corrections per step)."""
    pp = PredictorParams(alpha, dt, final_time, nl_tol, nl_maxiter, int(stop_by_norm),
                         int(steady), time_tol, max_steps)
    o = GmresOpts(kmax, lmax, tol, precision, 0.0, 0, None, None, engine, orth)
    hist = np.zeros(max(1, hist_cap), np.int32)
    info = PredictorInfo()
    bh = blocks.h if blocks is not None else None
    rp = rot.ptr if rot is not None else None
    tp = theta.ptr if theta is not None else None
    if edges is None:
        rc = lib.fg_predictor_old_euler(mesh.h, C.byref(ep), op.h, bh, int(mode), presc.ptr, rp,
                                        tp, C.byref(pp), C.byref(o), u.ptr, hist.ctypes.data,
                                        hist_cap, C.byref(info))
    else:
        rc = lib.fg_predictor_old_euler_ede(mesh.h, C.byref(ep), edges.h, op.h, bh, int(mode),
                                            presc.ptr, rp, tp, C.byref(pp), C.byref(o), u.ptr,
                                            hist.ctypes.data, hist_cap, C.byref(info))
    _check(rc, "fg_predictor_old_euler", ok=(FG_OK, FG_ERR_NOCONV))
    return info, hist[:min(info.steps, hist_cap)].copy()


# ---------------------------------------------------------------------------
# mesh structures (host; Preprocess.c / Fill_LM.c / csr_ / ede_Initialization.c)
# ---------------------------------------------------------------------------
def number_nodes(type_):
    type_ = np.ascontiguousarray(type_, np.int32)
    idv = np.zeros(type_.size, np.int32)
    neq = lib.fg_number_nodes(type_.size, type_, idv)
    return neq, idv


def fill_lm(neq, vertex, type_, idv, nnoel=3):
    vertex = np.ascontiguousarray(vertex, np.int32)
    lm = np.zeros(vertex.size, np.int32)
    _check(lib.fg_fill_lm(neq, vertex.size // nnoel, nnoel, vertex,
                          np.ascontiguousarray(type_, np.int32), idv, lm), "fg_fill_lm")
    return lm


def csr_pattern(neq, lm):
    """-> (nnz, IA, JA, csr_by_elem)"""
    lm = np.ascontiguousarray(lm, np.int32)
    nel = lm.size // 3
    nnz = C.c_longlong()
    _check(lib.fg_csr_pattern(neq, nel, lm, None, None, None, C.byref(nnz)), "csr_pattern")
    IA, JA = np.zeros(neq + 1, np.int32), np.zeros(nnz.value, np.int32)
    cbe = np.zeros(9 * nel, np.int32)
    _check(lib.fg_csr_pattern(neq, nel, lm, IA.ctypes.data, JA.ctypes.data, cbe.ctypes.data,
                              C.byref(nnz)), "csr_pattern")
    return nnz.value, IA, JA, cbe


def ede_pattern(neq, lm):
    """-> (nedge, order, edge_by_elem, lm2)"""
    lm = np.ascontiguousarray(lm, np.int32)
    nel = lm.size // 3
    ne = _i()
    _check(lib.fg_ede_pattern(neq, nel, lm, None, None, None, C.byref(ne)), "ede_pattern")
    order, ebe = np.zeros(3 * nel, np.int32), np.zeros(3 * nel, np.int32)
    lm2 = np.zeros(2 * ne.value, np.int32)
    _check(lib.fg_ede_pattern(neq, nel, lm, order.ctypes.data, ebe.ctypes.data,
                              lm2.ctypes.data, C.byref(ne)), "ede_pattern")
    return ne.value, order, ebe, lm2


def structured_mesh2d(n: int):
    """Unit square, n x n cells split along the "direita" diagonal with the
    node and element numbering of Gerador_de_Malhas/Structured_Mesh/geraMalha.c:112-176;
    boundary nodes Type 0 (Dirichlet), inner nodes Type 1.
    -> (x, y, type, vertex[nel*3])"""
    W = n + 1
    i = np.arange(W, dtype=np.float64)
    h = 1.0 / n
    x = np.tile(0.0 + i * h, W)
    y = np.repeat(0.0 + i * h, W)
    bnd = (np.abs(x - 1.0) <= 1e-15) | (np.abs(x) <= 1e-15) | \
          (np.abs(y - 1.0) <= 1e-15) | (np.abs(y) <= 1e-15)
    type_ = np.where(bnd, 0, 1).astype(np.int32)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (jj * W + ii).ravel().astype(np.int64)
    v = np.empty((n * n, 2, 3), np.int64)
    v[:, 0, 0], v[:, 0, 1], v[:, 0, 2] = a, a + 1, a + 1 + W
    v[:, 1, 0], v[:, 1, 1], v[:, 1, 2] = a, a + 1 + W, a + W
    return x, y, type_, v.reshape(-1).astype(np.int32)


class Mesh2D:
    """fg_mesh2d: device copy of a P1 triangle mesh + SUPG assembly."""

    def __init__(self, ctx: Context, x, y, type_, vertex):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        type_ = np.ascontiguousarray(type_, np.int32)
        vertex = np.ascontiguousarray(vertex, np.int32)
        h = _vp()
        _check(lib.fg_mesh2d_create(ctx.h, x.size, vertex.size // 3, x, y, type_, vertex,
                                    C.byref(h)), "fg_mesh2d_create")
        self.ctx, self.h = ctx, h
        neq, nel, nn = _i(), _i(), _i()
        lib.fg_mesh2d_info(h, C.byref(neq), C.byref(nel), C.byref(nn))
        self.neq, self.nel, self.nnodes = neq.value, nel.value, nn.value

    def lm(self):
        out = np.zeros(3 * self.nel, np.int32)
        _check(lib.fg_mesh2d_get_lm(self.h, out), "get_lm")
        return out

    def build_supg(self, prob: Supg2dProblem, ke: DeviceArray, F: DeviceArray | None,
                   u: DeviceArray | None = None):
        """Build_K_F_SUPG on the device: ke (nel*9), F (neq+1)."""
        _check(lib.fg_supg2d_build(self.h, C.byref(prob), u.ptr if u is not None else None,
                                   ke.ptr, F.ptr if F is not None else None), "fg_supg2d_build")

    def build_transient(self, prob: Supg2dProblem, alpha: float, dt: float, U: DeviceArray,
                        dU: DeviceArray, A: DeviceArray, R: DeviceArray):
        """Build_M_K_R_SUPG on the device: A (nel*9) = me + alpha dt ke,
        R (neq+1) = F - M dU - K U; U, dU nodal (nnodes)."""
        _check(lib.fg_supg2d_transient_build(self.h, C.byref(prob), float(alpha), float(dt),
                                             U.ptr, dU.ptr, A.ptr, R.ptr),
               "fg_supg2d_transient_build")

    def close(self):
        if self.h and self.ctx.h:
            lib.fg_mesh2d_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble_csr(op: Operator, csr_by_elem, ke: DeviceArray, k: int = 3):
    cbe = np.ascontiguousarray(csr_by_elem, np.int32)
    _check(lib.fg_op_assemble_csr_k(op.h, cbe.size // (k * k), k, cbe, ke.ptr), "assemble_csr")


def csr_pattern_k(neq, lm, k):
    """csr_Initialization for k dofs per element -> (nnz, IA, JA, csr_by_elem)"""
    lm = np.ascontiguousarray(lm, np.int32)
    nel = lm.size // k
    nnz = C.c_longlong()
    _check(lib.fg_csr_pattern_k(neq, nel, k, lm, None, None, None, C.byref(nnz)), "csr_pattern_k")
    IA, JA = np.zeros(neq + 1, np.int32), np.zeros(nnz.value, np.int32)
    cbe = np.zeros(k * k * nel, np.int32)
    _check(lib.fg_csr_pattern_k(neq, nel, k, lm, IA.ctypes.data, JA.ctypes.data,
                                cbe.ctypes.data, C.byref(nnz)), "csr_pattern_k")
    return nnz.value, IA, JA, cbe


def rcm(neq, IA, JA):
    """reverse Cuthill-McKee (symrcm.c): perm[c] = old index of new equation c"""
    IA = np.ascontiguousarray(IA, np.int32)
    JA = np.ascontiguousarray(JA, np.int32)
    perm = np.zeros(max(neq, 1), np.int32)
    _check(lib.fg_rcm(neq, IA, JA, perm), "fg_rcm")
    return perm[:neq]


def csr_bandwidth(neq, IA, JA):
    bw = _i()
    _check(lib.fg_csr_bandwidth(neq, np.ascontiguousarray(IA, np.int32),
                                np.ascontiguousarray(JA, np.int32), C.byref(bw)), "bandwidth")
    return bw.value


def renumber_lm(neq, lm, perm):
    """lm with equations renumbered by a permutation perm[new] = old (the
    sink neq stays the sink), as csr_Initialization.c renumbers Node[].id"""
    inv = np.empty(neq + 1, np.int64)
    inv[np.asarray(perm, np.int64)] = np.arange(neq)
    inv[neq] = neq
    return inv[np.asarray(lm, np.int64)].astype(np.int32)


class Supg3dProblem(C.Structure):
    """fg_supg3d_problem: DIFCONV (kappa = A, beta = (1,1,1), sigma = 0,
    manufactured f) or constant coefficients."""
    _fields_ = [("kind", C.c_int), ("A", C.c_double), ("kappa", C.c_double),
                ("bx", C.c_double), ("by", C.c_double), ("bz", C.c_double),
                ("sigma", C.c_double), ("f", C.c_double)]


def structured_mesh3d(n: int):
    """Unit cube, n^3 cells, each split into the 6 tetrahedra around its main
    diagonal (conforming Kuhn split).  Node (i,j,k) -> i + (n+1)(j + (n+1)k);
    boundary nodes uType 0 (prescribed), inner nodes uType 1 (unknown).
    Elements cell by cell (i fastest), the axis permutations (0,1,2), (0,2,1),
    (1,0,2), (1,2,0), (2,0,1), (2,1,0) in that order, the last two vertices
    swapped for the odd ones so every volume is positive.
    -> (x, y, z, utype, vertex[nel*4])"""
    N = n + 1
    g = np.arange(N, dtype=np.float64) / n
    z, y, x = [a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")]
    ii = np.arange(N)
    kk, jj, i3 = [a.ravel() for a in np.meshgrid(ii, ii, ii, indexing="ij")]
    inner = (i3 > 0) & (i3 < n) & (jj > 0) & (jj < n) & (kk > 0) & (kk < n)
    utype = inner.astype(np.int32)
    ck, cj, ci = [a.ravel().astype(np.int64) for a in
                  np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")]
    v0 = ci + N * (cj + N * ck)
    stride = np.array([1, N, N * N], np.int64)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    V = np.empty((n ** 3, 6, 4), np.int64)
    for p, (a, b, c) in enumerate(perms):
        w1 = v0 + stride[a]
        w2 = w1 + stride[b]
        w3 = w2 + stride[c]
        if p in (1, 2, 5):
            w2, w3 = w3, w2
        V[:, p, 0], V[:, p, 1], V[:, p, 2], V[:, p, 3] = v0, w1, w2, w3
    return x.copy(), y.copy(), z.copy(), utype, V.reshape(-1).astype(np.int32)


class Mesh3D:
    """fg_mesh3d: device copy of a P1 tetrahedral mesh + 3D SUPG assembly."""

    def __init__(self, ctx: Context, x, y, z, utype, vertex):
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        utype = np.ascontiguousarray(utype, np.int32)
        vertex = np.ascontiguousarray(vertex, np.int32)
        h = _vp()
        _check(lib.fg_mesh3d_create(ctx.h, x.size, vertex.size // 4, x, y, z, utype, vertex,
                                    C.byref(h)), "fg_mesh3d_create")
        self.ctx, self.h = ctx, h
        neq, nel, nn = _i(), _i(), _i()
        lib.fg_mesh3d_info(h, C.byref(neq), C.byref(nel), C.byref(nn))
        self.neq, self.nel, self.nnodes = neq.value, nel.value, nn.value

    def lm(self):
        out = np.zeros(4 * self.nel, np.int32)
        _check(lib.fg_mesh3d_get_lm(self.h, out), "get_lm")
        return out

    def build_supg(self, prob: Supg3dProblem, ke: DeviceArray, F: DeviceArray | None):
        """Build_SUPG (3D) on the device: ke (nel*16), F (neq+1)."""
        _check(lib.fg_supg3d_build(self.h, C.byref(prob), ke.ptr,
                                   F.ptr if F is not None else None), "fg_supg3d_build")

    def close(self):
        if self.h and self.ctx.h:
            lib.fg_mesh3d_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble_ede(op: Operator, order, edge_by_elem, ke: DeviceArray):
    order = np.ascontiguousarray(order, np.int32)
    ebe = np.ascontiguousarray(edge_by_elem, np.int32)
    _check(lib.fg_op_assemble_ede(op.h, order.size // 3, order, ebe, ke.ptr), "assemble_ede")


class Euler2D:
    """fg_euler2d: device copy of a 4-dof triangle mesh + the Euler SUPG
    element build (Build_M_K_F_SUPG, CAU)."""

    def __init__(self, ctx: Context, x, y, vertex, lm, neq):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        vertex = np.ascontiguousarray(vertex, np.int32)
        lm = np.ascontiguousarray(lm, np.int32)
        h = _vp()
        _check(lib.fg_euler2d_create(ctx.h, x.size, vertex.size // 3, int(neq), x, y, vertex, lm,
                                     C.byref(h)), "fg_euler2d_create")
        self.ctx, self.h = ctx, h
        self.nnodes, self.nel, self.neq = x.size, vertex.size // 3, int(neq)

    def build(self, params: EulerParams, U: DeviceArray, dU: DeviceArray, A: DeviceArray,
              R: DeviceArray, rot: DeviceArray | None = None, theta: DeviceArray | None = None):
        _check(lib.fg_euler2d_build(self.h, C.byref(params), U.ptr, dU.ptr,
                                    rot.ptr if rot is not None else None,
                                    theta.ptr if theta is not None else None, A.ptr, R.ptr),
               "fg_euler2d_build")

    def close(self):
        if self.h and self.ctx.h:
            lib.fg_euler2d_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# on-disk formats (host): parameter and mesh files of the drivers
# ---------------------------------------------------------------------------
# field order and kind of each driver's parameter file (Preprocess.c)
PARAM_FIELDS = {
    # 01_SS_Transport_Equation/2D/01_CommonFiles/Preprocess.c:44-61
    "ss2d": [("Experiments", "s"), ("ProblemTitle", "s"), ("SolverTolerance", "d"),
             ("NonLinearTolerance", "d"), ("SolverMaxIter", "i"),
             ("KrylovBasisVectorsQuantity", "i"), ("Solver", "s"), ("Preconditioner", "s"),
             ("Scaling", "s"), ("reordering", "s"), ("MatrixVectorProductScheme", "s"),
             ("StabilizationForm", "s"), ("ShockCapture", "s"), ("h_Shock", "s"),
             ("OutputFlow", "s"), ("nnodes", "i"), ("nel", "i")],
    # 03_Euler_Equations/2D/01_CommonFiles/Preprocess.c:45-74
    "euler2d": [("Experiments", "s"), ("ProblemTitle", "s"), ("SolverTolerance", "d"),
                ("NonLinearTolerance", "d"), ("TimeIntegrationTolerance", "d"),
                ("StabilizationTolerance", "d"), ("LinearMaxIter", "i"),
                ("KrylovBasisVectorsQuantity", "i"), ("Solver", "s"), ("Preconditioner", "p"),
                ("Scaling", "s"), ("reordering", "s"), ("TimeIntegration", "s"),
                ("StopAtSteadyState", "s"), ("Alpha", "d"), ("DeltaT", "d"),
                ("FinalTime", "d"), ("Dimensionless", "s"), ("Mach", "d"),
                ("NonLinearMaxIter", "i"), ("StopMulticorrection", "s"),
                ("MatrixVectorProductScheme", "s"), ("StabilizationForm", "s"),
                ("ShockCapture", "s"), ("invY0", "d"), ("invY1", "d"), ("invY2", "d"),
                ("invY3", "d"), ("nnodes", "i"), ("nel", "i")],
    # 01_SS_Transport_Equation/3D/01_CommonFiles/Preprocess.c:45-83
    "ss3d": [("Experiments", "s"), ("ProblemTitle", "s"), ("NumberMesh", "i"),
             ("OriginMesh", "s"), ("TypeMesh", "s"), ("ConstApli", "d"),
             ("ExactSolution", "s"), ("InitialSolution", "s"), ("Solver", "s"),
             ("KrylovBasisVectorsQuantity", "i"), ("SolverMaxIter", "i"),
             ("SolverTolerance", "d"), ("Preconditioner", "s"), ("Reordering", "s"),
             ("MatrixVectorProductScheme", "s"), ("StabilizationForm", "s"),
             ("UseDamping", "s"), ("ComputeResidual", "s"), ("UsePeclet", "s"),
             ("TauMacro", "s"), ("TauMicro", "s"), ("tolGradU", "d"), ("wMacro", "d"),
             ("wMicro", "d"), ("tetha", "d"), ("ErrorType", "s"), ("NonLinearTolerance", "d"),
             ("NonLinearMaxIter", "i"), ("TipoH", "i"), ("OutputFlow", "s"), ("Vmax", "d"),
             ("xInf", "d"), ("xSup", "d"), ("yInf", "d"), ("ySup", "d"), ("zInf", "d"),
             ("zSup", "d"), ("nnodes", "i"), ("nel", "i")],
}

# (ndim, node integers, vertices per element, element integers) of each mesh format
MESH_FORMATS = {"ss2d": (2, 1, 3, 1), "euler2d": (2, 4, 3, 1), "ss3d": (3, 1, 4, 1)}


def read_param_file(path, driver):
    """the driver's parameter file as {field: value} (typed as Preprocess.c reads it)"""
    fields = PARAM_FIELDS[driver]
    kinds = "".join(k for _, k in fields)
    buf = C.create_string_buffer(1 << 16)
    n = _i()
    _check(lib.fg_param_file_read(os.fsencode(path), kinds.encode(), buf, len(buf), C.byref(n)),
           "fg_param_file_read")
    raw = buf.raw.split(b"\0")[:n.value]
    out = {}
    for (name, kind), v in zip(fields, raw):
        v = v.decode()
        out[name] = float(v) if kind == "d" else int(v) if kind == "i" else v
    return out


def read_mesh_file(path, driver, one_based=False):
    """(coords [nnodes, ndim], node integers [nnodes, nint], vertex [nel*nvert],
    element integers [nel, ncols]) of a driver's mesh .dat file; one_based:
    the vertices are numbered from 1 (the 3D driver's GMSH meshes,
    Preprocess.c:189-193) and are returned from 0"""
    ndim, nint, nvert, ecols = MESH_FORMATS[driver]
    nn, ne = _i(), _i()
    p = os.fsencode(path)
    _check(lib.fg_mesh_file_read(p, ndim, nint, nvert, ecols, C.byref(nn), C.byref(ne), None,
                                 None, None, None), "fg_mesh_file_read")
    xyz = np.zeros((nn.value, ndim))
    ni = np.zeros((nn.value, nint), np.int32)
    v = np.zeros(ne.value * nvert, np.int32)
    ei = np.zeros((ne.value, ecols), np.int32)
    _check(lib.fg_mesh_file_read(p, ndim, nint, nvert, ecols, C.byref(nn), C.byref(ne),
                                 xyz.ctypes.data, ni.ctypes.data, v.ctypes.data,
                                 ei.ctypes.data), "fg_mesh_file_read")
    if one_based:
        v -= 1
    if v.size and (v.min() < 0 or v.max() >= nn.value):
        raise ValueError(f"{path}: element vertex outside [0, {nn.value})")
    return xyz, ni, v, ei
