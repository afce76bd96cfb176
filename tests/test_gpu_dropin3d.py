"""Drop-in check for the steady 3D driver: its solve call sequence
(precond_setup on the host, then the solver setSolver.c selects: gmres for
Preconditioner NOT, pgmres with ILUp_precond for ILU0 on CSR) driven through
fem_codes_amd/adapters/ss3d_adapter.c by oracle/dropin3d_harness.c, which
links the reference's own ILUp_precond_setup / ILUp_precond / NO_precond
(oracle/_ref/libss3d_dropin.so, built where the FEM_CODES tree is present).

System: the oracle's Kuhn mesh 8^3 (3072 tetrahedra, 343 equations) with
the DIFCONV SUPG matrices, EBE and CSR storage.  Against the oracle:
products within 1e-14 relative; gmres / pgmres with the oracle's step
count, X within 1e-11 relative, ResGMRES = the final residual estimate
(<= tol * ||F||) and ContGMRES = SolverIterations = the step count; the
host ILU(0) setup's preconditioned F equal to the oracle's bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libss3d_dropin.so")
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def system():
    o = O.Oracle()
    x, y, z, t, v = o.mesh3d(8)
    neq, idv = o.number_nodes(t)
    nel = len(v) // 4
    lm = o.fill_lm(neq, nel, v, t, idv, nnoel=4)
    ke, F = o.build_supg3d(x, y, z, t, v, lm, neq, O.prob3d(O.P3_DIFCONV, A=1.0))
    nnz, IA, JA, cbe = o.csr_init_k(neq, nel, 4, lm)
    AA = o.csr_assembly_k(nel, 4, nnz, ke, cbe)
    return dict(o=o, neq=neq, nel=nel, lm=lm, ke=ke, F=F, nnz=nnz, IA=IA, JA=JA, AA=AA)


def _lib():
    L = C.CDLL(LIB)
    L.d3_setup.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, _ip, _dp, C.c_char_p,
                           C.c_char_p, C.c_int, C.c_int, C.c_double]
    L.d3_matvec.argtypes = [_dp, _dp]
    L.d3_solve.argtypes = [_dp, _dp, _dp, _dp]
    return L


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


needs_lib = pytest.mark.skipif(not os.path.exists(LIB),
                               reason="oracle/_ref/libss3d_dropin.so is built only where the "
                                      "FEM_CODES tree is")


@needs_lib
@pytest.mark.parametrize("scheme,precond", [("EBE", "NOT"), ("CSR", "NOT"), ("CSR", "ILU0")])
def test_dropin3d_solve(system, scheme, precond):
    s, o = system, system["o"]
    L = _lib()
    neq, tol = s["neq"], 1e-10
    L.d3_setup(neq, s["nel"], s["lm"], s["ke"], s["nnz"], s["IA"], s["JA"], s["AA"],
               scheme.encode(), precond.encode(), 30, 200, tol)
    P = np.random.default_rng(3).standard_normal(neq + 1)
    P[neq] = 0.0
    Q = np.zeros(neq + 1)
    assert L.d3_matvec(P.copy(), Q) == 0
    if scheme == "EBE":
        Qo = o.ebemv(neq, s["nel"], 4, s["ke"], s["lm"], P.copy())
    else:
        Qo = o.csrmv(neq, s["AA"], s["JA"], s["IA"], P.copy())
    assert _rel(Q[:neq], Qo[:neq]) <= 1e-14 and Q[neq] == 0.0

    u, Fpre, st = np.zeros(neq + 1), np.zeros(neq + 1), np.zeros(3)
    assert L.d3_solve(s["F"].copy(), u, Fpre, st) == 0
    mats = (s["ke"], s["lm"], 4, s["nel"]) if scheme == "EBE" else (s["AA"], s["JA"], s["IA"])
    if precond == "NOT":
        assert np.array_equal(Fpre, s["F"])
        Xo, steps, _, rho, _ = o.pgmres(neq, 30, 200, tol, s["F"], scheme, mats)
        assert st[2] == steps and st[0] == steps          # ContGMRES, SolverIterations
        assert 0.0 <= st[1] <= tol * np.linalg.norm(s["F"][:neq])   # ResGMRES
        assert abs(st[1] - rho[steps - 1]) <= 1e-10 * rho[steps - 1] + 1e-14 * rho[0]
    else:
        rc, LU, D = o.ilu0_factor(neq, s["AA"], s["JA"], s["IA"])
        Bo = o.ilu0_apply(neq, LU, s["JA"], s["IA"], D, s["F"])
        assert np.array_equal(Fpre, Bo)
        Xo, steps, _, _, _ = o.pgmres(neq, 30, 200, tol, Bo, scheme, mats,
                                      ilu=(LU, s["JA"], s["IA"], D))
        assert st[0] == steps
        assert st[1] == -1.0 and st[2] == -1.0   # pgmres.c leaves ResGMRES/ContGMRES alone
    assert _rel(u[:neq], Xo[:neq]) <= 1e-11
