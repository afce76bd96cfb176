"""Write the golden fixtures of the SS-2D solve path from the reference itself.

Runs in a container that holds the FEM_CODES tree: it drives the reference
routines compiled from their own sources (oracle/_ref/libref_ss2d.so, built
by ``make -C oracle``) and stores inputs and outputs as small .npz files next
to this script.  The fixtures are data only; the reference code does not
travel with them.

    python tests/golden/make_golden.py

Cases (structured unit-square meshes, geraMalha "direita" numbering):
  n in {8, 16}; problems CONST (beta=(cos t, sin t), k=1e-4, f=1),
  ROTATING (scale 1.3, k=1e-3) and VALIDACAO (the reference's manufactured
  solution problem); the three storage schemes EBE/EDE/CSR.
Each file holds: mesh, lm, scheme structures, the element / edge / CSR
matrix and F after Build_K_F_SUPG at u=0 and at a seeded nonzero u (shock
capturing active), Q = A P for a seeded P, the Diag setup (Diag, invDiag,
scaled F), the element Jacobi setup (EBE only), and pgmres solutions:
full solves (NOT and Diag) and first-cycle iterates X_k for k = 1..K.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

PROBLEMS = {
    "const": dict(kind=O.P_CONST, kappa=1e-4, gamma=0.0, f=1.0, bx=float(np.cos(2.1)),
                  by=float(np.sin(2.1)), scale=1.0, h_shock=O.H_2SQRTAREA),
    "rot": dict(kind=O.P_ROTATING, kappa=1e-3, gamma=0.0, f=1.0, bx=0.0, by=0.0,
                scale=1.3, h_shock=O.H_SQRTAREA),
    "validacao": dict(kind=O.P_VALIDACAO, kappa=1.0, gamma=1.0, f=0.0, bx=0.0, by=0.0,
                      scale=1.0, h_shock=O.H_2SQRTAREA),
}
KHIST = 12


def one_case(n, pname, scheme):
    o, r = O.Oracle(), O.Reference()
    x, y, t, v = o.mesh2d(n)
    neq = r.setup(x, y, t, v, scheme)
    p = O.prob2d(**PROBLEMS[pname])
    r.set_problem(p)
    rng = np.random.default_rng(1000 * n + len(pname) + len(scheme))
    out = dict(n=n, scheme=scheme, x=x, y=y, type=t, vertex=v, neq=neq,
               ids=r.ids(len(x)), lm=r.lm(), prob=np.array([PROBLEMS[pname][k] for k in
                                                           ("kind", "kappa", "gamma", "f", "bx", "by", "scale", "h_shock")]))
    if scheme == "CSR":
        out["IA"], out["JA"], out["csr_by_elem"] = r.csr()
    if scheme == "EDE":
        out["order"], out["edge_by_elem"], out["lm2"] = r.ede()
    u0 = np.zeros(neq + 1)
    A0, F0 = r.build(u0)
    u1 = np.zeros(neq + 1)
    u1[:neq] = rng.uniform(-1.0, 1.0, neq)
    A1, F1 = r.build(u1)
    out.update(A_u0=A0, F_u0=F0, u1=u1, A_u1=A1, F_u1=F1)
    r.set_matrix(A0)
    P = np.zeros(neq + 1)
    P[:neq] = rng.standard_normal(neq)
    out["P"] = P
    out["Q"] = r.matvec(P)
    # unpreconditioned GMRES: full solve and first-cycle iterates
    r.set_problem(p)
    r.set_matrix(A0)
    Xf, itf = r.pgmres(30, 200, 1e-8, F0)
    out.update(X_not=Xf, it_not=itf)
    Xk = np.zeros((KHIST, neq + 1))
    for k in range(1, KHIST + 1):
        Xk[k - 1], _ = r.pgmres(k, 1, 1e-30, F0)
    out["Xk_not"] = Xk
    # Diag preconditioner: setup scales F; solve with the scaled F
    Fd = r.precond_setup("Diag", F0)
    D, iD = r.invdiag()
    out.update(Diag=D, invDiag=iD, F_diag=Fd)
    Xd, itd = r.pgmres(30, 200, 1e-8, Fd)
    out.update(X_diag=Xd, it_diag=itd)
    if scheme == "EBE":
        r.set_problem(p)
        r.set_matrix(A0)
        Fj = r.precond_setup("Jacobi", F0)
        # apply out of place only: Jacobi_precond_EBE.c:19-20 zeroes z before
        # reading p, so pgmres's in-place call (pgmres.c:104) yields 0 and
        # the reference solve returns NaN after one step.
        out.update(invDe=r.invde(), F_jac=Fj, zjac=r.precond_apply(P))
    return out


def main():
    if not O.Reference.available():
        O.build()
    if not O.Reference.available():
        raise SystemExit("oracle/_ref not built: the FEM_CODES tree is required")
    for n in (8, 16):
        for pname in PROBLEMS:
            for scheme in ("EBE", "EDE", "CSR"):
                d = one_case(n, pname, scheme)
                fn = os.path.join(HERE, f"ss2d_n{n}_{pname}_{scheme.lower()}.npz")
                np.savez_compressed(fn, **d)
                print(fn, os.path.getsize(fn))


if __name__ == "__main__":
    main()
