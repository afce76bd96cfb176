/*
 * ss3d_adapter.c -- drop-in GPU objects for the steady 3D transport driver.
 *
 * Compiled with the driver's own flags and header (-DSSTransportEquation3D,
 * 01_SS_Transport_Equation/3D/01_CommonFiles/SSTransportEquation3D.h), this
 * file exports the reference symbols
 *
 *   int ebemv3Dndof1(ParametersType*, MatrixDataType*, FemStructsType*, double *P, double *Q)
 *                                          (EBE/ebemv3Dndof1.c:3-45)
 *   int csrmv  (...)                       (CSR/csrmv.c:3-24)
 *   int gmres  (ParametersType*, MatrixDataType*, FemStructsType*, FemFunctionsType*)
 *                                          (Solvers_and_Preconditioners/gmres.c:3-191)
 *   int pgmres (...)                       (Solvers_and_Preconditioners/pgmres.c:4-211)
 *
 * which the 3D driver binds in setMatrixVectorProductType.c:10-22 and
 * setSolver.c:4-17 (gmres for Preconditioner "NOT", pgmres otherwise).
 *
 * Semantics kept: Q overwritten, P[neq] = Q[neq] = 0 for the element
 * product; gmres and pgmres read KrylovBasisVectorsQuantity, SolverMaxIter
 * and SolverTolerance, solve A u = F with FemStructs->u zeroed first, and
 * add the Arnoldi step count to SolverIterations; gmres also stores the
 * final residual estimate in ResGMRES and its step count in ContGMRES
 * (gmres.c:177-178).  gmres.c runs no preconditioner; its pow(h, 2) equals
 * h*h exactly, so the unpreconditioned device GMRES is its drop-in.
 * pgmres supports the left preconditioners NO_precond and ILUp_precond
 * with Preconditioner "ILU0" (the CSR scheme, setPreconditioner.c:12-20):
 * the driver's ILUp_precond_setup has already left-preconditioned F on the
 * host; the device factorises the same CSR values (bit for bit the factors
 * of ILUp_precond_setup.c) and applies them inside the solve.  Anything else
 * is reported and returns 1: there is no CPU fallback.  The element and CSR
 * values are uploaded by the first solver call after the previous one (the
 * driver rebuilds them per nonlinear iteration) and by a product called
 * before any upload (Compute_Residual.c:18 after a solve reuses them; the
 * rule is in fga_core.h); the structure is uploaded once, keyed by its host
 * pointers.
 */
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include "SSTransportEquation3D.h"
#include "fga_core.h"

int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ILUp_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ebemv3Dndof1(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int csrmv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);

enum { SCH_EBE = 0, SCH_CSR = 2 };

static struct {
    int scheme, neq, nitems;
    const void *key;
    fg_op *op;
    fg_prec *ilu;
} g;

static void drop_op(void)
{
    if (g.ilu) fg_prec_destroy(g.ilu);
    if (g.op) fg_op_destroy(g.op);
    memset(&g, 0, sizeof g);
    fga_new_operator();
}

static int ensure_structure(int scheme, ParametersType *Pm, MatrixDataType *M,
                            FemStructsType *FS)
{
    if (fga_ensure_ctx()) return 1;
    int neq = Pm->neq;
    const void *key = scheme == SCH_EBE ? (const void *)FS->lmaux : (const void *)M->JA;
    int nitems = scheme == SCH_EBE ? Pm->nel : Pm->nnzero;
    if (g.op && g.scheme == scheme && g.neq == neq && g.key == key && g.nitems == nitems)
        return 0;
    drop_op();
    int rc = scheme == SCH_EBE
                 ? fg_op_ebe_create(fga_ctx, neq, Pm->nel, NNOEL * NDOF, FS->lmaux, 0, &g.op)
                 : fg_op_csr_create(fga_ctx, neq, Pm->nnzero, M->IA, M->JA, &g.op);
    if (rc != FG_OK) return fga_fail("operator");
    g.scheme = scheme; g.neq = neq; g.key = key; g.nitems = nitems;
    return 0;
}

static int upload_values(MatrixDataType *M)
{
    const double *vals = g.scheme == SCH_CSR ? M->AA : M->Aaux;
    if (fg_op_set_values_host(g.op, vals) != FG_OK) return fga_fail("values");
    fga_uploaded();
    return 0;
}

/* a solver entry: operator and the values of the current build */
static int ensure_op(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    if (fga_needs_build_upload()) return upload_values(M);
    return 0;
}

static int product(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                   double *P, double *Q)
{
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    if (fga_needs_product_upload() && upload_values(M)) return 1;
    if (fga_product(g.op, Pm->neq, P, Q)) return 1;
    if (scheme == SCH_EBE) P[Pm->neq] = 0.0;   /* ebemv3Dndof1.c mutates its input */
    return 0;
}

int ebemv3Dndof1(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EBE, Pm, M, FS, P, Q);
}

int csrmv(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_CSR, Pm, M, FS, P, Q);
}

static int scheme_of(FemFunctionsType *FF)
{
    if (FF->ProductMatrixVector == ebemv3Dndof1) return SCH_EBE;
    if (FF->ProductMatrixVector == csrmv) return SCH_CSR;
    printf("femgpu: unsupported matrix-vector product\n");
    return -1;
}

/* the device solve of A u = F; pc NULL or the ILU(0) of the CSR operator */
static int solve(ParametersType *Pm, FemStructsType *FS, fg_prec *pc, fg_gmres_info *info)
{
    fg_gmres_opts o;
    fga_opts(&o, Pm->KrylovBasisVectorsQuantity, Pm->SolverMaxIter, Pm->SolverTolerance);
    return fga_solve(g.op, pc, Pm->neq, &o, FS->F, FS->u, info);
}

int gmres(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, FemFunctionsType *FF)
{
    int scheme = scheme_of(FF);
    if (scheme < 0 || ensure_op(scheme, Pm, M, FS)) return 1;
    fg_gmres_info info;
    if (solve(Pm, FS, NULL, &info)) return 1;
    Pm->ResGMRES = info.rho;
    Pm->ContGMRES = info.steps;
    Pm->SolverIterations += info.steps;
    return 0;
}

int pgmres(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, FemFunctionsType *FF)
{
    int scheme = scheme_of(FF);
    if (scheme < 0) return 1;
    if (FF->precondR != NO_precond) {
        printf("femgpu pgmres: only NO_precond is supported as right preconditioner\n");
        return 1;
    }
    int ilu = FF->precond == ILUp_precond;
    if (!ilu && FF->precond != NO_precond) {
        printf("femgpu pgmres: unsupported left preconditioner\n");
        return 1;
    }
    if (ilu && (scheme != SCH_CSR || strcasecmp(Pm->Preconditioner, "ILU0") != 0)) {
        printf("femgpu pgmres: ILU(p) is supported for p = 0 on the CSR scheme\n");
        return 1;
    }
    if (ensure_op(scheme, Pm, M, FS)) return 1;
    fg_prec *pc = NULL;
    if (ilu) {
        if (!g.ilu && fg_prec_create(g.op, FG_PREC_ILU0, &g.ilu) != FG_OK) return fga_fail("ILU(0)");
        if (fg_prec_setup(g.ilu, NULL) != FG_OK) return fga_fail("ILU(0) factor");
        pc = g.ilu;
    }
    fg_gmres_info info;
    if (solve(Pm, FS, pc, &info)) return 1;
    Pm->SolverIterations += info.steps;
    return 0;
}
