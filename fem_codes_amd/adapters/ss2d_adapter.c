/*
 * ss2d_adapter.c -- drop-in GPU objects for the steady 2D transport driver.
 *
 * Compiled with the driver's own flags and header (-DSSTranspEquation2D,
 * 01_SS_Transport_Equation/2D/01_CommonFiles/SSTranspEquation.h), this file
 * exports the reference symbols
 *
 *   int ebemv (ParametersType*, MatrixDataType*, FemStructsType*, double *P, double *Q)
 *   int edemv (...)                        (EDE/edemv.c:4-39)
 *   int csrmv (...)                        (CSR/csrmv.c:3-24)
 *   int pgmres(ParametersType*, MatrixDataType*, FemStructsType*, FemFunctionsType*)
 *                                          (Solvers_and_Preconditioners/pgmres.c:4-211)
 *
 * so that a driver linked with this object (instead of ebemv.o, edemv.o,
 * csrmv.o, pgmres.o) and libfemgpu.so runs its products and its whole GMRES
 * solve on the GPU.  Everything else of the driver (Build_K_F_SUPG,
 * Diag_precond_*_setup, NO_precond, I/O) is unchanged.
 *
 * Semantics kept: Q is overwritten and P[neq] = Q[neq] = 0; pgmres reads
 * KrylovBasisVectorsQuantity, SolverMaxIter and SolverTolerance, solves
 * A u = F with FemStructs->u zeroed first, and adds its Arnoldi step count
 * to Parameters->SolverIterations.  The driver rebuilds the element / edge /
 * CSR values in every Picard iteration (Process.c:296-322) and then calls
 * the solver, so the values are uploaded by the first pgmres call after the
 * previous one and by a product called before any upload, not by every
 * product (the rule is in fga_core.h); the structure (lm, lm2, IA/JA) is
 * uploaded once and cached, keyed by its host pointers.
 * Supported left preconditioners inside pgmres: NO_precond, Diag_precond
 * (MatrixData->invDiag as computed by the driver's setup) and
 * Jacobi_precond_EBE (EBE scheme: the device element Jacobi of the same
 * values, applied correctly in place where the reference's in-place call
 * returns 0); precondR must be NO_precond.  Anything else is reported and
 * returns 1: there is no CPU fallback.  Errors print the library message and
 * return 1 (the reference exits; its callers ignore the return value).
 */
#include <stdio.h>
#include <string.h>

#include "SSTranspEquation.h"
#include "fga_core.h"

int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Diag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
/* weak: the SS-2D driver's own table never selects it (setPreconditioner.c
 * of that driver), so it is linked only where a caller provides it */
int Jacobi_precond_EBE(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *)
    __attribute__((weak));
int ebemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int edemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int csrmv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);

enum { SCH_EBE = 0, SCH_EDE = 1, SCH_CSR = 2 };

static struct {
    int scheme, neq, nitems;
    const void *key;          /* lm / lm2 / JA host pointer of the cached structure */
    fg_op *op;
    fg_prec *pc, *jac;
} g;

static void drop_op(void)
{
    if (g.pc) fg_prec_destroy(g.pc);
    if (g.jac) fg_prec_destroy(g.jac);
    if (g.op) fg_op_destroy(g.op);
    memset(&g, 0, sizeof g);
    fga_new_operator();
}

/* the operator of the requested scheme, structure cached by its host pointers */
static int ensure_structure(int scheme, ParametersType *Pm, MatrixDataType *M,
                            FemStructsType *FS)
{
    if (fga_ensure_ctx()) return 1;
    int neq = Pm->neq;
    const void *key = scheme == SCH_EBE ? (const void *)FS->lmaux
                    : scheme == SCH_EDE ? (const void *)FS->lm2aux : (const void *)M->JA;
    int nitems = scheme == SCH_EBE ? Pm->nel : scheme == SCH_EDE ? Pm->nedge : Pm->nnzero;
    if (g.op && g.scheme == scheme && g.neq == neq && g.key == key && g.nitems == nitems)
        return 0;
    drop_op();
    int rc = FG_OK;
    if (scheme == SCH_EBE)
        rc = fg_op_ebe_create(fga_ctx, neq, Pm->nel, NNOEL * NDOF, FS->lmaux, 0, &g.op);
    else if (scheme == SCH_EDE)
        rc = fg_op_ede_create(fga_ctx, neq, Pm->nedge, FS->lm2aux, 0, &g.op);
    else
        rc = fg_op_csr_create(fga_ctx, neq, Pm->nnzero, M->IA, M->JA, &g.op);
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

static int product(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                   double *P, double *Q)
{
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    if (fga_needs_product_upload() && upload_values(M)) return 1;
    if (fga_product(g.op, Pm->neq, P, Q)) return 1;
    if (scheme != SCH_CSR) P[Pm->neq] = 0.0;   /* ebemv.c:21 mutates its input */
    return 0;
}

int ebemv(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EBE, Pm, M, FS, P, Q);
}

int edemv(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EDE, Pm, M, FS, P, Q);
}

int csrmv(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_CSR, Pm, M, FS, P, Q);
}

int pgmres(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, FemFunctionsType *FF)
{
    int scheme;
    if (FF->ProductMatrixVector == ebemv) scheme = SCH_EBE;
    else if (FF->ProductMatrixVector == edemv) scheme = SCH_EDE;
    else if (FF->ProductMatrixVector == csrmv) scheme = SCH_CSR;
    else { printf("femgpu pgmres: unsupported matrix-vector product\n"); return 1; }
    if (FF->precondR != NO_precond) {
        printf("femgpu pgmres: only NO_precond is supported as right preconditioner\n");
        return 1;
    }
    const int jac = Jacobi_precond_EBE && FF->precond == Jacobi_precond_EBE;
    if (FF->precond != NO_precond && FF->precond != Diag_precond && !jac) {
        printf("femgpu pgmres: unsupported left preconditioner\n");
        return 1;
    }
    if (jac && scheme != SCH_EBE) {
        printf("femgpu pgmres: Jacobi_precond_EBE needs the EBE scheme\n");
        return 1;
    }
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    /* the driver rebuilt the values since the last solve (Process.c:296-322) */
    if (fga_needs_build_upload() && upload_values(M)) return 1;
    fg_prec *pc = NULL;
    if (FF->precond == Diag_precond) {
        if (!g.pc && fg_prec_create(g.op, FG_PREC_DIAG, &g.pc) != FG_OK) return fga_fail("prec");
        if (fg_prec_set_invdiag(g.pc, M->invDiag) != FG_OK) return fga_fail("invDiag");
        pc = g.pc;
    } else if (jac) {
        /* the element Jacobi of the uploaded values (the bits of the driver's
         * Jacobi_precond_EBE_setup, which has already preconditioned F) */
        if (!g.jac && fg_prec_create(g.op, FG_PREC_JACOBI, &g.jac) != FG_OK)
            return fga_fail("Jacobi");
        if (fg_prec_setup(g.jac, NULL) != FG_OK) return fga_fail("Jacobi setup");
        pc = g.jac;
    }
    fg_gmres_opts o;
    fga_opts(&o, Pm->KrylovBasisVectorsQuantity, Pm->SolverMaxIter, Pm->SolverTolerance);
    fg_gmres_info info;
    memset(FS->u, 0, sizeof(double) * (size_t)(Pm->neq + 1));
    if (fga_solve(g.op, pc, Pm->neq, &o, FS->F, FS->u, &info)) return 1;
    FS->u[Pm->neq] = 0.0;
    Pm->SolverIterations += info.steps;
    return 0;
}
