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
 * to Parameters->SolverIterations.  The element/edge/CSR values are
 * re-uploaded on every call because the driver rebuilds them in every
 * Picard iteration (Process.c:296-322); the structure (lm, lm2, IA/JA) is
 * uploaded once and cached, keyed by its host pointers.
 * Supported left preconditioners inside pgmres: NO_precond and Diag_precond
 * (MatrixData->invDiag as computed by the driver's setup); precondR must be
 * NO_precond.  Anything else is reported and returns 1: there is no CPU
 * fallback.  Errors print the library message and return 1 (the reference
 * exits; its callers ignore the return value).
 */
#include <stdio.h>
#include <string.h>

#include "SSTranspEquation.h"
#include "femgpu.h"

int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Diag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ebemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int edemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int csrmv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);

enum { SCH_EBE = 0, SCH_EDE = 1, SCH_CSR = 2 };

static fg_ctx *g_ctx;
static struct {
    int scheme, neq, nitems;
    const void *key;          /* lm / lm2 / JA host pointer of the cached structure */
    fg_op *op;
    fg_prec *pc;
    double *dP, *dQ;          /* neq+1 device vectors */
} g;

static int fail(const char *where)
{
    printf("femgpu %s: %s\n", where, fg_last_error());
    return 1;
}

static int ensure_ctx(void)
{
    if (g_ctx) return 0;
    if (fg_ctx_create(0, &g_ctx) != FG_OK) return fail("context");
    return 0;
}

static void drop_op(void)
{
    if (g.pc) fg_prec_destroy(g.pc);
    if (g.op) fg_op_destroy(g.op);
    if (g.dP) fg_free(g_ctx, g.dP);
    if (g.dQ) fg_free(g_ctx, g.dQ);
    memset(&g, 0, sizeof g);
}

/* the operator of the requested scheme, structure cached, values refreshed */
static int ensure_op(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    if (ensure_ctx()) return 1;
    int neq = Pm->neq;
    const void *key = scheme == SCH_EBE ? (const void *)FS->lmaux
                    : scheme == SCH_EDE ? (const void *)FS->lm2aux : (const void *)M->JA;
    int nitems = scheme == SCH_EBE ? Pm->nel : scheme == SCH_EDE ? Pm->nedge : Pm->nnzero;
    if (!g.op || g.scheme != scheme || g.neq != neq || g.key != key || g.nitems != nitems) {
        drop_op();
        int rc = FG_OK;
        if (scheme == SCH_EBE)
            rc = fg_op_ebe_create(g_ctx, neq, Pm->nel, NNOEL * NDOF, FS->lmaux, 0, &g.op);
        else if (scheme == SCH_EDE)
            rc = fg_op_ede_create(g_ctx, neq, Pm->nedge, FS->lm2aux, 0, &g.op);
        else
            rc = fg_op_csr_create(g_ctx, neq, Pm->nnzero, M->IA, M->JA, &g.op);
        if (rc != FG_OK) return fail("operator");
        if (fg_malloc(g_ctx, sizeof(double) * (neq + 1), (void **)&g.dP) != FG_OK ||
            fg_malloc(g_ctx, sizeof(double) * (neq + 1), (void **)&g.dQ) != FG_OK)
            return fail("vectors");
        g.scheme = scheme; g.neq = neq; g.key = key; g.nitems = nitems;
    }
    const double *vals = scheme == SCH_CSR ? M->AA : M->Aaux;
    if (fg_op_set_values_host(g.op, vals) != FG_OK) return fail("values");
    return 0;
}

static int product(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                   double *P, double *Q)
{
    if (ensure_op(scheme, Pm, M, FS)) return 1;
    size_t bytes = sizeof(double) * (Pm->neq + 1);
    if (fg_memcpy_h2d(g_ctx, g.dP, P, bytes) || fg_op_apply_f64(g.op, g.dP, g.dQ) ||
        fg_memcpy_d2h(g_ctx, Q, g.dQ, bytes))
        return fail("matvec");
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
    if (FF->precond != NO_precond && FF->precond != Diag_precond) {
        printf("femgpu pgmres: unsupported left preconditioner\n");
        return 1;
    }
    if (ensure_op(scheme, Pm, M, FS)) return 1;
    fg_prec *pc = NULL;
    if (FF->precond == Diag_precond) {
        if (!g.pc && fg_prec_create(g.op, FG_PREC_DIAG, &g.pc) != FG_OK) return fail("prec");
        if (fg_prec_set_invdiag(g.pc, M->invDiag) != FG_OK) return fail("invDiag");
        pc = g.pc;
    }
    int neq = Pm->neq;
    size_t bytes = sizeof(double) * (neq + 1);
    fg_gmres_opts o;
    memset(&o, 0, sizeof o);
    o.kmax = Pm->KrylovBasisVectorsQuantity;
    o.lmax = Pm->SolverMaxIter;
    o.tol = Pm->SolverTolerance;
    o.precision = FG_GMRES_FP64;
    fg_gmres_info info;
    if (fg_memcpy_h2d(g_ctx, g.dP, FS->F, bytes) != FG_OK) return fail("F");
    int rc = fg_gmres(g_ctx, g.op, pc, g.dP, g.dQ, &o, &info);
    if (rc != FG_OK && rc != FG_ERR_NOCONV) return fail("gmres");
    if (fg_memcpy_d2h(g_ctx, FS->u, g.dQ, sizeof(double) * neq) != FG_OK) return fail("u");
    Pm->SolverIterations += info.steps;
    return 0;
}
