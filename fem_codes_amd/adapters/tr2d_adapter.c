/*
 * tr2d_adapter.c -- drop-in GPU objects for the transient 2D transport
 * driver (02_Transport_Equation/2D).
 *
 * Compiled with the driver's own flags and header (-DTranspEquation2D,
 * 02_Transport_Equation/2D/01_CommonFiles/TranspEquation.h), this file
 * exports the reference symbols the driver binds in
 * setMatrixVectorProductType.c:6-17, setPreconditioner.c:6-28,
 * setScaling.c:6-28 and setSolver.c:8:
 *
 *   ebemv / edemv / csrmv (P*, M*, FS*, double *P, double *Q)
 *        MatrixVector_Operations/EBE/ebemv.c:4-43, EDE/edemv.c:4-39,
 *        CSR/csrmv.c:3-24
 *   Diag_precond_EBE_setup / _EDE_setup / _CSR_setup (P*, M*, FS*, int tag, double *F)
 *        Solvers_and_Preconditioners/EBE/Diag_precond_EBE_setup.c:3-35 and
 *        its EDE / CSR forms
 *   NO_scaling, Left_scaling_EBE/EDE/CSR, LeftRight_scaling_EBE/EDE/CSR,
 *   NO_unscaling, Left_unscaling
 *        Solvers_and_Preconditioners/scaling.c:4-180 (the whole object)
 *   pgmres (P*, M*, FS*, FF*, double *B, double *X)
 *        the 6-argument solver that Predictor_Old.c:111 calls and
 *        TranspEquation.h:133 types (FemOtherFunctions->solver = pgmres,
 *        setSolver.c:8); the reference tree has no implementation of it
 *
 * so that the driver linked with this object (instead of ebemv.o, edemv.o,
 * csrmv.o, pgmres.o, the three Diag setups and scaling.o) and libfemgpu.so
 * keeps its products, its diagonal setup and scaling, and its whole GMRES
 * solve on the GPU.  The driver's Build, Predictor_Old, criteria, NO_precond,
 * Diag_precond and NO_precond_setup stay the reference's.
 *
 * Semantics kept:
 *  - the products overwrite Q and write P[neq] = Q[neq] = 0 for the element
 *    and edge forms (ebemv.c:21, edemv.c:20);
 *  - the setups fill MatrixData->Diag / invDiag (caller-allocated, as in
 *    Preprocess) with the reference's bits and left-precondition F in place;
 *  - the scalings run the Diag setup on FemStructs->F and scale the stored
 *    values (on the device; the host matrix keeps its unscaled values, see
 *    fga_core.h) with invDiag[neq] = invDiag[neq-1] and, for LeftRight,
 *    Diag[neq] = Diag[neq-1] and F[i] *= sqrt(Diag[i]), as scaling.c does;
 *    Left_unscaling multiplies x[i] by sqrt(invDiag[i]) on the host;
 *  - pgmres(B, X): solves A X = B with B already left-preconditioned; reads
 *    KrylovBasisVectorsQuantity (restart length), LinearMaxIter (cycle cap)
 *    and SolverTolerance; adds its Arnoldi step count to
 *    Parameters->iterations (Block_scaling.c:20 of the Euler tree passes
 *    iterations + 1 as the next setup tag, so the count has to grow);
 *    left preconditioners NO_precond and Diag_precond (the device diagonal
 *    of the last setup, or MatrixData->invDiag uploaded when the setup was
 *    not this adapter's); precondR must be NO_precond.
 * Anything else is reported and returns 1: there is no CPU fallback.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include "TranspEquation.h"
#include "fga_core.h"

int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Diag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ebemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int edemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int csrmv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);

enum { SCH_EBE = 0, SCH_EDE = 1, SCH_CSR = 2 };

static struct {
    int scheme, neq, nitems;
    const void *key;          /* lm / lm2 / JA host pointer of the cached structure */
    fg_op *op;
    fg_prec *diag;            /* FG_PREC_DIAG of op                                 */
    int diag_ready;           /* diag holds the setup of the current build         */
} g;

static int scheme_of_name(const char *s)
{
    if (strcasecmp(s, "EBE") == 0) return SCH_EBE;
    if (strcasecmp(s, "EDE") == 0) return SCH_EDE;
    if (strcasecmp(s, "CSR") == 0) return SCH_CSR;
    return -1;
}

static void drop_op(void)
{
    if (g.diag) fg_prec_destroy(g.diag);
    if (g.op) fg_op_destroy(g.op);
    memset(&g, 0, sizeof g);
    fga_new_operator();
}

/* the operator of `scheme`: structure cached by its host pointers */
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
    g.diag_ready = 0;
    fga_uploaded();
    return 0;
}

/* operator and values for a setup, scaling or solver entry */
static int ensure_build(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
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

/* ---- the diagonal: setups and scalings ---------------------------------- */
static int diag_to_host(ParametersType *Pm, MatrixDataType *M)
{
    if (!M->Diag && !M->invDiag) return 0;
    int n = Pm->neq + 1;
    double *D = M->Diag, *iD = M->invDiag;
    double *tmpD = NULL, *tmpI = NULL;
    if (!D) D = tmpD = (double *)calloc(n, sizeof(double));
    if (!iD) iD = tmpI = (double *)calloc(n, sizeof(double));
    int rc = fg_prec_get_diag(g.diag, D, iD);
    free(tmpD);
    free(tmpI);
    return rc != FG_OK ? fga_fail("Diag") : 0;
}

/* Diag_precond_*_setup: Diag = sum of the stored diagonal entries,
 * invDiag = 1/Diag (invDiag[neq] = 1), F *= invDiag */
static int diag_setup(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                      double *F)
{
    if (ensure_build(scheme, Pm, M, FS)) return 1;
    if (!g.diag && fg_prec_create(g.op, FG_PREC_DIAG, &g.diag) != FG_OK)
        return fga_fail("Diag");
    if (fga_f_roundtrip(Pm->neq, F)) return 1;
    if (fg_prec_setup(g.diag, fga_dv[0]) != FG_OK) return fga_fail("Diag setup");
    if (fga_f_back(Pm->neq, F)) return 1;
    g.diag_ready = 1;
    return diag_to_host(Pm, M);
}

int Diag_precond_EBE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, int tag,
                           double *F)
{
    (void)tag;
    return diag_setup(SCH_EBE, Pm, M, FS, F);
}

int Diag_precond_EDE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, int tag,
                           double *F)
{
    (void)tag;
    return diag_setup(SCH_EDE, Pm, M, FS, F);
}

int Diag_precond_CSR_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, int tag,
                           double *F)
{
    (void)tag;
    return diag_setup(SCH_CSR, Pm, M, FS, F);
}

static int scaling(int scheme, int mode, ParametersType *Pm, MatrixDataType *M,
                   FemStructsType *FS)
{
    if (ensure_build(scheme, Pm, M, FS)) return 1;
    if (!g.diag && fg_prec_create(g.op, FG_PREC_DIAG, &g.diag) != FG_OK)
        return fga_fail("Diag");
    if (fga_f_roundtrip(Pm->neq, FS->F)) return 1;
    if (fg_scaling_diag(g.diag, mode, fga_dv[0]) != FG_OK) return fga_fail("scaling");
    if (fga_f_back(Pm->neq, FS->F)) return 1;
    /* the scaled system is solved without a preconditioner (setPreconditioner.c
     * rejects Diag with a scaling): the diagonal is not a setup of this build */
    g.diag_ready = 0;
    return diag_to_host(Pm, M);
}

int NO_scaling(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    (void)Pm; (void)M; (void)FS;
    return 0;
}

int Left_scaling_EBE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_EBE, FG_SCALING_LEFT, Pm, M, FS);
}

int Left_scaling_EDE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_EDE, FG_SCALING_LEFT, Pm, M, FS);
}

int Left_scaling_CSR(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_CSR, FG_SCALING_LEFT, Pm, M, FS);
}

int LeftRight_scaling_EBE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_EBE, FG_SCALING_LEFTRIGHT, Pm, M, FS);
}

int LeftRight_scaling_EDE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_EDE, FG_SCALING_LEFTRIGHT, Pm, M, FS);
}

int LeftRight_scaling_CSR(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    return scaling(SCH_CSR, FG_SCALING_LEFTRIGHT, Pm, M, FS);
}

int NO_unscaling(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *x)
{
    (void)Pm; (void)M; (void)FS; (void)x;
    return 0;
}

/* scaling.c:170-180 on the host solution: x[i] *= sqrt(invDiag[i]) */
int Left_unscaling(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *x)
{
    (void)FS;
    int neq = Pm->neq;
    const double *invDiag = M->invDiag;
    for (int i = 0; i < neq; i++) x[i] *= sqrt(invDiag[i]);
    return 0;
}

/* ---- the 6-argument solver ----------------------------------------------- */
int pgmres(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, FemFunctionsType *FF,
           double *B, double *X)
{
    int scheme;
    if (FF->mv == ebemv) scheme = SCH_EBE;
    else if (FF->mv == edemv) scheme = SCH_EDE;
    else if (FF->mv == csrmv) scheme = SCH_CSR;
    else {
        printf("femgpu pgmres: unsupported matrix-vector product\n");
        return 1;
    }
    if (scheme_of_name(Pm->MatrixVectorProductScheme) >= 0 &&
        scheme_of_name(Pm->MatrixVectorProductScheme) != scheme) {
        printf("femgpu pgmres: MatrixVectorProductScheme disagrees with FemFunctions->mv\n");
        return 1;
    }
    if (FF->precondR != NO_precond) {
        printf("femgpu pgmres: only NO_precond is supported as right preconditioner\n");
        return 1;
    }
    if (FF->precond != NO_precond && FF->precond != Diag_precond) {
        printf("femgpu pgmres: unsupported left preconditioner\n");
        return 1;
    }
    if (ensure_build(scheme, Pm, M, FS)) return 1;
    fg_prec *pc = NULL;
    if (FF->precond == Diag_precond) {
        if (!g.diag && fg_prec_create(g.op, FG_PREC_DIAG, &g.diag) != FG_OK)
            return fga_fail("Diag");
        /* the diagonal of a host setup: take MatrixData->invDiag */
        if (!g.diag_ready && fg_prec_set_invdiag(g.diag, M->invDiag) != FG_OK)
            return fga_fail("invDiag");
        pc = g.diag;
    }
    fg_gmres_opts o;
    fga_opts(&o, Pm->KrylovBasisVectorsQuantity, Pm->LinearMaxIter, Pm->SolverTolerance);
    fg_gmres_info info;
    if (fga_solve(g.op, pc, Pm->neq, &o, B, X, &info)) return 1;
    g.diag_ready = 0;
    Pm->iterations += info.steps;
    return 0;
}
