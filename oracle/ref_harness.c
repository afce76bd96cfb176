/*
 * ref_harness.c -- drives the FEM_CODES SS-2D reference routines, compiled
 * from their own sources by oracle/Makefile, through a flat C interface so
 * that tests can compare the oracle restatement (oracle.c) against the real
 * reference.  TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, never linked
 * into the product.
 *
 * The harness is compiled with -DSSTranspEquation2D against the reference's
 * own header (01_SS_Transport_Equation/2D/01_CommonFiles/SSTranspEquation.h),
 * so ParametersType / MatrixDataType / FemStructsType have the reference
 * layout.  It reproduces what Preprocess.c:38-248 allocates and what
 * Process.c:276-321 dispatches, without reading parameter or mesh files.
 * Problem callbacks for the synthetic problems are defined here; VALIDACAO
 * uses the reference's own VALIDACAO_*.c.
 */
#include "SSTranspEquation.h"
#include "../../../00_CommonFiles/MatrixVector_Operations/matvec.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/preconditioners.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/solvers.h"
#include "../../../00_CommonFiles/BLAS_Operations/ourBLAS.h"
#include "../VALIDACAO/validacao.h"

int Jacobi_precond_EBE(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Jacobi_precond_EBE_setup(ParametersType *, MatrixDataType *, FemStructsType *, int, double *);

static ParametersType P_;
static MatrixDataType M_;
static FemStructsType FS_;
static FemFunctionsType FF_;

/* ---- synthetic problem coefficients (configurable from the tests) ---- */
static int pk_kind;
static double pk_kappa, pk_gamma, pk_f, pk_bx, pk_by, pk_scale;

static double syn_cond(void) { return pk_kappa; }
static double syn_reac(void) { return pk_gamma; }
static double syn_font(double X, double Y, double k, double g, double bx, double by)
{
    (void)X; (void)Y; (void)k; (void)g; (void)bx; (void)by;
    return pk_f;
}
static void syn_vel(double X, double Y, double B[2])
{
    if (pk_kind == 1) { B[0] = (-Y + 0.5) * pk_scale; B[1] = (X - 0.5) * pk_scale; }
    else { B[0] = pk_bx; B[1] = pk_by; }
}
static double syn_upresc(double X, double Y) { (void)X; (void)Y; return 0.0; }
/* F_assembly.c:32-34 calls hflux unconditionally; its result is multiplied
 * by the integer quotient 1/6 = 0 and only for Type-2 nodes. */
static double zero_hflux(double X, double Y) { (void)X; (void)Y; return 0.0; }

/* ---- setup: Preprocess.c:160-248 without the file reading ------------ */
int ref_setup(int nnodes, int nel, const double *x, const double *y, const int *type,
              const int *vertex, const char *scheme, const char *reorder)
{
    memset(&P_, 0, sizeof P_); memset(&M_, 0, sizeof M_);
    memset(&FS_, 0, sizeof FS_); memset(&FF_, 0, sizeof FF_);
    strcpy(P_.MatrixVectorProductScheme, scheme);
    strcpy(P_.reordering, reorder);
    strcpy(P_.OutputFlow, "NO");
    P_.nnodes = nnodes; P_.nel = nel;

    NodeType *Node = calloc(nnodes, sizeof(NodeType));
    int neq = 0;
    for (int i = 0; i < nnodes; i++) {
        Node[i].x = x[i]; Node[i].y = y[i]; Node[i].Type = type[i];
        Node[i].id = type[i] >= 1 ? neq++ : -1;
    }
    ElementType *El = calloc(nel, sizeof(ElementType));
    for (int e = 0; e < nel; e++) {
        El[e].Vertex[0] = vertex[3 * e]; El[e].Vertex[1] = vertex[3 * e + 1];
        El[e].Vertex[2] = vertex[3 * e + 2]; El[e].Type = -1;
    }
    int **lm = calloc(nel, sizeof(int *));
    int *lmaux = calloc((size_t)nel * 3, sizeof(int));
    for (int e = 0; e < nel; e++) lm[e] = &lmaux[3 * e];
    FS_.F = calloc(neq + 1, sizeof(double));
    FS_.u = calloc(neq + 1, sizeof(double));
    M_.Diag = calloc(neq + 1, sizeof(double));
    M_.invDiag = calloc(neq + 1, sizeof(double));
    Fill_LM(neq, nel, lm, Node, El);
    FS_.lm = lm; FS_.lmaux = lmaux;
    P_.neq = neq;

    if (strcasecmp(scheme, "EBE") == 0) {
        M_.A = calloc(nel, sizeof(double *));
        M_.Aaux = calloc((size_t)nel * 9, sizeof(double));
        for (int e = 0; e < nel; e++) M_.A[e] = &M_.Aaux[9 * e];
    } else if (strcasecmp(scheme, "EDE") == 0) {
        int **order = calloc(nel, sizeof(int *)), **EbE;
        for (int e = 0; e < nel; e++) order[e] = calloc(3, sizeof(int));
        int **lm2 = lm, *lm2aux = lmaux;
        ede_Initialization(&P_, order, &lm2, &lm2aux, &EbE);
        int nedge = P_.nedge;
        M_.A = calloc(nedge + 1, sizeof(double *));
        M_.Aaux = calloc((size_t)(nedge + 1) * 4, sizeof(double));
        for (int g = 0; g < nedge + 1; g++) M_.A[g] = &M_.Aaux[4 * g];
        FS_.lm2 = lm2; FS_.lm2aux = lm2aux;
        M_.Scheme_by_Element = EbE; M_.order = order;
    } else if (strcasecmp(scheme, "CSR") == 0) {
        int *IA, *JA, *perm, *invperm, **CbE;
        csr_Initialization(&P_, Node, &JA, &IA, &perm, &invperm, &lm, &lmaux, &CbE);
        M_.AA = calloc(P_.nnzero + 1, sizeof(double));
        M_.IA = IA; M_.JA = JA; M_.Scheme_by_Element = CbE;
    } else {
        return -1;
    }
    FS_.Node = Node; FS_.Element = El;
    /* setMatrixVectorProductType.c:6-17 (ref_set_problem sets it again) */
    if (strcasecmp(scheme, "EBE") == 0) FF_.ProductMatrixVector = ebemv;
    else if (strcasecmp(scheme, "EDE") == 0) FF_.ProductMatrixVector = edemv;
    else FF_.ProductMatrixVector = csrmv;
    FF_.precond = NO_precond; FF_.precondR = NO_precond; FF_.precond_setup = NO_precond_setup;
    return neq;
}

void ref_sizes(int *neq, int *nnz, int *nedge)
{
    *neq = P_.neq; *nnz = P_.nnzero; *nedge = P_.nedge;
}
void ref_get_lm(int *lm) { memcpy(lm, FS_.lmaux, sizeof(int) * 3 * (size_t)P_.nel); }
void ref_get_ids(int *id) { for (int i = 0; i < P_.nnodes; i++) id[i] = FS_.Node[i].id; }
void ref_get_csr(int *IA, int *JA, int *csr_by_elem)
{
    memcpy(IA, M_.IA, sizeof(int) * (P_.neq + 1));
    memcpy(JA, M_.JA, sizeof(int) * P_.nnzero);
    for (int e = 0; e < P_.nel; e++) memcpy(csr_by_elem + 9 * e, M_.Scheme_by_Element[e], 9 * sizeof(int));
}
void ref_get_ede(int *order, int *edge_by_elem, int *lm2)
{
    for (int e = 0; e < P_.nel; e++) {
        memcpy(order + 3 * e, M_.order[e], 3 * sizeof(int));
        memcpy(edge_by_elem + 3 * e, M_.Scheme_by_Element[e], 3 * sizeof(int));
    }
    memcpy(lm2, FS_.lm2aux, sizeof(int) * 2 * (size_t)P_.nedge);
}

/* ---- problem and dispatch: setProblem.c, setMatrixVectorProductType.c,
 *      setPreconditioner.c, setStabilizationForm.c ----------------------- */
void ref_set_problem(int kind, double kappa, double gamma, double f, double bx, double by,
                     double scale, int hshock)
{
    pk_kind = kind; pk_kappa = kappa; pk_gamma = gamma; pk_f = f;
    pk_bx = bx; pk_by = by; pk_scale = scale;
    if (kind == 2) {
        FF_.Condutivity = VALIDACAO_Condutivity; FF_.Font = VALIDACAO_Font;
        FF_.Reaction = VALIDACAO_Reaction; FF_.Velocity = VALIDACAO_Velocity;
        FF_.upresc = VALIDACAO_upresc;
    } else {
        FF_.Condutivity = syn_cond; FF_.Font = syn_font; FF_.Reaction = syn_reac;
        FF_.Velocity = syn_vel; FF_.upresc = syn_upresc;
    }
    FF_.hflux = zero_hflux;
    FF_.ShockCapture = CAU_ShockCapture;
    FF_.h_shock = hshock == 1 ? h_shock_sqrtArea : h_shock_2sqrtArea;
    if (strcasecmp(P_.MatrixVectorProductScheme, "EBE") == 0) {
        FF_.assembly = ebe_assembly; FF_.ProductMatrixVector = ebemv;
    } else if (strcasecmp(P_.MatrixVectorProductScheme, "EDE") == 0) {
        FF_.assembly = ede_assembly; FF_.ProductMatrixVector = edemv;
    } else {
        FF_.assembly = csr_assembly; FF_.ProductMatrixVector = csrmv;
    }
    FF_.precond = NO_precond; FF_.precondR = NO_precond; FF_.precond_setup = NO_precond_setup;
}

/* Build_K_F_SUPG with the current iterate u (neq+1 values) */
void ref_build(const double *u)
{
    memcpy(FS_.u, u, sizeof(double) * (P_.neq + 1));
    Build_K_F_SUPG(&P_, &M_, &FS_, &FF_);
}
void ref_get_F(double *F) { memcpy(F, FS_.F, sizeof(double) * (P_.neq + 1)); }
void ref_set_F(const double *F) { memcpy(FS_.F, F, sizeof(double) * (P_.neq + 1)); }
void ref_get_u(double *u) { memcpy(u, FS_.u, sizeof(double) * (P_.neq + 1)); }
/* EBE: nel*9 ; EDE: (nedge+1)*4 ; CSR: nnz+1 */
void ref_get_matrix(double *A)
{
    if (strcasecmp(P_.MatrixVectorProductScheme, "EBE") == 0)
        memcpy(A, M_.Aaux, sizeof(double) * 9 * (size_t)P_.nel);
    else if (strcasecmp(P_.MatrixVectorProductScheme, "EDE") == 0)
        memcpy(A, M_.Aaux, sizeof(double) * 4 * (size_t)(P_.nedge + 1));
    else
        memcpy(A, M_.AA, sizeof(double) * (size_t)(P_.nnzero + 1));
}
void ref_set_matrix(const double *A)
{
    if (strcasecmp(P_.MatrixVectorProductScheme, "EBE") == 0)
        memcpy(M_.Aaux, A, sizeof(double) * 9 * (size_t)P_.nel);
    else if (strcasecmp(P_.MatrixVectorProductScheme, "EDE") == 0)
        memcpy(M_.Aaux, A, sizeof(double) * 4 * (size_t)(P_.nedge + 1));
    else
        memcpy(M_.AA, A, sizeof(double) * (size_t)(P_.nnzero + 1));
}

/* matvec through the dispatch pointer; P is mutated (P[neq]=0) like the reference */
void ref_matvec(double *Pv, double *Q) { FF_.ProductMatrixVector(&P_, &M_, &FS_, Pv, Q); }

/* preconditioner selection and setup; name "NOT", "Diag" or "Jacobi" */
int ref_precond_setup(const char *name, double *F)
{
    if (strcasecmp(name, "Diag") == 0) {
        FF_.precond = Diag_precond; FF_.precondR = NO_precond;
        if (strcasecmp(P_.MatrixVectorProductScheme, "EBE") == 0)
            FF_.precond_setup = Diag_precond_EBE_setup;
        else if (strcasecmp(P_.MatrixVectorProductScheme, "EDE") == 0)
            FF_.precond_setup = Diag_precond_EDE_setup;
        else
            FF_.precond_setup = Diag_precond_CSR_setup;
    } else if (strcasecmp(name, "Jacobi") == 0) {
        FF_.precond = Jacobi_precond_EBE; FF_.precondR = NO_precond;
        FF_.precond_setup = Jacobi_precond_EBE_setup;
    } else {
        FF_.precond = NO_precond; FF_.precondR = NO_precond;
        FF_.precond_setup = NO_precond_setup;
    }
    return FF_.precond_setup(&P_, &M_, &FS_, 1, F);
}
void ref_get_invdiag(double *Diag, double *invDiag)
{
    memcpy(Diag, M_.Diag, sizeof(double) * (P_.neq + 1));
    memcpy(invDiag, M_.invDiag, sizeof(double) * (P_.neq + 1));
}
void ref_get_invde(double *invDe) { memcpy(invDe, M_.invDeaux, sizeof(double) * 3 * (size_t)P_.nel); }
/* apply the selected left preconditioner; Jacobi_precond_EBE zeroes z first,
 * so p and z are distinct here */
void ref_precond_apply(double *p, double *z) { FF_.precond(&P_, &M_, &FS_, p, z); }

/* pgmres.c:4-211 on F (B) -> u (X) */
int ref_pgmres(int kmax, int lmax, double tol, const double *B, double *X)
{
    P_.KrylovBasisVectorsQuantity = kmax;
    P_.SolverMaxIter = lmax;
    P_.SolverTolerance = tol;
    P_.SolverIterations = 0;
    memcpy(FS_.F, B, sizeof(double) * (P_.neq + 1));
    pgmres(&P_, &M_, &FS_, &FF_);
    memcpy(X, FS_.u, sizeof(double) * (P_.neq + 1));
    return P_.SolverIterations;
}

/* BLAS-1 passthroughs */
double ref_ddot(int n, double *x, double *y) { return ddot(n, x, y); }

/* Reordering/symrcm.c REORDERING_SYMRCM on a caller CSR pattern: perm
 * (neq) receives P, IA/JA are permuted in place (MATRIX_ROW/COL_permutation,
 * reordering.c:238-308) and pT (nnz) receives the slot permutation */
void REORDERING_SYMRCM(ParametersType *, int *, int *, int *, int *);
void ref_symrcm(int neq, int nnz, int *IA, int *JA, int *perm, int *pT)
{
    ParametersType P;
    memset(&P, 0, sizeof P);
    P.neq = neq;
    P.nnzero = nnz;
    for (int i = 0; i < neq; i++) perm[i] = i;
    for (int q = 0; q < nnz; q++) pT[q] = q;
    REORDERING_SYMRCM(&P, JA, IA, perm, pT);
}

/* scaling.c:10-180 on the current matrix and F (after ref_set_matrix /
 * ref_set_F): name "Left" or "LeftRight"; the scaled matrix is read back
 * with ref_get_matrix, F with ref_get_F, Diag/invDiag with ref_get_invdiag */
int Left_scaling_EBE(ParametersType *, MatrixDataType *, FemStructsType *);
int Left_scaling_EDE(ParametersType *, MatrixDataType *, FemStructsType *);
int Left_scaling_CSR(ParametersType *, MatrixDataType *, FemStructsType *);
int LeftRight_scaling_EBE(ParametersType *, MatrixDataType *, FemStructsType *);
int LeftRight_scaling_EDE(ParametersType *, MatrixDataType *, FemStructsType *);
int LeftRight_scaling_CSR(ParametersType *, MatrixDataType *, FemStructsType *);
int Left_unscaling(ParametersType *, MatrixDataType *, FemStructsType *, double *);

void ref_scaling(const char *name)
{
    const char *s = P_.MatrixVectorProductScheme;
    const int lr = strcasecmp(name, "LeftRight") == 0;
    if (strcasecmp(s, "EBE") == 0) lr ? LeftRight_scaling_EBE(&P_, &M_, &FS_) : Left_scaling_EBE(&P_, &M_, &FS_);
    else if (strcasecmp(s, "EDE") == 0) lr ? LeftRight_scaling_EDE(&P_, &M_, &FS_) : Left_scaling_EDE(&P_, &M_, &FS_);
    else lr ? LeftRight_scaling_CSR(&P_, &M_, &FS_) : Left_scaling_CSR(&P_, &M_, &FS_);
}
void ref_unscaling(double *x) { Left_unscaling(&P_, &M_, &FS_, x); }
