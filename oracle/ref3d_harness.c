/*
 * ref3d_harness.c -- drives the FEM_CODES 3D steady-transport reference
 * routines (01_SS_Transport_Equation/3D), compiled from their own sources by
 * oracle/Makefile, through a flat C interface so that tests can compare the
 * oracle restatement (oracle.c) with the real reference.  TEST INFRASTRUCTURE
 * ONLY: built into oracle/_ref/, never linked into the product.
 *
 * Compiled with -DSSTransportEquation3D against the reference's own header,
 * so the structs have the reference layout.  It reproduces what
 * 3D/01_CommonFiles/Preprocess.c:130-262 sets up for an EBE run (node ids by
 * Fill_ID, lm by Fill_LM, the A/Aaux element storage) without reading files,
 * with the DIFCONV problem callbacks of 3D/DIFCONV/*.c.  The ILU(0) entry
 * points call CSR/ILUp_precond_setup.c and CSR/ILUp_precond.c on a CSR matrix
 * given by the caller.
 */
#include "SSTransportEquation3D.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/preconditioners.h"
#include "../../../00_CommonFiles/Solvers_and_Preconditioners/ilup.h"
#include "../DIFCONV/difconv.h"

void ebe_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[4]);

static ParametersType P3;
static MatrixDataType M3;
static FemStructsType FS3;
static FemFunctionsType FF3;

/* nodes/elements as Preprocess.c:136-199 reads them (GMSH node types:
 * uType 1 = unknown) and the EBE storage of Preprocess.c:240-262 */
int ref3d_setup(int nnodes, int nel, const double *x, const double *y, const double *z,
                const int *utype, const int *vertex, double cst)
{
    memset(&P3, 0, sizeof P3); memset(&M3, 0, sizeof M3);
    memset(&FS3, 0, sizeof FS3); memset(&FF3, 0, sizeof FF3);
    strcpy(P3.MatrixVectorProductScheme, "EBE");
    P3.nnodes = nnodes;
    P3.nel = nel;
    P3.ConstApli = cst;
    NodeType *Node = (NodeType *)calloc(nnodes, sizeof(NodeType));
    for (int i = 0; i < nnodes; i++) {
        Node[i].x = x[i]; Node[i].y = y[i]; Node[i].z = z[i];
        Node[i].uType = utype[i];
    }
    int neq = 0;
    Fill_ID(&neq, Node, nnodes);
    P3.neq = neq;
    ElementType *Element = (ElementType *)calloc(nel, sizeof(ElementType));
    for (int e = 0; e < nel; e++)
        for (int j = 0; j < 4; j++) Element[e].Vertex[j] = vertex[4 * e + j];
    int **lm = (int **)calloc(nel, sizeof(int *));
    int *lmaux = (int *)calloc((size_t)nel * 4, sizeof(int));
    for (int e = 0; e < nel; e++) lm[e] = &lmaux[4 * e];
    Fill_LM(neq, nel, lm, Node, Element);
    double **A = (double **)calloc(nel, sizeof(double *));
    double *Aaux = (double *)calloc((size_t)nel * 16, sizeof(double));
    for (int e = 0; e < nel; e++) A[e] = &Aaux[16 * e];
    M3.A = A;
    M3.Aaux = Aaux;
    FS3.Node = Node;
    FS3.Element = Element;
    FS3.lm = lm;
    FS3.lmaux = lmaux;
    FS3.F = (double *)calloc(neq + 1, sizeof(double));
    FS3.u = (double *)calloc(neq + 1, sizeof(double));
    FS3.CFF = (CoefFormFuncType *)calloc(nel, sizeof(CoefFormFuncType));
    FF3.kappa = DIFCONV_kappa;
    FF3.beta = DIFCONV_beta;
    FF3.sigma = DIFCONV_sigma;
    FF3.f = DIFCONV_f;
    FF3.upresc = DIFCONV_upresc;
    FF3.MatrixAssembly = ebe_assembly;
    return neq;
}

void ref3d_lm(int *lm_out)
{
    memcpy(lm_out, FS3.lmaux, sizeof(int) * 4 * (size_t)P3.nel);
}

/* Build_SUPG.c:22-250 ; ke_out[nel*16] = A[e][16], F_out[neq+1] */
void ref3d_build(double *ke_out, double *F_out)
{
    Build_SUPG(&P3, &M3, &FS3, &FF3);
    memcpy(ke_out, M3.Aaux, sizeof(double) * 16 * (size_t)P3.nel);
    memcpy(F_out, FS3.F, sizeof(double) * (P3.neq + 1));
}

/* ILUp_precond_setup.c:8-60 with "ILU0" on a caller CSR matrix (tag 1);
 * F (neq+1) is preconditioned in place as the setup does.  Returns the
 * factors in A's CSR slots (LU), the inverted pivots (D). */
static MatrixDataType MI;
static ParametersType PI_;
void ref3d_ilu0_setup(int neq, int nnz, double *AA, int *JA, int *IA, double *F, double *LU,
                      double *D)
{
    memset(&MI, 0, sizeof MI);
    memset(&PI_, 0, sizeof PI_);
    strcpy(PI_.Preconditioner, "ILU0");
    PI_.neq = neq;
    PI_.nnzero = nnz;
    MI.AA = AA;
    MI.JA = JA;
    MI.IA = IA;
    MI.Diag = (double *)calloc(neq + 1, sizeof(double));
    ILUp_precond_setup(&PI_, &MI, NULL, 1, F);
    SparILU *lu = MI.ILUp;
    for (int i = 0; i < neq; i++) {
        int nl = lu->L->nzcount[i], nu = lu->U->nzcount[i];
        for (int j = 0; j < nl; j++) {
            int col = lu->L->ja[i][j];
            for (int q = IA[i]; q < IA[i + 1]; q++)
                if (JA[q] == col) LU[q] = lu->L->ma[i][j];
        }
        for (int j = 0; j < nu; j++) {
            int col = lu->U->ja[i][j];
            for (int q = IA[i]; q < IA[i + 1]; q++)
                if (JA[q] == col) LU[q] = lu->U->ma[i][j];
        }
        D[i] = lu->D[i];
    }
}

/* ILUp_precond.c:4-49 with the factors of the last ref3d_ilu0_setup */
void ref3d_ilu0_apply(double *p, double *z)
{
    ILUp_precond(&PI_, &MI, NULL, p, z);
}

/* EBE/ebemv3Dndof1.c:3-45 on caller arrays: lm[nel*4] and the element
 * matrices A[nel*16] laid out as Preprocess.c:240-262 (row pointers into one
 * block); Q = A P, P and Q of neq+1 entries */
int ebemv3Dndof1(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
static ParametersType PE3;
static MatrixDataType ME3;
static FemStructsType FE3;
void ref3d_ebemv_setup(int neq, int nel, const int *lm, const double *A)
{
    memset(&PE3, 0, sizeof PE3); memset(&ME3, 0, sizeof ME3); memset(&FE3, 0, sizeof FE3);
    PE3.neq = neq;
    PE3.nel = nel;
    int **L = (int **)calloc(nel, sizeof(int *));
    int *Laux = (int *)calloc((size_t)nel * 4, sizeof(int));
    memcpy(Laux, lm, sizeof(int) * 4 * (size_t)nel);
    for (int e = 0; e < nel; e++) L[e] = &Laux[4 * e];
    double **Ar = (double **)calloc(nel, sizeof(double *));
    double *Aaux = (double *)calloc((size_t)nel * 16, sizeof(double));
    memcpy(Aaux, A, sizeof(double) * 16 * (size_t)nel);
    for (int e = 0; e < nel; e++) Ar[e] = &Aaux[16 * e];
    FE3.lm = L;
    FE3.lmaux = Laux;
    ME3.A = Ar;
    ME3.Aaux = Aaux;
}

void ref3d_ebemv(double *P, double *Q) { ebemv3Dndof1(&PE3, &ME3, &FE3, P, Q); }
