/*
 * reftr_harness.c -- drives the FEM_CODES transient 2D SUPG element build
 * (02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c with that
 * driver's CAU_ShockCapture.c, h_shock_2sqrtArea.c, eval_U_dU.c,
 * ebe_assembly.c and setzeros.c), compiled from their own sources with the
 * driver's header by oracle/Makefile, through a flat C interface.  TEST
 * INFRASTRUCTURE ONLY: built into oracle/_ref/, never linked into the product.
 *
 * The problem callbacks the driver would take from a problem module
 * (setProblem.c) are supplied here: constant diffusivity, reaction and
 * source, a constant velocity (kind 0) or a rotation about (bx, by) scaled
 * by `scale` (kind 3, the CONE/CONE_Velocity.c form), and a constant
 * prescribed value ud.
 */
#include "TranspEquation.h"

int Build_M_K_R_SUPG(ParametersType *, MatrixDataType *, FemStructsType *, FemFunctionsType *);
double CAU_ShockCapture(double, double, double, double, double, double, double, double, double,
                        double, double, double, double, double, double, double, double, double,
                        double);
double h_shock_2sqrtArea(double, double, double, double, double, double, double, double, double,
                         double, double, double);
void ebe_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[3]);

static ParametersType PT;
static MatrixDataType MT;
static FemStructsType FT;
static FemFunctionsType FF;
static int g_kind;
static double g_kappa, g_gamma, g_f, g_bx, g_by, g_scale, g_ud;

static double cb_k(void) { return g_kappa; }
static double cb_gamma(void) { return g_gamma; }
static double cb_font(double x, double y, double k, double g, double bx, double by)
{
    (void)x; (void)y; (void)k; (void)g; (void)bx; (void)by;
    return g_f;
}
static void cb_velocity(double x, double y, double B[2])
{
    if (g_kind == 3) {
        B[0] = -(y - g_by) * g_scale;
        B[1] = (x - g_bx) * g_scale;
    } else {
        B[0] = g_bx;
        B[1] = g_by;
    }
}
static double cb_upresc(double x, double y)
{
    (void)x; (void)y;
    return g_ud;
}

void reftr_setup(int nnodes, int nel, int neq, const double *x, const double *y,
                 const int *type, const int *id, const int *vertex, const int *lm, int kind,
                 double kappa, double gamma, double f, double bx, double by, double scale,
                 double ud, double alpha, double dt)
{
    memset(&PT, 0, sizeof PT); memset(&MT, 0, sizeof MT); memset(&FT, 0, sizeof FT);
    memset(&FF, 0, sizeof FF);
    g_kind = kind; g_kappa = kappa; g_gamma = gamma; g_f = f; g_bx = bx; g_by = by;
    g_scale = scale; g_ud = ud;
    PT.nnodes = nnodes; PT.nel = nel; PT.neq = neq;
    PT.Alpha = alpha; PT.DeltaT = dt;
    strcpy(PT.MatrixVectorProductScheme, "EBE");
    NodeType *Node = (NodeType *)calloc(nnodes, sizeof(NodeType));
    for (int i = 0; i < nnodes; i++) {
        Node[i].x = x[i]; Node[i].y = y[i]; Node[i].Type = type[i]; Node[i].id = id[i];
    }
    ElementType *El = (ElementType *)calloc(nel, sizeof(ElementType));
    for (int e = 0; e < nel; e++)
        for (int v = 0; v < 3; v++) El[e].Vertex[v] = vertex[3 * e + v];
    int **L = (int **)calloc(nel, sizeof(int *));
    int *Laux = (int *)calloc((size_t)nel * 3, sizeof(int));
    memcpy(Laux, lm, sizeof(int) * 3 * (size_t)nel);
    for (int e = 0; e < nel; e++) L[e] = &Laux[3 * e];
    double **Ar = (double **)calloc(nel, sizeof(double *));
    double *Aaux = (double *)calloc((size_t)nel * 9, sizeof(double));
    for (int e = 0; e < nel; e++) Ar[e] = &Aaux[9 * e];
    FT.Node = Node; FT.Element = El; FT.lm = L; FT.lmaux = Laux;
    FT.F = (double *)calloc(neq + 1, sizeof(double));
    FT.u = (double *)calloc(neq + 1, sizeof(double));
    FT.du = (double *)calloc(neq + 1, sizeof(double));
    MT.A = Ar; MT.Aaux = Aaux;
    FF.Condutivity = cb_k; FF.Reaction = cb_gamma; FF.Font = cb_font;
    FF.Velocity = cb_velocity; FF.upresc = cb_upresc;
    FF.h_shock = h_shock_2sqrtArea; FF.ShockCapture = CAU_ShockCapture;
    FF.assembly = ebe_assembly;
}

/* u, du: neq+1 ; A out nel*9 ; R out neq+1 */
void reftr_build(const double *u, const double *du, double *A, double *R)
{
    memcpy(FT.u, u, sizeof(double) * (PT.neq + 1));
    memcpy(FT.du, du, sizeof(double) * (PT.neq + 1));
    Build_M_K_R_SUPG(&PT, &MT, &FT, &FF);
    memcpy(A, MT.Aaux, sizeof(double) * 9 * (size_t)PT.nel);
    memcpy(R, FT.F, sizeof(double) * (PT.neq + 1));
}
