/*
 * refpred_harness.c -- drives the FEM_CODES time integrator
 * (00_CommonFiles/Time_Integration/Predictor_Old.c with Criteria.c, the
 * ddot/daxpy BLAS routines, NO_precond_setup.c, Diag_precond_EBE_setup.c and
 * scaling.c) over the transient 2D SUPG build of the 02 driver
 * (Build_M_K_R_SUPG.c and its helpers), all compiled from their own sources
 * with the driver's header by oracle/Makefile.  TEST INFRASTRUCTURE ONLY:
 * built into oracle/_ref/, never linked into the product.
 *
 * Supplied here, as a problem module and the driver's Preprocess would:
 * the problem callbacks (as in reftr_harness.c), InitialSolution (u = u0),
 * and the 6-argument solver, which the reference tree declares but does not
 * implement (TranspEquation.h:133): a direct dense solve of the assembled,
 * optionally Diag-scaled system (dense_solve.h), exact on the small meshes
 * the tests use.  StopCriteria / StopTimeIntegration are the reference's
 * StopByNorm or StopByIterations and StopBySteadyState or StopByTime,
 * wrapped to record each step's correction count and to cap the number of
 * steps.
 *
 * Built a second time with -DFEMGPU_DROPIN and linked with the drop-in
 * adapter fem_codes_amd/adapters/tr2d_adapter.c instead of the reference's
 * Diag setups and scaling.c (oracle/_ref/libtr2d_dropin.so): the solver is
 * then the adapter's 6-argument pgmres on the GPU, FemFunctions->mv is the
 * adapter's ebemv and the preconditioner inside the solve is Diag_precond
 * when the Diag setup is selected; everything else is the same loop.
 */
#include "TranspEquation.h"
#include "dense_solve.h"
#ifdef FEMGPU_DROPIN
int pgmres(ParametersType *, MatrixDataType *, FemStructsType *, FemFunctionsType *, double *,
           double *);
int ebemv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Diag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
#endif

int Predictor_Old(ParametersType *, MatrixDataType *, FemStructsType *, FemFunctionsType *,
                  FemOtherFunctionsType *);
int StopByIterations(ParametersType *, double, double, int);
int StopByNorm(ParametersType *, double, double, int);
int StopBySteadyState(ParametersType *, double *, double *, double);
int StopByTime(ParametersType *, double *, double *, double);
int Build_M_K_R_SUPG(ParametersType *, MatrixDataType *, FemStructsType *, FemFunctionsType *);
double CAU_ShockCapture(double, double, double, double, double, double, double, double, double,
                        double, double, double, double, double, double, double, double, double,
                        double);
double h_shock_2sqrtArea(double, double, double, double, double, double, double, double, double,
                         double, double, double);
void ebe_assembly(ParametersType *, MatrixDataType *, FemStructsType *, int, double (*)[3]);
int NO_precond_setup(ParametersType *, MatrixDataType *, FemStructsType *, int, double *);
int Diag_precond_EBE_setup(ParametersType *, MatrixDataType *, FemStructsType *, int, double *);
int NO_scaling(ParametersType *, MatrixDataType *, FemStructsType *);
int NO_unscaling(ParametersType *, MatrixDataType *, FemStructsType *, double *);

static ParametersType PT;
static MatrixDataType MT;
static FemStructsType FT;
static FemFunctionsType FF;
static FemOtherFunctionsType FO;
static int g_kind, g_diag, g_norm, g_steady, g_cap, g_steps, g_rc, g_ncorr;
static double g_kappa, g_gamma, g_f, g_bx, g_by, g_scale, g_ud;
static const double *g_u0;
static int *g_hist;
static int g_hist_cap;

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
static int cb_initial(ParametersType *P, NodeType *N, double *u)
{
    (void)N;
    memcpy(u, g_u0, sizeof(double) * (P->neq + 1));
    return 0;
}
static int cb_solver(ParametersType *P, MatrixDataType *M, FemStructsType *S, FemFunctionsType *F,
                     double *B, double *X)
{
    (void)F;
    int rc = orc_dense_solve_ebe(P->neq, P->nel, 3, M->Aaux, S->lmaux, g_diag ? M->invDiag : NULL,
                                 B, X);
    if (rc && !g_rc) g_rc = rc;
    return rc;
}
static int cb_stop_corr(ParametersType *P, double na, double nda, int i)
{
    int s = g_norm ? StopByNorm(P, na, nda, i) : StopByIterations(P, na, nda, i);
    g_ncorr = i;
    return s;
}
static int cb_stop_time(ParametersType *P, double *u, double *u_old, double t)
{
    if (g_steps < g_hist_cap) g_hist[g_steps] = g_ncorr;
    g_steps++;
    int s = g_steady ? StopBySteadyState(P, u, u_old, t) : StopByTime(P, u, u_old, t);
    if (!s && g_cap > 0 && g_steps >= g_cap) {
        P->CurrentTime = t;
        return 1;
    }
    return s;
}

void refpred_setup(int nnodes, int nel, int neq, const double *x, const double *y,
                   const int *type, const int *id, const int *vertex, const int *lm, int kind,
                   double kappa, double gamma, double f, double bx, double by, double scale,
                   double ud)
{
    memset(&PT, 0, sizeof PT); memset(&MT, 0, sizeof MT); memset(&FT, 0, sizeof FT);
    memset(&FF, 0, sizeof FF); memset(&FO, 0, sizeof FO);
    g_kind = kind; g_kappa = kappa; g_gamma = gamma; g_f = f; g_bx = bx; g_by = by;
    g_scale = scale; g_ud = ud;
    PT.nnodes = nnodes; PT.nel = nel; PT.neq = neq;
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
    MT.A = Ar; MT.Aaux = Aaux;
    MT.Diag = (double *)calloc(neq + 1, sizeof(double));
    MT.invDiag = (double *)calloc(neq + 1, sizeof(double));
    FF.Condutivity = cb_k; FF.Reaction = cb_gamma; FF.Font = cb_font;
    FF.Velocity = cb_velocity; FF.upresc = cb_upresc;
    FF.h_shock = h_shock_2sqrtArea; FF.ShockCapture = CAU_ShockCapture;
    FF.assembly = ebe_assembly;
    FF.InitialSolution = cb_initial;
    FF.StopCriteria = cb_stop_corr;
    FF.StopTimeIntegration = cb_stop_time;
    FF.scaling = NO_scaling;
    FF.unscaling = NO_unscaling;
    FO.Build = Build_M_K_R_SUPG;
    FO.solver = cb_solver;
#ifdef FEMGPU_DROPIN
    FO.solver = pgmres;
    FF.mv = ebemv;
    PT.KrylovBasisVectorsQuantity = 30;
    PT.LinearMaxIter = 2000;
    PT.SolverTolerance = 1e-13;
#endif
}

#ifdef FEMGPU_DROPIN
/* the GPU solver's parameters (KrylovBasisVectorsQuantity, LinearMaxIter,
 * SolverTolerance) and the Arnoldi steps it counted in Parameters->iterations */
void refpred_dropin_params(int kmax, int lmax, double tol)
{
    PT.KrylovBasisVectorsQuantity = kmax;
    PT.LinearMaxIter = lmax;
    PT.SolverTolerance = tol;
}
int refpred_dropin_iterations(void) { return PT.iterations; }
#endif

/* one run of Predictor_Old from u0 (neq+1).  diag: 1 = Diag_precond_EBE_setup,
 * 0 = NO_precond_setup; stop_by_norm / steady select the criteria; cap: most
 * time steps (0 = none).  Out: u (neq+1), hist[step] = corrections of that
 * step, *nsteps, *t_final.  Returns the solver's status (0 = ok). */
int refpred_run(const double *u0, double alpha, double dt, double final_time, double nl_tol,
                int nl_maxiter, int stop_by_norm, int steady, double time_tol,
                double stab_tol, int diag, int cap, double *u, int *hist, int hist_cap,
                int *nsteps, double *t_final)
{
    PT.Alpha = alpha; PT.DeltaT = dt; PT.FinalTime = final_time;
    PT.NonLinearTolerance = nl_tol; PT.NonLinearMaxIter = nl_maxiter;
    PT.TimeIntegrationTolerance = time_tol; PT.StabilizationTolerance = stab_tol;
    PT.CurrentTime = 0.0;
    g_u0 = u0; g_diag = diag; g_norm = stop_by_norm; g_steady = steady; g_cap = cap;
    g_steps = 0; g_rc = 0; g_ncorr = 0; g_hist = hist; g_hist_cap = hist_cap;
    FF.precond_setup = diag ? Diag_precond_EBE_setup : NO_precond_setup;
#ifdef FEMGPU_DROPIN
    FF.precond = diag ? Diag_precond : NO_precond;
    FF.precondR = NO_precond;
    PT.iterations = 0;
#endif
    Predictor_Old(&PT, &MT, &FT, &FF, &FO);
    memcpy(u, FT.u, sizeof(double) * (PT.neq + 1));
    *nsteps = g_steps;
    *t_final = PT.CurrentTime;
    return g_rc;
}
