/*
 * oracle.h -- CPU restatement of the FEM_CODES solve path (test infrastructure).
 *
 * THIS IS THE CHECKER, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load liboracle.so.  The GPU library
 * (libfemgpu.so) never links or calls anything declared here.
 *
 * Every function restates one reference routine in the same loop and
 * expression order so that, built with -O2 -ffp-contract=off, it reproduces
 * the reference bit for bit.  The reference citation (path:line, relative to
 * the FEM_CODES tree) is given on each declaration.  The interface is
 * struct-agnostic: flat arrays plus sizes instead of the per-driver
 * ParametersType / MatrixDataType / FemStructsType structs.
 *
 * Conventions kept from the reference:
 *   - every equation vector has neq+1 slots; slot neq is the Dirichlet sink;
 *   - lm[e*k + j] is the equation number of local dof j of element e, or neq;
 *   - element matrices are row-major k x k per element (AoS), A[e*k*k + r*k + c].
 *
 * Pinning status of each function is recorded in the header comment of
 * oracle.c and checked by tests/test_oracle_*.py.
 */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- BLAS-1 (00_CommonFiles/BLAS_Operations) ---------------- */
double orc_ddot(int n, const double *x, const double *y);      /* ddot.c:4-19   */
void   orc_daxpy(int n, double a, const double *x, double *y); /* daxpy.c:4-21  */
void   orc_dscal(int n, double a, double *x);                  /* dscal.c:4-21  */
void   orc_dcopy(int n, const double *x, double *y);           /* dcopy.c:4-19  */
void   orc_dzero(int n, double *x);                            /* dzero.c:3-18  */

/* ---------------- matvecs (00_CommonFiles/MatrixVector_Operations) -------- */
/* csrmv  CSR/csrmv.c:3-24 */
void orc_csrmv(int neq, const double *AA, const int *JA, const int *IA,
               const double *P, double *Q);
/* ebemv EBE/ebemv.c:4-43 (k=3), ebemv3Dndof1 EBE/ebemv3Dndof1.c:3-45 (k=4),
 * ebemvNDOF4 EBE/ebemvNDOF4.c:5-112 (k=12) and its 3D branch (k=16).
 * Writes P[neq]=0 as the reference does (ebemv.c:21). */
void orc_ebemv(int neq, int nel, int k, const double *A, const int *lm,
               double *P, double *Q);
/* edemv EDE/edemv.c:4-39 ; A has (nedge+1)*4 entries, lm2 has nedge*2 */
void orc_edemv(int neq, int nedge, const double *A, const int *lm2,
               double *P, double *Q);
/* single-precision twins used for the FP32 inner solve checks */

/* ---------------- preconditioners ----------------------------------------- */
/* Diag_precond.c:3-14 : z = p*invDiag (p may alias z) */
void orc_diag_apply(int neq, const double *invDiag, const double *p, double *z);
/* EBE/Diag_precond_EBE_setup.c:3-35 ; Diag,invDiag,F have neq+1 slots */
void orc_diag_setup_ebe(int neq, int nel, int k, const double *A, const int *lm,
                        double *Diag, double *invDiag, double *F);
/* EDE/Diag_precond_EDE_setup.c:3-33 */
void orc_diag_setup_ede(int neq, int nedge, const double *A, const int *lm2,
                        double *Diag, double *invDiag, double *F);
/* CSR/Diag_precond_CSR_setup.c:3-29 */
void orc_diag_setup_csr(int neq, const double *AA, const int *JA, const int *IA,
                        double *Diag, double *invDiag, double *F);
/* EBE/Jacobi_precond_EBE_setup.c:4-43 (k=3): invDe[e*3+j] = 1/A_e[j][j],
 * then F <- Jacobi_precond_EBE(F) */
void orc_jacobi_ebe_setup(int neq, int nel, const double *A, const int *lm,
                          double *invDe, double *F);
/* EBE/Jacobi_precond_EBE.c:8-47 ; p and z must not alias (the reference
 * zeroes z first, so an aliased call returns zero) */
void orc_jacobi_ebe_apply(int neq, int nel, const double *invDe, const int *lm,
                          const double *p, double *z);
/* JacobiDOF4_precond_EBE(_setup).c: the element Jacobi of k-dof elements */
void orc_jacobi_ebe_apply_k(int neq, int nel, int k, const double *invDe, const int *lm,
                            const double *p, double *z);
void orc_jacobi_ebe_setup_k(int neq, int nel, int k, const double *A, const int *lm,
                            double *invDe, double *F);

/* ---------------- restarted GMRES ----------------------------------------- */
/* operator callback: y = Op(x); x and y have neq+1 slots and may alias */
typedef void (*orc_op_fn)(void *ctx, double *x, double *y);

typedef struct {
    orc_op_fn matvec;   void *matvec_ctx;    /* FemFunctions->ProductMatrixVector */
    orc_op_fn precond;  void *precond_ctx;   /* FemFunctions->precond  (NULL = NO_precond) */
    orc_op_fn precondR; void *precondR_ctx;  /* FemFunctions->precondR (NULL = NO_precond) */
} orc_ops;

/* pgmres  Solvers_and_Preconditioners/pgmres.c:4-211.
 * kmax = KrylovBasisVectorsQuantity, lmax = SolverMaxIter, tol = SolverTolerance.
 * B,X have neq+1 slots.  Returns the Arnoldi step count ("cont", :94).
 * If rho_hist != NULL it receives rho after every Arnoldi step (at most
 * hist_cap entries); e0_hist (optional) receives e[0] at each restart.
 * *n_cycles receives the number of restart cycles l. */
int orc_pgmres(int neq, int kmax, int lmax, double tol,
               const double *B, double *X, const orc_ops *ops,
               double *rho_hist, int hist_cap, double *e0_hist, int *n_cycles);

/* ready-made operator contexts for the callbacks above */
typedef struct { int neq; const double *AA; const int *JA; const int *IA; } orc_csr_ctx;
typedef struct { int neq, nel, k; const double *A; const int *lm; } orc_ebe_ctx;
typedef struct { int neq, nedge; const double *A; const int *lm2; } orc_ede_ctx;
typedef struct { int neq; const double *invDiag; } orc_diag_ctx;
void orc_op_csr(void *ctx, double *x, double *y);
void orc_op_ebe(void *ctx, double *x, double *y);
void orc_op_ede(void *ctx, double *x, double *y);
void orc_op_diag(void *ctx, double *x, double *y);
typedef struct { int neq; const double *LU; const int *JA; const int *IA; const double *D; } orc_ilu0_ctx;
void orc_op_ilu0(void *ctx, double *x, double *y);   /* ILUp_precond.c:4-49 */

/* ---------------- 2D mesh, structures and SUPG assembly ------------------- */
/* Structured unit-square mesh, n x n cells, each cell split along the
 * "direita" diagonal exactly as Gerador_de_Malhas/Structured_Mesh/geraMalha.c:112-176
 * numbers nodes and elements.  Boundary nodes get Type 0 (Dirichlet), inner
 * nodes Type 1 (geraMalha.c:324-332 with print_bound, :17-20).
 * x,y,type: (n+1)^2 ; vertex: 2n^2 * 3. */
void orc_mesh2d_structured(int n, double *x, double *y, int *type, int *vertex);
/* node equation numbering Preprocess.c:73-81 : id = neq++ for Type>=1 else -1.
 * returns neq */
int  orc_number_nodes(int nnodes, const int *type, int *id);
/* Fill_LM.c:3-24 for NNOEL nodes per element and one dof per node */
void orc_fill_lm(int neq, int nel, int nnoel, const int *vertex, const int *type,
                 const int *id, int *lm);

/* csr_Initialization.c:5-218 (reordering "NOT").  Two-pass: first call with
 * JA==NULL returns nnz; second call fills IA[neq+1], JA[nnz] and
 * csr_by_elem[nel*9] (slot nnz = sink). */
int  orc_csr_init(int neq, int nel, const int *lm, int *IA, int *JA, int *csr_by_elem);
/* ede_Initialization.c:4-171.  First call with lm2==NULL returns nedge (after
 * the reference's nedge-- at :103).  Second call fills order[nel*3],
 * edge_by_elem[nel*3], lm2[nedge*2]. */
int  orc_ede_init(int neq, int nel, const int *lm, int *order, int *edge_by_elem, int *lm2);

/* problem coefficients (FemFunctions callbacks of the SS-2D driver) */
enum { ORC_P_CONST = 0,      /* beta=(bx,by), f const, u_D=0              */
       ORC_P_ROTATING = 1,   /* beta=scale*(-y+0.5, x-0.5), f const, u_D=0 */
       ORC_P_VALIDACAO = 2,  /* VALIDACAO/VALIDACAO_*.c                    */
       ORC_P_ROTATE_ABOUT = 3 /* beta=scale*(-(y-by), x-bx): rotation about
                                (bx,by), as 02 CONE/CONE_Velocity.c:3-12  */ };
enum { ORC_H_2SQRTAREA = 0,  /* h_shock_2sqrtArea.c:3-8                   */
       ORC_H_SQRTAREA = 1    /* h_shock_sqrtArea.c:3-60 (every branch 2*sqrt|A|) */ };
typedef struct {
    int kind;
    double kappa, gamma, f;
    double bx, by, scale;
    int h_shock;
} orc_prob2d;

void orc_prob2d_velocity(const orc_prob2d *pb, double x, double y, double beta[2]);
double orc_prob2d_font(const orc_prob2d *pb, double x, double y, double k, double gamma,
                       double bx, double by);

/* Build_K_F_SUPG.c:5-163 with F_assembly.c:3-51, CAU_ShockCapture.c:3-30.
 * Inputs: nodes (x,y,type,id), elements (vertex), u[neq+1] (current iterate,
 * Picard), problem.  Outputs: ke[nel*9] row-major, F[neq+1]. */
void orc_build_supg2d(int nnodes, int nel, int neq, const double *x, const double *y,
                      const int *type, const int *id, const int *vertex,
                      const double *u, const orc_prob2d *pb, double *ke, double *F);
/* 02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c:6-184 with the
 * CAU shock capture of that driver (CAU_ShockCapture.c:3-30, residual with
 * du) and h_shock_2sqrtArea.c:3-6: per element Ae = me + alpha dt ke
 * (A[nel*9], row-major, the EBE assembly) and the residual
 * R = F - M dU - K U assembled in element order (R[neq+1], zeroed first,
 * R[neq] = 0).  U, dU: nodal values [nnodes] (eval_U_dU.c: free nodes from
 * u/du, prescribed nodes upresc and 0). */
void orc_build_supg2d_transient(int nnodes, int nel, int neq, const double *x, const double *y,
                                const int *vertex, const int *lm, const double *U,
                                const double *dU, const orc_prob2d *pb, double alpha, double dt,
                                double *A, double *R);
/* ebe_assembly.c:3-18 is the identity on ke (A[e][9] = ke row-major).
 * ede_assembly.c:3-35 : Aede[(nedge+1)*4] from ke (zeroed first as setzeros.c). */
void orc_ede_assembly(int nel, int nedge, const double *ke, const int *order,
                      const int *edge_by_elem, double *Aede);
/* csr_assembly.c:3-24 : AA[nnz+1] from ke (zeroed first as setzeros.c) */
void orc_csr_assembly(int nel, int nnz, const double *ke, const int *csr_by_elem, double *AA);
/* the same for k x k element matrices (3D csr_assembly.c:3-30 for k = 4) */
void orc_csr_assembly_k(int nel, int k, int nnz, const double *ke, const int *csr_by_elem,
                        double *AA);
/* csr_Initialization.c for k local dofs per element (2D :5-218 with k=3,
 * 3D :80-256 with k=4; reordering "NOT"); csr_by_elem[nel*k*k] */
int  orc_csr_init_k(int neq, int nel, int k, const int *lm, int *IA, int *JA, int *csr_by_elem);

/* ---------------- ILU(0) (Solvers_and_Preconditioners/CSR) ---------------- */
/* ILUp_precond_setup.c:8-60 + ILUp :66-167 with p = 0: LU[nnz] in A's CSR
 * slots (L below, U above the diagonal), D[neq] inverted pivots.  Returns 0
 * or 1 + i for a zero pivot at row i. */
int  orc_ilu0_factor(int neq, const double *AA, const int *JA, const int *IA, double *LU,
                     double *D);
/* ILUp_precond.c:4-49 : z = U^{-1} L^{-1} p (z may alias p) */
void orc_ilu0_apply(int neq, const double *LU, const int *JA, const int *IA, const double *D,
                    const double *p, double *z);
void orc_ilu0_apply_f32(int neq, const float *LU, const int *JA, const int *IA, const float *D,
                        const float *p, float *z);

/* ---------------- 4-dof (Euler) products and node block preconditioners --- */
/* EBE/ebe2mvNDOF4.c:4-79 with BlockDiag2_precond.c (invBD != NULL) or
 * NOBlockDiag2_precond.c (BD != NULL); Id[nnodes*4] = equations of a node */
void orc_ebe2mv_ndof4(int neq, int nel, const double *A, const int *lm, int nnodes,
                      const int *Id, const double *invBD, const double *BD, double *P,
                      double *Q);
/* the same over items of k = 8 (edges: EDE/ede2mvNDOF4.c:4-61) or 12 */
void orc_ebe2mv_items(int neq, int nitems, int k, const double *A, const int *lm, int nnodes,
                      const int *Id, const double *invBD, const double *BD, double *P,
                      double *Q);
/* BlockDiag_precond_EBE_setup.c:88-145 matrix_Inverse (row-major 4x4) */
void orc_matrix_inverse4(const double *B, double *inv);
/* BlockDiag_precond.c:3-27 */
void orc_blockdiag_apply(int neq, int nnodes, const int *Id, const double *invBD,
                         const double *p, double *z);
/* BlockDiag_precond_EBE_setup.c:6-86: BD/invBD [nnodes*16], F <- BD^{-1} F */
void orc_blockdiag_setup_ebe(int neq, int nel, int nnodes, const int *vertex, const double *A,
                             const int *Id, double *BD, double *invBD, double *F);
/* node blocks of items of nv nodes (nv = 3: the above; nv = 2: the edges,
 * BlockDiag_precond_EDE_setup.c:6-128, EDE_matrix_Inverse = matrix_Inverse) */
void orc_blockdiag_setup_items(int neq, int nitems, int nv, int nnodes, const int *nodes,
                               const double *A, const int *Id, double *BD, double *invBD,
                               double *F);
/* 03_Euler_Equations/2D/01_CommonFiles/ede_Initialization.c:4-241: returns
 * nedge; order, edge_by_elem [nel*3]; lm2 [nedge*8] (NULL: count only) */
int orc_ede4_init(int nnodes, int nel, int neq, const int *vertex, const int *ids, int *order,
                  int *edge_by_elem, int *lm2);
/* BlockDiag_precond_EDE_setup.c:82-103: enodes [nedge*2] */
void orc_ede4_edge_nodes(int nel, const int *vertex, const int *order, const int *edge_by_elem,
                         int *enodes);
/* 03_Euler_Equations/2D/01_CommonFiles/ede_assembly.c:3-113: Aede [nedge*64] */
void orc_ede4_assembly(int nel, int nedge, const double *A12, const int *order,
                       const int *edge_by_elem, double *Aede);
/* 03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c:6-1073 with the
 * CAU operator (Delta_CAU.c:4-78), the flux Jacobians of
 * Ax_Ay_calculations.c (dimless = 1: the dimensionless form with reference
 * Mach number Mach) and the slip-wall rotation (rotation.c:3-45) of the
 * nodes with rot[node] != 0 by theta[node] (rot may be NULL).  Inputs: nodal
 * U, dU [nnodes*4] (eval_U_dU.c), lm [nel*12], constant gamma, alpha =
 * Alpha_Build, dt = DeltaT_Build, tol = AuxBuild->tolerance.  Outputs:
 * A [nel*144] = Me + alpha dt Ke per element (row-major, ebe_assembly.c) and
 * R [neq+1] = -(Me dU + Ke U) assembled in element order, R[neq] = 0. */
void orc_build_euler2d(int nnodes, int nel, int neq, const double *x, const double *y,
                       const int *vertex, const int *lm, const double *U, const double *dU,
                       const int *rot, const double *theta, double gamma, double Mach,
                       int dimless, double alpha, double dt, double tol, double *A, double *R);
/* Block_scaling.c:10-59 Block_scaling_EBE: A (nel*144) scaled in place,
 * ids[nnodes*4] = Node[].id (-1 prescribed) */
void orc_block_scaling_ebe(int neq, int nel, int nnodes, const int *vertex, const int *ids,
                           const int *Id, double *A, double *BD, double *invBD, double *F);

/* ---------------- reordering --------------------------------------------- */
/* Reordering/symrcm.c:4-242 REORDERING_SYMRCM: reverse Cuthill-McKee of a
 * CSR pattern (read as CCS, as the reference does).  P[c] = old index of
 * the equation numbered c; the reference then permutes rows and columns by
 * P (reordering.c MATRIX_ROW/COL_permutation). */
void orc_symrcm(int neq, const int *IA, const int *JA, int *P);

/* ---------------- 3D steady transport ------------------------------------ */
/* Kuhn-split unit cube (6 tets per cell); x,y,z,utype: (n+1)^3, vertex: 24 n^3 */
void orc_mesh3d_structured(int n, double *x, double *y, double *z, int *utype, int *vertex);
enum { ORC_P3_DIFCONV = 0,   /* DIFCONV/DIFCONV_*.c with A = ConstApli          */
       ORC_P3_CONST = 1 };   /* constant kappa, beta, sigma, f; u = 0 on boundary */
typedef struct {
    int kind;
    double A;                 /* DIFCONV: kappa = A and the A of DIFCONV_f        */
    double kappa, bx, by, bz, sigma, f;   /* ORC_P3_CONST                         */
} orc_prob3d;
double orc_prob3d_f(const orc_prob3d *pb, double x, double y, double z);
/* Build_SUPG.c:22-196 + Element_Configuration.c:4-58 + font_assembly.c:3-57.
 * ke[nel*16] row-major, F[neq+1] (zeroed first). */
void orc_build_supg3d(int nel, int neq, const double *x, const double *y, const double *z,
                      const int *utype, const int *vertex, const int *lm,
                      const orc_prob3d *pb, double *ke, double *F);

/* Time_Integration/Predictor_Old.c:4-145 (predictor, multicorrection) with
 * Criteria.c:25-67 (StopByIterations / StopByNorm, StopBySteadyState /
 * StopByTime) over the transient build above: eval_U_dU (free nodes from u
 * and a, prescribed nodes ud and 0), Build_M_K_R_SUPG, NO_scaling,
 * NO_precond_setup or Diag_precond_EBE_setup (diag = 1), the 6-argument
 * solver as the direct solve of dense_solve.h, NO_unscaling, a += Da,
 * u += alpha dt Da.  cap > 0 ends the run after cap time steps (the
 * harness's safety limit).  u0, u: neq+1; hist[step] = corrections of each
 * step.  Returns 0, or 1 if a solve met a zero pivot. */
typedef struct {
    double alpha, dt, final_time, nl_tol, time_tol;
    int nl_maxiter, stop_by_norm, steady, cap;
} orc_pred_params;
int orc_predictor_old_tr(int nnodes, int nel, int neq, const double *x, const double *y,
                         const int *type, const int *id, const int *vertex, const int *lm,
                         const orc_prob2d *pb, double ud, const orc_pred_params *pp, int diag,
                         const double *u0, double *u, int *hist, int hist_cap, int *nsteps,
                         double *t_final);
/* the same loop over the Euler build (orc_build_euler2d; eval_U_dU.c of the
 * 03 driver: free dofs from u and a, prescribed dofs presc[0..3] and 0),
 * NO_scaling or Block_scaling_EBE (block_scaling = 1), NO_precond_setup and
 * the direct solve.  ids[nnodes*4] = Node[].id (-1 prescribed); rot, theta
 * per node (may be NULL); stab_tol = StabilizationTolerance. */
int orc_predictor_old_eu(int nnodes, int nel, int neq, const double *x, const double *y,
                         const int *vertex, const int *lm, const int *ids, const double *presc,
                         const int *rot, const double *theta, double gamma, double mach,
                         int dimless, double stab_tol, const orc_pred_params *pp,
                         int block_scaling, const double *u0, double *u, int *hist, int hist_cap,
                         int *nsteps, double *t_final);
/* the same over the driver's other schemes: 0 EBE, 1 EBE2, 2 EDE, 3 EDE2
 * (the solve: the direct solve of the scheme's product matrix) */
int orc_predictor_old_eu_scheme(int nnodes, int nel, int neq, const double *x, const double *y,
                                const int *vertex, const int *lm, const int *ids,
                                const double *presc, const int *rot, const double *theta,
                                double gamma, double mach, int dimless, double stab_tol,
                                const orc_pred_params *pp, int block_scaling, int scheme,
                                const double *u0, double *u, int *hist, int hist_cap,
                                int *nsteps, double *t_final);

#ifdef __cplusplus
}
#endif
/* Solvers_and_Preconditioners/scaling.c:10-180: Left (leftright = 0) and
 * LeftRight scalings of item matrices (EBE k = 3, EDE k = 2) or CSR, after
 * the Diag setup of the same matrix (Diag, invDiag; F already times
 * invDiag); Left_unscaling */
void orc_scaling_items(int neq, int nitems, int k, double *A, const int *lm, const double *Diag,
                       const double *invDiag, int leftright, double *F);
void orc_scaling_csr(int neq, double *AA, const int *JA, const int *IA, const double *Diag,
                     const double *invDiag, int leftright, double *F);
void orc_left_unscaling(int neq, const double *invDiag, double *x);

/* the drivers' parameter and mesh file reads, with fscanf as Preprocess.c */
int orc_read_params(const char *path, const char *kinds, char *out, int cap);
int orc_read_mesh(const char *path, int ndim, int nint, int nvert, int ecols, int *nnodes,
                  int *nel, double *xyz, int *ni, int *vertex, int *ei);

#endif
