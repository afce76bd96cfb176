/*
 * femgpu.h -- C ABI of libfemgpu, the MI355X (gfx950) solve path of FEM_CODES.
 *
 * Plain C: opaque handles, raw pointers and sizes, int status codes.  No HIP,
 * torch or C++ types cross this boundary.  Device buffers are obtained with
 * fg_malloc (or any hipMalloc'd pointer on the context's device).
 *
 * Which reference routine each entry point replaces is cited on the
 * declaration (paths relative to the FEM_CODES tree).  The reference's own
 * struct-based signatures (csrmv/ebemv/edemv/pgmres/Diag_precond...) are
 * exported by the per-driver adapters built on top of this ABI (see
 * fem_codes_amd/adapters/), which map ParametersType / MatrixDataType /
 * FemStructsType fields onto these calls.
 *
 * Conventions kept from the reference:
 *   - equation vectors have neq+1 slots; slot neq is the Dirichlet sink
 *     and reads as zero; matvecs write P[neq] = Q[neq] = 0
 *     (MatrixVector_Operations/EBE/ebemv.c:21-22);
 *   - lm[e*k + j] is the equation of local dof j of element e, or neq;
 *   - element matrices are row-major k x k per element (AoS) on input,
 *     A[e*k*k + r*k + c] (EBE/ebemv.c:34-36, ebe_assembly.c:9-17).
 *
 * Errors: every call returns FG_OK (0) or a positive code; fg_last_error()
 * gives the message.  The reference prints and exits; its callers ignore the
 * return value, so the adapters print the message and return nonzero.
 * Threading: one context per host thread; calls on one context are
 * serialised on the context's stream.
 */
#ifndef FEMGPU_H
#define FEMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_OK             0
#define FG_ERR_HIP        1   /* a HIP runtime call failed                 */
#define FG_ERR_ARG        2   /* invalid argument or size                  */
#define FG_ERR_BREAKDOWN  3   /* GMRES breakdown (zero residual or h)      */
#define FG_ERR_NOCONV     4   /* solver hit its cycle limit (not fatal)    */
#define FG_ERR_NOMEM      5

typedef struct fg_ctx  fg_ctx;   /* device, stream, reduction scratch     */
typedef struct fg_op   fg_op;    /* a stored operator: EBE, EDE or CSR     */
typedef struct fg_prec fg_prec;  /* a preconditioner bound to an operator  */

const char *fg_last_error(void);
const char *fg_version(void);

/* ---- context and device memory ------------------------------------------ */
int fg_ctx_create(int device, fg_ctx **out);
int fg_ctx_destroy(fg_ctx *ctx);
int fg_ctx_sync(fg_ctx *ctx);
/* release the solver workspace the context keeps between solves (no
 * reference counterpart: fg_gmres and fg_gmres_batched keep their Krylov
 * basis and state in one per-context buffer, grown on demand, so the solves
 * of a time loop allocate nothing; this frees it until the next solve) */
int fg_ctx_trim(fg_ctx *ctx);
int fg_malloc(fg_ctx *ctx, size_t bytes, void **dptr);
int fg_free(fg_ctx *ctx, void *dptr);
int fg_memcpy_h2d(fg_ctx *ctx, void *dst, const void *src, size_t bytes);
int fg_memcpy_d2h(fg_ctx *ctx, void *dst, const void *src, size_t bytes);
int fg_memcpy_d2d(fg_ctx *ctx, void *dst, const void *src, size_t bytes);
int fg_memset(fg_ctx *ctx, void *dst, int value, size_t bytes);
/* element-wise precision casts on device vectors of n entries */
int fg_cast_f64_to_f32(fg_ctx *ctx, int n, const double *x, float *y);
int fg_cast_f32_to_f64(fg_ctx *ctx, int n, const float *x, double *y);

/* ---- operators ------------------------------------------------------------ */
/* Element-by-element operator, k = NNOEL*NDOF dofs per element (2..16).
 * Replaces the storage set up by Preprocess.c:181-193 and the products
 * ebemv (EBE/ebemv.c:4-43, k=3), ebemv3Dndof1 (EBE/ebemv3Dndof1.c:3-45, k=4),
 * ebemvNDOF4 (EBE/ebemvNDOF4.c:5-112 k=12, :113-253 k=16).
 * lm: host, nel*k.  elems_per_block: 0 = default (autotuned choice). */
int fg_op_ebe_create(fg_ctx *ctx, int neq, int nel, int k, const int *lm,
                     int elems_per_block, fg_op **out);
/* Edge-by-edge operator (EDE/edemv.c:4-39): nedge edges, lm2 host nedge*2,
 * 2x2 matrices; identical machinery to EBE with k=2. */
int fg_op_ede_create(fg_ctx *ctx, int neq, int nedge, const int *lm2,
                     int edges_per_block, fg_op **out);
/* The element / edge operator with the block plan's element ordering as a
 * tuning knob (no reference counterpart):
 *   FG_ORDER_BFS   blocks grown breadth-first over shared dofs from the
 *                  lowest-numbered free element (the default of the two
 *                  calls above);
 *   FG_ORDER_MESH  blocks of items_per_block consecutive items in the
 *                  caller's order (an ordering the caller chose, such as a
 *                  space-filling or RCM order of the elements, is kept).
 * Results agree to rounding (1e-14): a dof's contributions are summed in
 * element order inside a block and in block order across blocks.  edges = 1
 * builds the edge operator (k = 2, fg_op_ede_create). */
#define FG_ORDER_BFS  0
#define FG_ORDER_MESH 1
int fg_op_ebe_create_ex(fg_ctx *ctx, int neq, int nitems, int k, const int *lm,
                        int items_per_block, int ordering, int edges, fg_op **out);
/* CSR operator (CSR/csrmv.c:3-24); IA[neq+1], JA[nnz] host. */
int fg_op_csr_create(fg_ctx *ctx, int neq, int nnz, const int *IA, const int *JA, fg_op **out);
int fg_op_destroy(fg_op *op);

/* Matrix values.  EBE: nel*k*k row-major per element (MatrixData->A);
 * EDE: nedge*4 (the sink row nedge of MatrixData->A is not read);
 * CSR: nnz (MatrixData->AA without the sink slot).
 * The _host variant copies from host memory, the _dev variant from a device
 * buffer in the same AoS layout.  Both refresh the FP32 copy used by the
 * mixed-precision solver. */
int fg_op_set_values_host(fg_op *op, const double *vals);
int fg_op_set_values_dev(fg_op *op, const double *vals_dev);

/* Q = A P on device vectors of neq+1 entries; writes P[neq]=Q[neq]=0.
 * FP64 and FP32 forms (the FP32 form uses the FP32 copy of the values). */
int fg_op_apply_f64(fg_op *op, double *P, double *Q);
int fg_op_apply_f32(fg_op *op, float *P, float *Q);
/* shape queries */
int fg_op_info(const fg_op *op, int *neq, int *nitems, int *k, long long *nnz);
/* bytes one matvec must move (compulsory traffic, see metric definition):
 * EBE nel(k^2 s + 4k) + 2 neq s ; EDE nedge(4s+8) + 2 neq s ;
 * CSR nnz(s+4) + 4(neq+1) + 2 neq s, or nnz(s+2) + 2 neq s where the
 * operator stores its columns as 16-bit offsets from the row (every column
 * within 32767 of its row; the sliced product then reads 2 bytes of index
 * per entry and no row pointers) ; s = 8 (f64) or 4 (f32) */
double fg_op_algorithmic_bytes(const fg_op *op, int fp32);
/* time `reps` back-to-back applies with HIP events on the context stream;
 * returns the mean milliseconds per apply in *ms.  main_only = 1 times the
 * dominant kernel alone (EBE/EDE: the block kernel without the halo pass),
 * its launches back to back between one event pair, then one full product
 * so that Q is complete. */
int fg_op_time_apply(fg_op *op, int fp32, void *P, void *Q, int reps, int main_only,
                     double *ms);
/* EBE/EDE launch schedule (no reference counterpart: a tuning knob of the
 * blocked layout).  wg_per_cu = 0: one element block per workgroup.
 * wg_per_cu > 0: a persistent grid of wg_per_cu workgroups per CU, each
 * walking its blocks and loading the next block's element matrices while it
 * reduces the current one.  Results are identical either way. */
int fg_op_set_pipeline(fg_op *op, int wg_per_cu);
/* row sums of the 12 x 12 element products (no reference counterpart: a
 * kernel choice).  FG_ROWS_VALU (default): each thread forms its element's
 * row sums with vector multiply-adds.  FG_ROWS_MFMA: a wave takes 16
 * elements at a time and forms their row sums with 16x16x4 matrix-core
 * instructions (the diagonal of A_rows x [p_1 .. p_16]); one element block
 * per workgroup whatever fg_op_set_pipeline says.  Needs k = 12 and 256
 * elements per block; results agree to rounding (1e-14).  Set on the base
 * operator of an ebe2 operator, it applies to the off-block product too. */
#define FG_ROWS_VALU 0
#define FG_ROWS_MFMA 1
int fg_op_set_row_kernel(fg_op *op, int kind);
/* EDE product kernel (no reference counterpart: a kernel choice).
 * FG_EDGE_ROWS: the half-edges grouped by equation in edge order (64
 * equations per wavefront slice), each equation adding its terms
 * A0*p1 + A1*p2 / A2*p1 + A3*p2 in edge order -- edemv.c bit for bit (FP64).
 * FG_EDGE_BLOCKED: the blocked element kernel at k = 2 (halo sums in block
 * order, 1e-14).  FG_EDGE_AUTO (default): rows, the faster of the two in
 * both precisions as measured on MI355X.  All read the same values (the rows
 * are refreshed from the blocked copy after every change). */
#define FG_EDGE_ROWS    0
#define FG_EDGE_BLOCKED 1
#define FG_EDGE_AUTO    2
int fg_op_set_edge_kernel(fg_op *op, int kind);
/* Element / edge product kernel (no reference counterpart: a kernel choice).
 * FG_PRODUCT_SLICED: the sliced rows -- for EBE with k = 3 (ebemv.c:24-40)
 * and k = 4 (ebemv3Dndof1.c:24-42), every element row stored with its
 * equation in element order, 64 equations per wavefront slice, each equation
 * adding its element row sums (formed left to right as the reference does)
 * in element order: the reference's product bit for bit (FP64); for k = 12
 * (ebemvNDOF4.c:5-112) and k = 8 (edemvNDOF4.c:4-72) node rows, one node's
 * four equations per lane, where every node's dofs are four consecutive
 * equations or all prescribed; for EDE the half-edge rows of FG_EDGE_ROWS.
 * FG_PRODUCT_BLOCKED: the blocked element
 * kernel (1e-14).  FG_PRODUCT_AUTO (default): sliced rows where the operator
 * has them.  An element operator gets sliced rows when k is 3 or 4 and every
 * element's other columns lie within 16383 (k = 4: the third within 32767)
 * of its own equation; else it keeps the blocked kernel only and
 * FG_PRODUCT_SLICED is rejected.  At k = 4 the rows name their other columns
 * through per-equation neighbour tables when every equation has at most 15
 * distinct other columns (same result).  The choice is made per precision:
 * precisions = FG_PRODUCT_FP64, FG_PRODUCT_FP32 or both (the FP32 products of
 * the mixed-precision solver may use one kernel and the FP64 products the
 * other).  fg_op_product_kernel sets *sliced to 0 for a blocked kernel, 1
 * for sliced rows (CSR: the sliced rows, unless the pattern would carry more
 * than a quarter of padding), 2 for the k = 4 rows with neighbour tables. */
#define FG_PRODUCT_SLICED  0
#define FG_PRODUCT_BLOCKED 1
#define FG_PRODUCT_AUTO    2
#define FG_PRODUCT_FP64    1
#define FG_PRODUCT_FP32    2
int fg_op_set_product_kernel(fg_op *op, int precisions, int kind);
int fg_op_product_kernel(const fg_op *op, int fp32, int *sliced);

/* ---- preconditioners ------------------------------------------------------ */
#define FG_PREC_NONE    0   /* NO_precond.c / NO_precond_setup.c            */
#define FG_PREC_DIAG    1   /* Diag_precond.c + Diag_precond_{EBE,EDE,CSR}_setup.c */
#define FG_PREC_JACOBI  2   /* element Jacobi of an EBE operator of any k:      */
                            /* k = 3: EBE/Jacobi_precond_EBE{,_setup}.c:8-47    */
                            /* k = 12: EBE/JacobiDOF4_precond_EBE{,_setup}.c    */
                            /* :7-83 (the Euler driver's "Jacobi",              */
                            /* 03 setPreconditioner.c:89-92); z_g = sum over the */
                            /* elements of g, in element order, of p_g / A^e_jj; */
                            /* applied correctly in place (the reference zeroes */
                            /* z first, so its in-place call returns 0)         */
int fg_prec_create(fg_op *op, int kind, fg_prec **out);
/* setup: build from the operator's current values and left-precondition F
 * in place (device, neq+1 entries), as every reference *_setup does.
 * Diag: Diag=sum of element/edge diagonals or the CSR diagonal entry,
 * invDiag=1/Diag, F*=invDiag (Diag_precond_EBE_setup.c:18-33). */
int fg_prec_setup(fg_prec *pc, double *F);
/* z = M^{-1} p (p may alias z) */
int fg_prec_apply_f64(fg_prec *pc, const double *p, double *z);
int fg_prec_apply_f32(fg_prec *pc, const float *p, float *z);
/* load an invDiag computed by the caller (host, neq+1 entries; Diag kind
 * only): the drop-in path when the driver ran Diag_precond_*_setup itself
 * and passes MatrixData->invDiag (Diag_precond.c:6). */
int fg_prec_set_invdiag(fg_prec *pc, const double *invDiag_host);
/* copy Diag / invDiag (neq+1) to host for checks (Diag kind only) */
int fg_prec_get_diag(fg_prec *pc, double *Diag_host, double *invDiag_host);
int fg_prec_destroy(fg_prec *pc);

/* ILU(0) of a CSR operator: FG_PREC_ILU0 replaces ILUp_precond_setup.c:8-60
 * ("ILU0": p = 0, no fill-in) and ILUp_precond.c:4-49.  Setup factorises
 * the operator's current values (zero pivot: FG_ERR_BREAKDOWN, where the
 * reference prints and exits, ILUp_precond_setup.c:154-163) and left-
 * preconditions F; apply solves (LU) z = p.  The factors and solves are
 * level-scheduled on the device; the FP32 apply uses an FP32 copy of the
 * factors.  The preconditioner may serve a solve whose matvec operator is
 * another storage scheme of the same matrix (EBE matvec + CSR ILU(0)). */
#define FG_PREC_ILU0    3
/* dependency levels of the forward / backward sweeps and kernel segments */
int fg_prec_ilu_info(fg_prec *pc, int *nlev_lower, int *nlev_upper, int *nsegments);
/* launch schedule (no reference counterpart; a tuning knob): a level with
 * more than wide_rows rows gets a launch of its own, runs of narrower levels
 * share one workgroup (default 8192).  Results are identical either way. */
int fg_prec_ilu_schedule(fg_prec *pc, int wide_rows);
/* triangular solves of the apply (no reference counterpart: an option of
 * this build): sweeps = 0 solves L and U exactly (level-scheduled, the
 * reference's result bit for bit); sweeps = s > 0 replaces each solve by s
 * Jacobi sweeps, y <- p - L y and x <- D (y - U x), every sweep one fully
 * parallel pass.  s sweeps are a fixed linear operator, exact once s reaches
 * the number of levels; orderings such as RCM make the exact solves
 * sequential over thousands of levels, the sweeps do not.  The setup
 * preconditions F with the same solves, so set this before fg_prec_setup. */
int fg_prec_ilu_sweeps(fg_prec *pc, int sweeps);
/* factors to host: LU[nnz] in the operator's CSR slots (diagonal slot
 * unused), D[neq] inverted pivots */
int fg_prec_ilu_get(fg_prec *pc, double *LU_host, double *D_host);

/* Node block-diagonal of a 4-dof, 3-node element operator (k = 12, Euler
 * 2D; or of a 2-node edge operator, k = 8, with `vertex` = the edge nodes).  FG_PREC_BLOCKDIAG replaces BlockDiag_precond_EBE_setup.c:6-86 (setup)
 * and BlockDiag_precond.c:3-27 (apply).  vertex[nel*3]: the element's nodes,
 * in the element order of the operator's lm; ids[nnodes*4]: Node[].id of
 * every node (< 0 or == neq: prescribed).  Setup sums the same-node 4x4
 * blocks of the element matrices in element order (I_diag,
 * BlockDiag_precond_EBE_setup.c:18-19), inverts every node block with the
 * reference's closed-form cofactor expressions (matrix_Inverse, :88-145,
 * entry 14 included as written there) and left-preconditions F when F is
 * not NULL (F == NULL gives the block diagonal alone, as
 * EBE/NOBlockDiag2_precond_EBE_setup.c does).  Apply: z_node = inv_node
 * p_node; z may alias p; the sink slot reads as 0.  Nodes without a free
 * dof are skipped (their blocks are left 0). */
#define FG_PREC_BLOCKDIAG 4
int fg_prec_blockdiag_create(fg_op *op, int nnodes, const int *vertex, const int *ids,
                             fg_prec **out);
/* BD, invBD to host, [nnodes*16] row-major node blocks (MatrixData->
 * BlockDiagAux layout); either pointer may be NULL */
int fg_prec_blockdiag_get(fg_prec *pc, double *BD_host, double *invBD_host);
/* Block_scaling.c:10-59 Block_scaling_EBE: runs the block-diagonal setup
 * (F left-preconditioned, F may be NULL), then rewrites the operator's
 * element matrices in place: the prescribed dofs of each node removed
 * (matrix_Bondary_Conditions, :73-92) and every node block row
 * left-multiplied by that node's inverse block (matrix_Scaling, :95-118).
 * The block rows of nodes without a free dof (they only feed the sink slot)
 * become 0. */
int fg_block_scaling_ebe(fg_prec *blockdiag, double *F);
/* ebe2mvNDOF4 (EBE/ebe2mvNDOF4.c:4-79): a product operator over the
 * element matrices of `ebe12` that skips each row node's own 4x4 block and
 * then applies BlockDiag2_precond.c:3-27 (form FG_EBE2_BLOCKDIAG:
 * Q = P + invBD Q, node by node) or NOBlockDiag2_precond.c:3-25 (form
 * FG_EBE2_NOBLOCKDIAG: Q += BD P), with the blocks of `blockdiag` (set up
 * first).  The new operator shares ebe12's storage and blockdiag's blocks;
 * destroy it before them.  It can drive fg_gmres like any operator. */
#define FG_EBE2_BLOCKDIAG   0
#define FG_EBE2_NOBLOCKDIAG 1
int fg_op_ebe2_create(fg_op *ebe12, fg_prec *blockdiag, int form, fg_op **out);
/* element matrices back to host, the fg_op_set_values_host layout (EBE/EDE
 * operators; elements whose dofs are all prescribed are not stored and read 0) */
int fg_op_get_values_host(fg_op *op, double *vals);

/* ---- on-disk formats (host) ------------------------------------------------ */
/* Parameter file: the value of each of the first strlen(kinds) fields, in
 * order, as NUL-separated strings in values[cap]; kinds: one char per field,
 * 's', 'd', 'i' (the "%s|%lf|%d\t:%[^\n]" lines of each driver's
 * Preprocess.c, e.g. 01_SS_Transport_Equation/2D/01_CommonFiles/
 * Preprocess.c:44-61) or 'p' (value up to the first tab: the Euler
 * driver's Preconditioner line, 03 Preprocess.c:54).  FG_ERR_ARG on a short
 * or unreadable file. */
int fg_param_file_read(const char *path, const char *kinds, char *values, int cap, int *count);
/* Mesh file <title>_<nnodes>_<nel>.dat (Preprocess.c:71-99 of the 2D
 * scalar drivers; 03 Preprocess.c:80-100; 3D Preprocess.c:132-150): nnodes;
 * per node ndim doubles and nint integers; nel; per element nvert vertices
 * and etype_cols integers.  Call with xyz == NULL for the sizes, then with
 * arrays of those sizes (node_ints / vertex / elem_ints may be NULL). */
int fg_mesh_file_read(const char *path, int ndim, int nint, int nvert, int etype_cols,
                      int *nnodes, int *nel, double *xyz, int *node_ints, int *vertex,
                      int *elem_ints);

/* ---- diagonal scalings (Solvers_and_Preconditioners/scaling.c) ------------ */
/* Left_scaling_EBE/EDE/CSR (scaling.c:10-73) and LeftRight_scaling_EBE/EDE/
 * CSR (:76-162), with `diag` the FG_PREC_DIAG of the operator: its setup
 * (Diag_precond_*_setup, F *= invDiag) runs first, then the stored values
 * are scaled in place (FP32 copy refreshed): Left: row i times invDiag[i];
 * LeftRight: entry (i, j) times sqrt(invDiag[i] invDiag[j]) and F[i] times
 * sqrt(Diag[i]).  F: device, neq+1.  setScaling.c pairs LeftRight with
 * Left_unscaling (:170-180): x[i] *= sqrt(invDiag[i]), i < neq. */
#define FG_SCALING_LEFT      1
#define FG_SCALING_LEFTRIGHT 2
int fg_scaling_diag(fg_prec *diag, int mode, double *F_dev);
int fg_unscaling_left(fg_prec *diag, double *x_dev);

/* ---- edge-by-edge structure of the 4-dof (Euler) driver ------------------- */
/* 03_Euler_Equations/2D/01_CommonFiles/ede_Initialization.c:4-190: edges
 * numbered in order of first appearance (elements ascending; within an
 * element the pairs (I1,I2), (I1,I3), (I2,I3) of its sorted vertices),
 * order[nel*3] (local index of the smallest, middle, largest vertex),
 * edge_by_elem[nel*3] (EDGE_by_Element), lm2[nedge*8] (the smaller node's 4
 * dofs, then the other's; id -1 -> neq), edge_nodes[nedge*2].
 * ids[nnodes*4]: Node[].id.  The edges are the items of a k = 8 operator
 * (fg_op_ebe_create(ctx, neq, nedge, 8, lm2, ...)): edemvNDOF4.c:4-69; its
 * node block-diagonal is fg_prec_blockdiag_create(op8, nnodes, edge_nodes,
 * ids) (BlockDiag_precond_EDE_setup.c, EDE_matrix_Inverse as the EBE
 * inverse, entry 14 included) and ede2mvNDOF4.c is fg_op_ebe2_create on
 * them.  FG_ERR_ARG when an edge has more than two elements or an id lies
 * outside [-1, neq). */
typedef struct fg_ede4 fg_ede4;
int fg_ede4_create(fg_ctx *ctx, int nnodes, int nel, int neq, const int *vertex, const int *ids,
                   fg_ede4 **out);
int fg_ede4_info(const fg_ede4 *g, int *nedge);
int fg_ede4_get(const fg_ede4 *g, int *order, int *edge_by_elem, int *lm2, int *edge_nodes);
/* ede_assembly.c:3-113 of the 03 driver: Aede_dev[nedge*64] (row-major 8x8
 * edge matrices) from the element matrices A12_dev[nel*144]: the off-diagonal
 * node blocks whole, the node-diagonal blocks times 1/2, added in element
 * order into zeroed edge matrices (bit for bit). */
int fg_ede4_assemble(fg_ede4 *g, const double *A12_dev, double *Aede_dev);
int fg_ede4_destroy(fg_ede4 *g);

/* ---- mesh structures (host, once per mesh) -------------------------------- */
/* Node numbering Preprocess.c:73-81 (id = neq++ for Type>=1, else -1) and
 * Fill_LM.c:3-24 (lm[e*nnoel+j] = id of vertex j, or neq).  Returns neq. */
int fg_number_nodes(int nnodes, const int *type, int *id);
int fg_fill_lm(int neq, int nel, int nnoel, const int *vertex, const int *type,
               const int *id, int *lm);
/* csr_Initialization.c:5-218 (reordering NOT) for 3-dof elements: sorted
 * column pattern of every row and the element -> CSR slot map.  Call with
 * JA == NULL to get nnz; then IA[neq+1], JA[nnz], csr_by_elem[nel*9] (sink = nnz). */
int fg_csr_pattern(int neq, int nel, const int *lm, int *IA, int *JA, int *csr_by_elem,
                   long long *nnz);
/* ede_Initialization.c:4-171: edges (I1<I2) of every element numbered row by
 * row; order[nel*3], edge_by_elem[nel*3], lm2[nedge*2].  Call with lm2 == NULL
 * to get nedge (after the reference's nedge-- at :103). */
int fg_ede_pattern(int neq, int nel, const int *lm, int *order, int *edge_by_elem, int *lm2,
                   int *nedge);

/* ---- SUPG element assembly on the device ----------------------------------- */
/* Build_K_F_SUPG.c:5-163 + F_assembly.c:3-51 + CAU_ShockCapture.c:3-30 for
 * problems with u_D = 0.  Problem kinds mirror the FemFunctions callbacks:
 *   FG_P_CONST     beta = (bx,by), f const, kappa, gamma
 *   FG_P_ROTATING  beta = scale*(-y+0.5, x-0.5), f const, kappa, gamma
 *   FG_P_VALIDACAO 01_SS_Transport_Equation/2D/VALIDACAO/VALIDACAO_*.c
 *   FG_P_ROTATE_ABOUT beta = scale*(-(y-by), x-bx), a rotation about (bx,by)
 *                  as 02_Transport_Equation/2D/CONE/CONE_Velocity.c:3-12   */
#define FG_P_CONST      0
#define FG_P_ROTATING   1
#define FG_P_VALIDACAO  2
#define FG_P_ROTATE_ABOUT 3
#define FG_H_2SQRTAREA  0   /* h_shock_2sqrtArea.c */
#define FG_H_SQRTAREA   1   /* h_shock_sqrtArea.c  */
typedef struct {
    int kind;
    double kappa, gamma, f, bx, by, scale;
    int h_shock;
} fg_supg2d_problem;

typedef struct fg_mesh2d fg_mesh2d;   /* device copy of a P1 triangle mesh */
/* x,y,type: nnodes ; vertex: nel*3 (host).  Numbers the nodes and fills lm. */
int fg_mesh2d_create(fg_ctx *ctx, int nnodes, int nel, const double *x, const double *y,
                     const int *type, const int *vertex, fg_mesh2d **out);
int fg_mesh2d_info(const fg_mesh2d *m, int *neq, int *nel, int *nnodes);
int fg_mesh2d_get_lm(const fg_mesh2d *m, int *lm_host);   /* nel*3 */
int fg_mesh2d_destroy(fg_mesh2d *m);
/* ke_dev: nel*9 element matrices, row-major (the EBE layout of
 * ebe_assembly.c:9-17); F_dev: neq+1, assembled in element order.
 * u_dev: current iterate (neq+1) for the shock-capturing term, or NULL = 0. */
int fg_supg2d_build(fg_mesh2d *m, const fg_supg2d_problem *pb, const double *u_dev,
                    double *ke_dev, double *F_dev);
/* Transient SUPG: replaces 02_Transport_Equation/2D/01_CommonFiles/
 * Build_M_K_R_SUPG.c:6-184 with that driver's CAU shock capture
 * (CAU_ShockCapture.c:3-30) and h_shock_2sqrtArea.c.  U_dev, dU_dev: nodal
 * values [nnodes] of u and du/dt (eval_U_dU.c: free nodes from u/du,
 * prescribed nodes upresc and 0).  A_dev: nel*9 element matrices
 * me + alpha*dt*ke, row-major (ebe_assembly.c); R_dev: neq+1, the residual
 * F - M dU - K U assembled in element order, R[neq] = 0. */
int fg_supg2d_transient_build(fg_mesh2d *m, const fg_supg2d_problem *pb, double alpha,
                              double dt, const double *U_dev, const double *dU_dev,
                              double *A_dev, double *R_dev);
/* csr_assembly.c:3-24 / ede_assembly.c:3-35 on the device: the stored
 * operator receives the sum of element contributions in element order. */
int fg_op_assemble_csr(fg_op *csr, int nel, const int *csr_by_elem, const double *ke_dev);
int fg_op_assemble_ede(fg_op *ede, int nel, const int *order, const int *edge_by_elem,
                       const double *ke_dev);

/* ---- Euler 2D element assembly on the device ------------------------------- */
/* Replaces 03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c:6-1073
 * (SUPG + the CAU operator of Delta_CAU.c:4-78, flux Jacobians of
 * Ax_Ay_calculations.c, slip-wall rotation.c:3-45).  The mesh object holds
 * the triangles (vertex[nel*3]), the 12 equation numbers of every element
 * (lm[nel*12], Fill_LM.c of that driver: DOF fastest per node, prescribed =
 * neq) and the equation -> element-entry incidence used to assemble R. */
typedef struct fg_euler2d fg_euler2d;
typedef struct {
    double gamma;         /* FemFunctions->gamma (constant over the mesh)          */
    double mach;          /* Parameters->Mach (reference Mach number)               */
    int dimensionless;    /* 1: dimensionless_Ax_Ay_calculations, 0: dimensional    */
    double alpha, dt;     /* Parameters->Alpha_Build, DeltaT_Build                  */
    double tolerance;     /* FemStructs->AuxBuild->tolerance (CAU switch-off)       */
} fg_euler2d_params;
int fg_euler2d_create(fg_ctx *ctx, int nnodes, int nel, int neq, const double *x,
                      const double *y, const int *vertex, const int *lm, fg_euler2d **out);
int fg_euler2d_destroy(fg_euler2d *m);
/* U_dev, dU_dev: nodal conserved variables and their time derivatives
 * [nnodes*4] (eval_U_dU.c); rot_dev [nnodes] (nonzero: Node[].v1Type == -1)
 * and theta_dev [nnodes] (BC_theta at the node), both may be NULL.
 * A_dev: nel*144 = Me + alpha*dt*Ke per element, row-major
 * (ebe_assembly.c); R_dev: neq+1, -(Me dU + Ke U) assembled in element
 * order, R[neq] = 0.  pow(v,2) is evaluated as v*v. */
int fg_euler2d_build(fg_euler2d *m, const fg_euler2d_params *p, const double *U_dev,
                     const double *dU_dev, const int *rot_dev, const double *theta_dev,
                     double *A_dev, double *R_dev);

/* ---- reordering (host, once per mesh) -------------------------------------- */
/* Reordering/symrcm.c:4-242 REORDERING_SYMRCM: reverse Cuthill-McKee
 * numbering of a CSR pattern with the reference's tie-breaking.
 * perm[c] = old index of the equation numbered c (the row order
 * reordering.c MATRIX_ROW_permutation applies; renumber ids by the inverse,
 * as csr_Initialization.c does for Node[].id). */
int fg_rcm(int neq, const int *IA, const int *JA, int *perm);
/* reordering.c MATRIX_bandwidth: max |i - j| over the pattern */
int fg_csr_bandwidth(int neq, const int *IA, const int *JA, int *bandwidth);

/* ---- 3D steady transport (01_SS_Transport_Equation/3D) -------------------- */
/* csr_Initialization.c for k dofs per element (2D: k = 3, :5-218 of the 2D
 * file; 3D: k = 4, 3D csr_Initialization.c:80-256, reordering NOT):
 * csr_by_elem[nel*k*k], slot nnz = sink.  Two calls as fg_csr_pattern. */
int fg_csr_pattern_k(int neq, int nel, int k, const int *lm, int *IA, int *JA,
                     int *csr_by_elem, long long *nnz);
/* csr_assembly.c for k x k element matrices (3D csr_assembly.c:3-30) */
int fg_op_assemble_csr_k(fg_op *csr, int nel, int k, const int *csr_by_elem,
                         const double *ke_dev);

/* problem callbacks of the 3D driver:
 *   FG_P3_DIFCONV  3D/DIFCONV/DIFCONV_*.c: kappa = A, beta = (1,1,1),
 *                  sigma = 0, manufactured f, u = 0 on the boundary
 *   FG_P3_CONST    constant kappa, beta, sigma, f; u = 0 on the boundary */
#define FG_P3_DIFCONV 0
#define FG_P3_CONST   1
typedef struct {
    int kind;
    double A;                               /* DIFCONV: ConstApli          */
    double kappa, bx, by, bz, sigma, f;     /* FG_P3_CONST                 */
} fg_supg3d_problem;

typedef struct fg_mesh3d fg_mesh3d;   /* device copy of a P1 tetrahedral mesh */
/* x,y,z,utype: nnodes (uType 1 = unknown, Fill_ID.c) ; vertex: nel*4 (host) */
int fg_mesh3d_create(fg_ctx *ctx, int nnodes, int nel, const double *x, const double *y,
                     const double *z, const int *utype, const int *vertex, fg_mesh3d **out);
int fg_mesh3d_info(const fg_mesh3d *m, int *neq, int *nel, int *nnodes);
int fg_mesh3d_get_lm(const fg_mesh3d *m, int *lm_host);   /* nel*4, Fill_LM.c */
int fg_mesh3d_destroy(fg_mesh3d *m);
/* Build_SUPG.c:22-196 + Element_Configuration.c:4-58 + font_assembly.c:3-57:
 * ke_dev nel*16 row-major (ebe_assembly.c:11-15), F_dev neq+1 in element
 * order. */
int fg_supg3d_build(fg_mesh3d *m, const fg_supg3d_problem *pb, double *ke_dev, double *F_dev);

/* ---- restarted GMRES -------------------------------------------------------*/
#define FG_GMRES_FP64     0   /* plain FP64: the drop-in for pgmres            */
#define FG_GMRES_FP32_IR  1   /* FP32 Krylov cycle, FP64 residual/update/LSQ  */

typedef struct {
    int    kmax;        /* KrylovBasisVectorsQuantity (pgmres.c:14)           */
    int    lmax;        /* SolverMaxIter: restart-cycle cap (pgmres.c:15,184) */
    double tol;         /* SolverTolerance: rho <= tol*||B|| (pgmres.c:49-53)  */
    int    precision;   /* FG_GMRES_FP64 or FG_GMRES_FP32_IR                   */
    double inner_tol;   /* FP32-IR: a cycle stops at rho <= inner_tol*e0 (0 = 1e-5) */
    int    hist_cap;    /* capacity of rho_hist below (0 = none)               */
    double *rho_hist;   /* host: rho after every Arnoldi step                  */
    double *e0_hist;    /* host, lmax+1 entries: ||r|| at every restart, or NULL
                           (kernel-per-operation engine only)                  */
    int    engine;      /* FG_ENGINE_AUTO / _KERNELS / _FUSED / _SMALL / _GRAPH */
    int    orth;        /* FP32-IR inner cycle: FG_ORTH_MGS (0, pgmres.c:107-114)
                           or FG_ORTH_CGS (classical Gram-Schmidt: every h_{j,i}
                           from one pass, one update pass; an option of this
                           build).  The FP64 solve is always MGS.            */
} fg_gmres_opts;

#define FG_ORTH_MGS 0
#define FG_ORTH_CGS 1

/* engines: AUTO selects SMALL where it applies, else KERNELS.
 * SMALL: the whole FP64 solve in one 1024-thread workgroup (systems of at
 * most 4096 equations; CSR, or EBE/EDE with k = 2, 3, 4; NONE, DIAG or
 * element JACOBI); same steps and stopping rules as KERNELS, products
 * accumulated in the reference's element/row order.  FG_ERR_ARG when it
 * does not apply.
 * KERNELS: one kernel per vector operation, host sync once per cycle.
 * FUSED: the whole solve in one persistent cooperative kernel (EBE/EDE with
 * k <= 4 or CSR, NONE or DIAG preconditioner, neq <= 32 rows per thread of
 * a two-workgroups-per-CU grid); returns FG_ERR_ARG when it does not apply.
 * Its grid-wide barriers cost more than kernel boundaries on a multi-XCD
 * device, so it is kept as an option, not a default.
 * GRAPH: KERNELS with each restart cycle (the scaling of the first basis
 * vector, the kmax Arnoldi steps, the back substitution) captured once as a
 * hipGraph and replayed: the same kernels, arguments and results as
 * KERNELS, without one host launch per kernel.  The captured cycle is kept
 * in the context and reused by later solves with the same operator,
 * preconditioner, kmax, precision and orthogonalisation while no operator
 * or preconditioner has been created, destroyed or reconfigured since (the
 * time loops' corrections replay one capture); an ILU(0) preconditioner
 * with exact level-scheduled solves runs as KERNELS.  No reference
 * counterpart. */
#define FG_ENGINE_AUTO     0
#define FG_ENGINE_KERNELS  1
#define FG_ENGINE_FUSED    2
#define FG_ENGINE_SMALL    3
#define FG_ENGINE_GRAPH    4

typedef struct {
    int    steps;       /* Arnoldi steps ("cont", pgmres.c:94)                 */
    int    cycles;      /* restart cycles ("l", pgmres.c:182)                  */
    double rho;         /* last residual estimate                              */
    double eps;         /* tol*||B||                                           */
    double true_res;    /* ||B - M^{-1}A X|| recomputed in FP64 at the end      */
    int    status;      /* FG_OK, FG_ERR_NOCONV, FG_ERR_BREAKDOWN              */
    double ms;          /* wall time of the solve on the stream (hipEvent)     */
} fg_gmres_info;

/* pgmres (Solvers_and_Preconditioners/pgmres.c:4-211): solve M^{-1}A X = B
 * with B already left-preconditioned by fg_prec_setup; X (neq+1, device)
 * is overwritten (the reference zeroes it first, :46).  pc may be NULL
 * (NO_precond).  Returns FG_OK on convergence. */
int fg_gmres(fg_ctx *ctx, fg_op *op, fg_prec *pc, const double *B, double *X,
             const fg_gmres_opts *opts, fg_gmres_info *info);

/* fg_gmres_small_noalloc -- no counterpart in the reference: an entry point
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * solve as fg_gmres with engine FG_ENGINE_SMALL (FP64, at most 4096
 * equations, CSR or EBE/EDE with k = 2..4, NONE / DIAG / element JACOBI;
 * same steps, same bits), but its workspace (Krylov basis, results,
 * histories, per-solve transposed copies) lives in a fixed device array of
 * 8 MiB per GPU, held until the process ends, instead of allocated memory;
 * it calls no allocating routine.  FG_ERR_ARG when the solve does not fit
 * that array (kmax, the number of elements and the history lengths set its
 * size) or the engine does not apply.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other host threads, other
 * streams) race. */
int fg_gmres_small_noalloc(fg_ctx *ctx, fg_op *op, fg_prec *pc, const double *B, double *X,
                           const fg_gmres_opts *opts, fg_gmres_info *info);

/* Batched restarted GMRES: nprob independent systems solved at once (no
 * reference counterpart as one call; each problem follows pgmres.c:4-211 as
 * fg_gmres does).  The systems are one block-diagonal operator `op` (and its
 * preconditioner): problem p owns equations [seg[p], seg[p+1]) (host array,
 * seg[0] = 0, seg[nprob] = neq); element blocks must not couple problems
 * (concatenate the per-problem lm with equation offsets, one shared sink
 * slot neq).  B, X: device vectors of neq+1 entries, B already
 * left-preconditioned.  Every problem has its own Krylov basis, Hessenberg
 * matrix, rotations and stopping test tol * ||B_p|| and its own cycle count
 * (opts->lmax each); the problems advance in lockstep and a finished
 * problem drops out of the kernels.  Precision FP64 (MGS) or FP32-IR (inner
 * cycle MGS, or CGS with opts->orth = FG_ORTH_CGS); opts->rho_hist /
 * e0_hist and engine are ignored.  graph = 1 captures one
 * restart cycle into a hipGraph and replays it every cycle (preconditioner
 * NONE, Diag or element Jacobi).  info[nprob] (may be NULL).  Returns FG_OK
 * when every problem converged, else FG_ERR_NOCONV or FG_ERR_BREAKDOWN. */
int fg_gmres_batched(fg_ctx *ctx, fg_op *op, fg_prec *pc, int nprob, const int *seg,
                     const double *B, double *X, const fg_gmres_opts *opts, int graph,
                     fg_gmres_info *info);

/* ---- time integration ------------------------------------------------------ */
/* Time_Integration/Predictor_Old.c:4-145 with the stopping rules of
 * Criteria.c:25-72, on the device: u, a (du), Da, R and the element
 * matrices stay in HBM; only the norms that steer the loops reach the host.
 * Per time step: u_old = u, u += (1 - alpha) dt a, a = 0 (:57-62); then
 * corrections (:68-125) Build, precond_setup, solver, a += Da,
 * u += alpha dt Da until StopByNorm (||Da|| < nl_tol ||a|| or
 * i >= nl_maxiter) or StopByIterations (i >= nl_maxiter); steps until
 * StopBySteadyState (time_tol ||u|| > ||u - u_old|| or |T - t| < dt) or
 * StopByTime (|T - t| < 1e-15; without the steady-state test the loop also
 * ends, with FG_ERR_NOCONV, once t > T + dt/2, where the reference would run
 * on without end).  The reference declares the 6-argument
 * solver(P, M, FS, FF, B, X) but has no implementation of it; here it is
 * fg_gmres with `gmres` options (a non-converged solve is reported in
 * `status` and the loop goes on, as a solver returning without convergence
 * would). */
typedef struct {
    double alpha, dt;        /* Parameters->Alpha, DeltaT                          */
    double final_time;       /* FinalTime                                          */
    double nl_tol;           /* NonLinearTolerance                                 */
    int    nl_maxiter;       /* NonLinearMaxIter                                   */
    int    stop_by_norm;     /* StopMulticorrection: 1 "NORM", 0 "ITERATION"       */
    int    steady;           /* StopAtSteadyState: 1 "YES", 0 "NOT"                */
    double time_tol;         /* TimeIntegrationTolerance                           */
    int    max_steps;        /* > 0: end after this many steps with FG_ERR_NOCONV  */
                             /* (no reference counterpart: StopByTime never ends   */
                             /* when t misses T by more than 1e-15)                */
} fg_predictor_params;
typedef struct {
    int    steps;            /* time steps ("passo")                               */
    int    corrections;      /* correction iterations over all steps              */
    long long krylov_steps;  /* Arnoldi steps over all solves                      */
    int    unconverged;      /* solves that returned FG_ERR_NOCONV                 */
    double t;                /* final time (Parameters->CurrentTime)               */
    double ms;               /* wall time                                          */
    int    status;           /* FG_OK, FG_ERR_NOCONV (max_steps, t > T + dt/2) or  */
                             /* an error                                           */
} fg_predictor_info;
/* The transient 2D transport driver (02_Transport_Equation): Build =
 * fg_supg2d_transient_build on `m` (Build_M_K_R_SUPG.c, nodal U, dU from
 * eval_U_dU.c with the prescribed values upresc_dev[nnodes]); A: the EBE
 * (k = 3) operator on m's lm that receives the element matrices; pc: NULL
 * (NO_precond_setup) or a preconditioner of A (Diag_precond_EBE_setup,
 * element Jacobi), set up on R before every solve; no scaling (NO_scaling,
 * NO_unscaling).  u_dev: neq+1, the initial solution in, u out.
 * corr_hist[step] (host, may be NULL): corrections of each step. */
int fg_predictor_old_transport(fg_mesh2d *m, const fg_supg2d_problem *pb, fg_op *A, fg_prec *pc,
                               const double *upresc_dev, const fg_predictor_params *p,
                               const fg_gmres_opts *gmres, double *u_dev, int *corr_hist,
                               int hist_cap, fg_predictor_info *info);
/* The Euler driver (03_Euler_Equations): Build = fg_euler2d_build on `m`
 * (Build_M_K_F_SUPG.c; alpha and dt of `ep` replaced by those of `p`, as
 * Predictor_Old.c:27-28 sets Alpha_Build / DeltaT_Build; nodal U, dU from
 * eval_U_dU.c with the prescribed values presc_dev[nnodes*4]); A: the EBE
 * (k = 12) operator on m's lm; `mode` (setScaling.c / setPreconditioner.c
 * of that driver):
 *   FG_PRED_EU_NOT            NO_scaling, NO_precond_setup;
 *   FG_PRED_EU_BLOCK_SCALING  Block_scaling_EBE (A and R scaled in place,
 *                             Block_scaling.c:10-59), no preconditioner;
 *   FG_PRED_EU_BLOCKDIAG      BlockDiag_precond_EBE_setup on R and
 *                             BlockDiag_precond inside GMRES;
 *   FG_PRED_EU_TWO            the EBE2 scheme: A = fg_op_ebe2_create(op12,
 *                             blocks, FG_EBE2_BLOCKDIAG) (ebe2mvNDOF4),
 *                             BlockDiag_precond_EBE_setup on R, NO_precond;
 * `blocks`: a FG_PREC_BLOCKDIAG of the k = 12 operator (NULL for
 * FG_PRED_EU_NOT).  rot_dev, theta_dev: slip-wall nodes and angles as for
 * fg_euler2d_build. */
#define FG_PRED_EU_NOT           0
#define FG_PRED_EU_BLOCK_SCALING 1
#define FG_PRED_EU_BLOCKDIAG     2
#define FG_PRED_EU_TWO           3
int fg_predictor_old_euler(fg_euler2d *m, const fg_euler2d_params *ep, fg_op *A, fg_prec *blocks,
                           int mode, const double *presc_dev, const int *rot_dev,
                           const double *theta_dev, const fg_predictor_params *p,
                           const fg_gmres_opts *gmres, double *u_dev, int *corr_hist,
                           int hist_cap, fg_predictor_info *info);
/* The same loop over the driver's edge-by-edge scheme (Preprocess.c:163-192,
 * setMatrixVectorProductType.c:17-22 of 03_Euler_Equations): each build's
 * element matrices are assembled into the edges (fg_ede4_assemble) and A is
 * the k = 8 operator on the edges' lm2 (EDE, edemvNDOF4) or, with
 * FG_PRED_EU_TWO, fg_op_ebe2_create(op8, blocks, FG_EBE2_BLOCKDIAG) (EDE2,
 * ede2mvNDOF4).  Modes NOT, BLOCKDIAG (BlockDiag_precond_EDE_setup on R,
 * BlockDiag_precond in GMRES) and TWO; block scaling is EBE only
 * (setScaling.c:10-16).  `blocks`: fg_prec_blockdiag_create(op8, nnodes,
 * edge_nodes, ids). */
int fg_predictor_old_euler_ede(fg_euler2d *m, const fg_euler2d_params *ep, fg_ede4 *edges,
                               fg_op *A, fg_prec *blocks, int mode, const double *presc_dev,
                               const int *rot_dev, const double *theta_dev,
                               const fg_predictor_params *p, const fg_gmres_opts *gmres,
                               double *u_dev, int *corr_hist, int hist_cap,
                               fg_predictor_info *info);

#ifdef __cplusplus
}
#endif
#endif
