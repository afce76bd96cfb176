// fg_internal.h -- shared internals of libfemgpu (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "femgpu.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void fg_set_error(const char *fmt, ...);

#define FG_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            fg_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,            \
                         hipGetErrorString(e_));                                 \
            return FG_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

#define FG_CHECK(expr)                                                           \
    do {                                                                         \
        int r_ = (expr);                                                         \
        if (r_ != FG_OK) return r_;                                              \
    } while (0)

#define FG_ARG(cond, msg)                                                        \
    do {                                                                         \
        if (!(cond)) {                                                           \
            fg_set_error("invalid argument: %s", msg);                           \
            return FG_ERR_ARG;                                                   \
        }                                                                        \
    } while (0)

// launch + immediate launch-error check
#define FG_LAUNCH(kern, grid, block, shm, stream, ...)                           \
    do {                                                                         \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, stream,           \
                           __VA_ARGS__);                                         \
        FG_HIP(hipGetLastError());                                               \
    } while (0)

constexpr int FG_TPB = 256;          // threads per block for every kernel
constexpr int FG_RED_MAX = 4096;     // max blocks of a grid reduction
constexpr int FG_KMAX = 64;          // max GMRES restart length
constexpr int FG_SMALL_MAX = 4096;   // equations of a system the one-workgroup solver takes
// vectors indexed through 32-bit byte offsets (8-byte entries) stay below this
constexpr int FG_OFF32_MAX = 1 << 28;
// per-GPU scratch of the allocation-free GMRES (fg_gmres_small_noalloc)
constexpr int FG_NOALLOC_SCRATCH_BYTES = 8 << 20;
// looks up the device address of that scratch for `device` (current)
int fg_noalloc_init(int device);

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// Configuration generation: bumped whenever an operator or preconditioner is
// created or destroyed or a kernel knob of one changes.  A captured solver
// cycle (hipGraph) records the generation and is replayed only while it is
// unchanged, so a graph never runs with the pointers or launch shapes of an
// object that is gone or was reconfigured.
unsigned long long fg_cfg_bump();
unsigned long long fg_cfg_gen();

// one captured restart cycle of the kernel engine (fg_gmres, FG_ENGINE_GRAPH)
struct fg_cycle_graph {
    unsigned long long cfg = 0, ws_gen = 0;
    const void *op = nullptr, *pc = nullptr;
    int n = 0, kmax = 0, ir = 0, cgs = 0;
    int pp_after = 0;                 // parity of the partial buffers after the cycle
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
};
void fg_cycle_graph_clear(fg_cycle_graph &G);

struct fg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    int red_blocks = 2048;           // grid of vector kernels / reductions
    double *red_part = nullptr;      // [FG_RED_MAX] partial sums
    unsigned *red_ticket = nullptr;  // [4] arrival tickets (self-resetting)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // solver workspace (Krylov basis, state, partial sums), grown on demand
    // and kept until fg_ctx_trim / fg_ctx_destroy, so repeated solves (the
    // corrections of a time loop) allocate nothing
    void *ws = nullptr;
    size_t ws_bytes = 0;
    unsigned long long ws_gen = 0;   // bumped when the workspace moves or is freed
    fg_cycle_graph cg;               // the last captured GMRES cycle
    // Arnoldi steps of the first FG_CYCLE_MEMO restart cycles of the last
    // kernel-engine solve (the last slot: every later cycle), the first launch
    // chunk of the same cycle of the next solve
    int cycle_steps[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
constexpr int FG_CYCLE_MEMO = 8;

// the context's solver workspace, at least `bytes` long (256-byte aligned);
// growing it waits for the stream first.  One user at a time: the solvers
// that take it (fg_gmres, fg_gmres_batched) do not nest.
int fg_ctx_workspace(fg_ctx *c, size_t bytes, void **p);
// carve consecutive 256-byte aligned pieces out of a workspace
struct fg_carver {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    }
};

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
// FG_OP_EBE2: off-block product of a k = 12 EBE operator + node pass
enum { FG_OP_EBE = 0, FG_OP_EDE = 1, FG_OP_CSR = 2, FG_OP_EBE2 = 3 };

// ---- entry pairs in sliced rows: a value column of a slice of width w fills
// w blocks of 64, entry k of lane l at k*64 + l; the paired form puts a
// lane's entries 2j and 2j+1 side by side, at 2j*64 + 2l + (k & 1), so that
// one 16-byte load reads both FP64 values (the last entry of an odd width
// alone, at k*64 + l).  Slices of up to 8 entries only (their rows are
// unrolled whole, so the compiler merges each pair's loads); wider slices
// stay plain.  fg_pair_shift is the paired position minus the plain one.
// Paired: the FP64 values of the CSR rows and of the triangle element rows
// (fg_rows_shift, K = 3; fg_ebe_rows.hip).
__host__ __device__ inline int fg_pair_shift(int k, int l, int w)
{
    return (w > 8 || ((w & 1) && k == w - 1)) ? 0 : l - (k & 1) * 63;
}
__host__ __device__ inline int fg_rows_shift(int K, int k, int l, int w)
{
    return K != 3 ? 0 : fg_pair_shift(k, l, w);
}
// per element row: the column stride w*64 and the FP64 shift in one int
__host__ __device__ inline int fg_rows_wst(int w, int shift) { return w * 64 * 128 + shift + 64; }
__host__ __device__ inline long long fg_rows_wst_stride(int wst) { return wst >> 7; }
__host__ __device__ inline int fg_rows_wst_shift(int wst) { return (wst & 127) - 64; }

struct fg_op {
    fg_op() { fg_cfg_bump(); }
    ~fg_op() { fg_cfg_bump(); }
    fg_ctx *ctx = nullptr;
    int kind = FG_OP_EBE;
    int neq = 0;
    int nitems = 0;        // elements, edges or rows
    int k = 0;             // dofs per item (EBE/EDE)
    long long nnz = 0;     // CSR nonzeros / EBE stored values

    // ---- blocked element layout (EBE / EDE) ----
    int nblk = 0, ept = 0, max_nloc = 0, nkept = 0;
    size_t smem64 = 0, smem32 = 0;  // LDS bytes per block (FP64 / FP32 values)
    int pipe_wpc = 0;            // 0: a block per workgroup; > 0: persistent, WGs per CU
    int rows_mfma = 0;           // k = 12: element row sums on the matrix cores
    int *d_eoff = nullptr;       // [nblk+1] first element (blocked order) of each block
    int *d_doff = nullptr;       // [nblk+1] first local dof of each block
    int *d_ibase = nullptr;      // [nblk+1] first incidence of each block
    uint16_t *d_ioff = nullptr;  // per block nd+1 incidence offsets (at doff[b] + b)
    uint16_t *d_inc = nullptr;   // incidence LDS slots r*ne + i, per block
    int *d_dofs = nullptr;       // [ndloc] global dof of each local dof
    int *d_slots = nullptr;      // [ndloc] -1: complete in block; else halo slot
    uint16_t *d_lml = nullptr;   // [k*nkept] local dof indices, [k][ne_b] per block
    double *d_K64 = nullptr;     // [k*k*nkept] values, [k*k][ne_b] per block
    float *d_K32 = nullptr;
    long long *d_kpos = nullptr; // [nitems] position of element e in the blocked order, -1 if dropped
    int *d_kstride = nullptr;    // [nitems] element count of element e's block
    int nhalo = 0, nslots = 0, nzero = 0;
    int *d_hdofs = nullptr, *d_hoff = nullptr, *d_zdofs = nullptr;
    int *d_rowmap = nullptr;     // [neq] -1: written by its block, >= 0: halo index, -2: untouched
    double *d_hpart64 = nullptr;
    float *d_hpart32 = nullptr;
    // incidence (dof -> element-local entries), element order: diagonal positions
    int *d_inc_off = nullptr;    // [neq+1]
    long long *d_inc_diag = nullptr;  // [ninc] index of K_jj in d_K64 for that incidence
    int *d_lm = nullptr;         // [nitems*k] original lm (AoS) for element-local preconditioners

    // ---- CSR ----
    int *d_IA = nullptr, *d_JA = nullptr;
    double *d_AA = nullptr;
    float *d_AA32 = nullptr;
    int nrb = 0;
    int *d_rb = nullptr;         // [nrb+1] row-block boundaries
    std::vector<int> h_IA, h_JA;
    // sliced rows (fg_sell.h) of the CSR product and of the EDE half-edge
    // rows: slice offsets, entry columns (-1: padding), values (EDE: pairs)
    int nslice = 0;
    long long nslot = 0;
    int *d_soff = nullptr, *d_scol = nullptr;
    short *d_soff16 = nullptr;   // CSR: columns as 16-bit offsets from the row (when all fit)
    long long *d_spos = nullptr; // CSR: slot of each entry, FP64 pair shift in the top
                                 // byte (null: row-block product)
    double *d_sv64 = nullptr;
    float *d_sv32 = nullptr;
    // EBE k = 3, 4 element rows (fg_ebe_rows.hip): packed column offsets,
    // the third offset (k = 4), source element * k + row (-1: padding)
    unsigned *d_ridx = nullptr;
    short *d_ridx2 = nullptr;
    // k = 4 neighbour form (the product's index when every row has at most
    // 15 distinct other columns): per slot lr | 3 x 4-bit neighbour index
    // (15: the sink; 0xffff: padding), per row its neighbours as offsets
    // from the equation in slices of their own (-32768: padding)
    unsigned short *d_rcode = nullptr;
    int *d_nsoff = nullptr;
    short *d_noff = nullptr;
    // k = 8, 12 node rows (fg_ebe_node.hip): nodes (four consecutive
    // equations from nbase) per lane; d_ridx then holds one word per node
    // entry, d_rsrc its item * (k/4) + node position, d_rvpos per (item,
    // node) the entry's value block plus the lane (-1: a prescribed node)
    int nnode = 0;
    int *d_nbase = nullptr;
    int *d_rsrc = nullptr;
    long long *d_rvpos = nullptr; // [nel*k] FP32 value position of row (e, lr), -1: no row
    int *d_rwst = nullptr;        // [nel*k] fg_rows_wst: column stride, FP64 shift

    // ---- systems of at most FG_SMALL_MAX equations: lane-coalesced lists for
    // the one-workgroup solver, entry c of row g at c*neq + g (-1: none) ----
    int sm_mi = 0;               // entries per row (max over the rows)
    int *d_sm_iT = nullptr;      // EBE/EDE: r*nitems + e of the c-th element of row g
                                 // (element order); CSR: JA of the c-th entry of row g
    int *d_sm_lmT = nullptr;     // EBE/EDE: lm transposed, [k][nitems]
    // EDE (k = 2): half-edge rows (fg_ede_rows.hip), a second copy of the
    // edge values ordered by equation, kept in step with the blocked copy
    int rows_kind[2] = {FG_PRODUCT_AUTO, FG_PRODUCT_AUTO}; // FP64, FP32 products: FG_PRODUCT_SLICED / _BLOCKED / _AUTO
    int rows_stale = 1;          // the rows' values lag the blocked values
    // (sliced: d_soff / d_scol / d_sv64 / d_sv32 below the CSR fields)
    long long nhalf = 0;         // half-edges (rows' entries)
    long long *d_hsrc = nullptr; // [nslot] (blocked value index << 1) | end, -1: padding

    // ---- EBE2 (shares the storage of `base` and the blocks of `bd`) ----
    fg_op *base = nullptr;
    struct fg_prec *bd = nullptr;
    int form = 0;
};

struct fg_ilu;
int fg_ilu_create(fg_op *op, fg_ilu **out);
void fg_ilu_destroy(fg_ilu *u);
int fg_ilu_factor(fg_ilu *u, fg_op *op);
int fg_ilu_apply(fg_ilu *u, fg_ctx *c, int fp32, const void *p, void *z, const int *skip);
int fg_ilu_get(const fg_ilu *u, fg_ctx *c, double *LU, double *D);
void fg_ilu_levels(const fg_ilu *u, int *nlev_lower, int *nlev_upper, int *nseg);
void fg_ilu_set_wide(fg_ilu *u, int wide_rows);
int fg_ilu_set_sweeps(fg_ilu *u, int sweeps);

struct fg_nodeblk;
void fg_nodeblk_destroy(fg_nodeblk *b);
int fg_nodeblk_setup(struct fg_prec *pc, double *F);
int fg_nodeblk_apply(struct fg_prec *pc, int fp32, const void *p, void *z, const int *skip);
int fg_ebe2_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip);
int fg_mesh2d_eval_U_dU(const fg_mesh2d *m, const double *u, const double *du,
                        const double *upresc, double *U, double *dU);
fg_ctx *fg_mesh2d_ctx(const fg_mesh2d *m);
int fg_euler2d_eval_U_dU(const fg_euler2d *m, const double *u, const double *du,
                         const double *presc, double *U, double *dU);
void fg_euler2d_sizes(const fg_euler2d *m, fg_ctx **c, int *nnodes, int *nel, int *neq);
void fg_ede4_sizes(const fg_ede4 *g, fg_ctx **c, int *nel, int *nedge);
double fg_ebe2_bytes(const fg_op *op, int fp32);

struct fg_prec {
    fg_prec() { fg_cfg_bump(); }
    ~fg_prec() { fg_cfg_bump(); }
    fg_op *op = nullptr;
    int kind = FG_PREC_NONE;
    fg_ilu *ilu = nullptr;       // FG_PREC_ILU0
    fg_nodeblk *nb = nullptr;    // FG_PREC_BLOCKDIAG
    float *d_jsum32 = nullptr;   // element Jacobi as a diagonal (FP32 inner cycle)
    double *d_diag = nullptr, *d_invdiag = nullptr;
    float *d_invdiag32 = nullptr;
    double *d_invde = nullptr;   // [nel*3] element-local Jacobi
    float *d_invde32 = nullptr;
    double *d_tmp64 = nullptr;   // scratch vector for out-of-place element Jacobi
    float *d_tmp32 = nullptr;
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// XCD-aware block index: workgroups are dispatched round-robin over the 8
// XCDs; remap so that each XCD receives a contiguous range of logical blocks
// (neighbouring blocks share halo dofs and P entries in that XCD's L2).
__device__ __forceinline__ int fg_xcd_remap(int bid, int nblk)
{
    const int NX = 8;
    int per = nblk / NX, rem = nblk % NX;
    int x = bid % NX, q = bid / NX;
    // XCD x owns logical blocks [x*per + min(x,rem), ...) of size per + (x<rem)
    int base = x * per + (x < rem ? x : rem);
    int cnt = per + (x < rem ? 1 : 0);
    if (q < cnt) return base + q;
    return bid;  // unreachable when gridDim.x == nblk
}

template <typename T>
__device__ __forceinline__ T fg_wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// deterministic block sum; result valid in thread 0
template <typename T>
__device__ __forceinline__ T fg_block_sum(T v, T *ws /* [FG_TPB/64] shared */)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = fg_wave_sum(v);
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
        r = ws[0];
#pragma unroll
        for (int w = 1; w < FG_TPB / 64; w++) r = r + ws[w];
    }
    __syncthreads();
    return r;
}

// Grid reduction: every block contributes `v` (its block partial is computed
// here); the last block to arrive sums the partials in block order and
// returns true in all of its threads with the total in *total (thread 0).
// The ticket resets itself, so consecutive kernels can reuse it.
__device__ __forceinline__ bool fg_grid_sum(double v, double *part, unsigned *ticket,
                                            double *total)
{
    __shared__ double ws[FG_TPB / 64];
    __shared__ bool last;
    double b = fg_block_sum(v, ws);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = b;
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += FG_TPB) acc = acc + part[i];
    acc = fg_block_sum(acc, ws);
    if (threadIdx.x == 0) {
        *total = acc;
        *ticket = 0u;
    }
    return true;
}

// ---------------------------------------------------------------------------
// internal host entry points shared between translation units
// ---------------------------------------------------------------------------
// skip: optional device flag; when *skip != 0 the launched kernels return at once
int fg_op_apply_any(fg_op *op, int fp32, void *P, void *Q, const int *skip = nullptr);
// EDE half-edge rows (fg_ede_rows.hip)
int fg_ede_rows_build(fg_op *op, const int *lm2, const long long *kpos);
int fg_ede_rows_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip);
// the CSR operator's AA / AA32 were written: refresh the sliced copy
int fg_csr_values_changed(fg_op *op);
void fg_ede_rows_free(fg_op *op);
inline bool fg_ede_rows_used(const fg_op *op, int fp32)
{
    return op->d_hsrc && op->rows_kind[fp32 ? 1 : 0] != FG_PRODUCT_BLOCKED;
}
int fg_ebe_rows_build(fg_op *op, const int *lm, const long long *kpos);
int fg_ebe_rows_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip);
// the element-row product skipping the slices whose problem is flagged:
// slice s (equations 64s .. 64s+63) belongs to problem sprob[s] (-1: to
// several), skipped when pflag[sprob[s] * pstride] != 0
int fg_ebe_rows_apply_masked(fg_op *op, int fp32, void *P, void *Q, const int *skip,
                             const int *sprob, const int *pflag, int pstride);
// AoS element values (device) -> blocked layout and rows in one pass
int fg_ebe_rows_set_values(fg_op *op, const double *A);
// the sums of each equation's element diagonals (Diag setup) from the rows
int fg_ebe_rows_diag(fg_op *op, double *D);
// refresh row copies that lag the blocked values (element and edge rows)
int fg_ebe_rows_refresh(fg_op *op);
int fg_ede_rows_refresh(fg_op *op);
// run every pending refresh of an operator's derived copies now (before a
// graph capture or replay, which must not depend on host-side flags)
int fg_op_prepare(fg_op *op);
int fg_ilu_sweep_count(const fg_ilu *u);
// fg_supg2d_transient_build with the element matrices written into the
// k = 3 EBE operator of the mesh (blocked values and element rows) instead of
// an AoS array
int fg_supg2d_transient_build_op(fg_mesh2d *m, const fg_supg2d_problem *pb, double alpha,
                                 double dt, const double *U, const double *dU, fg_op *op,
                                 double *R);
void fg_ebe_rows_free(fg_op *op);
int fg_ebe_node_build(fg_op *op, const int *lm, const long long *kpos);
int fg_ebe_node_set_values(fg_op *op, const double *A);
int fg_ebe_node_refresh(fg_op *op);
int fg_ebe_node_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip);
// EBE k = 3, 4 (element rows) and 8, 12 (node rows): the row product in use
// for that precision
inline bool fg_ebe_rows_used(const fg_op *op, int fp32)
{
    return op->kind == FG_OP_EBE && op->d_ridx && op->rows_kind[fp32 ? 1 : 0] != FG_PRODUCT_BLOCKED;
}
int fg_prec_apply_any(fg_prec *pc, int fp32, const void *p, void *z, const int *skip = nullptr);
int fg_vec_grid(fg_ctx *ctx, long long n);
