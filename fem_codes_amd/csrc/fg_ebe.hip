// fg_ebe.hip -- element-by-element (and edge-by-edge) operator.
//
// Reference: MatrixVector_Operations/EBE/ebemv.c:4-43 (k=3),
// ebemv3Dndof1.c:3-45 (k=4), ebemvNDOF4.c:5-253 (k=12,16) and
// EDE/edemv.c:4-39 (edges, k=2).  The reference walks the elements in order
// and scatters each element's k row sums into Q.
//
// MI355X design: the elements are cut into blocks of up to 256*EPT
// mesh-adjacent elements (breadth-first growth over shared dofs).  One
// workgroup owns one block:
//   1. the block's distinct dofs (sorted) are gathered from P into LDS;
//   2. every thread streams its elements' k*k values and u16 local dof
//      indices (both stored per block as [entry][element], values in groups
//      of four entries when k*k allows, so each load is coalesced and wide)
//      and forms the k row sums in registers;
//   3. the row sums go to LDS, and every local dof then adds the row sums
//      that belong to it through a per-block incidence list (u16 LDS slots,
//      elements ascending): a gather instead of a scatter, so there are no
//      write conflicts, one barrier, and the result does not depend on
//      thread timing (bitwise reproducible);
//   4. dofs touched only by this block are written to Q; dofs shared with
//      other blocks write a partial sum to a halo slot, and a small second
//      kernel adds the partials of each shared dof in block order.
// No atomics anywhere; HBM traffic is the element values, the u16 indices,
// one gather of P per block-local dof and one write of Q.
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "fg_ebe_body.h"
#include "fg_internal.h"

// ---------------------------------------------------------------------------
// host: block plan
// ---------------------------------------------------------------------------
namespace {

struct Plan {
    int nblk = 0, max_nloc = 0, max_ne = 0, max_ninc = 0, nkept = 0;
    std::vector<int> eoff{0}, doff{0}, ibase{0};
    std::vector<uint16_t> ioff, inc;
    std::vector<int> dofs, slots;
    std::vector<uint16_t> lml;
    std::vector<long long> kpos;   // [nel] blocked base index (value q at fg_ebe_vidx), -1 dropped
    std::vector<int> kstride;      // [nel] group stride: elements in the block x group size
    std::vector<int> hdofs, hoff{0}, zdofs;
    int nslots = 0;
};

// ordering FG_ORDER_BFS: blocks grown breadth-first over shared dofs;
// FG_ORDER_MESH: runs of eb consecutive (kept) elements in the caller's order
int build_plan(int neq, int nel, int k, const int *lm, int eb, int ordering, Plan &pl)
{
    // elements whose dofs are all Dirichlet add nothing to Q (their rows go
    // to the sink slot) and are dropped from the operator
    std::vector<int> state(nel, -1);
    for (int e = 0; e < nel; e++) {
        bool any = false;
        for (int j = 0; j < k; j++) {
            int g = lm[(size_t)e * k + j];
            if (g < 0 || g > neq) {
                fg_set_error("lm[%d][%d]=%d outside [0,%d]", e, j, g, neq);
                return FG_ERR_ARG;
            }
            any |= (g != neq);
        }
        if (!any) state[e] = -2;
    }
    // dof -> element incidence, elements ascending
    std::vector<int> ioff(neq + 1, 0);
    for (int e = 0; e < nel; e++)
        if (state[e] == -1)
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)e * k + j];
                if (g != neq) ioff[g + 1]++;
            }
    for (int g = 0; g < neq; g++) ioff[g + 1] += ioff[g];
    std::vector<int> inc(ioff[neq]);
    {
        std::vector<int> fill(ioff.begin(), ioff.end() - 1);
        for (int e = 0; e < nel; e++)
            if (state[e] == -1)
                for (int j = 0; j < k; j++) {
                    int g = lm[(size_t)e * k + j];
                    if (g != neq) inc[fill[g]++] = e;
                }
    }
    // breadth-first blocks
    std::vector<int> order;
    order.reserve(nel);
    std::vector<int> mark(nel, -1), queue;
    queue.reserve(8 * (size_t)eb);
    int seed = 0;
    if (ordering == FG_ORDER_MESH) {
        int cnt = 0;
        for (int e = 0; e < nel; e++) {
            if (state[e] != -1) continue;
            if (cnt == 0) pl.nblk++;
            state[e] = pl.nblk - 1;
            order.push_back(e);
            if (++cnt == eb) {
                pl.eoff.push_back((int)order.size());
                cnt = 0;
            }
        }
        if (cnt > 0) pl.eoff.push_back((int)order.size());
        seed = nel;
    }
    for (;;) {
        while (seed < nel && state[seed] != -1) seed++;
        if (seed >= nel) break;
        int b = pl.nblk++;
        queue.clear();
        size_t head = 0;
        queue.push_back(seed);
        mark[seed] = b;
        int cnt = 0;
        while (head < queue.size() && cnt < eb) {
            int e = queue[head++];
            state[e] = b;
            order.push_back(e);
            cnt++;
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)e * k + j];
                if (g == neq) continue;
                for (int q = ioff[g]; q < ioff[g + 1]; q++) {
                    int e2 = inc[q];
                    if (state[e2] == -1 && mark[e2] != b) {
                        mark[e2] = b;
                        queue.push_back(e2);
                    }
                }
            }
        }
        pl.eoff.push_back((int)order.size());
    }
    pl.nkept = (int)order.size();
    const int nb = pl.nblk;

    // per block: local dofs, local indices, dof -> (row, element) incidences
    std::vector<int> locid(neq, -1);
    std::vector<int> blk_of_dofcount(neq, 0);
    pl.lml.resize((size_t)k * pl.nkept);
    pl.kpos.assign(nel, -1);
    pl.kstride.assign(nel, 0);
    pl.inc.reserve((size_t)k * pl.nkept);
    std::vector<int> ld, icnt;
    for (int b = 0; b < nb; b++) {
        int e0 = pl.eoff[b], ne = pl.eoff[b + 1] - e0;
        ld.clear();
        for (int i = 0; i < ne; i++) {
            int e = order[e0 + i];
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)e * k + j];
                if (g != neq && locid[g] == -1) {
                    locid[g] = -2;
                    ld.push_back(g);
                }
            }
        }
        std::sort(ld.begin(), ld.end());
        int nloc = (int)ld.size();
        if (nloc >= 65535) {
            fg_set_error("block with %d local dofs exceeds u16 indexing", nloc);
            return FG_ERR_ARG;
        }
        for (int l = 0; l < nloc; l++) locid[ld[l]] = l;
        pl.max_nloc = std::max(pl.max_nloc, nloc);
        pl.max_ne = std::max(pl.max_ne, ne);
        if ((long long)k * ne >= 65535) {
            fg_set_error("block with %d elements of %d dofs exceeds u16 indexing", ne, k);
            return FG_ERR_ARG;
        }
        // local indices, elements in breadth-first order
        for (int i = 0; i < ne; i++) {
            int e = order[e0 + i];
            pl.kpos[e] = (long long)k * k * e0 + (long long)i * fg_ebe_vec(k);
            pl.kstride[e] = ne * fg_ebe_vec(k);
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)e * k + j];
                pl.lml[(size_t)k * e0 + (size_t)j * ne + i] = (uint16_t)(g == neq ? nloc : locid[g]);
            }
        }
        // incidences of each local dof as LDS slots j*ne + i, elements ascending
        icnt.assign(nloc + 1, 0);
        for (int i = 0; i < ne; i++)
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)order[e0 + i] * k + j];
                if (g != neq) icnt[locid[g] + 1]++;
            }
        for (int l = 0; l < nloc; l++) icnt[l + 1] += icnt[l];
        const int ninc = icnt[nloc];
        // each block's list starts on a 16-byte boundary (copied to LDS with
        // 16-byte loads); the padding entries are never read as slots
        const int ninc_pad = (ninc + 7) / 8 * 8;
        pl.max_ninc = std::max(pl.max_ninc, ninc_pad);
        const size_t ib = pl.inc.size();
        pl.inc.resize(ib + ninc_pad, 0);
        {
            std::vector<int> fp(icnt.begin(), icnt.end() - 1);
            for (int i = 0; i < ne; i++)
                for (int j = 0; j < k; j++) {
                    int g = lm[(size_t)order[e0 + i] * k + j];
                    if (g != neq) pl.inc[ib + fp[locid[g]]++] = (uint16_t)(j * ne + i);
                }
        }
        for (int l = 0; l <= nloc; l++) pl.ioff.push_back((uint16_t)icnt[l]);
        pl.ibase.push_back((int)pl.inc.size());
        for (int l = 0; l < nloc; l++) {
            pl.dofs.push_back(ld[l]);
            blk_of_dofcount[ld[l]]++;
            locid[ld[l]] = -1;
        }
        pl.doff.push_back((int)pl.dofs.size());
    }
    // shared dofs -> halo slots (blocks in ascending order per dof)
    std::vector<int> hidx(neq, -1);
    for (int g = 0; g < neq; g++) {
        if (blk_of_dofcount[g] > 1) {
            hidx[g] = (int)pl.hdofs.size();
            pl.hdofs.push_back(g);
            pl.hoff.push_back(pl.hoff.back() + blk_of_dofcount[g]);
        } else if (blk_of_dofcount[g] == 0) {
            pl.zdofs.push_back(g);
        }
    }
    pl.nslots = pl.hoff.back();
    std::vector<int> used(pl.hdofs.size(), 0);
    pl.slots.resize(pl.dofs.size());
    for (size_t l = 0; l < pl.dofs.size(); l++) {
        int g = pl.dofs[l];
        int h = hidx[g];
        pl.slots[l] = h < 0 ? -1 : pl.hoff[h] + used[h]++;
    }
    return FG_OK;
}

template <typename T>
int upload(const std::vector<T> &v, T **d)
{
    size_t n = v.size() ? v.size() : 1;
    FG_HIP(hipMalloc(d, n * sizeof(T)));
    if (!v.empty()) FG_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// device: apply
// ---------------------------------------------------------------------------
template <int K, int EPT, typename T, int BS>
__global__ __launch_bounds__(FG_TPB) void k_ebe_apply(FgEbeLayout L, const T *__restrict__ Kv,
                                                      const T *__restrict__ P, T *__restrict__ Q,
                                                      T *__restrict__ hpart, const int *skip)
{
    if (skip && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    const int b = fg_xcd_remap(blockIdx.x, L.nblk);
    fg_ebe_block<K, EPT, T, BS>(b, L, Kv, P, Q, hpart, smem_raw);
}

// persistent form: workgroup w walks logical blocks w, w+G, ...  While it
// finishes block b (row sums, incidence sums, stores) the index lists and
// element values of its next block are already in flight: they are issued
// before the finish, after b's own loads, so every wait of the finish (for
// b's values) leaves them outstanding.  The next block's P gathers follow
// the finish.  Past the last block the loads read the last block again
// (unconditional: exact counter waits).  G is a multiple of 8 so the
// XCD-contiguous remap stays aligned with the hardware's round-robin.
template <int K, int EPT, typename T, int BS>
__global__ __launch_bounds__(FG_TPB) void k_ebe_apply_pipe(FgEbeLayout L,
                                                           const T *__restrict__ Kv,
                                                           const T *__restrict__ P,
                                                           T *__restrict__ Q,
                                                           T *__restrict__ hpart, const int *skip)
{
    if (skip && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    const int nblk = L.nblk;
    int s = blockIdx.x;
    if (s >= nblk) return;
    FgEbeRegs<K, EPT, T> cur;
    // block numbers are wavefront-uniform: readfirstlane keeps the layout
    // reads (eoff, doff, ibase) scalar loads, which the vector memory
    // counter does not track
    {
        FgEbeIdx<K, EPT, T> x;
        const int b = __builtin_amdgcn_readfirstlane(fg_xcd_remap(s, nblk));
        fg_ebe_idx_load<K, EPT, T>(b, L, cur, x);
        fg_ebe_load<K, EPT, T>(b, L, Kv, cur);
        fg_ebe_fill<K, EPT, T>(b, L, P, smem_raw, cur, x);
    }
    for (;;) {
        const int b = __builtin_amdgcn_readfirstlane(fg_xcd_remap(s, nblk));
        const int sn = s + (int)gridDim.x;
        const int bn = __builtin_amdgcn_readfirstlane(fg_xcd_remap(sn < nblk ? sn : nblk - 1, nblk));
        // this block's extents, read before the next block's loads are issued
        // (later reads would wait for those loads: the kernel's stores may
        // alias the layout arrays as far as the compiler knows)
        const FgEbeBlk B = fg_ebe_blk(L, b);
        FgEbeRegs<K, EPT, T> nxt;
        FgEbeIdx<K, EPT, T> xn;
        fg_ebe_idx_load<K, EPT, T>(bn, L, nxt, xn);
        fg_ebe_load<K, EPT, T>(bn, L, Kv, nxt);
        fg_ebe_finish<K, EPT, T, BS>(B, L, Kv, Q, hpart, smem_raw, cur);
        if (sn >= nblk) break;
        __syncthreads();   // LDS is reused by the next block
        fg_ebe_fill<K, EPT, T>(bn, L, P, smem_raw, nxt, xn);
        cur = nxt;
        s = sn;
    }
}

// k = 12 with the row sums on the matrix cores (fg_op_set_row_kernel)
template <typename T, int BS>
__global__ __launch_bounds__(FG_TPB) void k_ebe_apply_mfma12(FgEbeLayout L,
                                                             const T *__restrict__ Kv,
                                                             const T *__restrict__ P,
                                                             T *__restrict__ Q,
                                                             T *__restrict__ hpart, const int *skip)
{
    if (skip && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    const int b = fg_xcd_remap(blockIdx.x, L.nblk);
    FgEbeRegs<12, 1, T> r;
    fg_ebe_gather<12, 1, T>(b, L, P, smem_raw, r);
    fg_ebe_rows_mfma12<T, BS>(b, L, Kv, smem_raw);
    __syncthreads();
    fg_ebe_reduce<12, 1, T>(fg_ebe_blk(L, b), L, Q, hpart, smem_raw, r);
}

// shared dofs: partials in block order; dofs no element touches: 0; sink: 0
template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_ebe_halo(
    int nh, const int *__restrict__ hdofs, const int *__restrict__ hoff,
    const T *__restrict__ hpart, int nz, const int *__restrict__ zdofs, int neq,
    T *__restrict__ P, T *__restrict__ Q, const int *skip)
{
    if (skip && *skip) return;
    const int h = blockIdx.x * FG_TPB + threadIdx.x;
    if (h < nh) {
        int q0 = hoff[h], q1 = hoff[h + 1];
        T s = hpart[q0];
        for (int q = q0 + 1; q < q1; q++) s = s + hpart[q];
        Q[hdofs[h]] = s;
    } else if (h < nh + nz) {
        Q[zdofs[h - nh]] = T(0);
    }
    if (h == 0) {
        Q[neq] = T(0);
        P[neq] = T(0);
    }
}

// AoS (element-major) values -> blocked layout, FP64 and FP32 copies
__global__ void k_ebe_scatter_values(long long n, int kk, const double *__restrict__ A,
                                     const long long *__restrict__ kpos,
                                     const int *__restrict__ kstride,
                                     double *__restrict__ K64, float *__restrict__ K32)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        long long e = t / kk;
        int q = (int)(t - e * kk);
        long long p = kpos[e];
        if (p < 0) continue;
        long long dst = fg_ebe_vidx(p, kstride[e], q, kk % 4 == 0 ? 4 : 1);
        double v = A[t];
        K64[dst] = v;
        K32[dst] = (float)v;
    }
}

template <int K, int EPT, typename T, int BS>
static int launch_apply_k(fg_op *op, const T *Kv, T *P, T *Q, T *hp, const int *skip)
{
    fg_ctx *c = op->ctx;
    const size_t shm = sizeof(T) == 8 ? op->smem64 : op->smem32;
    const FgEbeLayout L = fg_ebe_layout(op);
    if constexpr (K == 12 && EPT == 1) {
        if (op->rows_mfma) {
            FG_LAUNCH((k_ebe_apply_mfma12<T, BS>), op->nblk, FG_TPB, shm, c->stream, L, Kv, P, Q,
                      hp, skip);
            return FG_OK;
        }
    }
    if (op->pipe_wpc > 0) {
        long long g = (long long)c->num_cu * op->pipe_wpc;
        g = g / 8 * 8;
        if (g < 8) g = 8;
        if (g > op->nblk) g = op->nblk;
        FG_LAUNCH((k_ebe_apply_pipe<K, EPT, T, BS>), (int)g, FG_TPB, shm, c->stream, L, Kv, P, Q,
                  hp, skip);
        return FG_OK;
    }
    FG_LAUNCH((k_ebe_apply<K, EPT, T, BS>), op->nblk, FG_TPB, shm, c->stream, L, Kv, P, Q, hp,
              skip);
    return FG_OK;
}

template <int K, typename T, int BS = 0>
static int launch_apply_ept(fg_op *op, const T *Kv, T *P, T *Q, T *hp, const int *skip)
{
    switch (op->ept) {
    case 1: return launch_apply_k<K, 1, T, BS>(op, Kv, P, Q, hp, skip);
    case 2: return launch_apply_k<K, 2, T, BS>(op, Kv, P, Q, hp, skip);
    case 4: return launch_apply_k<K, 4, T, BS>(op, Kv, P, Q, hp, skip);
    case 8: return launch_apply_k<K, 8, T, BS>(op, Kv, P, Q, hp, skip);
    }
    fg_set_error("unsupported elements-per-thread %d", op->ept);
    return FG_ERR_ARG;
}

template <typename T>
static int ebe_apply(fg_op *op, const T *Kv, T *P, T *Q, T *hp, const int *skip)
{
    fg_ctx *c = op->ctx;
    if (op->nblk > 0) {
        switch (op->k) {
        case 2: FG_CHECK((launch_apply_ept<2, T>(op, Kv, P, Q, hp, skip))); break;
        case 3: FG_CHECK((launch_apply_ept<3, T>(op, Kv, P, Q, hp, skip))); break;
        case 4: FG_CHECK((launch_apply_ept<4, T>(op, Kv, P, Q, hp, skip))); break;
        case 8: FG_CHECK((launch_apply_ept<8, T>(op, Kv, P, Q, hp, skip))); break;
        case 12: FG_CHECK((launch_apply_ept<12, T>(op, Kv, P, Q, hp, skip))); break;
        case 16: FG_CHECK((launch_apply_ept<16, T>(op, Kv, P, Q, hp, skip))); break;
        default: fg_set_error("unsupported k=%d", op->k); return FG_ERR_ARG;
        }
    }
    int n = op->nhalo + op->nzero;
    int g = (n + FG_TPB - 1) / FG_TPB;
    if (g < 1) g = 1;
    FG_LAUNCH((k_ebe_halo<T>), g, FG_TPB, 0, c->stream, op->nhalo, op->d_hdofs, op->d_hoff,
              hp, op->nzero, op->d_zdofs, op->neq, P, Q, skip);
    return FG_OK;
}

// the block kernel alone (timing of the dominant kernel)
int fg_ebe_apply_main(fg_op *op, int fp32, void *P, void *Q)
{
    if (op->kind == FG_OP_EDE && fg_ede_rows_used(op, fp32))
        return fg_ede_rows_apply(op, fp32, P, Q, nullptr);
    if (fg_ebe_rows_used(op, fp32)) return fg_ebe_rows_apply(op, fp32, P, Q, nullptr);
    if (op->nblk == 0) return FG_OK;
    if (fp32) {
        switch (op->k) {
        case 2: return launch_apply_ept<2, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        case 3: return launch_apply_ept<3, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        case 4: return launch_apply_ept<4, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        case 8: return launch_apply_ept<8, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        case 12: return launch_apply_ept<12, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        case 16: return launch_apply_ept<16, float>(op, op->d_K32, (float *)P, (float *)Q, op->d_hpart32, nullptr);
        }
    } else {
        switch (op->k) {
        case 2: return launch_apply_ept<2, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        case 3: return launch_apply_ept<3, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        case 4: return launch_apply_ept<4, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        case 8: return launch_apply_ept<8, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        case 12: return launch_apply_ept<12, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        case 16: return launch_apply_ept<16, double>(op, op->d_K64, (double *)P, (double *)Q, op->d_hpart64, nullptr);
        }
    }
    fg_set_error("unsupported k=%d", op->k);
    return FG_ERR_ARG;
}

int fg_ebe_apply_f64(fg_op *op, double *P, double *Q, const int *skip)
{
    if (op->kind == FG_OP_EDE && fg_ede_rows_used(op, 0)) return fg_ede_rows_apply(op, 0, P, Q, skip);
    if (fg_ebe_rows_used(op, 0)) return fg_ebe_rows_apply(op, 0, P, Q, skip);
    return ebe_apply<double>(op, op->d_K64, P, Q, op->d_hpart64, skip);
}
int fg_ebe_apply_f32(fg_op *op, float *P, float *Q, const int *skip)
{
    if (op->kind == FG_OP_EDE && fg_ede_rows_used(op, 1)) return fg_ede_rows_apply(op, 1, P, Q, skip);
    if (fg_ebe_rows_used(op, 1)) return fg_ebe_rows_apply(op, 1, P, Q, skip);
    return ebe_apply<float>(op, op->d_K32, P, Q, op->d_hpart32, skip);
}

// off-block product of a 4-dof-per-node operator (12 dofs: 3 nodes, 8 dofs:
// an edge's 2 nodes): the row sums skip the row node's own 4x4 block
// (ebe2mvNDOF4.c:24-70, ede2mvNDOF4.c:42-51); Q complete, P and Q sink slots 0
int fg_ebe_apply_offblock(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    FG_ARG(op->k == 12 || op->k == 8, "off-block product needs k = 8 or 12");
    fg_ctx *c = op->ctx;
    if (op->nblk > 0) {
        if (op->k == 12) {
            if (fp32)
                FG_CHECK((launch_apply_ept<12, float, 4>(op, op->d_K32, (float *)P, (float *)Q,
                                                         op->d_hpart32, skip)));
            else
                FG_CHECK((launch_apply_ept<12, double, 4>(op, op->d_K64, (double *)P,
                                                          (double *)Q, op->d_hpart64, skip)));
        } else {
            if (fp32)
                FG_CHECK((launch_apply_ept<8, float, 4>(op, op->d_K32, (float *)P, (float *)Q,
                                                        op->d_hpart32, skip)));
            else
                FG_CHECK((launch_apply_ept<8, double, 4>(op, op->d_K64, (double *)P,
                                                         (double *)Q, op->d_hpart64, skip)));
        }
    }
    int n = op->nhalo + op->nzero;
    int g = (n + FG_TPB - 1) / FG_TPB;
    if (g < 1) g = 1;
    if (fp32)
        FG_LAUNCH((k_ebe_halo<float>), g, FG_TPB, 0, c->stream, op->nhalo, op->d_hdofs,
                  op->d_hoff, op->d_hpart32, op->nzero, op->d_zdofs, op->neq, (float *)P,
                  (float *)Q, skip);
    else
        FG_LAUNCH((k_ebe_halo<double>), g, FG_TPB, 0, c->stream, op->nhalo, op->d_hdofs,
                  op->d_hoff, op->d_hpart64, op->nzero, op->d_zdofs, op->neq, (double *)P,
                  (double *)Q, skip);
    return FG_OK;
}

// blocked layout -> AoS (element-major) FP64 values; dropped elements read 0
__global__ void k_ebe_gather_values(long long n, int kk, const long long *__restrict__ kpos,
                                    const int *__restrict__ kstride,
                                    const double *__restrict__ K64, double *__restrict__ A)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        long long e = t / kk;
        int q = (int)(t - e * kk);
        long long p = kpos[e];
        A[t] = p < 0 ? 0.0 : K64[fg_ebe_vidx(p, kstride[e], q, kk % 4 == 0 ? 4 : 1)];
    }
}

int fg_ebe_get_values_dev(fg_op *op, double *A)
{
    fg_ctx *c = op->ctx;
    long long n = (long long)op->nitems * op->k * op->k;
    if (n == 0) return FG_OK;
    FG_LAUNCH(k_ebe_gather_values, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, op->k * op->k,
              op->d_kpos, op->d_kstride, op->d_K64, A);
    return FG_OK;
}

// ---------------------------------------------------------------------------
// host API
// ---------------------------------------------------------------------------
static int default_ept(int k) { return fg_ebe_default_ept(k); }

static int ebe_create(fg_ctx *c, int neq, int nitems, int k, const int *lm, int per_block,
                      int ordering, int kind, fg_op **out)
{
    FG_ARG(ordering == FG_ORDER_BFS || ordering == FG_ORDER_MESH, "ordering");
    // sizes first (host-only checks): the kernels address P, Q and the halo
    // partials through 32-bit byte offsets (fg_at), so every index times 8
    // must stay below 2^32
    FG_ARG(neq >= 0 && nitems >= 0, "sizes");
    FG_ARG(neq < FG_OFF32_MAX, "neq must be below 2^28 (32-bit byte offsets of P and Q)");
    FG_ARG((long long)nitems * 16 < (1LL << 31), "nitems * k must fit in an int");
    FG_ARG(k == 2 || k == 3 || k == 4 || k == 8 || k == 12 || k == 16,
           "k must be 2, 3, 4, 8, 12 or 16");
    FG_ARG(c && out && lm, "ctx/out/lm");
    int ept = per_block > 0 ? (per_block + FG_TPB - 1) / FG_TPB : default_ept(k);
    if (ept > 4) ept = 8;
    else if (ept > 2) ept = 4;
    FG_HIP(hipSetDevice(c->device));
    Plan pl;
    FG_CHECK(build_plan(neq, nitems, k, lm, FG_TPB * ept, ordering, pl));
    FG_ARG(pl.nslots < FG_OFF32_MAX, "halo partial slots must stay below 2^28");
    fg_op *op = new fg_op();
    op->ctx = c;
    op->kind = kind;
    op->neq = neq;
    op->nitems = nitems;
    op->k = k;
    op->ept = ept;
    op->nblk = pl.nblk;
    op->max_nloc = pl.max_nloc;
    op->nkept = pl.nkept;
    op->nnz = (long long)k * k * pl.nkept;
    op->nhalo = (int)pl.hdofs.size();
    op->nslots = pl.nslots;
    op->nzero = (int)pl.zdofs.size();
    op->smem64 = fg_ebe_smem_bytes<double>(k, ept, pl.max_nloc, pl.max_ne, pl.max_ninc);
    op->smem32 = fg_ebe_smem_bytes<float>(k, ept, pl.max_nloc, pl.max_ne, pl.max_ninc);
    if (op->smem64 > 64 * 1024) {
        size_t shm = op->smem64;
        delete op;
        fg_set_error("block needs %zu bytes of LDS; use fewer elements per block", shm);
        return FG_ERR_ARG;
    }
    FG_CHECK(upload(pl.eoff, &op->d_eoff));
    FG_CHECK(upload(pl.doff, &op->d_doff));
    FG_CHECK(upload(pl.ibase, &op->d_ibase));
    FG_CHECK(upload(pl.ioff, &op->d_ioff));
    FG_CHECK(upload(pl.inc, &op->d_inc));
    FG_CHECK(upload(pl.dofs, &op->d_dofs));
    FG_CHECK(upload(pl.slots, &op->d_slots));
    FG_CHECK(upload(pl.lml, &op->d_lml));
    FG_CHECK(upload(pl.kpos, &op->d_kpos));
    FG_CHECK(upload(pl.hdofs, &op->d_hdofs));
    FG_CHECK(upload(pl.hoff, &op->d_hoff));
    FG_CHECK(upload(pl.zdofs, &op->d_zdofs));
    {
        std::vector<int> rowmap(neq, -1);
        for (size_t h = 0; h < pl.hdofs.size(); h++) rowmap[pl.hdofs[h]] = (int)h;
        for (int g : pl.zdofs) rowmap[g] = -2;
        FG_CHECK(upload(rowmap, &op->d_rowmap));
    }
    FG_CHECK(upload(pl.kstride, &op->d_kstride));
    size_t nv = (size_t)k * k * (pl.nkept ? pl.nkept : 1);
    FG_HIP(hipMalloc(&op->d_K64, nv * sizeof(double)));
    FG_HIP(hipMalloc(&op->d_K32, nv * sizeof(float)));
    FG_HIP(hipMemset(op->d_K64, 0, nv * sizeof(double)));
    FG_HIP(hipMemset(op->d_K32, 0, nv * sizeof(float)));
    size_t ns = pl.nslots ? pl.nslots : 1;
    FG_HIP(hipMalloc(&op->d_hpart64, ns * sizeof(double)));
    FG_HIP(hipMalloc(&op->d_hpart32, ns * sizeof(float)));
    // original lm (element-local preconditioners) and dof incidence with the
    // blocked position of each diagonal entry, elements ascending
    std::vector<int> lmv(lm, lm + (size_t)nitems * k);
    FG_CHECK(upload(lmv, &op->d_lm));
    std::vector<int> ioff(neq + 1, 0);
    for (int e = 0; e < nitems; e++)
        if (pl.kpos[e] >= 0)
            for (int j = 0; j < k; j++) {
                int g = lm[(size_t)e * k + j];
                if (g != neq) ioff[g + 1]++;
            }
    for (int g = 0; g < neq; g++) ioff[g + 1] += ioff[g];
    std::vector<long long> idiag(ioff[neq]);
    const bool small = neq <= FG_SMALL_MAX;
    int mi = 0;
    for (int g = 0; g < neq; g++) mi = std::max(mi, ioff[g + 1] - ioff[g]);
    std::vector<int> iT(small ? (size_t)mi * neq : 0, -1);
    {
        std::vector<int> fill(ioff.begin(), ioff.end() - 1);
        for (int e = 0; e < nitems; e++)
            if (pl.kpos[e] >= 0)
                for (int j = 0; j < k; j++) {
                    int g = lm[(size_t)e * k + j];
                    if (g == neq) continue;
                    if (small) iT[(size_t)(fill[g] - ioff[g]) * neq + g] = j * nitems + e;
                    idiag[fill[g]++] = fg_ebe_vidx(pl.kpos[e], pl.kstride[e], j * k + j,
                                                   fg_ebe_vec(k));
                }
    }
    FG_CHECK(upload(ioff, &op->d_inc_off));
    FG_CHECK(upload(idiag, &op->d_inc_diag));
    if (kind == FG_OP_EDE) FG_CHECK(fg_ede_rows_build(op, lm, pl.kpos.data()));
    else FG_CHECK(fg_ebe_rows_build(op, lm, pl.kpos.data()));
    if (small) {
        std::vector<int> lmT((size_t)k * nitems);
        for (int e = 0; e < nitems; e++)
            for (int j = 0; j < k; j++) lmT[(size_t)j * nitems + e] = lm[(size_t)e * k + j];
        op->sm_mi = mi;
        FG_CHECK(upload(iT, &op->d_sm_iT));
        FG_CHECK(upload(lmT, &op->d_sm_lmT));
    }
    *out = op;
    return FG_OK;
}

extern "C" int fg_op_ebe_create(fg_ctx *c, int neq, int nel, int k, const int *lm,
                                int elems_per_block, fg_op **out)
{
    return ebe_create(c, neq, nel, k, lm, elems_per_block, FG_ORDER_BFS, FG_OP_EBE, out);
}

extern "C" int fg_op_ede_create(fg_ctx *c, int neq, int nedge, const int *lm2,
                                int edges_per_block, fg_op **out)
{
    return ebe_create(c, neq, nedge, 2, lm2, edges_per_block, FG_ORDER_BFS, FG_OP_EDE, out);
}

extern "C" int fg_op_ebe_create_ex(fg_ctx *c, int neq, int nitems, int k, const int *lm,
                                   int items_per_block, int ordering, int edges, fg_op **out)
{
    FG_ARG(!edges || k == 2, "an edge operator has k = 2");
    return ebe_create(c, neq, nitems, k, lm, items_per_block, ordering,
                      edges ? FG_OP_EDE : FG_OP_EBE, out);
}

int fg_ebe_set_values_dev(fg_op *op, const double *A)
{
    fg_ctx *c = op->ctx;
    long long n = (long long)op->nitems * op->k * op->k;
    if (n == 0) return FG_OK;
    if (op->kind == FG_OP_EBE && op->d_rvpos) return fg_ebe_rows_set_values(op, A);
    int g = fg_vec_grid(c, n);
    FG_LAUNCH(k_ebe_scatter_values, g, FG_TPB, 0, c->stream, n, op->k * op->k, A, op->d_kpos,
              op->d_kstride, op->d_K64, op->d_K32);
    op->rows_stale = 1;
    return FG_OK;
}

void fg_ebe_free(fg_op *op)
{
    if (op->kind == FG_OP_EDE) fg_ede_rows_free(op);
    else fg_ebe_rows_free(op);
    hipFree(op->d_eoff); hipFree(op->d_doff); hipFree(op->d_dofs);
    hipFree(op->d_ibase); hipFree(op->d_ioff); hipFree(op->d_inc);
    hipFree(op->d_slots); hipFree(op->d_lml); hipFree(op->d_K64); hipFree(op->d_K32);
    hipFree(op->d_kpos); hipFree(op->d_hdofs); hipFree(op->d_hoff); hipFree(op->d_zdofs);
    hipFree(op->d_hpart64); hipFree(op->d_hpart32); hipFree(op->d_inc_off);
    hipFree(op->d_inc_diag); hipFree(op->d_lm); hipFree(op->d_kstride); hipFree(op->d_rowmap);
    hipFree(op->d_sm_iT); hipFree(op->d_sm_lmT);
}

// ---- diagonal of an element/edge operator, element order per dof ---------
// Diag_precond_EBE_setup.c:30-39 / Diag_precond_EDE_setup.c:30-39
__global__ void k_ebe_diag(int neq, const int *__restrict__ ioff,
                           const long long *__restrict__ idiag, const double *__restrict__ K64,
                           double *__restrict__ D)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g < neq; g += gridDim.x * FG_TPB) {
        double s = 0.0;
        for (int q = ioff[g]; q < ioff[g + 1]; q++) s = s + K64[idiag[q]];
        D[g] = s;
    }
}

int fg_ebe_diag(fg_op *op, double *D)
{
    fg_ctx *c = op->ctx;
    if (op->neq == 0) return FG_OK;
    // element rows: one equation per lane, the same values in the same order
    if (op->kind == FG_OP_EBE && op->d_ridx && op->k <= 4) return fg_ebe_rows_diag(op, D);
    FG_LAUNCH(k_ebe_diag, fg_vec_grid(c, op->neq), FG_TPB, 0, c->stream, op->neq,
              op->d_inc_off, op->d_inc_diag, op->d_K64, D);
    return FG_OK;
}

// ---- element-local Jacobi (EBE/Jacobi_precond_EBE.c:8-47, k = 3;
// EBE/JacobiDOF4_precond_EBE.c:7-83, k = 12; any k the operator has) --------
// z_g = sum over elements e containing g, in element order, of p_g / A^e_jj.
// invde holds 1/A^e_jj per incidence (same order as the diagonal incidence).
__global__ void k_ebe_jacobi_setup(int ninc, const long long *__restrict__ idiag,
                                   const double *__restrict__ K64, double *__restrict__ invde,
                                   float *__restrict__ invde32)
{
    for (int q = blockIdx.x * FG_TPB + threadIdx.x; q < ninc; q += gridDim.x * FG_TPB) {
        double v = 1.0 / K64[idiag[q]];
        invde[q] = v;
        invde32[q] = (float)v;
    }
}

template <typename T>
__global__ void k_ebe_jacobi_apply(int neq, const int *__restrict__ ioff,
                                   const T *__restrict__ invde, const T *__restrict__ p,
                                   T *__restrict__ z, const int *skip)
{
    if (skip && *skip) return;
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g < neq; g += gridDim.x * FG_TPB) {
        T pg = p[g], s = T(0);
        for (int q = ioff[g]; q < ioff[g + 1]; q++) s = s + pg * invde[q];
        z[g] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) z[neq] = T(0);
}

int fg_ebe_jacobi_setup(fg_op *op, double **invde, float **invde32)
{
    fg_ctx *c = op->ctx;
    int ninc = 0;
    FG_HIP(hipMemcpy(&ninc, op->d_inc_off + op->neq, sizeof(int), hipMemcpyDeviceToHost));
    if (!*invde) FG_HIP(hipMalloc(invde, sizeof(double) * (ninc ? ninc : 1)));
    if (!*invde32) FG_HIP(hipMalloc(invde32, sizeof(float) * (ninc ? ninc : 1)));
    if (ninc)
        FG_LAUNCH(k_ebe_jacobi_setup, fg_vec_grid(c, ninc), FG_TPB, 0, c->stream, ninc,
                  op->d_inc_diag, op->d_K64, *invde, *invde32);
    return FG_OK;
}

// the element Jacobi operator is diagonal: z_g = p_g * sum_e 1/A^e_gg.  The
// FP32 inner cycle of the mixed-precision solver applies it as that diagonal
// (summed in FP64, rounded once), fused into its first vector pass.
__global__ void k_ebe_jacobi_sum(int neq, const int *__restrict__ ioff,
                                 const double *__restrict__ invde, float *__restrict__ d32)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g <= neq; g += gridDim.x * FG_TPB) {
        double s = 0.0;
        if (g < neq)
            for (int q = ioff[g]; q < ioff[g + 1]; q++) s = s + invde[q];
        d32[g] = (float)s;
    }
}

int fg_ebe_jacobi_diag32(fg_op *op, const double *invde, float *d32)
{
    fg_ctx *c = op->ctx;
    FG_LAUNCH(k_ebe_jacobi_sum, fg_vec_grid(c, op->neq + 1), FG_TPB, 0, c->stream, op->neq,
              op->d_inc_off, invde, d32);
    return FG_OK;
}

int fg_ebe_jacobi_apply(fg_op *op, int fp32, const void *invde, const void *p, void *z,
                        const int *skip)
{
    fg_ctx *c = op->ctx;
    int g = fg_vec_grid(c, op->neq);
    if (fp32)
        FG_LAUNCH(k_ebe_jacobi_apply<float>, g, FG_TPB, 0, c->stream, op->neq, op->d_inc_off,
                  (const float *)invde, (const float *)p, (float *)z, skip);
    else
        FG_LAUNCH(k_ebe_jacobi_apply<double>, g, FG_TPB, 0, c->stream, op->neq, op->d_inc_off,
                  (const double *)invde, (const double *)p, (double *)z, skip);
    return FG_OK;
}
