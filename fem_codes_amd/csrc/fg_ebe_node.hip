// fg_ebe_node.hip -- the 4-dof-per-node item products (k = 12: triangles,
// k = 8: edges) as node rows.
//
// Reference: MatrixVector_Operations/EBE/ebemvNDOF4.c:5-112 (k = 12) and
// EDE/edemvNDOF4.c:4-72 (k = 8).  For item I with dofs i_1..i_k (k/4 nodes,
// four consecutive dofs each) and row-major values A[I], Q[i_r] += A[I][k r]
// * p_1 + ... + A[I][k r + k-1] * p_k (the row sum formed left to right, then
// added to Q), after Q = 0 and P[neq] = 0: equation r receives one row sum
// per item that holds it, in item order.
//
// MI355X design: the four equations of a node move together, so a node is
// one lane of a wavefront slice (fg_sell.h): its entries are the items that
// hold it, in item order, each carrying the node's 4 item rows (4k values)
// and one 32-bit word -- the node's position in the item (2 bits) and the
// first equations of the item's other nodes as 15-bit offsets from the
// lane's own (the Dirichlet sink: -16384).  A lane loads its own node's 4 P
// values once, streams each entry's values with 16-byte loads, gathers the 4
// P values of each other node, forms the 4 row sums left to right in column
// order and adds them in item order: the reference's product bit for bit (FMA
// contraction disabled).  No LDS, no halo pass.
//
// Built when every item's node groups are either four consecutive
// equations or all prescribed, the groups do not overlap, every equation lies
// in one, and every offset fits its field (the Euler drivers number the four
// variables of a free node consecutively); else the operator keeps the
// blocked kernel only.
#include <algorithm>
#include <vector>

#include "fg_ebe_body.h"
#include "fg_internal.h"
#include "fg_sell.h"

namespace {

// k dofs per item, k/4 nodes, 4k values per entry (4 rows x k columns)
constexpr unsigned NODE_PAD = 0u;      // padding (no entry has a zero first offset)
constexpr int NOFF_SINK = -16384;
constexpr int NOFF_MAX = 16383;

__device__ __forceinline__ int nfield_a(unsigned w) { return ((int)(w << 15)) >> 17; }
__device__ __forceinline__ int nfield_b(unsigned w) { return ((int)w) >> 17; }

// value j = v*k + c (row v of the node, column c) of an entry whose block of
// 4k*64 values starts at E, for lane l: 16-byte groups, coalesced over the
// lanes (4 FP32 or 2 FP64 values per group)
template <typename T>
__host__ __device__ inline long long node_vpos(long long E, int l, int j)
{
    constexpr int VW = 16 / (int)sizeof(T);
    return E + ((long long)(j / VW) * FG_SELL_W + l) * VW + j % VW;
}

template <typename T>
struct VecOf;
template <>
struct VecOf<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct VecOf<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename T, int NK>
__global__ __launch_bounds__(FG_TPB) void k_ebe_node_rows(int nslice, int nnode,
                                                          const int *__restrict__ soff,
                                                          const int *__restrict__ nbase,
                                                          const unsigned *__restrict__ word,
                                                          const T *__restrict__ V,
                                                          T *__restrict__ P, T *__restrict__ Q,
                                                          int neq, const int *skip)
{
    if (skip && *skip) return;
    using TV = typename VecOf<T>::type;
    constexpr int VW = 16 / (int)sizeof(T);
    constexpr int NV = 4 * NK;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    if (s < nslice) {
        const int i = s * FG_SELL_W + (int)lane;
        // idle lanes use node 0's equations (all their entries are padding)
        const int b = i < nnode ? nbase[i] : 0;
        T pown[4];
#pragma unroll
        for (int v = 0; v < 4; v++) pown[v] = P[b + v];
        const int o0 = soff[s], w = soff[s + 1] - o0;
        T acc[4] = {T(0), T(0), T(0), T(0)};
        for (int k = 0; k < w; k++) {
            const unsigned wd = __builtin_nontemporal_load(&word[(long long)(o0 + k) * FG_SELL_W + lane]);
            const TV *E = reinterpret_cast<const TV *>(V + (long long)(o0 + k) * NV * FG_SELL_W) + lane;
            T a[NV];
#pragma unroll
            for (int g = 0; g < NV / VW; g++) {
                const TV x = __builtin_nontemporal_load(&E[g * FG_SELL_W]);
#pragma unroll
                for (int j = 0; j < VW; j++) a[g * VW + j] = x[j];
            }
            // the other nodes' P (0 for the sink); padding reads the own node
            const int oa = nfield_a(wd), ob = NK == 12 ? nfield_b(wd) : NOFF_SINK;
            T pa[4], pb[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const T ta = P[(oa == NOFF_SINK ? b : b + oa) + v];
                pa[v] = oa == NOFF_SINK ? T(0) : ta;
                if constexpr (NK == 12) {
                    const T tb = P[(ob == NOFF_SINK ? b : b + ob) + v];
                    pb[v] = ob == NOFF_SINK ? T(0) : tb;
                } else {
                    pb[v] = T(0);
                }
            }
            const int a0 = (int)(wd & 3u);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T sum = T(0);
#pragma unroll
                for (int c = 0; c < NK; c++) {
                    // column c belongs to the item's node g = c/4: this
                    // lane's node, else the first or second other node
                    const int g = c / 4;
                    const T pc = g == a0 ? pown[c % 4]
                                         : ((g < a0 ? g : g - 1) == 0 ? pa[c % 4] : pb[c % 4]);
                    sum = c == 0 ? a[v * NK] * pc : sum + a[v * NK + c] * pc;
                }
                acc[v] = wd == NODE_PAD ? acc[v] : acc[v] + sum;
            }
        }
        if (i < nnode) {
#pragma unroll
            for (int v = 0; v < 4; v++) Q[b + v] = acc[v];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Q[neq] = T(0);
        P[neq] = T(0);
    }
}

// the node rows' values from the blocked values; one thread per (slice, lane)
template <int NK>
__global__ void k_ebe_node_mirror(int nslice, const int *__restrict__ soff,
                                  const int *__restrict__ src, const long long *__restrict__ kpos,
                                  const int *__restrict__ kstride, const double *__restrict__ K64,
                                  const float *__restrict__ K32, double *__restrict__ V64,
                                  float *__restrict__ V32)
{
    const long long nt = (long long)nslice * FG_SELL_W;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < nt;
         t += (long long)gridDim.x * FG_TPB) {
        const int s = (int)(t / FG_SELL_W), lane = (int)(t % FG_SELL_W);
        const int o0 = soff[s], w = soff[s + 1] - o0;
        for (int k = 0; k < w; k++) {
            const long long q = (long long)(o0 + k) * FG_SELL_W + lane;
            const int sv = src[q];
            if (sv < 0) continue;
            constexpr int NN = NK / 4;
            const int e = sv / NN, a0 = sv - NN * e;
            const long long p = kpos[e];
            const int st = kstride[e];
            const long long E = (long long)(o0 + k) * 4 * NK * FG_SELL_W;
            for (int j = 0; j < 4 * NK; j++) {
                const int r = 4 * a0 + j / NK, c = j % NK;
                const long long vi = fg_ebe_vidx(p, st, r * NK + c, fg_ebe_vec(NK));
                V64[node_vpos<double>(E, lane, j)] = K64[vi];
                V32[node_vpos<float>(E, lane, j)] = K32[vi];
            }
        }
    }
}

// AoS item values -> the blocked layout and the node rows; one thread per
// (item, row): the row's k values, written to their blocked positions and (a
// free node) to the entry of its node (npos: the entry's value block plus the
// lane)
template <int NK>
__global__ void k_ebe_node_scatter(long long n, const double *__restrict__ A,
                                   const long long *__restrict__ kpos,
                                   const int *__restrict__ kstride,
                                   const long long *__restrict__ npos,
                                   double *__restrict__ K64, float *__restrict__ K32,
                                   double *__restrict__ V64, float *__restrict__ V32)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const long long e = t / NK;
        const int r = (int)(t - e * NK);
        const long long p = kpos[e];
        if (p < 0) continue;
        const int st = kstride[e];
        double v[NK];
#pragma unroll
        for (int c = 0; c < NK; c++) v[c] = A[t * NK + c];
#pragma unroll
        for (int c = 0; c < NK; c++) {
            const long long d = fg_ebe_vidx(p, st, r * NK + c, fg_ebe_vec(NK));
            K64[d] = v[c];
            K32[d] = (float)v[c];
        }
        const long long np = npos[(NK / 4) * e + r / 4];
        if (np >= 0) {
            const int lane = (int)(np % FG_SELL_W);
            const long long E = np - lane;
#pragma unroll
            for (int c = 0; c < NK; c++) {
                const int j = (r % 4) * NK + c;
                V64[node_vpos<double>(E, lane, j)] = v[c];
                V32[node_vpos<float>(E, lane, j)] = (float)v[c];
            }
        }
    }
}

template <typename T>
int upload(const std::vector<T> &v, T **d)
{
    const size_t n = v.size() ? v.size() : 1;
    FG_HIP(hipMalloc(d, n * sizeof(T)));
    if (!v.empty()) FG_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

}  // namespace

int fg_ebe_node_build(fg_op *op, const int *lm, const long long *kpos)
{
    const int neq = op->neq, nel = op->nitems, NK = op->k, NN = NK / 4, NV = 4 * NK;
    if (op->kind != FG_OP_EBE || (NK != 8 && NK != 12) || neq == 0 || neq >= (1 << 28))
        return FG_OK;
    // node groups: four consecutive equations or all prescribed; every
    // equation in exactly one group
    std::vector<int> owner((size_t)neq, -1);
    for (int e = 0; e < nel; e++) {
        if (kpos[e] < 0) continue;
        for (int a = 0; a < NN; a++) {
            const int *g = lm + (size_t)e * NK + 4 * a;
            if (g[0] == neq && g[1] == neq && g[2] == neq && g[3] == neq) continue;
            if (g[0] < 0 || g[0] > neq - 4) return FG_OK;
            for (int v = 0; v < 4; v++) {
                if (g[v] != g[0] + v) return FG_OK;
                if (owner[g[v]] >= 0 && owner[g[v]] != g[0]) return FG_OK;
                owner[g[v]] = g[0];
            }
        }
    }
    for (int r = 0; r < neq; r++)
        if (owner[r] < 0) return FG_OK;
    // nodes in equation order
    std::vector<int> nbase, nidx((size_t)neq, -1);
    for (int r = 0; r < neq; r += 4) {
        if (owner[r] != r) return FG_OK;
        nidx[r] = (int)nbase.size();
        nbase.push_back(r);
    }
    const int nnode = (int)nbase.size();
    std::vector<int> cnt((size_t)nnode + 1, 0);
    for (int e = 0; e < nel; e++) {
        if (kpos[e] < 0) continue;
        for (int a = 0; a < NN; a++) {
            const int b = lm[(size_t)e * NK + 4 * a];
            if (b != neq) cnt[nidx[b] + 1]++;
        }
    }
    for (int i = 0; i < nnode; i++) cnt[i + 1] += cnt[i];
    const fg_sell_layout L = fg_sell_plan(nnode, cnt.data());
    std::vector<unsigned> word(L.nslot, NODE_PAD);
    std::vector<int> src(L.nslot, -1);
    std::vector<long long> npos((size_t)nel * NN, -1);
    std::vector<int> fill((size_t)nnode, 0);
    for (int e = 0; e < nel; e++) {
        if (kpos[e] < 0) continue;
        const int *le = lm + (size_t)e * NK;
        for (int a = 0; a < NN; a++) {
            const int b = le[4 * a];
            if (b == neq) continue;
            int off[2] = {0, 0}, m = 0;
            for (int a2 = 0; a2 < NN; a2++) {
                if (a2 == a) continue;
                const int b2 = le[4 * a2];
                const long long d = (long long)b2 - b;
                if (b2 != neq && (d == 0 || d < -NOFF_MAX || d > NOFF_MAX)) return FG_OK;
                off[m++] = b2 == neq ? NOFF_SINK : (int)d;
            }
            const int i = nidx[b], k = fill[i]++;
            const long long q = L.slot(i, k);
            word[q] = (unsigned)a | ((unsigned)(off[0] & 0x7fff) << 2) |
                      ((unsigned)(off[1] & 0x7fff) << 17);
            src[q] = NN * e + a;
            npos[(size_t)e * NN + a] = (q - i % FG_SELL_W) * NV + i % FG_SELL_W;
        }
    }
    op->nslice = L.nslice;
    op->nslot = L.nslot;
    op->nnode = nnode;
    FG_CHECK(upload(L.soff, &op->d_soff));
    FG_CHECK(upload(nbase, &op->d_nbase));
    FG_CHECK(upload(word, &op->d_ridx));
    FG_CHECK(upload(src, &op->d_rsrc));
    FG_CHECK(upload(npos, &op->d_rvpos));
    const size_t nv = (size_t)NV * (L.nslot ? L.nslot : 1);
    FG_HIP(hipMalloc(&op->d_sv64, sizeof(double) * nv));
    FG_HIP(hipMalloc(&op->d_sv32, sizeof(float) * nv));
    FG_HIP(hipMemset(op->d_sv64, 0, sizeof(double) * nv));
    FG_HIP(hipMemset(op->d_sv32, 0, sizeof(float) * nv));
    op->rows_stale = 1;
    return FG_OK;
}

int fg_ebe_node_set_values(fg_op *op, const double *A)
{
    fg_ctx *c = op->ctx;
    const long long n = (long long)op->nitems * op->k;
    if (n == 0) return FG_OK;
    if (op->k == 12)
        FG_LAUNCH(k_ebe_node_scatter<12>, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, A,
                  op->d_kpos, op->d_kstride, op->d_rvpos, op->d_K64, op->d_K32, op->d_sv64,
                  op->d_sv32);
    else
        FG_LAUNCH(k_ebe_node_scatter<8>, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, A,
                  op->d_kpos, op->d_kstride, op->d_rvpos, op->d_K64, op->d_K32, op->d_sv64,
                  op->d_sv32);
    op->rows_stale = 0;
    return FG_OK;
}

int fg_ebe_node_refresh(fg_op *op)
{
    if (!op->d_ridx || !op->rows_stale) return FG_OK;
    fg_ctx *c = op->ctx;
    const long long nt = (long long)op->nslice * FG_SELL_W;
    if (op->k == 12)
        FG_LAUNCH(k_ebe_node_mirror<12>, fg_vec_grid(c, nt), FG_TPB, 0, c->stream, op->nslice,
                  op->d_soff, op->d_rsrc, op->d_kpos, op->d_kstride, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    else
        FG_LAUNCH(k_ebe_node_mirror<8>, fg_vec_grid(c, nt), FG_TPB, 0, c->stream, op->nslice,
                  op->d_soff, op->d_rsrc, op->d_kpos, op->d_kstride, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    op->rows_stale = 0;
    return FG_OK;
}

int fg_ebe_node_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    fg_ctx *c = op->ctx;
    FG_CHECK(fg_ebe_node_refresh(op));
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int g = (op->nslice + SPW - 1) / SPW;
#define FG_NODE(T, KK, VV)                                                                     \
    FG_LAUNCH((k_ebe_node_rows<T, KK>), g, FG_TPB, 0, c->stream, op->nslice, op->nnode,       \
              op->d_soff, op->d_nbase, op->d_ridx, (const T *)VV, (T *)P, (T *)Q, op->neq,    \
              skip)
    if (op->k == 12) {
        if (fp32) FG_NODE(float, 12, op->d_sv32);
        else FG_NODE(double, 12, op->d_sv64);
    } else {
        if (fp32) FG_NODE(float, 8, op->d_sv32);
        else FG_NODE(double, 8, op->d_sv64);
    }
#undef FG_NODE
    return FG_OK;
}
