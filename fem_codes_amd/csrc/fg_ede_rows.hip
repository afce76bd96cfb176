// fg_ede_rows.hip -- the edge-by-edge product as half-edge rows.
//
// Reference: MatrixVector_Operations/EDE/edemv.c:4-39.  The reference walks
// the edges in order and adds, for edge I with dofs (i1, i2) and values
// A[I][0..3],
//     Q[i1] += A[I][0]*P[i1] + A[I][1]*P[i2];
//     Q[i2] += A[I][2]*P[i1] + A[I][3]*P[i2];
// after Q = 0 and P[neq] = 0.  Equation r therefore receives one term per
// half-edge (edge, end) at r, in edge order.
//
// MI355X design: the half-edges are stored grouped by equation, in edge
// order, each with its two coefficients ordered (own dof, other dof) and the
// other dof's index -- 2s + 4 bytes per half-edge, the reference's 4s + 8
// per edge -- in wavefront slices (fg_sell.h): 64 equations per slice, one
// per lane, the k-th half-edges of a slice's equations side by side.  Each
// lane loads its equation's P once, streams its half-edges (coefficient
// pairs as one vector load) with every load of the row issued at once,
// gathers P of their other ends, forms each term t = c_own*p_own +
// c_other*p_other -- the two rounded products of the reference added once,
// so t is the reference's term bit for bit -- and adds the terms in edge
// order starting from 0: Q equals the reference's edemv bit for bit (with
// FMA contraction disabled).  No LDS, no barrier, no atomics, no halo pass.
//
// The blocked layout of the same operator (fg_ebe.hip) stays the primary
// copy of the values (assembly, scaling, preconditioner setups and the
// one-workgroup solver read it); the rows are refreshed from it by one
// gather kernel before the first product after any change.
#include <algorithm>
#include <vector>

#include "fg_internal.h"
#include "fg_sell.h"

template <typename T> struct EdePair;
template <> struct EdePair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct EdePair<float> { typedef float type __attribute__((ext_vector_type(2))); };

// rows' coefficient pairs from the blocked values (edge values at p .. p+3:
// end 0 owns (A0, A1), end 1 owns (A3, A2)); src = (p << 1) | end, -1 for a
// padding slot
__global__ void k_ede_rows_mirror(long long nslot, const long long *__restrict__ hsrc,
                                  const double *__restrict__ K64, const float *__restrict__ K32,
                                  double *__restrict__ hv64, float *__restrict__ hv32)
{
    for (long long h = (long long)blockIdx.x * FG_TPB + threadIdx.x; h < nslot;
         h += (long long)gridDim.x * FG_TPB) {
        const long long sv = hsrc[h];
        if (sv < 0) continue;
        const long long p = sv >> 1;
        const int a = (int)(sv & 1);
        const long long own = p + (a ? 3 : 0), oth = p + (a ? 2 : 1);
        hv64[2 * h] = K64[own];
        hv64[2 * h + 1] = K64[oth];
        hv32[2 * h] = K32[own];
        hv32[2 * h + 1] = K32[oth];
    }
}

// one half-edge of a sliced equation row: the other end's dof (-1: padding,
// neq: the Dirichlet sink, whose P reads as 0 -- edemv.c:20 sets P[neq] = 0)
template <typename TV>
struct HalfEdgeSlot {
    using T = TV;
    using V2 = typename EdePair<T>::type;
    struct Rec {
        int c;
        V2 v;
        T p;
    };
    const int *__restrict__ col;
    const V2 *__restrict__ val;
    const T *__restrict__ P;
    T pown;
    int neq;
    // col and val start at the lane-0 slot of the wavefront's slice
    // (wave-uniform bases), q = k*64 + lane
    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec r;
        r.c = __builtin_nontemporal_load(&col[q]);
        r.p = T(0);
        return r;
    }
    __device__ __forceinline__ void stream(Rec &r, unsigned q) const
    {
        r.v = __builtin_nontemporal_load(&val[q]);
    }
    __device__ __forceinline__ void gather(Rec &r) const
    {
        // a 32-bit byte offset from P (rows are built for neq < 2^28)
        const unsigned b = (unsigned)(r.c < 0 ? 0 : r.c) * (unsigned)sizeof(T);
        const T t = *(const T *)((const char *)P + b);
        r.p = r.c < neq ? t : T(0);
    }
    __device__ __forceinline__ T term(const Rec &r) const
    {
        const T own = r.v.x * pown;
        const T oth = r.v.y * r.p;
        return r.c < 0 ? T(0) : own + oth;
    }
    __device__ __forceinline__ T add(T acc, const Rec &r) const { return acc + term(r); }
};

// four slices (one per wavefront) per workgroup
template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_ede_rows(int nslice, const int *__restrict__ soff,
                                                     const int *__restrict__ scol,
                                                     const T *__restrict__ sval,
                                                     T *__restrict__ P, T *__restrict__ Q,
                                                     int neq, const int *skip)
{
    if (skip && *skip) return;
    using V2 = typename EdePair<T>::type;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    if (s < nslice) {
        const int r = s * FG_SELL_W + (int)lane;
        // lanes past the last equation read the sink's P; all their slots
        // are padding
        const T pown = P[r < neq ? r : neq];
        const int o0 = soff[s];
        const long long q0 = (long long)o0 * FG_SELL_W;
        const HalfEdgeSlot<T> row{scol + q0, reinterpret_cast<const V2 *>(sval) + q0, P, pown,
                                  neq};
        const T acc = fg_sell_run(row, lane, soff[s + 1] - o0);
        if (r < neq) Q[r] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Q[neq] = T(0);
        P[neq] = T(0);
    }
}

template <typename T>
static int upload(const std::vector<T> &v, T **d)
{
    size_t n = v.size() ? v.size() : 1;
    FG_HIP(hipMalloc(d, n * sizeof(T)));
    if (!v.empty()) FG_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

int fg_ede_rows_build(fg_op *op, const int *lm2, const long long *kpos)
{
    const int neq = op->neq, nedge = op->nitems;
    // the product addresses P by a 32-bit byte offset
    if (neq >= (1 << 28)) return FG_OK;
    std::vector<int> rIA((size_t)neq + 1, 0);
    for (int e = 0; e < nedge; e++) {
        if (kpos[e] < 0) continue;
        for (int a = 0; a < 2; a++) {
            const int r = lm2[2 * (size_t)e + a];
            if (r != neq) rIA[r + 1]++;
        }
    }
    for (int r = 0; r < neq; r++) rIA[r + 1] += rIA[r];
    const fg_sell_layout L = fg_sell_plan(neq, rIA.data());
    std::vector<int> scol(L.nslot, -1);
    std::vector<long long> ssrc(L.nslot, -1);
    {
        // each equation's half-edges in edge order
        std::vector<int> fill((size_t)neq, 0);
        for (int e = 0; e < nedge; e++) {
            if (kpos[e] < 0) continue;
            for (int a = 0; a < 2; a++) {
                const int r = lm2[2 * (size_t)e + a];
                if (r == neq) continue;
                const long long q = L.slot(r, fill[r]++);
                scol[q] = lm2[2 * (size_t)e + 1 - a];
                ssrc[q] = (kpos[e] << 1) | a;
            }
        }
    }
    op->nslice = L.nslice;
    op->nslot = L.nslot;
    op->nhalf = rIA[neq];
    FG_CHECK(upload(L.soff, &op->d_soff));
    FG_CHECK(upload(scol, &op->d_scol));
    FG_CHECK(upload(ssrc, &op->d_hsrc));
    const size_t nv = 2 * (size_t)(L.nslot ? L.nslot : 1);
    FG_HIP(hipMalloc(&op->d_sv64, sizeof(double) * nv));
    FG_HIP(hipMalloc(&op->d_sv32, sizeof(float) * nv));
    FG_HIP(hipMemset(op->d_sv64, 0, sizeof(double) * nv));
    FG_HIP(hipMemset(op->d_sv32, 0, sizeof(float) * nv));
    op->rows_stale = 1;
    return FG_OK;
}

int fg_ede_rows_refresh(fg_op *op)
{
    if (!op->d_hsrc || !op->rows_stale) return FG_OK;
    fg_ctx *c = op->ctx;
    if (op->nslot > 0)
        FG_LAUNCH(k_ede_rows_mirror, fg_vec_grid(c, op->nslot), FG_TPB, 0, c->stream,
                  op->nslot, op->d_hsrc, op->d_K64, op->d_K32, op->d_sv64, op->d_sv32);
    op->rows_stale = 0;
    return FG_OK;
}

int fg_ede_rows_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    fg_ctx *c = op->ctx;
    FG_CHECK(fg_ede_rows_refresh(op));
    if (op->nslice == 0) {
        // no equation: Q[neq] = P[neq] = 0
        if (fp32) {
            FG_HIP(hipMemsetAsync((float *)Q + op->neq, 0, sizeof(float), c->stream));
            FG_HIP(hipMemsetAsync((float *)P + op->neq, 0, sizeof(float), c->stream));
        } else {
            FG_HIP(hipMemsetAsync((double *)Q + op->neq, 0, sizeof(double), c->stream));
            FG_HIP(hipMemsetAsync((double *)P + op->neq, 0, sizeof(double), c->stream));
        }
        return FG_OK;
    }
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int g = (op->nslice + SPW - 1) / SPW;
    if (fp32)
        FG_LAUNCH(k_ede_rows<float>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                  op->d_scol, (const float *)op->d_sv32, (float *)P, (float *)Q, op->neq, skip);
    else
        FG_LAUNCH(k_ede_rows<double>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                  op->d_scol, (const double *)op->d_sv64, (double *)P, (double *)Q, op->neq,
                  skip);
    return FG_OK;
}

void fg_ede_rows_free(fg_op *op)
{
    hipFree(op->d_soff); hipFree(op->d_scol); hipFree(op->d_hsrc);
    hipFree(op->d_sv64); hipFree(op->d_sv32);
}
