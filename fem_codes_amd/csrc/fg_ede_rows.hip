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
// order, each with its two coefficients ordered (own dof, other dof), the
// other dof's index, and the equation's index inside its row block (one
// byte) -- 2s + 5 bytes per half-edge, the reference's 4s + 8 per edge.  A
// row block (at most 256 equations, 2048 half-edges) is one workgroup: the
// block's own P entries are staged in LDS, every thread streams eight
// half-edges (coefficient pairs as one vector load) and gathers P of their
// other ends, forms each term t = c_own*p_own + c_other*p_other -- the two
// rounded products of the reference added once, so t is the reference's
// term bit for bit -- into LDS, and each equation then adds its terms in
// edge order starting from 0: Q equals the reference's edemv bit for bit
// (with FMA contraction disabled).  No atomics, no halo pass.
//
// The blocked layout of the same operator (fg_ebe.hip) stays the primary
// copy of the values (assembly, scaling, preconditioner setups and the
// one-workgroup solver read it); the rows are refreshed from it by one
// gather kernel before the first product after any change.
#include <algorithm>
#include <vector>

#include "fg_internal.h"

constexpr int FG_EDE_NNZB = 2048;
constexpr int FG_EDE_NPT = FG_EDE_NNZB / FG_TPB;

template <typename T> struct EdePair;
template <> struct EdePair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct EdePair<float> { typedef float type __attribute__((ext_vector_type(2))); };

// rows' coefficient pairs from the blocked values (edge values at p .. p+3:
// end 0 owns (A0, A1), end 1 owns (A3, A2))
__global__ void k_ede_rows_mirror(long long nh, const int *__restrict__ hsrc,
                                  const double *__restrict__ K64, const float *__restrict__ K32,
                                  double *__restrict__ hv64, float *__restrict__ hv32)
{
    for (long long h = (long long)blockIdx.x * FG_TPB + threadIdx.x; h < nh;
         h += (long long)gridDim.x * FG_TPB) {
        const int sv = hsrc[h];
        const long long p = sv >> 1;
        const int a = sv & 1;
        const long long own = p + (a ? 3 : 0), oth = p + (a ? 2 : 1);
        hv64[2 * h] = K64[own];
        hv64[2 * h + 1] = K64[oth];
        hv32[2 * h] = K32[own];
        hv32[2 * h + 1] = K32[oth];
    }
}

template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_ede_rows(int nrb, const int *__restrict__ rb,
                                                     const int *__restrict__ rIA,
                                                     const int *__restrict__ hcol,
                                                     const uint8_t *__restrict__ hrow,
                                                     const T *__restrict__ hv,
                                                     T *__restrict__ P, T *__restrict__ Q,
                                                     int neq, const int *skip)
{
    if (skip && *skip) return;
    using V2 = typename EdePair<T>::type;
    __shared__ T s[FG_EDE_NNZB];
    __shared__ T sp[FG_TPB];
    __shared__ T ws[FG_TPB / 64];
    __shared__ T dummy;
    const int b = fg_xcd_remap(blockIdx.x, nrb);
    const int r0 = rb[b], r1 = rb[b + 1];   // r1 > r0
    const int j0 = rIA[r0], j1 = rIA[r1];
    const V2 *hv2 = reinterpret_cast<const V2 *>(hv);
    // this thread's equation (clamped: every load issued on every path)
    const int r = r0 + (int)threadIdx.x;
    const int rc = r < r1 ? r : r1 - 1;
    const int ra = rIA[rc], re = rIA[rc + 1];
    const T pr = P[rc];
    if (j1 - j0 <= FG_EDE_NNZB) {
        if (j1 > j0) {
            int col[FG_EDE_NPT];
            V2 c[FG_EDE_NPT];
            uint8_t lr[FG_EDE_NPT];
            T po[FG_EDE_NPT];
#pragma unroll
            for (int q = 0; q < FG_EDE_NPT; q++) {
                const int j = j0 + (int)threadIdx.x + q * FG_TPB;
                const int jc = j < j1 ? j : j1 - 1;
                col[q] = __builtin_nontemporal_load(&hcol[jc]);
                c[q] = __builtin_nontemporal_load(&hv2[jc]);
                lr[q] = __builtin_nontemporal_load(&hrow[jc]);
            }
            // the sink's P reads as 0 (edemv.c:20 sets P[neq] = 0)
#pragma unroll
            for (int q = 0; q < FG_EDE_NPT; q++) {
                const T t = P[col[q]];
                po[q] = col[q] < neq ? t : T(0);
            }
            sp[threadIdx.x] = pr;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < FG_EDE_NPT; q++) {
                const int j = (int)threadIdx.x + q * FG_TPB;
                const T own = c[q].x * sp[lr[q]];
                const T oth = c[q].y * po[q];
                *(j < j1 - j0 ? &s[j] : &dummy) = own + oth;
            }
        }
        __syncthreads();
        if (r < r1) {
            T acc = T(0);
            for (int j = ra - j0; j < re - j0; j++) acc = acc + s[j];
            Q[r] = acc;
        }
    } else {   // one equation with more half-edges than a block holds
        T acc = T(0);
        for (int j = j0 + threadIdx.x; j < j1; j += FG_TPB) {
            const int o = hcol[j];
            const T p2 = o < neq ? P[o] : T(0);
            acc = acc + (hv[2 * j] * pr + hv[2 * j + 1] * p2);
        }
        acc = fg_block_sum(acc, ws);
        if (threadIdx.x == 0) Q[r0] = acc;
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
    std::vector<int> rIA((size_t)neq + 1, 0);
    for (int e = 0; e < nedge; e++) {
        if (kpos[e] < 0) continue;
        for (int a = 0; a < 2; a++) {
            const int r = lm2[2 * (size_t)e + a];
            if (r != neq) rIA[r + 1]++;
        }
    }
    for (int r = 0; r < neq; r++) rIA[r + 1] += rIA[r];
    const long long nh = rIA[neq];
    std::vector<int> hcol(nh), hsrc(nh);
    {
        std::vector<int> fill(rIA.begin(), rIA.end() - 1);
        for (int e = 0; e < nedge; e++) {
            if (kpos[e] < 0) continue;
            for (int a = 0; a < 2; a++) {
                const int r = lm2[2 * (size_t)e + a];
                if (r == neq) continue;
                const int h = fill[r]++;
                hcol[h] = lm2[2 * (size_t)e + 1 - a];
                hsrc[h] = (int)((kpos[e] << 1) | a);
            }
        }
    }
    // row blocks: at most 256 equations and FG_EDE_NNZB half-edges (a longer
    // row gets a block of its own)
    std::vector<int> rb{0};
    std::vector<uint8_t> hrow(nh);
    int rows = 0, nz = 0;
    for (int r = 0; r < neq; r++) {
        const int rn = rIA[r + 1] - rIA[r];
        if (rows > 0 && (rows == FG_TPB || nz + rn > FG_EDE_NNZB)) {
            rb.push_back(r);
            rows = 0;
            nz = 0;
        }
        for (int h = rIA[r]; h < rIA[r + 1]; h++) hrow[h] = (uint8_t)rows;
        rows++;
        nz += rn;
    }
    if (rows > 0) rb.push_back(neq);
    op->nrrb = (int)rb.size() - 1;
    op->nhalf = nh;
    FG_CHECK(upload(rb, &op->d_rrb));
    FG_CHECK(upload(rIA, &op->d_rIA));
    FG_CHECK(upload(hcol, &op->d_hcol));
    FG_CHECK(upload(hsrc, &op->d_hsrc));
    FG_CHECK(upload(hrow, &op->d_hrow));
    FG_HIP(hipMalloc(&op->d_hv64, sizeof(double) * 2 * (nh ? nh : 1)));
    FG_HIP(hipMalloc(&op->d_hv32, sizeof(float) * 2 * (nh ? nh : 1)));
    op->edge_rows = FG_EDGE_AUTO;
    op->rows_stale = 1;
    return FG_OK;
}

int fg_ede_rows_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    fg_ctx *c = op->ctx;
    if (op->rows_stale) {
        if (op->nhalf > 0)
            FG_LAUNCH(k_ede_rows_mirror, fg_vec_grid(c, op->nhalf), FG_TPB, 0, c->stream,
                      op->nhalf, op->d_hsrc, op->d_K64, op->d_K32, op->d_hv64, op->d_hv32);
        op->rows_stale = 0;
    }
    if (op->nrrb == 0) {
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
    if (fp32)
        FG_LAUNCH(k_ede_rows<float>, op->nrrb, FG_TPB, 0, c->stream, op->nrrb, op->d_rrb,
                  op->d_rIA, op->d_hcol, op->d_hrow, (const float *)op->d_hv32, (float *)P,
                  (float *)Q, op->neq, skip);
    else
        FG_LAUNCH(k_ede_rows<double>, op->nrrb, FG_TPB, 0, c->stream, op->nrrb, op->d_rrb,
                  op->d_rIA, op->d_hcol, op->d_hrow, (const double *)op->d_hv64, (double *)P,
                  (double *)Q, op->neq, skip);
    return FG_OK;
}

void fg_ede_rows_free(fg_op *op)
{
    hipFree(op->d_rrb); hipFree(op->d_rIA); hipFree(op->d_hcol); hipFree(op->d_hsrc);
    hipFree(op->d_hrow); hipFree(op->d_hv64); hipFree(op->d_hv32);
}
