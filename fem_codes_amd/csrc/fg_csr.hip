// fg_csr.hip -- CSR operator (MatrixVector_Operations/CSR/csrmv.c:3-24).
//
// The product runs on a sliced copy of the rows (fg_sell.h): 64 rows per
// wavefront slice, lane = row, the k-th entries of a slice's rows stored
// side by side, so every JA/AA load is one coalesced wavefront access and
// each lane sums its row in registers (the FP64 values with each lane's
// entries paired, one 16-byte load for two; fg_pair_shift).  The per-row summation order is the
// reference's (csrmv.c:15-16: Q[i]=0, then Q[i] += AA[j]*P[JA[j]] for
// ascending j), so with FMA contraction disabled the FP64 result is bitwise
// the reference's.  The reference-order arrays IA/JA/AA stay the primary
// copy (assembly, scaling, ILU(0), the diagonal setup and the one-workgroup
// solver read them); every write of AA refreshes the sliced values.
//
// A pattern whose slices would carry more than a quarter of padding (rows of
// very different lengths side by side) keeps the row-block ("CSR-stream")
// product instead: blocks of at most 256 rows and FG_CSR_NNZB nonzeros, the
// products staged in LDS, each thread summing one row in slot order; a row
// longer than a block gets a block of its own and a tree sum.
#include <algorithm>
#include <vector>

#include "fg_internal.h"
#include "fg_sell.h"

constexpr int FG_CSR_NNZB = 2048;

// nonzeros each thread stages per row block
constexpr int FG_CSR_NPT = FG_CSR_NNZB / FG_TPB;

// One row block per workgroup.  Latency, not bytes, bounds this product
// (counters: HBM traffic 1.05x the algorithmic bytes), so every thread
// issues all of its JA / AA loads of the block at once, then all of its P
// gathers, before it stores a product: FG_CSR_NPT independent load chains in
// flight per thread instead of one.  The row bounds of the thread's row are
// loaded before the barrier.
template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_csr_stream(
    int nrb, const int *__restrict__ rb, const int *__restrict__ IA,
    const int *__restrict__ JA, const T *__restrict__ AA, const T *__restrict__ P,
    T *__restrict__ Q, int neq, const int *skip)
{
    if (skip && *skip) return;
    __shared__ T s[FG_CSR_NNZB];
    __shared__ T ws[FG_TPB / 64];
    const int b = fg_xcd_remap(blockIdx.x, nrb);
    const int r0 = rb[b], r1 = rb[b + 1];
    const int j0 = IA[r0], j1 = IA[r1];
    // blocks hold at most FG_TPB rows: one row per thread in the sum phase
    const int r = r0 + (int)threadIdx.x;
    int ra = 0, re = 0;
    if (r < r1) {
        ra = IA[r];
        re = IA[r + 1];
    }
    if (j1 - j0 <= FG_CSR_NNZB) {
        // AA and JA are read once per product: non-temporal, so that they do
        // not evict the P entries other row blocks gather again.  Slots past
        // the block's last nonzero load that nonzero again and store to a
        // dummy LDS word: every load is issued on every path (no branch per
        // load), so the memory counter waits only for what each step needs.
        __shared__ T dummy;
        if (j1 > j0) {
            int col[FG_CSR_NPT];
            T a[FG_CSR_NPT], p[FG_CSR_NPT];
#pragma unroll
            for (int q = 0; q < FG_CSR_NPT; q++) {
                const int j = j0 + (int)threadIdx.x + q * FG_TPB;
                const int jc = j < j1 ? j : j1 - 1;
                col[q] = __builtin_nontemporal_load(&JA[jc]);
                a[q] = __builtin_nontemporal_load(&AA[jc]);
            }
#pragma unroll
            for (int q = 0; q < FG_CSR_NPT; q++) p[q] = P[col[q]];
#pragma unroll
            for (int q = 0; q < FG_CSR_NPT; q++) {
                const int j = (int)threadIdx.x + q * FG_TPB;
                *(j < j1 - j0 ? &s[j] : &dummy) = a[q] * p[q];
            }
        }
        __syncthreads();
        if (r < r1) {
            T acc = T(0);
            for (int j = ra - j0; j < re - j0; j++) acc = acc + s[j];
            Q[r] = acc;
        }
    } else {
        T acc = T(0);
        for (int j = j0 + threadIdx.x; j < j1; j += FG_TPB) acc = acc + AA[j] * P[JA[j]];
        acc = fg_block_sum(acc, ws);
        if (threadIdx.x == 0) Q[r0] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) Q[neq] = T(0);
}

// four slices (one per wavefront) per workgroup
template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_csr_sliced(int nslice, const int *__restrict__ soff,
                                                       const int *__restrict__ scol,
                                                       const short *__restrict__ soff16,
                                                       const T *__restrict__ sval,
                                                       const T *__restrict__ P,
                                                       T *__restrict__ Q, int neq, const int *skip)
{
    if (skip && *skip) return;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    if (s < nslice) {
        const int o0 = soff[s], w = soff[s + 1] - o0;
        const long long q0 = (long long)o0 * FG_SELL_W;
        const int r = s * FG_SELL_W + (int)lane;
        // columns as 16-bit offsets from the row where they fit, else indices
        const T acc = soff16
                          ? fg_sell_run(fg_sell_csr16<T>{soff16 + q0, sval + q0, P,
                                                         r < neq ? r : neq - 1, w},
                                        lane, w)
                          : fg_sell_run(fg_sell_csr<T>{scol + q0, sval + q0, P, w}, lane, w);
        if (r < neq) Q[r] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) Q[neq] = T(0);
}

// AA / AA32 (reference order) -> the sliced copies (FP64 entry-paired)
__global__ void k_csr_mirror(int nnz, const long long *__restrict__ pos,
                             const double *__restrict__ AA, const float *__restrict__ AA32,
                             double *__restrict__ s64, float *__restrict__ s32)
{
    for (int j = blockIdx.x * FG_TPB + threadIdx.x; j < nnz; j += gridDim.x * FG_TPB) {
        const long long q = fg_sell_packed_slot(pos[j]);
        s64[q + fg_sell_packed_shift(pos[j])] = AA[j];
        s32[q] = AA32[j];
    }
}

int fg_csr_values_changed(fg_op *op)
{
    if (!op->d_spos || op->nnz == 0) return FG_OK;
    fg_ctx *c = op->ctx;
    FG_LAUNCH(k_csr_mirror, fg_vec_grid(c, op->nnz), FG_TPB, 0, c->stream, (int)op->nnz,
              op->d_spos, op->d_AA, op->d_AA32, op->d_sv64, op->d_sv32);
    return FG_OK;
}

int fg_csr_apply(fg_op *op, int fp32, const void *P, void *Q, const int *skip)
{
    fg_ctx *c = op->ctx;
    if (op->nrb == 0) {
        if (fp32) FG_HIP(hipMemsetAsync((float *)Q + op->neq, 0, sizeof(float), c->stream));
        else FG_HIP(hipMemsetAsync((double *)Q + op->neq, 0, sizeof(double), c->stream));
        return FG_OK;
    }
    if (op->d_spos) {
        constexpr int SPW = FG_TPB / FG_SELL_W;
        const int g = (op->nslice + SPW - 1) / SPW;
        if (fp32)
            FG_LAUNCH(k_csr_sliced<float>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_scol, op->d_soff16, (const float *)op->d_sv32, (const float *)P,
                      (float *)Q, op->neq, skip);
        else
            FG_LAUNCH(k_csr_sliced<double>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_scol, op->d_soff16, (const double *)op->d_sv64, (const double *)P,
                      (double *)Q, op->neq, skip);
        return FG_OK;
    }
    if (fp32)
        FG_LAUNCH(k_csr_stream<float>, op->nrb, FG_TPB, 0, c->stream, op->nrb, op->d_rb,
                  op->d_IA, op->d_JA, (const float *)op->d_AA32, (const float *)P, (float *)Q,
                  op->neq, skip);
    else
        FG_LAUNCH(k_csr_stream<double>, op->nrb, FG_TPB, 0, c->stream, op->nrb, op->d_rb,
                  op->d_IA, op->d_JA, (const double *)op->d_AA, (const double *)P,
                  (double *)Q, op->neq, skip);
    return FG_OK;
}

extern "C" int fg_op_csr_create(fg_ctx *c, int neq, int nnz, const int *IA, const int *JA,
                                fg_op **out)
{
    FG_ARG(c && out && IA && (JA || nnz == 0), "ctx/out/IA/JA");
    FG_ARG(neq >= 0 && nnz >= 0 && IA[0] == 0 && IA[neq] == nnz, "IA[0]=0 and IA[neq]=nnz");
    for (int r = 0; r < neq; r++)
        FG_ARG(IA[r + 1] >= IA[r], "IA must be non-decreasing");
    for (int j = 0; j < nnz; j++)
        FG_ARG(JA[j] >= 0 && JA[j] <= neq, "JA entries must lie in [0, neq]");
    FG_HIP(hipSetDevice(c->device));
    std::vector<int> rb{0};
    int rows = 0, nz = 0;
    for (int r = 0; r < neq; r++) {
        int rn = IA[r + 1] - IA[r];
        if (rows > 0 && (rows == FG_TPB || nz + rn > FG_CSR_NNZB)) {
            rb.push_back(r);
            rows = 0;
            nz = 0;
        }
        rows++;
        nz += rn;
    }
    if (rows > 0) rb.push_back(neq);
    fg_op *op = new fg_op();
    op->ctx = c;
    op->kind = FG_OP_CSR;
    op->neq = neq;
    op->nitems = neq;
    op->nnz = nnz;
    op->nrb = (int)rb.size() - 1;
    op->h_IA.assign(IA, IA + neq + 1);
    op->h_JA.assign(JA, JA + nnz);
    FG_HIP(hipMalloc(&op->d_rb, sizeof(int) * rb.size()));
    FG_HIP(hipMemcpy(op->d_rb, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice));
    FG_HIP(hipMalloc(&op->d_IA, sizeof(int) * (neq + 1)));
    FG_HIP(hipMemcpy(op->d_IA, IA, sizeof(int) * (neq + 1), hipMemcpyHostToDevice));
    FG_HIP(hipMalloc(&op->d_JA, sizeof(int) * (nnz ? nnz : 1)));
    if (nnz) FG_HIP(hipMemcpy(op->d_JA, JA, sizeof(int) * nnz, hipMemcpyHostToDevice));
    FG_HIP(hipMalloc(&op->d_AA, sizeof(double) * (nnz ? nnz : 1)));
    FG_HIP(hipMalloc(&op->d_AA32, sizeof(float) * (nnz ? nnz : 1)));
    FG_HIP(hipMemset(op->d_AA, 0, sizeof(double) * (nnz ? nnz : 1)));
    FG_HIP(hipMemset(op->d_AA32, 0, sizeof(float) * (nnz ? nnz : 1)));
    {
        // the sliced copy, unless its padding exceeds a quarter of the entries
        const fg_sell_layout L = fg_sell_plan(neq, IA);
        // (the product addresses P by a 32-bit byte offset)
        if (nnz > 0 && neq < (1 << 28) && L.nslot <= (long long)nnz + nnz / 4 + 4096) {
            std::vector<int> scol(L.nslot, -1);
            std::vector<short> soff16(L.nslot, FG_SELL_OFF_PAD);
            bool fits16 = true;
            std::vector<long long> spos(nnz);
            for (int r = 0; r < neq; r++) {
                const int s = r / FG_SELL_W, w = L.soff[s + 1] - L.soff[s];
                for (int j = IA[r]; j < IA[r + 1]; j++) {
                    const int k = j - IA[r];
                    const long long q = L.slot(r, k);
                    scol[q] = JA[j];
                    const int d = JA[j] - r;
                    if (d < -32767 || d > 32767) fits16 = false;
                    else soff16[q] = (short)d;
                    spos[j] = fg_sell_pack(q, fg_pair_shift(k, r % FG_SELL_W, w));
                }
            }
            op->nslice = L.nslice;
            op->nslot = L.nslot;
            FG_HIP(hipMalloc(&op->d_soff, sizeof(int) * L.soff.size()));
            FG_HIP(hipMemcpy(op->d_soff, L.soff.data(), sizeof(int) * L.soff.size(),
                             hipMemcpyHostToDevice));
            FG_HIP(hipMalloc(&op->d_scol, sizeof(int) * L.nslot));
            FG_HIP(hipMemcpy(op->d_scol, scol.data(), sizeof(int) * L.nslot,
                             hipMemcpyHostToDevice));
            if (fits16) {
                FG_HIP(hipMalloc(&op->d_soff16, sizeof(short) * L.nslot));
                FG_HIP(hipMemcpy(op->d_soff16, soff16.data(), sizeof(short) * L.nslot,
                                 hipMemcpyHostToDevice));
            }
            FG_HIP(hipMalloc(&op->d_spos, sizeof(long long) * nnz));
            FG_HIP(hipMemcpy(op->d_spos, spos.data(), sizeof(long long) * nnz,
                             hipMemcpyHostToDevice));
            FG_HIP(hipMalloc(&op->d_sv64, sizeof(double) * L.nslot));
            FG_HIP(hipMalloc(&op->d_sv32, sizeof(float) * L.nslot));
            FG_HIP(hipMemset(op->d_sv64, 0, sizeof(double) * L.nslot));
            FG_HIP(hipMemset(op->d_sv32, 0, sizeof(float) * L.nslot));
        }
    }
    if (neq <= FG_SMALL_MAX) {
        // the one-workgroup solver's transposed column lists
        int mi = 0;
        for (int r = 0; r < neq; r++) mi = std::max(mi, IA[r + 1] - IA[r]);
        std::vector<int> iT((size_t)mi * neq, -1);
        for (int r = 0; r < neq; r++)
            for (int j = IA[r]; j < IA[r + 1]; j++) iT[(size_t)(j - IA[r]) * neq + r] = JA[j];
        op->sm_mi = mi;
        FG_HIP(hipMalloc(&op->d_sm_iT, sizeof(int) * (iT.empty() ? 1 : iT.size())));
        if (!iT.empty())
            FG_HIP(hipMemcpy(op->d_sm_iT, iT.data(), sizeof(int) * iT.size(),
                             hipMemcpyHostToDevice));
    }
    *out = op;
    return FG_OK;
}

void fg_csr_free(fg_op *op)
{
    hipFree(op->d_rb); hipFree(op->d_IA); hipFree(op->d_JA); hipFree(op->d_AA); hipFree(op->d_AA32);
    hipFree(op->d_sm_iT);
    hipFree(op->d_soff); hipFree(op->d_scol); hipFree(op->d_soff16); hipFree(op->d_spos);
    hipFree(op->d_sv64);
    hipFree(op->d_sv32);
}

// CSR/Diag_precond_CSR_setup.c:10-14: the last slot of row i with column i
__global__ void k_csr_diag(int neq, const int *__restrict__ IA, const int *__restrict__ JA,
                           const double *__restrict__ AA, double *__restrict__ D)
{
    for (int r = blockIdx.x * FG_TPB + threadIdx.x; r < neq; r += gridDim.x * FG_TPB) {
        double d = D[r];
        for (int j = IA[r]; j < IA[r + 1]; j++)
            if (JA[j] == r) d = AA[j];
        D[r] = d;
    }
}

int fg_csr_diag(fg_op *op, double *D)
{
    fg_ctx *c = op->ctx;
    if (op->neq == 0) return FG_OK;
    FG_LAUNCH(k_csr_diag, fg_vec_grid(c, op->neq), FG_TPB, 0, c->stream, op->neq, op->d_IA,
              op->d_JA, op->d_AA, D);
    return FG_OK;
}
