// fg_gmres_fused.hip -- restarted GMRES as one persistent cooperative kernel.
//
// Same algorithm and operation order as fg_gmres.hip (Solvers_and_Preconditioners/
// pgmres.c:4-211), restructured for the MI355X:
//
//  * one cooperative launch runs every restart cycle and every Arnoldi step;
//    the host only launches and reads the result.  All workgroups are
//    co-resident (hipLaunchCooperativeKernel refuses a grid that is not);
//  * each thread owns VPT rows (row = global thread id + v * total threads,
//    so every vector access is coalesced) and keeps its slice of the Arnoldi
//    vector w in registers for the whole modified Gram-Schmidt sweep: step j
//    reads u_{j+1} once from HBM (instead of w, u_j, u_{j+1} and a write of w
//    per j in the one-kernel-per-operation form);
//  * the matrix-vector product is the blocked LDS kernel body of fg_ebe.hip
//    (or the CSR row-block body), the workgroups stride over the blocks; the
//    halo partials of shared dofs are summed by the thread that owns the row
//    when it loads w, so no separate halo pass is needed;
//  * dot products are grid reductions: block partials, then the last
//    workgroup to arrive sums them in block order (deterministic) and
//    releases the others; every workgroup then applies the Givens rotation
//    itself on identical values, so all control flow is uniform;
//  * barriers spin with a bounded count and an abort flag: a stalled grid
//    ends the kernel with an error status instead of hanging.
//
// FP32-IR: the residual matvec, r = B - M^{-1} A X, the Hessenberg least
// squares and the X update are FP64; the basis, the inner matvec and the
// preconditioner are FP32 (dot products accumulate in FP64).
#include <math.h>
#include <string.h>

#include "fg_ebe_body.h"
#include "fg_internal.h"

namespace {

constexpr int KM = FG_KMAX;
constexpr unsigned long long SPIN_LIMIT = 1ull << 24;   // ~1 s of s_sleep(2)

struct Bar {
    unsigned count, gen, abort, pad;
};

struct Res {                 // results written by workgroup 0
    double rho, e0, eps;
    int steps, cycles, status, pad;
};

template <typename TB>
struct Args {
    int neq, kmax, lmax, ir, hist_cap, vpt;
    long long ld;
    double eps_final, inner_tol;
    // blocked element operator
    FgEbeLayout L;
    const double *K64;
    const TB *KB;
    double *hp64;
    TB *hpB;
    const int *rowmap;       // -1: plain row, >= 0: halo index, -2: no element
    const int *hoff;
    // CSR operator
    int nrb;
    const int *rb, *IA, *JA;
    const double *AA64;
    const TB *AAB;
    // preconditioner (Diag) or null
    const double *invD64;
    const TB *invDB;
    // vectors
    const double *B;
    double *X, *W64;
    TB *W, *U;
    // scratch / results
    double *part, *slot;
    Bar *bar;
    Res *res;
    double *rho_hist;
    double *H;               // [G][(KM+1)*KM] per-workgroup Hessenberg copies
};

__device__ __forceinline__ unsigned ld_acq(unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// thread 0 waits for the generation to move past g; false on abort/timeout
__device__ bool bar_wait(Bar *b, unsigned g)
{
    unsigned long long spins = 0;
    while (ld_acq(&b->gen) == g) {
        if (ld_acq(&b->abort)) return false;
        __builtin_amdgcn_s_sleep(2);
        if (++spins > SPIN_LIMIT) {
            atomicExch(&b->abort, 1u);
            return false;
        }
    }
    return true;
}

// grid-wide deterministic sum; every thread returns the total.  ok=false
// after an abort.
__device__ double grid_sum(double v, Bar *bar, double *part, double *slot, bool &ok)
{
    __shared__ double ws[FG_TPB / 64];
    __shared__ int s_last, s_ok;
    __shared__ unsigned s_gen;
    __shared__ double s_tot;
    const int G = gridDim.x;
    double b = fg_block_sum(v, ws);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = b;
        s_gen = ld_acq(&bar->gen);
        __threadfence();
        unsigned t = atomicAdd(&bar->count, 1u);
        s_last = (t == (unsigned)G - 1);
        s_ok = 1;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        double acc = 0.0;
        for (int i = threadIdx.x; i < G; i += FG_TPB)
            acc = acc + __hip_atomic_load(&part[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc = fg_block_sum(acc, ws);
        if (threadIdx.x == 0) {
            __hip_atomic_store(slot, acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            atomicAdd(&bar->gen, 1u);
            s_tot = acc;
        }
    } else if (threadIdx.x == 0) {
        s_ok = bar_wait(bar, s_gen);
        __threadfence();
        s_tot = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    __threadfence();
    ok = s_ok != 0;
    return s_tot;
}

__device__ bool grid_sync(Bar *bar, double *part, double *slot)
{
    bool ok;
    grid_sum(0.0, bar, part, slot, ok);
    return ok;
}

// ---- matvec phase: blocked EBE body, workgroups stride over blocks ---------
template <int K, int EPT, typename T>
__device__ void ebe_phase(const FgEbeLayout &L, const T *__restrict__ Kv,
                          const T *__restrict__ P, T *__restrict__ Q, T *__restrict__ hpart,
                          unsigned char *smem)
{
    const int G = gridDim.x;
    for (int t = blockIdx.x; t < L.nblk; t += G) {
        // logical index t is congruent to blockIdx.x mod 8 (G is a multiple of 8):
        // the remap gives each XCD a contiguous range of blocks
        const int b = fg_xcd_remap(t, L.nblk);
        __syncthreads();   // LDS reuse across blocks
        fg_ebe_block<K, EPT, T>(b, L, Kv, P, Q, hpart, smem);
    }
}

// ---- matvec phase: CSR row blocks ----------------------------------------
constexpr int NNZB = 2048;
template <typename T>
__device__ void csr_phase(int nrb, const int *__restrict__ rb, const int *__restrict__ IA,
                          const int *__restrict__ JA, const T *__restrict__ AA,
                          const T *__restrict__ P, T *__restrict__ Q, unsigned char *smem)
{
    __shared__ T ws[FG_TPB / 64];
    T *s = reinterpret_cast<T *>(smem);
    const int G = gridDim.x;
    for (int t = blockIdx.x; t < nrb; t += G) {
        const int b = fg_xcd_remap(t, nrb);
        const int r0 = rb[b], r1 = rb[b + 1];
        const int j0 = IA[r0], j1 = IA[r1];
        __syncthreads();
        if (j1 - j0 <= NNZB) {
            for (int j = j0 + threadIdx.x; j < j1; j += FG_TPB) s[j - j0] = AA[j] * P[JA[j]];
            __syncthreads();
            for (int r = r0 + threadIdx.x; r < r1; r += FG_TPB) {
                T acc = T(0);
                const int a = IA[r] - j0, e = IA[r + 1] - j0;
                for (int j = a; j < e; j++) acc = acc + s[j];
                Q[r] = acc;
            }
        } else {
            T acc = T(0);
            for (int j = j0 + threadIdx.x; j < j1; j += FG_TPB) acc = acc + AA[j] * P[JA[j]];
            acc = fg_block_sum(acc, ws);
            if (threadIdx.x == 0) Q[r0] = acc;
        }
    }
}

template <int K, typename T>
__device__ __forceinline__ void matvec_phase(const Args<T> &a, const T *Kv, const T *AA,
                                             const T *P, T *Q, T *hpart, unsigned char *smem)
{
    if constexpr (K == 0)
        csr_phase<T>(a.nrb, a.rb, a.IA, a.JA, AA, P, Q, smem);
    else
        ebe_phase<K, fg_ebe_default_ept(K), T>(a.L, Kv, P, Q, hpart, smem);
}

// value of row `row` of a product computed by matvec_phase
template <typename T>
__device__ __forceinline__ T row_value(const int *rowmap, const int *hoff, const T *Q,
                                       const T *hpart, int row)
{
    if (!rowmap) return Q[row];
    const int h = rowmap[row];
    if (h == -1) return Q[row];
    if (h == -2) return T(0);
    int q0 = hoff[h], q1 = hoff[h + 1];
    T s = hpart[q0];
    for (int q = q0 + 1; q < q1; q++) s = s + hpart[q];
    return s;
}

template <int K, typename TB>
__global__ __launch_bounds__(FG_TPB) void k_gmres_fused(Args<TB> A)
{
    const int VPT = A.vpt;
    // dynamic LDS: the matvec phase's block scratch, reused by the vector
    // phases for this workgroup's slice of w ([VPT][256], conflict-free)
    extern __shared__ unsigned char smem[];
    TB *sw = reinterpret_cast<TB *>(smem);
    // this workgroup's copy of the Hessenberg matrix, H[j*KM + i] (every
    // workgroup computes the same values; each keeps its own copy)
    double *sH = A.H + (size_t)blockIdx.x * (KM + 1) * KM;
    __shared__ double s_cs[KM], s_sn[KM], s_e[KM + 1], s_y[KM];
    __shared__ double s_bcast[4];
    const int n = A.neq;
    // rows of this thread: g32 + v*T (32-bit: neq < 2^31), so every vector
    // access is a uniform base plus one per-thread offset
    const int T = gridDim.x * FG_TPB;
    const unsigned g32 = blockIdx.x * FG_TPB + threadIdx.x;
    const long long gt = g32;
    const bool w0 = blockIdx.x == 0 && threadIdx.x == 0;
    double *R64 = A.W64;
    Args<double> A64;   // the FP64 view for the residual matvec
    A64.L = A.L;
    A64.nrb = A.nrb; A64.rb = A.rb; A64.IA = A.IA; A64.JA = A.JA;
    Bar *bar = A.bar;
    bool ok = true;
    int steps = 0, cycles = 0, status = FG_ERR_NOCONV;
    double rho = 0.0, e0 = 0.0;

    for (int l = 0; l < A.lmax && ok; l++) {
        // ---- r = B - M^{-1} A X (FP64) ; e0 = ||r|| ----
        matvec_phase<K, double>(A64, A.K64, A.AA64, A.X, A.W64, A.hp64, smem);
        ok = grid_sync(bar, A.part, A.slot);
        if (!ok) break;
        // r is kept in the thread's own rows of W64 (read back by the same
        // thread after the reduction), not in registers
        double acc = 0.0;
#pragma unroll 4
        for (int v = 0; v < VPT; v++) {
            const unsigned row = g32 + (unsigned)(v * T);
            if (row < (unsigned)n) {
                double q = row_value<double>(A.rowmap, A.hoff, A.W64, A.hp64, (int)row);
                if (A.invD64) q = q * A.invD64[row];
                double rr = A.B[row] - q;
                acc = acc + rr * rr;
                R64[row] = rr;
            }
        }
        e0 = sqrt(grid_sum(acc, bar, A.part, A.slot, ok));
        if (!ok) break;
        if (e0 != e0) {   // NaN residual: a singular or non-finite system
            status = FG_ERR_BREAKDOWN;
            break;
        }
        if (!(e0 > A.eps_final) && (A.ir || e0 == 0.0)) {
            rho = e0;
            status = FG_OK;
            break;
        }
        double eps = A.eps_final;
        if (A.ir && A.inner_tol * e0 > eps) eps = A.inner_tol * e0;
        const double sc0 = 1.0 / e0;
#pragma unroll 4
        for (int v = 0; v < VPT; v++) {
            const unsigned row = g32 + (unsigned)(v * T);
            if (row < (unsigned)n) A.U[row] = (TB)(sc0 * R64[row]);
        }
        if (gt == 0) A.U[n] = TB(0);
        if (threadIdx.x == 0) s_e[0] = e0;
        rho = e0;
        ok = grid_sync(bar, A.part, A.slot);
        if (!ok) break;

        int m = 0;
        for (int i = 0; i < A.kmax; i++) {
            const TB *Ui = A.U + (size_t)i * A.ld;
            matvec_phase<K, TB>(A, A.KB, A.AAB, Ui, A.W, A.hpB, smem);
            ok = grid_sync(bar, A.part, A.slot);
            if (!ok) break;
            // w = M^{-1} (A u_i) ; h_0i = (u_0, w)
            acc = 0.0;
#pragma unroll 4
            for (int v = 0; v < VPT; v++) {
                const unsigned row = g32 + (unsigned)(v * T);
                sw[v * FG_TPB + threadIdx.x] = TB(0);
                if (row < (unsigned)n) {
                    TB q = row_value<TB>(A.rowmap, A.hoff, A.W, A.hpB, (int)row);
                    if (A.invDB) q = q * A.invDB[row];
                    sw[v * FG_TPB + threadIdx.x] = q;
                    acc = acc + (double)A.U[row] * (double)q;
                }
            }
            // h holds the latest reduction, h_{j,i}; every thread has it
            double h = grid_sum(acc, bar, A.part, A.slot, ok);
            if (!ok) break;
            if (threadIdx.x == 0) sH[0 * KM + i] = h;
            // modified Gram-Schmidt, pgmres.c:107-114
            for (int j = 0; j < i; j++) {
                const TB *Uj = A.U + (size_t)j * A.ld;
                const TB *Uj1 = Uj + A.ld;
                const TB a = (TB)(-h);
                acc = 0.0;
#pragma unroll 4
                for (int v = 0; v < VPT; v++) {
                    const unsigned row = g32 + (unsigned)(v * T);
                    if (row < (unsigned)n) {
                        TB x = sw[v * FG_TPB + threadIdx.x] + a * Uj[row];
                        sw[v * FG_TPB + threadIdx.x] = x;
                        acc = acc + (double)Uj1[row] * (double)x;
                    }
                }
                h = grid_sum(acc, bar, A.part, A.slot, ok);
                if (!ok) break;
                if (threadIdx.x == 0) sH[(j + 1) * KM + i] = h;
            }
            if (!ok) break;
            {
                const TB a = (TB)(-h);
                acc = 0.0;
#pragma unroll 4
                for (int v = 0; v < VPT; v++) {
                    const unsigned row = g32 + (unsigned)(v * T);
                    if (row < (unsigned)n) {
                        TB x = sw[v * FG_TPB + threadIdx.x] + a * Ui[row];
                        sw[v * FG_TPB + threadIdx.x] = x;
                        acc = acc + (double)x * (double)x;
                    }
                }
            }
            double hn = sqrt(grid_sum(acc, bar, A.part, A.slot, ok));
            if (!ok) break;
            // Givens update, identical in every workgroup (pgmres.c:124-159)
            if (threadIdx.x == 0) {
                sH[(i + 1) * KM + i] = hn;
                for (int j = 0; j <= i - 1; j++) {
                    double h1 = s_cs[j] * sH[j * KM + i] + s_sn[j] * sH[(j + 1) * KM + i];
                    double h2 = -s_sn[j] * sH[j * KM + i] + s_cs[j] * sH[(j + 1) * KM + i];
                    sH[j * KM + i] = h1;
                    sH[(j + 1) * KM + i] = h2;
                }
                double rr = sqrt(sH[i * KM + i] * sH[i * KM + i] +
                                 sH[(i + 1) * KM + i] * sH[(i + 1) * KM + i]);
                s_cs[i] = sH[i * KM + i] / rr;
                s_sn[i] = sH[(i + 1) * KM + i] / rr;
                sH[i * KM + i] = rr;
                sH[(i + 1) * KM + i] = 0.0;
                s_e[i + 1] = -s_sn[i] * s_e[i];
                s_e[i] = s_cs[i] * s_e[i];
                s_bcast[0] = fabs(s_e[i + 1]);
            }
            __syncthreads();
            rho = s_bcast[0];
            if (w0 && A.rho_hist && steps < A.hist_cap) A.rho_hist[steps] = rho;
            steps++;
            m = i + 1;
            const bool stop = !(rho > eps && i + 1 < A.kmax) || hn == 0.0;
            if (stop) break;
            const TB s = (TB)(1.0 / hn);
            TB *U1 = A.U + (size_t)(i + 1) * A.ld;
#pragma unroll 4
            for (int v = 0; v < VPT; v++) {
                const unsigned row = g32 + (unsigned)(v * T);
                if (row < (unsigned)n) U1[row] = s * sw[v * FG_TPB + threadIdx.x];
            }
            if (gt == 0) U1[n] = TB(0);
            ok = grid_sync(bar, A.part, A.slot);
            if (!ok) break;
        }
        if (!ok) break;
        // back substitution (pgmres.c:165-175), identical in every workgroup
        if (threadIdx.x == 0) {
            int i = m - 1;
            s_y[i] = s_e[i] / sH[i * KM + i];
            for (int j = i - 1; j >= 0; j--) {
                double soma = 0.0;
                for (int q = j + 1; q <= i; q++) soma = soma + sH[j * KM + q] * s_y[q];
                s_y[j] = (s_e[j] - soma) / sH[j * KM + j];
            }
        }
        __syncthreads();
        // X += sum_j u_j y_j (pgmres.c:177-181)
#pragma unroll 4
        for (int v = 0; v < VPT; v++) {
            const unsigned row = g32 + (unsigned)(v * T);
            if (row < (unsigned)n) {
                double x = A.X[row];
                for (int j = 0; j < m; j++) x = x + (double)A.U[(size_t)j * A.ld + row] * s_y[j];
                A.X[row] = x;
            }
        }
        cycles++;
        ok = grid_sync(bar, A.part, A.slot);
        if (!ok) break;
        if (!A.ir && !(rho > A.eps_final)) {
            status = FG_OK;
            break;
        }
    }
    if (w0) {
        A.res->rho = A.ir ? e0 : rho;
        A.res->e0 = e0;
        A.res->steps = steps;
        A.res->cycles = cycles;
        A.res->status = ok ? status : FG_ERR_HIP;
    }
}

template <int K, typename TB>
int launch(fg_ctx *c, Args<TB> &a, int G, size_t shm)
{
    void *args[] = {&a};
    FG_HIP(hipLaunchCooperativeKernel((const void *)k_gmres_fused<K, TB>, dim3(G),
                                      dim3(FG_TPB), args, shm, c->stream));
    return FG_OK;
}

template <int K, typename TB>
int occupancy(int *nb, size_t shm)
{
    FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_gmres_fused<K, TB>, FG_TPB, shm));
    return FG_OK;
}

#define FG_K_SWITCH(TB, CALL)                                                          \
    switch (kk) {                                                                      \
    case 0: CALL(0, TB); break;                                                        \
    case 2: CALL(2, TB); break;                                                        \
    case 3: CALL(3, TB); break;                                                        \
    case 4: CALL(4, TB); break;                                                        \
    default: fg_set_error("fused GMRES: unsupported k=%d", kk); return FG_ERR_ARG;      \
    }

#define FG_CALL_OCC(K, TB) FG_CHECK((occupancy<K, TB>(&nb, shm)))
#define FG_CALL_LAUNCH(K, TB) FG_CHECK((launch<K, TB>(c, a, G, shm)))

template <typename TB>
int run_fused(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
              const fg_gmres_opts *o, double eps, fg_gmres_info *info, int *handled)
{
    *handled = 0;
    const int n = op->neq;
    const int kk = op->kind == FG_OP_CSR ? 0 : op->k;
    if (!(kk == 0 || kk == 2 || kk == 3 || kk == 4)) return FG_OK;
    // the block body is compiled for the default block size of each k
    if (kk != 0 && op->ept != fg_ebe_default_ept(kk)) return FG_OK;
    if (pc && pc->kind != FG_PREC_DIAG) return FG_OK;
    // LDS: the matvec block scratch (EBE: P values, row sums and incidence
    // lists of a block; CSR: NNZB doubles), reused for the workgroup's slice
    // of w (VPT*256 values)
    const size_t shm_mv = kk == 0 ? sizeof(double) * NNZB : op->smem64;
    // grid: 2 (else 1) workgroups per CU, a multiple of 8 (one range per XCD);
    // rows per thread: the power of two that covers neq
    int vpt = 0, nb = 0, G = 0;
    size_t shm = 0;
    for (int wpc = 2; wpc >= 1 && !G; wpc--) {
        int g = (c->num_cu * wpc / 8) * 8;
        if (g < 8) continue;
        long long need = ((long long)n + (long long)g * FG_TPB - 1) / ((long long)g * FG_TPB);
        vpt = 1;
        while (vpt < need) vpt *= 2;
        if (vpt > 32) continue;
        shm = shm_mv;
        size_t shm_w = sizeof(TB) * (size_t)vpt * FG_TPB;
        if (shm_w > shm) shm = shm_w;
        if (shm > 72 * 1024) continue;
        FG_K_SWITCH(TB, FG_CALL_OCC);
        if (nb >= wpc) G = g;
    }
    if (!G) return FG_OK;   // too large: the kernel-per-operation path

    const bool ir = sizeof(TB) == 4;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;
    Args<TB> a;
    memset(&a, 0, sizeof a);
    a.neq = n; a.kmax = o->kmax; a.lmax = o->lmax; a.ir = ir; a.ld = ld; a.vpt = vpt;
    a.eps_final = eps;
    a.inner_tol = ir ? (o->inner_tol > 0.0 ? o->inner_tol : 1e-5) : 0.0;
    a.hist_cap = o->hist_cap;
    void *buf = nullptr;
    size_t szU = sizeof(TB) * ld * (o->kmax + 1);
    size_t szW = sizeof(TB) * ld, szW64 = sizeof(double) * ld;
    size_t szP = sizeof(double) * G, szH = sizeof(double) * (o->hist_cap > 0 ? o->hist_cap : 1);
    size_t off_U = 0, off_W = off_U + szU, off_W64 = off_W + szW, off_P = off_W64 + szW64;
    size_t off_S = off_P + szP, off_B = off_S + 64, off_R = off_B + 64, off_H = off_R + 64;
    size_t off_HH = off_H + szH;
    size_t total = off_HH + sizeof(double) * (size_t)G * (KM + 1) * KM;
    FG_HIP(hipMalloc(&buf, total));
    char *base = (char *)buf;
    a.U = (TB *)(base + off_U);
    a.W = (TB *)(base + off_W);
    a.W64 = (double *)(base + off_W64);
    a.part = (double *)(base + off_P);
    a.slot = (double *)(base + off_S);
    a.bar = (Bar *)(base + off_B);
    a.res = (Res *)(base + off_R);
    a.rho_hist = o->hist_cap > 0 ? (double *)(base + off_H) : nullptr;
    a.H = (double *)(base + off_HH);
    FG_HIP(hipMemsetAsync(buf, 0, total, c->stream));
    if (kk == 0) {
        a.nrb = op->nrb; a.rb = op->d_rb; a.IA = op->d_IA; a.JA = op->d_JA;
        a.AA64 = op->d_AA; a.AAB = ir ? (const TB *)op->d_AA32 : (const TB *)op->d_AA;
    } else {
        a.L = fg_ebe_layout(op);
        a.K64 = op->d_K64; a.KB = ir ? (const TB *)op->d_K32 : (const TB *)op->d_K64;
        a.hp64 = op->d_hpart64; a.hpB = ir ? (TB *)op->d_hpart32 : (TB *)op->d_hpart64;
        a.rowmap = op->d_rowmap; a.hoff = op->d_hoff;
    }
    if (pc) {
        a.invD64 = pc->d_invdiag;
        a.invDB = ir ? (const TB *)pc->d_invdiag32 : (const TB *)pc->d_invdiag;
    }
    a.B = B;
    a.X = X;
    if (!ir) a.W = (TB *)a.W64;
    FG_HIP(hipMemsetAsync(X, 0, sizeof(double) * n, c->stream));
    FG_HIP(hipEventRecord(c->ev0, c->stream));
    FG_K_SWITCH(TB, FG_CALL_LAUNCH);
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    FG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    Res r;
    FG_HIP(hipMemcpy(&r, a.res, sizeof r, hipMemcpyDeviceToHost));
    if (o->rho_hist && o->hist_cap > 0) {
        int nh = r.steps < o->hist_cap ? r.steps : o->hist_cap;
        if (nh > 0)
            FG_HIP(hipMemcpy(o->rho_hist, a.rho_hist, sizeof(double) * nh, hipMemcpyDeviceToHost));
    }
    hipFree(buf);
    if (r.status == FG_ERR_HIP) {
        fg_set_error("fused GMRES: a grid barrier timed out");
        return FG_ERR_HIP;
    }
    memset(info, 0, sizeof *info);
    info->steps = r.steps;
    info->cycles = r.cycles;
    info->rho = r.rho;
    info->eps = eps;
    info->ms = ms;
    info->status = r.status;
    *handled = 1;
    return FG_OK;
}

}  // namespace

// entry from fg_gmres: returns FG_OK with *handled = 0 when the fused path
// does not apply (operator kind, preconditioner or size)
int fg_gmres_fused(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, double eps, fg_gmres_info *info, int *handled)
{
    if (o->precision == FG_GMRES_FP32_IR)
        return run_fused<float>(c, op, pc, B, X, o, eps, info, handled);
    return run_fused<double>(c, op, pc, B, X, o, eps, info, handled);
}
