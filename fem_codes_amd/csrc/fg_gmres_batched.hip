// fg_gmres_batched.hip -- many independent restarted GMRES solves at once.
//
// Reference: Solvers_and_Preconditioners/pgmres.c:4-211, applied to each of
// nprob independent systems (the reference solves one system per call).
//
// The systems are stored as ONE block-diagonal operator: problem p owns the
// equations [seg[p], seg[p+1]) of every vector, and the element blocks of the
// operator never mix problems.  One matvec launch therefore serves every
// problem.  Each problem keeps its own Krylov basis (the rows of the shared
// basis vectors that belong to it), Hessenberg matrix, rotations, residual
// estimate and stopping test (eps_p = tol * ||B_p||).  The problems advance
// in lockstep through the Arnoldi steps; a problem whose cycle has ended, or
// that has finished, drops out of the vector kernels.
//
// Vector kernels run one workgroup per chunk of BCH entries of one problem;
// a reduction writes one partial per chunk, and every chunk of the next
// kernel sums the partials of its own problem in chunk order (the
// consumer-side scheme of the single-system solver, segmented by problem).
// The first chunk of each problem owns that problem's state (Hessenberg
// column, Givens rotation), as block 0 does for one system.  The per-problem
// stopping logic of the host loop of the single-system solver runs on the
// device (k_b_cycle_start / k_b_cycle_end), so a restart cycle has no host
// synchronisation inside; it can be captured once into a hipGraph and
// replayed every cycle (graph = 1).  The host reads a 16-byte record per
// problem once per cycle.
//
// Precision: FG_GMRES_FP64 (modified Gram-Schmidt, as pgmres.c) or
// FG_GMRES_FP32_IR (FP32 basis / matvec / preconditioner inside, FP64
// residual, Hessenberg least squares and X), as fg_gmres; its inner cycle
// orthogonalises with MGS or, FG_ORTH_CGS, classical Gram-Schmidt (one pass
// for all coefficients of a step, one update pass).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fg_internal.h"
#include "fg_sell.h"

namespace {

constexpr int LDH = FG_KMAX + 1;
constexpr int BU = 4;                       // entries per thread per sweep
constexpr int BCH = FG_TPB * BU * 2;        // entries per chunk (two sweeps)

enum { FIN_NO = 0, FIN_OK = 1, FIN_BREAKDOWN = 2, FIN_NOCONV = 3 };

struct BState {
    double H[LDH * LDH];
    double cs[FG_KMAX], sn[FG_KMAX], e[FG_KMAX + 1], y[FG_KMAX];
    double eps, eps_final, rho, scale, e0;
    int i, done, steps, breakdown, conv, fin, cycles, kmax;
};

// what the host reads back per problem each cycle
struct BRec {
    double e0;
    int fin;
    int steps;
};

// the masked product reads st[p].done with a stride of whole ints
static_assert(sizeof(BState) % sizeof(int) == 0, "BState size in ints");

__device__ __forceinline__ double &Hij(BState *s, int j, int i) { return s->H[j * LDH + i]; }

// chunk table: chunk c covers entries [cs, ce) of problem cp
struct Chunks {
    const int *cs, *ce, *cp;
    const int *pc0, *pc1;   // chunks of problem p: [pc0[p], pc1[p])
};

// this chunk's partial -> part[blockIdx.x]
__device__ __forceinline__ void b_put(double v, double *__restrict__ part)
{
    __shared__ double ws[FG_TPB / 64];
    v = fg_block_sum(v, ws);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// total of problem p's partials in chunk order, in every thread
__device__ __forceinline__ double b_sum(const double *__restrict__ part, int c0, int c1)
{
    __shared__ double ws[FG_TPB / 64];
    __shared__ double tot;
    double a = 0.0;
    for (int q = c0 + threadIdx.x; q < c1; q += FG_TPB) a = a + part[q];
    a = fg_block_sum(a, ws);
    if (threadIdx.x == 0) tot = a;
    __syncthreads();
    return tot;
}

#define BSWEEP(k0, C)                                                            \
    const int cs_ = (C).cs[blockIdx.x], ce_ = (C).ce[blockIdx.x];               \
    for (int k0 = cs_ + threadIdx.x; k0 < ce_; k0 += FG_TPB * BU)

// ||B_p||^2 partials
__global__ __launch_bounds__(FG_TPB) void k_b_norm_part(Chunks C, const double *__restrict__ B,
                                                        double *__restrict__ pout)
{
    double acc = 0.0;
    BSWEEP(k0, C) {
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) acc = acc + B[k] * B[k];
        }
    }
    b_put(acc, pout);
}

// per-problem initial state (one block per problem)
__global__ __launch_bounds__(FG_TPB) void k_b_init(Chunks C, BState *st, int kmax, double tol,
                                                   const double *__restrict__ pin)
{
    const int p = blockIdx.x;
    const double nb2 = b_sum(pin, C.pc0[p], C.pc1[p]);
    if (threadIdx.x) return;
    BState *s = st + p;
    s->eps_final = tol * sqrt(nb2);
    s->eps = s->eps_final;
    s->steps = 0;
    s->done = 0;
    s->i = 0;
    s->breakdown = 0;
    s->conv = 0;
    s->fin = C.pc0[p] == C.pc1[p] ? FIN_OK : FIN_NO;   // an empty problem is solved
    s->cycles = 0;
    s->kmax = kmax;
}

// r = B - W (W = M^{-1} A X), partials of ||r||^2 (finished problems skip)
__global__ __launch_bounds__(FG_TPB) void k_b_res_part(Chunks C, const BState *st,
                                                       const double *__restrict__ B,
                                                       double *__restrict__ W,
                                                       double *__restrict__ pout, int all)
{
    if (!all && st[C.cp[blockIdx.x]].fin) return;
    double acc = 0.0;
    BSWEEP(k0, C) {
        double b[BU], w[BU];
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            const int kc = k < ce_ ? k : ce_ - 1;
            b[q] = B[kc];
            w[q] = W[kc];
        }
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) {
                const double r = b[q] - w[q];
                W[k] = r;
                acc = acc + r * r;
            }
        }
    }
    b_put(acc, pout);
}

// e0 = ||r||, the stopping tests before a cycle and the state of a new cycle
// (fg_gmres's host loop, per problem; one block per problem).  final = 1:
// only record e0 (the true residual after the solve).
__global__ __launch_bounds__(FG_TPB) void k_b_cycle_start(Chunks C, BState *st, BRec *rec,
                                                          const double *__restrict__ pin,
                                                          double inner_tol, int ir, int final)
{
    const int p = blockIdx.x;
    BState *s = st + p;
    const int fin0 = s->fin;
    double tot = 0.0;
    if (final || !fin0) tot = b_sum(pin, C.pc0[p], C.pc1[p]);
    if (threadIdx.x) return;
    if (final) {
        rec[p].e0 = sqrt(tot);
        return;
    }
    s->i = 0;
    if (fin0) {
        s->done = 1;
        rec[p].fin = fin0;
        rec[p].steps = s->steps;
        return;
    }
    const double e0 = sqrt(tot);
    s->e[0] = e0;
    s->e0 = e0;
    s->rho = e0;
    s->scale = 1.0 / e0;
    rec[p].e0 = e0;
    if (!(e0 <= 1.0e300)) {
        // a non-finite residual (NaN or Inf in B, A or X): breakdown, as the
        // single-system solver reports it
        s->fin = FIN_BREAKDOWN;
        s->done = 1;
        rec[p].fin = s->fin;
        rec[p].steps = s->steps;
        return;
    }
    double eps = s->eps_final;
    if (inner_tol > 0.0 && inner_tol * e0 > eps) eps = inner_tol * e0;
    s->eps = eps;
    s->conv = !(e0 > s->eps_final);
    s->breakdown = 0;
    s->done = s->conv;
    // FP32-IR: converged when the true FP64 residual meets eps_final;
    // FP64: a zero residual returns X as it is (fg_gmres)
    if ((ir && s->conv) || (!ir && e0 == 0.0)) {
        if (!ir) s->rho = 0.0;
        s->fin = FIN_OK;
        s->done = 1;
    }
    rec[p].e0 = e0;
    rec[p].fin = s->fin;
    rec[p].steps = s->steps;
}

// the tests after a cycle (one thread per problem)
__global__ void k_b_cycle_end(int np, BState *st, int lmax, int ir)
{
    const int p = blockIdx.x * FG_TPB + threadIdx.x;
    if (p >= np) return;
    BState *s = st + p;
    if (s->fin) return;
    s->cycles += 1;
    if (s->breakdown) s->fin = FIN_BREAKDOWN;
    else if (!ir && !(s->rho > s->eps)) s->fin = FIN_OK;
    else if (s->cycles >= lmax) s->fin = FIN_NOCONV;
}

// u_0 = scale_p * W64 (FP64 residual into the basis type)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_scale0(Chunks C, const BState *st,
                                                     const double *__restrict__ W64,
                                                     TB *__restrict__ U, int sink)
{
    const BState *s = st + C.cp[blockIdx.x];
    if (s->done) return;
    const double sc = s->scale;
    BSWEEP(k0, C) {
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) U[k] = (TB)(sc * W64[k]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) U[sink] = TB(0);
}

// [w = D w] ; partials of (u_0, w)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_dot0(Chunks C, const BState *st,
                                                   const TB *__restrict__ invD, TB *__restrict__ W,
                                                   const TB *__restrict__ U0,
                                                   double *__restrict__ pout)
{
    if (st[C.cp[blockIdx.x]].done) return;
    double acc = 0.0;
    BSWEEP(k0, C) {
        TB w[BU], u[BU], d[BU];
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            const int kc = k < ce_ ? k : ce_ - 1;
            w[q] = W[kc];
            u[q] = U0[kc];
            d[q] = invD ? invD[kc] : TB(1);
        }
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) {
                TB x = w[q];
                if (invD) {
                    x = x * d[q];
                    W[k] = x;
                }
                acc = acc + (double)u[q] * (double)x;
            }
        }
    }
    b_put(acc, pout);
}

// h_{j,i} = sum ; w -= h_{j,i} u_j ; partials of (u_{j+1}, w)  (or, last = 1,
// partials of ||w||^2 with u_j = u_i)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_axpy(Chunks C, BState *st, int i, int j,
                                                   TB *__restrict__ W, const TB *__restrict__ Uj,
                                                   const TB *__restrict__ Uj1,
                                                   const double *__restrict__ pin,
                                                   double *__restrict__ pout, int last)
{
    const int p = C.cp[blockIdx.x];
    BState *s = st + p;
    if (s->done) return;
    const double h = b_sum(pin, C.pc0[p], C.pc1[p]);
    if (blockIdx.x == C.pc0[p] && threadIdx.x == 0) Hij(s, j, i) = h;
    const TB a = (TB)(-h);
    double acc = 0.0;
    BSWEEP(k0, C) {
        TB w[BU], u[BU], v[BU];
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            const int kc = k < ce_ ? k : ce_ - 1;
            w[q] = W[kc];
            u[q] = Uj[kc];
            v[q] = last ? TB(0) : Uj1[kc];
        }
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) {
                const TB x = w[q] + a * u[q];
                W[k] = x;
                acc = last ? acc + (double)x * (double)x : acc + (double)v[q] * (double)x;
            }
        }
    }
    b_put(acc, pout);
}

// Classical Gram-Schmidt for the FP32 inner cycle (FG_ORTH_CGS): every
// chunk keeps its BCH entries of w in registers (8 per thread).
constexpr int BNV = BCH / FG_TPB;

// [w = D w] ; partials of (u_j, w) for j = 0..i -> pout[j * nch + chunk]
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_cgs_dots(Chunks C, const BState *st, int i, int nch,
                                                       const TB *__restrict__ invD,
                                                       TB *__restrict__ W,
                                                       const TB *__restrict__ U, long long ld,
                                                       double *__restrict__ pout)
{
    if (st[C.cp[blockIdx.x]].done) return;
    __shared__ double red[FG_KMAX + 1][FG_TPB / 64];
    const int cs = C.cs[blockIdx.x], ce = C.ce[blockIdx.x];
    const int base = cs + (int)threadIdx.x;
    // loads past the chunk's end read its last entry (issued on every path:
    // exact counter waits) and are zeroed by a select
    TB w[BNV];
#pragma unroll
    for (int q = 0; q < BNV; q++) {
        const int k = base + q * FG_TPB;
        const TB t = W[k < ce ? k : ce - 1];
        w[q] = k < ce ? t : TB(0);
    }
    if (invD) {
        TB d[BNV];
#pragma unroll
        for (int q = 0; q < BNV; q++) {
            const int k = base + q * FG_TPB;
            d[q] = invD[k < ce ? k : ce - 1];
        }
#pragma unroll
        for (int q = 0; q < BNV; q++) {
            const int k = base + q * FG_TPB;
            if (k < ce) {
                w[q] = w[q] * d[q];
                W[k] = w[q];
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j <= i; j++) {
        const TB *Uj = U + (size_t)j * ld;
        TB u[BNV];
#pragma unroll
        for (int q = 0; q < BNV; q++) {
            const int k = base + q * FG_TPB;
            u[q] = Uj[k < ce ? k : ce - 1];
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < BNV; q++) acc = acc + (double)u[q] * (double)w[q];
        acc = fg_wave_sum(acc);
        if (lane == 0) red[j][wid] = acc;
    }
    __syncthreads();
    for (int j = threadIdx.x; j <= i; j += FG_TPB) {
        double v = red[j][0];
#pragma unroll
        for (int q = 1; q < FG_TPB / 64; q++) v = v + red[j][q];
        pout[(size_t)j * nch + blockIdx.x] = v;
    }
}

// h_{j,i} = problem p's chunk partials of coefficient j, in chunk order (one
// block per problem)
__global__ __launch_bounds__(FG_TPB) void k_b_cgs_reduce(Chunks C, BState *st, int i, int nch,
                                                         const double *__restrict__ pin)
{
    const int p = blockIdx.x;
    BState *s = st + p;
    if (s->done) return;
    for (int j = 0; j <= i; j++) {
        const double h = b_sum(pin + (size_t)j * nch, C.pc0[p], C.pc1[p]);
        if (threadIdx.x == 0) Hij(s, j, i) = h;
    }
}

// w -= sum_j h_{j,i} u_j (j ascending) ; partials of ||w||^2
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_cgs_update(Chunks C, const BState *st, int i,
                                                         TB *__restrict__ W,
                                                         const TB *__restrict__ U, long long ld,
                                                         double *__restrict__ pout)
{
    const BState *s = st + C.cp[blockIdx.x];
    if (s->done) return;
    __shared__ TB a[FG_KMAX + 1];
    for (int j = threadIdx.x; j <= i; j += FG_TPB) a[j] = (TB)(-s->H[j * LDH + i]);
    __syncthreads();
    const int cs = C.cs[blockIdx.x], ce = C.ce[blockIdx.x];
    const int base = cs + (int)threadIdx.x;
    TB w[BNV];
#pragma unroll
    for (int q = 0; q < BNV; q++) {
        const int k = base + q * FG_TPB;
        w[q] = W[k < ce ? k : ce - 1];
    }
    for (int j = 0; j <= i; j++) {
        const TB *Uj = U + (size_t)j * ld;
        const TB aj = a[j];
        TB u[BNV];
#pragma unroll
        for (int q = 0; q < BNV; q++) {
            const int k = base + q * FG_TPB;
            u[q] = Uj[k < ce ? k : ce - 1];
        }
#pragma unroll
        for (int q = 0; q < BNV; q++) w[q] = w[q] + aj * u[q];
    }
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < BNV; q++) {
        const int k = base + q * FG_TPB;
        if (k < ce) {
            W[k] = w[q];
            acc = acc + (double)w[q] * (double)w[q];
        }
    }
    b_put(acc, pout);
}

// h_{i+1,i} = ||w|| ; u_{i+1} = w / h_{i+1,i} ; the first chunk of each
// problem applies the Givens rotations of its column (pgmres.c:118-163)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_givens(Chunks C, BState *st, int i,
                                                     const TB *__restrict__ W, TB *__restrict__ Un,
                                                     const double *__restrict__ pin, int sink)
{
    const int p = C.cp[blockIdx.x];
    BState *s = st + p;
    if (blockIdx.x == 0 && threadIdx.x == 0) Un[sink] = TB(0);
    __shared__ int s_done;
    if (threadIdx.x == 0)
        s_done = __hip_atomic_load(&s->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (s_done) return;
    const double hn = sqrt(b_sum(pin, C.pc0[p], C.pc1[p]));
    const double sc = 1.0 / hn;
    BSWEEP(k0, C) {
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) Un[k] = sizeof(TB) == 8 ? (TB)(sc * (double)W[k]) : (TB)((TB)sc * W[k]);
        }
    }
    if (blockIdx.x != C.pc0[p]) return;
    __syncthreads();   // every thread of this block has read `done`
    if (threadIdx.x) return;
    Hij(s, i + 1, i) = hn;
    s->scale = sc;
    for (int j = 0; j <= i - 1; j++) {
        double h1 = s->cs[j] * Hij(s, j, i) + s->sn[j] * Hij(s, j + 1, i);
        double h2 = -s->sn[j] * Hij(s, j, i) + s->cs[j] * Hij(s, j + 1, i);
        Hij(s, j, i) = h1;
        Hij(s, j + 1, i) = h2;
    }
    double r = sqrt(Hij(s, i, i) * Hij(s, i, i) + Hij(s, i + 1, i) * Hij(s, i + 1, i));
    s->cs[i] = Hij(s, i, i) / r;
    s->sn[i] = Hij(s, i + 1, i) / r;
    Hij(s, i, i) = r;
    Hij(s, i + 1, i) = 0.0;
    s->e[i + 1] = -s->sn[i] * s->e[i];
    s->e[i] = s->cs[i] * s->e[i];
    const double rho = fabs(s->e[i + 1]);
    s->rho = rho;
    s->steps += 1;
    s->i = i + 1;
    if (hn == 0.0) {
        if (rho > s->eps) s->breakdown = 1;
        s->done = 1;
    }
    if (!(rho > s->eps && i + 1 < s->kmax)) s->done = 1;
}

// back substitution (pgmres.c:165-175), one thread per problem
__global__ void k_b_backsub(int np, BState *st)
{
    const int p = blockIdx.x * FG_TPB + threadIdx.x;
    if (p >= np) return;
    BState *s = st + p;
    if (s->fin) return;
    const int i = s->i - 1;
    if (i < 0) return;
    s->y[i] = s->e[i] / Hij(s, i, i);
    for (int j = i - 1; j >= 0; j--) {
        double soma = 0.0;
        for (int q = j + 1; q <= i; q++) soma = soma + Hij(s, j, q) * s->y[q];
        s->y[j] = (s->e[j] - soma) / Hij(s, j, j);
    }
}

// X += sum_j u_j y_j (pgmres.c:177-181)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_b_xupdate(Chunks C, const BState *st,
                                                      double *__restrict__ X,
                                                      const TB *__restrict__ U, long long ld)
{
    const BState *s = st + C.cp[blockIdx.x];
    if (s->fin) return;
    const int m = s->i;
    BSWEEP(k0, C) {
#pragma unroll
        for (int q = 0; q < BU; q++) {
            const int k = k0 + q * FG_TPB;
            if (k < ce_) {
                double x = X[k];
                for (int j = 0; j < m; j++) x = x + (double)U[(size_t)j * ld + k] * s->y[j];
                X[k] = x;
            }
        }
    }
}

// zero X of every problem (pgmres.c:46)
__global__ void k_b_zero(long long n, double *X)
{
    for (long long k = (long long)blockIdx.x * FG_TPB + threadIdx.x; k < n;
         k += (long long)gridDim.x * FG_TPB)
        X[k] = 0.0;
}

struct Parts {
    double *buf[2];
    int cur;
    double *in() const { return buf[cur]; }
    double *out() const { return buf[cur ^ 1]; }
    void flip() { cur ^= 1; }
};

// one restart cycle of every active problem: u_0 = r/||r||, kmax lockstep
// Arnoldi steps, back substitution, X update, the post-cycle tests.  No host
// synchronisation and identical arguments every cycle (the partial buffers
// restart at the same parity), so it can be captured once and replayed.
template <typename TB>
int b_cycle(fg_ctx *c, fg_op *op, fg_prec *pc, const Chunks &C, int nch, int np, BState *st,
            TB *U, long long ld, TB *W, const double *W64, double *X, int kmax, int lmax, int ir,
            Parts &pp, double *cgparts, const int *sprob)
{
    const int n = op->neq, fp32 = sizeof(TB) == 4;
    const TB *invD = nullptr;
    bool diag = pc && pc->kind == FG_PREC_DIAG;
    if (diag) invD = fp32 ? (const TB *)pc->d_invdiag32 : (const TB *)pc->d_invdiag;
    // the FP32 inner cycle applies the element Jacobi as its diagonal
    if (fp32 && pc && pc->kind == FG_PREC_JACOBI && pc->d_jsum32) {
        diag = true;
        invD = (const TB *)pc->d_jsum32;
    }
    pp.cur = 0;
    FG_LAUNCH(k_b_scale0<TB>, nch, FG_TPB, 0, c->stream, C, (const BState *)st, W64, U, n);
    for (int i = 0; i < kmax; i++) {
        TB *Ui = U + (size_t)i * ld;
        // the product skips the slices of problems whose cycle has ended
        // (their rows of W are not read again in this cycle)
        if (sprob && fg_ebe_rows_used(op, fp32))
            FG_CHECK(fg_ebe_rows_apply_masked(op, fp32, Ui, W, nullptr, sprob, &st->done,
                                              (int)(sizeof(BState) / sizeof(int))));
        else
            FG_CHECK(fg_op_apply_any(op, fp32, Ui, W, nullptr));
        if (pc && !diag) FG_CHECK(fg_prec_apply_any(pc, fp32, W, W, nullptr));
        if (cgparts) {
            // CGS: all coefficients from one pass, one update pass
            FG_LAUNCH(k_b_cgs_dots<TB>, nch, FG_TPB, 0, c->stream, C, (const BState *)st, i, nch,
                      invD, W, (const TB *)U, ld, cgparts);
            FG_LAUNCH(k_b_cgs_reduce, np, FG_TPB, 0, c->stream, C, st, i, nch,
                      (const double *)cgparts);
            FG_LAUNCH(k_b_cgs_update<TB>, nch, FG_TPB, 0, c->stream, C, (const BState *)st, i, W,
                      (const TB *)U, ld, pp.out());
            pp.flip();
            FG_LAUNCH(k_b_givens<TB>, nch, FG_TPB, 0, c->stream, C, st, i, (const TB *)W,
                      U + (size_t)(i + 1) * ld, (const double *)pp.in(), n);
            continue;
        }
        FG_LAUNCH(k_b_dot0<TB>, nch, FG_TPB, 0, c->stream, C, (const BState *)st, invD, W,
                  (const TB *)U, pp.out());
        pp.flip();
        for (int j = 0; j <= i; j++) {
            const int last = j == i;
            FG_LAUNCH(k_b_axpy<TB>, nch, FG_TPB, 0, c->stream, C, st, i, j, W,
                      (const TB *)(U + (size_t)j * ld),
                      (const TB *)(last ? U : U + (size_t)(j + 1) * ld),
                      (const double *)pp.in(), pp.out(), last);
            pp.flip();
        }
        FG_LAUNCH(k_b_givens<TB>, nch, FG_TPB, 0, c->stream, C, st, i, (const TB *)W,
                  U + (size_t)(i + 1) * ld, (const double *)pp.in(), n);
    }
    const int gp = (np + FG_TPB - 1) / FG_TPB;
    FG_LAUNCH(k_b_backsub, gp, FG_TPB, 0, c->stream, np, st);
    FG_LAUNCH(k_b_xupdate<TB>, nch, FG_TPB, 0, c->stream, C, (const BState *)st, X,
              (const TB *)U, ld);
    FG_LAUNCH(k_b_cycle_end, gp, FG_TPB, 0, c->stream, np, st, lmax, ir);
    return FG_OK;
}

// W64 = M^{-1} A X ; r = B - W64 ; per-problem start of a cycle (or, final,
// the true residual norms)
int b_residual(fg_ctx *c, fg_op *op, fg_prec *pc, const Chunks &C, int nch, int np, BState *st,
               BRec *rec, const double *B, double *X, double *W64, double inner, int ir,
               int final, Parts &pp)
{
    FG_CHECK(fg_op_apply_any(op, 0, X, W64));
    if (pc) FG_CHECK(fg_prec_apply_any(pc, 0, W64, W64));
    pp.cur = 0;
    FG_LAUNCH(k_b_res_part, nch, FG_TPB, 0, c->stream, C, (const BState *)st, B, W64, pp.out(),
              final);
    pp.flip();
    FG_LAUNCH(k_b_cycle_start, np, FG_TPB, 0, c->stream, C, st, rec, (const double *)pp.in(),
              inner, ir, final);
    return FG_OK;
}

}  // namespace

extern "C" int fg_gmres_batched(fg_ctx *c, fg_op *op, fg_prec *pc, int nprob, const int *seg,
                                const double *B, double *X, const fg_gmres_opts *o, int graph,
                                fg_gmres_info *info)
{
    FG_ARG(c && op && seg && B && X && o, "ctx/op/seg/B/X/opts");
    FG_ARG(nprob >= 1, "nprob >= 1");
    FG_ARG(o->kmax >= 1 && o->kmax <= FG_KMAX, "kmax in [1,64]");
    FG_ARG(o->lmax >= 1, "lmax >= 1");
    FG_ARG(o->precision == FG_GMRES_FP64 || o->precision == FG_GMRES_FP32_IR, "precision");
    FG_ARG(o->orth == FG_ORTH_MGS || (o->orth == FG_ORTH_CGS &&
                                       o->precision == FG_GMRES_FP32_IR),
           "orth: MGS, or CGS for the FP32 inner cycle");
    if (pc && pc->kind == FG_PREC_NONE) pc = nullptr;
    FG_ARG(!pc || (pc->op->neq == op->neq && pc->op->ctx == c),
           "preconditioner built on an operator of another size or context");
    FG_ARG(seg[0] == 0 && seg[nprob] == op->neq, "seg[0] = 0 and seg[nprob] = neq");
    for (int p = 0; p < nprob; p++) FG_ARG(seg[p + 1] >= seg[p], "seg ascending");
    // graph replay: the cycle's kernels only (no copies or fills in it)
    FG_ARG(!graph || !pc || pc->kind == FG_PREC_DIAG || pc->kind == FG_PREC_JACOBI,
           "graph replay supports the Diag and element Jacobi preconditioners");
    FG_HIP(hipSetDevice(c->device));
    const int n = op->neq, kmax = o->kmax, np = nprob;
    const bool ir = o->precision == FG_GMRES_FP32_IR;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;
    const size_t sb = ir ? sizeof(float) : sizeof(double);

    // chunk table
    std::vector<int> hcs, hce, hcp, hpc0(np), hpc1(np);
    for (int p = 0; p < np; p++) {
        hpc0[p] = (int)hcs.size();
        for (int k = seg[p]; k < seg[p + 1]; k += BCH) {
            hcs.push_back(k);
            hce.push_back(k + BCH < seg[p + 1] ? k + BCH : seg[p + 1]);
            hcp.push_back(p);
        }
        hpc1[p] = (int)hcs.size();
    }
    const int nch = (int)hcs.size();

    struct Work {
        BState *st = nullptr;
        BRec *rec = nullptr;
        int *tab = nullptr;
        void *U = nullptr, *W = nullptr;
        double *W64 = nullptr, *parts = nullptr;
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        ~Work()
        {
            if (ge) hipGraphExecDestroy(ge);
            if (g) hipGraphDestroy(g);
        }
    } wk;
    // slice -> problem (equations 64s .. 64s+63; -1 when they belong to
    // several problems), for the masked element-row product
    const int nsl = (n + FG_SELL_W - 1) / FG_SELL_W;
    const int ntab = 3 * (nch ? nch : 1) + 2 * np + nsl;
    // buffers carved from the context's solver workspace (kept between calls)
    fg_carver cv;
    const size_t o_st = cv.take(sizeof(BState) * np), o_rec = cv.take(sizeof(BRec) * np);
    const size_t o_tab = cv.take(sizeof(int) * ntab), o_W64 = cv.take(sizeof(double) * ld);
    const size_t o_parts = cv.take(sizeof(double) * 2 * (nch ? nch : 1));
    const size_t o_U = cv.take(sb * ld * (kmax + 1)), o_W = cv.take(sb * ld);
    const bool cgs = o->orth == FG_ORTH_CGS;
    const size_t o_cg = cgs ? cv.take(sizeof(double) * (kmax + 1) * (nch ? nch : 1)) : 0;
    char *wsb = nullptr;
    FG_CHECK(fg_ctx_workspace(c, cv.off, (void **)&wsb));
    wk.st = (BState *)(wsb + o_st);
    wk.rec = (BRec *)(wsb + o_rec);
    wk.tab = (int *)(wsb + o_tab);
    wk.W64 = (double *)(wsb + o_W64);
    wk.parts = (double *)(wsb + o_parts);
    wk.U = wsb + o_U;
    wk.W = wsb + o_W;
    double *cgparts = cgs ? (double *)(wsb + o_cg) : nullptr;
    {
        std::vector<int> tab(ntab, 0);
        int *t = tab.data();
        std::copy(hcs.begin(), hcs.end(), t);
        std::copy(hce.begin(), hce.end(), t + nch);
        std::copy(hcp.begin(), hcp.end(), t + 2 * nch);
        std::copy(hpc0.begin(), hpc0.end(), t + 3 * nch);
        std::copy(hpc1.begin(), hpc1.end(), t + 3 * nch + np);
        int *sp = t + 3 * nch + 2 * np;
        for (int q = 0, p = 0; q < nsl; q++) {
            const int r0 = q * FG_SELL_W, r1 = std::min(n, r0 + FG_SELL_W) - 1;
            while (p < np && seg[p + 1] <= r0) p++;      // problem of r0
            sp[q] = seg[p + 1] > r1 ? p : -1;
        }
        FG_HIP(hipMemcpyAsync(wk.tab, tab.data(), sizeof(int) * ntab, hipMemcpyHostToDevice,
                              c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
    }
    Chunks C{wk.tab, wk.tab + nch, wk.tab + 2 * nch, wk.tab + 3 * nch, wk.tab + 3 * nch + np};
    const int *sprob = wk.tab + 3 * nch + 2 * np;
    FG_HIP(hipMemsetAsync(wk.U, 0, sb * ld * (kmax + 1), c->stream));
    FG_HIP(hipMemsetAsync(wk.W, 0, sb * ld, c->stream));
    FG_HIP(hipMemsetAsync(wk.W64, 0, sizeof(double) * ld, c->stream));
    Parts pp{{wk.parts, wk.parts + (nch ? nch : 1)}, 0};
    const double inner = ir ? (o->inner_tol > 0.0 ? o->inner_tol : 1e-5) : 0.0;

    FG_HIP(hipEventRecord(c->ev0, c->stream));
    // X = 0 (pgmres.c:46) ; eps_p = tol * ||B_p|| (:49-53)
    FG_LAUNCH(k_b_zero, fg_vec_grid(c, n), FG_TPB, 0, c->stream, (long long)n, X);
    if (nch) {
        FG_LAUNCH(k_b_norm_part, nch, FG_TPB, 0, c->stream, C, B, pp.out());
        pp.flip();
    }
    FG_LAUNCH(k_b_init, np, FG_TPB, 0, c->stream, C, wk.st, kmax, o->tol,
              (const double *)pp.in());

    std::vector<BRec> rec(np);
    std::vector<int> cycles(np, 0);
    int iter = 0;
    for (;;) {
        FG_CHECK(b_residual(c, op, pc, C, nch, np, wk.st, wk.rec, B, X, wk.W64, inner, ir, 0, pp));
        FG_HIP(hipMemcpyAsync(rec.data(), wk.rec, sizeof(BRec) * np, hipMemcpyDeviceToHost,
                              c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        int active = 0;
        for (int p = 0; p < np; p++) active += rec[p].fin == FIN_NO;
        if (!active || iter >= o->lmax) break;
        if (graph) {
            if (!wk.ge) {
                FG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int rc = ir ? b_cycle<float>(c, op, pc, C, nch, np, wk.st, (float *)wk.U, ld,
                                             (float *)wk.W, wk.W64, X, kmax, o->lmax, 1, pp,
                                             cgparts, sprob)
                            : b_cycle<double>(c, op, pc, C, nch, np, wk.st, (double *)wk.U, ld,
                                              (double *)wk.W, wk.W64, X, kmax, o->lmax, 0, pp,
                                              nullptr, sprob);
                hipError_t e = hipStreamEndCapture(c->stream, &wk.g);
                FG_CHECK(rc);
                if (e != hipSuccess) {
                    fg_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
                    return FG_ERR_HIP;
                }
                FG_HIP(hipGraphInstantiate(&wk.ge, wk.g, nullptr, nullptr, 0));
            }
            FG_HIP(hipGraphLaunch(wk.ge, c->stream));
        } else if (ir) {
            FG_CHECK(b_cycle<float>(c, op, pc, C, nch, np, wk.st, (float *)wk.U, ld,
                                    (float *)wk.W, wk.W64, X, kmax, o->lmax, 1, pp, cgparts,
                                    sprob));
        } else {
            FG_CHECK(b_cycle<double>(c, op, pc, C, nch, np, wk.st, (double *)wk.U, ld,
                                     (double *)wk.W, wk.W64, X, kmax, o->lmax, 0, pp, nullptr,
                                     sprob));
        }
        iter++;
    }
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    FG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));

    int worst = FG_OK;
    std::vector<BState> hs(np);
    FG_HIP(hipMemcpy(hs.data(), wk.st, sizeof(BState) * np, hipMemcpyDeviceToHost));
    for (int p = 0; p < np; p++) {
        if (hs[p].fin == FIN_BREAKDOWN) worst = FG_ERR_BREAKDOWN;
        else if (hs[p].fin != FIN_OK && worst == FG_OK) worst = FG_ERR_NOCONV;
    }
    if (info) {
        // true preconditioned residual of every returned X (FP64)
        FG_CHECK(b_residual(c, op, pc, C, nch, np, wk.st, wk.rec, B, X, wk.W64, 0.0, ir, 1, pp));
        FG_HIP(hipMemcpyAsync(rec.data(), wk.rec, sizeof(BRec) * np, hipMemcpyDeviceToHost,
                              c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        for (int p = 0; p < np; p++) {
            fg_gmres_info &f = info[p];
            memset(&f, 0, sizeof f);
            f.steps = hs[p].steps;
            f.cycles = hs[p].cycles;
            f.rho = ir ? hs[p].e0 : hs[p].rho;
            f.eps = hs[p].eps_final;
            f.ms = ms;
            f.true_res = rec[p].e0;
            const int fin = hs[p].fin;
            f.status = fin == FIN_OK ? FG_OK : fin == FIN_BREAKDOWN ? FG_ERR_BREAKDOWN
                                                                    : FG_ERR_NOCONV;
        }
    }
    return worst;
}
