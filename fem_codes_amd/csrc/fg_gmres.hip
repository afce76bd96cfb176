// fg_gmres.hip -- restarted GMRES on the device.
//
// Reference: Solvers_and_Preconditioners/pgmres.c:4-211 (left preconditioner,
// modified Gram-Schmidt Arnoldi, Givens QR of the Hessenberg column, back
// substitution, X += sum_j u_j y_j).
//
// Device-resident design: the basis, the Hessenberg matrix, the rotations and
// the residual estimate all live in HBM; the host only enqueues kernels and
// reads a few scalars once per restart cycle.  Each Arnoldi step is
//   matvec (+halo) -> [precondition ; partials of (u_0, w)]
//                  -> [h_ji = sum ; w -= h_ji u_j ; partials of (u_{j+1}, w)]
//                  -> [h_ii = sum ; w -= h_ii u_i ; partials of ||w||^2]
//                  -> [h_{i+1,i} = sqrt(sum) ; u_{i+1} = w / h_{i+1,i} ; Givens]
// Reductions are finished by the consumer: each kernel writes one partial
// per block, and every block of the next kernel sums all partials in the
// same fixed order (so every block holds the identical, deterministic
// total).  No fences or atomics sit on the critical path -- the kernel
// boundary orders the partials -- which matters on a multi-XCD device where
// an agent-scope fence writes back and invalidates the L2.  Consecutive
// reductions alternate between two partial buffers.  Every per-step kernel
// returns at once when the cycle has already converged, so a whole cycle is
// enqueued without host synchronisation.  Operation order inside each step
// follows pgmres.c so the FP64 solver reproduces the reference's iterates up
// to dot-product rounding.
//
// FP32-IR: the same cycle with an FP32 basis, FP32 matvec and FP32
// preconditioner; the residual r = B - M^{-1}A X at every restart, the
// Hessenberg least squares (H, Givens, y) and the update of X stay FP64.
// Dot products of FP32 vectors accumulate in FP64.
#include <math.h>
#include <string.h>

#include "fg_internal.h"

namespace {

constexpr int LDH = FG_KMAX + 1;

struct GmState {
    double H[LDH * LDH];   // H[j*LDH + i]
    double cs[FG_KMAX], sn[FG_KMAX], e[FG_KMAX + 1], y[FG_KMAX];
    double rhoh[FG_KMAX];  // rho after step i of the current cycle
    double hc[FG_KMAX + 1];  // CGS: the coefficients h_{j,i} of the current step
    double eps;            // convergence threshold of the current cycle
    double eps_final;      // tol * ||B||
    double rho;            // |e_{i+1}| after the last step
    double scale;          // 1/norm used by the next normalisation
    double e0;             // residual norm at the start of the cycle
    int i;                 // next Arnoldi step of the cycle
    int done;              // cycle finished (converged or i == kmax)
    int steps;             // total Arnoldi steps ("cont")
    int kmax;
    int breakdown;         // e0 == 0 or h_{i+1,i} == 0 with rho > eps
    int conv;              // e0 <= eps_final at the start of a cycle
};

__device__ __forceinline__ double &Hij(GmState *s, int j, int i) { return s->H[j * LDH + i]; }

// ---- grid of the vector kernels and consumer-side reductions ---------------
constexpr int VG_MAX = 1024;   // blocks of every vector kernel of a solve
constexpr int VU = 4;          // entries per thread per sweep (loads in flight)

int vgrid(long long n)
{
    long long g = (n + FG_TPB * VU - 1) / (FG_TPB * VU);
    return (int)(g < 1 ? 1 : g > VG_MAX ? VG_MAX : g);
}

// this block's partial of a grid reduction -> part[blockIdx.x]
__device__ __forceinline__ void put_part(double v, double *__restrict__ part)
{
    __shared__ double ws[FG_TPB / 64];
    v = fg_block_sum(v, ws);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// total of the previous kernel's G partials, summed in the same order by
// every block; returned in every thread
__device__ __forceinline__ double sum_parts(const double *__restrict__ part, int G)
{
    __shared__ double ws[FG_TPB / 64];
    __shared__ double tot;
    double a = 0.0;
    for (int q = threadIdx.x; q < G; q += FG_TPB) a = a + part[q];
    a = fg_block_sum(a, ws);
    if (threadIdx.x == 0) tot = a;
    __syncthreads();
    return tot;
}

// first entry of this thread's sweep; sweeps step by gridDim.x*FG_TPB*VU and
// cover entries base + q*FG_TPB, q < VU
#define VSWEEP(base, n)                                                          \
    for (int base = blockIdx.x * FG_TPB * VU + threadIdx.x; base < (n);          \
         base += gridDim.x * FG_TPB * VU)

// ---- cycle start: r = B - W (W = M^{-1} A X), partials of ||r||^2 ----------
__global__ __launch_bounds__(FG_TPB) void k_res_part(int n, const double *__restrict__ B,
                                                     double *__restrict__ W,
                                                     double *__restrict__ pout)
{
    double acc = 0.0;
    VSWEEP(base, n) {
        double b[VU], w[VU];
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            const int kc = k < n ? k : n - 1;
            b[q] = B[kc];
            w[q] = W[kc];
        }
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) {
                const double r = b[q] - w[q];
                W[k] = r;
                acc = acc + r * r;
            }
        }
    }
    put_part(acc, pout);
}

// e0 = ||r|| and the state of a new cycle (one block)
__global__ __launch_bounds__(FG_TPB) void k_cycle_start(GmState *st, const double *__restrict__ pin,
                                                        int G, double inner_tol)
{
    const double tot = sum_parts(pin, G);
    if (threadIdx.x == 0) {
        double e0 = sqrt(tot);
        st->e[0] = e0;
        st->e0 = e0;
        st->rho = e0;
        st->scale = 1.0 / e0;
        st->i = 0;
        double eps = st->eps_final;
        if (inner_tol > 0.0 && inner_tol * e0 > eps) eps = inner_tol * e0;
        st->eps = eps;
        st->conv = !(e0 > st->eps_final);
        st->breakdown = 0;
        st->done = st->conv;
    }
}

// u = scale * w (dscal(neq, 1.0/e, u), pgmres.c:85 and :121)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_scale_to(int n, const double *__restrict__ W64,
                                                     const TB *__restrict__ Wb,
                                                     TB *__restrict__ U, const GmState *st,
                                                     int honor_done)
{
    if (honor_done && st->done) return;
    const double s = st->scale;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        if (W64) U[k] = (TB)(s * W64[k]);
        else U[k] = (TB)((TB)s * Wb[k]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) U[n] = TB(0);
}

// w = M^{-1} w (Diag, optional) ; partials of h_{0,i} = (u_0, w)
// (pgmres.c:104, :110)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_prec_dot0(int n, const TB *__restrict__ invD,
                                                      TB *__restrict__ W,
                                                      const TB *__restrict__ U0,
                                                      const GmState *st, double *__restrict__ pout)
{
    if (st->done) return;
    double acc = 0.0;
    VSWEEP(base, n) {
        TB w[VU], u[VU], d[VU];
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            const int kc = k < n ? k : n - 1;
            w[q] = W[kc];
            u[q] = U0[kc];
            d[q] = invD ? invD[kc] : TB(1);
        }
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) {
                TB x = w[q];
                if (invD) {
                    x = x * d[q];
                    W[k] = x;
                }
                acc = acc + (double)u[q] * (double)x;
            }
        }
    }
    put_part(acc, pout);
}

// h_{j,i} = sum ; w -= h_{j,i} u_j ; partials of h_{j+1,i} = (u_{j+1}, w)
// (pgmres.c:107-114)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_axpy_dot(int n, int i, int j, TB *__restrict__ W,
                                                     const TB *__restrict__ Uj,
                                                     const TB *__restrict__ Uj1, GmState *st,
                                                     const double *__restrict__ pin, int G,
                                                     double *__restrict__ pout)
{
    if (st->done) return;
    const double h = sum_parts(pin, G);
    if (blockIdx.x == 0 && threadIdx.x == 0) Hij(st, j, i) = h;
    const TB a = (TB)(-h);
    double acc = 0.0;
    VSWEEP(base, n) {
        TB w[VU], u[VU], v[VU];
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            const int kc = k < n ? k : n - 1;
            w[q] = W[kc];
            u[q] = Uj[kc];
            v[q] = Uj1[kc];
        }
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) {
                const TB x = w[q] + a * u[q];
                W[k] = x;
                acc = acc + (double)v[q] * (double)x;
            }
        }
    }
    put_part(acc, pout);
}

// h_{i,i} = sum ; w -= h_{i,i} u_i ; partials of ||w||^2   (pgmres.c:113-118)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_axpy_norm(int n, int i, TB *__restrict__ W,
                                                      const TB *__restrict__ Ui, GmState *st,
                                                      const double *__restrict__ pin, int G,
                                                      double *__restrict__ pout)
{
    if (st->done) return;
    const double h = sum_parts(pin, G);
    if (blockIdx.x == 0 && threadIdx.x == 0) Hij(st, i, i) = h;
    const TB a = (TB)(-h);
    double acc = 0.0;
    VSWEEP(base, n) {
        TB w[VU], u[VU];
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            const int kc = k < n ? k : n - 1;
            w[q] = W[kc];
            u[q] = Ui[kc];
        }
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) {
                const TB x = w[q] + a * u[q];
                W[k] = x;
                acc = acc + (double)x * (double)x;
            }
        }
    }
    put_part(acc, pout);
}

// h_{i+1,i} = ||w|| ; u_{i+1} = w / h_{i+1,i} ; Givens rotation of column i
// and the new residual estimate (pgmres.c:118-163).  Block 0 owns the state:
// when it ends the cycle other blocks may see `done` late or early (each
// block reads it once), which only affects u_{i+1}, a vector the finished
// cycle never reads.
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_scale_givens(int n, int i, const TB *__restrict__ W,
                                                         TB *__restrict__ Un, GmState *st,
                                                         const double *__restrict__ pin, int G)
{
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = __hip_atomic_load(&st->done, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (s_done) return;
    const double hn = sqrt(sum_parts(pin, G));
    const double s = 1.0 / hn;
    VSWEEP(base, n) {
        TB w[VU];
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            w[q] = W[k < n ? k : n - 1];
        }
#pragma unroll
        for (int q = 0; q < VU; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) Un[k] = sizeof(TB) == 8 ? (TB)(s * (double)w[q]) : (TB)((TB)s * w[q]);
        }
    }
    if (blockIdx.x != 0) return;
    __syncthreads();   // every thread of block 0 has read `done`
    if (threadIdx.x == 0) {
        Un[n] = TB(0);
        Hij(st, i + 1, i) = hn;
        st->scale = s;
        for (int j = 0; j <= i - 1; j++) {
            double h1 = st->cs[j] * Hij(st, j, i) + st->sn[j] * Hij(st, j + 1, i);
            double h2 = -st->sn[j] * Hij(st, j, i) + st->cs[j] * Hij(st, j + 1, i);
            Hij(st, j, i) = h1;
            Hij(st, j + 1, i) = h2;
        }
        double r = sqrt(Hij(st, i, i) * Hij(st, i, i) + Hij(st, i + 1, i) * Hij(st, i + 1, i));
        st->cs[i] = Hij(st, i, i) / r;
        st->sn[i] = Hij(st, i + 1, i) / r;
        Hij(st, i, i) = r;
        Hij(st, i + 1, i) = 0.0;
        st->e[i + 1] = -st->sn[i] * st->e[i];
        st->e[i] = st->cs[i] * st->e[i];
        double rho = fabs(st->e[i + 1]);
        st->rho = rho;
        st->rhoh[i] = rho;
        st->steps += 1;
        st->i = i + 1;
        // h_{i+1,i} = 0: the Krylov space is invariant; the reference would
        // divide by zero (pgmres.c:121).  Stop the cycle; flag it unless the
        // residual estimate already meets the tolerance.
        if (hn == 0.0) {
            if (rho > st->eps) st->breakdown = 1;
            st->done = 1;
        }
        if (!(rho > st->eps && i + 1 < st->kmax)) st->done = 1;
    }
}

// ---- classical Gram-Schmidt (FP32-IR inner cycle, FG_ORTH_CGS) ------------
// All coefficients h_{j,i} = (u_j, w), j <= i, come from one pass over w and
// the basis, then w -= sum_j h_{j,i} u_j in a second pass that also forms
// ||w||^2.  Each thread keeps CG_NV entries of w in registers for both
// passes; the grid therefore covers n in one sweep.
constexpr int CG_NV = 16;
constexpr int CG_GMAX = 4096;

int cgs_grid(long long n)
{
    long long g = (n + (long long)FG_TPB * CG_NV - 1) / ((long long)FG_TPB * CG_NV);
    if (g < 1) g = 1;
    return g <= CG_GMAX ? (int)g : -1;
}

// [w = M^{-1} w (Diag)] ; partials of (u_j, w), j = 0..i -> pout[j*G + block]
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_cgs_dots(int n, int i, const TB *__restrict__ invD,
                                                     TB *__restrict__ W,
                                                     const TB *__restrict__ U, long long ld,
                                                     const GmState *st, double *__restrict__ pout)
{
    if (st->done || n <= 0) return;
    __shared__ double red[FG_KMAX + 1][FG_TPB / 64];
    const int G = gridDim.x;
    const int base = blockIdx.x * FG_TPB * CG_NV + threadIdx.x;
    // loads past n read entry n - 1 (issued on every path: exact counter
    // waits) and are zeroed by a select
    TB w[CG_NV];
#pragma unroll
    for (int q = 0; q < CG_NV; q++) {
        const int k = base + q * FG_TPB;
        const TB t = W[k < n ? k : n - 1];
        w[q] = k < n ? t : TB(0);
    }
    if (invD) {
        TB d[CG_NV];
#pragma unroll
        for (int q = 0; q < CG_NV; q++) {
            const int k = base + q * FG_TPB;
            d[q] = invD[k < n ? k : n - 1];
        }
#pragma unroll
        for (int q = 0; q < CG_NV; q++) {
            const int k = base + q * FG_TPB;
            if (k < n) {
                w[q] = w[q] * d[q];
                W[k] = w[q];
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j <= i; j++) {
        const TB *Uj = U + (size_t)j * ld;
        TB u[CG_NV];
#pragma unroll
        for (int q = 0; q < CG_NV; q++) {
            const int k = base + q * FG_TPB;
            u[q] = Uj[k < n ? k : n - 1];
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < CG_NV; q++) acc = acc + (double)u[q] * (double)w[q];
        acc = fg_wave_sum(acc);
        if (lane == 0) red[j][wid] = acc;
    }
    __syncthreads();
    for (int j = threadIdx.x; j <= i; j += FG_TPB) {
        double s = red[j][0];
#pragma unroll
        for (int v = 1; v < FG_TPB / 64; v++) s = s + red[j][v];
        pout[(size_t)j * G + blockIdx.x] = s;
    }
}

// h_{j,i} = sum over the G block partials of coefficient j (one block each)
__global__ __launch_bounds__(FG_TPB) void k_cgs_reduce(int i, const double *__restrict__ pin,
                                                       int G, GmState *st)
{
    if (st->done) return;
    __shared__ double ws[FG_TPB / 64];
    const int j = blockIdx.x;
    double a = 0.0;
    for (int b = threadIdx.x; b < G; b += FG_TPB) a = a + pin[(size_t)j * G + b];
    a = fg_block_sum(a, ws);
    if (threadIdx.x == 0) {
        st->hc[j] = a;
        Hij(st, j, i) = a;
    }
}

// w -= sum_j h_{j,i} u_j (j ascending) ; partials of ||w||^2
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_cgs_update(int n, int i, TB *__restrict__ W,
                                                       const TB *__restrict__ U, long long ld,
                                                       const GmState *st,
                                                       double *__restrict__ pout)
{
    if (st->done || n <= 0) return;
    __shared__ TB a[FG_KMAX + 1];
    for (int j = threadIdx.x; j <= i; j += FG_TPB) a[j] = (TB)(-st->hc[j]);
    __syncthreads();
    const int base = blockIdx.x * FG_TPB * CG_NV + threadIdx.x;
    TB w[CG_NV];
#pragma unroll
    for (int q = 0; q < CG_NV; q++) {
        const int k = base + q * FG_TPB;
        w[q] = W[k < n ? k : n - 1];
    }
    for (int j = 0; j <= i; j++) {
        const TB *Uj = U + (size_t)j * ld;
        const TB aj = a[j];
        TB u[CG_NV];
#pragma unroll
        for (int q = 0; q < CG_NV; q++) {
            const int k = base + q * FG_TPB;
            u[q] = Uj[k < n ? k : n - 1];
        }
#pragma unroll
        for (int q = 0; q < CG_NV; q++) w[q] = w[q] + aj * u[q];
    }
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < CG_NV; q++) {
        const int k = base + q * FG_TPB;
        if (k < n) {
            W[k] = w[q];
            acc = acc + (double)w[q] * (double)w[q];
        }
    }
    put_part(acc, pout);
}

// back substitution (pgmres.c:165-175)
__global__ void k_backsub(GmState *st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int i = st->i - 1;
    if (i < 0) return;
    st->y[i] = st->e[i] / Hij(st, i, i);
    for (int j = i - 1; j >= 0; j--) {
        double soma = 0.0;
        for (int q = j + 1; q <= i; q++) soma = soma + Hij(st, j, q) * st->y[q];
        st->y[j] = (st->e[j] - soma) / Hij(st, j, j);
    }
}

// X += sum_j u_j y_j, basis index outer per entry (pgmres.c:177-181)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_xupdate(int n, double *__restrict__ X,
                                                    const TB *__restrict__ U, long long ld,
                                                    const GmState *st)
{
    const int m = st->i;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        double x = X[k];
        for (int j = 0; j < m; j++) x = x + (double)U[(size_t)j * ld + k] * st->y[j];
        X[k] = x;
    }
}

__global__ void k_state_init(GmState *st, int kmax, double eps_final)
{
    if (threadIdx.x || blockIdx.x) return;
    st->kmax = kmax;
    st->eps_final = eps_final;
    st->eps = eps_final;
    st->steps = 0;
    st->done = 0;
    st->i = 0;
    st->breakdown = 0;
    st->conv = 0;
}

// ||B||^2 into st->tmp-free slot: reuse e0 as scratch
__global__ __launch_bounds__(FG_TPB) void k_norm2(int n, const double *__restrict__ B,
                                                  double *out, double *part, unsigned *ticket)
{
    double acc = 0.0;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB)
        acc = acc + B[k] * B[k];
    double tot;
    if (fg_grid_sum(acc, part, ticket, &tot) && threadIdx.x == 0) *out = tot;
}

struct HostState {
    double rho, e0, eps, eps_final;
    int steps, i, done, breakdown, conv;
    double rhoh[FG_KMAX];
};

int read_state(fg_ctx *c, GmState *st, HostState &h)
{
    static thread_local GmState tmp;
    FG_HIP(hipMemcpyAsync(&tmp, st, sizeof(GmState), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    h.rho = tmp.rho; h.e0 = tmp.e0; h.eps = tmp.eps; h.eps_final = tmp.eps_final;
    h.steps = tmp.steps; h.i = tmp.i; h.done = tmp.done; h.breakdown = tmp.breakdown;
    h.conv = tmp.conv;
    memcpy(h.rhoh, tmp.rhoh, sizeof h.rhoh);
    return FG_OK;
}

// residual W64 = M^{-1} A X, then r = B - W64 and e0
// the two partial buffers of the consumer-side reductions: a producer writes
// out() and flips, the next kernel reads in()
struct Parts {
    double *buf[2];
    int cur;
    int G;
    double *in() const { return buf[cur]; }
    double *out() const { return buf[cur ^ 1]; }
    void flip() { cur ^= 1; }
};

int residual(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X, double *W64,
             GmState *st, double inner_tol, Parts &pp)
{
    int n = op->neq;
    FG_CHECK(fg_op_apply_any(op, 0, X, W64));
    if (pc) FG_CHECK(fg_prec_apply_any(pc, 0, W64, W64));
    FG_LAUNCH(k_res_part, pp.G, FG_TPB, 0, c->stream, n, B, W64, pp.out());
    pp.flip();
    FG_LAUNCH(k_cycle_start, 1, FG_TPB, 0, c->stream, st, (const double *)pp.in(), pp.G,
              inner_tol);
    return FG_OK;
}

// one Arnoldi cycle with basis type TB, enqueued without host sync
template <typename TB>
int cycle(fg_ctx *c, fg_op *op, fg_prec *pc, TB *U, long long ld, TB *W, GmState *st, int i0,
          int i1, Parts &pp, double *cgparts)
{
    int n = op->neq;
    const int g = pp.G;
    const int fp32 = sizeof(TB) == 4;
    const TB *invD = nullptr;
    bool diag = pc && pc->kind == FG_PREC_DIAG;
    if (diag) invD = fp32 ? (const TB *)pc->d_invdiag32 : (const TB *)pc->d_invdiag;
    // the FP32 inner cycle applies the element Jacobi as its diagonal, fused
    // into the first vector pass (the FP64 residual keeps the exact form)
    if (fp32 && pc && pc->kind == FG_PREC_JACOBI && pc->d_jsum32) {
        diag = true;
        invD = (const TB *)pc->d_jsum32;
    }
    const int *skip = &st->done;
    const int gc = cgparts ? cgs_grid(n) : 0;
    for (int i = i0; i < i1; i++) {
        TB *Ui = U + (size_t)i * ld;
        FG_CHECK(fg_op_apply_any(op, fp32, Ui, W, skip));
        if (pc && !diag) FG_CHECK(fg_prec_apply_any(pc, fp32, W, W, skip));
        if (cgparts) {
            // classical Gram-Schmidt: all h_{j,i} from one pass, one update pass
            FG_LAUNCH(k_cgs_dots<TB>, gc, FG_TPB, 0, c->stream, n, i, invD, W, (const TB *)U,
                      ld, (const GmState *)st, cgparts);
            FG_LAUNCH(k_cgs_reduce, i + 1, FG_TPB, 0, c->stream, i, (const double *)cgparts, gc,
                      st);
            FG_LAUNCH(k_cgs_update<TB>, gc, FG_TPB, 0, c->stream, n, i, W, (const TB *)U, ld,
                      (const GmState *)st, pp.out());
            pp.flip();
            FG_LAUNCH(k_scale_givens<TB>, g, FG_TPB, 0, c->stream, n, i, (const TB *)W,
                      U + (size_t)(i + 1) * ld, st, (const double *)pp.in(), gc);
            continue;
        }
        FG_LAUNCH(k_prec_dot0<TB>, g, FG_TPB, 0, c->stream, n, invD, W, (const TB *)U,
                  (const GmState *)st, pp.out());
        pp.flip();
        for (int j = 0; j < i; j++) {
            FG_LAUNCH(k_axpy_dot<TB>, g, FG_TPB, 0, c->stream, n, i, j, W,
                      (const TB *)(U + (size_t)j * ld), (const TB *)(U + (size_t)(j + 1) * ld),
                      st, (const double *)pp.in(), g, pp.out());
            pp.flip();
        }
        FG_LAUNCH(k_axpy_norm<TB>, g, FG_TPB, 0, c->stream, n, i, W, (const TB *)Ui, st,
                  (const double *)pp.in(), g, pp.out());
        pp.flip();
        FG_LAUNCH(k_scale_givens<TB>, g, FG_TPB, 0, c->stream, n, i, (const TB *)W,
                  U + (size_t)(i + 1) * ld, st, (const double *)pp.in(), g);
    }
    return FG_OK;
}

}  // namespace

int fg_gmres_fused(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, double eps, fg_gmres_info *info, int *handled);
bool fg_gmres_small_applies(const fg_op *op, const fg_prec *pc, const fg_gmres_opts *o);
int fg_gmres_small(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, fg_gmres_info *info);

extern "C" int fg_gmres(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                        const fg_gmres_opts *o, fg_gmres_info *info)
{
    FG_ARG(c && op && B && X && o, "ctx/op/B/X/opts");
    FG_ARG(o->kmax >= 1 && o->kmax <= FG_KMAX, "kmax in [1,64]");
    FG_ARG(o->lmax >= 1, "lmax >= 1");
    FG_ARG(o->precision == FG_GMRES_FP64 || o->precision == FG_GMRES_FP32_IR, "precision");
    FG_ARG(o->engine >= FG_ENGINE_AUTO && o->engine <= FG_ENGINE_GRAPH, "engine");
    if (pc && pc->kind == FG_PREC_NONE) pc = nullptr;
    FG_ARG(!pc || (pc->op->neq == op->neq && pc->op->ctx == c),
           "preconditioner built on an operator of another size or context");
    FG_HIP(hipSetDevice(c->device));
    // systems of at most FG_SMALL_MAX equations: the one-workgroup engine
    if (o->engine == FG_ENGINE_SMALL || o->engine == FG_ENGINE_AUTO) {
        const bool ok = fg_gmres_small_applies(op, pc, o);
        if (ok) return fg_gmres_small(c, op, pc, B, X, o, info);
        if (o->engine == FG_ENGINE_SMALL) {
            fg_set_error("the one-workgroup GMRES engine does not apply to this "
                         "operator/preconditioner/size/precision");
            return FG_ERR_ARG;
        }
    }
    const int n = op->neq, kmax = o->kmax;
    const bool ir = o->precision == FG_GMRES_FP32_IR;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;
    const size_t sb = ir ? sizeof(float) : sizeof(double);
    const bool fused = o->engine == FG_ENGINE_FUSED;
    FG_ARG(o->orth == FG_ORTH_MGS || o->orth == FG_ORTH_CGS, "orth");
    FG_ARG(o->orth == FG_ORTH_MGS || ir, "classical Gram-Schmidt applies to the FP32-IR cycle");
    // CGS needs the grid of cgs_grid(n) blocks to cover n in one sweep
    const bool cgs = ir && o->orth == FG_ORTH_CGS && !fused && cgs_grid(n) > 0;

    // device workspace: carved from the context's solver workspace, which
    // persists between calls (no allocation or clearing per solve: every
    // entry of the basis, W and W64 is written before it is read, and the
    // state is reset by k_state_init / k_cycle_start)
    constexpr int PMAX = VG_MAX > CG_GMAX ? VG_MAX : CG_GMAX;   // partials per buffer
    fg_carver cv;
    const size_t o_st = cv.take(sizeof(GmState)), o_W64 = cv.take(sizeof(double) * ld);
    const size_t o_nb2 = cv.take(sizeof(double)), o_parts = cv.take(sizeof(double) * 2 * PMAX);
    const size_t o_cg = cgs ? cv.take(sizeof(double) * (FG_KMAX + 1) * CG_GMAX) : 0;
    const size_t o_U = fused ? 0 : cv.take(sb * ld * (kmax + 1));
    const size_t o_W = fused ? 0 : cv.take(sb * ld);
    char *wsb = nullptr;
    FG_CHECK(fg_ctx_workspace(c, cv.off, (void **)&wsb));
    GmState *st = (GmState *)(wsb + o_st);
    double *W64 = (double *)(wsb + o_W64), *nb2 = (double *)(wsb + o_nb2);
    double *parts = (double *)(wsb + o_parts);
    double *cgbuf = cgs ? (double *)(wsb + o_cg) : nullptr;
    Parts pp{{parts, parts + PMAX}, 0, vgrid(n)};

    if (fused) {
        // eps = tol * ||B|| (pgmres.c:49-53), then the persistent solver
        FG_LAUNCH(k_norm2, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, B, nb2, c->red_part,
                  c->red_ticket);
        double h = 0.0;
        FG_HIP(hipMemcpyAsync(&h, nb2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        const double eps = o->tol * sqrt(h);
        fg_gmres_info loc;
        int handled = 0;
        FG_CHECK(fg_gmres_fused(c, op, pc, B, X, o, eps, &loc, &handled));
        if (!handled) {
            fg_set_error("the fused GMRES engine does not apply to this operator/preconditioner/size");
            return FG_ERR_ARG;
        }
        if (info) {
            *info = loc;
            // true preconditioned residual of the returned X (FP64)
            FG_LAUNCH(k_state_init, 1, 1, 0, c->stream, st, o->kmax, eps);
            HostState h2{};
            FG_CHECK(residual(c, op, pc, B, X, W64, st, 0.0, pp));
            FG_CHECK(read_state(c, st, h2));
            info->true_res = h2.e0;
        }
        return loc.status == FG_ERR_NOCONV ? FG_ERR_NOCONV : loc.status;
    }
    // the multi-kernel engine (FG_ENGINE_AUTO and FG_ENGINE_KERNELS)
    void *U = wsb + o_U, *W = wsb + o_W;

    FG_HIP(hipEventRecord(c->ev0, c->stream));
    // X = 0 (pgmres.c:46) ; eps = tol*||B|| (:49-53)
    FG_HIP(hipMemsetAsync(X, 0, sizeof(double) * n, c->stream));
    FG_LAUNCH(k_norm2, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, B, nb2, c->red_part,
              c->red_ticket);
    double hnb2 = 0.0;
    FG_HIP(hipMemcpyAsync(&hnb2, nb2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    double eps = o->tol * sqrt(hnb2);
    FG_LAUNCH(k_state_init, 1, 1, 0, c->stream, st, kmax, eps);
    double inner = 0.0;
    if (ir) inner = o->inner_tol > 0.0 ? o->inner_tol : 1e-5;

    HostState hs{};
    int cycles = 0, rc = FG_OK;
    const int g = pp.G;
    // graph replay of the cycle: not for exact level-scheduled ILU(0) solves
    // (thousands of launches per apply, each level its own grid)
    const bool use_graph = o->engine == FG_ENGINE_GRAPH && n > 0 &&
                           !(pc && pc->kind == FG_PREC_ILU0 && fg_ilu_sweep_count(pc->ilu) == 0);
    if (use_graph) FG_CHECK(fg_op_prepare(op));
    for (;;) {
        FG_CHECK(residual(c, op, pc, B, X, W64, st, inner, pp));
        FG_CHECK(read_state(c, st, hs));
        if (o->e0_hist && cycles <= o->lmax) o->e0_hist[cycles] = hs.e0;
        // FP64: the reference tests rho only after a cycle (do-while, :184),
        // except that a zero residual (B = 0) returns X = 0 instead of the
        // reference's 0/0 (pgmres.c:85);
        // FP32-IR: converged when the true FP64 residual is below eps.
        if (hs.e0 != hs.e0) { rc = FG_ERR_BREAKDOWN; break; }   // non-finite system
        if (ir && hs.conv) break;
        if (!ir && hs.e0 == 0.0) { hs.rho = 0.0; break; }
        int steps_before = hs.steps;
        // u_0, the kmax Arnoldi steps, the back substitution: every argument
        // is a workspace or operator pointer, every per-solve quantity lives
        // in the device state, so the sequence can be captured and replayed
        auto start = [&]() -> int {
            if (ir)
                FG_LAUNCH(k_scale_to<float>, g, FG_TPB, 0, c->stream, n, W64,
                          (const float *)nullptr, (float *)U, st, 0);
            else
                FG_LAUNCH(k_scale_to<double>, g, FG_TPB, 0, c->stream, n, W64,
                          (const double *)nullptr, (double *)U, st, 0);
            return FG_OK;
        };
        auto steps = [&](int i0, int i1) -> int {
            if (ir)
                return cycle<float>(c, op, pc, (float *)U, ld, (float *)W, st, i0, i1, pp, cgbuf);
            return cycle<double>(c, op, pc, (double *)U, ld, (double *)W, st, i0, i1, pp,
                                 nullptr);
        };
        auto body = [&]() -> int {
            FG_CHECK(start());
            FG_CHECK(steps(0, kmax));
            FG_LAUNCH(k_backsub, 1, 1, 0, c->stream, st);
            return FG_OK;
        };
        if (use_graph) {
            // the cycle writes a partial buffer before it reads one, so it can
            // start at either parity: start at 0, as the capture did
            pp.cur = 0;
            fg_cycle_graph &G = c->cg;
            const unsigned long long cfg = fg_cfg_gen();
            const bool hit = G.ge && G.cfg == cfg && G.ws_gen == c->ws_gen && G.op == op &&
                             G.pc == pc && G.n == n && G.kmax == kmax && G.ir == (int)ir &&
                             G.cgs == (int)cgs;
            if (!hit) {
                fg_cycle_graph_clear(G);
                FG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                const int brc = body();
                hipGraph_t gr = nullptr;
                const hipError_t e = hipStreamEndCapture(c->stream, &gr);
                if (brc != FG_OK) {
                    if (gr) hipGraphDestroy(gr);
                    return brc;
                }
                if (e != hipSuccess) {
                    fg_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
                    return FG_ERR_HIP;
                }
                G.g = gr;
                FG_HIP(hipGraphInstantiate(&G.ge, G.g, nullptr, nullptr, 0));
                G.cfg = cfg;
                G.ws_gen = c->ws_gen;
                G.op = op;
                G.pc = pc;
                G.n = n;
                G.kmax = kmax;
                G.ir = ir;
                G.cgs = cgs;
                G.pp_after = pp.cur;
            }
            FG_HIP(hipGraphLaunch(G.ge, c->stream));
            pp.cur = G.pp_after;
        } else {
            // the steps in chunks: the first as long as this cycle of the
            // previous solve on the context (the corrections of a time loop
            // repeat their cycles' lengths; FIRST0 when there is none), then
            // chunks of CHUNK, reading the device's `done` flag after each, so
            // that few steps past the end of the cycle are launched (those
            // return at once, but still cost a dispatch of every kernel of a
            // step)
            constexpr int CHUNK = 6, FIRST0 = 12;
            FG_CHECK(start());
            const int slot = cycles < FG_CYCLE_MEMO ? cycles : FG_CYCLE_MEMO - 1;
            int first = c->cycle_steps[slot] > 0 ? c->cycle_steps[slot] : FIRST0;
            first = first < kmax ? first : kmax;
            int i0 = 0;
            for (int i1 = first;; i1 = (i0 + CHUNK < kmax ? i0 + CHUNK : kmax)) {
                FG_CHECK(steps(i0, i1));
                i0 = i1;
                if (i0 >= kmax) break;
                int done = 0;
                FG_HIP(hipMemcpyAsync(&done, &st->done, sizeof(int), hipMemcpyDeviceToHost,
                                      c->stream));
                FG_HIP(hipStreamSynchronize(c->stream));
                if (done) break;
            }
            FG_LAUNCH(k_backsub, 1, 1, 0, c->stream, st);
        }
        if (ir)
            FG_LAUNCH(k_xupdate<float>, g, FG_TPB, 0, c->stream, n, X, (const float *)U, ld, st);
        else
            FG_LAUNCH(k_xupdate<double>, g, FG_TPB, 0, c->stream, n, X, (const double *)U, ld,
                      st);
        FG_CHECK(read_state(c, st, hs));
        if (!use_graph) {
            const int slot = cycles < FG_CYCLE_MEMO ? cycles : FG_CYCLE_MEMO - 1;
            c->cycle_steps[slot] = hs.steps - steps_before;
        }
        cycles++;
        if (o->rho_hist && o->hist_cap > 0)
            for (int s = steps_before; s < hs.steps && s < o->hist_cap; s++)
                o->rho_hist[s] = hs.rhoh[s - steps_before];
        if (hs.breakdown) { rc = FG_ERR_BREAKDOWN; break; }
        if (!ir && !(hs.rho > eps)) break;
        if (cycles >= o->lmax) { rc = FG_ERR_NOCONV; break; }
    }
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    FG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ir && !hs.conv && rc == FG_OK) rc = FG_ERR_NOCONV;

    if (info) {
        memset(info, 0, sizeof *info);
        info->steps = hs.steps;
        info->cycles = cycles;
        info->rho = ir ? hs.e0 : hs.rho;
        info->eps = eps;
        info->ms = ms;
        info->status = rc;
        // true preconditioned residual of the returned X (FP64); a converged
        // FP32-IR solve has just computed it for this X (its stopping test)
        if (ir && hs.conv) {
            info->true_res = hs.e0;
        } else {
            HostState h2{};
            FG_CHECK(residual(c, op, pc, B, X, W64, st, 0.0, pp));
            FG_CHECK(read_state(c, st, h2));
            info->true_res = h2.e0;
        }
    }
    return rc == FG_ERR_NOCONV ? FG_ERR_NOCONV : rc;
}
