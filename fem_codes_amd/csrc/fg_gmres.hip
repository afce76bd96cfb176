// fg_gmres.hip -- restarted GMRES on the device.
//
// Reference: Solvers_and_Preconditioners/pgmres.c:4-211 (left preconditioner,
// modified Gram-Schmidt Arnoldi, Givens QR of the Hessenberg column, back
// substitution, X += sum_j u_j y_j).
//
// Device-resident design: the basis, the Hessenberg matrix, the rotations and
// the residual estimate all live in HBM; the host only enqueues kernels and
// reads a few scalars once per restart cycle.  Each Arnoldi step is
//   matvec (+halo) -> [precondition + h_0i]        (fused)
//                  -> [w -= h_ji u_j ; h_{j+1,i}] (fused, j = 0..i-1)
//                  -> [w -= h_ii u_i ; ||w|| ; Givens update in the last block]
//                  -> [u_{i+1} = w / h_{i+1,i}]
// with deterministic grid reductions (block partials summed in block order
// by the last block to arrive).  Every per-step kernel returns at once when
// the cycle has already converged, so a whole cycle is enqueued without host
// synchronisation.  Operation order inside each step follows pgmres.c so the
// FP64 solver reproduces the reference's iterates up to dot-product rounding.
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

// ---- cycle start: r = B - W (W = M^{-1} A X), e0 = ||r|| -------------------
__global__ __launch_bounds__(FG_TPB) void k_res_init(int n, const double *__restrict__ B,
                                                     double *__restrict__ W, GmState *st,
                                                     double *part, unsigned *ticket,
                                                     double inner_tol)
{
    double acc = 0.0;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        double r = B[k] - W[k];
        W[k] = r;
        acc = acc + r * r;
    }
    double tot;
    if (fg_grid_sum(acc, part, ticket, &tot) && threadIdx.x == 0) {
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

// w = M^{-1} w (Diag, optional) ; h_{0,i} = (u_0, w)   (pgmres.c:104, :110)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_prec_dot0(int n, const TB *__restrict__ invD,
                                                      TB *__restrict__ W,
                                                      const TB *__restrict__ U0, GmState *st,
                                                      double *part, unsigned *ticket)
{
    if (st->done) return;
    double acc = 0.0;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        TB w = W[k];
        if (invD) {
            w = w * invD[k];
            W[k] = w;
        }
        acc = acc + (double)U0[k] * (double)w;
    }
    double tot;
    if (fg_grid_sum(acc, part, ticket, &tot) && threadIdx.x == 0) Hij(st, 0, st->i) = tot;
}

// w -= h_{j,i} u_j ; h_{j+1,i} = (u_{j+1}, w)   (pgmres.c:107-114)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_axpy_dot(int n, int j, TB *__restrict__ W,
                                                     const TB *__restrict__ Uj,
                                                     const TB *__restrict__ Uj1, GmState *st,
                                                     double *part, unsigned *ticket)
{
    if (st->done || j >= st->i) return;
    const int i = st->i;
    const TB a = (TB)(-Hij(st, j, i));
    double acc = 0.0;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        TB w = W[k] + a * Uj[k];
        W[k] = w;
        acc = acc + (double)Uj1[k] * (double)w;
    }
    double tot;
    if (fg_grid_sum(acc, part, ticket, &tot) && threadIdx.x == 0) Hij(st, j + 1, i) = tot;
}

// w -= h_{i,i} u_i ; h_{i+1,i} = ||w|| ; Givens (pgmres.c:113-163)
template <typename TB>
__global__ __launch_bounds__(FG_TPB) void k_axpy_norm_givens(int n, TB *__restrict__ W,
                                                             const TB *__restrict__ Ui,
                                                             GmState *st, double *part,
                                                             unsigned *ticket)
{
    if (st->done) return;
    const int i = st->i;
    const TB a = (TB)(-Hij(st, i, i));
    double acc = 0.0;
    for (int k = blockIdx.x * FG_TPB + threadIdx.x; k < n; k += gridDim.x * FG_TPB) {
        TB w = W[k] + a * Ui[k];
        W[k] = w;
        acc = acc + (double)w * (double)w;
    }
    double tot;
    if (fg_grid_sum(acc, part, ticket, &tot) && threadIdx.x == 0) {
        double hn = sqrt(tot);
        Hij(st, i + 1, i) = hn;
        st->scale = 1.0 / hn;
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
int residual(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X, double *W64,
             GmState *st, double inner_tol)
{
    int n = op->neq;
    FG_CHECK(fg_op_apply_any(op, 0, X, W64));
    if (pc) FG_CHECK(fg_prec_apply_any(pc, 0, W64, W64));
    FG_LAUNCH(k_res_init, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, B, W64, st,
              c->red_part, c->red_ticket, inner_tol);
    return FG_OK;
}

// one Arnoldi cycle with basis type TB, enqueued without host sync
template <typename TB>
int cycle(fg_ctx *c, fg_op *op, fg_prec *pc, TB *U, long long ld, TB *W, GmState *st, int kmax)
{
    int n = op->neq;
    int g = fg_vec_grid(c, n);
    const int fp32 = sizeof(TB) == 4;
    const TB *invD = nullptr;
    bool diag = pc && pc->kind == FG_PREC_DIAG;
    if (diag) invD = fp32 ? (const TB *)pc->d_invdiag32 : (const TB *)pc->d_invdiag;
    const int *skip = &st->done;
    for (int i = 0; i < kmax; i++) {
        TB *Ui = U + (size_t)i * ld;
        FG_CHECK(fg_op_apply_any(op, fp32, Ui, W, skip));
        if (pc && !diag) FG_CHECK(fg_prec_apply_any(pc, fp32, W, W, skip));
        FG_LAUNCH(k_prec_dot0<TB>, g, FG_TPB, 0, c->stream, n, invD, W, U, st, c->red_part,
                  c->red_ticket);
        for (int j = 0; j < i; j++)
            FG_LAUNCH(k_axpy_dot<TB>, g, FG_TPB, 0, c->stream, n, j, W, U + (size_t)j * ld,
                      U + (size_t)(j + 1) * ld, st, c->red_part, c->red_ticket);
        FG_LAUNCH(k_axpy_norm_givens<TB>, g, FG_TPB, 0, c->stream, n, W, Ui, st, c->red_part,
                  c->red_ticket);
        FG_LAUNCH(k_scale_to<TB>, g, FG_TPB, 0, c->stream, n, (const double *)nullptr,
                  (const TB *)W, U + (size_t)(i + 1) * ld, st, 1);
    }
    return FG_OK;
}

}  // namespace

int fg_gmres_fused(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, double eps, fg_gmres_info *info, int *handled);

extern "C" int fg_gmres(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                        const fg_gmres_opts *o, fg_gmres_info *info)
{
    FG_ARG(c && op && B && X && o, "ctx/op/B/X/opts");
    FG_ARG(o->kmax >= 1 && o->kmax <= FG_KMAX, "kmax in [1,64]");
    FG_ARG(o->lmax >= 1, "lmax >= 1");
    FG_ARG(o->precision == FG_GMRES_FP64 || o->precision == FG_GMRES_FP32_IR, "precision");
    FG_ARG(o->engine >= FG_ENGINE_AUTO && o->engine <= FG_ENGINE_FUSED, "engine");
    if (pc && pc->kind == FG_PREC_NONE) pc = nullptr;
    FG_HIP(hipSetDevice(c->device));
    if (o->engine != FG_ENGINE_KERNELS) {
        // eps = tol * ||B|| (pgmres.c:49-53), then the persistent solver
        const int n = op->neq;
        double *nb2 = nullptr;
        FG_HIP(hipMalloc(&nb2, sizeof(double)));
        FG_LAUNCH(k_norm2, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, B, nb2, c->red_part,
                  c->red_ticket);
        double h = 0.0;
        FG_HIP(hipMemcpyAsync(&h, nb2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FG_HIP(hipStreamSynchronize(c->stream));
        hipFree(nb2);
        const double eps = o->tol * sqrt(h);
        fg_gmres_info loc;
        int handled = 0;
        FG_CHECK(fg_gmres_fused(c, op, pc, B, X, o, eps, &loc, &handled));
        if (handled) {
            if (info) {
                *info = loc;
                // true preconditioned residual of the returned X (FP64)
                GmState *st = nullptr;
                double *W64 = nullptr;
                FG_HIP(hipMalloc(&st, sizeof(GmState)));
                FG_HIP(hipMalloc(&W64, sizeof(double) * ((size_t)n + 1)));
                FG_LAUNCH(k_state_init, 1, 1, 0, c->stream, st, o->kmax, eps);
                HostState h2{};
                int rc = residual(c, op, pc, B, X, W64, st, 0.0);
                if (rc == FG_OK) rc = read_state(c, st, h2);
                hipFree(st);
                hipFree(W64);
                FG_CHECK(rc);
                info->true_res = h2.e0;
            }
            return loc.status == FG_ERR_NOCONV ? FG_ERR_NOCONV : loc.status;
        }
        if (o->engine == FG_ENGINE_FUSED) {
            fg_set_error("the fused GMRES engine does not apply to this operator/preconditioner/size");
            return FG_ERR_ARG;
        }
    }
    const int n = op->neq, kmax = o->kmax;
    const bool ir = o->precision == FG_GMRES_FP32_IR;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;
    const size_t sb = ir ? sizeof(float) : sizeof(double);

    GmState *st = nullptr;
    void *U = nullptr, *W = nullptr;
    double *W64 = nullptr, *nb2 = nullptr;
    FG_HIP(hipMalloc(&st, sizeof(GmState)));
    FG_HIP(hipMalloc(&U, sb * ld * (kmax + 1)));
    FG_HIP(hipMalloc(&W, sb * ld));
    FG_HIP(hipMalloc(&W64, sizeof(double) * ld));
    FG_HIP(hipMalloc(&nb2, sizeof(double)));
    FG_HIP(hipMemsetAsync(U, 0, sb * ld * (kmax + 1), c->stream));
    FG_HIP(hipMemsetAsync(W, 0, sb * ld, c->stream));
    FG_HIP(hipMemsetAsync(W64, 0, sizeof(double) * ld, c->stream));

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
    int g = fg_vec_grid(c, n);
    for (;;) {
        FG_CHECK(residual(c, op, pc, B, X, W64, st, inner));
        FG_CHECK(read_state(c, st, hs));
        if (o->e0_hist && cycles <= o->lmax) o->e0_hist[cycles] = hs.e0;
        // FP64: the reference tests rho only after a cycle (do-while, :184),
        // except that a zero residual (B = 0) returns X = 0 instead of the
        // reference's 0/0 (pgmres.c:85);
        // FP32-IR: converged when the true FP64 residual is below eps.
        if (ir && hs.conv) break;
        if (!ir && hs.e0 == 0.0) { hs.rho = 0.0; break; }
        int steps_before = hs.steps;
        if (ir) {
            FG_LAUNCH(k_scale_to<float>, g, FG_TPB, 0, c->stream, n, W64, (const float *)nullptr,
                      (float *)U, st, 0);
            FG_CHECK(cycle<float>(c, op, pc, (float *)U, ld, (float *)W, st, kmax));
            FG_LAUNCH(k_backsub, 1, 1, 0, c->stream, st);
            FG_LAUNCH(k_xupdate<float>, g, FG_TPB, 0, c->stream, n, X, (const float *)U, ld, st);
        } else {
            FG_LAUNCH(k_scale_to<double>, g, FG_TPB, 0, c->stream, n, W64,
                      (const double *)nullptr, (double *)U, st, 0);
            FG_CHECK(cycle<double>(c, op, pc, (double *)U, ld, (double *)W, st, kmax));
            FG_LAUNCH(k_backsub, 1, 1, 0, c->stream, st);
            FG_LAUNCH(k_xupdate<double>, g, FG_TPB, 0, c->stream, n, X, (const double *)U, ld,
                      st);
        }
        cycles++;
        FG_CHECK(read_state(c, st, hs));
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
        // true preconditioned residual of the returned X (FP64)
        HostState h2{};
        FG_CHECK(residual(c, op, pc, B, X, W64, st, 0.0));
        FG_CHECK(read_state(c, st, h2));
        info->true_res = h2.e0;
    }
    hipFree(st); hipFree(U); hipFree(W); hipFree(W64); hipFree(nb2);
    return rc == FG_ERR_NOCONV ? FG_ERR_NOCONV : rc;
}
