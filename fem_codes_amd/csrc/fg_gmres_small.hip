// fg_gmres_small.hip -- restarted GMRES of a small system in one workgroup.
//
// Same algorithm, operation order and stopping rules as the kernel-per-
// operation engine of fg_gmres.hip (Solvers_and_Preconditioners/pgmres.c:
// 4-211, FP64, modified Gram-Schmidt), for systems of at most FG_SMALL_MAX
// equations.  At that size every vector fits in one XCD's L2 and a solve is
// bound by the latency between dependent operations, not by bandwidth: a
// kernel per operation pays one dispatch per dot product (i + 4 per Arnoldi
// step), a multi-workgroup persistent kernel pays one cross-XCD barrier per
// dot product.  Here a single 1024-thread workgroup runs the whole solve:
//
//  * thread t owns rows t + v*1024 (v < VPT <= 4) and keeps its slice of
//    the Arnoldi vector w, of u_j and of u_{j+1} in registers for the whole
//    Gram-Schmidt sweep (VPT = 4 takes 122-128 VGPRs without spilling, the
//    budget of a 1024-thread workgroup);
//  * a dot product is a DPP wavefront reduction (register crossbar; the
//    shuffle form goes through LDS six times), one LDS exchange and one
//    workgroup barrier; every thread sums the 16 wavefront totals in the
//    same order, so all threads hold the same value and branch uniformly;
//  * the slice of u_{j+1} is loaded before the reduction of h_{j,i}, so the
//    L2 latency of the next basis vector hides behind the barrier;
//  * the product is bit for bit the reference's accumulation: EBE/EDE
//    (ebemv.c:26-39, edemv.c:24-35) in two phases, the row sums of every
//    element into an L2-resident scratch array, a barrier, then each row
//    adds its elements' row sums in element order (no atomics); CSR rows
//    (csrmv.c:18-21) in column order;
//  * one CU's vector L1 serves every access, so its address throughput, not
//    bandwidth, limits the product: every list is stored transposed (entry
//    c of row g at c*n + g, entry q of element e at q*nel + e) so that the
//    64 lanes of a wavefront read one contiguous run.  The element and CSR
//    values are copied into that order once per solve;
//  * Diag (Diag_precond.c) and element Jacobi (Jacobi_precond_EBE.c) are
//    row-local and are applied to each product row in registers.
#include <math.h>
#include <string.h>

#include "fg_ebe_body.h"
#include "fg_internal.h"

namespace {

constexpr int STPB = 1024;          // threads of the one workgroup
constexpr int NWAVE = STPB / 64;
constexpr int KM = FG_KMAX;
enum { PC_NONE = 0, PC_DIAG = 1, PC_JACOBI = 2 };

struct SRes {
    double rho, e0, eps, true_res;
    int steps, cycles, status, ne0;   // ne0: entries written to the e0 history
};

struct SArgs {
    int n, nel, mi, kmax, lmax, hist_cap, pc;
    long long ld;
    double tol;
    const int *iT;          // [mi][n] transposed entry lists (EBE/EDE: rs slots; CSR: JA)
    // CSR
    const int *IA;
    const double *AA;
    double *AAT;            // [mi][n] values in iT order (per solve)
    // EBE / EDE
    const int *ioff, *lmT, *kstride;
    const long long *kpos;
    const double *Kv;       // blocked values
    double *KeT;            // [K*K][nel] values, entry-major (per solve)
    double *rs;             // [K][nel] element row sums of the current product
    // preconditioners
    const double *invD;     // Diag: [n]
    const double *invde;    // element Jacobi: per incidence, ioff order
    double *invdeT;         // [mi][n] the same in iT order (per solve)
    // vectors
    const double *B;
    double *X, *U;
    double *rh, *e0h;       // rho after every step, ||r|| at every restart
    SRes *res;
};

// x from another lane by a DPP pattern (register crossbar, no LDS); lanes
// outside the row mask or without a source lane read 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_from(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// sum over the wavefront, valid in lane 63: partial sums of 2, 4, 8, 16
// consecutive lanes within each row of 16 (row_shr 1, 2, 4, 8), then row 1
// and 3 add the last lane of the row below (row_bcast:15), rows 2 and 3 add
// lane 31 (row_bcast:31).  A fixed order, six dependent steps.
__device__ __forceinline__ double wave_total(double v)
{
    v = v + dpp_from<0x111, 0xF>(v);
    v = v + dpp_from<0x112, 0xF>(v);
    v = v + dpp_from<0x114, 0xF>(v);
    v = v + dpp_from<0x118, 0xF>(v);
    v = v + dpp_from<0x142, 0xA>(v);
    v = v + dpp_from<0x143, 0xC>(v);
    return v;
}

// deterministic workgroup sum, returned in every thread: the 16 wavefront
// totals summed in a fixed order (four chains of four, then pairs).  Two LDS buffers used alternately: a
// thread writes buffer b again only after the barrier of the following
// reduction, which every thread reaches after reading b.
__device__ __forceinline__ double wg_sum(double v, double (*ws)[NWAVE], int &par)
{
    v = wave_total(v);
    double *b = ws[par];
    par ^= 1;
    if ((threadIdx.x & 63) == 63) b[threadIdx.x >> 6] = v;
    __syncthreads();
    double t[4];
#pragma unroll
    for (int g = 0; g < 4; g++) t[g] = ((b[4 * g] + b[4 * g + 1]) + b[4 * g + 2]) + b[4 * g + 3];
    return (t[0] + t[1]) + (t[2] + t[3]);
}

// per-solve copies in lane-coalesced order (entry c of row g at c*n + g,
// entry q of element e at q*nel + e): element values from the blocked
// layout, CSR values, element-Jacobi factors; then a barrier
template <int K>
__device__ __forceinline__ void layout_phase(const SArgs &a)
{
    const int n = a.n, mi = a.mi;
    if constexpr (K == 0) {
        for (int r = threadIdx.x; r < n; r += STPB) {
            const int j0 = a.IA[r], len = a.IA[r + 1] - j0;
            for (int c = 0; c < mi; c++) fg_at(a.AAT, (unsigned)(c * n + r)) = c < len ? a.AA[j0 + c] : 0.0;
        }
    } else {
        constexpr int V = fg_ebe_vec(K);
        for (int e = threadIdx.x; e < a.nel; e += STPB) {
            const long long kp = a.kpos[e];
            const int ks = a.kstride[e];
#pragma unroll
            for (int q = 0; q < K * K; q++)
                fg_at(a.KeT, (unsigned)(q * a.nel + e)) = kp >= 0 ? a.Kv[fg_ebe_vidx(kp, ks, q, V)] : 0.0;
        }
        if (a.pc == PC_JACOBI)
            for (int g = threadIdx.x; g < n; g += STPB) {
                const int t0 = a.ioff[g], len = a.ioff[g + 1] - t0;
                for (int c = 0; c < mi; c++)
                    fg_at(a.invdeT, (unsigned)(c * n + g)) = c < len ? a.invde[t0 + c] : 0.0;
            }
    }
    __syncthreads();
}

// element row sums of A*P (ebemv.c:34-36, left to right over the columns)
// into rs[r*nel + e], every element; EU elements per thread in flight, the
// values and dof numbers of all of them loaded before the gathers of P
template <int K, int EU>
__device__ __forceinline__ void element_rows(const SArgs &a, const double *__restrict__ P)
{
    const int nel = a.nel;
    for (int e0 = threadIdx.x; e0 < nel; e0 += EU * STPB) {
        int d[EU][K];
        double kv[EU][K * K];
#pragma unroll
        for (int u = 0; u < EU; u++) {
            const int e = e0 + u * STPB;
            const bool in = e < nel;
#pragma unroll
            for (int b = 0; b < K; b++) d[u][b] = in ? fg_at(a.lmT, (unsigned)(b * nel + e)) : a.n;
#pragma unroll
            for (int q = 0; q < K * K; q++)
                kv[u][q] = in ? fg_at(a.KeT, (unsigned)(q * nel + e)) : 0.0;
        }
        double p[EU][K];
#pragma unroll
        for (int u = 0; u < EU; u++)
#pragma unroll
            for (int b = 0; b < K; b++)   // P[neq] = 0 (ebemv.c:21)
                p[u][b] = d[u][b] == a.n ? 0.0 : fg_at(P, (unsigned)d[u][b]);
#pragma unroll
        for (int u = 0; u < EU; u++) {
            const int e = e0 + u * STPB;
            if (e >= nel) continue;
#pragma unroll
            for (int r = 0; r < K; r++) {
                double sr = kv[u][r * K] * p[u][0];
#pragma unroll
                for (int b = 1; b < K; b++) sr = sr + kv[u][r * K + b] * p[u][b];
                fg_at(a.rs, (unsigned)(r * nel + e)) = sr;
            }
        }
    }
}

// rows tid + v*STPB of M^{-1} A P.  EBE/EDE: each row adds the row sums of
// its elements in element order (after element_rows and a barrier); CSR:
// AA[j]*P[JA[j]] in column order (csrmv.c:18-21).  Entry c of every row
// comes from the transposed lists, so a wavefront's loads are contiguous;
// CU entries per row are in flight at once, and each row still adds its
// terms in its own order.
template <int K, int VPT>
__device__ __forceinline__ void apply_rows(const SArgs &a, const double *__restrict__ P,
                                           double (&q)[VPT])
{
    constexpr int CU = VPT >= 4 ? 2 : 8 / VPT;
    const int n = a.n, mi = a.mi;
    if constexpr (K != 0) {
        element_rows<K, (K == 4 ? 1 : 2)>(a, P);
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < VPT; v++) q[v] = 0.0;
    for (int c = 0; c < mi; c += CU) {
        double t[VPT][CU];
        bool ok[VPT][CU];
#pragma unroll
        for (int v = 0; v < VPT; v++)
#pragma unroll
            for (int u = 0; u < CU; u++) {
                const int row = threadIdx.x + v * STPB;
                const unsigned at = (unsigned)((c + u) * n + row);
                const int idx = (row < n && c + u < mi) ? fg_at(a.iT, at) : -1;
                ok[v][u] = idx >= 0;
                t[v][u] = 0.0;
                if (idx >= 0) {
                    if constexpr (K == 0)
                        t[v][u] = fg_at(a.AAT, at) * (idx == n ? 0.0 : fg_at(P, (unsigned)idx));
                    else
                        t[v][u] = fg_at(a.rs, (unsigned)idx);
                }
            }
#pragma unroll
        for (int v = 0; v < VPT; v++)
#pragma unroll
            for (int u = 0; u < CU; u++)
                if (ok[v][u]) q[v] = q[v] + t[v][u];
    }
    if (a.pc == PC_DIAG) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = threadIdx.x + v * STPB;
            if (row < n) q[v] = q[v] * fg_at(a.invD, (unsigned)row);
        }
    } else if (a.pc == PC_JACOBI) {
        // Jacobi_precond_EBE.c: z_g = sum over the elements of g, in element
        // order, of p_g / A^e_gg
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = threadIdx.x + v * STPB;
            if (row >= n) continue;
            double z = 0.0;
            for (int c = 0; c < mi; c++) {
                const unsigned at = (unsigned)(c * n + row);
                if (fg_at(a.iT, at) >= 0) z = z + q[v] * fg_at(a.invdeT, at);
            }
            q[v] = z;
        }
    }
}

template <int K, int VPT>
__global__ __launch_bounds__(STPB) void k_gmres_small(SArgs a)
{
    __shared__ double ws[2][NWAVE];
    __shared__ double sH[(KM + 1) * KM];   // H[j*KM + i]
    __shared__ double s_cs[KM], s_sn[KM], s_e[KM + 1], s_y[KM];
    __shared__ double s_rho;
    const int n = a.n, tid = threadIdx.x;
    const long long ld = a.ld;
    int par = 0;

    layout_phase<K>(a);
    // eps = tol * ||B|| (pgmres.c:49-53) ; X = 0 (:46), the sink slot too
    if (tid == 0) a.X[n] = 0.0;
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < VPT; v++) {
        const int row = tid + v * STPB;
        if (row < n) {
            const double b = fg_at(a.B, (unsigned)row);
            acc = acc + b * b;
            fg_at(a.X, (unsigned)row) = 0.0;
        }
    }
    const double eps = a.tol * sqrt(wg_sum(acc, ws, par));

    int steps = 0, cycles = 0, status = FG_OK, nres = 0;
    double rho = 0.0, e0 = 0.0;
    for (;;) {
        // r = B - M^{-1} A X ; e0 = ||r|| (X of every row written before the
        // barrier of the last reduction)
        double r[VPT];
        apply_rows<K, VPT>(a, a.X, r);
        acc = 0.0;
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = tid + v * STPB;
            r[v] = row < n ? fg_at(a.B, (unsigned)row) - r[v] : 0.0;
            acc = acc + r[v] * r[v];
        }
        e0 = sqrt(wg_sum(acc, ws, par));
        if (tid == 0 && a.e0h && cycles <= a.lmax) a.e0h[cycles] = e0;
        nres++;
        if (e0 != e0) {            // non-finite system
            status = FG_ERR_BREAKDOWN;
            break;
        }
        if (e0 == 0.0) {           // B = 0 or an exact X: X is the answer
            rho = 0.0;
            break;
        }
        rho = e0;
        bool breakdown = false;
        // a cycle that starts below eps takes no step (pgmres.c:92)
        if (e0 > eps) {
            const double sc = 1.0 / e0;
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const int row = tid + v * STPB;
                if (row < n) fg_at(a.U, (unsigned)row) = sc * r[v];
            }
            if (tid == 0) {
                a.U[n] = 0.0;
                s_e[0] = e0;
            }
            __syncthreads();
            int m = 0;
            for (int i = 0; i < a.kmax; i++) {
                const double *Ui = a.U + (size_t)i * ld;   // uniform
                // w = M^{-1} A u_i ; h_{0,i} = (u_0, w)
                double w[VPT], cur[VPT], nxt[VPT];
                apply_rows<K, VPT>(a, Ui, w);
                acc = 0.0;
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int row = tid + v * STPB;
                    cur[v] = row < n ? fg_at(a.U, (unsigned)row) : 0.0;
                    acc = acc + cur[v] * w[v];
                }
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int row = tid + v * STPB;
                    nxt[v] = (i >= 1 && row < n) ? fg_at(a.U + ld, (unsigned)row) : 0.0;
                }
                double h = wg_sum(acc, ws, par);
                if (tid == 0) sH[i] = h;
                // modified Gram-Schmidt (pgmres.c:107-114): w -= h_{j,i} u_j,
                // then h_{j+1,i} = (u_{j+1}, w)
                for (int j = 0; j < i; j++) {
                    acc = 0.0;
#pragma unroll
                    for (int v = 0; v < VPT; v++) {
                        w[v] = w[v] + (-h) * cur[v];
                        cur[v] = nxt[v];
                        acc = acc + cur[v] * w[v];
                    }
                    if (j + 2 <= i) {
                        const double *Un = a.U + (size_t)(j + 2) * ld;
#pragma unroll
                        for (int v = 0; v < VPT; v++) {
                            const int row = tid + v * STPB;
                            nxt[v] = row < n ? fg_at(Un, (unsigned)row) : 0.0;
                        }
                    }
                    h = wg_sum(acc, ws, par);
                    if (tid == 0) sH[(j + 1) * KM + i] = h;
                }
                // w -= h_{i,i} u_i ; h_{i+1,i} = ||w|| (:115-118)
                acc = 0.0;
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    w[v] = w[v] + (-h) * cur[v];
                    acc = acc + w[v] * w[v];
                }
                const double hn = sqrt(wg_sum(acc, ws, par));
                // u_{i+1} = w / h_{i+1,i} (:121; unused when the cycle stops)
                const double s = 1.0 / hn;
                double *U1 = a.U + (size_t)(i + 1) * ld;
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int row = tid + v * STPB;
                    if (row < n) fg_at(U1, (unsigned)row) = s * w[v];
                }
                // Givens rotations (:124-159)
                if (tid == 0) {
                    U1[n] = 0.0;
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
                    s_rho = fabs(s_e[i + 1]);
                    if (a.rh && steps < a.hist_cap) a.rh[steps] = s_rho;
                }
                __syncthreads();   // rho, and u_{i+1} for the next product
                rho = s_rho;
                steps++;
                m = i + 1;
                // h_{i+1,i} = 0: the Krylov space is invariant (the reference
                // divides by zero, :121); stop, flag it unless rho meets eps
                if (hn == 0.0) {
                    breakdown = rho > eps;
                    break;
                }
                if (!(rho > eps && i + 1 < a.kmax)) break;
            }
            // back substitution (:165-175), X += sum_j u_j y_j (:177-181)
            if (tid == 0) {
                const int i = m - 1;
                s_y[i] = s_e[i] / sH[i * KM + i];
                for (int j = i - 1; j >= 0; j--) {
                    double soma = 0.0;
                    for (int q = j + 1; q <= i; q++) soma = soma + sH[j * KM + q] * s_y[q];
                    s_y[j] = (s_e[j] - soma) / sH[j * KM + j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < VPT; v++) {
                const int row = tid + v * STPB;
                if (row < n) {
                    double x = fg_at(a.X, (unsigned)row);
                    for (int j = 0; j < m; j++)
                        x = x + fg_at(a.U + (size_t)j * ld, (unsigned)row) * s_y[j];
                    fg_at(a.X, (unsigned)row) = x;
                }
            }
            __syncthreads();   // X of every row before the next residual
        }
        cycles++;
        if (breakdown) {
            status = FG_ERR_BREAKDOWN;
            break;
        }
        if (!(rho > eps)) break;
        if (cycles >= a.lmax) {
            status = FG_ERR_NOCONV;
            break;
        }
    }
    // true preconditioned residual of the returned X
    __syncthreads();
    {
        double t[VPT];
        apply_rows<K, VPT>(a, a.X, t);
        acc = 0.0;
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = tid + v * STPB;
            if (row < n) {
                const double d = fg_at(a.B, (unsigned)row) - t[v];
                acc = acc + d * d;
            }
        }
    }
    const double tr = sqrt(wg_sum(acc, ws, par));
    if (tid == 0) {
        a.res->rho = rho;
        a.res->e0 = e0;
        a.res->eps = eps;
        a.res->true_res = tr;
        a.res->steps = steps;
        a.res->cycles = cycles;
        a.res->status = status;
        a.res->ne0 = nres <= a.lmax + 1 ? nres : a.lmax + 1;
    }
}

template <int K>
int launch_k(fg_ctx *c, const SArgs &a, int vpt)
{
    switch (vpt) {
    case 1: FG_LAUNCH((k_gmres_small<K, 1>), 1, STPB, 0, c->stream, a); break;
    case 2: FG_LAUNCH((k_gmres_small<K, 2>), 1, STPB, 0, c->stream, a); break;
    case 4: FG_LAUNCH((k_gmres_small<K, 4>), 1, STPB, 0, c->stream, a); break;
    default: fg_set_error("small GMRES: rows per thread %d", vpt); return FG_ERR_ARG;
    }
    return FG_OK;
}

}  // namespace

// does the one-workgroup engine take this solve?
bool fg_gmres_small_applies(const fg_op *op, const fg_prec *pc, const fg_gmres_opts *o)
{
    if (o->precision != FG_GMRES_FP64) return false;
    const int n = op->neq;
    if (n < 1 || n > FG_SMALL_MAX || o->kmax > KM) return false;
    if (!op->d_sm_iT) return false;   // built at creation for systems of this size
    if (op->kind == FG_OP_CSR) {
        if (!op->d_IA) return false;
    } else if (op->kind == FG_OP_EBE || op->kind == FG_OP_EDE) {
        if (!(op->k == 2 || op->k == 3 || op->k == 4)) return false;
    } else {
        return false;
    }
    if (pc && !(pc->kind == FG_PREC_DIAG || (pc->kind == FG_PREC_JACOBI && pc->d_invde)))
        return false;
    return true;
}

namespace {

// arguments and workspace layout of one solve; buffers are bound by bind()
struct SPlan {
    SArgs a;
    int vpt;
    long long ne0;
    size_t off_R, off_H, off_E, off_S, off_K, off_A, off_J, bytes;
};

void plan_small(fg_op *op, fg_prec *pc, const double *B, double *X, const fg_gmres_opts *o,
                SPlan &sp)
{
    const int n = op->neq;
    int vpt = 1;
    while (vpt * STPB < n) vpt *= 2;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;
    SArgs &a = sp.a;
    memset(&a, 0, sizeof a);
    a.n = n; a.kmax = o->kmax; a.lmax = o->lmax; a.ld = ld; a.tol = o->tol;
    a.hist_cap = o->rho_hist ? o->hist_cap : 0;
    a.mi = op->sm_mi;
    a.iT = op->d_sm_iT;
    if (op->kind == FG_OP_CSR) {
        a.IA = op->d_IA; a.AA = op->d_AA;
    } else {
        a.nel = op->nitems;
        a.ioff = op->d_inc_off; a.lmT = op->d_sm_lmT;
        a.kpos = op->d_kpos; a.kstride = op->d_kstride; a.Kv = op->d_K64;
    }
    a.pc = PC_NONE;
    if (pc && pc->kind == FG_PREC_DIAG) {
        a.pc = PC_DIAG;
        a.invD = pc->d_invdiag;
    } else if (pc && pc->kind == FG_PREC_JACOBI) {
        a.pc = PC_JACOBI;
        a.invde = pc->d_invde;
    }
    a.B = B;
    a.X = X;
    sp.vpt = vpt;
    // workspace: basis, results, histories, per-solve transposed copies
    sp.ne0 = o->e0_hist ? (long long)o->lmax + 1 : 0;
    const size_t szU = sizeof(double) * ld * (o->kmax + 1);
    const size_t szH = sizeof(double) * (a.hist_cap > 0 ? a.hist_cap : 0);
    const size_t szE = sizeof(double) * sp.ne0;
    const bool csr = op->kind == FG_OP_CSR;
    const size_t nmi = (size_t)op->sm_mi * n;
    const size_t szS = csr ? 0 : sizeof(double) * op->nitems * op->k;             // rs
    const size_t szK = csr ? 0 : sizeof(double) * op->nitems * op->k * op->k;     // KeT
    const size_t szA = csr ? sizeof(double) * nmi : 0;                              // AAT
    const size_t szJ = a.pc == PC_JACOBI ? sizeof(double) * nmi : 0;                // invdeT
    fg_carver cv;
    cv.take(szU);
    sp.off_R = cv.take(sizeof(SRes));
    sp.off_H = cv.take(szH);
    sp.off_E = cv.take(szE);
    sp.off_S = cv.take(szS);
    sp.off_K = cv.take(szK);
    sp.off_A = cv.take(szA);
    sp.off_J = cv.take(szJ);
    sp.bytes = cv.off;
    (void)szH;
}

// point the solve's buffers into `base` (sp.bytes long); every entry the
// kernel reads it writes first in the same solve, so nothing is cleared
void bind(SPlan &sp, char *base, fg_op *op)
{
    SArgs &a = sp.a;
    const bool csr = op->kind == FG_OP_CSR;
    a.U = (double *)base;
    a.res = (SRes *)(base + sp.off_R);
    a.rh = a.hist_cap > 0 ? (double *)(base + sp.off_H) : nullptr;
    a.e0h = sp.ne0 > 0 ? (double *)(base + sp.off_E) : nullptr;
    a.rs = csr ? nullptr : (double *)(base + sp.off_S);
    a.KeT = csr ? nullptr : (double *)(base + sp.off_K);
    a.AAT = csr ? (double *)(base + sp.off_A) : nullptr;
    a.invdeT = a.pc == PC_JACOBI ? (double *)(base + sp.off_J) : nullptr;
}

int run_small(fg_ctx *c, fg_op *op, const fg_gmres_opts *o, SPlan &sp, fg_gmres_info *info)
{
    const SArgs &a = sp.a;
    FG_HIP(hipEventRecord(c->ev0, c->stream));
    if (op->kind == FG_OP_CSR) FG_CHECK(launch_k<0>(c, a, sp.vpt));
    else if (op->k == 2) FG_CHECK(launch_k<2>(c, a, sp.vpt));
    else if (op->k == 3) FG_CHECK(launch_k<3>(c, a, sp.vpt));
    else FG_CHECK(launch_k<4>(c, a, sp.vpt));
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    FG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    SRes r;
    FG_HIP(hipMemcpyAsync(&r, a.res, sizeof r, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    if (a.rh) {
        const int nh = r.steps < a.hist_cap ? r.steps : a.hist_cap;
        if (nh > 0)
            FG_HIP(hipMemcpyAsync(o->rho_hist, a.rh, sizeof(double) * nh, hipMemcpyDeviceToHost,
                                  c->stream));
    }
    if (a.e0h) {
        const long long ne = r.ne0 < sp.ne0 ? r.ne0 : sp.ne0;
        if (ne > 0)
            FG_HIP(hipMemcpyAsync(o->e0_hist, a.e0h, sizeof(double) * ne,
                                  hipMemcpyDeviceToHost, c->stream));
    }
    FG_HIP(hipStreamSynchronize(c->stream));
    if (info) {
        memset(info, 0, sizeof *info);
        info->steps = r.steps;
        info->cycles = r.cycles;
        info->rho = r.rho;
        info->eps = r.eps;
        info->true_res = r.true_res;
        info->ms = ms;
        info->status = r.status;
    }
    return r.status == FG_ERR_NOCONV ? FG_ERR_NOCONV : r.status;
}

// ---- the allocation-free variant -----------------------------------------
// Its workspace is this file-scope device array (one copy per GPU, held until
// the process ends).  In host code `scratch` names a host-side stand-in, not
// device memory: the device address of each GPU's copy is looked up once per
// device (fg_noalloc_init, called by fg_ctx_create with that device current)
// and only that address is passed to kernels.
constexpr int NOALLOC_MAXDEV = 64;
static __device__ __align__(256) char scratch[FG_NOALLOC_SCRATCH_BYTES];
static void *g_scratch_dev[NOALLOC_MAXDEV];

}  // namespace

int fg_noalloc_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAXDEV) {
        fg_set_error("device %d beyond the %d devices the scratch table holds", device,
                     NOALLOC_MAXDEV);
        return FG_ERR_ARG;
    }
    if (g_scratch_dev[device]) return FG_OK;
    void *ptr = nullptr;
    FG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = ptr;
    return FG_OK;
}

int fg_gmres_small(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, fg_gmres_info *info)
{
    SPlan sp;
    plan_small(op, pc, B, X, o, sp);
    // the context's solver workspace: kept between solves
    void *buf = nullptr;
    FG_CHECK(fg_ctx_workspace(c, sp.bytes, &buf));
    bind(sp, (char *)buf, op);
    return run_small(c, op, o, sp, info);
}

extern "C" int fg_gmres_small_noalloc(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B,
                                      double *X, const fg_gmres_opts *o, fg_gmres_info *info)
{
    FG_ARG(c && op && B && X && o, "ctx/op/B/X/opts");
    FG_ARG(o->kmax >= 1 && o->kmax <= FG_KMAX, "kmax in [1,64]");
    FG_ARG(o->lmax >= 1, "lmax >= 1");
    if (pc && pc->kind == FG_PREC_NONE) pc = nullptr;
    FG_ARG(!pc || (pc->op->neq == op->neq && pc->op->ctx == c),
           "preconditioner built on an operator of another size or context");
    FG_ARG(op->ctx == c, "operator of another context");
    if (!fg_gmres_small_applies(op, pc, o)) {
        fg_set_error("invalid argument: the allocation-free GMRES takes FP64 solves of at most "
                     "%d equations, CSR or EBE/EDE with k = 2..4, and no, Diag or element "
                     "Jacobi preconditioning", FG_SMALL_MAX);
        return FG_ERR_ARG;
    }
    SPlan sp;
    plan_small(op, pc, B, X, o, sp);
    if (sp.bytes > (size_t)FG_NOALLOC_SCRATCH_BYTES) {
        fg_set_error("invalid argument: the solve needs %zu bytes of workspace, more than the "
                     "%d bytes of the allocation-free variant's scratch (fewer elements, a "
                     "shorter restart or shorter histories fit)", sp.bytes,
                     FG_NOALLOC_SCRATCH_BYTES);
        return FG_ERR_ARG;
    }
    FG_ARG(c->device >= 0 && c->device < NOALLOC_MAXDEV && g_scratch_dev[c->device],
           "the scratch address of this device was not looked up (context not created by "
           "fg_ctx_create)");
    FG_HIP(hipSetDevice(c->device));
    bind(sp, (char *)g_scratch_dev[c->device], op);
    return run_small(c, op, o, sp, info);
}
