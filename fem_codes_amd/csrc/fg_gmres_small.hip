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
//  * a dot product is a wavefront reduction plus one LDS exchange and one
//    workgroup barrier; every thread sums the 16 wavefront partials in the
//    same order, so all threads hold the same value and branch uniformly;
//  * the slice of u_{j+1} is loaded before the reduction of h_{j,i}, so the
//    L2 latency of the next basis vector hides behind the barrier;
//  * the product is bit for bit the reference's accumulation: EBE/EDE
//    (ebemv.c:26-39, edemv.c:24-35) in two phases, the row sums of every
//    element into an L2-resident scratch array, a barrier, then each row
//    adds its elements' row sums in element order (no atomics); CSR rows
//    (csrmv.c:18-21) in column order;
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
    int n, nel, kmax, lmax, hist_cap, pc;
    long long ld;
    double tol;
    // CSR
    const int *IA, *JA;
    const double *AA;
    // EBE / EDE (blocked values, element-order incidences)
    const int *ioff, *iea, *lm, *kstride;
    const long long *kpos;
    const double *Kv;
    double *rs;             // [nel*K] element row sums of the current product
    // preconditioners
    const double *invD;     // Diag: [n]
    const double *invde;    // element Jacobi: per incidence, ioff order
    // vectors
    const double *B;
    double *X, *U;
    double *rh, *e0h;       // rho after every step, ||r|| at every restart
    SRes *res;
};

// deterministic workgroup sum, returned in every thread.  Two LDS buffers
// used alternately: a thread writes buffer b again only after the barrier of
// the following reduction, which every thread reaches after reading b.
__device__ __forceinline__ double wg_sum(double v, double (*ws)[NWAVE], int &par)
{
    v = fg_wave_sum(v);
    double *b = ws[par];
    par ^= 1;
    if ((threadIdx.x & 63) == 0) b[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = b[0];
#pragma unroll
    for (int w = 1; w < NWAVE; w++) t = t + b[w];
    return t;
}

// element row sums of A*P (ebemv.c:34-36, left to right over the columns)
// into rs[e*K + r], every element of the system; EU elements per thread in
// flight
template <int K, int EU>
__device__ __forceinline__ void element_rows(const SArgs &a, const double *__restrict__ P)
{
    constexpr int V = fg_ebe_vec(K);
    for (int e0 = threadIdx.x; e0 < a.nel; e0 += EU * STPB) {
        long long kp[EU];
        int ks[EU], d[EU][K];
#pragma unroll
        for (int u = 0; u < EU; u++) {
            const int e = e0 + u * STPB;
            kp[u] = e < a.nel ? a.kpos[e] : -1;   // -1: every dof prescribed, not stored
            ks[u] = e < a.nel ? a.kstride[e] : 0;
#pragma unroll
            for (int b = 0; b < K; b++) d[u][b] = e < a.nel ? a.lm[e * K + b] : a.n;
        }
        double p[EU][K];
#pragma unroll
        for (int u = 0; u < EU; u++)
#pragma unroll
            for (int b = 0; b < K; b++)
                p[u][b] = d[u][b] == a.n ? 0.0 : P[d[u][b]];   // P[neq] = 0 (ebemv.c:21)
#pragma unroll
        for (int u = 0; u < EU; u++) {
            if (kp[u] < 0) continue;
            const int e = e0 + u * STPB;
#pragma unroll
            for (int r = 0; r < K; r++) {
                double sr = a.Kv[fg_ebe_vidx(kp[u], ks[u], r * K, V)] * p[u][0];
#pragma unroll
                for (int b = 1; b < K; b++)
                    sr = sr + a.Kv[fg_ebe_vidx(kp[u], ks[u], r * K + b, V)] * p[u][b];
                a.rs[e * K + r] = sr;
            }
        }
    }
}

// rows tid + v*STPB of M^{-1} A P.  EBE/EDE: each row adds the row sums of
// its elements in element order (after element_rows and a barrier); CSR:
// AA[j]*P[JA[j]] in column order (csrmv.c:18-21).  The entry loop runs in
// chunks of CU with one trip count for all VPT rows, so the loads of every
// row and chunk are in flight together; each row still adds its terms in
// its own order.
template <int K, int VPT>
__device__ __forceinline__ void apply_rows(const SArgs &a, const double *__restrict__ P,
                                           double (&q)[VPT])
{
    constexpr int CU = VPT >= 4 ? (K >= 3 ? 1 : 2) : 8 / VPT;   // terms in flight per row
    const int n = a.n;
    if constexpr (K != 0) {
        element_rows<K, (VPT >= 4 ? 1 : 2)>(a, P);   // VPT = 4: no spill
        __syncthreads();
    }
    const int *off_arr = K == 0 ? a.IA : a.ioff;
    int off[VPT], cnt[VPT], cmax = 0;
#pragma unroll
    for (int v = 0; v < VPT; v++) {
        const int row = threadIdx.x + v * STPB;
        off[v] = row < n ? off_arr[row] : 0;
        cnt[v] = row < n ? off_arr[row + 1] - off[v] : 0;
        cmax = cnt[v] > cmax ? cnt[v] : cmax;
        q[v] = 0.0;
    }
    for (int c = 0; c < cmax; c += CU) {
        double t[VPT][CU];
#pragma unroll
        for (int v = 0; v < VPT; v++)
#pragma unroll
            for (int u = 0; u < CU; u++) {
                const int j = off[v] + c + u;
                t[v][u] = 0.0;
                if (c + u < cnt[v]) {
                    if constexpr (K == 0) t[v][u] = a.AA[j] * P[a.JA[j]];
                    else t[v][u] = a.rs[a.iea[j]];
                }
            }
#pragma unroll
        for (int v = 0; v < VPT; v++)
#pragma unroll
            for (int u = 0; u < CU; u++)
                if (c + u < cnt[v]) q[v] = q[v] + t[v][u];
    }
    if (a.pc == PC_DIAG) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = threadIdx.x + v * STPB;
            if (row < n) q[v] = q[v] * a.invD[row];
        }
    } else if (a.pc == PC_JACOBI) {
        // Jacobi_precond_EBE.c: z_g = sum over the elements of g, in element
        // order, of p_g / A^e_gg
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const int row = threadIdx.x + v * STPB;
            if (row >= n) continue;
            double z = 0.0;
            for (int t = a.ioff[row]; t < a.ioff[row + 1]; t++) z = z + q[v] * a.invde[t];
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

    // eps = tol * ||B|| (pgmres.c:49-53) ; X = 0 (:46)
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < VPT; v++) {
        const int row = tid + v * STPB;
        if (row < n) {
            const double b = a.B[row];
            acc = acc + b * b;
            a.X[row] = 0.0;
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
            r[v] = row < n ? a.B[row] - r[v] : 0.0;
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
                if (row < n) a.U[row] = sc * r[v];
            }
            if (tid == 0) {
                a.U[n] = 0.0;
                s_e[0] = e0;
            }
            __syncthreads();
            int m = 0;
            for (int i = 0; i < a.kmax; i++) {
                const double *Ui = a.U + (size_t)i * ld;
                // w = M^{-1} A u_i ; h_{0,i} = (u_0, w)
                double w[VPT], cur[VPT], nxt[VPT];
                apply_rows<K, VPT>(a, Ui, w);
                acc = 0.0;
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int row = tid + v * STPB;
                    cur[v] = row < n ? a.U[row] : 0.0;
                    acc = acc + cur[v] * w[v];
                }
#pragma unroll
                for (int v = 0; v < VPT; v++) {
                    const int row = tid + v * STPB;
                    nxt[v] = (i >= 1 && row < n) ? a.U[ld + row] : 0.0;
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
                            nxt[v] = row < n ? Un[row] : 0.0;
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
                    if (row < n) U1[row] = s * w[v];
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
                    double x = a.X[row];
                    for (int j = 0; j < m; j++) x = x + a.U[(size_t)j * ld + row] * s_y[j];
                    a.X[row] = x;
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
                const double d = a.B[row] - t[v];
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
    if (op->kind == FG_OP_CSR) {
        if (!op->d_IA) return false;
    } else if (op->kind == FG_OP_EBE || op->kind == FG_OP_EDE) {
        if (!(op->k == 2 || op->k == 3 || op->k == 4) || !op->d_inc_ea) return false;
    } else {
        return false;
    }
    if (pc && !(pc->kind == FG_PREC_DIAG || (pc->kind == FG_PREC_JACOBI && pc->d_invde)))
        return false;
    return true;
}

int fg_gmres_small(fg_ctx *c, fg_op *op, fg_prec *pc, const double *B, double *X,
                   const fg_gmres_opts *o, fg_gmres_info *info)
{
    const int n = op->neq;
    int vpt = 1;
    while (vpt * STPB < n) vpt *= 2;
    const long long ld = ((long long)n + 1 + 63) / 64 * 64;

    SArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.kmax = o->kmax; a.lmax = o->lmax; a.ld = ld; a.tol = o->tol;
    a.hist_cap = o->rho_hist ? o->hist_cap : 0;
    if (op->kind == FG_OP_CSR) {
        a.IA = op->d_IA; a.JA = op->d_JA; a.AA = op->d_AA;
    } else {
        a.nel = op->nitems;
        a.ioff = op->d_inc_off; a.iea = op->d_inc_ea; a.lm = op->d_lm;
        a.kpos = op->d_kpos; a.kstride = op->d_kstride; a.Kv = op->d_K64;
    }
    a.pc = PC_NONE;
    if (pc && pc->kind == FG_PREC_DIAG) {
        a.pc = PC_DIAG;
        a.invD = pc->d_invdiag;
    } else if (pc && pc->kind == FG_PREC_JACOBI) {
        a.pc = PC_JACOBI;
        a.invde = pc->d_invde;
        a.ioff = op->d_inc_off;
    }
    a.B = B;
    a.X = X;

    // workspace: basis, results, histories
    const long long ne0 = o->e0_hist ? (long long)o->lmax + 1 : 0;
    const size_t szU = sizeof(double) * ld * (o->kmax + 1);
    const size_t szH = sizeof(double) * (a.hist_cap > 0 ? a.hist_cap : 0);
    const size_t szE = sizeof(double) * ne0;
    const size_t szS = op->kind == FG_OP_CSR ? 0 : sizeof(double) * op->nitems * op->k;
    const size_t off_R = szU, off_H = off_R + 256, off_E = off_H + (szH + 255) / 256 * 256;
    const size_t off_S = off_E + (szE + 255) / 256 * 256;
    void *buf = nullptr;
    FG_HIP(hipMalloc(&buf, off_S + szS));
    struct Free {
        void *p;
        ~Free() { hipFree(p); }
    } fr{buf};
    char *base = (char *)buf;
    a.U = (double *)base;
    a.res = (SRes *)(base + off_R);
    a.rh = szH ? (double *)(base + off_H) : nullptr;
    a.e0h = szE ? (double *)(base + off_E) : nullptr;
    a.rs = szS ? (double *)(base + off_S) : nullptr;

    FG_HIP(hipEventRecord(c->ev0, c->stream));
    if (op->kind == FG_OP_CSR) FG_CHECK(launch_k<0>(c, a, vpt));
    else if (op->k == 2) FG_CHECK(launch_k<2>(c, a, vpt));
    else if (op->k == 3) FG_CHECK(launch_k<3>(c, a, vpt));
    else FG_CHECK(launch_k<4>(c, a, vpt));
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    FG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    SRes r;
    FG_HIP(hipMemcpy(&r, a.res, sizeof r, hipMemcpyDeviceToHost));
    if (a.rh) {
        const int nh = r.steps < a.hist_cap ? r.steps : a.hist_cap;
        if (nh > 0)
            FG_HIP(hipMemcpy(o->rho_hist, a.rh, sizeof(double) * nh, hipMemcpyDeviceToHost));
    }
    if (a.e0h) {
        const long long ne = r.ne0 < ne0 ? r.ne0 : ne0;
        if (ne > 0) FG_HIP(hipMemcpy(o->e0_hist, a.e0h, sizeof(double) * ne, hipMemcpyDeviceToHost));
    }
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
