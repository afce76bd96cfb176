// fg_time.hip -- the predictor/multicorrector time integrator on the device.
//
// Time_Integration/Predictor_Old.c:4-145 with Criteria.c:25-72.  The loop
// control stays on the host, as in the reference; the state it updates (u,
// a = du, Da, u_old, R, the element matrices, nodal U/dU) lives in HBM for
// the whole run, and each correction moves two doubles to the host: ||a||^2
// and ||Da||^2 for the stopping test (two more per step for the steady-state
// test).  Every vector update is one pass over neq entries, fused with the
// partial sums of the norms that follow it; the partial sums are added in
// block order by one workgroup, so a run is deterministic.
#include <math.h>
#include <string.h>

#include <chrono>

#include "fg_internal.h"

namespace {

constexpr int PG = 1024;   // blocks of the vector kernels (partial sums)

// predictor (:57-62): u_old = u ; u += ((1 - alpha) dt) a ; a = 0
__global__ __launch_bounds__(FG_TPB) void k_predict(int n, double c, double *__restrict__ u,
                                                    double *__restrict__ u_old,
                                                    double *__restrict__ a)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB) {
        const double ui = u[i];
        u_old[i] = ui;
        u[i] = ui + c * a[i];
        a[i] = 0.0;
    }
}

// corrector (:115-119): a += 1 Da ; u += (alpha dt) Da (daxpy.c:4-21) ;
// partial sums of a.a and Da.Da
__global__ __launch_bounds__(FG_TPB) void k_correct(int n, double s, const double *__restrict__ Da,
                                                    double *__restrict__ a,
                                                    double *__restrict__ u,
                                                    double *__restrict__ part)
{
    __shared__ double ws[FG_TPB / 64];
    double sa = 0.0, sd = 0.0;
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB) {
        const double d = Da[i];
        const double ai = a[i] + 1.0 * d;
        a[i] = ai;
        u[i] = u[i] + s * d;
        sa = sa + ai * ai;
        sd = sd + d * d;
    }
    sa = fg_block_sum(sa, ws);
    sd = fg_block_sum(sd, ws);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sa;
        part[gridDim.x + blockIdx.x] = sd;
    }
}

// StopBySteadyState (:44-63): partial sums of (u - u_old).(u - u_old), u.u
__global__ __launch_bounds__(FG_TPB) void k_steady(int n, const double *__restrict__ u,
                                                   const double *__restrict__ u_old,
                                                   double *__restrict__ part)
{
    __shared__ double ws[FG_TPB / 64];
    double sdiff = 0.0, su = 0.0;
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB) {
        const double ui = u[i], di = ui - u_old[i];
        sdiff = sdiff + di * di;
        su = su + ui * ui;
    }
    sdiff = fg_block_sum(sdiff, ws);
    su = fg_block_sum(su, ws);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sdiff;
        part[gridDim.x + blockIdx.x] = su;
    }
}

// the two totals of G block partials each, in block order (one workgroup)
__global__ __launch_bounds__(FG_TPB) void k_sum2(const double *__restrict__ part, int G,
                                                 double *__restrict__ out)
{
    __shared__ double ws[FG_TPB / 64];
    for (int k = 0; k < 2; k++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < G; b += FG_TPB) acc = acc + part[k * G + b];
        acc = fg_block_sum(acc, ws);
        if (threadIdx.x == 0) out[k] = acc;
    }
}

struct Work {
    void *buf = nullptr;
    ~Work() { hipFree(buf); }
};

int sums(fg_ctx *c, const double *part, int G, double *dev_out, double host_out[2])
{
    FG_LAUNCH(k_sum2, 1, FG_TPB, 0, c->stream, part, G, dev_out);
    FG_HIP(hipMemcpyAsync(host_out, dev_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
}

// Predictor_Old's two loops around one correction `build(u, a, R)` that
// leaves the element matrices in A (and R, possibly scaled) on the stream
template <typename Build>
int predictor_loop(fg_ctx *c, int neq, size_t work_doubles, fg_op *A, fg_prec *pc,
                   const fg_predictor_params *p, const fg_gmres_opts *gm, double *u,
                   int *corr_hist, int hist_cap, fg_predictor_info *info, double **work,
                   Build build)
{
    FG_ARG(!pc || pc->op == A, "the preconditioner must be built on A");
    FG_ARG(p->dt > 0.0 && p->nl_maxiter >= 1, "dt > 0 and nl_maxiter >= 1");
    FG_HIP(hipSetDevice(c->device));
    // workspace: a, Da, u_old, R (neq+1), partial sums, two totals, then the
    // caller's share (nodal U/dU, element matrices)
    const size_t nv = (size_t)neq + 1;
    Work w;
    const size_t bytes = sizeof(double) * (4 * nv + 2 * PG + 2 + work_doubles);
    FG_HIP(hipMalloc(&w.buf, bytes));
    double *a = (double *)w.buf, *Da = a + nv, *u_old = Da + nv, *R = u_old + nv;
    double *part = R + nv, *tot = part + 2 * PG;
    *work = tot + 2;
    FG_HIP(hipMemsetAsync(w.buf, 0, bytes, c->stream));

    const int G = fg_vec_grid(c, neq) < PG ? fg_vec_grid(c, neq) : PG;
    const double alpha = p->alpha, dt = p->dt;
    const double cpred = (1.0 - alpha) * dt, scorr = alpha * dt;
    fg_predictor_info inf;
    memset(&inf, 0, sizeof inf);
    // host clock: the loop synchronises at every correction anyway, and the
    // context's event pair is taken by fg_gmres
    FG_HIP(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    double t = 0.0;
    int passo = 0;
    for (;;) {
        passo++;
        t += dt;
        if (neq > 0) FG_LAUNCH(k_predict, G, FG_TPB, 0, c->stream, neq, cpred, u, u_old, a);
        int i = 0;
        for (;;) {
            i++;
            FG_CHECK(build(u, a, R));
            if (pc) FG_CHECK(fg_prec_setup(pc, R));
            fg_gmres_info gi;
            const int rc = fg_gmres(c, A, pc, R, Da, gm, &gi);
            if (rc == FG_ERR_NOCONV) inf.unconverged++;
            else if (rc != FG_OK) return rc;
            inf.krylov_steps += gi.steps;
            double s2[2] = {0.0, 0.0};
            if (neq > 0) {
                FG_LAUNCH(k_correct, G, FG_TPB, 0, c->stream, neq, scorr, Da, a, u, part);
                FG_CHECK(sums(c, part, G, tot, s2));
            }
            const double norm_a = sqrt(s2[0]), norm_Da = sqrt(s2[1]);
            const bool stop = p->stop_by_norm ? (norm_Da < p->nl_tol * norm_a || i >= p->nl_maxiter)
                                              : (i >= p->nl_maxiter);
            if (stop) break;
        }
        inf.corrections += i;
        if (corr_hist && passo - 1 < hist_cap) corr_hist[passo - 1] = i;
        bool stop_time;
        if (p->steady) {
            double s2[2] = {0.0, 0.0};
            if (neq > 0) {
                FG_LAUNCH(k_steady, G, FG_TPB, 0, c->stream, neq, u, u_old, part);
                FG_CHECK(sums(c, part, G, tot, s2));
            }
            stop_time = p->time_tol * sqrt(s2[1]) > sqrt(s2[0]) || fabs(p->final_time - t) < dt;
        } else {
            stop_time = fabs(p->final_time - t) < 1e-15;
        }
        if (stop_time) break;
        // StopByTime ends only when t lands within 1e-15 of the final time;
        // accumulated steps usually miss it by more (dt = 1e-4, T = 1 drifts
        // ~2e-14), and the reference then never stops.  Once t has passed the
        // final time by half a step the loop ends with FG_ERR_NOCONV; every
        // run the reference itself ends, ends here in the same step.
        if (!p->steady && t > p->final_time + 0.5 * dt) {
            inf.status = FG_ERR_NOCONV;
            break;
        }
        if (p->max_steps > 0 && passo >= p->max_steps) {
            inf.status = FG_ERR_NOCONV;
            break;
        }
    }
    FG_HIP(hipStreamSynchronize(c->stream));
    inf.steps = passo;
    inf.t = t;
    inf.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (info) *info = inf;
    return inf.status;
}

}  // namespace

extern "C" int fg_predictor_old_transport(fg_mesh2d *m, const fg_supg2d_problem *pb, fg_op *A,
                                          fg_prec *pc, const double *upresc,
                                          const fg_predictor_params *p,
                                          const fg_gmres_opts *gm, double *u,
                                          int *corr_hist, int hist_cap,
                                          fg_predictor_info *info)
{
    FG_ARG(m && pb && A && upresc && p && gm && u, "mesh/problem/operator/upresc/params/opts/u");
    int neq = 0, nel = 0, nn = 0;
    FG_CHECK(fg_mesh2d_info(m, &neq, &nel, &nn));
    fg_ctx *c = fg_mesh2d_ctx(m);
    FG_ARG(A->ctx == c && A->kind == FG_OP_EBE && A->k == 3 && A->neq == neq && A->nitems == nel,
           "A must be the k = 3 EBE operator of the mesh's lm");
    double *work = nullptr;
    double *const *wp = &work;
    const double alpha = p->alpha, dt = p->dt;
    // Build_M_K_R_SUPG.c with eval_U_dU.c of the 02 driver; NO_scaling
    auto build = [&](const double *uu, const double *aa, double *R) -> int {
        double *U = *wp, *dU = U + nn;
        FG_CHECK(fg_mesh2d_eval_U_dU(m, uu, aa, upresc, U, dU));
        // the element matrices go straight into A's storage (the values
        // fg_op_set_values_dev would store from the AoS array)
        return fg_supg2d_transient_build_op(m, pb, alpha, dt, U, dU, A, R);
    };
    return predictor_loop(c, neq, 2 * (size_t)nn + 1, A, pc, p, gm, u, corr_hist, hist_cap,
                          info, &work, build);
}

// the Euler driver's loop over its element (k = 12) or edge (k = 8, edges !=
// NULL) scheme; A is the product operator: the EBE/EDE operator itself or,
// for FG_PRED_EU_TWO, the ebe2/ede2 operator over it
static int euler_predictor(fg_euler2d *m, const fg_euler2d_params *ep, fg_ede4 *edges, fg_op *A,
                           fg_prec *blocks, int mode, const double *presc, const int *rot,
                           const double *theta, const fg_predictor_params *p,
                           const fg_gmres_opts *gm, double *u, int *corr_hist, int hist_cap,
                           fg_predictor_info *info)
{
    FG_ARG(m && ep && A && presc && p && gm && u, "mesh/params/operator/presc/params/opts/u");
    FG_ARG(mode == FG_PRED_EU_NOT || mode == FG_PRED_EU_BLOCK_SCALING ||
               mode == FG_PRED_EU_BLOCKDIAG || mode == FG_PRED_EU_TWO, "mode");
    FG_ARG(!(edges && mode == FG_PRED_EU_BLOCK_SCALING),
           "block scaling is defined for the element scheme only (setScaling.c)");
    fg_op *V = A;   // the operator that holds the matrices
    if (mode == FG_PRED_EU_TWO) {
        FG_ARG(A->kind == FG_OP_EBE2 && A->form == FG_EBE2_BLOCKDIAG && A->bd == blocks,
               "FG_PRED_EU_TWO needs A = fg_op_ebe2_create(base, blocks, FG_EBE2_BLOCKDIAG)");
        V = A->base;
    }
    FG_ARG(mode == FG_PRED_EU_NOT || (blocks && blocks->kind == FG_PREC_BLOCKDIAG && blocks->op == V),
           "block scaling and the block-diagonal preconditioner need a FG_PREC_BLOCKDIAG of A");
    fg_ctx *c = nullptr;
    int nn = 0, nel = 0, neq = 0, nedge = 0;
    fg_euler2d_sizes(m, &c, &nn, &nel, &neq);
    if (edges) {
        fg_ctx *ce = nullptr;
        int nel_e = 0;
        fg_ede4_sizes(edges, &ce, &nel_e, &nedge);
        FG_ARG(ce == c && nel_e == nel, "the edges must be those of the Euler mesh");
        FG_ARG(V->ctx == c && V->kind == FG_OP_EBE && V->k == 8 && V->neq == neq &&
                   V->nitems == nedge,
               "A must be the k = 8 operator of the edges' lm2");
    } else {
        FG_ARG(V->ctx == c && V->kind == FG_OP_EBE && V->k == 12 && V->neq == neq &&
                   V->nitems == nel,
               "A must be the k = 12 EBE operator of the Euler mesh's lm");
    }
    const size_t ne = (size_t)nel * 144, nd = (size_t)nedge * 64;
    double *work = nullptr;
    double *const *wp = &work;
    // Build_M_K_F_SUPG with Alpha_Build / DeltaT_Build set by the predictor
    // (Predictor_Old.c:27-28); the CAU tolerance and the rest from ep
    fg_euler2d_params bp = *ep;
    bp.alpha = p->alpha;
    bp.dt = p->dt;
    auto build = [&](const double *uu, const double *aa, double *R) -> int {
        double *U = *wp, *dU = U + 4 * (size_t)nn, *ke = dU + 4 * (size_t)nn, *kd = ke + ne;
        FG_CHECK(fg_euler2d_eval_U_dU(m, uu, aa, presc, U, dU));
        FG_CHECK(fg_euler2d_build(m, &bp, U, dU, rot, theta, ke, R));
        if (edges) {
            // ede_assembly.c: the element matrices into the edge matrices
            FG_CHECK(fg_ede4_assemble(edges, ke, kd));
            FG_CHECK(fg_op_set_values_dev(V, kd));
        } else {
            FG_CHECK(fg_op_set_values_dev(V, ke));
        }
        // Block_scaling_EBE (Block_scaling.c:10-59): A and R scaled in place
        if (mode == FG_PRED_EU_BLOCK_SCALING) FG_CHECK(fg_block_scaling_ebe(blocks, R));
        // EBE2/EDE2 with BlockDiag: the setup scales R, the solve has no
        // preconditioner (setPreconditioner.c:27-34)
        if (mode == FG_PRED_EU_TWO) FG_CHECK(fg_prec_setup(blocks, R));
        return FG_OK;
    };
    const size_t wd = 8 * (size_t)nn + ne + nd;
    return predictor_loop(c, neq, wd ? wd : 1, A, mode == FG_PRED_EU_BLOCKDIAG ? blocks : nullptr,
                          p, gm, u, corr_hist, hist_cap, info, &work, build);
}

extern "C" int fg_predictor_old_euler(fg_euler2d *m, const fg_euler2d_params *ep, fg_op *A,
                                      fg_prec *blocks, int mode, const double *presc,
                                      const int *rot, const double *theta,
                                      const fg_predictor_params *p, const fg_gmres_opts *gm,
                                      double *u, int *corr_hist, int hist_cap,
                                      fg_predictor_info *info)
{
    return euler_predictor(m, ep, nullptr, A, blocks, mode, presc, rot, theta, p, gm, u,
                           corr_hist, hist_cap, info);
}

extern "C" int fg_predictor_old_euler_ede(fg_euler2d *m, const fg_euler2d_params *ep,
                                          fg_ede4 *edges, fg_op *A, fg_prec *blocks, int mode,
                                          const double *presc, const int *rot,
                                          const double *theta, const fg_predictor_params *p,
                                          const fg_gmres_opts *gm, double *u, int *corr_hist,
                                          int hist_cap, fg_predictor_info *info)
{
    FG_ARG(edges, "edges");
    return euler_predictor(m, ep, edges, A, blocks, mode, presc, rot, theta, p, gm, u,
                           corr_hist, hist_cap, info);
}
