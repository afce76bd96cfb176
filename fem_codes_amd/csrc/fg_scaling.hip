// fg_scaling.hip -- the diagonal scalings of Solvers_and_Preconditioners/
// scaling.c:10-180 on the device: Left (row i times invDiag[i]) and
// LeftRight (entry (i, j) times sqrt(invDiag[i] invDiag[j]), then F[i] times
// sqrt(Diag[i])), for the EBE / EDE (blocked values, each entry read and
// written once) and CSR (row loop) operators, and Left_unscaling
// (x[i] *= sqrt(invDiag[i])).  Diag and invDiag come from the Diag
// preconditioner's setup, which also multiplies F by invDiag, as the
// reference's scalings call Diag_precond_*_setup first.  The reference sets
// invDiag[neq] = invDiag[neq - 1] before scaling; entries of prescribed rows
// and columns (equation neq) read that value here too.
#include <math.h>

#include "fg_internal.h"
#include "fg_ebe_body.h"

namespace {

__device__ __forceinline__ double inv_at(const double *__restrict__ iD, int g, int neq)
{
    return iD[g < neq ? g : neq - 1];
}

__global__ void k_scale_items(long long n, int k, int neq, int mode, const int *__restrict__ lm,
                              const long long *__restrict__ kpos,
                              const int *__restrict__ kstride, const double *__restrict__ iD,
                              double *__restrict__ K64, float *__restrict__ K32)
{
    const int kk = k * k;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const long long e = t / kk;
        const int q = (int)(t - e * kk);
        const long long p = kpos[e];
        if (p < 0) continue;
        const int r = q / k, cc = q - r * k;
        const double di = inv_at(iD, lm[e * k + r], neq);
        const long long dst = fg_ebe_vidx(p, kstride[e], q, kk % 4 == 0 ? 4 : 1);
        double v = K64[dst];
        if (mode == FG_SCALING_LEFT) v = v * di;
        else v = v * sqrt(di * inv_at(iD, lm[e * k + cc], neq));
        K64[dst] = v;
        K32[dst] = (float)v;
    }
}

__global__ void k_scale_csr(int neq, int mode, const int *__restrict__ IA,
                            const int *__restrict__ JA, const double *__restrict__ iD,
                            double *__restrict__ AA, float *__restrict__ AA32)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < neq; i += gridDim.x * FG_TPB) {
        const double di = iD[i];
        for (int j = IA[i]; j < IA[i + 1]; j++) {
            double v = AA[j];
            if (mode == FG_SCALING_LEFT) v = v * di;
            else v = v * sqrt(di * inv_at(iD, JA[j], neq));
            AA[j] = v;
            if (AA32) AA32[j] = (float)v;
        }
    }
}

// F[i] *= sqrt(Diag[i]) (LeftRight) or x[i] *= sqrt(invDiag[i]) (unscaling)
__global__ void k_mul_sqrt(int neq, const double *__restrict__ d, double *__restrict__ x)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < neq; i += gridDim.x * FG_TPB)
        x[i] = x[i] * sqrt(d[i]);
}

}  // namespace

extern "C" int fg_scaling_diag(fg_prec *pc, int mode, double *F)
{
    FG_ARG(pc && pc->kind == FG_PREC_DIAG, "the Diag preconditioner of the operator");
    FG_ARG(mode == FG_SCALING_LEFT || mode == FG_SCALING_LEFTRIGHT, "mode");
    FG_ARG(F, "F (the reference scales FemStructs->F)");
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    FG_ARG(op->kind == FG_OP_EBE || op->kind == FG_OP_EDE || op->kind == FG_OP_CSR,
           "EBE, EDE or CSR operator");
    FG_ARG(op->neq >= 1, "neq >= 1 (the reference reads invDiag[neq - 1])");
    // Diag_precond_*_setup: Diag, invDiag, F *= invDiag
    FG_CHECK(fg_prec_setup(pc, F));
    const int neq = op->neq;
    if (op->kind == FG_OP_CSR) {
        FG_LAUNCH(k_scale_csr, fg_vec_grid(c, neq), FG_TPB, 0, c->stream, neq, mode, op->d_IA,
                  op->d_JA, pc->d_invdiag, op->d_AA, op->d_AA32);
        FG_CHECK(fg_csr_values_changed(op));
    } else {
        FG_ARG(op->d_lm, "operator without its element lm");
        const long long n = (long long)op->nitems * op->k * op->k;
        if (n > 0)
            FG_LAUNCH(k_scale_items, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, op->k, neq,
                      mode, op->d_lm, op->d_kpos, op->d_kstride, pc->d_invdiag, op->d_K64,
                      op->d_K32);
        op->rows_stale = 1;
    }
    if (mode == FG_SCALING_LEFTRIGHT)
        FG_LAUNCH(k_mul_sqrt, fg_vec_grid(c, neq), FG_TPB, 0, c->stream, neq, pc->d_diag, F);
    return FG_OK;
}

extern "C" int fg_unscaling_left(fg_prec *pc, double *x)
{
    FG_ARG(pc && pc->kind == FG_PREC_DIAG && x, "Diag preconditioner / x");
    fg_ctx *c = pc->op->ctx;
    FG_LAUNCH(k_mul_sqrt, fg_vec_grid(c, pc->op->neq), FG_TPB, 0, c->stream, pc->op->neq,
              pc->d_invdiag, x);
    return FG_OK;
}
