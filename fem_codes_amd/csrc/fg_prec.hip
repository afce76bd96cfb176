// fg_prec.hip -- preconditioners bound to an operator.
//
// Diag  : Solvers_and_Preconditioners/Diag_precond.c:3-14 with the setups
//         EBE/Diag_precond_EBE_setup.c:3-35, EDE/Diag_precond_EDE_setup.c:3-33,
//         CSR/Diag_precond_CSR_setup.c:3-29.
// Jacobi: EBE/Jacobi_precond_EBE.c:8-47 + EBE/Jacobi_precond_EBE_setup.c:4-43
//         (element-local diagonal, k=3 element operators).
// ILU(0): fg_ilu.hip.  Node block-diagonal (k = 12): fg_block.hip.
#include "fg_internal.h"

int fg_ebe_diag(fg_op *op, double *D);
int fg_csr_diag(fg_op *op, double *D);
int fg_ebe_jacobi_setup(fg_op *op, double **invde, float **invde32);
int fg_ebe_jacobi_diag32(fg_op *op, const double *invde, float *d32);
int fg_ebe_jacobi_apply(fg_op *op, int fp32, const void *invde, const void *p, void *z,
                        const int *skip);

// invDiag = 1/Diag ; F *= invDiag (first neq entries) ; EBE: invDiag[neq] = 1
__global__ void k_diag_finish(int neq, const double *__restrict__ D, double *__restrict__ iD,
                              float *__restrict__ iD32, double *__restrict__ F, int sink_one)
{
    for (int r = blockIdx.x * FG_TPB + threadIdx.x; r < neq; r += gridDim.x * FG_TPB) {
        double v = 1.0 / D[r];
        iD[r] = v;
        iD32[r] = (float)v;
        if (F) F[r] = F[r] * v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        iD[neq] = sink_one ? 1.0 : 0.0;
        iD32[neq] = sink_one ? 1.0f : 0.0f;
    }
}

template <typename T>
__global__ void k_diag_apply(int neq, const T *__restrict__ iD, const T *p, T *z, const int *skip)
{
    if (skip && *skip) return;
    for (int r = blockIdx.x * FG_TPB + threadIdx.x; r < neq; r += gridDim.x * FG_TPB)
        z[r] = p[r] * iD[r];
}

extern "C" int fg_prec_create(fg_op *op, int kind, fg_prec **out)
{
    FG_ARG(op && out, "op/out");
    FG_ARG(kind == FG_PREC_NONE || kind == FG_PREC_DIAG || kind == FG_PREC_JACOBI ||
               kind == FG_PREC_ILU0, "kind");
    FG_ARG(kind != FG_PREC_JACOBI || op->kind == FG_OP_EBE, "Jacobi needs an EBE operator");
    FG_ARG(kind != FG_PREC_ILU0 || op->kind == FG_OP_CSR, "ILU(0) needs a CSR operator");
    FG_ARG(op->kind != FG_OP_EBE2, "bind preconditioners to the k = 12 EBE operator");
    FG_HIP(hipSetDevice(op->ctx->device));
    fg_prec *pc = new fg_prec();
    pc->op = op;
    pc->kind = kind;
    size_t n = (size_t)op->neq + 1;
    if (kind == FG_PREC_ILU0) {
        int rc = fg_ilu_create(op, &pc->ilu);
        if (rc != FG_OK) {
            delete pc;
            return rc;
        }
    }
    if (kind == FG_PREC_DIAG) {
        FG_HIP(hipMalloc(&pc->d_diag, n * sizeof(double)));
        FG_HIP(hipMalloc(&pc->d_invdiag, n * sizeof(double)));
        FG_HIP(hipMalloc(&pc->d_invdiag32, n * sizeof(float)));
        FG_HIP(hipMemset(pc->d_diag, 0, n * sizeof(double)));
        FG_HIP(hipMemset(pc->d_invdiag, 0, n * sizeof(double)));
        FG_HIP(hipMemset(pc->d_invdiag32, 0, n * sizeof(float)));
    }
    if (kind == FG_PREC_JACOBI) {
        FG_HIP(hipMalloc(&pc->d_tmp64, n * sizeof(double)));
        FG_HIP(hipMalloc(&pc->d_tmp32, n * sizeof(float)));
    }
    *out = pc;
    return FG_OK;
}

extern "C" int fg_prec_setup(fg_prec *pc, double *F)
{
    FG_ARG(pc, "prec");
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    int neq = op->neq;
    if (pc->kind == FG_PREC_NONE) return FG_OK;
    if (pc->kind == FG_PREC_BLOCKDIAG) return fg_nodeblk_setup(pc, F);
    if (pc->kind == FG_PREC_ILU0) {
        // ILUp_precond_setup.c:53-56: factorise, then F <- (LU)^{-1} F
        FG_CHECK(fg_ilu_factor(pc->ilu, op));
        if (F) FG_CHECK(fg_ilu_apply(pc->ilu, c, 0, F, F, nullptr));
        return FG_OK;
    }
    if (pc->kind == FG_PREC_DIAG) {
        if (op->kind == FG_OP_CSR) FG_CHECK(fg_csr_diag(op, pc->d_diag));
        else FG_CHECK(fg_ebe_diag(op, pc->d_diag));
        FG_HIP(hipMemsetAsync(pc->d_diag + neq, 0, sizeof(double), c->stream));
        FG_LAUNCH(k_diag_finish, fg_vec_grid(c, neq), FG_TPB, 0, c->stream, neq, pc->d_diag,
                  pc->d_invdiag, pc->d_invdiag32, F, op->kind == FG_OP_EBE ? 1 : 0);
        return FG_OK;
    }
    // Jacobi: per-incidence 1/A^e_jj, then F <- Jacobi(F) (setup :76-84)
    FG_CHECK(fg_ebe_jacobi_setup(op, &pc->d_invde, &pc->d_invde32));
    if (!pc->d_jsum32) FG_HIP(hipMalloc(&pc->d_jsum32, sizeof(float) * (neq + 1)));
    FG_CHECK(fg_ebe_jacobi_diag32(op, pc->d_invde, pc->d_jsum32));
    if (F) {
        FG_HIP(hipMemcpyAsync(pc->d_tmp64, F, sizeof(double) * (neq + 1),
                              hipMemcpyDeviceToDevice, c->stream));
        FG_CHECK(fg_ebe_jacobi_apply(op, 0, pc->d_invde, pc->d_tmp64, F, nullptr));
    }
    return FG_OK;
}

int fg_prec_apply_any(fg_prec *pc, int fp32, const void *p, void *z, const int *skip)
{
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    int neq = op->neq;
    if (pc->kind == FG_PREC_NONE) {
        if (p != z) {
            size_t s = fp32 ? sizeof(float) : sizeof(double);
            FG_HIP(hipMemcpyAsync(z, p, s * neq, hipMemcpyDeviceToDevice, c->stream));
        }
        return FG_OK;
    }
    if (pc->kind == FG_PREC_DIAG) {
        int g = fg_vec_grid(c, neq);
        if (fp32)
            FG_LAUNCH(k_diag_apply<float>, g, FG_TPB, 0, c->stream, neq, pc->d_invdiag32,
                      (const float *)p, (float *)z, skip);
        else
            FG_LAUNCH(k_diag_apply<double>, g, FG_TPB, 0, c->stream, neq, pc->d_invdiag,
                      (const double *)p, (double *)z, skip);
        return FG_OK;
    }
    if (pc->kind == FG_PREC_ILU0) return fg_ilu_apply(pc->ilu, c, fp32, p, z, skip);
    if (pc->kind == FG_PREC_BLOCKDIAG) return fg_nodeblk_apply(pc, fp32, p, z, skip);
    // element Jacobi reads p[g] and writes z[g] in the same thread: in place is safe
    return fg_ebe_jacobi_apply(op, fp32, fp32 ? (const void *)pc->d_invde32 : pc->d_invde, p, z,
                               skip);
}

extern "C" int fg_prec_apply_f64(fg_prec *pc, const double *p, double *z)
{
    FG_ARG(pc && p && z, "prec/p/z");
    return fg_prec_apply_any(pc, 0, p, z);
}

extern "C" int fg_prec_apply_f32(fg_prec *pc, const float *p, float *z)
{
    FG_ARG(pc && p && z, "prec/p/z");
    return fg_prec_apply_any(pc, 1, p, z);
}

__global__ void k_to_f32(int n, const double *__restrict__ a, float *__restrict__ b)
{
    for (int r = blockIdx.x * FG_TPB + threadIdx.x; r < n; r += gridDim.x * FG_TPB)
        b[r] = (float)a[r];
}

extern "C" int fg_prec_set_invdiag(fg_prec *pc, const double *iDh)
{
    FG_ARG(pc && pc->kind == FG_PREC_DIAG && iDh, "Diag preconditioner / invDiag");
    fg_ctx *c = pc->op->ctx;
    int n = pc->op->neq + 1;
    FG_HIP(hipMemcpyAsync(pc->d_invdiag, iDh, sizeof(double) * n, hipMemcpyHostToDevice,
                          c->stream));
    FG_LAUNCH(k_to_f32, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, pc->d_invdiag,
              pc->d_invdiag32);
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
}

extern "C" int fg_prec_get_diag(fg_prec *pc, double *Dh, double *iDh)
{
    FG_ARG(pc && pc->kind == FG_PREC_DIAG, "Diag preconditioner");
    fg_ctx *c = pc->op->ctx;
    size_t n = (size_t)pc->op->neq + 1;
    FG_HIP(hipStreamSynchronize(c->stream));
    if (Dh) FG_HIP(hipMemcpy(Dh, pc->d_diag, n * sizeof(double), hipMemcpyDeviceToHost));
    if (iDh) FG_HIP(hipMemcpy(iDh, pc->d_invdiag, n * sizeof(double), hipMemcpyDeviceToHost));
    return FG_OK;
}

extern "C" int fg_prec_destroy(fg_prec *pc)
{
    if (!pc) return FG_OK;
    hipStreamSynchronize(pc->op->ctx->stream);
    hipFree(pc->d_diag); hipFree(pc->d_invdiag); hipFree(pc->d_invdiag32);
    hipFree(pc->d_invde); hipFree(pc->d_invde32); hipFree(pc->d_tmp64); hipFree(pc->d_tmp32);
    hipFree(pc->d_jsum32);
    fg_ilu_destroy(pc->ilu);
    fg_nodeblk_destroy(pc->nb);
    delete pc;
    return FG_OK;
}

extern "C" int fg_prec_ilu_info(fg_prec *pc, int *nlev_lower, int *nlev_upper, int *nsegments)
{
    FG_ARG(pc && pc->kind == FG_PREC_ILU0, "ILU(0) preconditioner");
    fg_ilu_levels(pc->ilu, nlev_lower, nlev_upper, nsegments);
    return FG_OK;
}

extern "C" int fg_prec_ilu_schedule(fg_prec *pc, int wide_rows)
{
    FG_ARG(pc && pc->kind == FG_PREC_ILU0, "ILU(0) preconditioner");
    FG_ARG(wide_rows >= 0, "wide_rows >= 0");
    fg_ilu_set_wide(pc->ilu, wide_rows);
    fg_cfg_bump();
    return FG_OK;
}

extern "C" int fg_prec_ilu_sweeps(fg_prec *pc, int sweeps)
{
    FG_ARG(pc && pc->kind == FG_PREC_ILU0, "ILU(0) preconditioner");
    FG_ARG(sweeps >= 0 && sweeps <= 1000, "sweeps in [0,1000]");
    fg_cfg_bump();
    return fg_ilu_set_sweeps(pc->ilu, sweeps);
}

extern "C" int fg_prec_ilu_get(fg_prec *pc, double *LU_host, double *D_host)
{
    FG_ARG(pc && pc->kind == FG_PREC_ILU0, "ILU(0) preconditioner");
    return fg_ilu_get(pc->ilu, pc->op->ctx, LU_host, D_host);
}
