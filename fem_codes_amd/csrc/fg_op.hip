// fg_op.hip -- operator-independent entry points: values, apply, timing.
#include "fg_internal.h"

int fg_ebe_apply_f64(fg_op *op, double *P, double *Q, const int *skip);
int fg_ebe_apply_f32(fg_op *op, float *P, float *Q, const int *skip);
int fg_ebe_set_values_dev(fg_op *op, const double *A);
int fg_ebe_apply_main(fg_op *op, int fp32, void *P, void *Q);
void fg_ebe_free(fg_op *op);
int fg_csr_apply(fg_op *op, int fp32, const void *P, void *Q, const int *skip);
void fg_csr_free(fg_op *op);

__global__ void k_copy_cast(long long n, const double *__restrict__ a, double *__restrict__ d64,
                            float *__restrict__ d32)
{
    for (long long i = (long long)blockIdx.x * FG_TPB + threadIdx.x; i < n;
         i += (long long)gridDim.x * FG_TPB) {
        double v = a[i];
        d64[i] = v;
        d32[i] = (float)v;
    }
}

static long long op_nvals(const fg_op *op)
{
    if (op->kind == FG_OP_CSR) return op->nnz;
    return (long long)op->nitems * op->k * op->k;
}

int fg_ebe_get_values_dev(fg_op *op, double *A);

extern "C" int fg_op_get_values_host(fg_op *op, double *vals)
{
    FG_ARG(op && (op->kind == FG_OP_EBE || op->kind == FG_OP_EDE), "EBE/EDE operator");
    fg_ctx *c = op->ctx;
    long long n = op_nvals(op);
    if (n == 0) return FG_OK;
    FG_ARG(vals, "vals");
    double *tmp = nullptr;
    FG_HIP(hipMalloc(&tmp, sizeof(double) * n));
    int rc = fg_ebe_get_values_dev(op, tmp);
    if (rc == FG_OK) {
        hipError_t e = hipMemcpyAsync(vals, tmp, sizeof(double) * n, hipMemcpyDeviceToHost,
                                      c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            fg_set_error("fg_op_get_values_host: %s", hipGetErrorString(e));
            rc = FG_ERR_HIP;
        }
    }
    hipFree(tmp);
    return rc;
}

extern "C" int fg_op_set_values_dev(fg_op *op, const double *vals)
{
    FG_ARG(op, "op");
    FG_ARG(op->kind != FG_OP_EBE2, "an ebe2 operator reads the values of its base operator");
    fg_ctx *c = op->ctx;
    long long n = op_nvals(op);
    if (n == 0) return FG_OK;
    FG_ARG(vals, "vals");
    if (op->kind == FG_OP_CSR) {
        FG_LAUNCH(k_copy_cast, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, vals, op->d_AA,
                  op->d_AA32);
        return fg_csr_values_changed(op);
    }
    return fg_ebe_set_values_dev(op, vals);
}

extern "C" int fg_op_set_values_host(fg_op *op, const double *vals)
{
    FG_ARG(op, "op");
    FG_ARG(op->kind != FG_OP_EBE2, "an ebe2 operator reads the values of its base operator");
    fg_ctx *c = op->ctx;
    long long n = op_nvals(op);
    if (n == 0) return FG_OK;
    FG_ARG(vals, "vals");
    double *tmp = nullptr;
    FG_HIP(hipMalloc(&tmp, sizeof(double) * n));
    FG_HIP(hipMemcpyAsync(tmp, vals, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    int rc = fg_op_set_values_dev(op, tmp);
    FG_HIP(hipStreamSynchronize(c->stream));
    FG_HIP(hipFree(tmp));
    return rc;
}

int fg_op_apply_any(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    if (op->kind == FG_OP_CSR) return fg_csr_apply(op, fp32, P, Q, skip);
    if (op->kind == FG_OP_EBE2) return fg_ebe2_apply(op, fp32, P, Q, skip);
    if (fp32) return fg_ebe_apply_f32(op, (float *)P, (float *)Q, skip);
    return fg_ebe_apply_f64(op, (double *)P, (double *)Q, skip);
}

extern "C" int fg_op_apply_f64(fg_op *op, double *P, double *Q)
{
    FG_ARG(op && P && Q, "op/P/Q");
    return fg_op_apply_any(op, 0, P, Q);
}

extern "C" int fg_op_apply_f32(fg_op *op, float *P, float *Q)
{
    FG_ARG(op && P && Q, "op/P/Q");
    return fg_op_apply_any(op, 1, P, Q);
}

extern "C" int fg_op_info(const fg_op *op, int *neq, int *nitems, int *k, long long *nnz)
{
    FG_ARG(op, "op");
    if (neq) *neq = op->neq;
    if (nitems) *nitems = op->nitems;
    if (k) *k = op->k;
    if (nnz) *nnz = op->nnz;
    return FG_OK;
}

extern "C" int fg_op_set_pipeline(fg_op *op, int wg_per_cu)
{
    FG_ARG(op, "op");
    FG_ARG(op->kind == FG_OP_EBE || op->kind == FG_OP_EDE, "pipeline applies to EBE/EDE operators");
    FG_ARG(wg_per_cu >= 0 && wg_per_cu <= 16, "wg_per_cu in [0,16]");
    op->pipe_wpc = wg_per_cu;
    fg_cfg_bump();
    return FG_OK;
}

extern "C" int fg_op_set_row_kernel(fg_op *op, int kind)
{
    FG_ARG(op, "op");
    FG_ARG(kind == FG_ROWS_VALU || kind == FG_ROWS_MFMA, "kind: FG_ROWS_VALU or FG_ROWS_MFMA");
    if (kind == FG_ROWS_MFMA)
        FG_ARG(op->kind == FG_OP_EBE && op->k == 12 && op->ept == 1,
               "matrix-core rows need a 12-dof element operator with 256 elements per block");
    op->rows_mfma = kind == FG_ROWS_MFMA;
    fg_cfg_bump();
    return FG_OK;
}

extern "C" int fg_op_set_edge_kernel(fg_op *op, int kind)
{
    FG_ARG(op, "op");
    FG_ARG(op->kind == FG_OP_EDE && op->k == 2, "edge kernels apply to EDE operators");
    FG_ARG(kind == FG_EDGE_ROWS || kind == FG_EDGE_BLOCKED || kind == FG_EDGE_AUTO,
           "kind: FG_EDGE_ROWS, FG_EDGE_BLOCKED or FG_EDGE_AUTO");
    op->rows_kind[0] = op->rows_kind[1] = kind;
    fg_cfg_bump();
    return FG_OK;
}

extern "C" int fg_op_set_product_kernel(fg_op *op, int precisions, int kind)
{
    FG_ARG(op, "op");
    FG_ARG(op->kind == FG_OP_EBE || op->kind == FG_OP_EDE,
           "the product kernel choice applies to EBE and EDE operators");
    FG_ARG(precisions >= 1 && precisions <= (FG_PRODUCT_FP64 | FG_PRODUCT_FP32),
           "precisions: FG_PRODUCT_FP64, FG_PRODUCT_FP32 or both");
    FG_ARG(kind == FG_PRODUCT_SLICED || kind == FG_PRODUCT_BLOCKED || kind == FG_PRODUCT_AUTO,
           "kind: FG_PRODUCT_SLICED, FG_PRODUCT_BLOCKED or FG_PRODUCT_AUTO");
    FG_ARG(kind != FG_PRODUCT_SLICED || op->d_ridx || op->d_hsrc,
           "this operator has no sliced rows (k other than 2, 3, 4, or column offsets "
           "beyond the row fields)");
    if (precisions & FG_PRODUCT_FP64) op->rows_kind[0] = kind;
    if (precisions & FG_PRODUCT_FP32) op->rows_kind[1] = kind;
    fg_cfg_bump();
    return FG_OK;
}

extern "C" int fg_op_product_kernel(const fg_op *op, int fp32, int *sliced)
{
    FG_ARG(op && sliced, "op/sliced");
    *sliced = (op->kind == FG_OP_EDE && fg_ede_rows_used(op, fp32)) ||
              fg_ebe_rows_used(op, fp32) || (op->kind == FG_OP_CSR && op->d_spos);
    if (*sliced && op->kind == FG_OP_EBE && op->d_rcode) *sliced = 2;
    return FG_OK;
}

extern "C" double fg_op_algorithmic_bytes(const fg_op *op, int fp32)
{
    if (!op) return 0.0;
    double s = fp32 ? 4.0 : 8.0;
    double neq = op->neq;
    if (op->kind == FG_OP_EBE2) return fg_ebe2_bytes(op, fp32);
    if (op->kind == FG_OP_CSR) {
        // the 16-bit column form moves 2 index bytes per entry and no IA
        if (op->d_spos && op->d_soff16) return (double)op->nnz * (s + 2.0) + 2.0 * neq * s;
        return (double)op->nnz * (s + 4.0) + 4.0 * (neq + 1.0) + 2.0 * neq * s;
    }
    double n = op->nitems, k = op->k;
    if (op->kind == FG_OP_EDE) return n * (4.0 * s + 8.0) + 2.0 * neq * s;
    return n * (k * k * s + 4.0 * k) + 2.0 * neq * s;
}

extern "C" int fg_op_time_apply(fg_op *op, int fp32, void *P, void *Q, int reps, int main_only,
                                double *ms)
{
    FG_ARG(op && P && Q && reps > 0 && ms, "op/P/Q/reps/ms");
    fg_ctx *c = op->ctx;
    FG_CHECK(fg_op_apply_any(op, fp32, P, Q));   // warm
    if (!main_only || op->kind == FG_OP_CSR || op->kind == FG_OP_EBE2) {
        FG_HIP(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; r++) FG_CHECK(fg_op_apply_any(op, fp32, P, Q));
        FG_HIP(hipEventRecord(c->ev1, c->stream));
        FG_HIP(hipEventSynchronize(c->ev1));
        float t = 0.f;
        FG_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
        *ms = (double)t / reps;
        return FG_OK;
    }
    // the block kernel alone, launched back to back between two events (each
    // launch rewrites the same Q entries and halo partials, so repeats are
    // idempotent), then one full product so that Q is complete
    FG_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++) FG_CHECK(fg_ebe_apply_main(op, fp32, P, Q));
    FG_HIP(hipEventRecord(c->ev1, c->stream));
    FG_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    FG_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    FG_CHECK(fg_op_apply_any(op, fp32, P, Q));
    FG_HIP(hipStreamSynchronize(c->stream));
    *ms = (double)t / reps;
    return FG_OK;
}

extern "C" int fg_op_destroy(fg_op *op)
{
    if (!op) return FG_OK;
    hipStreamSynchronize(op->ctx->stream);
    if (op->kind == FG_OP_CSR) fg_csr_free(op);
    else if (op->kind != FG_OP_EBE2) fg_ebe_free(op);
    delete op;
    return FG_OK;
}

int fg_op_prepare(fg_op *op)
{
    if (op->kind == FG_OP_EBE2) return fg_op_prepare(op->base);
    if (op->kind == FG_OP_EDE) return fg_ede_rows_refresh(op);
    if (op->kind == FG_OP_EBE) return fg_ebe_rows_refresh(op);
    return FG_OK;
}
