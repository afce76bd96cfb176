// fg_block.hip -- node block-diagonal of 4-dof, 3-node element operators
// (Euler 2D): the block preconditioner, the block scaling of the element
// matrices and the node pass of the off-block product.
//
//   setup  : BlockDiag_precond_EBE_setup.c:6-86 (node blocks summed in
//            element order, closed-form inverse matrix_Inverse :88-145, F
//            left-preconditioned)
//   apply  : BlockDiag_precond.c:3-27
//   scaling: Block_scaling.c:10-59 (boundary conditions :73-92, block rows
//            times the inverse block of their node :95-118)
//   ebe2   : ebe2mvNDOF4.c:4-79 = off-block element product (fg_ebe.hip)
//            + BlockDiag2_precond.c:3-27 or NOBlockDiag2_precond.c:3-25
//
// Device layout: node blocks and node dofs are stored entry-major,
// BD[16][nnodes], Id[4][nnodes], so one thread per node makes every load of
// a wavefront contiguous.  The node -> element incidences (element, vertex
// slot) are kept in ascending element order, the order the reference adds
// the blocks in.  Reads of the sink slot neq return 0.
#include <vector>

#include "fg_ebe_body.h"
#include "fg_internal.h"

int fg_ebe_apply_offblock(fg_op *op, int fp32, void *P, void *Q, const int *skip);

struct fg_nodeblk {
    int nnodes = 0, nel = 0;
    int k = 12, nv = 3;       // dofs and nodes per item: 12 / 3 (elements), 8 / 2 (edges)
    int *d_Id = nullptr;      // [4][nnodes] equation of dof d of node I (neq: prescribed)
    int *d_vert = nullptr;    // [nv][nel] nodes of each item
    int *d_noff = nullptr;    // [nnodes+1] incidence offsets
    int *d_ninc = nullptr;    // element*4 + vertex slot, element ascending
    double *d_BD = nullptr, *d_iBD = nullptr;   // [16][nnodes]
    float *d_BD32 = nullptr, *d_iBD32 = nullptr;
    int setup_done = 0;
};

namespace {

template <typename T>
int upload(const std::vector<T> &v, T **d)
{
    size_t n = v.size() ? v.size() : 1;
    FG_HIP(hipMalloc(d, n * sizeof(T)));
    if (!v.empty()) FG_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

// matrix_Inverse (BlockDiag_precond_EBE_setup.c:88-145): the cofactor
// inverse of a row-major 4x4 block in the reference's expression grouping
// (so the FP64 result matches it bit for bit).  Entry 14 is written there as
// c*j*m - b*g*m - c*e*n + a*g*n + b*e*o - a*f*o; the cofactor of the inverse
// has c*f*m in its first term.  It is kept as the reference has it.
__device__ __forceinline__ void inverse4(const double *B, double *R)
{
    const double a = B[0], b = B[1], c = B[2], d = B[3], e = B[4], f = B[5], g = B[6],
                 h = B[7], i = B[8], j = B[9], k = B[10], l = B[11], m = B[12], n = B[13],
                 o = B[14], p = B[15];
    const double det = (a*f*k*p - a*f*l*o - a*g*j*p + a*g*l*n + a*h*j*o - a*h*k*n
                        -b*e*k*p  + b*e*l*o + b*g*i*p - b*g*l*m - b*h*i*o + b*h*k*m +
                        c*(e*j*p - e*l*n - f*i*p + f*l*m + h*i*n - h*j*m) +
                        d*(-e*j*o + e*k*n + f*i*o - f*k*m - g*i*n + g*j*m));
    const double q = 1.0 / det;
    R[0] = q*(-h*k*n + g*l*n + h*j*o - f*l*o - g*j*p + f*k*p);
    R[1] = q*(d*k*n - c*l*n - d*j*o + b*l*o + c*j*p - b*k*p);
    R[2] = q*(-d*g*n + c*h*n + d*f*o - b*h*o - c*f*p + b*g*p);
    R[3] = q*(d*g*j - c*h*j - d*f*k + b*h*k + c*f*l - b*g*l);
    R[4] = q*(h*k*m - g*l*m - h*i*o + e*l*o + g*i*p - e*k*p);
    R[5] = q*(-d*k*m + c*l*m + d*i*o - a*l*o - c*i*p + a*k*p);
    R[6] = q*(d*g*m - c*h*m - d*e*o + a*h*o + c*e*p - a*g*p);
    R[7] = q*(-d*g*i + c*h*i + d*e*k - a*h*k - c*e*l + a*g*l);
    R[8] = q*(-h*j*m + f*l*m + h*i*n - e*l*n - f*i*p + e*j*p);
    R[9] = q*(d*j*m - b*l*m - d*i*n + a*l*n + b*i*p - a*j*p);
    R[10] = q*(-d*f*m + b*h*m + d*e*n - a*h*n - b*e*p + a*f*p);
    R[11] = q*(d*f*i - b*h*i - d*e*j + a*h*j + b*e*l - a*f*l);
    R[12] = q*(g*j*m - f*k*m - g*i*n + e*k*n + f*i*o - e*j*o);
    R[13] = q*(-c*j*m + b*k*m + c*i*n - a*k*n - b*i*o + a*j*o);
    R[14] = q*(c*j*m - b*g*m - c*e*n + a*g*n + b*e*o - a*f*o);
    R[15] = q*(-c*f*i + b*g*i + c*e*j - a*g*j - b*e*k + a*f*k);
}

__device__ __forceinline__ bool node_free(const int *Id, int nn, int I, int neq)
{
    return Id[I] != neq || Id[nn + I] != neq || Id[2 * nn + I] != neq || Id[3 * nn + I] != neq;
}

// node blocks: BD_I = sum over the items e containing I (ascending) of the
// 4x4 block of A_e at (4v, 4v); entry (r, c) of that block is item entry
// 4(k+1) v + k r + c (I_diag: 52 v + 12 r + c for the 12-dof element,
// BlockDiag_precond_EBE_setup.c:18-19; 36 v + 8 r + c for the 8-dof edge,
// BlockDiag_precond_EDE_setup.c:22).  Then the inverse, and FP32 copies.
__global__ void k_bd_setup(int nn, int neq, int k, const int *__restrict__ Id,
                           const int *__restrict__ noff, const int *__restrict__ ninc,
                           const long long *__restrict__ kpos, const int *__restrict__ kstride,
                           const double *__restrict__ K64, double *__restrict__ BD,
                           double *__restrict__ iBD, float *__restrict__ BD32,
                           float *__restrict__ iBD32)
{
    for (int I = blockIdx.x * FG_TPB + threadIdx.x; I < nn; I += gridDim.x * FG_TPB) {
        double s[16], r[16];
        for (int q = 0; q < 16; q++) s[q] = 0.0;
        const bool act = node_free(Id, nn, I, neq);
        if (act) {
            for (int t = noff[I]; t < noff[I + 1]; t++) {
                const int e = ninc[t] >> 2, v = ninc[t] & 3;
                const long long p = kpos[e];
                if (p < 0) continue;
                const long long ks = kstride[e];
#pragma unroll
                for (int q = 0; q < 16; q++)
                    s[q] = s[q] + K64[fg_ebe_vidx(p, ks, 4 * (k + 1) * v + k * (q >> 2) + (q & 3),
                                                  4)];
            }
            inverse4(s, r);
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) r[q] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            BD[(size_t)q * nn + I] = s[q];
            iBD[(size_t)q * nn + I] = r[q];
            BD32[(size_t)q * nn + I] = (float)s[q];
            iBD32[(size_t)q * nn + I] = (float)r[q];
        }
    }
}

template <typename T>
__device__ __forceinline__ T sink0(const T *x, int g, int neq) { return g == neq ? T(0) : x[g]; }

// BlockDiag_precond.c:11-21: z_I = M_I p_I, row sums left to right
template <typename T>
__global__ void k_bd_apply(int nn, int neq, const int *__restrict__ Id, const T *__restrict__ M,
                           const T *p, T *z, const int *skip)
{
    if (skip && *skip) return;
    for (int I = blockIdx.x * FG_TPB + threadIdx.x; I < nn; I += gridDim.x * FG_TPB) {
        int g[4];
        T x[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            g[d] = Id[(size_t)d * nn + I];
            x[d] = sink0(p, g[d], neq);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            if (g[rr] == neq) continue;
            const T *m = M + (size_t)(4 * rr) * nn + I;
            z[g[rr]] = m[0] * x[0] + m[nn] * x[1] + m[2 * (size_t)nn] * x[2] +
                       m[3 * (size_t)nn] * x[3];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) z[neq] = T(0);
}

// node pass of ebe2mvNDOF4.c:72-75 after the off-block product Q:
//   FORM 0, BlockDiag2_precond.c:11-20:   Q_I = P_I + invBD_I Q_I
//   FORM 1, NOBlockDiag2_precond.c:11-20: Q_I = Q_I + BD_I P_I
template <typename T, int FORM>
__global__ void k_ebe2_nodes(int nn, int neq, const int *__restrict__ Id,
                             const T *__restrict__ M, const T *__restrict__ P, T *__restrict__ Q,
                             const int *skip)
{
    if (skip && *skip) return;
    for (int I = blockIdx.x * FG_TPB + threadIdx.x; I < nn; I += gridDim.x * FG_TPB) {
        int g[4];
        T x[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            g[d] = Id[(size_t)d * nn + I];
            x[d] = FORM == 0 ? sink0(Q, g[d], neq) : sink0(P, g[d], neq);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            if (g[rr] == neq) continue;
            const T *m = M + (size_t)(4 * rr) * nn + I;
            if (FORM == 0)
                Q[g[rr]] = P[g[rr]] + m[0] * x[0] + m[nn] * x[1] + m[2 * (size_t)nn] * x[2] +
                           m[3 * (size_t)nn] * x[3];
            else
                Q[g[rr]] = Q[g[rr]] + (m[0] * x[0] + m[nn] * x[1] + m[2 * (size_t)nn] * x[2] +
                                       m[3 * (size_t)nn] * x[3]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) Q[neq] = T(0);
}

// Block_scaling.c:27-56, one thread per (element, block row v, block column
// w).  After matrix_Bondary_Conditions (:73-92) the block (v, w) reads
//   1 on the diagonal of row i when v == w and dof i of node v is prescribed,
//   0 in row i when dof i of node v is prescribed,
//   0 in column j when dof j of node w is prescribed,
//   A otherwise;
// matrix_Scaling (:95-118) then writes sum_j invBD_v[i][j] M[j][k], j
// ascending from a zero start.
__global__ void k_block_scale(int nel, int nn, int neq, const int *__restrict__ Id,
                              const int *__restrict__ vert, const double *__restrict__ iBD,
                              const long long *__restrict__ kpos,
                              const int *__restrict__ kstride, double *__restrict__ K64,
                              float *__restrict__ K32)
{
    const long long n = (long long)nel * 9;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const int e = (int)(t / 9), vw = (int)(t - (long long)e * 9);
        const int v = vw / 3, w = vw - 3 * v;
        const long long p = kpos[e];
        if (p < 0) continue;
        const long long ks = kstride[e];
        const int Iv = vert[(size_t)v * nel + e], Iw = vert[(size_t)w * nel + e];
        bool pv[4], pw[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            pv[d] = Id[(size_t)d * nn + Iv] == neq;
            pw[d] = Id[(size_t)d * nn + Iw] == neq;
        }
        double M[16];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                double a = K64[fg_ebe_vidx(p, ks, 12 * (4 * v + i) + 4 * w + j, 4)];
                if (pv[i] || pw[j]) a = (v == w && i == j && pv[i]) ? 1.0 : 0.0;
                M[4 * i + j] = a;
            }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double m[4];
#pragma unroll
            for (int j = 0; j < 4; j++) m[j] = iBD[(size_t)(4 * i + j) * nn + Iv];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < 4; j++) s = s + m[j] * M[4 * j + k];
                const long long dst = fg_ebe_vidx(p, ks, 12 * (4 * v + i) + 4 * w + k, 4);
                K64[dst] = s;
                K32[dst] = (float)s;
            }
        }
    }
}

// entry-major [16][nn] -> row-major node blocks [nn][16]
__global__ void k_bd_to_aos(int nn, const double *__restrict__ S, double *__restrict__ A)
{
    const long long n = (long long)nn * 16;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const long long I = t >> 4;
        const int q = (int)(t & 15);
        A[t] = S[(size_t)q * nn + I];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void fg_nodeblk_destroy(fg_nodeblk *b)
{
    if (!b) return;
    hipFree(b->d_Id); hipFree(b->d_vert); hipFree(b->d_noff); hipFree(b->d_ninc);
    hipFree(b->d_BD); hipFree(b->d_iBD); hipFree(b->d_BD32); hipFree(b->d_iBD32);
    delete b;
}

extern "C" int fg_prec_blockdiag_create(fg_op *op, int nnodes, const int *vertex, const int *ids,
                                        fg_prec **out)
{
    FG_ARG(op && vertex && ids && out, "op/vertex/ids/out");
    FG_ARG(op->kind == FG_OP_EBE && (op->k == 12 || op->k == 8),
           "the node block-diagonal needs a k = 12 (element) or k = 8 (edge) operator");
    FG_ARG(nnodes > 0, "nnodes > 0");
    const int nel = op->nitems, neq = op->neq, K = op->k, NV = K / 4;
    for (long long t = 0; t < (long long)NV * nel; t++)
        FG_ARG(vertex[t] >= 0 && vertex[t] < nnodes, "vertex out of range");
    std::vector<int> Id(4 * (size_t)nnodes), seen(neq, 0);
    for (int I = 0; I < nnodes; I++)
        for (int d = 0; d < 4; d++) {
            int g = ids[4 * (size_t)I + d];
            FG_ARG(g <= neq, "node dof id > neq");
            if (g < 0) g = neq;
            if (g < neq) {
                FG_ARG(!seen[g], "two node dofs share an equation");
                seen[g] = 1;
            }
            Id[(size_t)d * nnodes + I] = g;
        }
    // the operator's lm must hold these node dofs (Fill_LM.c:7-24 / the edge
    // lm of ede_Initialization.c: DOF fastest per node)
    {
        std::vector<int> lm((size_t)K * nel);
        if (nel) FG_HIP(hipMemcpy(lm.data(), op->d_lm, lm.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int e = 0; e < nel; e++)
            for (int v = 0; v < NV; v++)
                for (int d = 0; d < 4; d++)
                    FG_ARG(lm[(size_t)K * e + 4 * v + d] ==
                               Id[(size_t)d * nnodes + vertex[(size_t)NV * e + v]],
                           "lm does not match vertex/ids (node dofs, DOF fastest)");
    }
    std::vector<int> vert((size_t)NV * nel), noff(nnodes + 1, 0);
    for (int e = 0; e < nel; e++)
        for (int v = 0; v < NV; v++) {
            vert[(size_t)v * nel + e] = vertex[(size_t)NV * e + v];
            noff[vertex[(size_t)NV * e + v] + 1]++;
        }
    for (int I = 0; I < nnodes; I++) noff[I + 1] += noff[I];
    std::vector<int> ninc(noff[nnodes]);
    {
        std::vector<int> fill(noff.begin(), noff.end() - 1);
        for (int e = 0; e < nel; e++)
            for (int v = 0; v < NV; v++) ninc[fill[vertex[(size_t)NV * e + v]]++] = 4 * e + v;
    }
    FG_ARG(nel < (1 << 29), "too many elements");
    FG_HIP(hipSetDevice(op->ctx->device));
    fg_nodeblk *b = new fg_nodeblk();
    b->nnodes = nnodes;
    b->nel = nel;
    b->k = K;
    b->nv = NV;
    int rc = FG_OK;
    if ((rc = upload(Id, &b->d_Id)) != FG_OK || (rc = upload(vert, &b->d_vert)) != FG_OK ||
        (rc = upload(noff, &b->d_noff)) != FG_OK || (rc = upload(ninc, &b->d_ninc)) != FG_OK) {
        fg_nodeblk_destroy(b);
        return rc;
    }
    const size_t nb = 16 * (size_t)nnodes;
    if (hipMalloc(&b->d_BD, nb * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->d_iBD, nb * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->d_BD32, nb * sizeof(float)) != hipSuccess ||
        hipMalloc(&b->d_iBD32, nb * sizeof(float)) != hipSuccess) {
        fg_nodeblk_destroy(b);
        fg_set_error("node blocks: out of device memory");
        return FG_ERR_NOMEM;
    }
    hipMemset(b->d_BD, 0, nb * sizeof(double));
    hipMemset(b->d_iBD, 0, nb * sizeof(double));
    hipMemset(b->d_BD32, 0, nb * sizeof(float));
    hipMemset(b->d_iBD32, 0, nb * sizeof(float));
    fg_prec *pc = new fg_prec();
    pc->op = op;
    pc->kind = FG_PREC_BLOCKDIAG;
    pc->nb = b;
    *out = pc;
    return FG_OK;
}

int fg_nodeblk_setup(fg_prec *pc, double *F)
{
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    fg_nodeblk *b = pc->nb;
    FG_LAUNCH(k_bd_setup, fg_vec_grid(c, b->nnodes), FG_TPB, 0, c->stream, b->nnodes, op->neq,
              b->k, b->d_Id, b->d_noff, b->d_ninc, op->d_kpos, op->d_kstride, op->d_K64, b->d_BD,
              b->d_iBD, b->d_BD32, b->d_iBD32);
    b->setup_done = 1;
    if (F)
        FG_LAUNCH(k_bd_apply<double>, fg_vec_grid(c, b->nnodes), FG_TPB, 0, c->stream,
                  b->nnodes, op->neq, b->d_Id, b->d_iBD, (const double *)F, F, nullptr);
    return FG_OK;
}

int fg_nodeblk_apply(fg_prec *pc, int fp32, const void *p, void *z, const int *skip)
{
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    fg_nodeblk *b = pc->nb;
    const int g = fg_vec_grid(c, b->nnodes);
    if (fp32)
        FG_LAUNCH(k_bd_apply<float>, g, FG_TPB, 0, c->stream, b->nnodes, op->neq, b->d_Id,
                  b->d_iBD32, (const float *)p, (float *)z, skip);
    else
        FG_LAUNCH(k_bd_apply<double>, g, FG_TPB, 0, c->stream, b->nnodes, op->neq, b->d_Id,
                  b->d_iBD, (const double *)p, (double *)z, skip);
    return FG_OK;
}

extern "C" int fg_prec_blockdiag_get(fg_prec *pc, double *BD_host, double *invBD_host)
{
    FG_ARG(pc && pc->kind == FG_PREC_BLOCKDIAG, "block-diagonal preconditioner");
    fg_ctx *c = pc->op->ctx;
    fg_nodeblk *b = pc->nb;
    const size_t nb = 16 * (size_t)b->nnodes;
    double *tmp = nullptr;
    FG_HIP(hipMalloc(&tmp, nb * sizeof(double)));
    const double *src[2] = {b->d_BD, b->d_iBD};
    double *dst[2] = {BD_host, invBD_host};
    for (int w = 0; w < 2; w++) {
        if (!dst[w]) continue;
        hipLaunchKernelGGL(k_bd_to_aos, dim3(fg_vec_grid(c, (long long)nb)), dim3(FG_TPB), 0,
                           c->stream, b->nnodes, src[w], tmp);
        hipError_t e = hipMemcpyAsync(dst[w], tmp, nb * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            hipFree(tmp);
            fg_set_error("fg_prec_blockdiag_get: %s", hipGetErrorString(e));
            return FG_ERR_HIP;
        }
    }
    FG_HIP(hipFree(tmp));
    return FG_OK;
}

extern "C" int fg_block_scaling_ebe(fg_prec *pc, double *F)
{
    FG_ARG(pc && pc->kind == FG_PREC_BLOCKDIAG, "block-diagonal preconditioner");
    FG_ARG(pc->nb->k == 12, "Block_scaling_EBE needs the 12-dof element operator "
                            "(the reference scales only the EBE scheme, setScaling.c:11-18)");
    fg_op *op = pc->op;
    fg_ctx *c = op->ctx;
    fg_nodeblk *b = pc->nb;
    FG_CHECK(fg_nodeblk_setup(pc, F));
    const long long n = 9LL * b->nel;
    if (n == 0) return FG_OK;
    FG_LAUNCH(k_block_scale, fg_vec_grid(c, n), FG_TPB, 0, c->stream, b->nel, b->nnodes,
              op->neq, b->d_Id, b->d_vert, b->d_iBD, op->d_kpos, op->d_kstride, op->d_K64,
              op->d_K32);
    op->rows_stale = 1;
    return FG_OK;
}

// ---- ebe2mvNDOF4 operator ---------------------------------------------------
extern "C" int fg_op_ebe2_create(fg_op *base, fg_prec *bd, int form, fg_op **out)
{
    FG_ARG(base && bd && out, "op/blockdiag/out");
    FG_ARG(base->kind == FG_OP_EBE && (base->k == 12 || base->k == 8),
           "ebe2 needs a k = 12 (element) or k = 8 (edge) operator");
    FG_ARG(bd->kind == FG_PREC_BLOCKDIAG && bd->op == base,
           "ebe2 needs the block-diagonal preconditioner of the same operator");
    FG_ARG(form == FG_EBE2_BLOCKDIAG || form == FG_EBE2_NOBLOCKDIAG, "form");
    fg_op *op = new fg_op();
    op->ctx = base->ctx;
    op->kind = FG_OP_EBE2;
    op->neq = base->neq;
    op->nitems = base->nitems;
    op->k = base->k;
    op->nnz = base->nnz;
    op->base = base;
    op->bd = bd;
    op->form = form;
    *out = op;
    return FG_OK;
}

int fg_ebe2_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    fg_op *base = op->base;
    fg_nodeblk *b = op->bd->nb;
    fg_ctx *c = op->ctx;
    if (!b->setup_done) {
        fg_set_error("ebe2 product: the block-diagonal preconditioner is not set up");
        return FG_ERR_ARG;
    }
    FG_CHECK(fg_ebe_apply_offblock(base, fp32, P, Q, skip));
    const int g = fg_vec_grid(c, b->nnodes);
    const int nn = b->nnodes, neq = op->neq;
    if (op->form == FG_EBE2_BLOCKDIAG) {
        if (fp32)
            FG_LAUNCH((k_ebe2_nodes<float, 0>), g, FG_TPB, 0, c->stream, nn, neq, b->d_Id,
                      b->d_iBD32, (const float *)P, (float *)Q, skip);
        else
            FG_LAUNCH((k_ebe2_nodes<double, 0>), g, FG_TPB, 0, c->stream, nn, neq, b->d_Id,
                      b->d_iBD, (const double *)P, (double *)Q, skip);
    } else {
        if (fp32)
            FG_LAUNCH((k_ebe2_nodes<float, 1>), g, FG_TPB, 0, c->stream, nn, neq, b->d_Id,
                      b->d_BD32, (const float *)P, (float *)Q, skip);
        else
            FG_LAUNCH((k_ebe2_nodes<double, 1>), g, FG_TPB, 0, c->stream, nn, neq, b->d_Id,
                      b->d_BD, (const double *)P, (double *)Q, skip);
    }
    return FG_OK;
}

// compulsory bytes of one ebe2 product: the off-block item product (k*k -
// 16 nv values and k equation numbers per item: 96 of 144 for an element, 32
// of 64 for an edge; P read, Q written) and the node pass (16 block values +
// 4 dof ids per node, P and Q read, Q written)
double fg_ebe2_bytes(const fg_op *op, int fp32)
{
    const double s = fp32 ? 4.0 : 8.0;
    const double nel = op->nitems, neq = op->neq, nn = op->bd->nb->nnodes;
    const double k = op->k, nv = op->k / 4;
    return nel * ((k * k - 16.0 * nv) * s + 4.0 * k) + 2.0 * neq * s + nn * (16.0 * s + 16.0) +
           3.0 * neq * s;
}
