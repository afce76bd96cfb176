// fg_ilu.hip -- ILU(0) preconditioner of a CSR operator on the device.
//
// Reference: Solvers_and_Preconditioners/CSR/ILUp_precond_setup.c:8-60
// (ILUp :66-167 with p = 0: ilup.c LEVEL_OF_FILL_lu adds no fill-in at level
// 0, so L and U have the pattern of A below and above the diagonal) and
// CSR/ILUp_precond.c:4-49 (forward solve with unit L, backward solve with U
// and the inverted pivots).
//
// The factors live in A's CSR slots (LU[nnz]; the diagonal slot unused) and
// D[neq] holds the inverted pivots, FP64 plus an FP32 copy for the mixed-
// precision solver.  Row i of the factorisation and of the forward solve
// depends on the rows of its L pattern; of the backward solve on the rows of
// its U pattern.  The host groups rows into dependency levels (level scheduling)
// and cuts the level sequence into segments:
//   wide levels  -> one launch per level, a thread per row;
//   runs of narrow levels -> one workgroup of 1024 threads that walks the
//      levels in order with a barrier between levels (reorderings such as RCM
//      produce thousands of narrow levels; a launch each would cost more than
//      the work).
// Within a row every update is applied in the reference's order (L entries
// ascending, U entries of the pivot row ascending), so the factors and the
// solves are bitwise those of the reference routines.
#include <string.h>

#include <algorithm>
#include <vector>

#include "fg_internal.h"
#include "fg_sell.h"

namespace {

constexpr int CHAIN_TPB = 1024;
constexpr int WIDE_LEVEL = 8192;   // rows: above this a level gets its own launch

struct Seg {
    int wide;      // 1: one level, thread per row ; 0: level range in one workgroup
    int l0, l1;    // level range [l0, l1)
    int r0, r1;    // row-list range [r0, r1)
};

struct Sched {
    int nlev = 0;
    int *d_rows = nullptr;   // rows ordered by level
    int *d_loff = nullptr;   // [nlev+1]
    std::vector<int> loff;   // host copy
    std::vector<Seg> segs;
};

// cut the level sequence into wide-level launches and narrow-level runs
void make_segs(Sched &s, int wide_rows)
{
    s.segs.clear();
    const std::vector<int> &loff = s.loff;
    for (int l = 0; l < s.nlev;) {
        if (loff[l + 1] - loff[l] > wide_rows) {
            s.segs.push_back(Seg{1, l, l + 1, loff[l], loff[l + 1]});
            l++;
        } else {
            int l1 = l;
            while (l1 < s.nlev && loff[l1 + 1] - loff[l1] <= wide_rows) l1++;
            s.segs.push_back(Seg{0, l, l1, loff[l], loff[l1]});
            l = l1;
        }
    }
}

}  // namespace

struct fg_ilu {
    int neq = 0, nnz = 0;
    const int *d_IA = nullptr, *d_JA = nullptr;   // the operator's pattern
    int *d_lend = nullptr;     // [neq] end of the L part of each row
    int *d_ubeg = nullptr;     // [neq] start of the U part
    int *d_dpos = nullptr;     // [neq] diagonal slot or -1
    double *d_LU = nullptr, *d_D = nullptr;
    float *d_LU32 = nullptr, *d_D32 = nullptr;
    int *d_err = nullptr;      // first zero pivot row + 1
    Sched lo, up;
    int sweeps = 0;            // 0: exact level-scheduled solves; s > 0: s Jacobi sweeps
    double *d_s64[2] = {nullptr, nullptr};   // sweep scratch
    float *d_s32[2] = {nullptr, nullptr};
    // strictly lower / upper parts as sliced rows (fg_sell.h) for the sweeps
    struct Tri {
        int nslice = 0;
        long long nslot = 0;
        int *d_soff = nullptr, *d_JA = nullptr;
        short *d_off16 = nullptr;   // columns as 16-bit offsets from the row (else d_JA)
        long long *d_map = nullptr;   // [nslot] factor slot of each entry (FG_SELL_NOSRC:
                                      // padding), FP64 pair shift in the top byte
        double *d_v64 = nullptr;
        float *d_v32 = nullptr;
    } lt, ut;
};

namespace {

int build_sched(int neq, const std::vector<int> &IA, const std::vector<int> &JA,
                const std::vector<int> &lend, const std::vector<int> &ubeg, bool lower,
                Sched &s)
{
    std::vector<int> lev(neq, 0);
    int nlev = 0;
    if (lower) {
        for (int i = 0; i < neq; i++) {
            int l = 0;
            for (int q = IA[i]; q < lend[i]; q++) l = std::max(l, lev[JA[q]] + 1);
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    } else {
        for (int i = neq - 1; i >= 0; i--) {
            int l = 0;
            for (int q = ubeg[i]; q < IA[i + 1]; q++) l = std::max(l, lev[JA[q]] + 1);
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    }
    std::vector<int> loff(nlev + 1, 0), rows(neq);
    for (int i = 0; i < neq; i++) loff[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) loff[l + 1] += loff[l];
    {
        std::vector<int> fill(loff.begin(), loff.end() - 1);
        for (int i = 0; i < neq; i++) rows[fill[lev[i]]++] = i;
    }
    s.nlev = nlev;
    s.loff = loff;
    make_segs(s, WIDE_LEVEL);
    FG_HIP(hipMalloc(&s.d_rows, sizeof(int) * (neq ? neq : 1)));
    FG_HIP(hipMalloc(&s.d_loff, sizeof(int) * (nlev + 1)));
    if (neq) FG_HIP(hipMemcpy(s.d_rows, rows.data(), sizeof(int) * neq, hipMemcpyHostToDevice));
    FG_HIP(hipMemcpy(s.d_loff, loff.data(), sizeof(int) * (nlev + 1), hipMemcpyHostToDevice));
    return FG_OK;
}

struct IluView {
    const int *IA, *JA, *lend, *ubeg, *dpos;
};

// one row of ILUp (p = 0): copy, eliminate with the previous rows, invert the pivot
__device__ __forceinline__ void factor_row(int i, const IluView &v, const double *__restrict__ AA,
                                           double *__restrict__ LU, double *__restrict__ D,
                                           int *err)
{
    const int q0 = v.IA[i], q1 = v.IA[i + 1], dp = v.dpos[i], le = v.lend[i];
    double Di = 0.0;
    for (int q = q0; q < q1; q++) {
        if (q == dp) Di = AA[q];
        else LU[q] = AA[q];
    }
    for (int q = q0; q < le; q++) {
        const int jrow = v.JA[q];
        const double l = LU[q] * D[jrow];
        LU[q] = l;
        int p = q + 1;
        for (int t = v.ubeg[jrow]; t < v.IA[jrow + 1]; t++) {
            const int col = v.JA[t];
            while (p < q1 && v.JA[p] < col) p++;
            if (p >= q1) break;
            if (v.JA[p] == col) {
                if (col == i) Di -= l * LU[t];
                else LU[p] -= l * LU[t];
            }
        }
    }
    if (Di == 0.0 && *err == 0) *err = i + 1;
    D[i] = 1.0 / Di;
}

__global__ __launch_bounds__(FG_TPB) void k_ilu_factor_wide(IluView v, const int *__restrict__ rows,
                                                            int r0, int r1,
                                                            const double *__restrict__ AA,
                                                            double *__restrict__ LU,
                                                            double *__restrict__ D, int *err)
{
    const int r = r0 + blockIdx.x * FG_TPB + threadIdx.x;
    if (r < r1) factor_row(rows[r], v, AA, LU, D, err);
}

__global__ __launch_bounds__(CHAIN_TPB) void k_ilu_factor_chain(IluView v,
                                                                const int *__restrict__ rows,
                                                                const int *__restrict__ loff,
                                                                int l0, int l1,
                                                                const double *__restrict__ AA,
                                                                double *__restrict__ LU,
                                                                double *__restrict__ D, int *err)
{
    for (int l = l0; l < l1; l++) {
        for (int r = loff[l] + threadIdx.x; r < loff[l + 1]; r += CHAIN_TPB)
            factor_row(rows[r], v, AA, LU, D, err);
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void lrow(int i, const IluView &v, const T *__restrict__ LU,
                                     const T *p, T *z)
{
    T x = p[i];
    for (int q = v.IA[i]; q < v.lend[i]; q++) x = x - z[v.JA[q]] * LU[q];
    z[i] = x;
}

template <typename T>
__device__ __forceinline__ void urow(int i, const IluView &v, const T *__restrict__ LU,
                                     const T *__restrict__ D, T *z)
{
    T x = z[i];
    for (int q = v.ubeg[i]; q < v.IA[i + 1]; q++) x = x - z[v.JA[q]] * LU[q];
    z[i] = x * D[i];
}

template <typename T, bool UP>
__global__ __launch_bounds__(FG_TPB) void k_ilu_solve_wide(IluView v, const int *__restrict__ rows,
                                                           int r0, int r1, const T *LU, const T *D,
                                                           const T *p, T *z, const int *skip)
{
    if (skip && *skip) return;
    const int r = r0 + blockIdx.x * FG_TPB + threadIdx.x;
    if (r >= r1) return;
    if (UP) urow<T>(rows[r], v, LU, D, z);
    else lrow<T>(rows[r], v, LU, p, z);
}

template <typename T, bool UP>
__global__ __launch_bounds__(CHAIN_TPB) void k_ilu_solve_chain(IluView v,
                                                               const int *__restrict__ rows,
                                                               const int *__restrict__ loff,
                                                               int l0, int l1, const T *LU,
                                                               const T *D, const T *p, T *z,
                                                               const int *skip)
{
    if (skip && *skip) return;
    for (int l = l0; l < l1; l++) {
        for (int r = loff[l] + threadIdx.x; r < loff[l + 1]; r += CHAIN_TPB) {
            if (UP) urow<T>(rows[r], v, LU, D, z);
            else lrow<T>(rows[r], v, LU, p, z);
        }
        __syncthreads();
    }
}

__global__ void k_ilu_cast(long long n, const double *__restrict__ a, float *__restrict__ b)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB)
        b[t] = (float)a[t];
}

IluView view(const fg_ilu *u)
{
    return IluView{u->d_IA, u->d_JA, u->d_lend, u->d_ubeg, u->d_dpos};
}

int blocks_for(int rows) { return (rows + FG_TPB - 1) / FG_TPB; }

}  // namespace

// the strictly lower (lower = true) or upper part of A's pattern as sliced
// rows, with the factor slot of each entry
static int build_tri(int neq, const std::vector<int> &IA, const std::vector<int> &JA,
                     const std::vector<int> &lend, const std::vector<int> &ubeg, bool lower,
                     fg_ilu::Tri &t)
{
    std::vector<int> tIA(neq + 1, 0);
    for (int i = 0; i < neq; i++)
        tIA[i + 1] = tIA[i] + (lower ? lend[i] - IA[i] : IA[i + 1] - ubeg[i]);
    const fg_sell_layout L = fg_sell_plan(neq, tIA.data());
    std::vector<int> col(L.nslot, -1);
    std::vector<short> off16(L.nslot, FG_SELL_OFF_PAD);
    bool fits16 = true;
    // per slot: the factor entry (FG_SELL_NOSRC: padding) and the FP64 pair shift
    std::vector<long long> map(L.nslot);
    for (int s = 0; s < L.nslice; s++) {
        const int w = L.soff[s + 1] - L.soff[s];
        for (int k = 0; k < w; k++)
            for (int l = 0; l < FG_SELL_W; l++)
                map[(long long)(L.soff[s] + k) * FG_SELL_W + l] =
                    fg_sell_pack(FG_SELL_NOSRC, fg_pair_shift(k, l, w));
    }
    for (int i = 0; i < neq; i++) {
        const int q0 = lower ? IA[i] : ubeg[i], q1 = lower ? lend[i] : IA[i + 1];
        const int s = i / FG_SELL_W, w = L.soff[s + 1] - L.soff[s];
        for (int q = q0; q < q1; q++) {
            const long long at = L.slot(i, q - q0);
            col[at] = JA[q];
            const int d = JA[q] - i;
            if (d < -32767 || d > 32767) fits16 = false;
            else off16[at] = (short)d;
            map[at] = fg_sell_pack(q, fg_pair_shift(q - q0, i % FG_SELL_W, w));
        }
    }
    t.nslice = L.nslice;
    t.nslot = L.nslot;
    const size_t ns = t.nslot ? t.nslot : 1;
    FG_HIP(hipMalloc(&t.d_soff, sizeof(int) * L.soff.size()));
    FG_HIP(hipMalloc(&t.d_JA, sizeof(int) * ns));
    FG_HIP(hipMalloc(&t.d_map, sizeof(long long) * ns));
    FG_HIP(hipMalloc(&t.d_v64, sizeof(double) * ns));
    FG_HIP(hipMalloc(&t.d_v32, sizeof(float) * ns));
    FG_HIP(hipMemcpy(t.d_soff, L.soff.data(), sizeof(int) * L.soff.size(),
                     hipMemcpyHostToDevice));
    if (t.nslot && fits16) {
        FG_HIP(hipMalloc(&t.d_off16, sizeof(short) * ns));
        FG_HIP(hipMemcpy(t.d_off16, off16.data(), sizeof(short) * t.nslot,
                         hipMemcpyHostToDevice));
    }
    if (t.nslot) {
        FG_HIP(hipMemcpy(t.d_JA, col.data(), sizeof(int) * t.nslot, hipMemcpyHostToDevice));
        FG_HIP(hipMemcpy(t.d_map, map.data(), sizeof(long long) * t.nslot,
                         hipMemcpyHostToDevice));
    }
    return FG_OK;
}

static void free_tri(fg_ilu::Tri &t)
{
    hipFree(t.d_soff); hipFree(t.d_JA); hipFree(t.d_off16); hipFree(t.d_map); hipFree(t.d_v64);
    hipFree(t.d_v32);
}

__global__ void k_tri_gather(long long n, const long long *__restrict__ map,
                             const double *__restrict__ LU, double *__restrict__ v64,
                             float *__restrict__ v32)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const long long q = fg_sell_packed_slot(map[t]);
        const double a = q == FG_SELL_NOSRC ? 0.0 : LU[q];
        v64[t + fg_sell_packed_shift(map[t])] = a;   // FP64 entry-paired
        v32[t] = (float)a;
    }
}

int fg_ilu_create(fg_op *op, fg_ilu **out)
{
    FG_ARG(op && op->kind == FG_OP_CSR, "ILU(0) needs a CSR operator");
    const int neq = op->neq;
    const std::vector<int> &IA = op->h_IA, &JA = op->h_JA;
    fg_ilu *u = new fg_ilu();
    u->neq = neq;
    u->nnz = (int)op->nnz;
    u->d_IA = op->d_IA;
    u->d_JA = op->d_JA;
    std::vector<int> lend(neq), ubeg(neq), dpos(neq);
    for (int i = 0; i < neq; i++) {
        int q = IA[i];
        while (q < IA[i + 1] && JA[q] < i) q++;
        lend[i] = q;
        dpos[i] = (q < IA[i + 1] && JA[q] == i) ? q : -1;
        ubeg[i] = dpos[i] >= 0 ? q + 1 : q;
    }
    FG_HIP(hipMalloc(&u->d_lend, sizeof(int) * (neq ? neq : 1)));
    FG_HIP(hipMalloc(&u->d_ubeg, sizeof(int) * (neq ? neq : 1)));
    FG_HIP(hipMalloc(&u->d_dpos, sizeof(int) * (neq ? neq : 1)));
    if (neq) {
        FG_HIP(hipMemcpy(u->d_lend, lend.data(), sizeof(int) * neq, hipMemcpyHostToDevice));
        FG_HIP(hipMemcpy(u->d_ubeg, ubeg.data(), sizeof(int) * neq, hipMemcpyHostToDevice));
        FG_HIP(hipMemcpy(u->d_dpos, dpos.data(), sizeof(int) * neq, hipMemcpyHostToDevice));
    }
    size_t nz = u->nnz ? u->nnz : 1, nq = neq + 1;
    FG_HIP(hipMalloc(&u->d_LU, sizeof(double) * nz));
    FG_HIP(hipMalloc(&u->d_LU32, sizeof(float) * nz));
    FG_HIP(hipMalloc(&u->d_D, sizeof(double) * nq));
    FG_HIP(hipMalloc(&u->d_D32, sizeof(float) * nq));
    FG_HIP(hipMalloc(&u->d_err, sizeof(int)));
    FG_HIP(hipMemset(u->d_LU, 0, sizeof(double) * nz));
    FG_HIP(hipMemset(u->d_D, 0, sizeof(double) * nq));
    FG_CHECK(build_sched(neq, IA, JA, lend, ubeg, true, u->lo));
    FG_CHECK(build_sched(neq, IA, JA, lend, ubeg, false, u->up));
    FG_CHECK(build_tri(neq, IA, JA, lend, ubeg, true, u->lt));
    FG_CHECK(build_tri(neq, IA, JA, lend, ubeg, false, u->ut));
    *out = u;
    return FG_OK;
}

void fg_ilu_destroy(fg_ilu *u)
{
    if (!u) return;
    hipFree(u->d_lend); hipFree(u->d_ubeg); hipFree(u->d_dpos);
    hipFree(u->d_LU); hipFree(u->d_LU32); hipFree(u->d_D); hipFree(u->d_D32);
    hipFree(u->d_err);
    for (int k = 0; k < 2; k++) {
        hipFree(u->d_s64[k]);
        hipFree(u->d_s32[k]);
    }
    free_tri(u->lt);
    free_tri(u->ut);
    hipFree(u->lo.d_rows); hipFree(u->lo.d_loff); hipFree(u->up.d_rows); hipFree(u->up.d_loff);
    delete u;
}

void fg_ilu_set_wide(fg_ilu *u, int wide_rows)
{
    make_segs(u->lo, wide_rows);
    make_segs(u->up, wide_rows);
}

void fg_ilu_levels(const fg_ilu *u, int *nlev_lower, int *nlev_upper, int *nseg)
{
    if (nlev_lower) *nlev_lower = u->lo.nlev;
    if (nlev_upper) *nlev_upper = u->up.nlev;
    if (nseg) *nseg = (int)(u->lo.segs.size() + u->up.segs.size());
}

// numeric factorisation from the operator's current FP64 values
int fg_ilu_factor(fg_ilu *u, fg_op *op)
{
    fg_ctx *c = op->ctx;
    const IluView v = view(u);
    FG_HIP(hipMemsetAsync(u->d_err, 0, sizeof(int), c->stream));
    for (const Seg &s : u->lo.segs) {
        if (s.wide)
            FG_LAUNCH(k_ilu_factor_wide, blocks_for(s.r1 - s.r0), FG_TPB, 0, c->stream, v,
                      (const int *)u->lo.d_rows, s.r0, s.r1, (const double *)op->d_AA, u->d_LU,
                      u->d_D, u->d_err);
        else
            FG_LAUNCH(k_ilu_factor_chain, 1, CHAIN_TPB, 0, c->stream, v,
                      (const int *)u->lo.d_rows, (const int *)u->lo.d_loff, s.l0, s.l1,
                      (const double *)op->d_AA, u->d_LU, u->d_D, u->d_err);
    }
    int err = 0;
    FG_HIP(hipMemcpyAsync(&err, u->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    if (err) {
        fg_set_error("ILU(0): zero pivot at row %d", err - 1);
        return FG_ERR_BREAKDOWN;
    }
    if (u->nnz)
        FG_LAUNCH(k_ilu_cast, fg_vec_grid(c, u->nnz), FG_TPB, 0, c->stream, (long long)u->nnz,
                  u->d_LU, u->d_LU32);
    if (u->neq)
        FG_LAUNCH(k_ilu_cast, fg_vec_grid(c, u->neq), FG_TPB, 0, c->stream, (long long)u->neq,
                  u->d_D, u->d_D32);
    for (fg_ilu::Tri *t : {&u->lt, &u->ut})
        if (t->nslot)
            FG_LAUNCH(k_tri_gather, fg_vec_grid(c, t->nslot), FG_TPB, 0, c->stream, t->nslot,
                      (const long long *)t->d_map, (const double *)u->d_LU, t->d_v64,
                      t->d_v32);
    return FG_OK;
}

template <typename T>
static int ilu_apply_t(fg_ilu *u, fg_ctx *c, const T *LU, const T *D, const T *p, T *z,
                       const int *skip)
{
    const IluView v = view(u);
    for (const Seg &s : u->lo.segs) {
        if (s.wide)
            FG_LAUNCH((k_ilu_solve_wide<T, false>), blocks_for(s.r1 - s.r0), FG_TPB, 0, c->stream,
                      v, (const int *)u->lo.d_rows, s.r0, s.r1, LU, D, p, z, skip);
        else
            FG_LAUNCH((k_ilu_solve_chain<T, false>), 1, CHAIN_TPB, 0, c->stream, v,
                      (const int *)u->lo.d_rows, (const int *)u->lo.d_loff, s.l0, s.l1, LU, D, p,
                      z, skip);
    }
    for (const Seg &s : u->up.segs) {
        if (s.wide)
            FG_LAUNCH((k_ilu_solve_wide<T, true>), blocks_for(s.r1 - s.r0), FG_TPB, 0, c->stream,
                      v, (const int *)u->up.d_rows, s.r0, s.r1, LU, D, (const T *)z, z, skip);
        else
            FG_LAUNCH((k_ilu_solve_chain<T, true>), 1, CHAIN_TPB, 0, c->stream, v,
                      (const int *)u->up.d_rows, (const int *)u->up.d_loff, s.l0, s.l1, LU, D,
                      (const T *)z, z, skip);
    }
    return FG_OK;
}

// ---- approximate triangular solves: fixed Jacobi sweeps ---------------------
// forward  y_{m+1} = p - L y_m,          y_0 = p
// backward x_{m+1} = D (y - U x_m),      x_0 = D y
// A fixed sweep count is a fixed linear operator (a truncated Neumann series
// of the exact solve, exact once it reaches the number of levels), so GMRES
// can use it as a preconditioner; every sweep is one fully parallel pass.
// out[r] = base[r] - (T x)[r]            (UPPER = false: a forward sweep)
// out[r] = D[r] * (base[r] - (T x)[r])   (UPPER = true: a backward sweep)
// T is the strictly lower or upper part as sliced rows (fg_sell.h): one
// row per lane, every load of the row issued at once, the terms added in the
// row's entry order.
template <typename T, bool UPPER>
__global__ __launch_bounds__(FG_TPB) void k_tri_sweep(int nslice, const int *__restrict__ soff,
                                                      const int *__restrict__ JA,
                                                      const short *__restrict__ off16,
                                                      const T *__restrict__ V,
                                                      const T *__restrict__ x,
                                                      const T *__restrict__ base,
                                                      const T *__restrict__ D,
                                                      T *__restrict__ out, int n, const int *skip)
{
    if (skip && *skip) return;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    if (s < nslice) {
        const int r = s * FG_SELL_W + (int)lane;
        const int rc = r < n ? r : n - 1;
        const T br = base[rc];
        const T dr = UPPER ? D[rc] : T(0);
        const int o0 = soff[s], w = soff[s + 1] - o0;
        const long long q0 = (long long)o0 * FG_SELL_W;
        // columns as 16-bit offsets from the row where they fit, else indices
        const T acc = off16 ? fg_sell_run(fg_sell_csr16<T>{off16 + q0, V + q0, x, rc, w}, lane, w)
                            : fg_sell_run(fg_sell_csr<T>{JA + q0, V + q0, x, w}, lane, w);
        const T y = br - acc;
        if (r < n) out[r] = UPPER ? dr * y : y;
    }
}

template <typename T>
__global__ __launch_bounds__(FG_TPB) void k_dscale(int n, const T *__restrict__ D,
                                                   const T *__restrict__ y, T *__restrict__ out,
                                                   const int *skip)
{
    if (skip && *skip) return;
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < n; i += gridDim.x * FG_TPB)
        out[i] = D[i] * y[i];
}

template <typename T>
static const T *tri_vals(const fg_ilu::Tri &t);
template <>
const double *tri_vals<double>(const fg_ilu::Tri &t) { return t.d_v64; }
template <>
const float *tri_vals<float>(const fg_ilu::Tri &t) { return t.d_v32; }

static int sweep_grid(const fg_ilu::Tri &t)
{
    constexpr int SPW = FG_TPB / FG_SELL_W;
    return (t.nslice + SPW - 1) / SPW;
}

template <typename T>
static int ilu_sweeps_t(fg_ilu *u, fg_ctx *c, const T *D, const T *p, T *z, T *const s[2],
                        const int *skip)
{
    const int n = u->neq, S = u->sweeps;
    const fg_ilu::Tri &L = u->lt, &U = u->ut;
    // forward: y_1..y_S alternate s[0], s[1] (y_0 = p)
    const T *y = p;
    for (int m = 0; m < S; m++) {
        T *out = s[m & 1];
        if (L.nslice)
            FG_LAUNCH((k_tri_sweep<T, false>), sweep_grid(L), FG_TPB, 0, c->stream, L.nslice,
                      (const int *)L.d_soff, (const int *)L.d_JA, (const short *)L.d_off16,
                      tri_vals<T>(L), y, p, D, out, n, skip);
        y = out;
    }
    // backward: x_0..x_S between the free scratch and z, x_S in z (p is no
    // longer read, so z may alias it)
    T *other = (y == s[0]) ? s[1] : s[0];
    T *buf[2] = {(S & 1) ? other : z, (S & 1) ? z : other};   // x_m in buf[m & 1]
    FG_LAUNCH(k_dscale<T>, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, D, y, buf[0], skip);
    for (int m = 1; m <= S; m++)
        if (U.nslice)
            FG_LAUNCH((k_tri_sweep<T, true>), sweep_grid(U), FG_TPB, 0, c->stream, U.nslice,
                      (const int *)U.d_soff, (const int *)U.d_JA, (const short *)U.d_off16,
                      tri_vals<T>(U), (const T *)buf[(m - 1) & 1], y, D, buf[m & 1], n, skip);
    return FG_OK;
}

int fg_ilu_set_sweeps(fg_ilu *u, int sweeps)
{
    // the sweeps address vectors by 32-bit byte offsets
    if (sweeps > 0 && u->neq >= (1 << 28)) {
        fg_set_error("ILU(0) sweeps need fewer than 2^28 equations");
        return FG_ERR_ARG;
    }
    u->sweeps = sweeps;
    if (sweeps > 0 && !u->d_s64[0]) {
        const size_t n = (size_t)u->neq + 1;
        for (int k = 0; k < 2; k++) {
            FG_HIP(hipMalloc(&u->d_s64[k], n * sizeof(double)));
            FG_HIP(hipMalloc(&u->d_s32[k], n * sizeof(float)));
        }
    }
    return FG_OK;
}

// z = (LU)^{-1} p on device vectors (neq+1); p may alias z
int fg_ilu_apply(fg_ilu *u, fg_ctx *c, int fp32, const void *p, void *z, const int *skip)
{
    if (u->sweeps > 0) {
        if (fp32)
            return ilu_sweeps_t<float>(u, c, u->d_D32, (const float *)p, (float *)z, u->d_s32,
                                       skip);
        return ilu_sweeps_t<double>(u, c, u->d_D, (const double *)p, (double *)z, u->d_s64,
                                    skip);
    }
    if (fp32)
        return ilu_apply_t<float>(u, c, u->d_LU32, u->d_D32, (const float *)p, (float *)z, skip);
    return ilu_apply_t<double>(u, c, u->d_LU, u->d_D, (const double *)p, (double *)z, skip);
}

int fg_ilu_get(const fg_ilu *u, fg_ctx *c, double *LU, double *D)
{
    FG_HIP(hipStreamSynchronize(c->stream));
    if (LU && u->nnz) FG_HIP(hipMemcpy(LU, u->d_LU, sizeof(double) * u->nnz, hipMemcpyDeviceToHost));
    if (D && u->neq) FG_HIP(hipMemcpy(D, u->d_D, sizeof(double) * u->neq, hipMemcpyDeviceToHost));
    return FG_OK;
}

int fg_ilu_sweep_count(const fg_ilu *u) { return u->sweeps; }
