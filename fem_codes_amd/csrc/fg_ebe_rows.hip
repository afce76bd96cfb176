// fg_ebe_rows.hip -- the element-by-element product as element rows grouped
// by equation (k = 3 triangles, k = 4 tetrahedra).
//
// Reference: MatrixVector_Operations/EBE/ebemv.c:24-40 (k = 3) and
// EBE/ebemv3Dndof1.c:24-42 (k = 4).  The reference walks the elements in
// order and, for element I with dofs i_1..i_k and row-major values A[I],
//     Q[i_r] += A[I][k r] * p_1 + A[I][k r + 1] * p_2 + ... + A[I][k r + k-1] * p_k
// (the row sum formed left to right, then added to Q), after Q = 0 and
// P[neq] = 0.  Equation r therefore receives one row sum per element that
// holds it, in element order.
//
// MI355X design: every (element, local row) pair whose dof is free is stored
// with its equation, in element order, in wavefront slices (fg_sell.h): 64
// equations per slice, one per lane.  An entry holds the row's k values and
// the other k-1 columns as offsets from the equation (15-bit signed fields,
// two in one 32-bit word with the local row index; for k = 4 the third in a
// 16-bit word), so the index stream costs 4 (k = 3) or 6 (k = 4) bytes per
// row, against the reference's 4k bytes of lm per element -- the same
// compulsory bytes for k = 3.  A lane loads its equation's P once, streams
// its rows (all loads of a chunk issued at once), gathers the other columns'
// P, forms each row sum left to right exactly as the reference does, and
// adds the sums in element order starting from 0: Q is the reference's ebemv
// bit for bit (FMA contraction disabled).  No LDS, no halo pass, no atomics.
//
// The values of a slice are stored contiguously: column c of a slice of
// width w starting at slot column o0 fills blocks k_dofs*o0 + c*w .. + w-1 of
// 64 values, entry k of the lane at k*64 + lane, or for FP64 triangles with
// the entries paired (entries 2j and 2j+1 side by side, one 16-byte load for
// both; fg_rows_shift): 8% faster for the FP64 triangle product, slower for
// FP32 and for FP64 tetrahedra (rows of 24 entries in chunks of 12), as
// measured.  The blocked layout (fg_ebe.hip) stays the primary copy of the
// values; the rows are refreshed from it by one kernel before the first
// product after any change.  A mesh whose column offsets do not fit the
// fields keeps the blocked product only.
#include <algorithm>
#include <vector>

#include "fg_ebe_body.h"
#include "fg_internal.h"
#include "fg_sell.h"

namespace {

// A padding slot has both 15-bit fields 0 (and, k = 4, the 16-bit field 0):
// an element's other columns never equal its own (checked when the rows are
// built), so no row has that word, and the padding's gathers read the lane's
// own P, in bounds.
constexpr unsigned ROW_PAD = 0u;
constexpr int OFF_SINK = -16384;            // 15-bit field: the Dirichlet sink
constexpr int OFF_MAX = 16383;
constexpr int OFF2_SINK = -32768;           // 16-bit field (k = 4, third column)

__device__ __forceinline__ int field_a(unsigned w) { return ((int)(w << 15)) >> 17; }
__device__ __forceinline__ int field_b(unsigned w) { return ((int)w) >> 17; }

// one element row of an equation's sliced list.  The index and value
// pointers start at the lane-0 slot of the wavefront's slice (wave-uniform
// bases): a load is a scalar base plus the lane's 32-bit offset, entry k at
// lane + 64 k.
template <typename TV, int K>
struct ElemRowSlot {
    using T = TV;
    struct Rec {
        unsigned w;
        int c2;
        T v[K];
        T po[K - 1];   // P of the other columns, in column order
    };
    const unsigned *__restrict__ idx;
    const short *__restrict__ idx2;   // k = 4: the third other column
    const T *__restrict__ V[K];       // value column c of the slice
    const T *__restrict__ P;
    int r;              // the lane's equation (clamped to the sink for idle lanes)
    int w;              // the slice width
    T pown;

    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec e;
        e.w = __builtin_nontemporal_load(&idx[q]);
        e.c2 = 0;
        if constexpr (K == 4) e.c2 = (int)__builtin_nontemporal_load(&idx2[q]);
#pragma unroll
        for (int c = 0; c < K - 1; c++) e.po[c] = T(0);
        return e;
    }
    __device__ __forceinline__ void stream(Rec &e, unsigned q) const
    {
        // FP64: the entry-paired position (consecutive entries of a lane at
        // consecutive addresses, merged into 16-byte loads)
        const unsigned p = sizeof(T) == 8 ? q + fg_rows_shift(K, (int)(q / FG_SELL_W),
                                                              (int)(q % FG_SELL_W), w)
                                          : q;
#pragma unroll
        for (int c = 0; c < K; c++) e.v[c] = __builtin_nontemporal_load(&V[c][p]);
    }
    __device__ __forceinline__ T col_p(int off, int sink) const
    {
        // a 32-bit byte offset from P (the rows are built for neq < 2^28)
        const unsigned b = (unsigned)(off == sink ? r : r + off) * (unsigned)sizeof(T);
        const T t = *(const T *)((const char *)P + b);
        return off == sink ? T(0) : t;
    }
    __device__ __forceinline__ void gather(Rec &e) const
    {
        e.po[0] = col_p(field_a(e.w), OFF_SINK);
        e.po[1] = col_p(field_b(e.w), OFF_SINK);
        if constexpr (K == 4) e.po[K - 2] = col_p(e.c2, OFF2_SINK);
    }
    __device__ __forceinline__ T term(const Rec &e) const
    {
        const int lr = (int)(e.w & 3u);
        T s = T(0);
#pragma unroll
        for (int c = 0; c < K; c++) {
            // column c: the equation itself at c == lr, else the next other column
            const T oth = c == 0 ? e.po[0] : (c == K - 1 ? e.po[K - 2] : (c < lr ? e.po[c] : e.po[c - 1]));
            const T pc = c == lr ? pown : oth;
            s = c == 0 ? e.v[0] * pc : s + e.v[c] * pc;
        }
        return e.w == ROW_PAD ? T(0) : s;
    }
    __device__ __forceinline__ T add(T acc, const Rec &e) const { return acc + term(e); }
};

// four slices (one per wavefront) per workgroup.  Optional per-slice mask
// (the batched solver): slice s belongs to problem sprob[s] (-1: rows of
// several problems), whose flag pflag[sprob[s] * pstride] != 0 skips it.
template <typename T, int K>
__global__ __launch_bounds__(FG_TPB) void k_ebe_rows(int nslice, const int *__restrict__ soff,
                                                     const unsigned *__restrict__ idx,
                                                     const short *__restrict__ idx2,
                                                     const T *__restrict__ V,
                                                     T *__restrict__ P, T *__restrict__ Q,
                                                     int neq, const int *skip,
                                                     const int *__restrict__ sprob,
                                                     const int *pflag, int pstride)
{
    if (skip && *skip) return;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    bool run = s < nslice;
    if (run && sprob) {
        const int pb = sprob[s];
        run = pb < 0 || pflag[(long long)pb * pstride] == 0;
    }
    if (run) {
        const int r = s * FG_SELL_W + (int)lane;
        // lanes past the last equation use the sink; all their slots are padding
        const int rc = r < neq ? r : neq;
        const int o0 = soff[s], w = soff[s + 1] - o0;
        const long long q0 = (long long)o0 * FG_SELL_W;
        ElemRowSlot<T, K> row;
        row.idx = idx + q0;
        row.idx2 = K == 4 ? idx2 + q0 : idx2;
#pragma unroll
        for (int c = 0; c < K; c++) row.V[c] = V + K * q0 + (long long)c * w * FG_SELL_W;
        row.P = P;
        row.r = rc;
        row.w = w;
        row.pown = P[rc];
        // tetrahedra: rows of about 24 entries in chunks of 12 (two load
        // bursts), measured faster than chunks of 8 in both precisions
        const T acc = fg_sell_run<K == 4 ? 12 : FG_SELL_CHUNK>(row, lane, w);
        if (r < neq) Q[r] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Q[neq] = T(0);
        P[neq] = T(0);
    }
}

// Element rows with per-equation neighbour tables (built for tetrahedra).  A
// tetrahedron row names 3 other columns, 72 scattered P loads per equation of
// 24 rows, but an equation has only about 14 distinct neighbours: each lane
// gathers its neighbours' P once into its own column of an LDS table, and
// each entry names its other columns by 4-bit table indices (2 bytes per row
// instead of 6).  The row sums and their order are those of the offset rows:
// the same result bit for bit.
template <int K>
struct NbTable {
    static constexpr int N = K == 3 ? 8 : 16;   // entries; the last holds 0 (the sink)
    static constexpr int MAXNB = N - 1;          // neighbours per equation
};
constexpr unsigned short NB_PAD = 0xffffu;

template <typename TV, int K>
struct NeighRow {
    using T = TV;
    struct Rec {
        unsigned cd;
        T v[K];
    };
    const unsigned short *__restrict__ code;   // the slice's lane-0 slot
    const T *__restrict__ V[K];
    const T *tab;   // the lane's table (LDS): neighbour j at tab[j * 64]
    T pown;
    int w;
    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec e;
        e.cd = __builtin_nontemporal_load(&code[q]);
        return e;
    }
    __device__ __forceinline__ void stream(Rec &e, unsigned q) const
    {
        // FP64 triangles: the entry-paired positions (fg_rows_shift)
        const unsigned p = sizeof(T) == 8 ? q + fg_rows_shift(K, (int)(q / FG_SELL_W),
                                                              (int)(q % FG_SELL_W), w)
                                          : q;
#pragma unroll
        for (int c = 0; c < K; c++) e.v[c] = __builtin_nontemporal_load(&V[c][p]);
    }
    // the table reads wait for little: they are made where the term needs
    // them, so that a chunk holds no more registers than its loads
    __device__ __forceinline__ void gather(Rec &) const {}
    __device__ __forceinline__ T term(const Rec &e) const
    {
        const int lr = (int)(e.cd & 3u);
        T po[K - 1];
#pragma unroll
        for (int j = 0; j < K - 1; j++) po[j] = tab[((e.cd >> (2 + 4 * j)) & 15u) * FG_SELL_W];
        T s = T(0);
#pragma unroll
        for (int c = 0; c < K; c++) {
            const T oth = c == 0 ? po[0] : (c == K - 1 ? po[K - 2] : (c < lr ? po[c] : po[c - 1]));
            const T pc = c == lr ? pown : oth;
            s = c == 0 ? e.v[0] * pc : s + e.v[c] * pc;
        }
        return e.cd == NB_PAD ? T(0) : s;
    }
    __device__ __forceinline__ T add(T acc, const Rec &e) const { return acc + term(e); }
};

template <typename T, int K>
__global__ __launch_bounds__(FG_TPB) void k_ebe_nrows(int nslice, const int *__restrict__ soff,
                                                      const int *__restrict__ nsoff,
                                                      const unsigned short *__restrict__ code,
                                                      const short *__restrict__ noff,
                                                      const T *__restrict__ V,
                                                      T *__restrict__ P, T *__restrict__ Q,
                                                      int neq, const int *skip,
                                                      const int *__restrict__ sprob,
                                                      const int *pflag, int pstride)
{
    if (skip && *skip) return;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    constexpr int NB = NbTable<K>::MAXNB;
    __shared__ T tabs[SPW][NbTable<K>::N * FG_SELL_W];
    const int wid = (int)(threadIdx.x / FG_SELL_W);
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    bool run = s < nslice;
    if (run && sprob) {
        const int pb = sprob[s];
        run = pb < 0 || pflag[(long long)pb * pstride] == 0;
    }
    if (run) {
        const int r = s * FG_SELL_W + (int)lane;
        const int rc = r < neq ? r : neq;
        const int o0 = soff[s], w = soff[s + 1] - o0;
        const int n0 = nsoff[s], m = nsoff[s + 1] - n0;
        T *tab = &tabs[wid][lane];
        // the lane's neighbours: all gathers issued together (entries past
        // the slice's count re-read its last one; padding reads the own P);
        // only this lane reads its column of the table, so no barrier
        {
            const short *no = noff + (long long)n0 * FG_SELL_W;
            short o[NB];
            T t[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const int jj = j < m ? j : m - 1;
                o[j] = m > 0 ? __builtin_nontemporal_load(&no[(unsigned)jj * FG_SELL_W + lane])
                             : (short)-32768;
            }
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const unsigned b = (unsigned)(o[j] == -32768 ? rc : rc + o[j]) * (unsigned)sizeof(T);
                t[j] = *(const T *)((const char *)P + b);
            }
#pragma unroll
            for (int j = 0; j < NB; j++)
                if (j < m) tab[j * FG_SELL_W] = o[j] == -32768 ? T(0) : t[j];
            tab[NB * FG_SELL_W] = T(0);
        }
        const long long q0 = (long long)o0 * FG_SELL_W;
        NeighRow<T, K> row;
        row.code = code + q0;
#pragma unroll
        for (int c = 0; c < K; c++) row.V[c] = V + K * q0 + (long long)c * w * FG_SELL_W;
        row.tab = tab;
        row.pown = P[rc];
        row.w = w;
        // tetrahedron rows (about 24 entries) in whole chunks; triangle rows
        // (about 6) unrolled whole
        const T acc = K == 4 ? fg_sell_chunks<8>(row, lane, w) : fg_sell_run<8>(row, lane, w);
        if (r < neq) Q[r] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Q[neq] = T(0);
        P[neq] = T(0);
    }
}

// the rows' values from the blocked values; one thread per (slice, lane)
template <int K>
__global__ void k_ebe_rows_mirror(int nslice, const int *__restrict__ soff,
                                  const int *__restrict__ src, const long long *__restrict__ kpos,
                                  const int *__restrict__ kstride, const double *__restrict__ K64,
                                  const float *__restrict__ K32, double *__restrict__ V64,
                                  float *__restrict__ V32)
{
    const long long nt = (long long)nslice * FG_SELL_W;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < nt;
         t += (long long)gridDim.x * FG_TPB) {
        const int s = (int)(t / FG_SELL_W), lane = (int)(t % FG_SELL_W);
        const int o0 = soff[s], w = soff[s + 1] - o0;
        for (int k = 0; k < w; k++) {
            const long long q = (long long)(o0 + k) * FG_SELL_W + lane;
            const int sv = src[q];
            if (sv < 0) continue;
            const int e = sv / K, lr = sv - e * K;
            const long long p = kpos[e];
            const int st = kstride[e];
            const int sh = fg_rows_shift(K, k, lane, w);
#pragma unroll
            for (int c = 0; c < K; c++) {
                const long long vi = fg_ebe_vidx(p, st, lr * K + c, fg_ebe_vec(K));
                const long long d = ((long long)K * o0 + (long long)c * w + k) * FG_SELL_W + lane;
                V64[d + sh] = K64[vi];
                V32[d] = K32[vi];
            }
        }
    }
}

// AoS element values -> the blocked layout and the rows in one pass; one
// thread per (element, local row): the row's k values are read contiguously
// (consecutive threads, consecutive rows), written to their blocked positions
// and, when the row's dof is free, to its slot (FP32 value c at vpos + c *
// stride, FP64 shifted to the entry-paired position)
template <int K>
__global__ void k_ebe_scatter_rows(long long n, const double *__restrict__ A,
                                   const long long *__restrict__ kpos,
                                   const int *__restrict__ kstride,
                                   const long long *__restrict__ vpos,
                                   const int *__restrict__ wst, double *__restrict__ K64,
                                   float *__restrict__ K32, double *__restrict__ V64,
                                   float *__restrict__ V32)
{
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < n;
         t += (long long)gridDim.x * FG_TPB) {
        const long long e = t / K;
        const int lr = (int)(t - e * K);
        const long long p = kpos[e];
        if (p < 0) continue;
        const int st = kstride[e];
        const long long vp = vpos[t];
        const long long ws = fg_rows_wst_stride(wst[t]);
        const int sh = fg_rows_wst_shift(wst[t]);
        double v[K];
#pragma unroll
        for (int c = 0; c < K; c++) v[c] = A[t * K + c];
#pragma unroll
        for (int c = 0; c < K; c++) {
            const long long d = fg_ebe_vidx(p, st, lr * K + c, fg_ebe_vec(K));
            K64[d] = v[c];
            K32[d] = (float)v[c];
        }
        if (vp >= 0) {
#pragma unroll
            for (int c = 0; c < K; c++) {
                V64[vp + c * ws + sh] = v[c];
                V32[vp + c * ws] = (float)v[c];
            }
        }
    }
}

// the diagonal entry of each element row of an equation: FP64 value column
// lr of its slot (padding adds nothing); pointers from the slice's lane-0
// slot, q = k*64 + lane
template <int K>
struct DiagRow {
    using T = double;
    struct Rec {
        unsigned w;
        unsigned q;
        double v;
    };
    const unsigned *__restrict__ idx;
    const double *__restrict__ V;   // the slice's column 0
    int wstride, w;
    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec r;
        r.w = idx[q];
        r.q = q;
        r.v = 0.0;
        return r;
    }
    __device__ __forceinline__ void stream(Rec &, unsigned) const {}
    __device__ __forceinline__ void gather(Rec &r) const
    {
        const int sh = fg_rows_shift(K, (int)(r.q / FG_SELL_W), (int)(r.q % FG_SELL_W), w);
        r.v = V[(long long)r.q + sh + (long long)(r.w & 3u) * wstride];
    }
    __device__ __forceinline__ double add(double acc, const Rec &r) const
    {
        return r.w == ROW_PAD ? acc : acc + r.v;
    }
};

// D[g] = sum over g's element rows, in element order, of their diagonal
// entries (Diag_precond_EBE_setup.c:18-24), from the rows
template <int K>
__global__ __launch_bounds__(FG_TPB) void k_ebe_rows_diag(int nslice, const int *__restrict__ soff,
                                                          const unsigned *__restrict__ idx,
                                                          const double *__restrict__ V,
                                                          double *__restrict__ D, int neq)
{
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_sell_slice(fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW);
    const unsigned lane = threadIdx.x % FG_SELL_W;
    if (s >= nslice) return;
    const int o0 = soff[s], w = soff[s + 1] - o0;
    const long long q0 = (long long)o0 * FG_SELL_W;
    const DiagRow<K> row{idx + q0, V + K * q0, w * FG_SELL_W, w};
    const double d = fg_sell_run(row, lane, w);
    const int r = s * FG_SELL_W + (int)lane;
    if (r < neq) D[r] = d;
}

template <typename T>
int upload(const std::vector<T> &v, T **d)
{
    const size_t n = v.size() ? v.size() : 1;
    FG_HIP(hipMalloc(d, n * sizeof(T)));
    if (!v.empty()) FG_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

}  // namespace

// rows of an element operator with k = 3 or 4; leaves the operator without
// rows (blocked product only) when a column offset does not fit its field
int fg_ebe_rows_build(fg_op *op, const int *lm, const long long *kpos)
{
    const int neq = op->neq, nel = op->nitems, K = op->k;
    if (op->kind == FG_OP_EBE && (K == 8 || K == 12)) return fg_ebe_node_build(op, lm, kpos);
    if (op->kind != FG_OP_EBE || (K != 3 && K != 4) || neq == 0) return FG_OK;
    // the product addresses P by a 32-bit byte offset
    if (neq >= (1 << 28)) return FG_OK;
    std::vector<int> cnt((size_t)neq + 1, 0);
    for (int e = 0; e < nel; e++) {
        if (kpos[e] < 0) continue;
        for (int a = 0; a < K; a++) {
            const int r = lm[(size_t)e * K + a];
            if (r != neq) cnt[r + 1]++;
        }
    }
    for (int r = 0; r < neq; r++) cnt[r + 1] += cnt[r];
    const fg_sell_layout L = fg_sell_plan(neq, cnt.data());
    // fg_rows_wst holds the column stride w*64 in 24 bits
    for (int s = 0; s < L.nslice; s++)
        if (L.soff[s + 1] - L.soff[s] >= (1 << 18)) return FG_OK;
    std::vector<unsigned> idx(L.nslot, ROW_PAD);
    std::vector<short> idx2(K == 4 ? L.nslot : 0, 0);
    std::vector<int> src(L.nslot, -1);
    std::vector<long long> vpos((size_t)nel * K, -1);
    std::vector<int> wst((size_t)nel * K, 0);
    std::vector<int> fill((size_t)neq, 0);
    for (int e = 0; e < nel; e++) {
        if (kpos[e] < 0) continue;
        const int *le = lm + (size_t)e * K;
        for (int a = 0; a < K; a++) {
            const int r = le[a];
            if (r == neq) continue;
            int off[3] = {0, 0, 0}, m = 0;
            for (int c = 0; c < K; c++) {
                if (c == a) continue;
                const long long d = (long long)le[c] - r;
                const int lim = m == 2 ? 32767 : OFF_MAX;
                // an offset beyond its field, or a repeated dof: no rows
                if (le[c] != neq && (d == 0 || d < -lim || d > lim)) return FG_OK;
                off[m] = le[c] == neq ? (m == 2 ? OFF2_SINK : OFF_SINK) : (int)d;
                m++;
            }
            const int kr = fill[r]++;
            const long long q = L.slot(r, kr);
            idx[q] = (unsigned)a | ((unsigned)(off[0] & 0x7fff) << 2) |
                     ((unsigned)(off[1] & 0x7fff) << 17);
            if (K == 4) idx2[q] = (short)off[2];
            src[q] = e * K + a;
            const int s0 = r / FG_SELL_W, w = L.soff[s0 + 1] - L.soff[s0];
            vpos[(size_t)e * K + a] = q + (long long)(K - 1) * L.soff[s0] * FG_SELL_W;
            wst[(size_t)e * K + a] = fg_rows_wst(w, fg_rows_shift(K, kr, r % FG_SELL_W, w));
        }
    }
    // the neighbour form when every equation has at most nbmax distinct other
    // columns within 16-bit offsets (neighbours in order of first
    // appearance), else the offset rows
    // (tetrahedra only: for triangle rows the table costs more than the 6
    // gathers it saves per equation -- 2048^2, measured: 124.7 -> 128.3 us
    // FP64, 65.3 -> 67.0 us FP32)
    const int nbmax = NbTable<4>::MAXNB;
    bool neigh = K == 4;
    std::vector<unsigned short> code;
    std::vector<int> nsoff;
    std::vector<short> noff;
    if (neigh) {
        std::vector<int> nb((size_t)neq * nbmax), nbn((size_t)neq, 0);
        code.assign(L.nslot, NB_PAD);
        std::fill(fill.begin(), fill.end(), 0);
        for (int e = 0; e < nel && neigh; e++) {
            if (kpos[e] < 0) continue;
            const int *le = lm + (size_t)e * K;
            for (int a = 0; a < K && neigh; a++) {
                const int r = le[a];
                if (r == neq) continue;
                unsigned cd = (unsigned)a;
                int m = 0;
                for (int c = 0; c < K; c++) {
                    if (c == a) continue;
                    int j = nbmax;   // the sink
                    if (le[c] != neq) {
                        int *L0 = &nb[(size_t)r * nbmax];
                        for (j = 0; j < nbn[r] && L0[j] != le[c]; j++) {}
                        if (j == nbn[r]) {
                            const long long d = (long long)le[c] - r;
                            if (j == nbmax || d < -32767 || d > 32767) {
                                neigh = false;
                                break;
                            }
                            L0[nbn[r]++] = le[c];
                        }
                    }
                    cd |= (unsigned)j << (2 + 4 * m++);
                }
                code[L.slot(r, fill[r]++)] = (unsigned short)cd;
            }
        }
        if (neigh) {
            const int ns = L.nslice;
            nsoff.assign((size_t)ns + 1, 0);
            for (int s = 0; s < ns; s++) {
                int mx = 0;
                for (int r = s * FG_SELL_W; r < std::min(neq, (s + 1) * FG_SELL_W); r++)
                    mx = std::max(mx, nbn[r]);
                nsoff[s + 1] = nsoff[s] + mx;
            }
            noff.assign((size_t)nsoff[ns] * FG_SELL_W, (short)-32768);
            for (int r = 0; r < neq; r++)
                for (int j = 0; j < nbn[r]; j++)
                    noff[((size_t)nsoff[r / FG_SELL_W] + j) * FG_SELL_W + r % FG_SELL_W] =
                        (short)(nb[(size_t)r * nbmax + j] - r);
        }
    }
    op->nslice = L.nslice;
    op->nslot = L.nslot;
    FG_CHECK(upload(L.soff, &op->d_soff));
    FG_CHECK(upload(idx, &op->d_ridx));
    if (neigh) {
        FG_CHECK(upload(code, &op->d_rcode));
        FG_CHECK(upload(nsoff, &op->d_nsoff));
        FG_CHECK(upload(noff, &op->d_noff));
    } else if (K == 4) {
        FG_CHECK(upload(idx2, &op->d_ridx2));
    }
    FG_CHECK(upload(src, &op->d_rsrc));
    FG_CHECK(upload(vpos, &op->d_rvpos));
    FG_CHECK(upload(wst, &op->d_rwst));
    const size_t nv = (size_t)K * (L.nslot ? L.nslot : 1);
    FG_HIP(hipMalloc(&op->d_sv64, sizeof(double) * nv));
    FG_HIP(hipMalloc(&op->d_sv32, sizeof(float) * nv));
    FG_HIP(hipMemset(op->d_sv64, 0, sizeof(double) * nv));
    FG_HIP(hipMemset(op->d_sv32, 0, sizeof(float) * nv));
    op->rows_stale = 1;
    return FG_OK;
}

int fg_ebe_rows_set_values(fg_op *op, const double *A)
{
    if (op->k == 8 || op->k == 12) return fg_ebe_node_set_values(op, A);
    fg_ctx *c = op->ctx;
    const long long n = (long long)op->nitems * op->k;
    if (n == 0) return FG_OK;
    if (op->k == 3)
        FG_LAUNCH(k_ebe_scatter_rows<3>, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, A,
                  op->d_kpos, op->d_kstride, op->d_rvpos, op->d_rwst, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    else
        FG_LAUNCH(k_ebe_scatter_rows<4>, fg_vec_grid(c, n), FG_TPB, 0, c->stream, n, A,
                  op->d_kpos, op->d_kstride, op->d_rvpos, op->d_rwst, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    op->rows_stale = 0;
    return FG_OK;
}

int fg_ebe_rows_refresh(fg_op *op)
{
    if (op->k == 8 || op->k == 12) return fg_ebe_node_refresh(op);
    if (!op->d_ridx || !op->rows_stale) return FG_OK;
    fg_ctx *c = op->ctx;
    const long long nt = (long long)op->nslice * FG_SELL_W;
    if (op->k == 3)
        FG_LAUNCH(k_ebe_rows_mirror<3>, fg_vec_grid(c, nt), FG_TPB, 0, c->stream, op->nslice,
                  op->d_soff, op->d_rsrc, op->d_kpos, op->d_kstride, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    else
        FG_LAUNCH(k_ebe_rows_mirror<4>, fg_vec_grid(c, nt), FG_TPB, 0, c->stream, op->nslice,
                  op->d_soff, op->d_rsrc, op->d_kpos, op->d_kstride, op->d_K64, op->d_K32,
                  op->d_sv64, op->d_sv32);
    op->rows_stale = 0;
    return FG_OK;
}

int fg_ebe_rows_diag(fg_op *op, double *D)
{
    FG_CHECK(fg_ebe_rows_refresh(op));
    fg_ctx *c = op->ctx;
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int g = (op->nslice + SPW - 1) / SPW;
    if (op->k == 3)
        FG_LAUNCH(k_ebe_rows_diag<3>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                  op->d_ridx, (const double *)op->d_sv64, D, op->neq);
    else
        FG_LAUNCH(k_ebe_rows_diag<4>, g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                  op->d_ridx, (const double *)op->d_sv64, D, op->neq);
    return FG_OK;
}

int fg_ebe_rows_apply(fg_op *op, int fp32, void *P, void *Q, const int *skip)
{
    return fg_ebe_rows_apply_masked(op, fp32, P, Q, skip, nullptr, nullptr, 0);
}

int fg_ebe_rows_apply_masked(fg_op *op, int fp32, void *P, void *Q, const int *skip,
                             const int *sprob, const int *pflag, int pstride)
{
    // node rows (k = 8, 12): every slice (the mask is an optimisation only)
    if (op->k == 8 || op->k == 12) return fg_ebe_node_apply(op, fp32, P, Q, skip);
    fg_ctx *c = op->ctx;
    const int K = op->k;
    FG_CHECK(fg_ebe_rows_refresh(op));
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int g = (op->nslice + SPW - 1) / SPW;
    if (op->d_rcode) {
#define FG_NROWS(T, KK, VV)                                                                    \
    FG_LAUNCH((k_ebe_nrows<T, KK>), g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,          \
              op->d_nsoff, op->d_rcode, op->d_noff, (const T *)VV, (T *)P, (T *)Q, op->neq,   \
              skip, sprob, pflag, pstride)
        if (fp32) FG_NROWS(float, 4, op->d_sv32);
        else FG_NROWS(double, 4, op->d_sv64);
#undef FG_NROWS
        return FG_OK;
    }
    if (fp32) {
        if (K == 3)
            FG_LAUNCH((k_ebe_rows<float, 3>), g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_ridx, op->d_ridx2, (const float *)op->d_sv32, (float *)P,
                      (float *)Q, op->neq, skip, sprob, pflag, pstride);
        else
            FG_LAUNCH((k_ebe_rows<float, 4>), g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_ridx, op->d_ridx2, (const float *)op->d_sv32, (float *)P,
                      (float *)Q, op->neq, skip, sprob, pflag, pstride);
    } else {
        if (K == 3)
            FG_LAUNCH((k_ebe_rows<double, 3>), g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_ridx, op->d_ridx2, (const double *)op->d_sv64, (double *)P,
                      (double *)Q, op->neq, skip, sprob, pflag, pstride);
        else
            FG_LAUNCH((k_ebe_rows<double, 4>), g, FG_TPB, 0, c->stream, op->nslice, op->d_soff,
                      op->d_ridx, op->d_ridx2, (const double *)op->d_sv64, (double *)P,
                      (double *)Q, op->neq, skip, sprob, pflag, pstride);
    }
    return FG_OK;
}

void fg_ebe_rows_free(fg_op *op)
{
    hipFree(op->d_ridx); hipFree(op->d_ridx2); hipFree(op->d_rsrc); hipFree(op->d_rvpos);
    hipFree(op->d_rcode); hipFree(op->d_nsoff); hipFree(op->d_noff); hipFree(op->d_nbase);
    hipFree(op->d_rwst);
    if (op->kind == FG_OP_EBE) {
        hipFree(op->d_soff); hipFree(op->d_sv64); hipFree(op->d_sv32);
    }
}
