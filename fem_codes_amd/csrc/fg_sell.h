// fg_sell.h -- rows in wavefront slices ("sliced ELL"), shared by the CSR
// product and the EDE half-edge rows.
//
// Rows are cut into slices of 64 consecutive rows, one row per lane of a
// wavefront.  Slice s owns slot columns soff[s] .. soff[s+1]-1 (its widest
// row's entry count); entry k of the row at lane l sits at slot
// (soff[s] + k) * 64 + l, so the k-th entries of the slice's 64 rows are 64
// consecutive words: every load of a product is one fully coalesced
// wavefront access, issued without a branch, and each lane adds its row's
// terms in entry order in registers -- no LDS staging, no barrier, the same
// per-row summation order as a serial row loop.  Unused slots hold column -1
// (their loads are issued like the others; the term is masked to nothing).
#pragma once

#include <vector>

#include "fg_internal.h"

constexpr int FG_SELL_W = 64;       // rows per slice (one wavefront)
constexpr int FG_SELL_CHUNK = 8;    // entries per unrolled chunk of a wide row (default)

// host: slice layout of rows with entry counts rowptr[r+1]-rowptr[r]
struct fg_sell_layout {
    int nslice = 0;
    long long nslot = 0;
    std::vector<int> soff;          // [nslice+1] in units of 64 slots
    // slot of entry k of row r
    long long slot(int r, int k) const
    {
        return ((long long)soff[r / FG_SELL_W] + k) * FG_SELL_W + (r % FG_SELL_W);
    }
};

inline fg_sell_layout fg_sell_plan(int nrows, const int *rowptr)
{
    fg_sell_layout L;
    L.nslice = (nrows + FG_SELL_W - 1) / FG_SELL_W;
    L.soff.assign((size_t)L.nslice + 1, 0);
    for (int s = 0; s < L.nslice; s++) {
        int w = 0;
        const int r1 = (s + 1) * FG_SELL_W < nrows ? (s + 1) * FG_SELL_W : nrows;
        for (int r = s * FG_SELL_W; r < r1; r++) {
            const int n = rowptr[r + 1] - rowptr[r];
            w = n > w ? n : w;
        }
        L.soff[s + 1] = L.soff[s] + w;
    }
    L.nslot = (long long)L.soff[L.nslice] * FG_SELL_W;
    return L;
}

// Issues the loads of L consecutive entries of a row (slot `at` and every
// 64th after it), gathers, and folds the entries into acc in entry order.
// With CLAMP, entries at or beyond `nleft` re-read entry nleft-1 and add
// nothing (all loads still issued on every path).  Row::key(slot) returns an
// entry record holding the entry's index loads (issued first, for every
// entry), Row::stream(rec, slot) issues the loads that need no index (the
// values), Row::gather(rec) the loads that do, Row::add(acc, rec) folds the
// entry in (acc + its term for a product row; acc itself for a padding
// slot).
template <int L, bool CLAMP, typename Row, typename I>
__device__ __forceinline__ typename Row::T fg_sell_chunk(const Row &row, I at, int nleft,
                                                         typename Row::T acc)
{
    typename Row::Rec e[L];
    I q[L];
#pragma unroll
    for (int k = 0; k < L; k++) {
        const int kk = CLAMP ? (k < nleft ? k : nleft - 1) : k;
        q[k] = at + (I)kk * FG_SELL_W;
        e[k] = row.key(q[k]);
    }
#pragma unroll
    for (int k = 0; k < L; k++) row.stream(e[k], q[k]);
    // every load of the chunk that needs no index is issued before the first
    // wait for one (the scheduler would otherwise hold value loads back)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < L; k++) row.gather(e[k]);
#pragma unroll
    for (int k = 0; k < L; k++) {
        const typename Row::T a = row.add(acc, e[k]);
        acc = (!CLAMP || k < nleft) ? a : acc;
    }
    return acc;
}

// The row sum in whole chunks of C entries, then the rest (clamped): for rows
// that are almost never short (their registers sized for C entries only)
template <int C, typename Row, typename I>
__device__ __forceinline__ typename Row::T fg_sell_chunks(const Row &row, I at, int w)
{
    typename Row::T acc = typename Row::T(0);
    int k0 = 0;
    for (; k0 + C <= w; k0 += C)
        acc = fg_sell_chunk<C, false>(row, at + (I)k0 * FG_SELL_W, C, acc);
    if (k0 < w) acc = fg_sell_chunk<C, true>(row, at + (I)k0 * FG_SELL_W, w - k0, acc);
    return acc;
}

// The row sum of the lane whose first entry is slot `at` of a slice of
// width w: rows of up to 8 entries fully unrolled (one load burst), wider
// rows in chunks of C (the last one clamped).  `at` is an absolute slot (long long) or, for a Row
// whose pointers start at its slice (a wave-uniform base, so that every load
// is a scalar base plus the lane's 32-bit offset), the lane (unsigned).
template <int C = FG_SELL_CHUNK, typename Row, typename I>
__device__ __forceinline__ typename Row::T fg_sell_run(const Row &row, I at, int w)
{
    using T = typename Row::T;
    switch (w) {
    case 0: return T(0);
    case 1: return fg_sell_chunk<1, false>(row, at, 1, T(0));
    case 2: return fg_sell_chunk<2, false>(row, at, 2, T(0));
    case 3: return fg_sell_chunk<3, false>(row, at, 3, T(0));
    case 4: return fg_sell_chunk<4, false>(row, at, 4, T(0));
    case 5: return fg_sell_chunk<5, false>(row, at, 5, T(0));
    case 6: return fg_sell_chunk<6, false>(row, at, 6, T(0));
    case 7: return fg_sell_chunk<7, false>(row, at, 7, T(0));
    case 8: return fg_sell_chunk<8, false>(row, at, 8, T(0));
    default: return fg_sell_chunks<C>(row, at, w);
    }
}

// The row sum of lane `lane` of a slice of width w starting at slot column o0
template <typename Row>
__device__ __forceinline__ typename Row::T fg_sell_row(const Row &row, int o0, int w, int lane)
{
    return fg_sell_run(row, (long long)o0 * FG_SELL_W + lane, w);
}

// The wavefront's slice: s0 (wave-uniform) plus the wave's index in the
// workgroup, read back as a scalar so that the slice's offsets and width,
// and the branch on the width, are scalar.
__device__ __forceinline__ int fg_sell_slice(int s0)
{
    return __builtin_amdgcn_readfirstlane(s0 + (int)(threadIdx.x / FG_SELL_W));
}

// a sparse product row of a slice (CSR rows, the ILU(0) triangular parts):
// the column and value pointers start at the lane-0 slot of the wavefront's
// slice (wave-uniform bases, q = k*64 + lane; column -1: padding); FP64
// values entry-paired (fg_pair_shift), merged into 16-byte loads.  Columns
// and values are read once per product: non-temporal, so that they do not
// evict the x entries neighbouring slices gather again.
template <typename TV>
struct fg_sell_csr {
    using T = TV;
    struct Rec {
        int c;
        T v, p;
    };
    const int *__restrict__ col;
    const T *__restrict__ val;
    const T *__restrict__ x;
    int w;
    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec r;
        r.c = __builtin_nontemporal_load(&col[q]);
        r.p = T(0);
        return r;
    }
    __device__ __forceinline__ void stream(Rec &r, unsigned q) const
    {
        const unsigned p = sizeof(T) == 8 ? q + fg_pair_shift((int)(q / FG_SELL_W),
                                                              (int)(q % FG_SELL_W), w)
                                          : q;
        r.v = __builtin_nontemporal_load(&val[p]);
    }
    __device__ __forceinline__ void gather(Rec &r) const
    {
        // a 32-bit byte offset from x (sliced copies are built for n < 2^28)
        const unsigned b = (unsigned)(r.c < 0 ? 0 : r.c) * (unsigned)sizeof(T);
        r.p = *(const T *)((const char *)x + b);
    }
    __device__ __forceinline__ T add(T acc, const Rec &r) const
    {
        return acc + (r.c < 0 ? T(0) : r.v * r.p);
    }
};

// the same row with each column stored as a 16-bit offset from the lane's
// row (FG_SELL_OFF_PAD: padding): 2 bytes of index per entry instead of 4,
// for patterns whose columns all lie within 32767 of their row
constexpr short FG_SELL_OFF_PAD = -32768;
template <typename TV>
struct fg_sell_csr16 {
    using T = TV;
    struct Rec {
        int o;
        T v, p;
    };
    const short *__restrict__ off;
    const T *__restrict__ val;
    const T *__restrict__ x;
    int r;   // the lane's row (any row of the slice for idle lanes)
    int w;
    __device__ __forceinline__ Rec key(unsigned q) const
    {
        Rec e;
        e.o = __builtin_nontemporal_load(&off[q]);
        e.p = T(0);
        return e;
    }
    __device__ __forceinline__ void stream(Rec &e, unsigned q) const
    {
        const unsigned p = sizeof(T) == 8 ? q + fg_pair_shift((int)(q / FG_SELL_W),
                                                              (int)(q % FG_SELL_W), w)
                                          : q;
        e.v = __builtin_nontemporal_load(&val[p]);
    }
    __device__ __forceinline__ void gather(Rec &e) const
    {
        const unsigned b = (unsigned)(e.o == FG_SELL_OFF_PAD ? r : r + e.o) * (unsigned)sizeof(T);
        e.p = *(const T *)((const char *)x + b);
    }
    __device__ __forceinline__ T add(T acc, const Rec &e) const
    {
        return acc + (e.o == FG_SELL_OFF_PAD ? T(0) : e.v * e.p);
    }
};

// the slot of an entry with its FP64 pair shift in the top byte (the
// writers of paired values); FG_SELL_NOSRC in the low bits: no source entry
constexpr int FG_SELL_SHIFT_BIT = 56;
constexpr long long FG_SELL_NOSRC = (1LL << FG_SELL_SHIFT_BIT) - 1;
__host__ __device__ inline long long fg_sell_pack(long long q, int shift)
{
    return q | ((long long)(shift + 64) << FG_SELL_SHIFT_BIT);
}
__host__ __device__ inline long long fg_sell_packed_slot(long long pk)
{
    return pk & FG_SELL_NOSRC;
}
__host__ __device__ inline int fg_sell_packed_shift(long long pk)
{
    return (int)(pk >> FG_SELL_SHIFT_BIT) - 64;
}
