// fg_ebe_body.h -- one workgroup's share of an element-by-element product:
// the block body used by the stand-alone EBE/EDE kernels (fg_ebe.hip) and by
// the matvec phase of the persistent GMRES kernel (fg_gmres_fused.hip).
//
// Block b holds ne elements (breadth-first order) and nd distinct dofs.
// Values are stored per block as [k*k][ne] and local dof indices as [k][ne],
// so every global load is coalesced.  For each local dof the block also
// stores its incidences -- the (row r, element i) pairs that add into it, in
// element order -- as LDS slots r*ne + i.  The body has three phases so that
// a persistent kernel can overlap the loads of the next block with the work
// on the current one:
//   load   -- this thread's local indices (and, when they fit in registers,
//             its element values); independent of P;
//   gather -- the block's dofs of P, its incidence lists into LDS;
//   finish -- row sums of each element (reference order: left to right over
//             the columns, ebemv.c:34-36) into LDS; then every dof adds its
//             incidences in their stored order (deterministic, no atomics),
//             dofs complete in this block -> Q, shared dofs -> their halo
//             slot.
#pragma once

#include "fg_internal.h"

// default elements per thread (block = 256 * EPT elements), the measured
// optimum on MI355X: 1024 edges (k = 2), 512 triangles (k = 3), 1024
// tetrahedra (k = 4: 0.195 vs 0.213 ms for the FP32 product of 12.6M tets,
// fewer halo dofs per block), 256 elements of 12 or 16 dofs
constexpr int fg_ebe_default_ept(int k) { return k <= 2 ? 4 : k == 3 ? 2 : k == 4 ? 4 : 1; }

// entries per group of the blocked value layout.  When k*k is a multiple of
// 4 (k = 2, 4, 12, 16) a block stores its values as [k*k/4][ne][4]: the four
// consecutive entries q = 4g..4g+3 of an element sit together, so one 16-byte
// (FP32) or 32-byte (FP64) load fetches them, and a wavefront still reads
// one contiguous run per group.  Otherwise (k = 3) the layout is [k*k][ne].
__host__ __device__ constexpr int fg_ebe_vec(int k) { return (k * k) % 4 == 0 ? 4 : 1; }

// position of entry q of an element in the blocked values; kpos / kstride
// are the element's base (first entry) and group stride
__host__ __device__ inline long long fg_ebe_vidx(long long kpos, int kstride, int q, int V)
{
    return kpos + (long long)(q / V) * kstride + (q % V);
}

// device view of the blocked layout
struct FgEbeLayout {
    int nblk;
    const int *eoff;       // [nblk+1] first element of each block
    const int *doff;       // [nblk+1] first local dof of each block
    const int *dofs;       // global dof of each local dof
    const int *slots;      // -1: dof complete in its block, else halo slot
    const int *ibase;      // [nblk+1] first incidence of each block
    const uint16_t *lml;   // local dof indices, [k][ne] per block (nd = sink)
    const uint16_t *ioff;  // per block nd+1 incidence offsets, at doff[b] + b
    const uint16_t *inc;   // incidence LDS slots r*ne + i
};

// one block's extents (element, dof and incidence ranges)
struct FgEbeBlk {
    int e0, ne, d0, nd, i0, ninc;
};
__device__ __forceinline__ FgEbeBlk fg_ebe_blk(const FgEbeLayout &L, int b)
{
    FgEbeBlk B;
    B.e0 = L.eoff[b];
    B.ne = L.eoff[b + 1] - B.e0;
    B.d0 = L.doff[b];
    B.nd = L.doff[b + 1] - B.d0;
    B.i0 = L.ibase[b];
    B.ninc = L.ibase[b + 1] - B.i0;
    return B;
}

inline FgEbeLayout fg_ebe_layout(const fg_op *op)
{
    return FgEbeLayout{op->nblk, op->d_eoff, op->d_doff, op->d_dofs, op->d_slots,
                       op->d_ibase, op->d_lml, op->d_ioff, op->d_inc};
}

// LDS bytes one block needs (ninc: the block's incidence count padded to a
// multiple of 8): incidence lists and offsets first (16-byte aligned), then
// P values and element row sums
__host__ __device__ constexpr size_t fg_ebe_smem_idx_bytes(int nd, int ninc)
{
    return (2 * ((size_t)ninc + nd + 1) + 15) / 16 * 16;
}
// k >= 8 with one element per thread: the row sums take the LDS of the P
// values (each thread holds its element's P values in registers across a
// barrier), so a block needs max(nd + 1, k ne) values instead of the sum
__host__ __device__ constexpr bool fg_ebe_alias(int k, int ept) { return k >= 8 && ept == 1; }
template <typename T>
__host__ __device__ constexpr size_t fg_ebe_smem_bytes(int k, int ept, int nd, int ne, int ninc)
{
    const size_t np = (size_t)nd + 1, ny = (size_t)k * ne;
    return fg_ebe_smem_idx_bytes(nd, ninc) +
           sizeof(T) * (fg_ebe_alias(k, ept) ? (np > ny ? np : ny) : np + ny);
}

// x[i] through a block-uniform base (scalar registers) and a 32-bit byte
// offset, so each load or store needs one VGPR of address rather than a
// 64-bit pair.  Requires i * sizeof(T) < 2^32: block-local arrays, and the
// global vectors of neq + 1 <= 2^28 entries and the halo partials (fewer
// than 2^28 slots), both checked by ebe_create before any device work.
template <typename T>
__device__ __forceinline__ const T &fg_at(const T *x, unsigned i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(x) +
                                        i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T &fg_at(T *x, unsigned i)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(x) + i * (unsigned)sizeof(T));
}

// Element values and local indices are read once per product (hundreds of
// MB, far beyond L2 and the MALL): they are loaded non-temporally, so they
// do not evict the P entries that neighbouring blocks gather again.
// Measured on the 2048x2048 FP64 product: 157.8 -> 126.9 us per launch.
template <typename T>
__device__ __forceinline__ T fg_ld_nt(const T *x, unsigned i)
{
    return __builtin_nontemporal_load(
        reinterpret_cast<const T *>(reinterpret_cast<const char *>(x) + i * (unsigned)sizeof(T)));
}

// group g (4 consecutive values) of a grouped layout: one vector load,
// non-temporal for the 16-byte FP32 groups (the 32-byte FP64 groups of the
// 12x12 product measured 8% slower with the hint: 4.53 vs 4.94 TB/s)
template <typename T> struct FgVec4;
template <> struct FgVec4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct FgVec4<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <typename T>
__device__ __forceinline__ void fg_ld4(const T *x, unsigned g, T out[4])
{
    using V = typename FgVec4<T>::type;
    const V *a = reinterpret_cast<const V *>(reinterpret_cast<const char *>(x) +
                                             g * (unsigned)(4 * sizeof(T)));
    V v;
    if constexpr (sizeof(T) == 4) v = __builtin_nontemporal_load(a);
    else v = *a;
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    out[3] = v.w;
}

template <int K, int EPT, typename T>
struct FgEbeRegs {
    // element values held in registers: at most 36 FP64 / 72 FP32 values
    static constexpr bool PRE = EPT * K * K <= (sizeof(T) == 8 ? 36 : 72);
    // block dofs per thread whose global id and halo slot are kept in
    // registers from the gather to the end of the block (more are reloaded)
    static constexpr int ND = 2;
    uint16_t li[EPT][K];
    T kv[PRE ? EPT : 1][PRE ? K * K : 1];
    int gd[ND], sd[ND];
};

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_load(int b, const FgEbeLayout &L, const T *__restrict__ Kv,
                                            FgEbeRegs<K, EPT, T> &r)
{
    const int e0 = L.eoff[b];
    const unsigned ne = (unsigned)(L.eoff[b + 1] - e0);
    const T *Kb = Kv + (size_t)K * K * e0;
    const uint16_t *Lb = L.lml + (size_t)K * e0;
    // threads past the block's last element load its last element again
    // (ne >= 1), so every load is issued on every path and the memory counter
    // waits stay exact; their results are never stored
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const unsigned el0 = threadIdx.x + i * FG_TPB;
        const unsigned el = el0 < ne ? el0 : ne - 1;
#pragma unroll
        for (int j = 0; j < K; j++) r.li[i][j] = fg_ld_nt(Lb, j * ne + el);
        if constexpr (FgEbeRegs<K, EPT, T>::PRE) {
            if constexpr (fg_ebe_vec(K) == 4) {
#pragma unroll
                for (int g = 0; g < K * K / 4; g++) fg_ld4(Kb, g * ne + el, &r.kv[i][4 * g]);
            } else {
#pragma unroll
                for (int q = 0; q < K * K; q++) r.kv[i][q] = fg_ld_nt(Kb, q * ne + el);
            }
        }
    }
}

// LDS carve-up of a block: P values, element row sums, incidence lists
template <int K, int EPT, typename T>
struct FgEbeSmem {
    T *sP, *sY;
    uint16_t *sInc, *sIoff;
    __device__ __forceinline__ FgEbeSmem(unsigned char *smem, int nd, int ne, int ninc)
    {
        sInc = reinterpret_cast<uint16_t *>(smem);
        sIoff = sInc + ninc;
        sP = reinterpret_cast<T *>(smem + fg_ebe_smem_idx_bytes(nd, ninc));
        sY = fg_ebe_alias(K, EPT) ? sP : sP + nd + 1;
        (void)ne;
    }
};

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_gather(int b, const FgEbeLayout &L,
                                              const T *__restrict__ P, unsigned char *smem,
                                              FgEbeRegs<K, EPT, T> &r)
{
    const int d0 = L.doff[b], nd = L.doff[b + 1] - d0;
    const int e0 = L.eoff[b], ne = L.eoff[b + 1] - e0;
    const int i0 = L.ibase[b], ninc = L.ibase[b + 1] - i0;
    FgEbeSmem<K, EPT, T> S(smem, nd, ne, ninc);
    const uint16_t *io = L.ioff + d0 + b;
    const int *Db = L.dofs + d0, *Sb = L.slots + d0;
    // incidence lists: 16-byte aligned, padded to 8 entries (see the plan)
    const uint4 *Ib = reinterpret_cast<const uint4 *>(L.inc + i0);
    uint4 *sInc4 = reinterpret_cast<uint4 *>(S.sInc);
    for (unsigned q = threadIdx.x; (int)q < (ninc >> 3); q += FG_TPB) sInc4[q] = fg_at(Ib, q);
    constexpr int ND = FgEbeRegs<K, EPT, T>::ND;
#pragma unroll
    for (int u = 0; u < ND; u++) {
        const unsigned l = threadIdx.x + u * FG_TPB;
        if ((int)l < nd) {
            const int g = fg_at(Db, l);
            r.gd[u] = g;
            r.sd[u] = fg_at(Sb, l);
            S.sP[l] = fg_at(P, (unsigned)g);
            S.sIoff[l] = fg_at(io, l);
        }
    }
    for (unsigned l = threadIdx.x + ND * FG_TPB; (int)l < nd; l += FG_TPB) {
        S.sP[l] = fg_at(P, (unsigned)fg_at(Db, l));
        S.sIoff[l] = fg_at(io, l);
    }
    if (threadIdx.x == 0) {
        S.sP[nd] = T(0);
        S.sIoff[nd] = io[nd];
    }
    __syncthreads();
}

// BS > 0: the row sums leave out the BS x BS diagonal block of the row's
// node (the off-block product of ebe2mvNDOF4.c:52-67, BS = 4); those entries
// are not read.  First summed column of row rr:
template <int BS>
__device__ __forceinline__ constexpr int fg_ebe_c0(int rr) { return BS > 0 && rr < BS ? BS : 0; }

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_reduce(const FgEbeBlk &B, const FgEbeLayout &L,
                                              T *__restrict__ Q, T *__restrict__ hpart,
                                              unsigned char *smem, const FgEbeRegs<K, EPT, T> &r);

template <int K, int EPT, typename T, int BS = 0>
__device__ __forceinline__ void fg_ebe_finish(const FgEbeBlk &B, const FgEbeLayout &L,
                                              const T *__restrict__ Kv, T *__restrict__ Q,
                                              T *__restrict__ hpart, unsigned char *smem,
                                              const FgEbeRegs<K, EPT, T> &r)
{
    const int e0 = B.e0;
    const unsigned ne = (unsigned)B.ne;
    const int nd = B.nd;
    const int ninc = B.ninc;
    FgEbeSmem<K, EPT, T> S(smem, nd, ne, ninc);
    const T *Kb = Kv + (size_t)K * K * e0;
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const unsigned el = threadIdx.x + i * FG_TPB;
        T p[K];
        if (el < ne) {
#pragma unroll
            for (int j = 0; j < K; j++) p[j] = S.sP[r.li[i][j]];
        }
        // aliased LDS (EPT == 1, one pass of this loop): every thread has read
        // its P values before any row sum overwrites them
        if constexpr (fg_ebe_alias(K, EPT)) __syncthreads();
        if (el < ne) {
#pragma unroll
            for (int rr = 0; rr < K; rr++) {
                const int c0 = fg_ebe_c0<BS>(rr);
                T acc;
                if constexpr (FgEbeRegs<K, EPT, T>::PRE) {
                    acc = r.kv[i][rr * K + c0] * p[c0];
#pragma unroll
                    for (int c = c0 + 1; c < K; c++)
                        if (BS == 0 || c / BS != rr / BS) acc = acc + r.kv[i][rr * K + c] * p[c];
                } else if constexpr (fg_ebe_vec(K) == 4) {
                    // row rr = groups rr*K/4 .. (rr+1)*K/4 - 1, four columns each;
                    // with BS = 4 a group is one node block and is skipped whole
                    static_assert(K % 4 == 0 && (BS == 0 || BS == 4), "grouped rows");
                    acc = T(0);
#pragma unroll
                    for (int gq = 0; gq < K / 4; gq++) {
                        if (BS != 0 && gq == rr / 4) continue;
                        T t[4];
                        fg_ld4(Kb, (rr * (K / 4) + gq) * ne + el, t);
#pragma unroll
                        for (int m = 0; m < 4; m++) {
                            const int c = 4 * gq + m;
                            acc = c == c0 ? t[m] * p[c] : acc + t[m] * p[c];
                        }
                    }
                } else {
                    acc = fg_ld_nt(Kb, (rr * K + c0) * ne + el) * p[c0];
#pragma unroll
                    for (int c = c0 + 1; c < K; c++)
                        if (BS == 0 || c / BS != rr / BS)
                            acc = acc + fg_ld_nt(Kb, (rr * K + c) * ne + el) * p[c];
                }
                S.sY[rr * ne + el] = acc;
            }
        }
    }
    __syncthreads();
    fg_ebe_reduce<K, EPT, T>(B, L, Q, hpart, smem, r);
}

template <int K, int EPT, typename T, int BS = 0>
__device__ __forceinline__ void fg_ebe_finish(int b, const FgEbeLayout &L,
                                              const T *__restrict__ Kv, T *__restrict__ Q,
                                              T *__restrict__ hpart, unsigned char *smem,
                                              const FgEbeRegs<K, EPT, T> &r)
{
    fg_ebe_finish<K, EPT, T, BS>(fg_ebe_blk(L, b), L, Kv, Q, hpart, smem, r);
}

// the incidence sums of a block's row sums (in LDS): every local dof adds its
// incidences in their stored order; dofs complete in the block -> Q, shared
// dofs -> their halo slot
template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_reduce(const FgEbeBlk &B, const FgEbeLayout &L,
                                              T *__restrict__ Q, T *__restrict__ hpart,
                                              unsigned char *smem, const FgEbeRegs<K, EPT, T> &r)
{
    const unsigned ne = (unsigned)B.ne;
    const int d0 = B.d0, nd = B.nd;
    const int ninc = B.ninc;
    FgEbeSmem<K, EPT, T> S(smem, nd, ne, ninc);
    const int *Db = L.dofs + d0, *Sb = L.slots + d0;
    constexpr int ND = FgEbeRegs<K, EPT, T>::ND;
#pragma unroll
    for (int u = 0; u < ND; u++) {
        const int l = threadIdx.x + u * FG_TPB;
        if (l < nd) {
            const int s0 = S.sIoff[l], s1 = S.sIoff[l + 1];
            T acc = S.sY[S.sInc[s0]];
            for (int q = s0 + 1; q < s1; q++) acc = acc + S.sY[S.sInc[q]];
            if (r.sd[u] < 0)
                fg_at(Q, (unsigned)r.gd[u]) = acc;
            else
                fg_at(hpart, (unsigned)r.sd[u]) = acc;
        }
    }
    for (unsigned l = threadIdx.x + ND * FG_TPB; (int)l < nd; l += FG_TPB) {
        const int s0 = S.sIoff[l], s1 = S.sIoff[l + 1];
        T acc = S.sY[S.sInc[s0]];
        for (int q = s0 + 1; q < s1; q++) acc = acc + S.sY[S.sInc[q]];
        const int s = fg_at(Sb, l);
        if (s < 0)
            fg_at(Q, (unsigned)fg_at(Db, l)) = acc;
        else
            fg_at(hpart, (unsigned)s) = acc;
    }
}

// ---------------------------------------------------------------------------
// 12 x 12 element rows on the matrix cores (an alternative to the VALU row
// sums above for k = 12, one element per thread).  A wave takes tiles of 16
// elements; for row rr of the tile's elements, three 16x16x4 MFMAs (one per
// group of four columns) multiply A[i][c] = K_{e_i}[rr][4g+c] by
// B[c][j] = p_{e_j}[4g+c], and the wanted row sum of element i is the
// diagonal entry D[i][i].  Only 16 of the 256 outputs are used: the 12x12
// products are 0.25 flop per byte, and this form exists to measure whether
// moving the multiply-adds off the VALU frees issue slots for loads.
// Operand layouts (16x16x4, 64 lanes, lane = i + 16 c): A[i][c] and B[c][j]
// one value per lane.  Output, register v of lane j + 16 q (measured on
// gfx950 with a one-wave probe of both builtins): FP32 D[4 q + v][j], FP64
// D[q + 4 v][j].  So D[i][i] is in lane i + 16 (i / 4), register i % 4
// (FP32), or lane i + 16 (i % 4), register i / 4 (FP64).
// BS = 4: the row node's own 4x4 block (column group rr / 4) is skipped.
template <typename T> struct FgMfma16;
template <> struct FgMfma16<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // lane i + 16 c holds D[i][i] in register *v
    static __device__ __forceinline__ bool diag(unsigned i, unsigned c, unsigned *v)
    {
        *v = i >> 2;
        return (i & 3) == c;
    }
};
template <> struct FgMfma16<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ bool diag(unsigned i, unsigned c, unsigned *v)
    {
        *v = i & 3;
        return (i >> 2) == c;
    }
};

template <typename T, int BS>
__device__ __forceinline__ void fg_ebe_rows_mfma12(int b, const FgEbeLayout &L,
                                                   const T *__restrict__ Kv, unsigned char *smem)
{
    constexpr int K = 12, G = K / 4;
    constexpr int WAVES = FG_TPB / 64, NT = FG_TPB / 16 / WAVES;   // tiles per wave
    static_assert(fg_ebe_vec(K) == 4 && fg_ebe_alias(K, 1), "grouped, aliased layout");
    using M = FgMfma16<T>;
    const int e0 = L.eoff[b];
    const unsigned ne = (unsigned)(L.eoff[b + 1] - e0);   // <= FG_TPB (one element per thread)
    const int d0 = L.doff[b], nd = L.doff[b + 1] - d0;
    const int ninc = L.ibase[b + 1] - L.ibase[b];
    FgEbeSmem<K, 1, T> S(smem, nd, ne, ninc);
    const T *Kb = Kv + (size_t)K * K * e0;
    const uint16_t *Lb = L.lml + (size_t)K * e0;
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned i = lane & 15, c = lane >> 4;
    // B operands of every tile first: the row sums overwrite the P values
    T bp[NT][G];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const unsigned el = (wave + t * WAVES) * 16 + i;
#pragma unroll
        for (int g = 0; g < G; g++)
            bp[t][g] = el < ne ? S.sP[fg_ld_nt(Lb, (4 * g + c) * ne + el)] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const unsigned base = (wave + t * WAVES) * 16;
        if (base >= ne) break;   // wave-uniform: MFMAs run with the whole wave
        const unsigned el = base + i;
#pragma unroll
        for (int rr = 0; rr < K; rr++) {
            typename M::acc d = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int g = 0; g < G; g++) {
                if (BS != 0 && g == rr / 4) continue;
                const unsigned q = ((rr * G + g) * ne + el) * 4 + c;
                T a = T(0);
                if (el < ne) a = sizeof(T) == 4 ? fg_ld_nt(Kb, q) : fg_at(Kb, q);
                d = M::mma(a, bp[t][g], d);
            }
            unsigned v;
            if (M::diag(i, c, &v) && el < ne)
                S.sY[rr * ne + el] = v == 0 ? d[0] : v == 1 ? d[1] : v == 2 ? d[2] : d[3];
        }
    }
}

// load and gather of one block in the order the memory counter needs.  A
// wave waits for its loads in issue order (a wait for one load is a wait for
// every load issued before it), so the loads are issued as: the block's
// incidence lists, dof ids, halo slots and incidence offsets; then the P
// gathers (which wait for the dof ids only); then the element values and
// local indices -- and the LDS writes, which wait for the first two groups,
// come after all of them.  The value stream is thus in flight during the
// whole gather instead of being waited for by its first LDS write.
template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_prologue(int b, const FgEbeLayout &L,
                                                const T *__restrict__ Kv,
                                                const T *__restrict__ P, unsigned char *smem,
                                                FgEbeRegs<K, EPT, T> &r)
{
    const int d0 = L.doff[b], nd = L.doff[b + 1] - d0;
    const int e0 = L.eoff[b], ne = L.eoff[b + 1] - e0;
    const int i0 = L.ibase[b], ninc = L.ibase[b + 1] - i0;
    FgEbeSmem<K, EPT, T> S(smem, nd, ne, ninc);
    const uint16_t *io = L.ioff + d0 + b;
    const int *Db = L.dofs + d0, *Sb = L.slots + d0;
    const uint4 *Ib = reinterpret_cast<const uint4 *>(L.inc + i0);
    uint4 *sInc4 = reinterpret_cast<uint4 *>(S.sInc);
    constexpr int ND = FgEbeRegs<K, EPT, T>::ND;
    // incidence lists: at most K * 256 * EPT slots, in 16-byte groups
    constexpr int NI = (K * EPT + 7) / 8;
    const int ninc8 = ninc >> 3;
    // indices past a list's end are clamped to its last entry (every block
    // has at least one incidence group and one dof): all loads are issued on
    // every path, the surplus results are not stored
    uint4 ic[NI];
#pragma unroll
    for (int v = 0; v < NI; v++) {
        const int q = threadIdx.x + v * FG_TPB;
        ic[v] = fg_at(Ib, (unsigned)(q < ninc8 ? q : ninc8 - 1));
    }
    uint16_t iv[ND];
#pragma unroll
    for (int u = 0; u < ND; u++) {
        const unsigned l0 = threadIdx.x + u * FG_TPB;
        const unsigned l = (int)l0 < nd ? l0 : nd - 1;
        r.gd[u] = fg_at(Db, l);
        r.sd[u] = fg_at(Sb, l);
        iv[u] = fg_at(io, l);
    }
    const uint16_t iend = io[nd];
    T pv[ND];
#pragma unroll
    for (int u = 0; u < ND; u++) pv[u] = fg_at(P, (unsigned)r.gd[u]);
    fg_ebe_load<K, EPT, T>(b, L, Kv, r);
    // the LDS writes are unconditional (surplus lanes write a dummy slot), so
    // that the compiler keeps the loads above where they are instead of
    // sinking each into a branch next to its only store
    __shared__ uint4 dummy4;
    __shared__ T dummyT;
    __shared__ uint16_t dummy16;
#pragma unroll
    for (int v = 0; v < NI; v++) {
        const int q = threadIdx.x + v * FG_TPB;
        *(q < ninc8 ? &sInc4[q] : &dummy4) = ic[v];
    }
    for (int q = threadIdx.x + NI * FG_TPB; q < ninc8; q += FG_TPB) sInc4[q] = fg_at(Ib, (unsigned)q);
#pragma unroll
    for (int u = 0; u < ND; u++) {
        const unsigned l = threadIdx.x + u * FG_TPB;
        const bool in = (int)l < nd;
        *(in ? &S.sP[l] : &dummyT) = pv[u];
        *(in ? &S.sIoff[l] : &dummy16) = iv[u];
    }
    for (unsigned l = threadIdx.x + ND * FG_TPB; (int)l < nd; l += FG_TPB) {
        S.sP[l] = fg_at(P, (unsigned)fg_at(Db, l));
        S.sIoff[l] = fg_at(io, l);
    }
    if (threadIdx.x == 0) {
        S.sP[nd] = T(0);
        S.sIoff[nd] = iend;
    }
    __syncthreads();
}

// Split form of the prologue for the persistent schedule: the index loads
// of a block (into registers), then -- later, once the previous block is
// finished -- the P gathers and the LDS fill.
template <int K, int EPT, typename T>
struct FgEbeIdx {
    static constexpr int NI = (K * EPT + 7) / 8;
    static constexpr int ND = FgEbeRegs<K, EPT, T>::ND;
    uint4 ic[NI];
    uint16_t iv[ND];
    uint16_t iend;
};

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_idx_load(int b, const FgEbeLayout &L,
                                                FgEbeRegs<K, EPT, T> &r, FgEbeIdx<K, EPT, T> &x)
{
    const int d0 = L.doff[b], nd = L.doff[b + 1] - d0;
    const int i0 = L.ibase[b], ninc8 = (L.ibase[b + 1] - i0) >> 3;
    const uint16_t *io = L.ioff + d0 + b;
    const int *Db = L.dofs + d0, *Sb = L.slots + d0;
    const uint4 *Ib = reinterpret_cast<const uint4 *>(L.inc + i0);
#pragma unroll
    for (int v = 0; v < FgEbeIdx<K, EPT, T>::NI; v++) {
        const int q = threadIdx.x + v * FG_TPB;
        x.ic[v] = fg_at(Ib, (unsigned)(q < ninc8 ? q : ninc8 - 1));
    }
#pragma unroll
    for (int u = 0; u < FgEbeIdx<K, EPT, T>::ND; u++) {
        const unsigned l0 = threadIdx.x + u * FG_TPB;
        const unsigned l = (int)l0 < nd ? l0 : nd - 1;
        r.gd[u] = fg_at(Db, l);
        r.sd[u] = fg_at(Sb, l);
        x.iv[u] = fg_at(io, l);
    }
    x.iend = io[nd];
}

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_fill(int b, const FgEbeLayout &L, const T *__restrict__ P,
                                            unsigned char *smem, const FgEbeRegs<K, EPT, T> &r,
                                            const FgEbeIdx<K, EPT, T> &x)
{
    const int d0 = L.doff[b], nd = L.doff[b + 1] - d0;
    const int ne = L.eoff[b + 1] - L.eoff[b];
    const int i0 = L.ibase[b], ninc = L.ibase[b + 1] - i0, ninc8 = ninc >> 3;
    FgEbeSmem<K, EPT, T> S(smem, nd, ne, ninc);
    const uint16_t *io = L.ioff + d0 + b;
    const int *Db = L.dofs + d0;
    const uint4 *Ib = reinterpret_cast<const uint4 *>(L.inc + i0);
    uint4 *sInc4 = reinterpret_cast<uint4 *>(S.sInc);
    constexpr int NI = FgEbeIdx<K, EPT, T>::NI, ND = FgEbeIdx<K, EPT, T>::ND;
    T pv[ND];
#pragma unroll
    for (int u = 0; u < ND; u++) pv[u] = fg_at(P, (unsigned)r.gd[u]);
    __shared__ uint4 dummy4;
    __shared__ T dummyT;
    __shared__ uint16_t dummy16;
#pragma unroll
    for (int v = 0; v < NI; v++) {
        const int q = threadIdx.x + v * FG_TPB;
        *(q < ninc8 ? &sInc4[q] : &dummy4) = x.ic[v];
    }
    for (int q = threadIdx.x + NI * FG_TPB; q < ninc8; q += FG_TPB) sInc4[q] = fg_at(Ib, (unsigned)q);
#pragma unroll
    for (int u = 0; u < ND; u++) {
        const unsigned l = threadIdx.x + u * FG_TPB;
        const bool in = (int)l < nd;
        *(in ? &S.sP[l] : &dummyT) = pv[u];
        *(in ? &S.sIoff[l] : &dummy16) = x.iv[u];
    }
    for (unsigned l = threadIdx.x + ND * FG_TPB; (int)l < nd; l += FG_TPB) {
        S.sP[l] = fg_at(P, (unsigned)fg_at(Db, l));
        S.sIoff[l] = fg_at(io, l);
    }
    if (threadIdx.x == 0) {
        S.sP[nd] = T(0);
        S.sIoff[nd] = x.iend;
    }
    __syncthreads();
}

// the three phases back to back (one block per workgroup)
template <int K, int EPT, typename T, int BS = 0>
__device__ __forceinline__ void fg_ebe_block(int b, const FgEbeLayout &L, const T *__restrict__ Kv,
                                             const T *__restrict__ P, T *__restrict__ Q,
                                             T *__restrict__ hpart, unsigned char *smem)
{
    FgEbeRegs<K, EPT, T> r;
    fg_ebe_prologue<K, EPT, T>(b, L, Kv, P, smem, r);
    fg_ebe_finish<K, EPT, T, BS>(b, L, Kv, Q, hpart, smem, r);
}
