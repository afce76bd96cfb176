// fg_ebe_body.h -- one workgroup's share of an element-by-element product:
// the block body used by the stand-alone EBE/EDE kernel (fg_ebe.hip) and by
// the matvec phase of the persistent GMRES kernel (fg_gmres_fused.hip).
//
// Block b holds ne elements (colour-sorted) and nd distinct dofs.  Values are
// stored per block as [k*k][ne] and local dof indices as [k][ne], so every
// global load is coalesced.  Steps:
//   1. issue the loads of this thread's local indices (and, when they fit in
//      registers, its element values) -- independent of P;
//   2. gather the block's dofs of P into LDS, zero the LDS accumulators;
//   3. row sums of each element (reference order: left to right over the
//      columns, ebemv.c:34-36);
//   4. add the row sums into the accumulators one colour at a time (no two
//      elements of a colour share a dof, so no conflicts and a fixed order);
//   5. dofs complete in this block -> Q, shared dofs -> their halo slot.
#pragma once

#include "fg_internal.h"

// default elements per thread (block = 256 * EPT elements): keeps the element
// values of a thread (EPT * k * k) in registers for k <= 4
constexpr int fg_ebe_default_ept(int k) { return k <= 2 ? 8 : k == 3 ? 4 : k == 4 ? 2 : 1; }

template <int K, int EPT, typename T>
__device__ __forceinline__ void fg_ebe_block(
    int b, int maxc, const int *__restrict__ eoff, const int *__restrict__ doff,
    const uint16_t *__restrict__ coff, const int *__restrict__ dofs,
    const int *__restrict__ slots, const uint16_t *__restrict__ lml, const T *__restrict__ Kv,
    const T *__restrict__ P, T *__restrict__ Q, T *__restrict__ hpart, unsigned char *smem,
    uint16_t *s_cb)
{
    constexpr bool PRE = EPT * K * K <= 72;   // element values held in registers
    const int e0 = eoff[b], ne = eoff[b + 1] - e0;
    const int d0 = doff[b], nd = doff[b + 1] - d0;
    T *sP = reinterpret_cast<T *>(smem);
    T *sA = sP + nd + 1;
    const T *Kb = Kv + (size_t)K * K * e0;
    const uint16_t *Lb = lml + (size_t)K * e0;

    uint16_t li[EPT][K];
    T kv[PRE ? EPT : 1][PRE ? K * K : 1];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int el = threadIdx.x + i * FG_TPB;
        if (el < ne) {
#pragma unroll
            for (int j = 0; j < K; j++) li[i][j] = Lb[(size_t)j * ne + el];
            if constexpr (PRE) {
#pragma unroll
                for (int q = 0; q < K * K; q++) kv[i][q] = Kb[(size_t)q * ne + el];
            }
        }
    }
    for (int l = threadIdx.x; l < nd; l += FG_TPB) {
        sP[l] = P[dofs[d0 + l]];
        sA[l] = T(0);
    }
    if (threadIdx.x == 0) {
        sP[nd] = T(0);
        sA[nd] = T(0);
    }
    for (int c = threadIdx.x; c <= maxc; c += FG_TPB) s_cb[c] = coff[(size_t)b * (maxc + 1) + c];
    __syncthreads();

    T y[EPT][K];
    int col[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        const int el = threadIdx.x + i * FG_TPB;
        col[i] = -1;
        if (el < ne) {
            T p[K];
#pragma unroll
            for (int j = 0; j < K; j++) p[j] = sP[li[i][j]];
#pragma unroll
            for (int r = 0; r < K; r++) {
                T acc;
                if constexpr (PRE) {
                    acc = kv[i][r * K] * p[0];
#pragma unroll
                    for (int c = 1; c < K; c++) acc = acc + kv[i][r * K + c] * p[c];
                } else {
                    acc = Kb[(size_t)(r * K) * ne + el] * p[0];
#pragma unroll
                    for (int c = 1; c < K; c++) acc = acc + Kb[(size_t)(r * K + c) * ne + el] * p[c];
                }
                y[i][r] = acc;
            }
            int c = 0;
            while (s_cb[c + 1] <= el) c++;
            col[i] = c;
        }
    }
    for (int c = 0; c < maxc; c++) {
        if (s_cb[c] == s_cb[c + 1]) continue;   // empty colour: uniform over the block
#pragma unroll
        for (int i = 0; i < EPT; i++)
            if (col[i] == c) {
#pragma unroll
                for (int r = 0; r < K; r++) sA[li[i][r]] = sA[li[i][r]] + y[i][r];
            }
        __syncthreads();
    }
    for (int l = threadIdx.x; l < nd; l += FG_TPB) {
        const int s = slots[d0 + l];
        if (s < 0)
            Q[dofs[d0 + l]] = sA[l];
        else
            hpart[s] = sA[l];
    }
}
