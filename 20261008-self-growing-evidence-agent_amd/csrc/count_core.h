// count_core.h -- the wave64 AND+popcount core shared by every counting kernel.
//
// Restates the reference's per-node counter (getResFrequencyCudaKernel,
// code/engine_tensorfim/code/engine/src/kernel_bitmap.cu:412-442: for survivor e,
// sum over words of popc(row_e & AND_i row_{X_i})) for a chunk of CW words:
//   1. the workgroup ANDs the k prefix rows of the chunk once and stages the
//      prefix tidset in LDS (the reference re-reads the k prefix rows for every e);
//   2. each wave takes candidates round-robin and streams the candidate's chunk
//      with coalesced 64-bit loads (64 lanes x 8 B = 512 B per load), ANDs with
//      the LDS prefix and accumulates v_bcnt popcounts in a 32-bit lane register;
//   3. the 64 lane partials are reduced with DPP row rotations inside each row
//      of 16 lanes plus four readlanes, and lane 0 adds the sum into the
//      candidate's int32 count with one atomic (integer addition is exact and
//      order-free, so split-N partial sums give bit-identical totals).
#pragma once
#include "sgfim_impl.h"

// Sum of v over the 64 lanes of the wave; every lane must be active.
__device__ __forceinline__ uint32_t sg_wave_sum(uint32_t v)
{
    // DPP row_ror:n (0x120 + n) rotates within each row of 16 lanes.
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xf, 0xf, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

// Build the prefix tidset of words [w0, w0 + wn) in LDS: AND of the k rows
// (k == 0 gives all ones). items[] are row indices relative to seg_rows.
template <typename ItemT>
__device__ __forceinline__ void sg_build_prefix(uint64_t* __restrict__ pre,
                                                const uint64_t* __restrict__ seg_rows,
                                                int64_t stride, int64_t w0, int wn, int k,
                                                const ItemT* items)
{
    for (int w = threadIdx.x; w < wn; w += blockDim.x) {
        uint64_t v = ~0ull;
        for (int i = 0; i < k; i++) v &= seg_rows[(int64_t)items[i] * stride + w0 + w];
        pre[w] = v;
    }
}

// popcount(pre & row[w0 .. w0+wn)) for one candidate row, summed over the wave.
template <int CW>
__device__ __forceinline__ uint32_t sg_count_row(const uint64_t* __restrict__ pre,
                                                 const uint64_t* __restrict__ row, int wn)
{
    const int lane = threadIdx.x & 63;
    uint32_t acc = 0;
    if (wn == CW) {
        constexpr int PER = CW / 64;
        uint64_t x[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) x[q] = row[lane + 64 * q];
#pragma unroll
        for (int q = 0; q < PER; q++) acc += (uint32_t)__popcll(x[q] & pre[lane + 64 * q]);
    } else {
        int w = lane;
        for (; w + 192 < wn; w += 256) {
            uint64_t x0 = row[w], x1 = row[w + 64], x2 = row[w + 128], x3 = row[w + 192];
            acc += (uint32_t)__popcll(x0 & pre[w]) + (uint32_t)__popcll(x1 & pre[w + 64]) +
                   (uint32_t)__popcll(x2 & pre[w + 128]) + (uint32_t)__popcll(x3 & pre[w + 192]);
        }
        for (; w < wn; w += 64) acc += (uint32_t)__popcll(row[w] & pre[w]);
    }
    return sg_wave_sum(acc);
}
