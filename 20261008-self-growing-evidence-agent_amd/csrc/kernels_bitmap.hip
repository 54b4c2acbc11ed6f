// CSR transactions -> vertical uint64 tidset bitmaps, with the item-support epilogue.
//
// Restates buildBitmap (code/engine_tensorfim/code/engine/src/kernel_bitmap.cu:342-380):
// bit (t & 63) of word (t >> 6) of row x is set iff transaction t holds item x. The
// reference builds the bitmap on the host, one transaction at a time; here one
// workgroup owns a tile of TP_DI items x TP_TW words, ORs the bits of its 64*TP_TW
// transactions into LDS, and writes each item's TP_TW words as one contiguous
// 128-byte segment. The popcount of every written word is summed per item with a
// 16-lane shuffle reduction and one atomic per (item, tile), which yields the item
// supports of getItemsFrequenc (kernel_bitmap.cu:269-298) for free.
#include "sgfim_impl.h"

#define TP_TW 16
#define TP_DI 512

__global__ __launch_bounds__(256) void k_transpose(const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ items,
                                                   const int32_t* __restrict__ idmap,
                                                   int64_t n_txn, int32_t id_max,
                                                   uint64_t* __restrict__ rows, int32_t d,
                                                   int64_t Wp, int32_t* __restrict__ sup)
{
    __shared__ unsigned long long tile[TP_DI * TP_TW];
    const int64_t w0 = (int64_t)blockIdx.x * TP_TW;
    const int32_t i0 = blockIdx.y * TP_DI;
    const int32_t di = min(TP_DI, d - i0);
    for (int q = threadIdx.x; q < TP_DI * TP_TW; q += blockDim.x) tile[q] = 0ull;
    __syncthreads();
    const int64_t t0 = w0 * 64;
    for (int q = threadIdx.x; q < TP_TW * 64; q += blockDim.x) {
        const int64_t t = t0 + q;
        if (t >= n_txn) break;
        const int64_t a = off[t], b = off[t + 1];
        for (int64_t p = a; p < b; p++) {
            const int32_t id = items[p];
            if (id < 1 || id > id_max) continue;
            const int32_t x = idmap[id] - i0;
            if ((unsigned)x < (unsigned)di)
                atomicOr(&tile[x * TP_TW + (q >> 6)], 1ull << (q & 63));
        }
    }
    __syncthreads();
    // 256 threads = 16 items x 16 words per pass; each group of 16 lanes is one item.
    for (int base = 0; base < TP_DI; base += 16) {
        const int x = base + (threadIdx.x >> 4);
        const int wl = threadIdx.x & 15;
        uint64_t v = 0;
        const bool live = x < di;
        if (live) {
            v = tile[x * TP_TW + wl];
            rows[(int64_t)(i0 + x) * Wp + w0 + wl] = v;
        }
        uint32_t c = (uint32_t)__popcll(v);
        c += __shfl_xor(c, 8, 16);
        c += __shfl_xor(c, 4, 16);
        c += __shfl_xor(c, 2, 16);
        c += __shfl_xor(c, 1, 16);
        if (live && wl == 0 && c) atomicAdd(&sup[i0 + x], (int32_t)c);
    }
}

hipError_t sg_launch_transpose(const int64_t* off, const int32_t* items, const int32_t* idmap,
                               int64_t n_txn, int32_t id_max, uint64_t* rows, int32_t d,
                               int32_t dpad, int64_t Wp, int32_t* supports, hipStream_t s)
{
    (void)dpad;
    if (d <= 0 || Wp <= 0) return hipSuccess;
    dim3 grid((unsigned)(Wp / TP_TW), (unsigned)((d + TP_DI - 1) / TP_DI));
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, off, items, idmap, n_txn, id_max,
                       rows, d, Wp, supports);
    return hipGetLastError();
}

// Compact search universe: one wave per (kept row, segment) copies the row and its support.
__global__ __launch_bounds__(256) void k_gather_rows(const uint64_t* __restrict__ src, const int32_t* __restrict__ src_sup,
                                                     const SgSegment* __restrict__ ssegs, int32_t sdpad,
                                                     uint64_t* __restrict__ dst, int32_t* __restrict__ dst_sup,
                                                     const SgSegment* __restrict__ dsegs, int32_t ddpad,
                                                     const int32_t* __restrict__ keep, int32_t dc)
{
    const int w = blockIdx.y;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= dc) return;
    const SgSegment sg = ssegs[w], tg = dsegs[w];
    const int32_t x = keep[i];
    const uint64_t* a = src + sg.row_base + (int64_t)x * sg.stride;
    uint64_t* b = dst + tg.row_base + (int64_t)i * tg.stride;
    for (int64_t k = lane; k < sg.stride; k += 64) b[k] = a[k];
    if (lane == 0) dst_sup[(int64_t)w * ddpad + i] = src_sup[(int64_t)w * sdpad + x];
}

hipError_t sg_launch_gather_rows(const uint64_t* src, const int32_t* src_sup, const SgSegment* src_segs,
                                 int32_t src_dpad, uint64_t* dst, int32_t* dst_sup, const SgSegment* dst_segs,
                                 int32_t dst_dpad, const int32_t* keep, int32_t dc, int32_t nseg, hipStream_t s)
{
    if (dc <= 0 || nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((dc + 3) / 4), (unsigned)nseg), dim3(256), 0, s, src, src_sup,
                       src_segs, src_dpad, dst, dst_sup, dst_segs, dst_dpad, keep, dc);
    return hipGetLastError();
}

// Plain device copy (16-byte loads and stores, grid-stride): the on-box bandwidth probe
// and the known-size reference for hardware byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t sg_launch_copy_probe(const void* src, void* dst, int64_t bytes, hipStream_t s)
{
    const int64_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}
