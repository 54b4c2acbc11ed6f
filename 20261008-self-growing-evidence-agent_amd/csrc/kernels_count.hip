// Batched AND+popcount over explicit (prefix, candidate list) nodes: the kernel
// behind sgfim_count_batch, which replaces the reference's per-node counter call
// getResFrequencyCuda (code/engine_tensorfim/code/engine/src/kernel_bitmap.cu:451-536).
// The reference launches one grid per search node and synchronises its stream each
// time; here all nodes of a batch are cut into (node, chunk) work units that a
// persistent grid drains from one device work counter.
#include "count_core.h"

#include <mutex>

template <int CW>
__global__ __launch_bounds__(256) void k_count_nodes(const uint64_t* __restrict__ rows,
                                                     const SgSegment* __restrict__ segs,
                                                     const SgCountNode* __restrict__ nodes,
                                                     int32_t n_nodes, int64_t n_units,
                                                     const int32_t* __restrict__ cand,
                                                     int32_t* __restrict__ out,
                                                     int32_t* __restrict__ unit_counter)
{
    __shared__ uint64_t pre[CW];
    __shared__ int64_t s_unit;
    __shared__ int32_t s_items[SG_MAX_DEPTH];
    const int wave = threadIdx.x >> 6;
    const int nwave = blockDim.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int64_t)atomicAdd(unit_counter, 1);
        __syncthreads();
        const int64_t u = s_unit;
        if (u >= n_units) break;
        // node of unit u: last node with unit_off <= u
        int lo = 0, hi = n_nodes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (nodes[mid].unit_off <= u) lo = mid; else hi = mid - 1;
        }
        const SgCountNode& nd = nodes[lo];
        const SgSegment sg = segs[nd.seg];
        const int c = (int)(u - nd.unit_off);
        const int64_t w0 = (int64_t)c * CW;
        const int wn = (int)min((int64_t)CW, sg.words - w0);
        if (threadIdx.x < nd.k) s_items[threadIdx.x] = nd.items[threadIdx.x];
        __syncthreads();
        const uint64_t* seg_rows = rows + sg.row_base;
        sg_build_prefix(pre, seg_rows, sg.stride, w0, wn, nd.k, s_items);
        __syncthreads();
        for (int j = wave; j < nd.ncand; j += nwave) {
            const int32_t e = cand[nd.cand_off + j];
            const uint32_t s = sg_count_row<CW>(pre, seg_rows + (int64_t)e * sg.stride + w0, wn);
            if ((threadIdx.x & 63) == 0 && s) atomicAdd(&out[nd.cand_off + j], (int32_t)s);
        }
        __syncthreads();
    }
}

hipError_t sg_launch_count_nodes(const uint64_t* rows, const SgSegment* segs,
                                 const SgCountNode* nodes, int32_t n_nodes, int64_t n_units,
                                 const int32_t* cand, int32_t* out, int32_t chunk_words,
                                 int32_t* unit_counter, int32_t grid, hipStream_t s)
{
    if (n_units <= 0) return hipSuccess;
    const unsigned g = (unsigned)(grid > 0 ? grid : 2048);
    switch (chunk_words) {
    case 256:
        hipLaunchKernelGGL(k_count_nodes<256>, dim3(g), dim3(256), 0, s, rows, segs, nodes, n_nodes, n_units, cand, out, unit_counter);
        break;
    case 512:
        hipLaunchKernelGGL(k_count_nodes<512>, dim3(g), dim3(256), 0, s, rows, segs, nodes, n_nodes, n_units, cand, out, unit_counter);
        break;
    case 2048:
        hipLaunchKernelGGL(k_count_nodes<2048>, dim3(g), dim3(256), 0, s, rows, segs, nodes, n_nodes, n_units, cand, out, unit_counter);
        break;
    case 1024:
        hipLaunchKernelGGL(k_count_nodes<1024>, dim3(g), dim3(256), 0, s, rows, segs, nodes, n_nodes, n_units, cand, out, unit_counter);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ allocation-free scratch
// Device scratch of sgfim_count_batch_noalloc: node descriptors, compacted candidate ids,
// their counts and the work counter, in one static array per device (each GPU the library
// runs on holds its own copy until the process ends). Host code must not pass `scratch`
// itself to a kernel or copy (its name is a host-side stand-in); sg_count_scratch() looks
// up each device's address once and every caller uses that.
static __device__ __align__(256) unsigned char scratch[SG_NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SG_NOALLOC_MAX_DEVICES] = {};

hipError_t sg_count_scratch(int device, void** ptr)
{
    if (device < 0 || device >= SG_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}
