// Pair co-occurrence matrix C[i][j] = |T_i & T_j| with int8 MFMA on 0/1 tiles.
//
// The reference counts the pairs as the first wave of its 1-bit BMMA engine: A =
// the single-item root masks, B = all survivor bitmaps, C = popc(A_i AND B_j)
// (code/engine_tensorfim/code/engine/src/kernel_bmma.cu:636-768, driven by
// batchCountBMMA :772-872 over the roots pushed at :1224-1236). CDNA4 has no 1-bit
// matrix instruction, so each 16-bit slice of a bitmap word is widened in
// registers to 16 bytes of 0/1 (nibble * 0x00204081 & 0x01010101 places the four
// bits of a nibble in the low bit of four bytes) and multiplied with
// v_mfma_i32_32x32x32_i8, accumulating exactly in int32. The widened tiles never
// touch memory. A and B fragments use the same (lane, byte) -> transaction map,
// so the K order inside an instruction is irrelevant to the sum.
//
// Tiling: a workgroup of 4 waves owns a 128 x 128 output tile, each wave a 64 x 64
// quarter (2 x 2 MFMA tiles, 4 accumulators of 16 int32 per lane); grid.z splits
// the transactions (split-N) and partial tiles are added with integer atomics,
// which keeps the result bit-exact.
#include "sgfim_impl.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v4i sg_widen16(uint32_t b)
{
    v4i r;
    r[0] = (int)(((b) & 0xFu) * 0x00204081u & 0x01010101u);
    r[1] = (int)(((b >> 4) & 0xFu) * 0x00204081u & 0x01010101u);
    r[2] = (int)(((b >> 8) & 0xFu) * 0x00204081u & 0x01010101u);
    r[3] = (int)(((b >> 12) & 0xFu) * 0x00204081u & 0x01010101u);
    return r;
}

// Accumulate the 64 x 64 block (I0.., J0..) of one segment over words [wb, we).
__device__ __forceinline__ void sg_pair_block(const uint64_t* __restrict__ rows, int64_t stride,
                                              int I0, int J0, int64_t wb, int64_t we,
                                              int32_t* __restrict__ out, int ld)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, h = lane >> 5;
    const uint64_t* a0 = rows + (int64_t)(I0 + r) * stride;
    const uint64_t* a1 = rows + (int64_t)(I0 + 32 + r) * stride;
    const uint64_t* b0 = rows + (int64_t)(J0 + r) * stride;
    const uint64_t* b1 = rows + (int64_t)(J0 + 32 + r) * stride;
    v16i c00 = {}, c01 = {}, c10 = {}, c11 = {};
    for (int64_t w = wb; w < we; w++) {
        const uint64_t xa0 = a0[w], xa1 = a1[w], xb0 = b0[w], xb1 = b1[w];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int sh = 32 * s + 16 * h;
            const v4i A0 = sg_widen16((uint32_t)(xa0 >> sh) & 0xFFFFu);
            const v4i A1 = sg_widen16((uint32_t)(xa1 >> sh) & 0xFFFFu);
            const v4i B0 = sg_widen16((uint32_t)(xb0 >> sh) & 0xFFFFu);
            const v4i B1 = sg_widen16((uint32_t)(xb1 >> sh) & 0xFFFFu);
            c00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B1, c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B0, c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B1, c11, 0, 0, 0);
        }
    }
    // 32 x 32 int32 accumulator layout: lane l, element v holds
    // row 8*(v/4) + 4*(l/32) + v%4, column l%32.
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int row = 8 * (v >> 2) + 4 * h + (v & 3);
        const int col = r;
        if (c00[v]) atomicAdd(&out[(int64_t)(I0 + row) * ld + J0 + col], c00[v]);
        if (c01[v]) atomicAdd(&out[(int64_t)(I0 + row) * ld + J0 + 32 + col], c01[v]);
        if (c10[v]) atomicAdd(&out[(int64_t)(I0 + 32 + row) * ld + J0 + col], c10[v]);
        if (c11[v]) atomicAdd(&out[(int64_t)(I0 + 32 + row) * ld + J0 + 32 + col], c11[v]);
    }
}

// grid (dpad/128, dpad/128, splits); out must be zeroed.
__global__ __launch_bounds__(256) void k_pair_mfma(const uint64_t* __restrict__ rows, int64_t Wp,
                                                   int64_t W, int32_t dpad, int32_t* __restrict__ out,
                                                   int64_t words_per_split)
{
    const int wave = threadIdx.x >> 6;
    const int I0 = blockIdx.x * 128 + (wave >> 1) * 64;
    const int J0 = blockIdx.y * 128 + (wave & 1) * 64;
    const int64_t wb = (int64_t)blockIdx.z * words_per_split;
    const int64_t we = min(W, wb + words_per_split);
    if (wb >= we) return;
    sg_pair_block(rows, Wp, I0, J0, wb, we, out, dpad);
}

// Batched over segments: grid (tiles of the largest segment, splits, nseg).
__global__ __launch_bounds__(256) void k_pair_mfma_segs(const uint64_t* __restrict__ rows,
                                                        const SgSegment* __restrict__ segs,
                                                        int32_t* __restrict__ out, int32_t splits)
{
    const SgSegment sg = segs[blockIdx.z];
    const int tiles = sg.dpad / 128;
    const int ti = blockIdx.x / tiles, tj = blockIdx.x % tiles;
    if (ti >= tiles) return;
    const int wave = threadIdx.x >> 6;
    const int I0 = ti * 128 + (wave >> 1) * 64;
    const int J0 = tj * 128 + (wave & 1) * 64;
    const int64_t per = (sg.words + splits - 1) / splits;
    const int64_t wb = (int64_t)blockIdx.y * per;
    const int64_t we = min(sg.words, wb + per);
    if (wb >= we) return;
    sg_pair_block(rows + sg.row_base, sg.stride, I0, J0, wb, we, out + sg.pair_base, sg.dpad);
}

// VALU cross-check: one wave per (i, j) pair, 64-bit AND + popcount.
__global__ __launch_bounds__(256) void k_pair_valu(const uint64_t* __restrict__ rows, int64_t Wp,
                                                   int64_t W, int32_t dpad, int32_t* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= (int64_t)dpad * dpad) return;
    const int i = (int)(p / dpad), j = (int)(p % dpad);
    const uint64_t* a = rows + (int64_t)i * Wp;
    const uint64_t* b = rows + (int64_t)j * Wp;
    uint32_t acc = 0;
    for (int64_t w = lane; w < W; w += 64) acc += (uint32_t)__popcll(a[w] & b[w]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) out[p] = (int32_t)acc;
}

// ok[i][w] bit b = pair(i, 64w+b) >= thr, for 64w+b < d.
__global__ void k_pair_ok(const int32_t* __restrict__ pair, const SgSegment* __restrict__ segs,
                          int32_t max_d, uint64_t* __restrict__ ok)
{
    const SgSegment sg = segs[blockIdx.y];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)sg.d * sg.dw) return;
    const int i = (int)(q / sg.dw), w = (int)(q % sg.dw);
    const int32_t* pr = pair + sg.pair_base + (int64_t)i * sg.dpad;
    uint64_t v = 0;
    for (int b = 0; b < 64; b++) {
        const int j = w * 64 + b;
        if (j < sg.d && pr[j] >= sg.thr) v |= 1ull << b;
    }
    ok[sg.ok_base + (int64_t)i * sg.dw + w] = v;
}

static int64_t pair_splits(int32_t dpad, int64_t W)
{
    const int64_t tiles = (int64_t)(dpad / 128) * (dpad / 128);
    int64_t splits = (2048 + tiles - 1) / tiles;
    if (splits > W) splits = W > 0 ? W : 1;
    if (splits < 1) splits = 1;
    return splits;
}

hipError_t sg_launch_pair_mfma(const uint64_t* rows, int32_t dpad, int64_t Wp, int64_t W,
                               int32_t* pair, hipStream_t s)
{
    hipError_t e = hipMemsetAsync(pair, 0, sizeof(int32_t) * (size_t)dpad * dpad, s);
    if (e != hipSuccess || W <= 0) return e;
    const int64_t splits = pair_splits(dpad, W);
    const int64_t per = (W + splits - 1) / splits;
    dim3 grid(dpad / 128, dpad / 128, (unsigned)((W + per - 1) / per));
    hipLaunchKernelGGL(k_pair_mfma, grid, dim3(256), 0, s, rows, Wp, W, dpad, pair, per);
    return hipGetLastError();
}

hipError_t sg_launch_pair_mfma_segs(const uint64_t* rows, const SgSegment* segs, int32_t nseg,
                                    int32_t max_dpad, int64_t max_words, int32_t* pair, hipStream_t s)
{
    if (nseg <= 0) return hipSuccess;
    const int tiles = max_dpad / 128;
    int32_t splits = (int32_t)max_words / 64;
    if (splits < 1) splits = 1;
    if (splits > 64) splits = 64;
    dim3 grid((unsigned)(tiles * tiles), (unsigned)splits, (unsigned)nseg);
    hipLaunchKernelGGL(k_pair_mfma_segs, grid, dim3(256), 0, s, rows, segs, pair, splits);
    return hipGetLastError();
}

hipError_t sg_launch_pair_valu(const uint64_t* rows, int32_t dpad, int64_t Wp, int64_t W,
                               int32_t* pair, hipStream_t s)
{
    const int64_t pairs = (int64_t)dpad * dpad;
    hipLaunchKernelGGL(k_pair_valu, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, s, rows, Wp, W, dpad, pair);
    return hipGetLastError();
}

hipError_t sg_launch_pair_ok(const int32_t* pair, const SgSegment* segs, int32_t nseg,
                             int32_t max_d, uint64_t* ok, hipStream_t s)
{
    if (nseg <= 0 || max_d <= 0) return hipSuccess;
    const int64_t per = (int64_t)max_d * ((max_d + 63) / 64);
    dim3 grid((unsigned)((per + 255) / 256), (unsigned)nseg);
    hipLaunchKernelGGL(k_pair_ok, grid, dim3(256), 0, s, pair, segs, max_d, ok);
    return hipGetLastError();
}
