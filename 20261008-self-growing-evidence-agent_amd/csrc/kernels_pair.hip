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
// Tiling: a workgroup of 4 waves owns a 256 x 256 output tile and each wave a 128 x
// 128 quarter: 4 x 4 MFMA tiles whose 16 accumulators (256 int32 per lane) stay in the
// accumulation registers, so every widened 16-bit fragment feeds four instructions.
// The bit tiles of a K step (8 words = 512 transactions of 256 rows of A and of B) are
// loaded with coalesced 64-byte row segments into LDS, double-buffered through
// registers (the next step's loads are in flight while the current one multiplies).
// Only tiles on or above the diagonal are computed; a mirror pass fills the lower
// triangle. grid.y splits the transactions (split-N) and partial tiles are added with
// integer atomics, which keeps the result bit-exact.
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

#define PT_TILE 256   // output rows / columns of a workgroup tile
#define PT_K 8        // words per K step
#define PT_LDS 20     // LDS row stride in dwords (16 data + 4 pad; 16-byte aligned rows)

// Tile pair (ti, tj), ti <= tj, of the upper triangle of a T x T tile grid.
__device__ __forceinline__ void sg_tri_pair(int q, int T, int& ti, int& tj)
{
    int i = 0;
    while (q >= T - i) { q -= T - i; i++; }
    ti = i;
    tj = i + q;
}

// C[I0.., J0..] += A^T B over words [wb, we) of one segment (bit rows, 0/1 widened in
// registers). SEGS: the segment is segs[blockIdx.z], else `one`.
template <bool SEGS>
__global__ __launch_bounds__(256) void k_pair_tile(const uint64_t* __restrict__ rows_all,
                                                   const SgSegment* __restrict__ segs, SgSegment one,
                                                   int32_t* __restrict__ out_all, int32_t splits)
{
    __shared__ __attribute__((aligned(16))) uint32_t As[PT_TILE * PT_LDS];
    __shared__ __attribute__((aligned(16))) uint32_t Bs[PT_TILE * PT_LDS];
    const SgSegment sg = SEGS ? segs[blockIdx.z] : one;
    const int dpad = sg.dpad;
    const int T = (dpad + PT_TILE - 1) / PT_TILE;
    if ((int)blockIdx.x >= T * (T + 1) / 2) return;
    int ti, tj;
    sg_tri_pair((int)blockIdx.x, T, ti, tj);
    const int I0 = ti * PT_TILE, J0 = tj * PT_TILE;
    const int64_t per = ((sg.words + splits - 1) / splits + PT_K - 1) / PT_K * PT_K;
    const int64_t wb = (int64_t)blockIdx.y * per;
    const int64_t we = min(sg.words, wb + per);
    if (wb >= we) return;
    const uint64_t* rows = rows_all + sg.row_base;
    int32_t* out = out_all + sg.pair_base;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 31, h = lane >> 5;
    const int wr = wave >> 1, wc = wave & 1;
    // loader: thread t moves word (t & 7) of rows (t >> 3) + 32 i, i < 8, of A and B
    const int lq = tid & 7, lr = tid >> 3;
    uint64_t ga[8], gb[8];
    auto gload = [&](int64_t w0) {
        const bool wok = w0 + lq < we;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int row = lr + 32 * i;
            ga[i] = (wok && I0 + row < dpad) ? rows[(int64_t)(I0 + row) * sg.stride + w0 + lq] : 0ull;
            gb[i] = (wok && J0 + row < dpad) ? rows[(int64_t)(J0 + row) * sg.stride + w0 + lq] : 0ull;
        }
    };
    v16i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = v16i{};
    gload(wb);
    for (int64_t w0 = wb; w0 < we; w0 += PT_K) {
        __syncthreads();   // the previous step has finished reading the tiles
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int row = lr + 32 * i;
            As[row * PT_LDS + 2 * lq] = (uint32_t)ga[i];
            As[row * PT_LDS + 2 * lq + 1] = (uint32_t)(ga[i] >> 32);
            Bs[row * PT_LDS + 2 * lq] = (uint32_t)gb[i];
            Bs[row * PT_LDS + 2 * lq + 1] = (uint32_t)(gb[i] >> 32);
        }
        __syncthreads();
        if (w0 + PT_K < we) gload(w0 + PT_K);
#pragma unroll 1
        for (int g4 = 0; g4 < 4; g4++) {   // four dwords = four MFMA K steps
            uint4 av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = *(const uint4*)&As[(128 * wr + 32 * a + r) * PT_LDS + 4 * g4];
#pragma unroll
            for (int b = 0; b < 4; b++) bv[b] = *(const uint4*)&Bs[(128 * wc + 32 * b + r) * PT_LDS + 4 * g4];
#pragma unroll
            for (int dd = 0; dd < 4; dd++) {
                v4i A[4], B[4];
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    const uint32_t x = dd == 0 ? av[a].x : dd == 1 ? av[a].y : dd == 2 ? av[a].z : av[a].w;
                    A[a] = sg_widen16((x >> (16 * h)) & 0xFFFFu);
                }
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const uint32_t x = dd == 0 ? bv[b].x : dd == 1 ? bv[b].y : dd == 2 ? bv[b].z : bv[b].w;
                    B[b] = sg_widen16((x >> (16 * h)) & 0xFFFFu);
                }
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++) acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[a], B[b], acc[a][b], 0, 0, 0);
            }
        }
    }
    // 32 x 32 int32 accumulator layout: lane l, element v holds
    // row 8*(v/4) + 4*(l/32) + v%4, column l%32.
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = I0 + 128 * wr + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                const int col = J0 + 128 * wc + 32 * b + r;
                const int32_t x = acc[a][b][v];
                if (x && row < dpad && col < dpad) atomicAdd(&out[(int64_t)row * dpad + col], x);
            }
}

// Lower triangle from the upper one: C[i][j] = C[j][i] for i > j.
template <bool SEGS>
__global__ void k_pair_mirror(const SgSegment* __restrict__ segs, SgSegment one, int32_t* __restrict__ out_all)
{
    const SgSegment sg = SEGS ? segs[blockIdx.y] : one;
    int32_t* out = out_all + sg.pair_base;
    const int64_t n = (int64_t)sg.dpad * sg.dpad;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(p / sg.dpad), j = (int)(p % sg.dpad);
        if (i > j) out[p] = out[(int64_t)j * sg.dpad + i];
    }
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

// ok[i][w] bit b = pair(i, 64w+b) >= thr, for 64w+b < d. One wave per word: lane b
// tests column 64w+b (one coalesced 256-byte read) and the ballot is the word.
__global__ __launch_bounds__(256) void k_pair_ok(const int32_t* __restrict__ pair, const SgSegment* __restrict__ segs,
                                                 int32_t max_d, uint64_t* __restrict__ ok)
{
    const SgSegment sg = segs[blockIdx.y];
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;   // the same on every lane
    if (q >= (int64_t)sg.d * sg.dw) return;
    const int lane = threadIdx.x & 63;
    const int i = (int)(q / sg.dw), w = (int)(q % sg.dw);
    const int j = w * 64 + lane;
    const bool hit = j < sg.d && pair[sg.pair_base + (int64_t)i * sg.dpad + j] >= sg.thr;
    const uint64_t v = __ballot(hit);
    if (lane == 0) ok[sg.ok_base + (int64_t)i * sg.dw + w] = v;
}

// Transaction splits so that about two workgroups per CU (512) are in flight.
static int32_t pair_splits(int32_t tiles, int64_t words)
{
    int64_t splits = (512 + tiles - 1) / tiles;
    const int64_t steps = (words + PT_K - 1) / PT_K;   // at least one K step per split
    if (splits > steps) splits = steps > 0 ? steps : 1;
    return (int32_t)(splits < 1 ? 1 : splits);
}

// Zero n int32s (16-byte stores, grid-stride): the split-N tiles add into the matrix. A
// kernel launch costs the host less than a memset call, and it may sit in a captured graph.
__global__ __launch_bounds__(256) void k_zero32(int32_t* __restrict__ p, int64_t n)
{
    const int64_t n4 = n / 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        ((int4*)p)[i] = int4{0, 0, 0, 0};
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

hipError_t sg_launch_zero32(int32_t* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(1024, (n / 4 + 255) / 256 + 1);
    hipLaunchKernelGGL(k_zero32, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t sg_launch_pair_mfma(const uint64_t* rows, int32_t dpad, int64_t Wp, int64_t W,
                               int32_t* pair, hipStream_t s)
{
    hipError_t e = sg_launch_zero32(pair, (int64_t)dpad * dpad, s);
    if (e != hipSuccess || W <= 0) return e;
    SgSegment one{};
    one.row_base = 0;
    one.stride = Wp;
    one.words = W;
    one.dpad = dpad;
    one.pair_base = 0;
    const int T = (dpad + PT_TILE - 1) / PT_TILE;
    const int tiles = T * (T + 1) / 2;
    const int32_t splits = pair_splits(tiles, W);
    hipLaunchKernelGGL(k_pair_tile<false>, dim3((unsigned)tiles, (unsigned)splits, 1), dim3(256), 0, s, rows,
                       (const SgSegment*)nullptr, one, pair, splits);
    hipLaunchKernelGGL(k_pair_mirror<false>, dim3(256, 1), dim3(256), 0, s, (const SgSegment*)nullptr, one, pair);
    return hipGetLastError();
}

// Windows: every segment's matrix in one launch; `pair` must be zeroed.
hipError_t sg_launch_pair_mfma_segs(const uint64_t* rows, const SgSegment* segs, int32_t nseg,
                                    int32_t max_dpad, int64_t max_words, int32_t* pair, hipStream_t s)
{
    if (nseg <= 0) return hipSuccess;
    const int T = (max_dpad + PT_TILE - 1) / PT_TILE;
    const int tiles = T * (T + 1) / 2;
    int64_t splits = (512 + (int64_t)tiles * nseg - 1) / ((int64_t)tiles * nseg);
    const int64_t steps = (max_words + PT_K - 1) / PT_K;
    if (splits > steps) splits = steps;
    if (splits < 1) splits = 1;
    SgSegment none{};
    hipLaunchKernelGGL(k_pair_tile<true>, dim3((unsigned)tiles, (unsigned)splits, (unsigned)nseg), dim3(256), 0, s,
                       rows, segs, none, pair, (int32_t)splits);
    hipLaunchKernelGGL(k_pair_mirror<true>, dim3(64, (unsigned)nseg), dim3(256), 0, s, segs, none, pair);
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
    const int64_t per = (int64_t)max_d * ((max_d + 63) / 64);   // words; a wave each
    dim3 grid((unsigned)((per * 64 + 255) / 256), (unsigned)nseg);
    hipLaunchKernelGGL(k_pair_ok, grid, dim3(256), 0, s, pair, segs, max_d, ok);
    return hipGetLastError();
}
