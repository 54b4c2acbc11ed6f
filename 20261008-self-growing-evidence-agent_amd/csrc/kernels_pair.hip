// Pair co-occurrence matrix C[i][j] = |T_i & T_j| with MFMA on 0/1 tiles: int8 below,
// fp4 (the default) further down.
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
// Tiling: a workgroup of 4 waves owns a TT x TT output tile and each wave a TT/2 x TT/2
// quarter of (TT/64) x (TT/64) MFMA tiles whose accumulators stay in the accumulation
// registers, so every widened 16-bit fragment feeds TT/64 instructions. TT = 128 (2 x 2
// tiles, 64 int32 accumulators per lane, 144 VGPRs, 20 KB of LDS: three workgroups per
// CU) runs 184 us on 1M transactions x 512 items where TT = 256 (4 x 4 tiles, one
// workgroup per CU) ran 204 us. The bit tiles of a K step (8 words = 512 transactions of
// TT rows of A and of B) are loaded with coalesced 64-byte row segments into LDS,
// double-buffered through registers (the next step's loads are in flight while the
// current one multiplies).
// Only tiles on or above the diagonal are computed; a mirror pass fills the lower
// triangle. grid.y splits the transactions (split-N) and partial tiles are added with
// integer atomics, which keeps the result bit-exact. Workgroups are renumbered XCD-major
// (sg_xcd_major), as in the fp4 kernel below.
#include "sgfim_impl.h"
#include "f4_phase.h"

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

#define PT_TILE 128   // output rows / columns of a workgroup tile (the kernel also takes 256)
#define PT_SLOTS 768  // workgroups resident at once, three per CU (the grid fills whole rounds)
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

// Workgroup q of a one-segment grid of n: consecutive workgroup ids go to the eight
// XCDs in turn, so q is renumbered XCD-major; the tiles of one transaction split then
// share an XCD and its L2 (they read the same rows).
__device__ __forceinline__ int sg_xcd_major(int q, int n)
{
    const int x = q & 7, k = q >> 3;
    return x * (n >> 3) + min(x, n & 7) + k;
}

// C[I0.., J0..] += A^T B over words [wb, we) of one segment (bit rows, 0/1 widened in
// registers). SEGS: the segment is segs[blockIdx.z], else `one`.
template <bool SEGS, int TT>
__global__ __launch_bounds__(256) void k_pair_tile(const uint64_t* __restrict__ rows_all,
                                                   const SgSegment* __restrict__ segs, SgSegment one,
                                                   int32_t* __restrict__ out_all, int32_t splits)
{
    constexpr int NA = TT / 64;   // MFMA tiles per wave along each side (a wave owns TT/2 x TT/2)
    constexpr int LP = TT / 32;   // loader passes (32 rows each)
    __shared__ __attribute__((aligned(16))) uint32_t As[TT * PT_LDS];
    __shared__ __attribute__((aligned(16))) uint32_t Bs[TT * PT_LDS];
    const SgSegment sg = SEGS ? segs[blockIdx.z] : one;
    const int dpad = sg.dpad;
    const int T = (dpad + TT - 1) / TT;
    int bx = (int)blockIdx.x, by = (int)blockIdx.y;
    if (!SEGS) {
        const int q = sg_xcd_major(bx + (int)gridDim.x * by, (int)(gridDim.x * gridDim.y));
        bx = q % (int)gridDim.x;
        by = q / (int)gridDim.x;
    }
    if (bx >= T * (T + 1) / 2) return;
    int ti, tj;
    sg_tri_pair(bx, T, ti, tj);
    const int I0 = ti * TT, J0 = tj * TT;
    const int64_t per = ((sg.words + splits - 1) / splits + PT_K - 1) / PT_K * PT_K;
    const int64_t wb = (int64_t)by * per;
    const int64_t we = min(sg.words, wb + per);
    if (wb >= we) return;
    const uint64_t* rows = rows_all + sg.row_base;
    int32_t* out = out_all + sg.pair_base;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 31, h = lane >> 5;
    const int wr = wave >> 1, wc = wave & 1;
    // loader: thread t moves word (t & 7) of rows (t >> 3) + 32 i, i < LP, of A and B
    const int lq = tid & 7, lr = tid >> 3;
    uint64_t ga[LP], gb[LP];
    auto gload = [&](int64_t w0) {
        const bool wok = w0 + lq < we;
#pragma unroll
        for (int i = 0; i < LP; i++) {
            const int row = lr + 32 * i;
            ga[i] = (wok && I0 + row < dpad) ? rows[(int64_t)(I0 + row) * sg.stride + w0 + lq] : 0ull;
            gb[i] = (wok && J0 + row < dpad) ? rows[(int64_t)(J0 + row) * sg.stride + w0 + lq] : 0ull;
        }
    };
    v16i acc[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++) acc[a][b] = v16i{};
    gload(wb);
    for (int64_t w0 = wb; w0 < we; w0 += PT_K) {
        __syncthreads();   // the previous step has finished reading the tiles
#pragma unroll
        for (int i = 0; i < LP; i++) {
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
            uint4 av[NA], bv[NA];
#pragma unroll
            for (int a = 0; a < NA; a++) av[a] = *(const uint4*)&As[((TT / 2) * wr + 32 * a + r) * PT_LDS + 4 * g4];
#pragma unroll
            for (int b = 0; b < NA; b++) bv[b] = *(const uint4*)&Bs[((TT / 2) * wc + 32 * b + r) * PT_LDS + 4 * g4];
#pragma unroll
            for (int dd = 0; dd < 4; dd++) {
                v4i A[NA], B[NA];
#pragma unroll
                for (int a = 0; a < NA; a++) {
                    const uint32_t x = dd == 0 ? av[a].x : dd == 1 ? av[a].y : dd == 2 ? av[a].z : av[a].w;
                    A[a] = sg_widen16((x >> (16 * h)) & 0xFFFFu);
                }
#pragma unroll
                for (int b = 0; b < NA; b++) {
                    const uint32_t x = dd == 0 ? bv[b].x : dd == 1 ? bv[b].y : dd == 2 ? bv[b].z : bv[b].w;
                    B[b] = sg_widen16((x >> (16 * h)) & 0xFFFFu);
                }
#pragma unroll
                for (int a = 0; a < NA; a++)
#pragma unroll
                    for (int b = 0; b < NA; b++) acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[a], B[b], acc[a][b], 0, 0, 0);
            }
        }
    }
    // 32 x 32 int32 accumulator layout: lane l, element v holds
    // row 8*(v/4) + 4*(l/32) + v%4, column l%32.
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = I0 + (TT / 2) * wr + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                const int col = J0 + (TT / 2) * wc + 32 * b + r;
                const int32_t x = acc[a][b][v];
                if (x && row < dpad && col < dpad) atomicAdd(&out[(int64_t)row * dpad + col], x);
            }
}

// ------------------------------------------------------------------ fp4 MFMA tile
//
// The same matrix on v_mfma_scale_f32_32x32x64_f8f6f4 with e2m1 (fp4) operands, which
// the matrix cores run at twice the int8 rate (64 transactions per K step in the 32
// cycles the int8 instruction needs for 32). Every product is 0 or 1 and the fp32
// accumulators hold integers. A split spans at most PF_MAXW words = 2^20 transactions,
// so every partial sum is an integer of at most 21 bits, exact in fp32 (24-bit
// significand) with room to spare; it is converted to int32 before the atomics add the
// splits.
//
// Widening by bit phase (f4_phase.h): K step p (0..3) of a group of four bitmap words
// takes, for each word, the transactions at bit p of every nibble, as exact fp4 values
// whose scaled products are 1. One VALU instruction (two for phase 3) widens 8
// transactions, and lane (r, h) of K step p holds, in VGPR j, half h of word j's phase-p
// nibbles of row r.
//
// Tiling: a workgroup of 8 waves owns a TT x TT output tile, each wave TT/4 x TT/2 as
// (TT/128) x (TT/64) MFMA tiles. TT = 128 (1 x 2 tiles, 32 fp32 accumulators per lane,
// 100 VGPRs and 40 KB of LDS: two workgroups per CU, four waves per SIMD) runs 91 us on
// 1M transactions x 512 items where TT = 256 (2 x 4 tiles, 128 accumulators, one workgroup
// per CU) ran 117 us: a quarter of the split-N atomics per output element and twice the
// resident waves outweigh the operands' halved reuse. Bit tiles of PK words are
// staged in LDS with rows laid out [half][word], so a lane reads four K steps of its
// half with one 16-byte read; they are double-buffered (one barrier per stage: the next
// stage's tile is stored into the other buffer while this one is multiplied), and each
// four-word group's reads are issued before the previous group's matrix instructions.
// Upper triangle only, split-N partial tiles added with integer atomics.

#define PF_K 8                // words per stage (two groups of four words)
#define PF_TILE 128           // output rows / columns of a workgroup tile (the kernel also takes 256)
#define PF_SLOTS 512          // workgroups resident at once, two per CU (the grid fills whole rounds)
#define PF_MAXW (1ll << 14)   // words per split (2^20 transactions; a multiple of PF_K)

template <bool SEGS, int PK, int TT>
__global__ __launch_bounds__(512) void k_pair_f4(const uint64_t* __restrict__ rows_all,
                                                  const SgSegment* __restrict__ segs, SgSegment one,
                                                  int32_t* __restrict__ out_all, int32_t splits)
{
    constexpr int NT = 512;
    constexpr int MA = TT / 128;          // A fragments (32 rows each) per wave
    constexpr int NB = TT / 64;           // B fragments (32 columns each) per wave
    constexpr int LDW = 2 * PK + 4;       // LDS row stride in dwords: [half][word] + pad
    constexpr int LQ = PK;                // loader threads per row
    constexpr int LR = NT / LQ;           // rows per loader pass
    constexpr int LP = TT / LR;           // loader passes per matrix
    constexpr int G = PK / 4;             // four-word groups per stage
    __shared__ __attribute__((aligned(16))) uint32_t As[2][TT * LDW];
    __shared__ __attribute__((aligned(16))) uint32_t Bs[2][TT * LDW];
    const SgSegment sg = SEGS ? segs[blockIdx.z] : one;
    const int dpad = sg.dpad;
    const int T = (dpad + TT - 1) / TT;
    int bx = (int)blockIdx.x, by = (int)blockIdx.y;
    if (!SEGS) {
        const int q = sg_xcd_major(bx + (int)gridDim.x * by, (int)(gridDim.x * gridDim.y));
        bx = q % (int)gridDim.x;
        by = q / (int)gridDim.x;
    }
    if (bx >= T * (T + 1) / 2) return;
    int ti, tj;
    sg_tri_pair(bx, T, ti, tj);
    const int I0 = ti * TT, J0 = tj * TT;
    const int64_t per = ((sg.words + splits - 1) / splits + PK - 1) / PK * PK;
    const int64_t wb = (int64_t)by * per;
    const int64_t we = min(sg.words, wb + per);
    if (wb >= we) return;
    const uint64_t* rows = rows_all + sg.row_base;
    int32_t* out = out_all + sg.pair_base;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 31, h = lane >> 5;
    const int wr = wave >> 1, wc = wave & 1;
    // loader: thread t moves word (t % PK) of rows t / PK + LR i, i < LP, of A and B. Rows
    // past dpad (a partial last tile) are clamped to row dpad - 1: they only feed output
    // rows and columns that are never stored. A row's words are zero from the segment's
    // last word up to its stride (a multiple of 64), and a split's end is a multiple of PK
    // otherwise, so the loads need no guard.
    const int lq = tid % LQ, lr = tid / LQ;
    const uint64_t* pa[LP];
    const uint64_t* pb[LP];
#pragma unroll
    for (int i = 0; i < LP; i++) {
        pa[i] = rows + (int64_t)min(I0 + lr + LR * i, dpad - 1) * sg.stride + lq;
        pb[i] = rows + (int64_t)min(J0 + lr + LR * i, dpad - 1) * sg.stride + lq;
    }
    uint64_t ga[LP], gb[LP];
    auto gload = [&](int64_t w0) {
#pragma unroll
        for (int i = 0; i < LP; i++) {
            ga[i] = pa[i][w0];
            gb[i] = pb[i][w0];
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LP; i++) {
            const int row = lr + LR * i;
            As[buf][row * LDW + lq] = (uint32_t)ga[i];
            As[buf][row * LDW + PK + lq] = (uint32_t)(ga[i] >> 32);
            Bs[buf][row * LDW + lq] = (uint32_t)gb[i];
            Bs[buf][row * LDW + PK + lq] = (uint32_t)(gb[i] >> 32);
        }
    };
    uint4 av[2][MA], bv[2][NB];
    auto lread = [&](int buf, int g, int slot) {
#pragma unroll
        for (int a = 0; a < MA; a++)
            av[slot][a] = *(const uint4*)&As[buf][(32 * MA * wr + 32 * a + r) * LDW + PK * h + 4 * g];
#pragma unroll
        for (int b = 0; b < NB; b++)
            bv[slot][b] = *(const uint4*)&Bs[buf][((TT / 2) * wc + 32 * b + r) * LDW + PK * h + 4 * g];
    };
    v16f acc[MA][NB];
#pragma unroll
    for (int a = 0; a < MA; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = v16f{};
    auto mma = [&](int slot) {
#pragma unroll
        for (int dd = 0; dd < 4; dd++) {
            v8i A[MA], B[NB];
#pragma unroll
            for (int a = 0; a < MA; a++) A[a] = sg_f4_mask_frag(av[slot][a], dd);
#pragma unroll
            for (int b = 0; b < NB; b++) B[b] = sg_f4_mask_frag(bv[slot][b], dd);
            const int sc = dd == 0 ? sg_f4_scale<0>() : dd == 1 ? sg_f4_scale<1>() : sg_f4_scale<2>();
#pragma unroll
            for (int a = 0; a < MA; a++)
#pragma unroll
                for (int b = 0; b < NB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A[a], B[b], acc[a][b], 4, 4, 0, sc,
                                                                                 0, sc);
        }
    };
    gload(wb);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int64_t w0 = wb; w0 < we; w0 += PK) {
        const bool more = w0 + PK < we;
        if (more) gload(w0 + PK);
        lread(buf, 0, 0);
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (g + 1 < G) lread(buf, g + 1, (g + 1) & 1);
            mma(g & 1);
        }
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int a = 0; a < MA; a++)
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = I0 + 32 * MA * wr + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                const int col = J0 + (TT / 2) * wc + 32 * b + r;
                const int32_t x = (int32_t)acc[a][b][v];
                if (x && row < dpad && col < dpad) atomicAdd(&out[(int64_t)row * dpad + col], x);
            }
}

// ------------------------------------------------------------------ MFMA rate probe
// Matrix-core throughput with no memory traffic: every wave issues `iters` x 8
// independent 32 x 32 MFMAs (fp4 v_mfma_scale_f32_32x32x64_f8f6f4 or int8
// v_mfma_i32_32x32x32_i8) on register operands; one lane per workgroup stores a sum so
// the work is kept. What the pair kernels are measured against beside the published peak.
template <bool F4>
__global__ __launch_bounds__(256) void k_mfma_probe(float* __restrict__ out, int32_t iters, int32_t seed)
{
    v8i a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (int)(((uint32_t)seed * (threadIdx.x + i)) & 0x22222222u);
        b[i] = (int)(((uint32_t)seed ^ (threadIdx.x * 7u + i)) & 0x22222222u);
    }
    v16f acc[8];
    v16i iacc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        acc[j] = v16f{};
        iacc[j] = v16i{};
    }
    const v4i a4{a[0], a[1], a[2], a[3]}, b4{b[0], b[1], b[2], b[3]};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (F4) acc[j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[j], 4, 4, 0, 127, 0, 127);
            else iacc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a4, b4, iacc[j], 0, 0, 0);
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++)
#pragma unroll
        for (int v = 0; v < 16; v++) sum += F4 ? acc[j][v] : (float)iacc[j][v];
    if (threadIdx.x == 0) out[blockIdx.x] = sum;
}

hipError_t sg_launch_mfma_probe(float* out, int32_t blocks, int32_t iters, int32_t kind, hipStream_t s)
{
    if (kind == SG_PAIR_FP4)
        hipLaunchKernelGGL(k_mfma_probe<true>, dim3((unsigned)blocks), dim3(256), 0, s, out, iters, 3);
    else
        hipLaunchKernelGGL(k_mfma_probe<false>, dim3((unsigned)blocks), dim3(256), 0, s, out, iters, 3);
    return hipGetLastError();
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

// Transaction splits of `units` tile grids (tiles x segments). The grid is sized in whole
// dispatch rounds of the workgroups that are resident at once (three per CU for the
// int8 kernel, two for the fp4 one): a count just above a multiple of a round would
// leave a last round of a few workgroups with the chip waiting on them. The fp4 kernel's
// splits also stay within PF_MAXW words, where its fp32 sums are exact. `cap` > 0 limits
// the splits (never below what exactness needs).
static int32_t pair_splits(int64_t units, int64_t words, bool f4, int32_t cap)
{
    const int64_t slots = f4 ? PF_SLOTS : PT_SLOTS;   // workgroups resident at once
    const int64_t steps = std::max<int64_t>(1, (words + PT_K - 1) / PT_K);   // one K step per split at least
    const int64_t need = f4 ? std::max<int64_t>(1, (words + PF_MAXW - 1) / PF_MAXW) : 1;
    int64_t splits = std::max<int64_t>(need, slots / std::max<int64_t>(units, 1));
    if (units * splits > slots) {   // more than one round: fill the last one
        const int64_t rounds = (units * splits + slots - 1) / slots;
        splits = std::max(splits, rounds * slots / units);
    }
    if (cap > 0) splits = std::min<int64_t>(splits, std::max<int64_t>(cap, need));
    splits = std::min(splits, steps);
    return (int32_t)std::max<int64_t>(splits, 1);
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
                               int32_t* pair, int32_t kind, int32_t split_cap, hipStream_t s)
{
    hipError_t e = sg_launch_zero32(pair, (int64_t)dpad * dpad, s);
    if (e != hipSuccess || W <= 0) return e;
    SgSegment one{};
    one.row_base = 0;
    one.stride = Wp;
    one.words = W;
    one.dpad = dpad;
    one.pair_base = 0;
    const bool f4 = kind == SG_PAIR_FP4;
    const int tt = f4 ? PF_TILE : PT_TILE;
    const int T = (dpad + tt - 1) / tt;
    const int tiles = T * (T + 1) / 2;
    const int32_t splits = pair_splits(tiles, W, f4, split_cap);
    const dim3 grid((unsigned)tiles, (unsigned)splits, 1);
    if (f4)
        hipLaunchKernelGGL((k_pair_f4<false, PF_K, PF_TILE>), grid, dim3(512), 0, s, rows, (const SgSegment*)nullptr, one,
                           pair, splits);
    else
        hipLaunchKernelGGL((k_pair_tile<false, PT_TILE>), grid, dim3(256), 0, s, rows, (const SgSegment*)nullptr, one, pair,
                           splits);
    hipLaunchKernelGGL(k_pair_mirror<false>, dim3(256, 1), dim3(256), 0, s, (const SgSegment*)nullptr, one, pair);
    return hipGetLastError();
}

// Windows: every segment's matrix in one launch; `pair` must be zeroed.
hipError_t sg_launch_pair_mfma_segs(const uint64_t* rows, const SgSegment* segs, int32_t nseg,
                                    int32_t max_dpad, int64_t max_words, int32_t* pair, int32_t kind,
                                    int32_t split_cap, hipStream_t s)
{
    if (nseg <= 0) return hipSuccess;
    const bool f4 = kind == SG_PAIR_FP4;
    const int tt = f4 ? PF_TILE : PT_TILE;
    const int T = (max_dpad + tt - 1) / tt;
    const int tiles = T * (T + 1) / 2;
    const int32_t splits = pair_splits((int64_t)tiles * nseg, max_words, f4, split_cap);
    SgSegment none{};
    if (f4)
        hipLaunchKernelGGL((k_pair_f4<true, PF_K, PF_TILE>), dim3((unsigned)tiles, (unsigned)splits, (unsigned)nseg),
                           dim3(512), 0, s, rows, segs, none, pair, (int32_t)splits);
    else
        hipLaunchKernelGGL((k_pair_tile<true, PT_TILE>), dim3((unsigned)tiles, (unsigned)splits, (unsigned)nseg), dim3(256), 0,
                           s, rows, segs, none, pair, (int32_t)splits);
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
