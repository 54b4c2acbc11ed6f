// Proposal network for the anytime frontier: 24 features -> 256 hidden (ReLU) -> 1 logit.
//
// The reference has no learned component. This network only reorders the search
// frontier; every support is still counted exactly and the final MFI set does not
// depend on it. Weights are BF16, accumulation FP32, on v_mfma_f32_32x32x16_bf16:
//   forward  H = relu(X W1 + b1) (M = 32 nodes, N = 256 hidden, K = 24 padded to 32)
//            z = H . w2 + b2 (VALU dot + lane reduction), score = sigmoid(z)
//   backward dz = (sigmoid(z) - y) / B, dw2 = H^T dz, db2 = sum dz,
//            dH = dz w2^T * [H > 0], db1 = sum dH, dW1 = X^T dH (MFMA, K = nodes)
// Training runs in one workgroup so the gradient sums have a fixed order (bit-
// reproducible steps); Adam keeps FP32 master weights and writes BF16 copies.
//
// BF16 rounding is round-to-nearest-even of the FP32 value (finite inputs).
// sigmoid uses the device library's __expf.
#include "sgfim_impl.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t sg_bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float sg_bf16_float(uint16_t b)
{
    return __uint_as_float(((uint32_t)b) << 16);
}

__device__ __forceinline__ v8bf sg_pack8(const uint16_t* v)
{
    v8s s;
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] = (short)v[j];
    return __builtin_bit_cast(v8bf, s);
}

__device__ __forceinline__ float sg_sigmoid(float z)
{
    return 1.0f / (1.0f + __expf(-z));
}

// Load the A fragments of a 32-node tile: lane l holds node r0 + l%32, features
// 8h..8h+7 (first K half) and 16+8h..16+8h+7 (second K half), h = l/32.
__device__ __forceinline__ void sg_load_x(const float* __restrict__ X, int64_t n, int64_t r0,
                                          v8bf& a0, v8bf& a1)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = r0 + (lane & 31);
    const int h = lane >> 5;
    uint16_t u0[8], u1[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k0 = 8 * h + j, k1 = 16 + 8 * h + j;
        u0[j] = row < n ? sg_bf16_bits(X[row * SG_FEAT + k0]) : 0;
        u1[j] = (row < n && k1 < SG_FEAT) ? sg_bf16_bits(X[row * SG_FEAT + k1]) : 0;
    }
    a0 = sg_pack8(u0);
    a1 = sg_pack8(u1);
}

// As sg_load_x, with row r of the tile read from X[idx[r]] (idx may be null).
__device__ __forceinline__ void sg_load_x_idx(const float* __restrict__ X, const int32_t* __restrict__ idx,
                                              int64_t n, int64_t r0, v8bf& a0, v8bf& a1)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = r0 + (lane & 31);
    const int h = lane >> 5;
    const int64_t src = row < n ? (idx ? (int64_t)idx[row] : row) : 0;
    uint16_t u0[8], u1[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k0 = 8 * h + j, k1 = 16 + 8 * h + j;
        u0[j] = row < n ? sg_bf16_bits(X[src * SG_FEAT + k0]) : 0;
        u1[j] = (row < n && k1 < SG_FEAT) ? sg_bf16_bits(X[src * SG_FEAT + k1]) : 0;
    }
    a0 = sg_pack8(u0);
    a1 = sg_pack8(u1);
}

// Pre-activation tile nt (hidden units 32nt..32nt+31) of the 32-node tile.
// w1t is BF16 [SG_HID][32] (hidden-major, K padded with zeros).
__device__ __forceinline__ v16f sg_hidden_tile(const uint16_t* __restrict__ w1t, int nt, v8bf a0, v8bf a1)
{
    const int lane = threadIdx.x & 63;
    const int col = nt * 32 + (lane & 31);
    const int h = lane >> 5;
    const uint16_t* wr = w1t + col * 32;
    const v8bf b0 = sg_pack8(wr + 8 * h);
    const v8bf b1 = sg_pack8(wr + 16 + 8 * h);
    v16f acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    return acc;
}

// Row index of accumulator element v for this lane: 8*(v/4) + 4*(lane/32) + v%4.
__device__ __forceinline__ int sg_acc_row(int v) { return 8 * (v >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (v & 3); }

// z for the 16 rows this lane's half-wave holds (index v), after the 32-lane reduction.
__device__ __forceinline__ void sg_logits(const uint16_t* __restrict__ w1t, const float* __restrict__ b1,
                                          const uint16_t* __restrict__ w2, float b2, v8bf a0, v8bf a1,
                                          float* z)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 16; v++) z[v] = 0.0f;
    for (int nt = 0; nt < SG_HID / 32; nt++) {
        const v16f acc = sg_hidden_tile(w1t, nt, a0, a1);
        const int col = nt * 32 + (lane & 31);
        const float bb = b1[col], ww = sg_bf16_float(w2[col]);
#pragma unroll
        for (int v = 0; v < 16; v++) z[v] += fmaxf(acc[v] + bb, 0.0f) * ww;
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
        float t = z[v];
        for (int o = 16; o > 0; o >>= 1) t += __shfl_xor(t, o, 32);
        z[v] = t + b2;
    }
}

// ------------------------------------------------------------------ inference
__global__ __launch_bounds__(256) void k_mlp_score(const uint16_t* __restrict__ w1t, const float* __restrict__ b1,
                                                   const uint16_t* __restrict__ w2, const float* __restrict__ b2,
                                                   const float* __restrict__ X, int32_t n_max,
                                                   const int64_t* __restrict__ n_dev, float* __restrict__ out,
                                                   int32_t logits)
{
    const int64_t n = n_dev ? min((int64_t)n_max, *n_dev) : (int64_t)n_max;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t r0 = ((int64_t)blockIdx.x * 4 + wv) * 32; r0 < n; r0 += (int64_t)gridDim.x * 128) {
        v8bf a0, a1;
        sg_load_x(X, n, r0, a0, a1);
        float z[16];
        sg_logits(w1t, b1, w2, b2[0], a0, a1, z);
        if ((lane & 31) == 0) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int64_t row = r0 + sg_acc_row(v);
                if (row < n) out[row] = logits ? z[v] : sg_sigmoid(z[v]);
            }
        }
    }
}

hipError_t sg_launch_score(const SgScorer* sc, const float* feat, int32_t n_max,
                           const int64_t* n_dev, float* out, hipStream_t s)
{
    if (!sc || !sc->w1) return hipErrorInvalidValue;
    // the row count is read on the device (a round's newborns); a modest grid loops over
    // the tiles, so a round with few newborns does not pay for thousands of idle groups
    int64_t blocks = ((int64_t)n_max + 127) / 128;
    if (blocks > 256) blocks = 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_score, dim3((unsigned)blocks), dim3(256), 0, s, sc->w1, sc->b1, sc->w2, sc->b2,
                       feat, n_max, n_dev, out, 0);
    return hipGetLastError();
}

// ------------------------------------------------------------------ training step

// Write BF16 copies of the master weights.
__global__ void k_mlp_pack(SgTrainState T)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < SG_HID * 32; i += blockDim.x * gridDim.x) {
        const int col = i / 32, k = i % 32;
        T.w1t[i] = k < SG_FEAT ? sg_bf16_bits(T.W1[k * SG_HID + col]) : 0;
    }
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < SG_HID; i += blockDim.x * gridDim.x)
        T.w2b[i] = sg_bf16_bits(T.w2[i]);
}

// One fused step on a minibatch: forward, BCE loss, backward, and (if lr > 0) Adam.
// One workgroup of 4 waves; each wave walks 32-node tiles; sums are combined in a
// fixed order.
__global__ __launch_bounds__(256) void k_mlp_train_step(SgTrainState T, const float* __restrict__ Xall,
                                                        const float* __restrict__ Yall,
                                                        const int32_t* __restrict__ idx, int32_t n,
                                                        float lr, float beta1, float beta2, float eps,
                                                        int32_t step)
{
    __shared__ float s_grad[SG_NPARAM];
    __shared__ uint16_t s_dh[4][32][32];   // per-wave dH tile, BF16, [node][hidden]
    __shared__ uint16_t s_x[4][32][32];    // per-wave X tile, BF16, [node][feature]
    __shared__ float s_loss[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, c32 = lane & 31;
    for (int i = threadIdx.x; i < SG_NPARAM; i += blockDim.x) s_grad[i] = 0.0f;
    if (lane == 0) s_loss[wv] = 0.0f;
    __syncthreads();
    const float b2 = T.b2[0];
    const float inv_n = 1.0f / (float)n;
    // dW1 accumulators: 8 tiles of [32 features][32 hidden], element v <-> feature acc_row(v), hidden c32
    v16f gW[SG_HID / 32];
#pragma unroll
    for (int t = 0; t < SG_HID / 32; t++) gW[t] = (v16f){};
    float gw2[SG_HID / 32], gb1[SG_HID / 32];
#pragma unroll
    for (int t = 0; t < SG_HID / 32; t++) { gw2[t] = 0.0f; gb1[t] = 0.0f; }
    float gb2 = 0.0f, lsum = 0.0f;
    for (int64_t r0 = (int64_t)wv * 32; r0 < n; r0 += 128) {
        v8bf a0, a1;
        sg_load_x_idx(Xall, idx, n, r0, a0, a1);
        // stage the X tile (BF16) for the transposed operand of dW1
        {
            const int64_t row = r0 + c32;
            const int64_t src = row < n ? (idx ? (int64_t)idx[row] : row) : 0;
            for (int j = 0; j < 16; j++) {
                const int k = 16 * h + j;
                s_x[wv][c32][k] = (row < n && k < SG_FEAT) ? sg_bf16_bits(Xall[src * SG_FEAT + k]) : 0;
            }
        }
        float z[16];
        sg_logits(T.w1t, T.b1, T.w2b, b2, a0, a1, z);
        float dz[16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int64_t row = r0 + sg_acc_row(v);
            if (row < n) {
                const float y = Yall[idx ? (int64_t)idx[row] : row];
                const float s = sg_sigmoid(z[v]);
                dz[v] = (s - y) * inv_n;
                // softplus(z) - y z, stable form
                const float l = fmaxf(z[v], 0.0f) - y * z[v] + __logf(1.0f + __expf(-fabsf(z[v])));
                if (c32 == 0) lsum += l * inv_n;
            } else dz[v] = 0.0f;
        }
        if (c32 == 0) {
#pragma unroll
            for (int v = 0; v < 16; v++) gb2 += dz[v];
        }
        for (int nt = 0; nt < SG_HID / 32; nt++) {
            const v16f acc = sg_hidden_tile(T.w1t, nt, a0, a1);
            const int col = nt * 32 + c32;
            const float bb = T.b1[col], ww = sg_bf16_float(T.w2b[col]);
            float pw2 = 0.0f, pb1 = 0.0f;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const float hv = fmaxf(acc[v] + bb, 0.0f);
                pw2 += hv * dz[v];
                const float dh = hv > 0.0f ? dz[v] * ww : 0.0f;
                pb1 += dh;
                s_dh[wv][sg_acc_row(v)][c32] = sg_bf16_bits(dh);
            }
            gw2[nt] += pw2;
            gb1[nt] += pb1;
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            // dW1 tile nt += X^T dH over the 32 nodes: A[m=feature][k=node], B[k=node][n=hidden]
            uint16_t ua0[8], ua1[8], ub0[8], ub1[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                ua0[j] = s_x[wv][8 * h + j][c32];
                ua1[j] = s_x[wv][16 + 8 * h + j][c32];
                ub0[j] = s_dh[wv][8 * h + j][c32];
                ub1[j] = s_dh[wv][16 + 8 * h + j][c32];
            }
            gW[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sg_pack8(ua0), sg_pack8(ub0), gW[nt], 0, 0, 0);
            gW[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sg_pack8(ua1), sg_pack8(ub1), gW[nt], 0, 0, 0);
            __builtin_amdgcn_wave_barrier();
        }
    }
    // reduce per-lane partials: gw2/gb1 over the two half-waves, gb2/lsum over lanes 0 and 32
#pragma unroll
    for (int t = 0; t < SG_HID / 32; t++) {
        gw2[t] += __shfl_xor(gw2[t], 32, 64);
        gb1[t] += __shfl_xor(gb1[t], 32, 64);
    }
    gb2 += __shfl_xor(gb2, 32, 64);
    lsum += __shfl_xor(lsum, 32, 64);
    // combine waves in a fixed order
    for (int w = 0; w < 4; w++) {
        if (wv == w) {
            if (h == 0) {
                for (int t = 0; t < SG_HID / 32; t++) {
                    s_grad[SG_FEAT * SG_HID + t * 32 + c32] += gb1[t];
                    s_grad[SG_FEAT * SG_HID + SG_HID + t * 32 + c32] += gw2[t];
                }
            }
            if (lane == 0) {
                s_grad[SG_NPARAM - 1] += gb2;
                s_loss[0] += lsum;
            }
            for (int t = 0; t < SG_HID / 32; t++) {
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int f = sg_acc_row(v);
                    if (f < SG_FEAT) s_grad[f * SG_HID + t * 32 + c32] += gW[t][v];
                }
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < SG_NPARAM; i += blockDim.x) T.grad[i] = s_grad[i];
    if (threadIdx.x == 0) T.loss[0] = s_loss[0];
    if (lr > 0.0f) {
        const float bc1 = 1.0f - __powf(beta1, (float)step);
        const float bc2 = 1.0f - __powf(beta2, (float)step);
        for (int i = threadIdx.x; i < SG_NPARAM; i += blockDim.x) {
            const float g = s_grad[i];
            const float mm = beta1 * T.m[i] + (1.0f - beta1) * g;
            const float vv = beta2 * T.v[i] + (1.0f - beta2) * g * g;
            T.m[i] = mm;
            T.v[i] = vv;
            float* p;
            if (i < SG_FEAT * SG_HID) p = &T.W1[i];
            else if (i < SG_FEAT * SG_HID + SG_HID) p = &T.b1[i - SG_FEAT * SG_HID];
            else if (i < SG_FEAT * SG_HID + 2 * SG_HID) p = &T.w2[i - SG_FEAT * SG_HID - SG_HID];
            else p = &T.b2[0];
            *p -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
        }
    }
}

hipError_t sg_mlp_pack(const SgTrainState& T, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(32), dim3(256), 0, s, T);
    return hipGetLastError();
}

hipError_t sg_mlp_train_step(const SgTrainState& T, const float* X, const float* Y, const int32_t* idx, int32_t n,
                             float lr, float beta1, float beta2, float eps, int32_t step, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_train_step, dim3(1), dim3(256), 0, s, T, X, Y, idx, n, lr, beta1, beta2, eps, step);
    return hipGetLastError();
}

hipError_t sg_mlp_forward(const uint16_t* w1t, const float* b1, const uint16_t* w2b, const float* b2,
                          const float* X, int32_t n, float* out, int32_t logits, hipStream_t s)
{
    int64_t blocks = ((int64_t)n + 127) / 128;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mlp_score, dim3((unsigned)blocks), dim3(256), 0, s, w1t, b1, w2b, b2, X, n,
                       (const int64_t*)nullptr, out, logits);
    return hipGetLastError();
}
