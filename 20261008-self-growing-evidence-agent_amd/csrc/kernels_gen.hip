// Synthetic Zipf transaction set generated straight into the vertical bitmaps
// (the large dense-sparse workload: 10,000,000 transactions over 2,048 items).
//
// Specification (fixed here; the reference publishes no generator for it):
//   * item i (1-based rank) is present in a transaction independently with
//     probability q_i = min(1, c * i^-s), c chosen so that sum_i q_i = mean_len;
//   * planted itemset j (of n_planted) is ORed into a transaction with probability r_j;
//   * every draw is a counter-based hash: u(t, i) = top 53 bits of
//     mix64(seed ^ (t * G1) ^ ((i + 1) * G2)) / 2^53 for items, and the same with
//     (seed ^ P) and j for planted sets, where mix64 is the splitmix64 finaliser and
//     G1 = 0x9E3779B97F4A7C15, G2 = 0xC2B2AE3D27D4EB4F, P = 0x5851F42D4C957F2D;
//   * bit u < q_i (resp. u < r_j) is drawn independently per (transaction, item/set).
// The C oracle restates the same draws to build the CSR of small sets for parity.
// One thread builds one 64-bit word (64 transactions) of one item row, then adds the
// planted sets containing that item; row popcounts give the item supports.
#include "sgfim_impl.h"

__device__ __forceinline__ uint64_t sg_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0,1) with 53 bits, compared as integers: u < p  <=>  (h >> 11) < p * 2^53
__global__ __launch_bounds__(256) void k_gen_zipf(uint64_t* __restrict__ rows, int64_t Wp, int64_t n_txn,
                                                  int32_t d, const uint64_t* __restrict__ q53,
                                                  const uint64_t* __restrict__ r53,
                                                  const int32_t* __restrict__ item_sets_off,
                                                  const int32_t* __restrict__ item_sets, uint64_t seed,
                                                  int32_t* __restrict__ sup)
{
    const int64_t W = (n_txn + 63) / 64;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t i = (int32_t)(gid / Wp);
    const int64_t w = gid % Wp;
    if (i >= d) return;
    uint64_t v = 0;
    if (w < W) {
        const uint64_t thr = q53[i];
        const int64_t t0 = w * 64;
        const int nt = (int)min((int64_t)64, n_txn - t0);
        const uint64_t ik = (uint64_t)(i + 1) * 0xC2B2AE3D27D4EB4Full;
        for (int b = 0; b < nt; b++) {
            const uint64_t h = sg_mix64(seed ^ ((uint64_t)(t0 + b) * 0x9E3779B97F4A7C15ull) ^ ik);
            if ((h >> 11) < thr) v |= 1ull << b;
        }
        const uint64_t sp = seed ^ 0x5851F42D4C957F2Dull;
        for (int a = item_sets_off[i]; a < item_sets_off[i + 1]; a++) {
            const int j = item_sets[a];
            const uint64_t jk = (uint64_t)(j + 1) * 0xC2B2AE3D27D4EB4Full;
            const uint64_t rt = r53[j];
            for (int b = 0; b < nt; b++) {
                const uint64_t h = sg_mix64(sp ^ ((uint64_t)(t0 + b) * 0x9E3779B97F4A7C15ull) ^ jk);
                if ((h >> 11) < rt) v |= 1ull << b;
            }
        }
    }
    rows[(int64_t)i * Wp + w] = v;
    uint32_t c = (uint32_t)__popcll(v);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&sup[i], (int32_t)c);
}

hipError_t sg_launch_gen_zipf(uint64_t* rows, int64_t Wp, int64_t n_txn, int32_t d, const uint64_t* q53,
                              const uint64_t* r53, const int32_t* item_sets_off, const int32_t* item_sets,
                              uint64_t seed, int32_t* sup, hipStream_t s)
{
    const int64_t total = (int64_t)d * Wp;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_zipf, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, rows, Wp, n_txn, d, q53,
                       r53, item_sets_off, item_sets, seed, sup);
    return hipGetLastError();
}
