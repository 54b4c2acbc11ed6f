/*
 * zipf_ref.c -- CPU restatement of the synthetic dense-sparse (Zipf) generator --
 * TEST INFRASTRUCTURE ONLY (the product library generates the set on the GPU).
 *
 * Specification: item i (1..d) is present in a transaction independently with
 * probability q_i = min(1, c i^-s), c chosen (bisection) so that sum q_i = mean_len;
 * planted itemset j is ORed in with probability r_j. Draw for (transaction t, item i):
 * u = (mix64(seed ^ t*G1 ^ (i+1)*G2) >> 11) / 2^53 < q_i, with mix64 the splitmix64
 * finaliser, G1 = 0x9E3779B97F4A7C15, G2 = 0xC2B2AE3D27D4EB4F; planted sets use
 * seed ^ 0x5851F42D4C957F2D and (j+1) in place of (i+1). Planted sizes, items and
 * rates come from a splitmix64 stream started at seed ^ 0x243F6A8885A308D3.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static uint64_t zmix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double zsum(int32_t d, double s, double c)
{
    double t = 0.0;
    for (int32_t i = 1; i <= d; i++) { double v = c * pow((double)i, -s); t += v < 1.0 ? v : 1.0; }
    return t;
}

void zipf_spec_ref(int32_t d, double s, double mean_len, int32_t np, int32_t slo, int32_t shi,
                   double rlo, double rhi, uint64_t seed, double* q, int32_t* poff, int32_t* pit, double* prate)
{
    double lo = 0.0, hi = 1.0;
    while (zsum(d, s, hi) < mean_len) hi *= 2.0;
    for (int it = 0; it < 200; it++) {
        double m = 0.5 * (lo + hi);
        if (zsum(d, s, m) < mean_len) lo = m; else hi = m;
    }
    for (int32_t i = 1; i <= d; i++) { double v = hi * pow((double)i, -s); q[i - 1] = v < 1.0 ? v : 1.0; }
    uint64_t st = seed ^ 0x243F6A8885A308D3ULL;
    poff[0] = 0;
    for (int32_t j = 0; j < np; j++) {
        st += 0x9E3779B97F4A7C15ULL;
        int32_t sz = slo + (int32_t)(zmix(st) % (uint64_t)(shi - slo + 1));
        int32_t k = poff[j];
        while (k - poff[j] < sz) {
            st += 0x9E3779B97F4A7C15ULL;
            int32_t x = 1 + (int32_t)(zmix(st) % (uint64_t)d);
            int dup = 0;
            for (int32_t a = poff[j]; a < k; a++) dup |= pit[a] == x;
            if (!dup) pit[k++] = x;
        }
        poff[j + 1] = k;
        st += 0x9E3779B97F4A7C15ULL;
        prate[j] = rlo + (rhi - rlo) * ((double)(zmix(st) >> 11) * (1.0 / 9007199254740992.0));
    }
}

/* CSR of transactions [0, n): off [n+1], items (ascending ids per row); returns nnz.
 * Call once with items == NULL to size. */
int64_t zipf_csr_ref(int64_t n, int32_t d, const double* q, int32_t np, const int32_t* poff, const int32_t* pit,
                     const double* prate, uint64_t seed, int64_t* off, int32_t* items)
{
    const double two53 = 9007199254740992.0;
    uint64_t* q53 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)d);
    uint64_t* r53 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(np > 0 ? np : 1));
    for (int32_t i = 0; i < d; i++) q53[i] = (uint64_t)(q[i] * two53);
    for (int32_t j = 0; j < np; j++) r53[j] = (uint64_t)(prate[j] * two53);
    unsigned char* mark = (unsigned char*)calloc((size_t)d + 1, 1);
    const uint64_t sp = seed ^ 0x5851F42D4C957F2DULL;
    int64_t nnz = 0;
    if (off) off[0] = 0;
    for (int64_t t = 0; t < n; t++) {
        const uint64_t tk = (uint64_t)t * 0x9E3779B97F4A7C15ULL;
        for (int32_t i = 0; i < d; i++)
            mark[i + 1] = (zmix(seed ^ tk ^ ((uint64_t)(i + 1) * 0xC2B2AE3D27D4EB4FULL)) >> 11) < q53[i];
        for (int32_t j = 0; j < np; j++)
            if ((zmix(sp ^ tk ^ ((uint64_t)(j + 1) * 0xC2B2AE3D27D4EB4FULL)) >> 11) < r53[j])
                for (int32_t a = poff[j]; a < poff[j + 1]; a++) mark[pit[a]] = 1;
        for (int32_t i = 1; i <= d; i++)
            if (mark[i]) { if (items) items[nnz] = i; nnz++; }
        if (off) off[t + 1] = nnz;
    }
    free(q53); free(r53); free(mark);
    return nnz;
}

/* Vertical bitmap rows[(i-1) * W + (t >> 6)] of the same set (W = ceil(n / 64) words per
 * row), drawn straight into the row layout: each 64-transaction word is owned by one
 * OpenMP thread, so the draws need no synchronisation. The draws are those of
 * zipf_csr_ref, transaction by transaction. */
void zipf_bitmap_ref(int64_t n, int32_t d, const double* q, int32_t np, const int32_t* poff, const int32_t* pit,
                     const double* prate, uint64_t seed, int64_t W, uint64_t* rows)
{
    const double two53 = 9007199254740992.0;
    uint64_t* q53 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)d);
    uint64_t* r53 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(np > 0 ? np : 1));
    for (int32_t i = 0; i < d; i++) q53[i] = (uint64_t)(q[i] * two53);
    for (int32_t j = 0; j < np; j++) r53[j] = (uint64_t)(prate[j] * two53);
    const uint64_t sp = seed ^ 0x5851F42D4C957F2DULL;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t w = 0; w < W; w++) {
        for (int32_t i = 0; i < d; i++) {
            uint64_t v = 0;
            for (int b = 0; b < 64; b++) {
                const int64_t t = w * 64 + b;
                if (t >= n) break;
                const uint64_t tk = (uint64_t)t * 0x9E3779B97F4A7C15ULL;
                v |= (uint64_t)((zmix(seed ^ tk ^ ((uint64_t)(i + 1) * 0xC2B2AE3D27D4EB4FULL)) >> 11) < q53[i]) << b;
            }
            rows[(size_t)i * (size_t)W + (size_t)w] = v;
        }
        for (int32_t j = 0; j < np; j++) {
            uint64_t v = 0;
            for (int b = 0; b < 64; b++) {
                const int64_t t = w * 64 + b;
                if (t >= n) break;
                const uint64_t tk = (uint64_t)t * 0x9E3779B97F4A7C15ULL;
                v |= (uint64_t)((zmix(sp ^ tk ^ ((uint64_t)(j + 1) * 0xC2B2AE3D27D4EB4FULL)) >> 11) < r53[j]) << b;
            }
            for (int32_t a = poff[j]; a < poff[j + 1]; a++) rows[(size_t)(pit[a] - 1) * (size_t)W + (size_t)w] |= v;
        }
    }
    free(q53); free(r53);
}
