/*
 * batch_proto_kat.c -- regenerates the reference's seeded batch-prototype dataset
 * (the only known-answer test of this path that can be rebuilt) -- TEST
 * INFRASTRUCTURE ONLY; the product library never links it.
 *
 * Restates gen_data of code/engine_tensorfim/code/batch_proto/batch_proto.cu:193-230:
 *   xorshift64 (s ^= s<<13; s ^= s>>7; s ^= s<<17) seeded 0xC0FFEE123456789;
 *   80 patterns: size 2 + x % 5, freq (0.2 + 1.3 * (x % 1000) / 1000) / 100,
 *   distinct items x % 512 drawn until the pattern is full, then sorted;
 *   for every transaction t of N = 2^23 and every pattern p in order: the pattern
 *   fires iff (uint32)xs64() < (uint32)(freq_p * 2^32), setting its items' bits;
 *   then, per item j, (int)(N * 0.0002) noise bits at t = xs64() % N.
 * The published answer at minsup (int)(N * 0.15 / 100) = 12582
 * (code/engine_tensorfim/code/batch_proto/run_dispatch.txt:3-15):
 *   |F1..F7| = 248 / 585 / 587 / 350 / 117 / 17 / 0, total 1904, 80/80 planted
 *   patterns recalled with no false positive.
 * Bits are stored as uint64 words: bit t & 63 of word t >> 6, which on a
 * little-endian machine is the reference's uint32 layout (bit t & 31 of word t >> 5).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KAT_D 512
#define KAT_NPAT 80
#define KAT_MAXPAT 6

static inline uint64_t xs64(uint64_t* s)
{
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

/* rows: KAT_D x (n/64) uint64, zeroed by the callee. pat_items: KAT_NPAT x KAT_MAXPAT
 * (-1 padded), pat_size: KAT_NPAT, pat_freq: KAT_NPAT. n must be a multiple of 64. */
void kat_generate(int64_t n, uint64_t* rows, int32_t* pat_items, int32_t* pat_size, double* pat_freq)
{
    const int64_t W = n / 64;
    uint64_t seed = 0xC0FFEE123456789ULL;
    memset(rows, 0, sizeof(uint64_t) * (size_t)KAT_D * (size_t)W);
    for (int p = 0; p < KAT_NPAT; p++) {
        int sz = 2 + (int)(xs64(&seed) % (KAT_MAXPAT - 1));
        double freq = (0.2 + 1.3 * (double)(xs64(&seed) % 1000) / 1000.0) / 100.0;
        int used[KAT_MAXPAT];
        int k = 0;
        while (k < sz) {
            int it = (int)(xs64(&seed) % KAT_D);
            int dup = 0;
            for (int q = 0; q < k; q++) if (used[q] == it) dup = 1;
            if (!dup) used[k++] = it;
        }
        for (int a = 1; a < sz; a++) {  /* sort ascending */
            int v = used[a], b = a - 1;
            while (b >= 0 && used[b] > v) { used[b + 1] = used[b]; b--; }
            used[b + 1] = v;
        }
        for (int q = 0; q < KAT_MAXPAT; q++) pat_items[p * KAT_MAXPAT + q] = q < sz ? used[q] : -1;
        pat_size[p] = sz;
        pat_freq[p] = freq;
    }
    uint32_t thr[KAT_NPAT];
    for (int p = 0; p < KAT_NPAT; p++) thr[p] = (uint32_t)(pat_freq[p] * 4294967296.0);
    for (int64_t t = 0; t < n; t++) {
        const int64_t wd = t >> 6;
        const uint64_t bit = 1ULL << (t & 63);
        for (int p = 0; p < KAT_NPAT; p++) {
            if ((uint32_t)xs64(&seed) < thr[p]) {
                for (int q = 0; q < pat_size[p]; q++)
                    rows[(size_t)pat_items[p * KAT_MAXPAT + q] * W + wd] |= bit;
            }
        }
    }
    const int noise = (int)((double)n * 0.0002);
    for (int j = 0; j < KAT_D; j++) {
        for (int q = 0; q < noise; q++) {
            const uint64_t t = xs64(&seed) % (uint64_t)n;
            rows[(size_t)j * W + (t >> 6)] |= 1ULL << (t & 63);
        }
    }
}

/* CSR transactions (1-based ids j+1) from the bitmap: off has n+1 entries; items
 * must hold the total bit count (returned by kat_nnz). */
int64_t kat_nnz(const uint64_t* rows, int64_t n)
{
    const int64_t W = n / 64;
    int64_t s = 0;
    for (size_t i = 0; i < (size_t)KAT_D * (size_t)W; i++) s += __builtin_popcountll(rows[i]);
    return s;
}

void kat_csr(const uint64_t* rows, int64_t n, int64_t* off, int32_t* items)
{
    const int64_t W = n / 64;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int j = 0; j < KAT_D; j++)
        for (int64_t w = 0; w < W; w++) {
            uint64_t v = rows[(size_t)j * W + w];
            while (v) {
                const int b = __builtin_ctzll(v);
                v &= v - 1;
                cnt[w * 64 + b + 1]++;
            }
        }
    off[0] = 0;
    for (int64_t t = 0; t < n; t++) off[t + 1] = off[t] + cnt[t + 1];
    int64_t* pos = cnt;
    for (int64_t t = 0; t < n; t++) pos[t] = off[t];
    for (int j = 0; j < KAT_D; j++)      /* ascending item order inside each row */
        for (int64_t w = 0; w < W; w++) {
            uint64_t v = rows[(size_t)j * W + w];
            while (v) {
                const int b = __builtin_ctzll(v);
                v &= v - 1;
                items[pos[w * 64 + b]++] = j + 1;
            }
        }
    free(cnt);
}

/* Level-wise frequent-itemset enumeration with canonical extensions (j > last),
 * restating the search of batch_proto.cu:287-374 on the CPU: level_count[L] =
 * |F_L| for L = 1..7 (index 0 unused). Candidates are restricted to frequent
 * items, which by downward closure changes no count. Returns the total. */
int64_t kat_levelwise(const uint64_t* rows, int64_t n, int64_t minsup, int64_t* level_count,
                      int32_t* out_sets, int64_t out_cap)
{
    const int64_t W = n / 64;
    int f1[KAT_D], nf1 = 0;
    for (int j = 0; j < KAT_D; j++) {
        int64_t s = 0;
        for (int64_t w = 0; w < W; w++) s += __builtin_popcountll(rows[(size_t)j * W + w]);
        if (s >= minsup) f1[nf1++] = j;
    }
    for (int L = 0; L < 8; L++) level_count[L] = 0;
    level_count[1] = nf1;
    /* current level: itemsets of size L stored flat */
    int64_t ncur = nf1, total = nf1, nout = 0;
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf1);
    for (int i = 0; i < nf1; i++) cur[i] = f1[i];
    for (int i = 0; i < nf1 && nout < out_cap; i++) { out_sets[nout * 8] = 1; out_sets[nout * 8 + 1] = f1[i]; nout++; }
    for (int L = 2; L <= 7 && ncur > 0; L++) {
        int64_t cap = ncur * nf1;
        int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1) * L);
        int64_t nnext = 0;
#pragma omp parallel
        {
            uint64_t* mask = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W);
#pragma omp for schedule(dynamic, 1)
            for (int64_t p = 0; p < ncur; p++) {
                const int32_t* x = cur + p * (L - 1);
                for (int64_t w = 0; w < W; w++) {
                    uint64_t v = ~0ULL;
                    for (int q = 0; q < L - 1; q++) v &= rows[(size_t)x[q] * W + w];
                    mask[w] = v;
                }
                for (int i = 0; i < nf1; i++) {
                    const int j = f1[i];
                    if (j <= x[L - 2]) continue;
                    const uint64_t* r = rows + (size_t)j * W;
                    int64_t s = 0;
                    for (int64_t w = 0; w < W; w++) s += __builtin_popcountll(mask[w] & r[w]);
                    if (s >= minsup) {
                        int64_t slot;
#pragma omp atomic capture
                        slot = nnext++;
                        for (int q = 0; q < L - 1; q++) next[slot * L + q] = x[q];
                        next[slot * L + L - 1] = j;
                    }
                }
            }
            free(mask);
        }
        /* order-independent output: sort lexicographically */
        for (int64_t a = 1; a < nnext; a++) {
            int32_t tmp[8];
            memcpy(tmp, next + a * L, sizeof(int32_t) * L);
            int64_t b = a - 1;
            while (b >= 0) {
                int cmp = 0;
                for (int q = 0; q < L && !cmp; q++) cmp = (next[b * L + q] > tmp[q]) - (next[b * L + q] < tmp[q]);
                if (cmp <= 0) break;
                memcpy(next + (b + 1) * L, next + b * L, sizeof(int32_t) * L);
                b--;
            }
            memcpy(next + (b + 1) * L, tmp, sizeof(int32_t) * L);
        }
        for (int64_t a = 0; a < nnext && nout < out_cap; a++, nout++) {
            out_sets[nout * 8] = L;
            for (int q = 0; q < L; q++) out_sets[nout * 8 + 1 + q] = next[a * L + q];
        }
        level_count[L] = nnext;
        total += nnext;
        free(cur);
        cur = next;
        ncur = nnext;
    }
    free(cur);
    return total;
}
