/*
 * fim_ref.c -- CPU oracle for the MFI support-counting path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may call it; the product
 * path (the HIP library) never links or calls it.
 *
 * It restates, in plain C written for this repository, the reference engine's
 * CPU algorithm:
 *   - vertical bitmap build, bit (t & 63) of word (t >> 6) of row x set iff
 *     transaction t holds item x           code/engine_tensorfim/code/engine/src/kernel_bitmap.cu:342-380
 *   - the node counter getResFrequency:
 *     mask = AND of the prefix rows, res[e] = popcount(mask & row_e) for every
 *     survivor e, res[d'] = #{e : res[e] >= thr}                kernel_bitmap.cu:390-408
 *   - the search: roots are the survivors in stable descending-frequency
 *     order, a node X is emitted iff res[d'] == |X|, otherwise every dense id
 *     i < progress with res[i] >= thr becomes a child        kernel_bitmap.cu:841-894
 *   - the TensorFIM root rule (root iff freq/n > s, strict) kernel_bitmap.cu:816
 *     and the AnyFIM stage rule (roots = top-k of the stable frequency order,
 *     thr = k-th highest count)                AnyTime kernel.cu:844-882
 *
 * Two node counters are offered: REF_COUNT_ALL counts every survivor per node
 * exactly as getResFrequency does; REF_COUNT_INHERIT counts only the parent's
 * frequent extensions (downward closure makes the others infrequent), which
 * yields identical res[d'] and identical children, and is the CPU twin of the
 * device search's candidate inheritance.
 *
 * Threading: root subtrees are independent (the maximality test is local and
 * exact), so roots are distributed over OpenMP threads with dynamic
 * scheduling; results are sorted before they are returned, so the output does
 * not depend on the thread count.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_MAX_DEPTH 64

enum { REF_COUNT_ALL = 0, REF_COUNT_INHERIT = 1 };
enum { REF_ROOTS_ALL = 0, REF_ROOTS_STRICT = 1, REF_ROOTS_TOPK = 2 };

static inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

/* ---------------------------------------------------------------- vectorised AND + popcount
 * The node counter's inner loop, sum_w popcount(m[w] & r[w]) (kernel_bitmap.cu:400-405;
 * the CPU-vectorised miner does the same with a byte LUT, exp_e6_cpuvec.py:43-107). Two
 * implementations, chosen once at run time: AVX-512 VPOPCNTDQ (eight 64-bit popcounts
 * per instruction, four independent accumulators) on hosts that have it, else the scalar
 * popcnt loop. Both return the same integer; the tests check them against each other. */
static int64_t andpop_scalar(const uint64_t* m, const uint64_t* r, int64_t W)
{
    int64_t s = 0;
    for (int64_t w = 0; w < W; w++) s += popc64(m[w] & r[w]);
    return s;
}

__attribute__((target("avx512f,avx512vpopcntdq")))
static int64_t andpop_avx512(const uint64_t* m, const uint64_t* r, int64_t W)
{
    __m512i a0 = _mm512_setzero_si512(), a1 = a0, a2 = a0, a3 = a0;
    int64_t w = 0;
    for (; w + 32 <= W; w += 32) {
        a0 = _mm512_add_epi64(a0, _mm512_popcnt_epi64(_mm512_and_si512(_mm512_loadu_si512(m + w), _mm512_loadu_si512(r + w))));
        a1 = _mm512_add_epi64(a1, _mm512_popcnt_epi64(_mm512_and_si512(_mm512_loadu_si512(m + w + 8), _mm512_loadu_si512(r + w + 8))));
        a2 = _mm512_add_epi64(a2, _mm512_popcnt_epi64(_mm512_and_si512(_mm512_loadu_si512(m + w + 16), _mm512_loadu_si512(r + w + 16))));
        a3 = _mm512_add_epi64(a3, _mm512_popcnt_epi64(_mm512_and_si512(_mm512_loadu_si512(m + w + 24), _mm512_loadu_si512(r + w + 24))));
    }
    for (; w + 8 <= W; w += 8)
        a0 = _mm512_add_epi64(a0, _mm512_popcnt_epi64(_mm512_and_si512(_mm512_loadu_si512(m + w), _mm512_loadu_si512(r + w))));
    int64_t s = _mm512_reduce_add_epi64(_mm512_add_epi64(_mm512_add_epi64(a0, a1), _mm512_add_epi64(a2, a3)));
    for (; w < W; w++) s += popc64(m[w] & r[w]);
    return s;
}

/* dst[w] = a[w] & b[w] (the prefix masks); the compiler vectorises it for the host ISA */
__attribute__((target("avx512f")))
static void and_rows_avx512(uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t W)
{
    int64_t w = 0;
    for (; w + 8 <= W; w += 8)
        _mm512_storeu_si512(dst + w, _mm512_and_si512(_mm512_loadu_si512(a + w), _mm512_loadu_si512(b + w)));
    for (; w < W; w++) dst[w] = a[w] & b[w];
}

static void and_rows_scalar(uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t W)
{
    for (int64_t w = 0; w < W; w++) dst[w] = a[w] & b[w];
}

static int g_simd = -1;   /* -1: not chosen yet; 0: scalar; 1: AVX-512 VPOPCNTDQ */

static int simd_path(void)
{
    if (g_simd < 0) {
        __builtin_cpu_init();
        g_simd = (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vpopcntdq")) ? 1 : 0;
    }
    return g_simd;
}

/* The counter path in use (1 = AVX-512 VPOPCNTDQ, 0 = scalar popcnt). */
int ref_simd_path(void) { return simd_path(); }

/* Force the scalar path (0) or the host's best path (1; scalar where AVX-512 VPOPCNTDQ is
 * absent). Returns the path now in use. */
int ref_set_simd(int want)
{
    g_simd = -1;
    if (want) return simd_path();
    g_simd = 0;
    return 0;
}

static inline int64_t andpop(const uint64_t* m, const uint64_t* r, int64_t W)
{
    return simd_path() ? andpop_avx512(m, r, W) : andpop_scalar(m, r, W);
}

static inline void and_rows(uint64_t* dst, const uint64_t* a, const uint64_t* b, int64_t W)
{
    if (simd_path()) and_rows_avx512(dst, a, b, W);
    else and_rows_scalar(dst, a, b, W);
}

/* ---------------------------------------------------------------- bitmap build */
/* rows[(id-1) * W + (t >> 6)] |= 1 << (t & 63) for every item id in 1..n_items. */
void ref_bitmap(const int64_t* off, const int32_t* items, int64_t n_txn,
                int32_t n_items, int64_t W, uint64_t* rows)
{
    memset(rows, 0, sizeof(uint64_t) * (size_t)n_items * (size_t)W);
    for (int64_t t = 0; t < n_txn; t++)
        for (int64_t k = off[t]; k < off[t + 1]; k++) {
            int32_t id = items[k];
            if (id >= 1 && id <= n_items)
                rows[(size_t)(id - 1) * W + (t >> 6)] |= 1ULL << (t & 63);
        }
}

/* ---------------------------------------------------------------- counter */
/* out[j] = popcount(AND_{x in prefix} row_x  &  row_{cands[j]}); an empty prefix is all ones. */
void ref_count(const uint64_t* rows, int64_t W, const int32_t* prefix, int32_t k,
               const int32_t* cands, int32_t nc, int64_t* out)
{
    uint64_t* mask = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(W > 0 ? W : 1));
    for (int64_t w = 0; w < W; w++) mask[w] = ~0ULL;
    for (int32_t i = 0; i < k; i++) {
        const uint64_t* r = rows + (size_t)prefix[i] * W;
        for (int64_t w = 0; w < W; w++) mask[w] &= r[w];
    }
    for (int32_t j = 0; j < nc; j++) out[j] = andpop(mask, rows + (size_t)cands[j] * W, W);
    free(mask);
}

/* Every row's popcount. */
void ref_row_supports(const uint64_t* rows, int32_t d, int64_t W, int64_t* out)
{
    for (int32_t i = 0; i < d; i++) {
        int64_t s = 0;
        const uint64_t* r = rows + (size_t)i * W;
        for (int64_t w = 0; w < W; w++) s += popc64(r[w]);
        out[i] = s;
    }
}

/* Pair co-occurrence matrix out[i*d + j] = |T_i & T_j|. */
void ref_pair_matrix(const uint64_t* rows, int32_t d, int64_t W, int64_t* out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < d; i++)
        for (int32_t j = 0; j < d; j++) {
            const uint64_t* a = rows + (size_t)i * W;
            const uint64_t* b = rows + (size_t)j * W;
            int64_t s = 0;
            for (int64_t w = 0; w < W; w++) s += popc64(a[w] & b[w]);
            out[(size_t)i * d + j] = s;
        }
}

/* ---------------------------------------------------------------- result store */
typedef struct {
    int32_t* items;   /* concatenated itemsets (dense ids, insertion order) */
    int64_t* off;     /* n + 1 */
    int64_t* sup;     /* n */
    int64_t n, cap_sets, cap_items;
} ref_result;

static void res_init(ref_result* r)
{
    r->n = 0; r->cap_sets = 64; r->cap_items = 512;
    r->items = (int32_t*)malloc(sizeof(int32_t) * r->cap_items);
    r->off = (int64_t*)malloc(sizeof(int64_t) * (r->cap_sets + 1));
    r->sup = (int64_t*)malloc(sizeof(int64_t) * r->cap_sets);
    r->off[0] = 0;
}

static void res_push(ref_result* r, const int32_t* x, int k, int64_t sup)
{
    if (r->n == r->cap_sets) {
        r->cap_sets *= 2;
        r->off = (int64_t*)realloc(r->off, sizeof(int64_t) * (r->cap_sets + 1));
        r->sup = (int64_t*)realloc(r->sup, sizeof(int64_t) * r->cap_sets);
    }
    if (r->off[r->n] + k > r->cap_items) {
        while (r->off[r->n] + k > r->cap_items) r->cap_items *= 2;
        r->items = (int32_t*)realloc(r->items, sizeof(int32_t) * r->cap_items);
    }
    memcpy(r->items + r->off[r->n], x, sizeof(int32_t) * k);
    r->sup[r->n] = sup;
    r->off[r->n + 1] = r->off[r->n] + k;
    r->n++;
}

/* ---------------------------------------------------------------- search */
typedef struct {
    const uint64_t* rows;
    int32_t d;
    int64_t W;
    double thr;            /* a count c is frequent iff (double)c >= thr */
    int count_mode;
    int64_t max_nodes;     /* 0 = unlimited */
} ref_ctx;

typedef struct {
    uint64_t* masks;       /* (REF_MAX_DEPTH + 1) * W */
    int64_t* res;          /* (REF_MAX_DEPTH + 1) * d */
    int32_t* cand;         /* (REF_MAX_DEPTH + 1) * d : frequent-extension lists per depth */
    int32_t* ncand;        /* REF_MAX_DEPTH + 1 */
    int32_t x[REF_MAX_DEPTH + 1];
    ref_result out;
    int64_t nodes, counts;
    int overflow;
    /* optional node trace (single-threaded runs): one record per visited node */
    int64_t* tr;          /* [cap][6]: parent, depth, support, counted, is_mfi, last item */
    int64_t tr_n, tr_cap;
    int64_t tr_parent;    /* index of the node being expanded */
} ref_worker;

/* Count node x[0..k-1] whose tidset is masks[k-1]; candidates are all survivors
 * (REF_COUNT_ALL) or the parent's frequent extensions (REF_COUNT_INHERIT). */
static void visit(const ref_ctx* c, ref_worker* w, int k, int64_t freq)
{
    const int64_t W = c->W;
    const int32_t d = c->d;
    if (k > REF_MAX_DEPTH) { w->overflow = 1; return; }
    if (c->max_nodes && w->nodes >= c->max_nodes) { w->overflow = 2; return; }
    w->nodes++;
    int64_t me = -1;
    if (w->tr && w->tr_n < w->tr_cap) {
        me = w->tr_n++;
        int64_t* r = w->tr + me * 6;
        r[0] = w->tr_parent; r[1] = k; r[2] = freq; r[3] = 0; r[4] = 0; r[5] = w->x[k - 1];
    }
    const uint64_t* m = w->masks + (size_t)(k - 1) * W;
    int64_t* res = w->res + (size_t)k * d;
    int32_t* fl = w->cand + (size_t)k * d;
    int nf = 0;
    const int32_t* src; int nsrc;
    if (c->count_mode == REF_COUNT_INHERIT && k > 1) {
        src = w->cand + (size_t)(k - 1) * d;
        nsrc = w->ncand[k - 1];
    } else {
        src = NULL; nsrc = d;
    }
    for (int j = 0; j < nsrc; j++) {
        int32_t e = src ? src[j] : j;
        const int64_t s = andpop(m, c->rows + (size_t)e * W, W);
        res[e] = s;
        w->counts++;
        if ((double)s >= c->thr) fl[nf++] = e;
    }
    w->ncand[k] = nf;
    if (me >= 0) w->tr[me * 6 + 3] = nsrc;
    /* items of X are frequent extensions of X (res = freq(X)), so the node
     * is maximal iff the frequent-extension count equals |X| (kernel_bitmap.cu:870). */
    if (nf == k) {
        if (me >= 0) w->tr[me * 6 + 4] = 1;
        res_push(&w->out, w->x, k, freq);
        return;
    }
    int32_t progress = w->x[k - 1];
    /* children: frequent ids below progress; pushed ascending onto a LIFO in the
     * reference, visited here in descending id order -- the same tree. */
    for (int j = nf - 1; j >= 0; j--) {
        int32_t i = fl[j];
        if (i >= progress) continue;
        and_rows(w->masks + (size_t)k * W, m, c->rows + (size_t)i * W, W);
        w->x[k] = i;
        const int64_t saved = w->tr_parent;
        w->tr_parent = me;
        visit(c, w, k + 1, res[i]);
        w->tr_parent = saved;
        if (w->overflow) return;
        /* the recursion overwrote deeper levels only; level k lists stay valid */
    }
}

static int cmp_desc_freq_idx(const void* a, const void* b, void* ctx)
{
    const int64_t* f = (const int64_t*)ctx;
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    if (f[x] != f[y]) return f[x] > f[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

/* stable descending-frequency order of the survivors (ties: ascending id), kernel_bitmap.cu:276-283 */
static void freq_order(const int64_t* f, int32_t d, int32_t* order)
{
    for (int32_t i = 0; i < d; i++) order[i] = i;
    /* insertion sort keeps stability and d is small */
    for (int32_t i = 1; i < d; i++) {
        int32_t v = order[i]; int32_t j = i - 1;
        while (j >= 0 && cmp_desc_freq_idx(&order[j], &v, (void*)f) > 0) { order[j + 1] = order[j]; j--; }
        order[j + 1] = v;
    }
}

/*
 * Mine the rows (dense ids 0..d-1, every row already a survivor) at threshold thr.
 * root_mode: REF_ROOTS_ALL    -> every survivor is a root (exact MFI family)
 *            REF_ROOTS_STRICT -> root iff (double)freq / n_total > s   (TensorFIM rule)
 *            REF_ROOTS_TOPK   -> the first topk survivors of the stable frequency order
 * Returns a result handle (sets in dense ids, insertion order: root first).
 */
static int64_t* g_trace = 0;
static int64_t g_trace_cap = 0, g_trace_n = 0;

/* Record the visited nodes of the next ref_mine run (forced single-threaded) into
 * trace[cap][6] = (parent index or -1, depth, support, candidates counted, is MFI,
 * last item); ref_trace_count() returns the number recorded. */
void ref_trace_set(int64_t* trace, int64_t cap) { g_trace = trace; g_trace_cap = cap; g_trace_n = 0; }
int64_t ref_trace_count(void) { return g_trace_n; }

ref_result* ref_mine(const uint64_t* rows, int32_t d, int64_t W, double thr,
                     int root_mode, int64_t n_total, double s, int32_t topk,
                     int count_mode, int threads, int64_t max_nodes,
                     int64_t* stats /* [nodes, counts, overflow] */)
{
    if (g_trace) threads = 1;
    ref_ctx c = { rows, d, W, thr, count_mode, max_nodes };
    int64_t* f = (int64_t*)malloc(sizeof(int64_t) * (d > 0 ? d : 1));
    ref_row_supports(rows, d, W, f);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (d > 0 ? d : 1));
    freq_order(f, d, order);
    int32_t nroots = 0;
    for (int32_t i = 0; i < d; i++) {
        int32_t e = order[i];
        int ok;
        if (root_mode == REF_ROOTS_STRICT) ok = ((double)f[e] / (double)n_total) > s;
        else if (root_mode == REF_ROOTS_TOPK) ok = i < topk;
        else ok = (double)f[e] >= thr;
        if (ok) order[nroots++] = e;
    }
    if (threads <= 0) threads = 1;
    ref_worker* ws = (ref_worker*)calloc((size_t)threads, sizeof(ref_worker));
    for (int t = 0; t < threads; t++) {
        ws[t].masks = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(REF_MAX_DEPTH + 1) * (W > 0 ? W : 1));
        ws[t].res = (int64_t*)malloc(sizeof(int64_t) * (size_t)(REF_MAX_DEPTH + 2) * (d > 0 ? d : 1));
        ws[t].cand = (int32_t*)malloc(sizeof(int32_t) * (size_t)(REF_MAX_DEPTH + 2) * (d > 0 ? d : 1));
        ws[t].ncand = (int32_t*)malloc(sizeof(int32_t) * (REF_MAX_DEPTH + 2));
        res_init(&ws[t].out);
        ws[t].tr = g_trace; ws[t].tr_cap = g_trace_cap; ws[t].tr_n = 0;
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int32_t ri = 0; ri < nroots; ri++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        ref_worker* w = &ws[tid];
        if (w->overflow) continue;
        int32_t r = order[ri];
        w->tr_parent = -1;
        memcpy(w->masks, rows + (size_t)r * W, sizeof(uint64_t) * (size_t)W);
        w->x[0] = r;
        visit(&c, w, 1, f[r]);
    }
    ref_result* out = (ref_result*)malloc(sizeof(ref_result));
    res_init(out);
    int64_t nodes = 0, counts = 0, ovf = 0;
    for (int t = 0; t < threads; t++) {
        for (int64_t i = 0; i < ws[t].out.n; i++)
            res_push(out, ws[t].out.items + ws[t].out.off[i],
                     (int)(ws[t].out.off[i + 1] - ws[t].out.off[i]), ws[t].out.sup[i]);
        nodes += ws[t].nodes; counts += ws[t].counts;
        if (g_trace) g_trace_n = ws[t].tr_n;
        if (ws[t].overflow) ovf = ws[t].overflow;
        free(ws[t].masks); free(ws[t].res); free(ws[t].cand); free(ws[t].ncand);
        free(ws[t].out.items); free(ws[t].out.off); free(ws[t].out.sup);
    }
    free(ws); free(order); free(f);
    g_trace = 0;
    if (stats) { stats[0] = nodes; stats[1] = counts; stats[2] = ovf; }
    return out;
}

int64_t ref_result_n(const ref_result* r) { return r->n; }
int64_t ref_result_nitems(const ref_result* r) { return r->off[r->n]; }
void ref_result_copy(const ref_result* r, int32_t* items, int64_t* off, int64_t* sup)
{
    memcpy(items, r->items, sizeof(int32_t) * (size_t)r->off[r->n]);
    memcpy(off, r->off, sizeof(int64_t) * (size_t)(r->n + 1));
    memcpy(sup, r->sup, sizeof(int64_t) * (size_t)r->n);
}
void ref_result_free(ref_result* r)
{
    if (!r) return;
    free(r->items); free(r->off); free(r->sup); free(r);
}

/* ---------------------------------------------------------------- antichain check
 * Number of itemsets of a family that are proper subsets of another member (0 for an
 * antichain, as every family of maximal itemsets is). Itemsets are items[off[i] ..
 * off[i+1]) over ids 0..d-1 in any order. A superset Y of X contains X's item that lies
 * in the fewest members, so only that item's members are compared with X: first by a
 * 64-bit signature of item ids mod 64, then exactly on sorted copies. */
static int cmp_i32(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int64_t ref_antichain_violations(const int32_t* items, const int64_t* off, int64_t n, int32_t d)
{
    const int64_t m = off[n];
    int32_t* srt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    uint64_t* sig = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* cnt = (int64_t*)calloc((size_t)d + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const int64_t a = off[i], k = off[i + 1] - off[i];
        memcpy(srt + a, items + a, sizeof(int32_t) * (size_t)k);
        qsort(srt + a, (size_t)k, sizeof(int32_t), cmp_i32);
        uint64_t s = 0;
        for (int64_t j = 0; j < k; j++) s |= 1ULL << (srt[a + j] & 63);
        sig[i] = s;
    }
    for (int64_t j = 0; j < m; j++) cnt[srt[j] + 1]++;
    for (int32_t x = 0; x < d; x++) cnt[x + 1] += cnt[x];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d > 0 ? d : 1));
    int64_t* lst = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    for (int32_t x = 0; x < d; x++) fill[x] = cnt[x];
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = off[i]; j < off[i + 1]; j++) lst[fill[srt[j]]++] = i;
    int64_t bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (int64_t i = 0; i < n; i++) {
        const int64_t a = off[i], k = off[i + 1] - off[i];
        if (k == 0) { bad += n > 1; continue; }
        int32_t best = srt[a];
        for (int64_t j = a + 1; j < a + k; j++)
            if (cnt[srt[j] + 1] - cnt[srt[j]] < cnt[best + 1] - cnt[best]) best = srt[j];
        for (int64_t p = cnt[best]; p < cnt[best + 1]; p++) {
            const int64_t y = lst[p];
            const int64_t b = off[y], ky = off[y + 1] - off[y];
            if (y == i || ky <= k || (sig[i] & ~sig[y])) continue;
            int64_t u = 0, v = 0;
            while (u < k && v < ky) {
                if (srt[a + u] == srt[b + v]) { u++; v++; }
                else if (srt[a + u] > srt[b + v]) v++;
                else break;
            }
            if (u == k) { bad++; break; }
        }
    }
    free(srt); free(sig); free(cnt); free(fill); free(lst);
    return bad;
}
