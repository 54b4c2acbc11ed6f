// Device-resident MFI search: frontier, batched counting, maximality test, MFI ring.
//
// The reference keeps its search state on the host: a mutex-protected LIFO of
// struct FI {itemsSet, frequency, root, progress} (kernel_bitmap.cu:38-43,300-318),
// one counter call per popped node, and the node step of kernel_bitmap.cu:866-890:
//   emit X iff #{e : res[e] >= thr} == |X|; otherwise push X u {i} for every dense
//   id i < progress with res[i] >= thr, with frequency res[i] and progress i.
// Here the same tree is searched with the state in HBM:
//   arena      node records (items, support, tail bound, candidate mask)
//   buckets    64 priority stacks of arena indices; a round takes up to bmax nodes
//              from the highest non-empty buckets (LIFO inside a bucket)
//   free list  stack of free arena slots
//   ring       confirmed MFIs, drained by the host after every round
// Candidates: a node counts only the items that can still extend it. For the
// child Y = X u {e} that set is (F(X) \ {e}) & ok[e], where F(X) holds the
// frequent one-item extensions of X and ok[e] the items whose pair support with e
// reaches the threshold; by downward closure every frequent extension of Y lies
// in it, so the frequent-extension count -- and with it the reference's local,
// exact maximality test -- is unchanged while most infeasible counts are skipped.
// Roots are processed straight from the pair matrix (their counts are its rows).
//
// One round = k_select -> k_candlist -> k_count_search -> k_expand -> [score]
// -> k_push -> k_release; every launch has a fixed shape and reads its sizes from
// the control block, so the whole round can be captured once as a hipGraph.
#include "count_core.h"

#define SEL_THREADS 1024

__device__ __forceinline__ int sg_bucket_support(int32_t sup, int32_t thr)
{
    const float r = __log2f((float)sup / (float)thr);
    int b = (int)(8.0f * r);
    return b < 0 ? 0 : (b > SG_NBUCKETS - 1 ? SG_NBUCKETS - 1 : b);
}

__device__ __forceinline__ int sg_bucket_score(float s)
{
    int b = (int)(s * (float)SG_NBUCKETS);
    return b < 0 ? 0 : (b > SG_NBUCKETS - 1 ? SG_NBUCKETS - 1 : b);
}

// ------------------------------------------------------------------ init
__global__ void k_search_init(SgSearchDev S)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.cap) S.free_idx[i] = (int32_t)(S.cap - 1 - i);
    if (i < CTL_SLOTS) S.ctl[i] = 0;
    if (i == 0) S.ctl[CTL_FREE_TOP] = S.cap;
}

// ------------------------------------------------------------------ features
// 24 features of a newborn node Y = X u {e}, all taken from what the miner holds.
struct SgFeatIn {
    int32_t sup_y, sup_x, k_y, nf_x, ncand_y, tail_y, thr, n, d, e;
    int32_t sup_e, min_item, max_item, pair_min, nf_x_tail, root;
    float mean_log_item, sum_log_item;
};

__device__ __forceinline__ void sg_features(const SgFeatIn& f, float* out)
{
    const float lt = __log2f((float)f.thr);
    const float ln = __log2f((float)f.n);
    out[0] = __log2f((float)f.sup_y) - lt;
    out[1] = __log2f((float)f.sup_x) - lt;
    out[2] = (float)f.sup_y / (float)f.sup_x;
    out[3] = (float)f.k_y * (1.0f / 8.0f);
    out[4] = __log2f(1.0f + (float)f.nf_x);
    out[5] = __log2f(1.0f + (float)f.ncand_y);
    out[6] = __log2f(1.0f + (float)f.tail_y);
    out[7] = __log2f((float)f.sup_e) - lt;
    out[8] = __log2f((float)f.min_item) - lt;
    out[9] = __log2f((float)f.max_item) - lt;
    out[10] = f.mean_log_item - lt;
    out[11] = (float)f.pair_min / (float)f.sup_e;
    out[12] = (float)f.e / (float)f.d;
    out[13] = (float)f.tail_y / (float)(f.ncand_y > 0 ? f.ncand_y : 1);
    out[14] = (float)f.sup_y / (float)f.min_item;
    // log-lift of Y against independent items
    out[15] = (__log2f((float)f.sup_y) - ln) - (f.sum_log_item - (float)f.k_y * ln);
    out[16] = __log2f(1.0f + (float)f.nf_x_tail);
    out[17] = (float)f.ncand_y / (float)(f.nf_x > 0 ? f.nf_x : 1);
    out[18] = f.root ? 1.0f : 0.0f;
    out[19] = ln - lt;
    out[20] = (float)f.sup_y / (float)f.sup_e;
    out[21] = out[0] * out[3];
    out[22] = out[6] - out[5];
    out[23] = 1.0f;
}

// ------------------------------------------------------------------ child construction
// Writes child Y = X u {e} into arena slot `slot`. F is the parent's frequent-extension
// mask (dw words, LDS), items the parent's items (LDS).
__device__ void sg_make_child(const SgSearchDev& S, const SgSegment& sg, const uint64_t* __restrict__ ok,
                              const int32_t* __restrict__ item_sup, const int32_t* __restrict__ pair,
                              int32_t slot, const uint16_t* px, int k, int32_t sup_x, int nf_x,
                              int nf_x_tail, const uint64_t* F, int32_t e, int32_t sup_y, int seg,
                              float* feat_out, int64_t key, int64_t parent_key)
{
    const int dw = sg.dw;
    uint64_t* cm = S.cand + (int64_t)slot * S.dw_max;
    const uint64_t* oke = ok + sg.ok_base + (int64_t)e * dw;
    int ncand = 0, tail = 0;
    const int ew = e >> 6, eb = e & 63;
    for (int w = 0; w < dw; w++) {
        uint64_t v = F[w] & oke[w];
        if (w == ew) v &= ~(1ull << eb);
        cm[w] = v;
        const int c = __popcll(v);
        ncand += c;
        if (w < ew) tail += c;
        else if (w == ew) tail += __popcll(v & ((1ull << eb) - 1ull));
    }
    uint16_t* ci = S.items + (int64_t)slot * SG_MAX_DEPTH;
    int32_t mn = 0, mx = 0, pmin = 0x7fffffff;
    float sl = 0.0f;
    if (feat_out) {
        mn = item_sup[e];
        mx = item_sup[e];
        sl = __log2f((float)item_sup[e]);
        const int32_t* pe = pair + sg.pair_base + (int64_t)e * sg.dpad;
        for (int i = 0; i < k; i++) {
            const int32_t sx = item_sup[px[i]];
            mn = min(mn, sx);
            mx = max(mx, sx);
            sl += __log2f((float)sx);
            pmin = min(pmin, pe[px[i]]);
        }
    }
    for (int i = 0; i < k; i++) ci[i] = px[i];
    ci[k] = (uint16_t)e;
    SgNodeMeta m;
    m.k = k + 1;
    m.sup = sup_y;
    m.last = e;
    m.seg = seg;
    m.parent_sup = sup_x;
    m.parent_nf = nf_x;
    m.ncand = ncand;
    m.score = 0.0f;
    S.meta[slot] = m;
    S.key[slot] = key;
    S.pkey[slot] = parent_key;
    if (feat_out) {
        SgFeatIn f;
        f.sup_y = sup_y; f.sup_x = sup_x; f.k_y = k + 1; f.nf_x = nf_x; f.ncand_y = ncand;
        f.tail_y = tail; f.thr = sg.thr; f.n = sg.n; f.d = sg.d; f.e = e; f.sup_e = item_sup[e];
        f.min_item = mn; f.max_item = mx; f.pair_min = (pmin == 0x7fffffff ? item_sup[e] : pmin);
        f.nf_x_tail = nf_x_tail; f.root = (k == 1);
        f.sum_log_item = sl; f.mean_log_item = sl / (float)(k + 1);
        sg_features(f, feat_out);
    }
}

// Emit X into the MFI ring.
__device__ __forceinline__ void sg_emit(const SgSearchDev& S, const uint16_t* px, int k, int32_t sup,
                                        int seg, int64_t key)
{
    const int64_t pos = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_RING_POS], 1ull);
    const int64_t r = pos % S.ring_cap;
    uint16_t* ri = S.ring_items + r * SG_MAX_DEPTH;
    for (int i = 0; i < k; i++) ri[i] = px[i];
    int32_t* rm = S.ring_meta + r * 4;
    rm[0] = k;
    rm[1] = sup;
    rm[2] = seg;
    rm[3] = (int32_t)S.ctl[CTL_ROUND];
    if (S.trace_cap) {
        const int64_t t = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_TRACE_MFI], 1ull);
        if (t < S.trace_cap) S.trace_mfi[t] = key;
    }
}

// Shared per-wave expansion. F (LDS) holds the frequent one-item extensions of X,
// EQ (LDS) the equivalent ones, sup(X u {e}) == sup(X). Parent-equivalence pruning,
// exact for maximal itemsets because T(X) is contained in T(e) for e in EQ:
//   * an equivalent item outside the enumeration tail (id >= last) extends every
//     itemset of X's subtree without losing support, so the subtree holds no MFI;
//   * the equivalent tail items join every MFI of the subtree, so they are merged
//     into X (X' = X u EQ, T(X') = T(X): the counts of X serve X' unchanged) and
//     removed from F before X' is tested for maximality and expanded.
// Then X' is emitted iff F is empty, otherwise its frequent tail extensions become
// children (the reference's node step, kernel_bitmap.cu:870-889).
__device__ void sg_expand_common(const SgSearchDev& S, const SgSegment& sg, const uint64_t* ok,
                                 const int32_t* item_sup, const int32_t* pair, uint64_t* F,
                                 const uint64_t* EQ, uint16_t* px, int k, int32_t sup_x, int32_t last, int seg,
                                 const int32_t* ext_item, const int32_t* ext_cnt, int next,
                                 int64_t self_key)
{
    const int lane = threadIdx.x & 63;
    const int lw = last >> 6, lb = last & 63;
    int nf = 0, nft = 0, net = 0, nen = 0;
    for (int w = 0; w < sg.dw; w++) {
        const uint64_t low = w < lw ? ~0ull : (w == lw ? ((1ull << lb) - 1ull) : 0ull);
        const uint64_t v = F[w], q = EQ[w];
        nf += __popcll(v);
        nft += __popcll(v & low);
        net += __popcll(q & low);
        nen += __popcll(q & ~low);
    }
    if (nf == 0) {
        if (lane == 0) sg_emit(S, px, k, sup_x, seg, self_key);
        return;
    }
    if (nen > 0) return;
    if (net > 0) {
        if (k + net > SG_MAX_DEPTH) {
            if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 2ull);
            return;
        }
        if (lane == 0) {
            int kk = k;
            for (int w = 0; w < sg.dw; w++) {
                const uint64_t low = w < lw ? ~0ull : (w == lw ? ((1ull << lb) - 1ull) : 0ull);
                uint64_t q = EQ[w] & low;
                F[w] &= ~q;
                while (q) {
                    px[kk++] = (uint16_t)(w * 64 + (int)__builtin_ctzll(q));
                    q &= q - 1ull;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        k += net;
        nf -= net;
        nft -= net;
        if (nf == 0) {
            if (lane == 0) sg_emit(S, px, k, sup_x, seg, self_key);
            return;
        }
    }
    if (nft == 0) return;
    if (k >= SG_MAX_DEPTH) {
        if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 2ull);
        return;
    }
    // Reserve nft newborn entries and nft free slots with one atomic each. The round's
    // node budget (k_select) and the seed's bound (host) keep both within capacity; if
    // a counter ever overflows, nothing is written out of bounds: the reserved newborn
    // entries inside the table are marked -1 (k_push skips them) and the error flag
    // ends the run.
    int64_t base = 0, nb = 0;
    if (lane == 0) {
        nb = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_NEWBORN], (unsigned long long)nft);
        base = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_FREE_TOP], (unsigned long long)(-(int64_t)nft));
    }
    base = __shfl(base, 0, 64);
    nb = __shfl(nb, 0, 64);
    if (nb + nft > S.newborn_cap || base - nft < 0 || base > S.cap) {
        for (int t = lane; t < nft; t += 64)
            if (nb + t >= 0 && nb + t < S.newborn_cap) S.newborn[nb + t] = -1;
        if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 1ull);
        return;
    }
    int64_t kbase = 0;
    if (lane == 0) kbase = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_KEY], (unsigned long long)nft);
    kbase = __shfl(kbase, 0, 64);
    // children in ascending item order: one pass over the tail words, lane j taking bit j
    int tbase = 0;
    for (int w = 0; w < sg.dw; w++) {
        const uint64_t low = w < lw ? ~0ull : (w == lw ? ((1ull << lb) - 1ull) : 0ull);
        const uint64_t v = F[w] & low;
        if (v == 0ull) continue;
        if ((v >> lane) & 1ull) {
            const int t = tbase + __popcll(v & ((1ull << lane) - 1ull));
            const int e = w * 64 + lane;
            int32_t sup_y;
            if (ext_item) {
                // counts listed per candidate (ascending ids): locate e
                int lo = 0, hi = next - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (ext_item[mid] < e) lo = mid + 1; else hi = mid;
                }
                sup_y = ext_cnt[lo];
            } else {
                sup_y = ext_cnt[e];  // seed: pair-matrix row of the root
            }
            const int32_t slot = S.free_idx[base - 1 - t];
            float* fo = S.feat_on ? S.feat + (nb + t) * SG_FEAT : nullptr;
            sg_make_child(S, sg, ok, item_sup, pair, slot, px, k, sup_x, nf, nft, F, e, sup_y, seg, fo,
                          kbase + t, self_key);
            S.newborn[nb + t] = slot;
        }
        tbase += __popcll(v);
    }
}

// ------------------------------------------------------------------ seed (roots from the pair matrix)
// One wave per root (seg, r): F = ok[r] \ {r}; counts are the pair-matrix row.
__global__ __launch_bounds__(256) void k_seed(SgSearchDev S, const SgSegment* __restrict__ segs,
                                              const int32_t* __restrict__ pair,
                                              const uint64_t* __restrict__ ok,
                                              const int32_t* __restrict__ roots, int32_t n_roots,
                                              const int32_t* __restrict__ sup_all,
                                              const int64_t* __restrict__ seg_item_base)
{
    __shared__ uint64_t sF[4][64];
    __shared__ uint64_t sE[4][64];
    __shared__ uint16_t sX[4][SG_MAX_DEPTH];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + wv;
    const bool live = q < n_roots;
    int seg = 0, r = 0;
    sE[wv][lane] = 0ull;
    __syncthreads();
    if (live) {
        seg = roots[2 * q];
        r = roots[2 * q + 1];
        const SgSegment sg = segs[seg];
        const uint64_t* okr = ok + sg.ok_base + (int64_t)r * sg.dw;
        uint64_t v = 0ull;
        if (lane < sg.dw) {
            v = okr[lane];
            if (lane == (r >> 6)) v &= ~(1ull << (r & 63));
        }
        sF[wv][lane] = v;
        if (lane == 0) sX[wv][0] = (uint16_t)r;
        const int32_t* prow = pair + sg.pair_base + (int64_t)r * sg.dpad;
        const int32_t sr = sup_all[seg_item_base[seg] + r];
        for (int e = lane; e < sg.d; e += 64)
            if (e != r && prow[e] == sr) atomicOr((unsigned long long*)&sE[wv][e >> 6], 1ull << (e & 63));
    }
    __syncthreads();
    if (!live) return;
    const SgSegment sg = segs[seg];
    const int32_t* isup = sup_all + seg_item_base[seg];
    int64_t key = 0;
    if (lane == 0) key = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_KEY], 1ull);
    key = __shfl(key, 0, 64);
    sg_expand_common(S, sg, ok, isup, pair, sF[wv], sE[wv], sX[wv], 1, isup[r], r, seg, nullptr,
                     pair + sg.pair_base + (int64_t)r * sg.dpad, 0, key);
    if (lane == 0) atomicAdd((unsigned long long*)&S.ctl[CTL_NODES], 1ull);
}

__global__ void k_zero_newborn(SgSearchDev S)
{
    S.ctl[CTL_NEWBORN] = 0;
}

// ------------------------------------------------------------------ select
// Block-wide exclusive scan of one int64 per thread (SEL_THREADS threads).
__device__ __forceinline__ int64_t sg_block_excl_scan(int64_t v, int64_t* wsum, int64_t* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t a = 0;
        for (int w = 0; w < SEL_THREADS / 64; w++) { const int64_t t = wsum[w]; wsum[w] = a; a += t; }
        *total = a;
    }
    __syncthreads();
    const int64_t r = wsum[wv] + x - v;
    __syncthreads();
    return r;
}

// Take up to bmax nodes from the highest buckets whose candidate total fits the
// free-slot and newborn budgets; write candidate offsets and work-unit offsets.
__global__ __launch_bounds__(SEL_THREADS) void k_select(SgSearchDev S, const SgSegment* __restrict__ segs)
{
    __shared__ int32_t take[SG_NBUCKETS], start[SG_NBUCKETS + 1], tops[SG_NBUCKETS];
    __shared__ int64_t wsum[SEL_THREADS / 64];
    __shared__ int64_t s_budget, s_tot;
    __shared__ unsigned long long s_ctot, s_units, s_words;
    __shared__ int32_t s_total, s_cut;
    const int tid = threadIdx.x;
    if (tid < SG_NBUCKETS) tops[tid] = (int32_t)S.ctl[CTL_BUCKET_TOP0 + tid];
    __syncthreads();
    if (tid == 0) {
        int32_t taken = 0;
        for (int b = SG_NBUCKETS - 1; b >= 0; b--) {
            const int32_t top = tops[b];
            const int32_t t = min(top, S.bmax - taken);
            take[b] = t;
            taken += t;
        }
        int32_t acc = 0;
        for (int b = SG_NBUCKETS - 1; b >= 0; b--) { start[b] = acc; acc += take[b]; }
        start[SG_NBUCKETS] = acc;
        s_total = taken;
        s_cut = taken;
        s_ctot = 0; s_units = 0; s_words = 0;
        s_budget = min(S.ctl[CTL_FREE_TOP], S.newborn_cap);
        S.ctl[CTL_NEWBORN] = 0;
        S.ctl[CTL_UNIT_NEXT] = 0;
        S.ctl[CTL_T0] = 0x7fffffffffffffffll;
        S.ctl[CTL_T1] = 0;
    }
    __syncthreads();
    const int total = s_total;
    constexpr int PER = SG_BMAX / SEL_THREADS;
    int32_t nc[PER], nu[PER], nid[PER];
    int64_t sc = 0, su = 0;
    // phase 1: arena slots of this thread's positions (independent loads)
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = tid * PER + q;
        nid[q] = -1;
        if (i < total) {
            int b = SG_NBUCKETS - 1;
            while (!(i >= start[b] && i < start[b] + take[b])) b--;
            nid[q] = S.bucket[(int64_t)b * S.cap + (tops[b] - 1 - (i - start[b]))];
        }
    }
    // phase 2: their candidate counts and segments
    int32_t sgi[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        nc[q] = 0;
        sgi[q] = 0;
        if (nid[q] >= 0) {
            nc[q] = S.meta[nid[q]].ncand;
            sgi[q] = S.meta[nid[q]].seg;
        }
    }
    // phase 3: work units
#pragma unroll
    for (int q = 0; q < PER; q++) {
        nu[q] = 0;
        if (nid[q] >= 0) {
            const int64_t words = segs[sgi[q]].words;
            nu[q] = (int32_t)((words + S.chunk_words - 1) / S.chunk_words);
        }
        sc += nc[q];
        su += nu[q];
    }
    int64_t ec = sg_block_excl_scan(sc, wsum, &s_tot);
    int64_t eu = sg_block_excl_scan(su, wsum, &s_tot);
    const int64_t budget = s_budget;
    for (int q = 0; q < PER; q++) {
        const int i = tid * PER + q;
        if (i < total) {
            if (ec + nc[q] > budget) atomicMin(&s_cut, i);
            S.coff[i] = (int32_t)ec;
            S.unit_off[i] = eu;
            S.batch[i] = nid[q];
        }
        ec += nc[q];
        eu += nu[q];
    }
    __syncthreads();
    const int cut = s_cut;
    unsigned long long mc = 0, mu = 0, mw = 0;
    for (int q = 0; q < PER; q++) {
        const int i = tid * PER + q;
        if (i < cut) {
            mc += (unsigned long long)nc[q];
            mu += (unsigned long long)nu[q];
            mw += (unsigned long long)(nc[q] + 1) * (unsigned long long)segs[sgi[q]].words;
        }
    }
    if (mc) atomicAdd(&s_ctot, mc);
    if (mu) atomicAdd(&s_units, mu);
    if (mw) atomicAdd(&s_words, mw);
    __syncthreads();
    if (tid == 0) {
        if (cut == 0 && total > 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 1ull);
        for (int b = SG_NBUCKETS - 1; b >= 0; b--) {
            const int32_t lo = start[b], hi = start[b] + take[b];
            S.ctl[CTL_BUCKET_TOP0 + b] -= max(0, min(hi, cut) - lo);
        }
        S.coff[cut] = (int32_t)s_ctot;
        S.unit_off[cut] = (int64_t)s_units;
        S.ctl[CTL_BATCH_N] = cut;
        S.ctl[CTL_UNITS] = (int64_t)s_units;
        S.ctl[CTL_COUNTS] += (int64_t)s_ctot;
        S.ctl[CTL_WORDS] += (int64_t)s_words;
        S.ctl[CTL_NODES] += cut;
    }
}

// ------------------------------------------------------------------ candidate lists
// One wave per batch node: expand the candidate mask to an ascending id list, zero counts.
__global__ __launch_bounds__(256) void k_candlist(SgSearchDev S, const SgSegment* __restrict__ segs)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv;
    if (b >= (int)S.ctl[CTL_BATCH_N]) return;
    const int32_t slot = S.batch[b];
    const SgNodeMeta m = S.meta[slot];
    const int dw = segs[m.seg].dw;
    const uint64_t v = lane < dw ? S.cand[(int64_t)slot * S.dw_max + lane] : 0ull;
    const int c = __popcll(v);
    int x = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    int pos = S.coff[b] + x - c;
    uint64_t t = v;
    while (t) {
        const int bit = (int)__builtin_ctzll(t);
        t &= t - 1ull;
        S.clist[pos] = lane * 64 + bit;
        S.ccount[pos] = 0;
        pos++;
    }
}

// ------------------------------------------------------------------ counting
template <int CW>
__global__ __launch_bounds__(256) void k_count_search(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                      const SgSegment* __restrict__ segs, int32_t uniform)
{
    __shared__ uint64_t pre[CW];
    __shared__ int64_t s_unit;
    __shared__ int32_t s_items[SG_MAX_DEPTH];
    const int wave = threadIdx.x >> 6;
    const int nwave = blockDim.x >> 6;
    const int64_t n_units = S.ctl[CTL_UNITS];
    const int32_t nb = (int32_t)S.ctl[CTL_BATCH_N];
    if (threadIdx.x == 0 && n_units > 0)
        atomicMin((unsigned long long*)&S.ctl[CTL_T0], (unsigned long long)wall_clock64());
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_UNIT_NEXT], 1ull);
        __syncthreads();
        const int64_t u = s_unit;
        if (u >= n_units) break;
        int b, c;
        if (uniform) {
            // chunk-major order: all nodes' chunk c run together, so the candidate
            // rows' chunk c is shared in L2 across the batch
            c = (int)(u / nb);
            b = (int)(u % nb);
        } else {
            int lo = 0, hi = nb - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (S.unit_off[mid] <= u) lo = mid; else hi = mid - 1;
            }
            b = lo;
            c = (int)(u - S.unit_off[b]);
        }
        const int32_t slot = S.batch[b];
        const SgNodeMeta m = S.meta[slot];
        const SgSegment sg = segs[m.seg];
        const int64_t w0 = (int64_t)c * CW;
        const int wn = (int)min((int64_t)CW, sg.words - w0);
        if (threadIdx.x < m.k) s_items[threadIdx.x] = S.items[(int64_t)slot * SG_MAX_DEPTH + threadIdx.x];
        __syncthreads();
        const uint64_t* seg_rows = rows + sg.row_base;
        sg_build_prefix(pre, seg_rows, sg.stride, w0, wn, m.k, s_items);
        __syncthreads();
        const int32_t c0 = S.coff[b], nc = S.coff[b + 1] - c0;
        for (int j = wave; j < nc; j += nwave) {
            const int32_t e = S.clist[c0 + j];
            const uint32_t s = sg_count_row<CW>(pre, seg_rows + (int64_t)e * sg.stride + w0, wn);
            if ((threadIdx.x & 63) == 0 && s) atomicAdd(&S.ccount[c0 + j], (int32_t)s);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && n_units > 0) {
        __threadfence();
        atomicMax((unsigned long long*)&S.ctl[CTL_T1], (unsigned long long)wall_clock64());
    }
}

// ------------------------------------------------------------------ expand
__global__ __launch_bounds__(256) void k_expand(SgSearchDev S, const SgSegment* __restrict__ segs,
                                                const int32_t* __restrict__ pair,
                                                const uint64_t* __restrict__ ok,
                                                const int32_t* __restrict__ sup_all,
                                                const int64_t* __restrict__ seg_item_base)
{
    __shared__ uint64_t sF[4][64];
    __shared__ uint64_t sE[4][64];
    __shared__ uint16_t sX[4][SG_MAX_DEPTH];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv;
    const bool live = b < (int)S.ctl[CTL_BATCH_N];
    int32_t slot = 0;
    SgNodeMeta m = {};
    if (live) {
        slot = S.batch[b];
        m = S.meta[slot];
        sF[wv][lane] = 0ull;
        sE[wv][lane] = 0ull;
        if (lane < m.k) sX[wv][lane] = S.items[(int64_t)slot * SG_MAX_DEPTH + lane];
    }
    __syncthreads();
    if (!live) return;
    const SgSegment sg = segs[m.seg];
    const int32_t c0 = S.coff[b], nc = S.coff[b + 1] - c0;
    for (int j = lane; j < nc; j += 64) {
        const int32_t cnt = S.ccount[c0 + j];
        if (cnt >= sg.thr) {
            const int e = S.clist[c0 + j];
            atomicOr((unsigned long long*)&sF[wv][e >> 6], 1ull << (e & 63));
            if (cnt == m.sup) atomicOr((unsigned long long*)&sE[wv][e >> 6], 1ull << (e & 63));
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const int64_t key = S.key[slot];
    sg_expand_common(S, sg, ok, sup_all + seg_item_base[m.seg], pair, sF[wv], sE[wv], sX[wv], m.k, m.sup,
                     m.last, m.seg, S.clist + c0, S.ccount + c0, nc, key);
}

// ------------------------------------------------------------------ push newborns into buckets
__global__ void k_push(SgSearchDev S, const SgSegment* __restrict__ segs, int32_t order_mode,
                       const float* __restrict__ scores)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = S.ctl[CTL_NEWBORN];
    if (i >= n) return;
    const int32_t slot = S.newborn[i];
    if (slot < 0 || slot >= S.cap) return;
    const SgNodeMeta m = S.meta[slot];
    int b;
    if (order_mode == 1) {
        const float sc = scores[i];
        S.meta[slot].score = sc;
        b = sg_bucket_score(sc);
    } else if (order_mode == 2) {
        b = min(m.k, SG_NBUCKETS - 1);
    } else {
        b = sg_bucket_support(m.sup, segs[m.seg].thr);
    }
    const int32_t pos = (int32_t)atomicAdd((unsigned long long*)&S.ctl[CTL_BUCKET_TOP0 + b], 1ull);
    S.bucket[(int64_t)b * S.cap + pos] = slot;
    if (S.trace_cap) {
        const int64_t t = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_TRACE_BIRTH], 1ull);
        if (t < S.trace_cap) {
            S.trace_key[2 * t] = S.key[slot];
            S.trace_key[2 * t + 1] = S.pkey[slot];
            for (int f = 0; f < SG_FEAT; f++) S.trace_feat[t * SG_FEAT + f] = S.feat[i * SG_FEAT + f];
        }
    }
}

// ------------------------------------------------------------------ release batch slots
__global__ void k_release(SgSearchDev S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = (int)S.ctl[CTL_BATCH_N];
    if (i < nb) {
        const int64_t pos = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_FREE_TOP], 1ull);
        if (pos >= 0 && pos < S.cap) S.free_idx[pos] = S.batch[i];   // out of range only after an overflow
    }
    if (i == 0) {
        S.ctl[CTL_ROUND] += 1;
        int64_t front = 0;
        for (int b = 0; b < SG_NBUCKETS; b++) front += S.ctl[CTL_BUCKET_TOP0 + b];
        S.ctl[CTL_FRONTIER] = front;
    }
}

__global__ void k_frontier_sum(SgSearchDev S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int64_t front = 0;
        for (int b = 0; b < SG_NBUCKETS; b++) front += S.ctl[CTL_BUCKET_TOP0 + b];
        S.ctl[CTL_FRONTIER] = front;
    }
}

// ------------------------------------------------------------------ launchers
hipError_t sg_search_init(const SgSearchDev& S, hipStream_t s)
{
    const int64_t n = S.cap > CTL_SLOTS ? S.cap : CTL_SLOTS;
    hipLaunchKernelGGL(k_search_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S);
    return hipGetLastError();
}

static hipError_t sg_push_newborn(const SgSearchDev& S, const SgSegment* segs, int32_t order_mode,
                                  const SgScorer* scorer, hipStream_t s)
{
    if (order_mode == 1) {
        hipError_t e = sg_launch_score(scorer, S.feat, (int32_t)S.newborn_cap, &S.ctl[CTL_NEWBORN], S.scores, s);
        if (e != hipSuccess) return e;
    }
    const int64_t n = S.newborn_cap;
    hipLaunchKernelGGL(k_push, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, segs, order_mode,
                       (const float*)S.scores);
    return hipGetLastError();
}

hipError_t sg_search_seed(const SgSearchDev& S, const SgSegment* segs, const int32_t* pair,
                          const uint64_t* ok, const int32_t* roots, int32_t n_roots,
                          const int32_t* sup_all, const int64_t* seg_item_base, int32_t order_mode,
                          const SgScorer* scorer, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (n_roots > 0) {
        // newborn counter reset, then roots
        hipLaunchKernelGGL(k_zero_newborn, dim3(1), dim3(1), 0, s, S);
        hipLaunchKernelGGL(k_seed, dim3((unsigned)((n_roots + 3) / 4)), dim3(256), 0, s, S, segs, pair, ok,
                           roots, n_roots, sup_all, seg_item_base);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = sg_push_newborn(S, segs, order_mode, scorer, s)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_frontier_sum, dim3(1), dim3(64), 0, s, S);
    return hipGetLastError();
}

hipError_t sg_search_round(const SgSearchDev& S, const uint64_t* rows, const SgSegment* segs,
                           const int32_t* pair, const uint64_t* ok, const int32_t* sup_all,
                           const int64_t* seg_item_base, int32_t uniform, int32_t order_mode,
                           const SgScorer* scorer, hipEvent_t ev0, hipEvent_t ev1, hipStream_t s)
{
    hipError_t e;
    hipLaunchKernelGGL(k_select, dim3(1), dim3(SEL_THREADS), 0, s, S, segs);
    hipLaunchKernelGGL(k_candlist, dim3((unsigned)((S.bmax + 3) / 4)), dim3(256), 0, s, S, segs);
    if (ev0 && (e = hipEventRecord(ev0, s)) != hipSuccess) return e;
    switch (S.chunk_words) {
    case 256:
        hipLaunchKernelGGL(k_count_search<256>, dim3(S.count_grid), dim3(256), 0, s, S, rows, segs, uniform);
        break;
    case 512:
        hipLaunchKernelGGL(k_count_search<512>, dim3(S.count_grid), dim3(256), 0, s, S, rows, segs, uniform);
        break;
    case 2048:
        hipLaunchKernelGGL(k_count_search<2048>, dim3(S.count_grid), dim3(256), 0, s, S, rows, segs, uniform);
        break;
    default:
        hipLaunchKernelGGL(k_count_search<1024>, dim3(S.count_grid), dim3(256), 0, s, S, rows, segs, uniform);
        break;
    }
    if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((S.bmax + 3) / 4)), dim3(256), 0, s, S, segs, pair, ok,
                       sup_all, seg_item_base);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = sg_push_newborn(S, segs, order_mode, scorer, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_release, dim3((unsigned)((S.bmax + 255) / 256)), dim3(256), 0, s, S);
    return hipGetLastError();
}
