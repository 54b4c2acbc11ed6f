// Device-resident MFI search: frontier, batched counting, maximality test, MFI ring.
//
// The reference keeps its search state on the host: a mutex-protected LIFO of
// struct FI {itemsSet, frequency, root, progress} (kernel_bitmap.cu:38-43,300-318),
// one counter call per popped node, and the node step of kernel_bitmap.cu:866-890:
//   emit X iff #{e : res[e] >= thr} == |X|; otherwise push X u {i} for every dense
//   id i < progress with res[i] >= thr, with frequency res[i] and progress i.
// Here the same tree is searched with the state in HBM, by sibling groups:
//   record     an itemset X whose frequent one-item extensions U = F(X) are known and
//              whose children X u {e}, e in U below the tail bound `last`, are not yet
//              counted (items, support, U mask)
//   buckets    64 priority stacks of record indices; a round takes up to bmax records
//              from the highest non-empty buckets (LIFO inside a bucket)
//   free list  stack of free record slots
//   ring       confirmed MFIs, written to mapped host memory
// Processing a record counts all its children against all of U at once: the count
// matrix C[e][f] = |T(X) & T(e) & T(f)| (e a child, f in U) is a tile in which every
// candidate row chunk is loaded once and ANDed with up to 16 child prefixes held in
// LDS. By downward closure every frequent extension of X u {e} lies in U, so the
// reference's local maximality test (frequent-extension count == |X u {e}|) is
// exact on the tile's row; cells that cannot be frequent are skipped and stay 0: the
// pair (e, f) is infrequent, or f is not in F(P u {e}) of X's parent P (the sibling ring).
// Parent-equivalence pruning (exact): see sg_node_decide.
//
// One round = k_select -> k_candlist -> k_need_rows -> k_count_group -> k_expand_children
// -> [score] -> k_commit; every launch has a fixed shape and reads its sizes from the
// control block, so the round is captured as one hipGraph.
#include "count_core.h"
#include "f4_phase.h"

#define SEL_THREADS 1024

__device__ __forceinline__ int sg_bucket_support(int32_t sup, int32_t thr)
{
    const float r = __log2f((float)sup / (float)thr);
    int b = (int)(8.0f * r);
    return b < 0 ? 0 : (b > SG_NBUCKETS - 1 ? SG_NBUCKETS - 1 : b);
}

// Learned order: 64 buckets of the expected MFIs a record's round confirms per count,
// highest first. The network's score s estimates the fraction of the record's children its
// round confirms, so the round yields about s * ntail MFIs for ntail * (ncand - ntail / 2)
// counts; the key is log2(s) - 1.5 * log2(ncand - ntail / 2) over about 20 octaves (the
// exponent 1.5 favours cheap rounds a little more than the plain ratio does). On 500,000 x
// 256-item streams this confirms a quarter of the final MFIs in 0.76x and half in 0.94x
// the support order's time, against 0.88x and 1.05x ordering by s alone; on the 10M x
// 2048-item Zipf set the whole search takes 3% longer than by s alone. Orders that keep
// support buckets inside coarser score levels, which keep siblings together for L2, gave
// up more early coverage than they won back in speed.
__device__ __forceinline__ int sg_bucket_score(float s, int32_t ncand, int32_t ntail)
{
    const float cost = fmaxf((float)ncand - 0.5f * (float)ntail, 1.0f);
    const float key = __log2f(fmaxf(s, 1e-6f)) - 1.5f * __log2f(cost);
    const int b = (int)((key + 20.0f) * ((float)SG_NBUCKETS / 20.0f));
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


// ------------------------------------------------------------------ records
#define SG_CB 16      // children per counting tile

__device__ __forceinline__ uint64_t sg_low_mask(int w, int last)
{
    const int lw = last >> 6, lb = last & 63;
    return w < lw ? ~0ull : (w == lw ? ((1ull << lb) - 1ull) : 0ull);
}

#define SG_RC 4       // candidates held in registers per wave step

// Row-chunk ranges (units of `range` chunks of `chunk` words) of a window of `words` words.
// chunk is 128, 256 or 512 and rows hold fewer than 2^25 words, so the chunk count fits a
// 32-bit division.
__device__ __forceinline__ int32_t sg_ranges(int64_t words, int32_t chunk, int32_t range)
{
    const uint32_t nchunk = (uint32_t)((words + chunk - 1) >> __builtin_ctz((uint32_t)chunk));
    return (int32_t)((nchunk + (uint32_t)range - 1u) / (uint32_t)range);
}

__device__ __forceinline__ int32_t sg_wave_imin(int32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int32_t sg_wave_imax(int32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int32_t sg_wave_isum(int32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float sg_wave_fsum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// A sibling block at absolute offset off is intact while no later write has wrapped onto
// it: every word written so far lies below off + sib_cap.
__device__ __forceinline__ bool sg_sib_live(const SgSearchDev& S, int64_t off)
{
    return off >= 0 && S.ctl[CTL_SIB_NEXT] - off <= S.sib_cap;
}

// Offset of child e's row F(X u {e}) in the sibling block of X at off: the block's first
// row is U(X), and children are the items of U(X) in ascending order. One wave.
__device__ __forceinline__ int64_t sg_sib_row(const SgSearchDev& S, int64_t off, int32_t e, int dw)
{
    const int lane = threadIdx.x & 63;
    const uint64_t h = lane < dw ? S.sib[(off + lane) % S.sib_cap] : 0ull;
    const int w = e >> 6;
    const uint64_t part = lane < w ? h : (lane == w ? (h & ((1ull << (e & 63)) - 1ull)) : 0ull);
    return off + (int64_t)(1 + sg_wave_isum(__popcll(part))) * S.dw_max;
}

// Write record X (items px[0..k), support, tail bound `last`, extensions U in LDS) to
// slot. Called by all 64 lanes of a wave (dw <= 64 words of U, k <= 32 items: one per
// lane); lane 0 writes the metadata and the features.
__device__ void sg_make_record(const SgSearchDev& S, const SgSegment& sg, const int32_t* __restrict__ item_sup,
                               const int32_t* __restrict__ pair, int32_t slot, const uint16_t* px, int k,
                               int32_t sup, int32_t last, int seg, const uint64_t* U, int32_t sup_parent,
                               int32_t nf_parent, int32_t ntail_parent, int parent_is_root, float* feat_out,
                               int64_t key, int64_t parent_key, int64_t sib, int64_t proj, int32_t pk)
{
    const int lane = threadIdx.x & 63;
    const uint64_t v = lane < sg.dw ? U[lane] : 0ull;
    if (lane < sg.dw) S.cand[(int64_t)slot * S.dw_max + lane] = v;
    const int nu = sg_wave_isum(__popcll(v));
    const int nt = sg_wave_isum(__popcll(v & sg_low_mask(lane, last)));
    if (lane < k) S.items[(int64_t)slot * SG_MAX_DEPTH + lane] = px[lane];
    // features: minimum / maximum / log-sum of the item supports, smallest pair support with e
    int32_t mn = 0, mx = 0, pmin = 0;
    float sl = 0.0f;
    if (feat_out) {
        const int e = last;
        const int32_t x = lane < k ? (int32_t)px[lane] : -1;
        const int32_t sx = x >= 0 ? item_sup[x] : 0;
        mn = sg_wave_imin(x >= 0 ? sx : 0x7fffffff);
        mx = sg_wave_imax(x >= 0 ? sx : 0);
        sl = sg_wave_fsum(x >= 0 ? __log2f((float)sx) : 0.0f);
        pmin = sg_wave_imin((x >= 0 && x != e) ? pair[sg.pair_base + (int64_t)e * sg.dpad + x] : 0x7fffffff);
    }
    if (lane != 0) return;
    SgNodeMeta m;
    m.k = k;
    m.sup = sup;
    m.last = last;
    m.seg = seg;
    m.parent_sup = sup_parent;
    m.parent_nf = nf_parent;
    m.ncand = nu;
    m.ntail = nt;
    m.nblk = (nt + SG_CB - 1) / SG_CB;
    m.nrange = sg_ranges(sg.words, S.chunk_words, S.chunk_range);
    m.score = 0.0f;
    m.sib = sib;
    m.proj = proj;
    m.pk = pk;
    m.pad_ = 0;
    S.meta[slot] = m;
    S.key[slot] = key;
    S.pkey[slot] = parent_key;
    if (feat_out) {
        const int e = last;
        mn = min(mn, item_sup[e]);
        mx = max(mx, item_sup[e]);
        SgFeatIn f;
        f.sup_y = sup; f.sup_x = sup_parent > 0 ? sup_parent : sup; f.k_y = k; f.nf_x = nf_parent; f.ncand_y = nu;
        f.tail_y = nt; f.thr = sg.thr; f.n = sg.n; f.d = sg.d; f.e = e; f.sup_e = item_sup[e];
        f.min_item = mn; f.max_item = mx; f.pair_min = (pmin == 0x7fffffff ? item_sup[e] : pmin);
        f.nf_x_tail = ntail_parent; f.root = parent_is_root;
        f.sum_log_item = sl; f.mean_log_item = sl / (float)k;
        sg_features(f, feat_out);
    }
}

// Write MFI X at ring position pos (pinned host memory mapped into the device). Called by
// the whole wave: the items go out as one coalesced store (lane i writes item i) and the
// four meta fields as one 16-byte store, so an MFI costs two writes across the host link
// instead of one per item and field.
__device__ __forceinline__ void sg_emit_at(const SgSearchDev& S, int64_t pos, const uint16_t* px, int k, int32_t sup,
                                           int seg, int64_t parent_key)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = pos % S.ring_cap;
    uint16_t* ri = S.ring_items + r * SG_MAX_DEPTH;
    if (lane < k) ri[lane] = px[lane];
    if (lane != 0) return;
    int4 mv;
    mv.x = k;
    mv.y = sup;
    mv.z = seg;
    mv.w = (int32_t)S.ctl[CTL_ROUND];
    *(int4*)(S.ring_meta + r * 4) = mv;
    if (S.trace_cap) {
        const int64_t t = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_TRACE_MFI], 1ull);
        if (t < S.trace_cap) S.trace_mfi[t] = parent_key;
    }
}

enum { SG_ACT_NONE = 0, SG_ACT_EMIT = 1, SG_ACT_RECORD = 2 };

// Node step for one wave: X = px[0..k) with tail bound `last`, frequent one-item
// extensions F and equivalent ones EQ (sup(X u {e}) == sup(X)) in LDS.
// Parent-equivalence pruning, exact for maximal itemsets because T(X) is contained in
// T(e) for e in EQ:
//   * an equivalent item outside the enumeration tail (id >= last) extends every
//     itemset of X's subtree without losing support, so the subtree holds no MFI;
//   * the equivalent tail items join every MFI of the subtree, so they are merged into
//     X (X' = X u EQ, T(X') = T(X): the counts of X serve X' unchanged; F loses them).
// Then X' is an MFI iff F is empty (the reference's test, kernel_bitmap.cu:870);
// otherwise, if it has frequent tail extensions, it becomes one frontier record.
// Returns the action; k is updated to |X'|.
__device__ int sg_node_decide(const SgSearchDev& S, const SgSegment& sg, uint64_t* F, const uint64_t* EQ,
                              uint16_t* px, int& k, int32_t last)
{
    const int lane = threadIdx.x & 63;
    int nf = 0, nft = 0, net = 0, nen = 0;
    for (int w = 0; w < sg.dw; w++) {
        const uint64_t low = sg_low_mask(w, last);
        const uint64_t v = F[w], q = EQ[w];
        nf += __popcll(v);
        nft += __popcll(v & low);
        net += __popcll(q & low);
        nen += __popcll(q & ~low);
    }
    if (nf == 0) return SG_ACT_EMIT;
    if (nen > 0) return SG_ACT_NONE;
    if (net > 0) {
        if (k + net > SG_MAX_DEPTH) {
            if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 2ull);
            return SG_ACT_NONE;
        }
        if (lane == 0) {
            int kk = k;
            for (int w = 0; w < sg.dw; w++) {
                uint64_t q = EQ[w] & sg_low_mask(w, last);
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
        if (nf == 0) return SG_ACT_EMIT;
    }
    if (nft == 0) return SG_ACT_NONE;
    if (k >= SG_MAX_DEPTH) {   // its children would exceed the longest supported itemset
        if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 2ull);
        return SG_ACT_NONE;
    }
    return SG_ACT_RECORD;
}

// Take record slot number `rank` of a reservation: newborn entry nb, free-list top
// `free_top` (the value before the reservation). The round budget (k_select) and the
// seed bound (host) keep both in range; an overflow writes nothing out of bounds and
// stops the run with the error flag. Returns the slot or -1.
__device__ __forceinline__ int32_t sg_take_slot(const SgSearchDev& S, int64_t nb, int64_t free_top)
{
    if (nb >= S.newborn_cap || free_top < 1 || free_top > S.cap) {
        if (nb < S.newborn_cap) S.newborn[nb] = -1;
        atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 1ull);
        return -1;
    }
    return S.free_idx[free_top - 1];
}

// ------------------------------------------------------------------ seed (roots from the pair matrix)
// One wave per root (seg, r): F = ok[r] \ {r}; equivalents from the pair-matrix row. MFI
// ring positions, record slots and keys are reserved once per workgroup of 16 roots.
#define SG_SEED_WAVES 16
__global__ __launch_bounds__(SG_SEED_WAVES * 64) void k_seed(SgSearchDev S, const SgSegment* __restrict__ segs,
                                                             const int32_t* __restrict__ pair,
                                                             const uint64_t* __restrict__ ok,
                                                             const int32_t* __restrict__ roots, int32_t n_roots,
                                                             const int32_t* __restrict__ sup_all,
                                                             const int64_t* __restrict__ seg_item_base)
{
    __shared__ uint64_t sF[SG_SEED_WAVES][64];
    __shared__ uint64_t sE[SG_SEED_WAVES][64];
    __shared__ uint16_t sX[SG_SEED_WAVES][SG_MAX_DEPTH];
    __shared__ int32_t s_act[SG_SEED_WAVES];
    __shared__ int64_t s_ring, s_nb, s_top, s_key;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * SG_SEED_WAVES + wv;
    const bool live = q < n_roots;
    int seg = 0, r = 0, act = SG_ACT_NONE, k = 1;
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
    if (live) act = sg_node_decide(S, segs[seg], sF[wv], sE[wv], sX[wv], k, r);   // k grows by merged items
    if (lane == 0) s_act[wv] = act;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ne = 0, nr = 0, nl = 0;
        for (int w = 0; w < SG_SEED_WAVES; w++) {
            ne += s_act[w] == SG_ACT_EMIT;
            nr += s_act[w] == SG_ACT_RECORD;
            nl += blockIdx.x * SG_SEED_WAVES + w < n_roots;
        }
        atomicAdd((unsigned long long*)&S.ctl[CTL_NODES], (unsigned long long)nl);
        if (ne) s_ring = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_RING_POS], (unsigned long long)ne);
        if (nr) {
            s_nb = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_NEWBORN], (unsigned long long)nr);
            s_top = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_FREE_TOP], (unsigned long long)(-(int64_t)nr));
            s_key = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_KEY], (unsigned long long)nr);
        }
    }
    __syncthreads();
    if (act == SG_ACT_NONE) return;   // wave-uniform
    int rank = 0;
    for (int w = 0; w < wv; w++) rank += s_act[w] == act;
    const SgSegment sg = segs[seg];
    const int32_t* isup = sup_all + seg_item_base[seg];
    if (act == SG_ACT_EMIT) {
        sg_emit_at(S, s_ring + rank, sX[wv], k, isup[r], seg, -1);
    } else {
        const int64_t nb = s_nb + rank;
        int32_t slot = -1;
        if (lane == 0) slot = sg_take_slot(S, nb, s_top - rank);
        slot = __shfl(slot, 0, 64);
        if (slot >= 0) {
            float* fo = S.feat_on ? S.feat + nb * SG_FEAT : nullptr;
            sg_make_record(S, sg, isup, pair, slot, sX[wv], k, isup[r], r, seg, sF[wv], 0, 0, 0, 0, fo, s_key + rank, -1, -1,
                           -1, 0);
            if (lane == 0) S.newborn[nb] = slot;
        }
    }
}

__global__ void k_zero_newborn(SgSearchDev S)
{
    S.ctl[CTL_NEWBORN] = 0;
}

// ------------------------------------------------------------------ select
// Block-wide exclusive scans of N int64s per thread at once (SEL_THREADS threads, N no
// more than the block's waves). The values go through LDS transposed: wave n scans
// sequence n, each lane 16 consecutive entries (a serial prefix in registers, then one
// wave scan of the lane totals), so the block pays one 6-step wave scan per sequence
// instead of N per wave. sc holds N x (SEL_THREADS + SEL_THREADS / 16) entries (one pad
// per 16 spreads a lane's run over the banks); total[n] receives sequence n's sum.
#define SEL_SCAN_STRIDE (SEL_THREADS + SEL_THREADS / 16)
template <int N>
__device__ __forceinline__ void sg_block_excl_scan(int64_t (&v)[N], int64_t* sc, int64_t* total)
{
    static_assert(N <= SEL_THREADS / 64 && SEL_THREADS == 1024, "one wave per sequence, 16 entries per lane");
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int at = tid + (tid >> 4);
#pragma unroll
    for (int n = 0; n < N; n++) sc[n * SEL_SCAN_STRIDE + at] = v[n];
    __syncthreads();
    if (wv < N) {
        int64_t* q = sc + wv * SEL_SCAN_STRIDE + 17 * lane;   // entries 16 lane .. 16 lane + 15
        int64_t x[16];
        int64_t run = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            x[j] = run;
            run += q[j];
        }
        int64_t incl = run;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const int64_t base = incl - run;
#pragma unroll
        for (int j = 0; j < 16; j++) q[j] = base + x[j];
        if (lane == 63) total[wv] = incl;
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < N; n++) v[n] = sc[n * SEL_SCAN_STRIDE + at];
}

// ------------------------------------------------------------------ projected blocks
// A block p is usable this round when neither ring can wrap onto it before the round ends:
// the round allocates at most SG_BMAX descriptors and proj_round_words arena words, and
// (pn0, an0) are the ring positions before it. The descriptor of a usable block is p's.
// d_off is the arena position of block p's descriptor slot, read ahead.
__device__ __forceinline__ bool sg_block_usable(const SgSearchDev& S, int64_t p, int64_t pn0, int64_t an0,
                                                int64_t d_off)
{
    if (p < 0 || p >= pn0 || pn0 + SG_BMAX - p > S.pdesc_cap) return false;
    return an0 + S.proj_round_words - d_off <= S.parena_cap;
}

// Row of item f in block p (its rank in U of the block's itemset). Every item a record
// of the block's subtree names is in that U (U shrinks down the tree, and every later item
// of the record came from it); an item outside it would address past the block, so it
// flags an error (the search stops with SGFIM_E_DEVICE) and row 0 is used instead.
__device__ __forceinline__ int32_t sg_block_row(const SgSearchDev& S, int64_t p, int32_t f)
{
    const int64_t b = (p % S.pdesc_cap) * S.dw_max + (f >> 6);
    const uint64_t mk = S.pmask[b];
    if (f < 0 || f >= S.d_max || !((mk >> (f & 63)) & 1ull)) {
        atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
        return 0;
    }
    return S.prank[b] + __popcll(mk & ((1ull << (f & 63)) - 1ull));
}

// Words of a block over n transactions, and its row stride.
__device__ __forceinline__ int64_t sg_block_words(int64_t n) { return (n + 63) >> 6; }
__device__ __forceinline__ int64_t sg_block_stride(int64_t n) { return (sg_block_words(n) + 7) & ~7ll; }

// How one taken record is counted this round. The record counts on its inherited block
// while that block is usable, else on its window's rows; with projection on, a record
// whose support is at most 1/proj_factor of its view's transactions gets a block of its
// own, projected from that view this round, and counts on it.
struct SgPlan {
    int64_t view;      // block counted on (-1: the window's rows); a new block: its id
    int64_t src;       // a new block's source view (-1: the window's rows)
    int64_t words;     // words of the view's rows
    int64_t src_words; // words of the source view (new blocks)
    int64_t blk;       // arena words of a new block
    int64_t list;      // a new block's ptid entries: its transactions (gathered) or one per compress unit
    int64_t ents;      // a compacted new block's word entries (at most one per source word)
    int64_t units;     // a new block's gather or compress units
    int32_t pnew;      // a new block: 1 gathered bit by bit, 2 compacted word by word
    int32_t nr;        // row-chunk ranges of the view
};

// What the plan reads of a record, its window and its block's descriptor slot.
struct SgSelIn {
    int64_t proj;       // the record's block (-1: none)
    int64_t sg_words;   // words of the window's rows
    int64_t d_off;      // descriptor slot of the block: arena position, words, transactions
    int64_t d_words;
    int32_t sup, ncand, ntail, nblk;
    int32_t sg_n, d_n;
};

__device__ __forceinline__ SgPlan sg_plan_record(const SgSearchDev& S, const SgSelIn& m, int64_t pn0, int64_t an0)
{
    SgPlan P;
    P.view = -1;
    P.src = -1;
    P.words = m.sg_words;
    P.src_words = m.sg_words;
    P.blk = 0;
    P.list = 0;
    P.ents = 0;
    P.units = 0;
    P.pnew = 0;
    int64_t n_view = m.sg_n;
    if (S.proj_factor > 0) {
        if (sg_block_usable(S, m.proj, pn0, an0, m.d_off)) {
            P.view = m.proj;
            P.words = m.d_words;
            n_view = m.d_n;
        }
        // A record with one child gets no block of its own: the child can project itself
        // at no more cost when it is counted, and needs none if it is an MFI. A dense block
        // (built word by word: ~40 lane operations per source word and row, against 4 for
        // one count over that word) needs proj_mintail children to repay its build.
        const int64_t blk = (int64_t)m.ncand * sg_block_stride(m.sup);
        const bool dense = S.proj_dense > 0 && (int64_t)m.sup * S.proj_dense >= n_view;
        if (m.ntail >= (dense ? max(S.proj_mintail, 2) : 2) && (int64_t)m.sup * S.proj_factor <= n_view &&
            blk <= S.proj_max_block && m.sup <= S.ptid_cap) {
            P.src = P.view;
            P.src_words = P.words;
            P.blk = blk;
            P.words = sg_block_words(m.sup);
            P.view = -2;   // its id is assigned from the scan
            const int64_t ents = min((int64_t)m.sup, P.src_words);
            const int64_t cu = (P.words + SG_PC_WORDS - 1) / SG_PC_WORDS;
            if (dense && ents <= S.pent_cap && P.src_words < ((int64_t)1 << 24)) {
                P.pnew = 2;
                P.list = cu;
                P.ents = ents;
                P.units = cu * ((m.ncand + SG_PC_ROWS - 1) / SG_PC_ROWS);
            } else {
                P.pnew = 1;
                P.list = m.sup;
                P.units = ((int64_t)m.sup + SG_PG_TIDS - 1) / SG_PG_TIDS * ((m.ncand + SG_PG_ROWS - 1) / SG_PG_ROWS);
            }
        }
    }
    P.nr = sg_ranges(P.words, S.chunk_words, S.chunk_range);
    return P;
}

// Records per load group of the selection passes, and the group's loads: the fields the
// plan reads of the records nid[0 .. SEL_GROUP) (slot 0 for absent ones, nid < 0), then of
// their windows and block descriptor slots (window 0 and slot 0 for absent ones; no
// descriptors with projection off). No load waits on a branch of its own record, so the
// group's loads are in flight together. pn0m = pn0 % pdesc_cap: a usable block p lies
// less than pdesc_cap before pn0, so its slot is pn0m - (pn0 - p), wrapped once.
#define SEL_GROUP 2
__device__ __forceinline__ void sel_load_group(const SgSearchDev& S, const SgSegment* __restrict__ segs,
                                               const int32_t* nid, int64_t pn0, int64_t pn0m,
                                               SgSelIn (&m)[SEL_GROUP])
{
    int32_t seg[SEL_GROUP];
#pragma unroll
    for (int c = 0; c < SEL_GROUP; c++) {
        const SgNodeMeta* mp = S.meta + max(nid[c], 0);
        seg[c] = nid[c] >= 0 ? mp->seg : 0;
        m[c].proj = nid[c] >= 0 ? mp->proj : -1;
        m[c].sup = mp->sup;
        m[c].ncand = mp->ncand;
        m[c].ntail = mp->ntail;
        m[c].nblk = mp->nblk;
    }
#pragma unroll
    for (int c = 0; c < SEL_GROUP; c++) {
        m[c].sg_words = segs[seg[c]].words;
        m[c].sg_n = segs[seg[c]].n;
        m[c].d_off = 0;
        m[c].d_words = 0;
        m[c].d_n = 0;
        if (S.proj_factor > 0) {   // uniform
            const int64_t back = pn0 - m[c].proj;   // blocks since p (p >= 0 and this < pdesc_cap if usable)
            int64_t slot = 0;
            if (m[c].proj >= 0 && back > 0 && back < S.pdesc_cap) slot = pn0m - back + (pn0m < back ? S.pdesc_cap : 0);
            const SgProj* dp = S.pdesc + slot;
            m[c].d_off = dp->off;
            m[c].d_words = dp->words;
            m[c].d_n = dp->n;
        }
    }
}

// Take up to bmax records from the highest buckets, as many as fit the budgets: new
// records (at most one per child) within the free slots and newborn table, count
// cells within the cell table, new blocks within the round's transaction-list and arena
// budgets. Writes the U-list, child, cell and unit offsets and each record's view.
// Counting units: records on window rows run range-major across the batch (units
// [0, UNITS_A): all blocks of range 0, then range 1, ...), records on blocks record-major
// after them (a record's blocks, range by range). Sizes are parked in the offset arrays
// between the passes (each thread rereads only its own entries).
__global__ __launch_bounds__(SEL_THREADS) void k_select(SgSearchDev S, const SgSegment* __restrict__ segs)
{
    constexpr int NQ = 13;
    __shared__ int32_t tops[SG_NBUCKETS], ends[SG_NBUCKETS];
    __shared__ int64_t sc[NQ * SEL_SCAN_STRIDE];
    __shared__ int64_t tot[NQ];
    __shared__ int64_t s_budget;
    __shared__ int32_t s_total, s_cut, s_maxr;
    const int tid = threadIdx.x;
    static_assert(SG_NBUCKETS == 64, "one lane per bucket");
    if (tid < SG_NBUCKETS) {
        // lane o holds order o = 63 - b, the highest bucket first; ends[o] = records taken
        // through o = min(records in buckets of order <= o, bmax), an inclusive scan
        const int o = tid, b = SG_NBUCKETS - 1 - o;
        const int32_t t = (int32_t)S.ctl[CTL_BUCKET_TOP0 + b];
        tops[b] = t;
        int32_t cum = t;
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t y = __shfl_up(cum, d, 64);
            if (o >= d) cum += y;
        }
        const int32_t end = min(cum, S.bmax);
        ends[o] = end;
        const int32_t taken = __shfl(end, SG_NBUCKETS - 1, 64);
        if (o == 0) {
            s_total = taken;
            s_cut = taken;
        }
    }
    if (tid == 0) {
        s_maxr = 0;
        s_budget = min(S.ctl[CTL_FREE_TOP], S.newborn_cap);
        S.ctl[CTL_ROUND] += 1;   // MFIs of this round carry its number; the seed's carry 0
        S.ctl[CTL_NEWBORN] = 0;
        S.ctl[CTL_UNIT_NEXT] = 0;
        S.ctl[CTL_T0] = 0x7fffffffffffffffll;
        S.ctl[CTL_T1] = 0;
    }
    const int64_t pn0 = S.ctl[CTL_PROJ_NEXT], an0 = S.ctl[CTL_ARENA_NEXT];
    const int64_t pn0m = S.proj_factor > 0 ? pn0 % S.pdesc_cap : 0;
    __syncthreads();
    const int total = s_total;
    constexpr int PER = SG_BMAX / SEL_THREADS;
    // thread tid takes batch positions tid * pe .. tid * pe + pe - 1: contiguous, so the
    // scans below stay in batch order, and spread over every thread of the block
    const int pe = (total + SEL_THREADS - 1) / SEL_THREADS;
    // pass 1: record slots, then sizes. A thread's records go in groups of four whose loads
    // (metas, then their segments and block descriptors; absent records read slot 0) are
    // all issued before any is used, so a group waits on memory twice, not twice per record.
    int32_t nid[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = tid * pe + q;
        nid[q] = -1;
        if (q < pe && i < total) {
            int lo = 0, hi = SG_NBUCKETS - 1;   // first order o with ends[o] > i
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ends[mid] > i) hi = mid; else lo = mid + 1;
            }
            const int b = SG_NBUCKETS - 1 - lo;
            const int32_t pos = i - (lo > 0 ? ends[lo - 1] : 0);
            nid[q] = S.bucket[(int64_t)b * S.cap + (tops[b] - 1 - pos)];
        }
    }
    // sums: U items, children, cells, window-row blocks, sibling words, block units,
    // list entries, arena words, new blocks, transaction-list units, gather units, word
    // entries, compress units
    int64_t sm[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) sm[n] = 0;
    int32_t mr = 0;   // longest window (row-chunk ranges) among the taken records on window rows
#pragma unroll
    for (int g = 0; g < PER; g += SEL_GROUP) {
        if (g >= pe) break;   // block-uniform
        SgSelIn m[SEL_GROUP];
        sel_load_group(S, segs, nid + g, pn0, pn0m, m);
#pragma unroll
        for (int c = 0; c < SEL_GROUP; c++) {
            if (nid[g + c] < 0) continue;
            const SgPlan P = sg_plan_record(S, m[c], pn0, an0);
            const bool onb = P.view != -1;
            if (!onb) mr = max(mr, P.nr);
            sm[0] += m[c].ncand;
            sm[1] += m[c].ntail;
            sm[2] += (int64_t)m[c].ncand * m[c].ntail;
            sm[3] += onb ? 0 : m[c].nblk;
            sm[4] += (int64_t)(1 + m[c].ntail) * S.dw_max;
            sm[5] += onb ? (int64_t)m[c].nblk * P.nr : 0;
            if (P.pnew) {
                sm[6] += P.list;
                sm[7] += P.blk;
                sm[8] += 1;
                sm[9] += (P.src_words + SG_PT_WORDS - 1) / SG_PT_WORDS;
                sm[P.pnew == 2 ? 12 : 10] += P.units;
                sm[11] += P.ents;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PER; q++)
        if (nid[q] >= 0) S.batch[tid * pe + q] = nid[q];
    int64_t ex[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) ex[n] = sm[n];
    sg_block_excl_scan<NQ>(ex, sc, tot);
    const int64_t sib_base = S.ctl[CTL_SIB_NEXT];
    // pass 2: offsets; the cut is the first record past a budget (offsets only grow)
    const int64_t budget = s_budget;
#pragma unroll
    for (int g = 0; g < PER; g += SEL_GROUP) {
        if (g >= pe) break;   // block-uniform
        SgSelIn mg[SEL_GROUP];
        sel_load_group(S, segs, nid + g, pn0, pn0m, mg);
#pragma unroll
        for (int c = 0; c < SEL_GROUP; c++) {
            const int q = g + c;
            const int i = tid * pe + q;
            if (nid[q] >= 0) {
                const SgSelIn& m = mg[c];
                SgPlan P = sg_plan_record(S, m, pn0, an0);
                const int32_t nu = m.ncand, nt = m.ntail;
                const int64_t cells = (int64_t)nu * nt;
                if (ex[1] + nt > budget || ex[2] + cells > S.cell_cap ||
                    (P.pnew && (ex[6] + P.list > S.ptid_cap || ex[7] + P.blk > S.proj_round_words ||
                                ex[11] + P.ents > S.pent_cap)))
                    atomicMin(&s_cut, i);
                S.coff[i] = (int32_t)ex[0];
                S.choff[i] = (int32_t)ex[1];
                S.celloff[i] = ex[2];
                S.unit_off[i] = ex[3];
                // a block that would wrap onto another of this round gets none
                const int64_t sz = (int64_t)(1 + nt) * S.dw_max;
                S.sib_off[i] = ex[4] + sz <= S.sib_cap ? sib_base + ex[4] : -1;
                S.bunitB[i] = ex[5];
                S.bunitT[i] = ex[9];
                S.bunitG[i] = ex[10];
                S.bunitC[i] = ex[12];
                // ring positions before this record (the allocation at the cut); a new block
                // gets descriptor pn0 + ex[8] and arena words an0 + ex[7] onwards, and its
                // descriptor is written by k_candlist for the records before the cut only
                S.bpd[i] = pn0 + ex[8];
                S.bpa[i] = an0 + ex[7];
                if (P.pnew) {
                    P.view = pn0 + ex[8];
                    S.ptoff[i] = ex[6];
                    S.peoff[i] = ex[11];
                }
                S.bview[i] = P.view;
                S.bsrc[i] = P.src;
                S.bpnew[i] = P.pnew;
                S.bnr[i] = P.nr;
                ex[0] += nu;
                ex[1] += nt;
                ex[2] += cells;
                ex[3] += P.view == -1 ? m.nblk : 0;
                ex[4] += sz;
                ex[5] += P.view == -1 ? 0 : (int64_t)m.nblk * P.nr;
                if (P.pnew) {
                    ex[6] += P.list;
                    ex[7] += P.blk;
                    ex[8] += 1;
                    ex[9] += (P.src_words + SG_PT_WORDS - 1) / SG_PT_WORDS;
                    ex[P.pnew == 2 ? 12 : 10] += P.units;
                    ex[11] += P.ents;
                }
            }
        }
    }
    // window-row units = their blocks x the longest window's ranges (taken over every taken
    // record: a longer record past the cut only adds units that every workgroup skips)
    if (mr) atomicMax(&s_maxr, mr);
    __syncthreads();
    if (tid < SG_NBUCKETS) {   // release the taken records from their buckets
        const int cut = s_cut;
        const int o = tid, b = SG_NBUCKETS - 1 - o;
        const int32_t lo = o > 0 ? ends[o - 1] : 0, hi = ends[o];
        const int32_t n = max(0, min(hi, cut) - lo);
        if (n) S.ctl[CTL_BUCKET_TOP0 + b] -= n;
    }
    if (tid == 0) {
        const int cut = s_cut;
        if (cut == total) {   // every taken record fits: the end offsets are the totals
            S.coff[cut] = (int32_t)tot[0];
            S.choff[cut] = (int32_t)tot[1];
            S.celloff[cut] = tot[2];
            S.unit_off[cut] = tot[3];
            S.bunitB[cut] = tot[5];
            S.bunitT[cut] = tot[9];
            S.bunitG[cut] = tot[10];
            S.bunitC[cut] = tot[12];
        }
        if (cut == 0 && total > 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 1ull);
        const int64_t nch = S.choff[cut];
        S.ctl[CTL_BATCH_N] = cut;
        S.ctl[CTL_CHILDREN] = nch;
        S.ctl[CTL_CELLS] = S.celloff[cut];
        S.ctl[CTL_BLOCKS] = S.unit_off[cut];
        S.ctl[CTL_UNITS_A] = S.unit_off[cut] * (int64_t)s_maxr;
        S.ctl[CTL_UNITS] = S.unit_off[cut] * (int64_t)s_maxr + S.bunitB[cut];
        S.ctl[CTL_PROJ_TUNITS] = S.bunitT[cut];
        S.ctl[CTL_PROJ_GUNITS] = S.bunitG[cut];
        S.ctl[CTL_PROJ_CUNITS] = S.bunitC[cut];
        S.ctl[CTL_NODES] += nch;
        S.ctl[CTL_SIB_NEXT] = S.ctl[CTL_SIB_NEXT] + tot[4];   // blocks of records past the cut stay unused
        // the rings advance by the allocations of the records before the cut
        S.ctl[CTL_PROJ_NEXT] = cut == total ? pn0 + tot[8] : S.bpd[cut];
        S.ctl[CTL_ARENA_NEXT] = cut == total ? an0 + tot[7] : S.bpa[cut];
    }
}

// ------------------------------------------------------------------ U lists
// One wave per record: expand U into an ascending id list; the children are its
// first ntail entries (U is ascending and children are the items below `last`). Each
// entry also gets its row in the record's view (crow): the id itself on window rows, its
// rank in U of the block's itemset on a block (for a block projected this round that is
// the entry's own position). The wave maps its children and, on window rows, its child
// blocks back to the record (chrec, blkrec), so the counting and expansion kernels find
// a unit's record with one load. A record that gets a new block writes the block's
// descriptor, its U mask and rank prefix, and clears its transaction-list fill.
// The grid also clears the batch's count cells.
__global__ __launch_bounds__(256) void k_candlist(SgSearchDev S, const SgSegment* __restrict__ segs)
{
    const int64_t ncell = S.ctl[CTL_CELLS];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += (int64_t)gridDim.x * blockDim.x)
        S.cells[i] = 0;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv;
    if (b >= (int)S.ctl[CTL_BATCH_N]) return;
    const int32_t slot = S.batch[b];
    const SgNodeMeta m = S.meta[slot];
    const int dw = segs[m.seg].dw;
    const uint64_t v = lane < dw ? S.cand[(int64_t)slot * S.dw_max + lane] : 0ull;
    const int64_t sb = S.sib_off[b];
    if (sb >= 0 && lane < S.dw_max) S.sib[(sb + lane) % S.sib_cap] = v;   // U(X) heads X's sibling block
    const int c = __popcll(v);
    int x = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    const int64_t view = S.bview[b];
    const int pnew = S.bpnew[b];
    if (pnew) {   // the record's new block (descriptor slot and arena words allocated by k_select)
        const int64_t ds = view % S.pdesc_cap;
        if (lane < S.dw_max) {
            S.pmask[ds * S.dw_max + lane] = v;
            S.prank[ds * S.dw_max + lane] = x - c;
        }
        if (lane == 0) {
            SgProj D;
            D.off = S.bpa[b];
            D.words = sg_block_words(m.sup);
            D.stride = sg_block_stride(m.sup);
            D.n = m.sup;
            D.nrows = m.ncand;
            S.pdesc[ds] = D;
            S.pfill[b] = 0;
            // words the block costs: X's unapplied rows and the U rows' words touched at T(X)
            // in the source view (read), the block's rows (written)
            const int64_t src = S.bsrc[b];
            const int64_t sw = src >= 0 ? S.pdesc[src % S.pdesc_cap].words : segs[m.seg].words;
            const int kk = m.k - (src >= 0 ? m.pk : 0);
            const int64_t w = (int64_t)kk * sw + (int64_t)m.ncand * (min(sw, (int64_t)m.sup) + D.words);
            atomicAdd((unsigned long long*)&S.ctl[CTL_PROJS], 1ull);
            atomicAdd((unsigned long long*)&S.ctl[CTL_PROJ_WORDS], (unsigned long long)w);
            if (src < 0) atomicAdd((unsigned long long*)&S.ctl[CTL_PROJ_ROWS], 1ull);
            atomicAdd((unsigned long long*)&S.ctl[CTL_PROJ_ARENA], (unsigned long long)((int64_t)m.ncand * D.stride));
        }
    }
    // rank of this lane's first item in U of the view's block (inherited blocks)
    int32_t rk = 0;
    uint64_t bm = 0ull;
    if (view >= 0 && !pnew && lane < S.dw_max) {
        const int64_t q = (view % S.pdesc_cap) * S.dw_max + lane;
        rk = S.prank[q];
        bm = S.pmask[q];
    }
    int pos = S.coff[b] + x - c;
    uint64_t t = v;
    while (t) {
        const int bit = (int)__builtin_ctzll(t);
        t &= t - 1ull;
        const int32_t f = lane * 64 + bit;
        S.clist[pos] = f;
        int32_t r = f;
        if (pnew) {
            r = pos - S.coff[b];
        } else if (view >= 0) {
            r = rk + __popcll(bm & ((1ull << bit) - 1ull));
            if (!((bm >> bit) & 1ull)) {   // not in the block's U (see sg_block_row)
                atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
                r = 0;
            }
        }
        S.crow[pos] = r;
        pos++;
    }
    const int32_t c0 = S.choff[b];
    for (int j = lane; j < m.ntail; j += 64) {
        S.ccount[c0 + j] = 0;
        S.chrec[c0 + j] = b;
    }
    if (view == -1) {
        const int64_t u0 = S.unit_off[b];
        for (int i = lane; i < m.nblk; i += 64) S.blkrec[u0 + i] = b;
    }
}

// Largest batch position g < n with off[g] <= v, for a non-decreasing prefix off (the
// record whose range of units holds v; records without units share their successor's
// offset and are never returned for a v inside a range).
__device__ __forceinline__ int sg_find_record(const int64_t* __restrict__ off, int n, int64_t v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ projection
// Transaction lists of the round's new blocks: T(X) of each such record over its source
// view. Units are (record, SG_PT_WORDS source words); each wave takes 256 of them as 4
// steps of 64 coalesced words and ANDs X's rows the source has not applied. A gathered
// block's list holds the set bits' positions (ascending within the wave), appended at a
// position reserved with one atomic per workgroup. A compacted block's list holds one
// entry per nonzero word (SgPWord: the word, its mask and the block bit of its first
// transaction); one atomic per workgroup on the packed counter (entries << 40 |
// transactions) reserves the entries and bits together, so entries are in block-bit
// order. The entry holding the first bit of
// each compress unit is recorded in ptid. Every row of the block is then built in this
// list's order, so the order across waves is free.
__global__ __launch_bounds__(256) void k_proj_tids(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                  const SgSegment* __restrict__ segs)
{
    __shared__ int32_t s_row[SG_MAX_DEPTH];
    __shared__ int s_g;
    __shared__ int64_t s_local;
    __shared__ unsigned long long s_req[4], s_base[4];
    const int64_t nunits = S.ctl[CTL_PROJ_TUNITS];
    const int nb = (int)S.ctl[CTL_BATCH_N];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {   // workgroup-uniform
        __syncthreads();   // the previous unit is done with the LDS
        if (threadIdx.x == 0) {
            const int g = sg_find_record(S.bunitT, nb, u);
            s_g = g;
            s_local = u - S.bunitT[g];
        }
        __syncthreads();
        const int g = s_g;
        const int32_t slot = S.batch[g];
        const SgNodeMeta m = S.meta[slot];
        const int64_t src = S.bsrc[g];
        const uint64_t* vrows;
        int64_t vstride, vwords;
        int k0;
        if (src >= 0) {
            const SgProj D = S.pdesc[src % S.pdesc_cap];
            vrows = S.parena + D.off % S.parena_cap;
            vstride = D.stride;
            vwords = D.words;
            k0 = m.pk;
        } else {
            const SgSegment sg = segs[m.seg];
            vrows = rows + sg.row_base;
            vstride = sg.stride;
            vwords = sg.words;
            k0 = 0;
        }
        const int kk = m.k - k0;
        if (kk <= 0) {   // a record always has items its view has not applied
            if (threadIdx.x == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
            continue;
        }
        if (threadIdx.x < kk) {
            const int32_t f = S.items[(int64_t)slot * SG_MAX_DEPTH + k0 + threadIdx.x];
            s_row[threadIdx.x] = src >= 0 ? sg_block_row(S, src, f) : f;
        }
        __syncthreads();
        const int64_t wb = s_local * SG_PT_WORDS + (int64_t)wv * (SG_PT_WORDS / 4);
        uint64_t mv[SG_PT_WORDS / 256];
        int32_t ex[SG_PT_WORDS / 256];
        int32_t tot = 0;
#pragma unroll
        for (int i = 0; i < SG_PT_WORDS / 256; i++) {
            const int64_t w = wb + i * 64 + lane;
            uint64_t a = 0ull;
            if (w < vwords) {
                a = vrows[(int64_t)s_row[0] * vstride + w];
                for (int j = 1; j < kk; j++) a &= vrows[(int64_t)s_row[j] * vstride + w];
            }
            mv[i] = a;
            const int32_t cnt = __popcll(a);
            int32_t y = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t z = __shfl_up(y, o, 64);
                if (lane >= o) y += z;
            }
            ex[i] = tot + y - cnt;
            tot += __shfl(y, 63, 64);
        }
        // one reservation per workgroup: the four waves' requests are added in LDS and
        // thread 0 takes them with one atomic (wave order inside the workgroup's range)
        const bool cmp = S.bpnew[g] == 2;
        int32_t ne[SG_PT_WORDS / 256], el[SG_PT_WORDS / 256];
        int32_t etot = 0;
        if (cmp) {
#pragma unroll
            for (int i = 0; i < SG_PT_WORDS / 256; i++) {
                const uint64_t nz = __ballot(mv[i] != 0ull);
                ne[i] = __popcll(nz);
                el[i] = __popcll(nz & ((1ull << lane) - 1ull));
                etot += ne[i];
            }
        }
        if (lane == 0)
            s_req[wv] = cmp ? ((unsigned long long)etot << 40) | (unsigned long long)tot : (unsigned long long)tot;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long sum = s_req[0] + s_req[1] + s_req[2] + s_req[3];
            unsigned long long b = sum ? atomicAdd((unsigned long long*)&S.pfill[g], sum) : 0ull;
            for (int w = 0; w < 4; w++) {
                s_base[w] = b;
                b += s_req[w];
            }
        }
        __syncthreads();
        if (cmp) {
            const uint64_t base = s_base[wv];
            const int64_t e0 = (int64_t)(base >> 40), b0 = (int64_t)(base & ((1ull << 40) - 1ull));
            const int64_t ecap = min((int64_t)m.sup, vwords);
            SgPWord* ent = S.pent + S.peoff[g];
            int32_t* first = S.ptid + S.ptoff[g];
            int64_t eacc = 0;
#pragma unroll
            for (int i = 0; i < SG_PT_WORDS / 256; i++) {
                if (mv[i]) {
                    const int64_t e = e0 + eacc + el[i];
                    const int64_t p = b0 + ex[i];
                    const int64_t c = __popcll(mv[i]);
                    if (e < ecap && p + c <= m.sup) {
                        SgPWord E;
                        E.w = (int32_t)(wb + i * 64 + lane);
                        E.p = (int32_t)p;
                        E.m = mv[i];
                        ent[e] = E;
                        // the compress unit whose first bit is among this entry's bits
                        const int64_t u = (p + c - 1) / (64 * SG_PC_WORDS);
                        if (u * (64 * SG_PC_WORDS) >= p) first[u] = (int32_t)e;
                    }
                }
                eacc += ne[i];
            }
            continue;
        }
        const int64_t base = (int64_t)s_base[wv];
        int32_t* out = S.ptid + S.ptoff[g];
        const int64_t lim = m.sup;
#pragma unroll
        for (int i = 0; i < SG_PT_WORDS / 256; i++) {
            const int64_t w = wb + i * 64 + lane;
            uint64_t a = mv[i];
            int64_t pos = base + ex[i];
            while (a) {
                const int bit = (int)__builtin_ctzll(a);
                a &= a - 1ull;
                if (pos < lim) out[pos] = (int32_t)(w * 64 + bit);
                pos++;
            }
        }
    }
}

// Rows of the round's new blocks: row r of record X's block holds bit j = transaction
// T(X)[j] of the source view in the row of X's r-th item of U. One wave per unit
// (record, SG_PG_TIDS list positions = SG_PG_TIDS / 64 block words, up to SG_PG_ROWS
// rows): the wave loads the unit's list positions once (SG_PG_TIDS / 64 per lane), then
// per row gathers one bit per position and packs 64 of them per ballot; lane j keeps word
// j of the unit, so each row is one contiguous store. A list whose length differs from
// the record's support flags an error (the block is then not trusted: the search stops
// with SGFIM_E_DEVICE).
__global__ __launch_bounds__(256) void k_proj_gather(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                    const SgSegment* __restrict__ segs)
{
    constexpr int NJ = SG_PG_TIDS / 64;   // block words per unit row
    static_assert(NJ <= 64, "one lane per block word of a unit row");
    const int64_t nunits = S.ctl[CTL_PROJ_GUNITS];
    const int nb = (int)S.ctl[CTL_BATCH_N];
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t u = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < nunits; u += nwaves) {   // wave-uniform
        const int g = sg_find_record(S.bunitG, nb, u);
        const int64_t local = u - S.bunitG[g];
        const int32_t slot = S.batch[g];
        const SgNodeMeta m = S.meta[slot];
        const int nrg = (m.ncand + SG_PG_ROWS - 1) / SG_PG_ROWS;
        const int64_t tb = local / nrg;
        const int rg = (int)(local - tb * nrg);
        if (S.pfill[g] != (int64_t)m.sup) {
            if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
            continue;
        }
        const int64_t src = S.bsrc[g];
        const uint64_t* vrows;
        int64_t vstride;
        if (src >= 0) {
            const SgProj D = S.pdesc[src % S.pdesc_cap];
            vrows = S.parena + D.off % S.parena_cap;
            vstride = D.stride;
        } else {
            const SgSegment sg = segs[m.seg];
            vrows = rows + sg.row_base;
            vstride = sg.stride;
        }
        const SgProj N = S.pdesc[S.bview[g] % S.pdesc_cap];
        uint64_t* dst = S.parena + N.off % S.parena_cap;
        const int32_t* tl = S.ptid + S.ptoff[g];
        int32_t t[NJ];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int64_t p = tb * SG_PG_TIDS + j * 64 + lane;
            t[j] = p < m.sup ? tl[p] : -1;
        }
        const int32_t* ul = S.clist + S.coff[g];
        const int r1 = min(m.ncand, (rg + 1) * SG_PG_ROWS);
        const int64_t wo = tb * NJ + lane;   // the word of the block row this lane stores
        // four rows per memory round trip: their NJ bit loads per lane are all in flight
        // before the first ballot (32-bit loads: the word half that holds the bit)
        constexpr int RB = 4;
        for (int r = rg * SG_PG_ROWS; r < r1; r += RB) {
            const uint32_t* rp[RB];
#pragma unroll
            for (int q = 0; q < RB; q++) {
                const int32_t f = ul[min(r + q, r1 - 1)];
                rp[q] = (const uint32_t*)(vrows + (int64_t)(src >= 0 ? sg_block_row(S, src, f) : f) * vstride);
            }
            uint32_t v[RB][NJ];
#pragma unroll
            for (int q = 0; q < RB; q++)
#pragma unroll
                for (int j = 0; j < NJ; j++) v[q][j] = t[j] >= 0 ? rp[q][t[j] >> 5] : 0u;
#pragma unroll
            for (int q = 0; q < RB; q++) {
                uint64_t word = 0ull;
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const uint64_t bal = __ballot((v[q][j] >> (t[j] & 31)) & 1u);
                    if (lane == j) word = bal;
                }
                if (r + q < r1 && lane < NJ && wo < N.words) dst[(int64_t)(r + q) * N.stride + wo] = word;
            }
        }
    }
}

// Bit compress of x under mask m (the bits of x where m is set, packed to the low end), in
// the form with the six move masks of m precomputed (Hacker's Delight, 2nd ed., sec. 7-4,
// "compress" with the masks hoisted out of the loop; the masks cost ~70 operations once per
// mask). Stage i moves the bits of mv[i] down by 2^i and keeps the others:
// x = (x & ~mv[i]) | ((x & mv[i]) >> 2^i), on a 32-bit pair two ANDs, a funnel shift, a
// shift and two three-input logic instructions. Stage 0 keeps keep0 = m & ~mv[0] instead,
// so x needs no separate AND with m (after it x holds no bits outside the moved mask), and
// stage 5 only moves high bits into the low half: 32 instructions per word.
__device__ __forceinline__ void sg_compress_masks(uint64_t m, uint64_t (&mv)[6], uint64_t& keep0)
{
    uint64_t mk = ~m << 1;   // zeros of m, counted from the right
#pragma unroll
    for (int i = 0; i < 6; i++) {
        uint64_t mp = mk ^ (mk << 1);   // parallel suffix of mk
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        mp ^= mp << 32;
        const uint64_t v = mp & m;   // bits of m that move by 2^i
        mv[i] = v;
        if (i == 0) keep0 = m & ~v;
        m = (m ^ v) | (v >> (1 << i));
        mk &= ~mp;
    }
}

// One v_bitop3_b32 each; the immediate is the truth table over the operands' patterns
// 0xF0, 0xCC, 0xAA: (a & b) | c is 0xEA, (a & ~b) | c is 0xBA.
__device__ __forceinline__ uint32_t sg_and_or(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xEA);
}
__device__ __forceinline__ uint32_t sg_andn_or(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xBA);
}

__device__ __forceinline__ uint64_t sg_compress(uint64_t x, const uint64_t (&mv)[6], uint64_t keep0)
{
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const uint32_t ml = (uint32_t)mv[i], mh = (uint32_t)(mv[i] >> 32);
        const uint64_t t = ((uint64_t)(hi & mh) << 32 | (lo & ml)) >> (1 << i);
        if (i == 0) {
            lo = sg_and_or(lo, (uint32_t)keep0, (uint32_t)t);
            hi = sg_and_or(hi, (uint32_t)(keep0 >> 32), (uint32_t)(t >> 32));
        } else {
            lo = sg_andn_or(lo, ml, (uint32_t)t);
            hi = sg_andn_or(hi, mh, (uint32_t)(t >> 32));
        }
    }
    // stage 5 moves bits of the high half into the low one only
    const uint32_t mh5 = (uint32_t)(mv[5] >> 32);
    lo |= hi & mh5;
    hi &= ~mh5;
    return (uint64_t)hi << 32 | lo;
}

// Rows of the round's compacted blocks. One wave per unit (record, SG_PC_WORDS block words
// = bits [start, end), SG_PC_ROWS rows): starting at the word entry that holds bit start,
// each lane takes one entry (a nonzero source word w with mask m and first bit p), trims
// m to its bits inside the unit (an entry that straddles the unit's start or end), and for
// every row loads the source row's word w, compresses it under the trimmed mask and ORs
// the packed bits into the unit's words in LDS at bit p - start. Entries are in block-bit
// order, so the unit's entries are consecutive; the wave takes 64 at a time until they
// pass end. The unit's words are then stored whole, row by row.
__global__ __launch_bounds__(64) void k_proj_compress(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                     const SgSegment* __restrict__ segs)
{
    __shared__ unsigned long long s_out[SG_PC_ROWS][SG_PC_WORDS];
    constexpr int64_t BITS = 64 * SG_PC_WORDS;
    const int64_t nunits = S.ctl[CTL_PROJ_CUNITS];
    const int nb = (int)S.ctl[CTL_BATCH_N];
    const int lane = threadIdx.x;
    for (int64_t u = blockIdx.x; u < nunits; u += gridDim.x) {   // workgroup (= wave) uniform
        const int g = sg_find_record(S.bunitC, nb, u);
        const int64_t local = u - S.bunitC[g];
        const int32_t slot = S.batch[g];
        const SgNodeMeta m = S.meta[slot];
        const int nrg = (m.ncand + SG_PC_ROWS - 1) / SG_PC_ROWS;
        const int32_t ub = (int32_t)local / nrg;   // a record's units fit 32 bits (its block words / 16)
        const int rg = (int)local - ub * nrg;
        const int64_t src = S.bsrc[g];
        const uint64_t* vrows;
        int64_t vstride, vwords;
        if (src >= 0) {
            const SgProj D = S.pdesc[src % S.pdesc_cap];
            vrows = S.parena + D.off % S.parena_cap;
            vstride = D.stride;
            vwords = D.words;
        } else {
            const SgSegment sg = segs[m.seg];
            vrows = rows + sg.row_base;
            vstride = sg.stride;
            vwords = sg.words;
        }
        const uint64_t fill = (uint64_t)S.pfill[g];
        const int64_t nbits = (int64_t)(fill & ((1ull << 40) - 1ull)), nent = (int64_t)(fill >> 40);
        if (nbits != (int64_t)m.sup || nent > min((int64_t)m.sup, vwords)) {
            if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
            continue;
        }
        const SgProj N = S.pdesc[S.bview[g] % S.pdesc_cap];
        uint64_t* dst = S.parena + N.off % S.parena_cap;
        const SgPWord* ent = S.pent + S.peoff[g];
        const int64_t efirst = S.ptid[S.ptoff[g] + ub];
        if (efirst < 0 || efirst >= nent) {   // every unit's first bit has an entry
            if (lane == 0) atomicOr((unsigned long long*)&S.ctl[CTL_ERR], 4ull);
            continue;
        }
        const int64_t start = (int64_t)ub * BITS, end = min(start + BITS, (int64_t)m.sup);
        const int r0 = rg * SG_PC_ROWS;
        const int nr = min(m.ncand - r0, SG_PC_ROWS);
        // lane r < nr holds the source row offset of block row r0 + r
        int64_t roff = 0;
        if (lane < nr) {
            const int32_t f = S.clist[S.coff[g] + r0 + lane];
            roff = (int64_t)(src >= 0 ? sg_block_row(S, src, f) : f) * vstride;
        }
        for (int i = lane; i < SG_PC_ROWS * SG_PC_WORDS; i += 64) (&s_out[0][0])[i] = 0ull;
        __syncthreads();
        for (int64_t e0 = efirst; e0 < nent; e0 += 64) {   // wave-uniform
            const int64_t e = e0 + lane;
            SgPWord E;
            E.w = 0;
            E.p = 0;
            E.m = 0ull;
            if (e < nent) E = ent[e];
            const bool act = e < nent && E.p < end && E.m != 0ull;
            if (!__ballot(act)) break;
            // inactive lanes: no bits (mask 0), word 0 of the source row, LDS word 0
            uint64_t mk = act ? E.m : 0ull;
            const int c = __popcll(mk);
            const int64_t off = act ? (int64_t)E.p - start : 0;
            // bits before the unit's start (fewer than the entry's c; capped so that a
            // corrupt list cannot stretch the trim loop below: 64 clears any mask)
            const int hs = off < 0 ? (int)min(-off, (int64_t)64) : 0;
            const int64_t pos = off < 0 ? 0 : off;
            const int keep = act ? (int)min((int64_t)(c - hs), end - start - pos) : 0;   // >= 1 on active lanes
            // trim the mask to the kept bits: drop its lowest hs and then its highest
            // c - hs - keep set bits (only an entry that straddles the start or the end has any)
            for (int i = 0; i < hs; i++) mk &= mk - 1ull;
            for (int i = keep; i < c - hs; i++) mk ^= 1ull << (63 - __clzll(mk));
            const int q = (int)(pos >> 6), s = (int)(pos & 63);
            const bool spill = s != 0 && keep + s > 64;
            const int64_t w = act ? E.w : 0;
            uint64_t mvv[6], kp0;
            sg_compress_masks(mk, mvv, kp0);
            // source word w of block row rr (rows past the last reread the last row)
            auto load = [&](int rr) -> uint64_t {
                rr = min(rr, nr - 1);
                const int64_t ro = ((int64_t)__builtin_amdgcn_readlane((int)(roff >> 32), rr) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)roff, rr);
                return vrows[ro + w];
            };
            auto put = [&](int rr, uint64_t x) {
                const uint64_t y = sg_compress(x, mvv, kp0);
                atomicOr(&s_out[rr][q], (unsigned long long)(y << s));
                if (spill) atomicOr(&s_out[rr][q + 1], (unsigned long long)(y >> (64 - s)));
            };
            // two batches of RB rows in registers: one batch's loads are in flight while
            // the other's words are compressed
            constexpr int RB = 4;
            uint64_t xa[RB], xb[RB];
#pragma unroll
            for (int j = 0; j < RB; j++) xa[j] = load(j);
            for (int r = 0; r < nr; r += 2 * RB) {   // wave-uniform
#pragma unroll
                for (int j = 0; j < RB; j++) xb[j] = load(r + RB + j);
#pragma unroll
                for (int j = 0; j < RB; j++)
                    if (r + j < nr) put(r + j, xa[j]);
                if (r + RB >= nr) break;
#pragma unroll
                for (int j = 0; j < RB; j++) xa[j] = load(r + 2 * RB + j);
#pragma unroll
                for (int j = 0; j < RB; j++)
                    if (r + RB + j < nr) put(r + RB + j, xb[j]);
            }
        }
        __syncthreads();
        const int64_t nq = min((int64_t)SG_PC_WORDS, N.words - (int64_t)ub * SG_PC_WORDS);
        for (int i = lane; i < nr * SG_PC_WORDS; i += 64) {
            const int r = i / SG_PC_WORDS, qq = i - r * SG_PC_WORDS;
            if (qq < nq) dst[(int64_t)(r0 + r) * N.stride + (int64_t)ub * SG_PC_WORDS + qq] = s_out[r][qq];
        }
        __syncthreads();   // the next unit clears the LDS
    }
}

// ------------------------------------------------------------------ need rows
// One wave per child e of the round's records: the row of candidates e needs, F(P u {e})
// from the record's parent P's sibling block while it is intact, else e's pair-ok row.
// Every counting unit of the child's block and the child's expansion read it from here.
__global__ __launch_bounds__(256) void k_need_rows(SgSearchDev S, const SgSegment* __restrict__ segs,
                                                  const uint64_t* __restrict__ ok)
{
    const int64_t nchild = S.ctl[CTL_CHILDREN];
    const int lane = threadIdx.x & 63;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < nchild; q += stride) {   // wave-uniform
        const int g = S.chrec[q];
        const SgNodeMeta m = S.meta[S.batch[g]];
        const SgSegment sg = segs[m.seg];
        const int32_t e = S.clist[S.coff[g] + (int)(q - S.choff[g])];
        uint64_t v;
        if (sg_sib_live(S, m.sib)) {
            const int64_t r = sg_sib_row(S, m.sib, e, sg.dw);
            v = lane < sg.dw ? S.sib[(r + lane) % S.sib_cap] : 0ull;
        } else {
            v = lane < sg.dw ? ok[sg.ok_base + (int64_t)e * sg.dw + lane] : 0ull;
        }
        if (lane < S.dw_max) S.need[q * S.dw_max + lane] = v;
    }
}

// ------------------------------------------------------------------ counting tiles
// Sum of four per-lane values over the wave, transposed: on return lane 16*r holds the
// total of v[r] (every lane of row r holds it). Two 32-lane swaps, one 16-lane swap and
// a 16-lane DPP rotation tree: 10 VALU ops for four sums instead of four full trees.
__device__ __forceinline__ uint32_t sg_wave_sum4(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3)
{
    // permlane32_swap(a, b) -> {a.lo | b.lo, a.hi | b.hi} (halves of 32 lanes)
    const auto p02 = __builtin_amdgcn_permlane32_swap(v0, v2, false, false);
    const auto p13 = __builtin_amdgcn_permlane32_swap(v1, v3, false, false);
    const uint32_t s02 = p02[0] + p02[1];   // lanes 0-31: v0 halves, 32-63: v2 halves
    const uint32_t s13 = p13[0] + p13[1];   // lanes 0-31: v1 halves, 32-63: v3 halves
    // permlane16_swap(a, b) -> {a.r0 b.r0 a.r2 b.r2, a.r1 b.r1 a.r3 b.r3} (rows of 16 lanes)
    const auto q = __builtin_amdgcn_permlane16_swap(s02, s13, false, false);
    uint32_t t = q[0] + q[1];               // row r holds 16 partial sums of v[r]
    t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x128, 0xf, 0xf, false);
    t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x124, 0xf, 0xf, false);
    t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x122, 0xf, 0xf, false);
    t += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x121, 0xf, 0xf, false);
    return t;
}

#define SG_TILE_THREADS 512

// Popcounts of x & c over the PER words a lane holds (all PER, or the first qn).
template <int PER, bool FULL>
__device__ __forceinline__ uint32_t sg_andpop(const uint64_t (&x)[PER], const uint64_t (&c)[PER], int qn)
{
    uint32_t a = 0u;
#pragma unroll
    for (int q = 0; q < PER; q++)
        if (FULL || q < qn) a += (uint32_t)__popcll(x[q] & c[q]);
    return a;
}

// One wave step of a chunk: SG_RC listed candidates against the children that need
// them. Candidate r's row chunk is loaded once into registers (sg_step_load); for every
// child j that needs one of them, the child prefix is read from LDS once and ANDed +
// popcounted with each candidate that needs it; the four per-lane sums reduce together
// (sg_wave_sum4) and lane 16r adds candidate r's total to its cell (sg_step_count).
template <int PER>
__device__ __forceinline__ void sg_step_load(uint64_t (&x)[SG_RC][PER], const uint64_t* __restrict__ seg_rows,
                                             int64_t stride, int64_t w0, int wn, uint64_t nm,
                                             const int32_t (&fr)[SG_RC])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int rr = 0; rr < SG_RC; rr++) {
        const bool any = ((nm >> rr) & 0x1111111111111111ull) != 0ull;   // some child needs candidate rr
        const uint64_t* row = seg_rows + (int64_t)(fr[rr] < 0 ? 0 : fr[rr]) * stride + w0;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int w = lane + 64 * q;
            x[rr][q] = (any && w < wn) ? row[w] : 0ull;
        }
    }
}

template <int PER, bool FULL>
__device__ __forceinline__ void sg_step_count(const uint64_t (&x)[SG_RC][PER], int qn, uint64_t nm, int il_mine,
                                              const uint64_t (*cp)[PER * 64], int ncb, int32_t* cells, int j0, int nu)
{
    const int lane = threadIdx.x & 63;
    const int rl = lane >> 4;
    // keep the per-child masks inside the step: hoisted, their branch conditions would
    // not fit the scalar registers
    asm volatile("" : "+s"(nm));
    for (int j = 0; j < ncb; j++) {
        const uint32_t mj = (uint32_t)(nm >> (4 * j)) & 0xfu;   // candidates child j needs
        if (mj == 0u) continue;
        uint64_t c[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) c[q] = (FULL || q < qn) ? cp[j][lane + 64 * q] : 0ull;
        const uint32_t a0 = (mj & 1u) ? sg_andpop<PER, FULL>(x[0], c, qn) : 0u;
        const uint32_t a1 = (mj & 2u) ? sg_andpop<PER, FULL>(x[1], c, qn) : 0u;
        const uint32_t a2 = (mj & 4u) ? sg_andpop<PER, FULL>(x[2], c, qn) : 0u;
        const uint32_t a3 = (mj & 8u) ? sg_andpop<PER, FULL>(x[3], c, qn) : 0u;
        const uint32_t t = sg_wave_sum4(a0, a1, a2, a3);
        if ((lane & 15) == 0 && ((mj >> rl) & 1u) && t) atomicAdd(&cells[(int64_t)(j0 + j) * nu + il_mine], (int32_t)t);
    }
}

// Step i0 of the unit's listed candidates: lane (j, r) = (lane >> 2, lane & 3) holds
// whether child j needs candidate r, so the ballot keeps child j's four bits at 4j.
__device__ __forceinline__ void sg_step_plan(const uint32_t* s_cand, const int32_t* clist, int nc, int i0,
                                             uint64_t& nm, int32_t (&fr)[SG_RC], int& il_mine)
{
    const int lane = threadIdx.x & 63;
    const int rq = lane & 3, jq = lane >> 2;
    const uint32_t ent = i0 + rq < nc ? s_cand[i0 + rq] : 0u;
    const int32_t fl = ent ? clist[(int)(ent & 0xffffu)] : -1;
    nm = __ballot((ent >> (16 + jq)) & 1u);   // bit 4*j + r
#pragma unroll
    for (int rr = 0; rr < SG_RC; rr++) fr[rr] = __builtin_amdgcn_readlane(fl, rr);
    const uint32_t ent_mine = i0 + (lane >> 4) < nc ? s_cand[i0 + (lane >> 4)] : 0u;
    il_mine = (int)(ent_mine & 0xffffu);
}

// Unit (record g, block of up to SG_CB children, range of up to chunk_range row chunks).
// Set-up, once per unit: the children's need rows (k_need_rows) give each candidate f of
// U the mask of children e that need it (f in F(P u {e}) of the parent's sibling block,
// or the pair (e, f) frequent; f != e); candidates no child needs are dropped, the others
// listed in LDS. Then per chunk: the prefix of X and the
// prefixes of the children go to LDS, and each wave takes SG_RC listed candidates at a
// time (sg_tile_step); lane (j, r) = (lane >> 2, lane & 3) of a step holds whether
// child j needs candidate r, so the ballot keeps child j's four bits at 4j. One integer
// atomic per needed (child, candidate, chunk). Skipped cells stay 0, which the
// expansion reads as infrequent -- exact, since a skipped cell's itemset has an
// infrequent subset. The child prefixes' popcounts give the children's supports.
// Registers are capped for 6 waves per SIMD (three workgroups per CU, as the LDS of a
// 256-word chunk allows); 128-word chunks leave LDS for four workgroups, and ask for 8.
template <int CW>
__global__ __launch_bounds__(SG_TILE_THREADS) __attribute__((amdgpu_waves_per_eu(CW <= 128 ? 8 : 6, 8))) void k_count_group(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                                 const SgSegment* __restrict__ segs,
                                                                 const uint64_t* __restrict__ ok)
{
    constexpr int PER = CW / 64;
    constexpr int NW = SG_TILE_THREADS / 64;
    static_assert(SG_RC * SG_CB == 64, "one lane per (child, candidate) pair of a wave step");
    __shared__ uint64_t pre[CW];             // T(X) over the chunk
    __shared__ uint64_t cp[SG_CB][CW];       // child prefixes; the children's need rows during set-up
    __shared__ uint32_t s_cand[SG_DMAX];     // (children mask << 16) | position in U
    __shared__ int32_t s_g, s_cr, s_blk;
    __shared__ int32_t s_items[SG_MAX_DEPTH];
    __shared__ int32_t s_child[SG_CB];
    __shared__ int32_t s_nc;
    static_assert(SG_CB * 64 <= SG_CB * CW, "ok rows fit the prefix buffer");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t* okr = &cp[0][0];
    const int64_t n_units = S.ctl[CTL_UNITS];
    const int64_t n_blocks = S.ctl[CTL_BLOCKS];
    const int64_t units_a = S.ctl[CTL_UNITS_A];
    const int nb = (int)S.ctl[CTL_BATCH_N];
    if (n_units == 0) return;   // a round queued after the frontier emptied: nothing to count
    if (threadIdx.x == 0)
        atomicMin((unsigned long long*)&S.ctl[CTL_T0], (unsigned long long)wall_clock64());
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t u = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_UNIT_NEXT], 1ull);
            s_nc = 0;
            if (u >= n_units) {
                s_g = -1;
            } else if (u < units_a) {
                // records on window rows, range-major: all their child blocks work on
                // row-chunk range cr together, so the rows of one range stay in L2 while
                // every record uses them
                const int cr = (int)(u / n_blocks);
                const int64_t bq = u - (int64_t)cr * n_blocks;
                const int g = S.blkrec[bq];
                s_g = g;
                s_cr = cr;
                s_blk = (int)(bq - S.unit_off[g]);
            } else {
                // records on projected blocks, record-major: a record's blocks range by range
                const int64_t v = u - units_a;
                const int g = sg_find_record(S.bunitB, nb, v);
                const int64_t local = v - S.bunitB[g];
                const int nblk = S.meta[S.batch[g]].nblk;
                const int cr = (int)(local / nblk);
                s_g = g;
                s_cr = cr;
                s_blk = (int)(local - (int64_t)cr * nblk);
            }
        }
        __syncthreads();
        const int g = s_g;
        if (g < 0) break;
        const int cr = s_cr, blk = s_blk;
        const int32_t slot = S.batch[g];
        const SgNodeMeta m = S.meta[slot];
        if (cr >= S.bnr[g]) {   // past this record's (shorter) window
            __syncthreads();
            continue;
        }
        const SgSegment sg = segs[m.seg];
        // the record's view: its projected block (T(X) of the block's itemset is every
        // transaction of it; X's later items are ANDed) or the window's rows (all of X)
        const int64_t view = S.bview[g];
        const uint64_t* seg_rows;
        int64_t vstride, vwords;
        int k0;
        if (view >= 0) {
            const SgProj D = S.pdesc[view % S.pdesc_cap];
            seg_rows = S.parena + D.off % S.parena_cap;
            vstride = D.stride;
            vwords = D.words;
            k0 = S.bpnew[g] ? m.k : m.pk;
        } else {
            seg_rows = rows + sg.row_base;
            vstride = sg.stride;
            vwords = sg.words;
            k0 = 0;
        }
        const int kk = m.k - k0;
        const int64_t nchunk = (vwords + CW - 1) / CW;
        const int c0 = cr * S.chunk_range;
        const int c1 = (int)min((int64_t)(c0 + S.chunk_range), nchunk);
        const int j0 = blk * SG_CB;
        const int ncb = min(SG_CB, m.ntail - j0);
        const int32_t u0 = S.coff[g];
        const int nu = m.ncand;
        if (threadIdx.x < kk) {
            const int32_t f = S.items[(int64_t)slot * SG_MAX_DEPTH + k0 + threadIdx.x];
            s_items[threadIdx.x] = view >= 0 ? sg_block_row(S, view, f) : f;
        }
        if (threadIdx.x < ncb) s_child[threadIdx.x] = S.crow[u0 + j0 + threadIdx.x];
        __syncthreads();
        // set-up: the children's need rows (k_need_rows), then the list of needed candidates
        {
            const uint64_t* nr = S.need + (int64_t)(S.choff[g] + j0) * S.dw_max;
            for (int t = threadIdx.x; t < ncb * sg.dw; t += blockDim.x) {
                const int j = t / sg.dw, w = t - j * sg.dw;
                okr[j * 64 + w] = nr[(int64_t)j * S.dw_max + w];
            }
        }
        __syncthreads();
        // Candidate i needs counting for child j0 + j when the pair is frequent, except that
        // a candidate that is itself a child (i < ntail) is counted only below the
        // diagonal: |T(X u {e_a, e_b})| is one support, counted once in cell (a, b),
        // a > b, and read from there for cell (b, a).
        for (int i = threadIdx.x; i < nu; i += blockDim.x) {
            const int32_t f = S.clist[u0 + i];
            const int jlo = i < m.ntail ? i + 1 - j0 : 0;   // children j with j0 + j > i
            uint32_t mask = 0u;
            for (int j = max(jlo, 0); j < ncb; j++)
                if ((okr[j * 64 + (f >> 6)] >> (f & 63)) & 1ull) mask |= 1u << j;
            if (mask) s_cand[atomicAdd(&s_nc, 1)] = (mask << 16) | (uint32_t)i;
        }
        __syncthreads();
        const int nc = s_nc;
        int32_t* cells = S.cells + S.celloff[g];
        const int32_t* crow = S.crow + u0;   // rows of U's items in the view
        // the wave's first step takes the same candidates in every chunk of the unit
        const int i0f = wave * SG_RC;
        uint64_t nmf;
        int32_t frf[SG_RC];
        int ilf;
        sg_step_plan(s_cand, crow, nc, i0f, nmf, frf, ilf);
        for (int c = c0; c < c1; c++) {
            const int64_t w0 = (int64_t)c * CW;
            const int wn = (int)min((int64_t)CW, vwords - w0);
            const int qn = (wn + 63) >> 6;   // 64-word columns holding row words
            __syncthreads();   // every wave is done with the previous chunk's prefixes
            // child prefixes T(X) & T(e) (zero past the last word) and the children's
            // supports. Each wave loads its children's rows and, for its slice of the
            // chunk's words, the rows of X, all in one round trip; T(X) over the slices
            // meets in LDS.
            {
                uint64_t rv[2][PER];
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int j = wave + NW * jj;
                    const uint64_t* row = seg_rows + (int64_t)(j < ncb ? s_child[j] : 0) * vstride + w0;
#pragma unroll
                    for (int q = 0; q < PER; q++) {
                        const int w = lane + 64 * q;
                        rv[jj][q] = (j < ncb && w < wn) ? row[w] : 0ull;
                    }
                }
                constexpr int SW = CW / NW;   // words of a wave's slice
                if (lane < SW) {
                    const int w = wave * SW + lane;
                    uint64_t v = w < wn ? ~0ull : 0ull;
                    for (int i = 0; i < kk; i++)
                        if (w < wn) v &= seg_rows[(int64_t)s_items[i] * vstride + w0 + w];
                    pre[w] = v;
                }
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int j = wave + NW * jj;
                    if (j < ncb) {
                        uint32_t acc = 0;
#pragma unroll
                        for (int q = 0; q < PER; q++) {
                            const uint64_t v = pre[lane + 64 * q] & rv[jj][q];
                            cp[j][lane + 64 * q] = v;
                            acc += (uint32_t)__popcll(v);
                        }
                        const uint32_t sv = sg_wave_sum(acc);
                        if (lane == 0 && sv) atomicAdd(&S.ccount[S.choff[g] + j0 + j], (int32_t)sv);
                    }
                }
            }
            // the first step's candidates load while the other waves finish their prefixes
            uint64_t xf[SG_RC][PER];
            sg_step_load<PER>(xf, seg_rows, vstride, w0, wn, i0f < nc ? nmf : 0ull, frf);
            __syncthreads();
            if (i0f < nc) {
                if (qn == PER) sg_step_count<PER, true>(xf, qn, nmf, ilf, cp, ncb, cells, j0, nu);
                else sg_step_count<PER, false>(xf, qn, nmf, ilf, cp, ncb, cells, j0, nu);
            }
            for (int i0 = i0f + NW * SG_RC; i0 < nc; i0 += NW * SG_RC) {
                uint64_t nm;
                int32_t fr[SG_RC];
                int il;
                sg_step_plan(s_cand, crow, nc, i0, nm, fr, il);
                uint64_t x[SG_RC][PER];
                sg_step_load<PER>(x, seg_rows, vstride, w0, wn, nm, fr);
                if (qn == PER) sg_step_count<PER, true>(x, qn, nm, il, cp, ncb, cells, j0, nu);
                else sg_step_count<PER, false>(x, qn, nm, il, cp, ncb, cells, j0, nu);
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        atomicMax((unsigned long long*)&S.ctl[CTL_T1], (unsigned long long)wall_clock64());
    }
}

// ------------------------------------------------------------------ counting on fp4 MFMA
#define SG_F4G 2   // candidate groups of 16 per item: a child's masked operand feeds SG_F4G MFMAs
// The units of k_count_group on the matrix cores (sgfim_set_count_kernel). Units, set-up
// (need rows, listed candidates) and child prefixes (T(X) & T(e) of the block's 16
// children over a chunk, in LDS) are k_count_group's; the cells of a group of 16 listed
// candidates are then one 16 x 16 product of 0/1 matrices over the chunk's transactions,
// C[j][c] = |prefix_j & row_c|, on v_mfma_scale_f32_16x16x128_f8f6f4 with the bit-phase
// fp4 operands of f4_phase.h. Lane (r, q) = (lane & 15, lane >> 4) of a K step reads words
// 8s + 2q and 8s + 2q + 1 of child r's prefix (LDS) and of candidate r's row (the view);
// the four bit phases of those 128 bits are four K steps, 512 transactions per 8-word
// slice s. A wave takes items (SG_F4G candidate groups, range of slices) round-robin; a
// child's masked phase feeds the item's SG_F4G products. The MFMA computes all 16 x 16
// cells of a group and only the needed nonzero ones are added (one integer atomic per
// cell, chunk and slice range), so the cell matrix equals k_count_group's bit for bit; an
// item's fp32 sums are at most 2^15 (512 words x 64), exact. Registers are capped for four
// waves per SIMD (two workgroups per CU). It is slower than k_count_group on the benchmark
// workloads (the masks cost 5 VALU instructions per operand dword and phase set, reused
// only 16 or 32 times at 16 children per unit), so it is not the default.
template <int CW>
__global__ __launch_bounds__(SG_TILE_THREADS) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_count_f4(SgSearchDev S, const uint64_t* __restrict__ rows,
                                                                 const SgSegment* __restrict__ segs,
                                                                 const uint64_t* __restrict__ ok)
{
    constexpr int PER = CW / 64;
    constexpr int NW = SG_TILE_THREADS / 64;
    constexpr int CWP = CW + 2;   // prefix row stride: 16 bytes of pad spread a slice's 16 rows over the banks
    static_assert(SG_CB == 16 && NW == 8, "one 16 x 16 product per candidate group; two children per wave");
    __shared__ uint64_t pre[CW];                                        // T(X) over the chunk
    __shared__ __attribute__((aligned(16))) uint64_t cp[SG_CB][CWP];   // child prefixes; need rows during set-up
    __shared__ uint32_t s_cand[SG_DMAX];     // (children mask << 16) | position in U
    __shared__ int32_t s_g, s_cr, s_blk;
    __shared__ int32_t s_items[SG_MAX_DEPTH];
    __shared__ int32_t s_child[SG_CB];
    __shared__ int32_t s_nc;
    static_assert(SG_CB * 64 <= SG_CB * CWP, "need rows fit the prefix buffer");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t* okr = &cp[0][0];
    const int64_t n_units = S.ctl[CTL_UNITS];
    const int64_t n_blocks = S.ctl[CTL_BLOCKS];
    const int64_t units_a = S.ctl[CTL_UNITS_A];
    const int nb = (int)S.ctl[CTL_BATCH_N];
    if (n_units == 0) return;   // a round queued after the frontier emptied: nothing to count
    if (threadIdx.x == 0)
        atomicMin((unsigned long long*)&S.ctl[CTL_T0], (unsigned long long)wall_clock64());
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t u = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_UNIT_NEXT], 1ull);
            s_nc = 0;
            if (u >= n_units) {
                s_g = -1;
            } else if (u < units_a) {
                // records on window rows, range-major: all their child blocks work on
                // row-chunk range cr together, so the rows of one range stay in L2 while
                // every record uses them
                const int cr = (int)(u / n_blocks);
                const int64_t bq = u - (int64_t)cr * n_blocks;
                const int g = S.blkrec[bq];
                s_g = g;
                s_cr = cr;
                s_blk = (int)(bq - S.unit_off[g]);
            } else {
                // records on projected blocks, record-major: a record's blocks range by range
                const int64_t v = u - units_a;
                const int g = sg_find_record(S.bunitB, nb, v);
                const int64_t local = v - S.bunitB[g];
                const int nblk = S.meta[S.batch[g]].nblk;
                const int cr = (int)(local / nblk);
                s_g = g;
                s_cr = cr;
                s_blk = (int)(local - (int64_t)cr * nblk);
            }
        }
        __syncthreads();
        const int g = s_g;
        if (g < 0) break;
        const int cr = s_cr, blk = s_blk;
        const int32_t slot = S.batch[g];
        const SgNodeMeta m = S.meta[slot];
        if (cr >= S.bnr[g]) {   // past this record's (shorter) window
            __syncthreads();
            continue;
        }
        const SgSegment sg = segs[m.seg];
        // the record's view: its projected block (T(X) of the block's itemset is every
        // transaction of it; X's later items are ANDed) or the window's rows (all of X)
        const int64_t view = S.bview[g];
        const uint64_t* seg_rows;
        int64_t vstride, vwords;
        int k0;
        if (view >= 0) {
            const SgProj D = S.pdesc[view % S.pdesc_cap];
            seg_rows = S.parena + D.off % S.parena_cap;
            vstride = D.stride;
            vwords = D.words;
            k0 = S.bpnew[g] ? m.k : m.pk;
        } else {
            seg_rows = rows + sg.row_base;
            vstride = sg.stride;
            vwords = sg.words;
            k0 = 0;
        }
        const int kk = m.k - k0;
        const int64_t nchunk = (vwords + CW - 1) / CW;
        const int c0 = cr * S.chunk_range;
        const int c1 = (int)min((int64_t)(c0 + S.chunk_range), nchunk);
        const int j0 = blk * SG_CB;
        const int ncb = min(SG_CB, m.ntail - j0);
        const int32_t u0 = S.coff[g];
        const int nu = m.ncand;
        if (threadIdx.x < kk) {
            const int32_t f = S.items[(int64_t)slot * SG_MAX_DEPTH + k0 + threadIdx.x];
            s_items[threadIdx.x] = view >= 0 ? sg_block_row(S, view, f) : f;
        }
        if (threadIdx.x < ncb) s_child[threadIdx.x] = S.crow[u0 + j0 + threadIdx.x];
        __syncthreads();
        // set-up: the children's need rows (k_need_rows), then the list of needed candidates
        {
            const uint64_t* nr = S.need + (int64_t)(S.choff[g] + j0) * S.dw_max;
            for (int t = threadIdx.x; t < ncb * sg.dw; t += blockDim.x) {
                const int j = t / sg.dw, w = t - j * sg.dw;
                okr[j * 64 + w] = nr[(int64_t)j * S.dw_max + w];
            }
        }
        __syncthreads();
        // Candidate i needs counting for child j0 + j when the pair is frequent, except that
        // a candidate that is itself a child (i < ntail) is counted only below the
        // diagonal: |T(X u {e_a, e_b})| is one support, counted once in cell (a, b),
        // a > b, and read from there for cell (b, a).
        for (int i = threadIdx.x; i < nu; i += blockDim.x) {
            const int32_t f = S.clist[u0 + i];
            const int jlo = i < m.ntail ? i + 1 - j0 : 0;   // children j with j0 + j > i
            uint32_t mask = 0u;
            for (int j = max(jlo, 0); j < ncb; j++)
                if ((okr[j * 64 + (f >> 6)] >> (f & 63)) & 1ull) mask |= 1u << j;
            if (mask) s_cand[atomicAdd(&s_nc, 1)] = (mask << 16) | (uint32_t)i;
        }
        __syncthreads();
        const int nc = s_nc;
        int32_t* cells = S.cells + S.celloff[g];
        const int32_t* crow = S.crow + u0;   // rows of U's items in the view
        // items: (SG_F4G groups of 16 listed candidates, range of 8-word slices); kp slice
        // ranges per item so that small units still give every wave an item
        const int nsg = (nc + 16 * SG_F4G - 1) / (16 * SG_F4G);
        const int kp = nsg >= NW ? 1 : NW / max(nsg, 1);
        // lane (r, q): candidate r of its item's group (its need bits and row), slice quarter q
        const int r = lane & 15, q = lane >> 4;
        for (int c = c0; c < c1; c++) {
            const int64_t w0 = (int64_t)c * CW;
            const int wn = (int)min((int64_t)CW, vwords - w0);
            __syncthreads();   // every wave is done with the previous chunk's prefixes
            // child prefixes T(X) & T(e) (zero past the last word and for absent children)
            // and the children's supports, as in k_count_group
            {
                uint64_t rv[2][PER];
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int j = wave + NW * jj;
                    const uint64_t* row = seg_rows + (int64_t)(j < ncb ? s_child[j] : 0) * vstride + w0;
#pragma unroll
                    for (int qq = 0; qq < PER; qq++) {
                        const int w = lane + 64 * qq;
                        rv[jj][qq] = (j < ncb && w < wn) ? row[w] : 0ull;
                    }
                }
                constexpr int SW = CW / NW;   // words of a wave's slice
                if (lane < SW) {
                    const int w = wave * SW + lane;
                    uint64_t v = w < wn ? ~0ull : 0ull;
                    for (int i = 0; i < kk; i++)
                        if (w < wn) v &= seg_rows[(int64_t)s_items[i] * vstride + w0 + w];
                    pre[w] = v;
                }
                __syncthreads();
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int j = wave + NW * jj;
                    uint32_t acc = 0;
#pragma unroll
                    for (int qq = 0; qq < PER; qq++) {
                        const uint64_t v = pre[lane + 64 * qq] & rv[jj][qq];
                        cp[j][lane + 64 * qq] = v;   // zero for j >= ncb
                        acc += (uint32_t)__popcll(v);
                    }
                    if (j < ncb) {
                        const uint32_t sv = sg_wave_sum(acc);
                        if (lane == 0 && sv) atomicAdd(&S.ccount[S.choff[g] + j0 + j], (int32_t)sv);
                    }
                }
            }
            __syncthreads();
            const int nsl = (wn + 7) >> 3;   // 8-word slices holding row words
            for (int it = wave; it < nsg * kp; it += NW) {   // wave-uniform
                const int gs = it / kp, part = it - gs * kp;
                const int s0 = part * nsl / kp, s1 = (part + 1) * nsl / kp;
                // group gg of the item: listed candidates (4 gs + gg) x 16 + r
                uint32_t ent[SG_F4G];
                const uint64_t* brow[SG_F4G];
                bool have[SG_F4G];
#pragma unroll
                for (int gg = 0; gg < SG_F4G; gg++) {
                    const int ci = (SG_F4G * gs + gg) * 16 + r;
                    ent[gg] = ci < nc ? s_cand[ci] : 0u;
                    const int32_t fr = ent[gg] ? crow[(int)(ent[gg] & 0xffffu)] : -1;
                    have[gg] = fr >= 0;
                    brow[gg] = seg_rows + (int64_t)(fr < 0 ? 0 : fr) * vstride + w0;
                }
                // B: candidate r's words w, w + 1 of slice sl in each group (one 16-byte load:
                // rows, chunks and block bases are 16-byte aligned). A pair at or past wn, or
                // of an absent candidate, reads word 0 of the chunk and is zeroed (w < wn keeps
                // the pair inside the row's stride, and a word past the row meets a zero
                // prefix word).
                auto loadb = [&](uint4 (&bq)[SG_F4G], int sl) {
                    const int w = 8 * sl + 2 * q;
#pragma unroll
                    for (int gg = 0; gg < SG_F4G; gg++) {
                        const bool in = have[gg] && sl < s1 && w < wn;
                        const uint4 v = *(const uint4*)(brow[gg] + (in ? w : 0));
                        bq[gg] = in ? v : uint4{0u, 0u, 0u, 0u};
                    }
                };
                // per group two accumulators (phases 0, 2 and 1, 3): independent MFMA chains
                v4f acc[SG_F4G][2];
#pragma unroll
                for (int gg = 0; gg < SG_F4G; gg++) acc[gg][0] = acc[gg][1] = v4f{0.f, 0.f, 0.f, 0.f};
                // A: child r's words w, w + 1 (LDS, zero past the chunk's words), masked once per
                // phase for the item's four groups
                auto mult = [&](const uint4 (&bq)[SG_F4G], int sl) {
                    const uint4 a = *(const uint4*)&cp[r][8 * sl + 2 * q];
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        const v8i ap = sg_f4_mask_frag(a, p);
                        const int sc = p == 0 ? sg_f4_scale<0>() : p == 1 ? sg_f4_scale<1>() : sg_f4_scale<2>();
#pragma unroll
                        for (int gg = 0; gg < SG_F4G; gg++)
                            acc[gg][p & 1] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(ap, sg_f4_mask_frag(bq[gg], p),
                                                                                          acc[gg][p & 1], 4, 4, 0, sc, 0, sc);
                        __builtin_amdgcn_sched_barrier(0);   // one phase's operands live at a time
                    }
                };
                for (int sl = s0; sl < s1; sl++) {   // wave-uniform
                    uint4 bq[SG_F4G];
                    loadb(bq, sl);
                    mult(bq, sl);
                }
                // 16 x 16 output: lane (r, q) holds children 4q .. 4q + 3 against candidate r
#pragma unroll
                for (int gg = 0; gg < SG_F4G; gg++) {
                    const int il = (int)(ent[gg] & 0xffffu);
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int j = 4 * q + e;
                        const int32_t v = (int32_t)(acc[gg][0][e] + acc[gg][1][e]);   // integers below 2^24
                        if (j < ncb && ((ent[gg] >> (16 + j)) & 1u) && v)
                            atomicAdd(&cells[(int64_t)(j0 + j) * nu + il], v);
                    }
                }
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        atomicMax((unsigned long long*)&S.ctl[CTL_T1], (unsigned long long)wall_clock64());
    }
}

// ------------------------------------------------------------------ expansion
#define SG_EXP_WAVES 16
#define SG_EXP_PER 4   // most children per wave per workgroup step

// Workgroups of 16 waves stride over the batch's children, 16 x per per step, per =
// ceil(children / (16 x workgroups)) capped at 4: wave w takes children w x per ..
// w x per + per - 1 of the step one after another (small rounds keep one child per wave
// and all workgroups busy). Child j of record g is
// Y = X u {e}; its row of the count matrix gives F(Y) and EQ(Y) over U \ {e}; then the
// node step. MFI ring positions and record slots are reserved once per workgroup step
// (one atomic per counter for the step's children), and the cumulative counters once per
// workgroup. A step waits at its barriers for its slowest wave; four children per wave
// average that wait out on large rounds (one child per wave left the waves idle about
// half of the time on 500,000 x 256-item streams, whose children differ widely in
// candidates).
__global__ __launch_bounds__(SG_EXP_WAVES * 64) void k_expand_children(SgSearchDev S, const SgSegment* __restrict__ segs,
                                                                       const int32_t* __restrict__ pair,
                                                                       const uint64_t* __restrict__ ok,
                                                                       const int32_t* __restrict__ sup_all,
                                                                       const int64_t* __restrict__ seg_item_base)
{
    constexpr int NSM = SG_EXP_WAVES * SG_EXP_PER;
    __shared__ uint64_t sF[NSM][64];
    __shared__ uint64_t sE[NSM][64];
    __shared__ uint64_t sN[SG_EXP_WAVES][64];   // the child's need row (the counting tile's rule)
    __shared__ uint16_t sX[NSM][SG_MAX_DEPTH];
    __shared__ int32_t s_act[NSM], s_k[NSM], s_g[NSM], s_e[NSM], s_sup[NSM];
    __shared__ int64_t s_ring, s_nb, s_top, s_key;
    __shared__ unsigned long long s_counts, s_words;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nchild = S.ctl[CTL_CHILDREN];
    const int64_t lanes = (int64_t)gridDim.x * SG_EXP_WAVES;
    const int per = (int)min((int64_t)SG_EXP_PER, max((int64_t)1, (nchild + lanes - 1) / lanes));
    const int NS = SG_EXP_WAVES * per;   // children per step
    if (threadIdx.x == 0) { s_counts = 0; s_words = 0; }
    for (int64_t q0 = (int64_t)blockIdx.x * NS; q0 < nchild; q0 += (int64_t)gridDim.x * NS) {
#pragma unroll 1
        for (int c = 0; c < per; c++) {
            const int sl = wv * per + c;
            const int64_t q = q0 + sl;
            int act = SG_ACT_NONE, k = 0, g = 0;
            int32_t e = 0, sup_e = 0;
            sF[sl][lane] = 0ull;
            sE[sl][lane] = 0ull;
            if (q < nchild) {   // wave-uniform
                g = S.chrec[q];
                const int j = (int)(q - S.choff[g]);
                const int32_t slot = S.batch[g];
                const SgNodeMeta m = S.meta[slot];
                const SgSegment sg = segs[m.seg];
                const int32_t u0 = S.coff[g];
                const int nu = m.ncand;
                e = S.clist[u0 + j];
                sup_e = S.ccount[S.choff[g] + j];
                const int32_t* crow = S.cells + S.celloff[g] + (int64_t)j * nu;
                sN[wv][lane] = lane < sg.dw ? S.need[q * S.dw_max + lane] : 0ull;   // the counting tile's rule
                if (lane < m.k) sX[sl][lane] = S.items[(int64_t)slot * SG_MAX_DEPTH + lane];
                if (lane == 0) sX[sl][m.k] = (uint16_t)e;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                int needed = 0;
                for (int i = lane; i < nu; i += 64) {
                    if (i == j) continue;
                    const int32_t f = S.clist[u0 + i];
                    // a later child's support with e sits in that child's row (counted once)
                    const bool mirror = i < m.ntail && i > j;
                    const int32_t cnt = mirror ? crow[(int64_t)(i - j) * nu + j] : crow[i];
                    if (!mirror && ((sN[wv][f >> 6] >> (f & 63)) & 1ull)) needed++;
                    if (cnt >= sg.thr) {
                        atomicOr((unsigned long long*)&sF[sl][f >> 6], 1ull << (f & 63));
                        if (cnt == sup_e) atomicOr((unsigned long long*)&sE[sl][f >> 6], 1ull << (f & 63));
                    }
                }
                for (int o = 32; o > 0; o >>= 1) needed += __shfl_xor(needed, o, 64);
                if (lane == 0) {
                    // words per count: the row length of the view the record counted on
                    const int64_t view = S.bview[g];
                    const int64_t vw = view >= 0 ? S.pdesc[view % S.pdesc_cap].words : sg.words;
                    atomicAdd(&s_counts, (unsigned long long)needed);
                    atomicAdd(&s_words, (unsigned long long)(needed + 1) * (unsigned long long)vw);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // F(X u {e}) is child e's row of X's sibling block (before the node step takes
                // the equal-support items out of it)
                const int64_t sb = S.sib_off[g];
                if (sb >= 0 && lane < S.dw_max) S.sib[(sb + (int64_t)(1 + j) * S.dw_max + lane) % S.sib_cap] = sF[sl][lane];
                k = m.k + 1;
                act = sg_node_decide(S, sg, sF[sl], sE[sl], sX[sl], k, e);
                // the next child reuses sN[wv]
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (lane == 0) {
                s_act[sl] = act;
                s_k[sl] = k;
                s_g[sl] = g;
                s_e[sl] = e;
                s_sup[sl] = sup_e;
            }
        }
        __syncthreads();
        if (threadIdx.x < 4) {   // the four reservations in flight together
            int ne = 0, nr = 0;
            for (int w = 0; w < NS; w++) {
                ne += s_act[w] == SG_ACT_EMIT;
                nr += s_act[w] == SG_ACT_RECORD;
            }
            if (threadIdx.x == 0 && ne)
                s_ring = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_RING_POS], (unsigned long long)ne);
            if (threadIdx.x == 1 && nr)
                s_nb = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_NEWBORN], (unsigned long long)nr);
            if (threadIdx.x == 2 && nr)
                s_top = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_FREE_TOP], (unsigned long long)(-(int64_t)nr));
            if (threadIdx.x == 3 && nr)
                s_key = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_KEY], (unsigned long long)nr);
        }
        __syncthreads();
#pragma unroll 1
        for (int c = 0; c < per; c++) {
            const int sl = wv * per + c;
            const int act = s_act[sl];
            if (act == SG_ACT_NONE) continue;   // wave-uniform
            const int g = s_g[sl], k = s_k[sl];
            int rank = 0;
            for (int w = 0; w < sl; w++) rank += s_act[w] == act;
            const int32_t slot = S.batch[g];
            const SgNodeMeta m = S.meta[slot];
            const SgSegment sg = segs[m.seg];
            if (act == SG_ACT_EMIT) {
                sg_emit_at(S, s_ring + rank, sX[sl], k, s_sup[sl], m.seg, S.key[slot]);
            } else {
                const int64_t nbi = s_nb + rank;
                int32_t nslot = -1;
                if (lane == 0) nslot = sg_take_slot(S, nbi, s_top - rank);
                nslot = __shfl(nslot, 0, 64);
                if (nslot >= 0) {
                    float* fo = S.feat_on ? S.feat + nbi * SG_FEAT : nullptr;
                    // the child counts on the block its parent counted on; that block has
                    // applied the parent's first pk items (all of them for a block
                    // projected at the parent this round)
                    const int64_t view = S.bview[g];
                    const int32_t pk = view >= 0 ? (S.bpnew[g] ? m.k : m.pk) : 0;
                    sg_make_record(S, sg, sup_all + seg_item_base[m.seg], pair, nslot, sX[sl], k, s_sup[sl], s_e[sl],
                                   m.seg, sF[sl], m.sup, m.ncand, m.ntail, m.k == 1, fo, s_key + rank, S.key[slot],
                                   S.sib_off[g], view, pk);
                    if (lane == 0) S.newborn[nbi] = nslot;
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_counts) {
        atomicAdd((unsigned long long*)&S.ctl[CTL_COUNTS], s_counts);
        atomicAdd((unsigned long long*)&S.ctl[CTL_WORDS], s_words);
    }
}

// ------------------------------------------------------------------ commit: push newborns, release the batch
// Newborn records go onto their priority bucket; the batch's record slots go back to
// the free list (their children have all been expanded). The two index spaces are
// independent, so one grid-stride launch does both. The host sums the bucket tops.
__global__ __launch_bounds__(256) void k_commit(SgSearchDev S, const SgSegment* __restrict__ segs,
                                                int32_t order_mode, const float* __restrict__ scores,
                                                int32_t release)
{
    // bucket positions are reserved once per workgroup and bucket (an LDS histogram gives
    // each newborn its rank), not once per newborn
    __shared__ int32_t s_cnt[SG_NBUCKETS];
    __shared__ int64_t s_base[SG_NBUCKETS];
    __shared__ int64_t s_free;
    const int64_t n = min(S.ctl[CTL_NEWBORN], S.newborn_cap);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {   // workgroup-uniform
        const int64_t i = i0 + threadIdx.x;
        if (threadIdx.x < SG_NBUCKETS) s_cnt[threadIdx.x] = 0;
        __syncthreads();
        int b = -1, rank = 0;
        int32_t slot = -1;
        if (i < n) {
            slot = S.newborn[i];
            if (slot >= 0 && slot < S.cap) {
                const SgNodeMeta m = S.meta[slot];
                if (order_mode == 1) {
                    const float sc = scores[i];
                    S.meta[slot].score = sc;
                    b = sg_bucket_score(sc, m.ncand, m.ntail);
                } else if (order_mode == 2) {
                    b = min(m.k, SG_NBUCKETS - 1);
                } else {
                    b = sg_bucket_support(m.sup, segs[m.seg].thr);
                }
                rank = atomicAdd(&s_cnt[b], 1);
            }
        }
        __syncthreads();
        if (threadIdx.x < SG_NBUCKETS && s_cnt[threadIdx.x])
            s_base[threadIdx.x] = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_BUCKET_TOP0 + threadIdx.x],
                                                     (unsigned long long)s_cnt[threadIdx.x]);
        __syncthreads();
        if (b >= 0) {
            S.bucket[(int64_t)b * S.cap + s_base[b] + rank] = slot;
            if (S.trace_cap) {
                const int64_t t = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_TRACE_BIRTH], 1ull);
                if (t < S.trace_cap) {
                    S.trace_key[2 * t] = S.key[slot];
                    S.trace_key[2 * t + 1] = S.pkey[slot];
                    for (int f = 0; f < SG_FEAT; f++) S.trace_feat[t * SG_FEAT + f] = S.feat[i * SG_FEAT + f];
                }
            }
        }
        __syncthreads();   // s_cnt and s_base are reused by the next iteration
    }
    if (release) {   // the batch's record slots go back onto the free list, one reservation per workgroup
        const int64_t nb = S.ctl[CTL_BATCH_N];
        for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < nb; i0 += stride) {
            const int64_t cnt = min((int64_t)blockDim.x, nb - i0);
            if (threadIdx.x == 0) s_free = (int64_t)atomicAdd((unsigned long long*)&S.ctl[CTL_FREE_TOP],
                                                              (unsigned long long)cnt);
            __syncthreads();
            if (threadIdx.x < cnt) {
                const int64_t pos = s_free + threadIdx.x;
                if (pos >= 0 && pos < S.cap) S.free_idx[pos] = S.batch[i0 + threadIdx.x];   // out of range only after an overflow
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ launchers
// Resident workgroups per CU of the tile kernel (sizes its persistent grid).
int sg_count_group_blocks_per_cu(int32_t chunk_words, int32_t kind)
{
    int n = 0;
    hipError_t e;
    if (kind == SG_COUNT_F4) {
        switch (chunk_words) {
        case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_f4<128>, SG_TILE_THREADS, 0); break;
        case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_f4<512>, SG_TILE_THREADS, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_f4<256>, SG_TILE_THREADS, 0); break;
        }
    } else {
        switch (chunk_words) {
        case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_group<128>, SG_TILE_THREADS, 0); break;
        case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_group<512>, SG_TILE_THREADS, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_count_group<256>, SG_TILE_THREADS, 0); break;
        }
    }
    if (e != hipSuccess || n < 1) {
        (void)hipGetLastError();
        n = 1;
    }
    return n;
}

hipError_t sg_search_init(const SgSearchDev& S, hipStream_t s)
{
    const int64_t n = S.cap > CTL_SLOTS ? S.cap : CTL_SLOTS;
    hipLaunchKernelGGL(k_search_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S);
    return hipGetLastError();
}

static hipError_t sg_commit(const SgSearchDev& S, const SgSegment* segs, int32_t order_mode, const SgScorer* scorer,
                            int32_t release, hipStream_t s)
{
    if (order_mode == 1) {
        hipError_t e = sg_launch_score(scorer, S.feat, (int32_t)S.newborn_cap, &S.ctl[CTL_NEWBORN], S.scores, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_commit, dim3(1024), dim3(256), 0, s, S, segs, order_mode, (const float*)S.scores, release);
    return hipGetLastError();
}

hipError_t sg_search_seed(const SgSearchDev& S, const SgSegment* segs, const int32_t* pair,
                          const uint64_t* ok, const int32_t* roots, int32_t n_roots,
                          const int32_t* sup_all, const int64_t* seg_item_base, int32_t order_mode,
                          const SgScorer* scorer, hipStream_t s)
{
    if (n_roots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_newborn, dim3(1), dim3(1), 0, s, S);
    hipLaunchKernelGGL(k_seed, dim3((unsigned)((n_roots + SG_SEED_WAVES - 1) / SG_SEED_WAVES)), dim3(SG_SEED_WAVES * 64), 0,
                       s, S, segs, pair, ok, roots, n_roots,
                       sup_all, seg_item_base);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sg_commit(S, segs, order_mode, scorer, 0, s);
}

hipError_t sg_search_round(const SgSearchDev& S, const uint64_t* rows, const SgSegment* segs,
                           const int32_t* pair, const uint64_t* ok, const int32_t* sup_all,
                           const int64_t* seg_item_base, int32_t uniform, int32_t order_mode,
                           const SgScorer* scorer, const hipEvent_t* ev, hipStream_t s)
{
    (void)uniform;
    hipError_t e;
    // ev (direct launches only): SG_STAGES + 1 events bracketing the round's stages
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], s) : hipSuccess; };
    if ((e = mark(0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_select, dim3(1), dim3(SEL_THREADS), 0, s, S, segs);
    hipLaunchKernelGGL(k_candlist, dim3((unsigned)((S.bmax + 3) / 4)), dim3(256), 0, s, S, segs);
    if ((e = mark(1)) != hipSuccess) return e;
    if (S.proj_factor > 0) {
        hipLaunchKernelGGL(k_proj_tids, dim3(2048), dim3(256), 0, s, S, rows, segs);
        if ((e = mark(2)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_proj_gather, dim3(2048), dim3(256), 0, s, S, rows, segs);
        if ((e = mark(3)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_proj_compress, dim3(8192), dim3(64), 0, s, S, rows, segs);
    } else if ((e = mark(2)) != hipSuccess || (e = mark(3)) != hipSuccess) {
        return e;
    }
    if ((e = mark(4)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_need_rows, dim3(1024), dim3(256), 0, s, S, segs, ok);
    if ((e = mark(5)) != hipSuccess) return e;
    if (S.count_kind == SG_COUNT_F4) {
        switch (S.chunk_words) {
        case 128: hipLaunchKernelGGL(k_count_f4<128>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        case 512: hipLaunchKernelGGL(k_count_f4<512>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        default: hipLaunchKernelGGL(k_count_f4<256>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        }
    } else {
        switch (S.chunk_words) {
        case 128: hipLaunchKernelGGL(k_count_group<128>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        case 512: hipLaunchKernelGGL(k_count_group<512>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        default: hipLaunchKernelGGL(k_count_group<256>, dim3(S.count_grid), dim3(SG_TILE_THREADS), 0, s, S, rows, segs, ok); break;
        }
    }
    if ((e = mark(6)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_expand_children, dim3(1024), dim3(SG_EXP_WAVES * 64), 0, s, S, segs, pair, ok, sup_all,
                       seg_item_base);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(7)) != hipSuccess) return e;
    if ((e = sg_commit(S, segs, order_mode, scorer, 1, s)) != hipSuccess) return e;
    if ((e = mark(8)) != hipSuccess) return e;
    return hipGetLastError();
}
