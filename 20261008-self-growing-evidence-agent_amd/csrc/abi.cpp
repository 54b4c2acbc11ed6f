// abi.cpp -- host side of the sgfim C ABI (include/sgfim.h).
//
// A context owns one HIP stream on one device and every device buffer of one
// loaded transaction set (or one set of windows). The search is driven in rounds:
// each round is one replay of a captured hipGraph (or the same kernels launched
// directly), followed by one small device-to-host copy of the control block and,
// when MFIs were confirmed, of the new MFI ring entries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sgfim.h"
#include "sgfim_impl.h"

#define SGFIM_VERSION "sgfim 0.1 (gfx950)"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct sgfim_result {
    std::vector<int32_t> items;
    std::vector<int64_t> off{0};
    std::vector<int64_t> sup;
    std::vector<int32_t> round;
    std::vector<double> time_s;
    std::vector<int32_t> seg;
};

// Device buffers of the search, cached per context and reused by later runs that fit.
struct SearchBufs {
    SgSearchDev D{};
    std::vector<DevBuf> bufs;
    hipGraph_t graph[2] = {nullptr, nullptr};       // one captured round per frontier order
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int32_t graph_feat[2] = {0, 0};                 // the D.feat_on each round was captured with
    int32_t graph_ck[2] = {0, 0};                   // the D.count_kind each round was captured with
    uint64_t graph_gen[2] = {0, 0};                 // the search universe each round was captured on
    hipEvent_t ev[SG_STAGES + 1] = {};             // around the round's stages (direct launches)
    int64_t* h_ctl = nullptr;                       // pinned: processed copy of the control block, then
                                                    // two snapshots of rounds in flight
    hipEvent_t snap_ev[2] = {nullptr, nullptr};     // a round and its snapshot have completed
    uint16_t* h_ring_items = nullptr;               // MFI ring, pinned host memory mapped into the
    int32_t* h_ring_meta = nullptr;                 // device: kernels write MFIs straight to the host
    int32_t* d_roots = nullptr;
    int64_t roots_cap = 0;
};

// The item universe a search runs on: the loaded items, or a compact copy of the items
// frequent at the search's thresholds (when more than SG_DMAX distinct items are loaded).
// Ids inside the search are dense ids of this universe.
struct Universe {
    int32_t d = 0, dpad = 0;
    int64_t W = 0, Wp = 0;
    uint64_t* rows = nullptr;           // rows arena (segment row_base / stride)
    int32_t* sup = nullptr;             // [nseg][dpad] item supports
    int32_t* pair = nullptr;            // [nseg][dpad][dpad] pair matrix
    bool* pair_valid = nullptr;
    uint64_t* ok = nullptr;             // pair-ok masks, one block per search segment
    const int64_t* h_sup_seg = nullptr; // host [nseg][d]
    const SgSegment* segs = nullptr;    // host segment table
    SgSegment* d_seg = nullptr;         // device segment table
    const int32_t* ids = nullptr;       // dense -> original id
    uint64_t gen = 0;                   // changes whenever these buffers do
};

// One search run.
struct SearchState {
    bool active = false;
    bool done = false;
    int status = SGFIM_OK;
    sgfim_mine_args args{};
    int order_mode = 0;
    bool use_graph = true;
    int64_t ring_read = 0;
    int64_t rounds = 0;
    int64_t pair_counts = 0;
    std::chrono::steady_clock::time_point t0;
    sgfim_result pending;            // MFIs confirmed but not yet taken
    sgfim_round_stats stats{};
    bool cancel = false;
    int32_t n_roots = 0;
    double count_ms = 0.0;           // device wall-clock duration of the counting kernel
    double event_ms = 0.0;           // HIP-event duration (direct launches only)
    double stage_ms[SG_STAGES] = {};   // HIP-event duration of each round stage (direct launches only)
    int64_t count_launches = 0;
    // rounds queued on the stream and not yet processed (their snapshot slots are
    // [pipe_head, pipe_head + pipe_n) mod 2); they carry over from one step call to the next
    int pipe_n = 0, pipe_head = 0;
    // queued rounds already counted in count_launches (the rounds left queued when the
    // search ended: they run after the one that emptied the frontier and take nothing)
    int pipe_counted = 0;
    SearchBufs* B = nullptr;         // the context's buffers used by this run
    Universe U;                      // the items it searches
    // host sources of the run's asynchronous uploads, alive until the next run
    std::vector<SgSegment> vsegs;
    std::vector<int64_t> vibase;
    std::vector<int32_t> roots;
};

static void pipe_settle(sgfim_ctx* c);

struct sgfim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t mem_cap = 0;
    int64_t mem_used = 0;
    int wall_khz = 0;                   // wall_clock64() rate
    int n_cu = 256;                     // compute units of the device
    std::string dev_name;               // gcnArchName of the device
    int32_t tile_cw = 0, tile_cr = 0;   // counting-tile shape of the search (0: default)
    int64_t sib_words = 0;              // sibling-ring words of the search (0: default)
    void* scratch = nullptr;            // this device's static count-batch scratch (sg_count_scratch)
    int32_t proj_factor = 4;            // projected blocks: factor (0: off)
    int64_t proj_arena = 0;             // projected-block arena words (0: default)
    int32_t proj_dense = 32;            // blocks at least 1/proj_dense dense are compacted word by word (0: never)
    int32_t proj_mintail = 16;          // children a record needs for such a block
    sgfim_opts opts{};
    // loaded data
    bool loaded = false;
    int64_t n_txn = 0;
    int32_t d = 0, dpad = 0;
    int64_t W = 0, Wp = 0;
    std::vector<int32_t> ids;           // dense -> original id
    std::vector<int64_t> h_sup;         // item supports of window 0
    std::vector<int64_t> h_sup_seg;     // item supports per window [nseg][d]
    int32_t nseg = 0;                   // windows (segments)
    std::vector<SgSegment> segs;        // host copy of the segment table
    int64_t rows_words = 0;             // words in the rows arena
    std::vector<int64_t> win_thr;       // per-window min_count for the next run (empty: args.min_count)
    uint64_t* d_rows = nullptr;
    int32_t* d_sup = nullptr;
    int32_t* d_pair = nullptr;
    bool pair_valid = false;
    int32_t pair_kind = SG_PAIR_FP4;    // pair-matrix kernel (sgfim_set_pair_kernel)
    int32_t pair_split_cap = 0;         // its transaction splits at most (0: automatic)
    int32_t count_kind = SG_COUNT_VALU;   // counting kernel of the searches (sgfim_set_count_kernel)
    hipEvent_t pair_ev[2] = {nullptr, nullptr};   // around the latest pair-matrix computation
    bool pair_timed = false;
    int32_t* h_pair_pin = nullptr;      // pinned staging buffer for pair-matrix read-back (dpad x dpad)
    int64_t h_pair_pin_n = 0;
    char* h_stage = nullptr;            // pinned staging of a search's uploads (segments, item bases, roots)
    size_t h_stage_cap = 0;
    uint64_t* d_ok = nullptr;           // pair-ok masks, one block per search segment (vseg_cap blocks)
    SgSegment* d_seg = nullptr;         // data segments (windows) + one for count_batch
    int64_t* d_seg_item_base = nullptr;
    // Search segments: one per window, or one per threshold of a single set (a batched
    // multi-threshold search); stable device arrays so captured rounds stay valid.
    SgSegment* d_vseg = nullptr;
    int64_t* d_vseg_ibase = nullptr;
    int32_t vseg_cap = 0;
    std::vector<int64_t> multi_thr;     // thresholds of the next run on a single set (empty: one)
    std::vector<int32_t> multi_topk;    // per-threshold root counts for SGFIM_ROOT_TOPK (empty: args)
    std::vector<DevBuf> data_bufs;
    std::vector<int32_t> sensor_items;  // [windows][sensors] items of the latest sensor load
    // proposal network
    bool scorer_set = false;
    SgScorer scorer{};
    std::vector<DevBuf> scorer_bufs;
    // trace
    bool trace = false;
    int64_t trace_cap = 0;
    // search
    SearchState* S = nullptr;
    SearchBufs* SB = nullptr;
    int64_t cap_hint = 0;               // frontier capacity learned from an earlier overflow
    sgfim_round_stats last{};
    // compact search universe (more than SG_DMAX distinct items loaded): the items frequent
    // at the latest such search's thresholds, gathered from the loaded rows
    struct {
        std::vector<int32_t> keep;      // compact -> loaded dense id
        std::vector<int32_t> ids;       // compact -> original id
        std::vector<int64_t> h_sup_seg; // [nseg][d]
        std::vector<SgSegment> segs;
        int32_t d = 0, dpad = 0;
        int64_t rows_cap = 0, sup_cap = 0, pair_cap = 0, ok_cap = 0;
        int32_t seg_cap = 0;
        uint64_t* rows = nullptr;
        int32_t* sup = nullptr;
        int32_t* pair = nullptr;
        uint64_t* ok = nullptr;
        SgSegment* d_seg = nullptr;
        bool pair_valid = false;
        bool built = false;
        uint64_t gen = 0;
        std::vector<DevBuf> bufs;
    } cu;
    uint64_t gen_counter = 0;           // source of universe generations
    uint64_t loaded_gen = 0;            // generation of the loaded items' buffers
};

namespace {

int fail(sgfim_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), SGFIM_E_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                       \
    } while (0)

int dev_alloc(sgfim_ctx* c, std::vector<DevBuf>& owner, void** p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (c->mem_cap > 0 && c->mem_used + (int64_t)bytes > c->mem_cap)
        return fail(c, SGFIM_E_NOMEM, "device memory cap of %lld bytes exceeded (need %zu more, %lld in use)",
                    (long long)c->mem_cap, bytes, (long long)c->mem_used);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, SGFIM_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    owner.push_back(DevBuf{*p, bytes});
    c->mem_used += (int64_t)bytes;
    return SGFIM_OK;
}

template <typename T>
int dev_alloc_t(sgfim_ctx* c, std::vector<DevBuf>& owner, T** p, size_t count)
{
    void* q = nullptr;
    int rc = dev_alloc(c, owner, &q, count * sizeof(T));
    *p = (T*)q;
    return rc;
}

void dev_free_all(sgfim_ctx* c, std::vector<DevBuf>& owner)
{
    for (auto& b : owner) {
        if (b.p) (void)hipFree(b.p);
        c->mem_used -= (int64_t)b.bytes;
    }
    owner.clear();
}

void search_free(sgfim_ctx* c)
{
    delete c->S;
    c->S = nullptr;
}

// Pinned staging of at least `bytes` for a search's uploads. Called after the search's
// stream is synchronized, so earlier uploads from it have completed.
int stage_reserve(sgfim_ctx* c, size_t bytes)
{
    if (bytes <= c->h_stage_cap) return SGFIM_OK;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->h_stage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, (size_t)1 << 16);
    if (hipHostMalloc((void**)&c->h_stage, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->h_stage = nullptr;
        return fail(c, SGFIM_E_NOMEM, "pinned staging allocation failed");
    }
    c->h_stage_cap = cap;
    return SGFIM_OK;
}

// Drop captured round g: its kernel arguments name the loaded data's buffers.
void graph_free(SearchBufs* B, int g)
{
    if (B->exec[g]) (void)hipGraphExecDestroy(B->exec[g]);
    if (B->graph[g]) (void)hipGraphDestroy(B->graph[g]);
    B->exec[g] = nullptr;
    B->graph[g] = nullptr;
}

void bufs_free(sgfim_ctx* c)
{
    SearchBufs* B = c->SB;
    if (!B) return;
    for (int i = 0; i < 2; i++) graph_free(B, i);
    if (B->h_ctl) (void)hipHostFree(B->h_ctl);
    if (B->h_ring_items) (void)hipHostFree(B->h_ring_items);
    if (B->h_ring_meta) (void)hipHostFree(B->h_ring_meta);
    for (int i = 0; i <= SG_STAGES; i++)
        if (B->ev[i]) (void)hipEventDestroy(B->ev[i]);
    for (int i = 0; i < 2; i++)
        if (B->snap_ev[i]) (void)hipEventDestroy(B->snap_ev[i]);
    dev_free_all(c, B->bufs);
    delete B;
    c->SB = nullptr;
}

// Frees the loaded data. The search buffers (frontier, buckets, pinned MFI ring) stay for
// the next search on new data when they fit it; only the captured rounds, whose arguments
// name the old data's buffers, are dropped.
void universe_free(sgfim_ctx* c)
{
    auto& u = c->cu;
    dev_free_all(c, u.bufs);
    u.rows = nullptr;
    u.sup = nullptr;
    u.pair = nullptr;
    u.ok = nullptr;
    u.d_seg = nullptr;
    u.rows_cap = u.sup_cap = u.pair_cap = u.ok_cap = 0;
    u.seg_cap = 0;
    u.keep.clear();
    u.pair_valid = false;
    u.built = false;
}

void data_free(sgfim_ctx* c)
{
    search_free(c);
    if (c->SB)
        for (int i = 0; i < 2; i++) graph_free(c->SB, i);
    universe_free(c);
    c->loaded_gen = ++c->gen_counter;
    dev_free_all(c, c->data_bufs);
    c->d_rows = nullptr;
    c->d_sup = nullptr;
    c->d_pair = nullptr;
    c->d_ok = nullptr;
    c->d_seg = nullptr;
    c->d_seg_item_base = nullptr;
    c->d_vseg = nullptr;
    c->d_vseg_ibase = nullptr;
    c->vseg_cap = 0;
    c->multi_thr.clear();
    c->multi_topk.clear();
    c->pair_valid = false;
    c->win_thr.clear();
    c->sensor_items.clear();
    c->loaded = false;
}

int32_t pick_chunk(const sgfim_ctx* c, int64_t W)
{
    if (c->opts.chunk_words == 256 || c->opts.chunk_words == 512 || c->opts.chunk_words == 1024 ||
        c->opts.chunk_words == 2048)
        return c->opts.chunk_words;
    if (W <= 256) return 256;
    if (W <= 4096) return 512;
    return 1024;
}

// Row-chunk words of one counting tile of the search (the tile kernel holds 16 child
// prefixes of one chunk in LDS; 256 words keep three 512-thread workgroups resident per CU),
// and chunks per counting unit.
int32_t search_chunk(const sgfim_ctx* c)
{
    if (c->tile_cw) return c->tile_cw;
    const int32_t cw = c->opts.chunk_words;
    return (cw == 128 || cw == 256 || cw == 512) ? cw : 256;
}

int32_t search_range(const sgfim_ctx* c) { return c->tile_cr ? c->tile_cr : 4; }

// The loaded items as a search universe (pair matrix and pair-ok masks as allocated so far).
Universe loaded_universe(sgfim_ctx* c)
{
    Universe U;
    U.d = c->d;
    U.dpad = c->dpad;
    U.W = c->W;
    U.Wp = c->Wp;
    U.rows = c->d_rows;
    U.sup = c->d_sup;
    U.pair = c->d_pair;
    U.pair_valid = &c->pair_valid;
    U.ok = c->d_ok;
    U.h_sup_seg = c->h_sup_seg.data();
    U.segs = c->segs.data();
    U.d_seg = c->d_seg;
    U.ids = c->ids.data();
    U.gen = c->loaded_gen;
    return U;
}

Universe compact_universe(sgfim_ctx* c)
{
    auto& u = c->cu;
    Universe U;
    U.d = u.d;
    U.dpad = u.dpad;
    U.W = c->W;
    U.Wp = c->Wp;
    U.rows = u.rows;
    U.sup = u.sup;
    U.pair = u.pair;
    U.pair_valid = &u.pair_valid;
    U.ok = u.ok;
    U.h_sup_seg = u.h_sup_seg.data();
    U.segs = u.segs.data();
    U.d_seg = u.d_seg;
    U.ids = u.ids.data();
    U.gen = u.gen;
    return U;
}

// Pair matrix of universe U (unless still valid), timed with the context's pair events.
int ensure_pair_u(sgfim_ctx* c, const Universe& U)
{
    if (*U.pair_valid) return SGFIM_OK;
    const bool timed = c->pair_ev[0] && hipEventRecord(c->pair_ev[0], c->stream) == hipSuccess;
    if (c->nseg == 1) {
        HIPCHK(c, sg_launch_pair_mfma(U.rows, U.dpad, U.Wp, U.W, U.pair, c->pair_kind, c->pair_split_cap,
                                      c->stream));
    } else {
        int64_t maxw = 0;
        for (int32_t w = 0; w < c->nseg; w++) maxw = std::max(maxw, U.segs[w].words);
        HIPCHK(c, sg_launch_zero32(U.pair, (int64_t)U.dpad * U.dpad * c->nseg, c->stream));
        HIPCHK(c, sg_launch_pair_mfma_segs(U.rows, U.d_seg, c->nseg, U.dpad, maxw, U.pair, c->pair_kind,
                                           c->pair_split_cap, c->stream));
    }
    c->pair_timed = timed && hipEventRecord(c->pair_ev[1], c->stream) == hipSuccess;
    *U.pair_valid = true;
    return SGFIM_OK;
}

// Pair matrix of the loaded items.
int ensure_pair(sgfim_ctx* c)
{
    if (c->pair_valid) return SGFIM_OK;
    if (!c->d_pair) {
        int rc = dev_alloc_t(c, c->data_bufs, &c->d_pair, (size_t)c->dpad * c->dpad * c->nseg);
        if (rc) return rc;
    }
    return ensure_pair_u(c, loaded_universe(c));
}

// Compact search universe of the loaded items `keep` (ascending dense ids, at most SG_DMAX):
// their rows and supports in every window, gathered on the device. The same items as the
// current copy keep it (and its pair matrix); buffers only grow.
int compact_build(sgfim_ctx* c, std::vector<int32_t>& keep)
{
    auto& u = c->cu;
    if (u.built && u.keep == keep) return SGFIM_OK;
    const int32_t dc = (int32_t)keep.size();
    const int32_t dcpad = std::max(128, (dc + 127) / 128 * 128);
    const int32_t dwc = (dc + 63) / 64;
    const int32_t nseg = c->nseg;
    std::vector<SgSegment> segs(c->segs.begin(), c->segs.begin() + nseg);
    int64_t base = 0;
    for (int32_t w = 0; w < nseg; w++) {
        SgSegment& g = segs[w];
        g.row_base = base;
        base += (int64_t)dcpad * g.stride;
        g.d = dc;
        g.dpad = dcpad;
        g.dw = dwc;
        g.pair_base = (int64_t)w * dcpad * dcpad;
        g.ok_base = (int64_t)w * dc * dwc;
    }
    const int64_t sup_need = (int64_t)dcpad * nseg, pair_need = (int64_t)dcpad * dcpad * nseg;
    const int64_t ok_need = (int64_t)dc * dwc * c->vseg_cap;
    if (base > u.rows_cap || sup_need > u.sup_cap || pair_need > u.pair_cap || ok_need > u.ok_cap ||
        nseg > u.seg_cap) {
        universe_free(c);
        int rc;
        if ((rc = dev_alloc_t(c, u.bufs, &u.rows, (size_t)base)) || (rc = dev_alloc_t(c, u.bufs, &u.sup, (size_t)sup_need)) ||
            (rc = dev_alloc_t(c, u.bufs, &u.pair, (size_t)pair_need)) || (rc = dev_alloc_t(c, u.bufs, &u.ok, (size_t)ok_need)) ||
            (rc = dev_alloc_t(c, u.bufs, &u.d_seg, (size_t)nseg))) {
            universe_free(c);
            return rc;
        }
        u.rows_cap = base;
        u.sup_cap = sup_need;
        u.pair_cap = pair_need;
        u.ok_cap = ok_need;
        u.seg_cap = nseg;
    }
    u.built = false;
    std::vector<DevBuf> tmp;
    int32_t* d_keep = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, tmp, &d_keep, (size_t)dc))) return rc;
    hipError_t e = hipMemcpyAsync(d_keep, keep.data(), (size_t)dc * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(u.d_seg, segs.data(), (size_t)nseg * sizeof(SgSegment), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(u.rows, 0, (size_t)base * 8, c->stream);   // pad rows stay zero
    if (e == hipSuccess) e = hipMemsetAsync(u.sup, 0, (size_t)sup_need * 4, c->stream);
    if (e == hipSuccess)
        e = sg_launch_gather_rows(c->d_rows, c->d_sup, c->d_seg, c->dpad, u.rows, u.sup, u.d_seg, dcpad, d_keep, dc, nseg,
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free_all(c, tmp);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "compacting the search's items failed: %s", hipGetErrorString(e));
    u.h_sup_seg.assign((size_t)dc * nseg, 0);
    u.ids.resize(dc);
    for (int32_t i = 0; i < dc; i++) u.ids[i] = c->ids[keep[i]];
    for (int32_t w = 0; w < nseg; w++)
        for (int32_t i = 0; i < dc; i++) u.h_sup_seg[(size_t)w * dc + i] = c->h_sup_seg[(size_t)w * c->d + keep[i]];
    u.keep = keep;
    u.segs = std::move(segs);
    u.d = dc;
    u.dpad = dcpad;
    u.pair_valid = false;
    u.built = true;
    u.gen = ++c->gen_counter;
    return SGFIM_OK;
}

// Search-segment tables for up to max(windows, 256) segments (allocated with the data).
int alloc_vsegs(sgfim_ctx* c)
{
    c->vseg_cap = std::max<int32_t>(c->nseg, 256);
    int rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_vseg, (size_t)c->vseg_cap))) return rc;
    return dev_alloc_t(c, c->data_bufs, &c->d_vseg_ibase, (size_t)c->vseg_cap);
}

double now_s(const SearchState* S)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t0).count();
}

// Copy MFIs [ring_read, pos) from the ring into S->pending. The ring lives in mapped host
// memory and the stream has been synchronised, so entries are read in place. The outputs
// are sized once (a prefix over the itemset lengths); large hand-overs are filled by
// several threads, each converting a contiguous range of entries.
int drain_ring(sgfim_ctx* c, SearchState* S, int64_t pos, double t)
{
    const int64_t n = pos - S->ring_read;
    if (n <= 0) return SGFIM_OK;
    const int64_t cap = S->B->D.ring_cap;
    if (n > cap) return fail(c, SGFIM_E_CAPACITY, "MFI ring overflow (%lld new entries, capacity %lld)",
                             (long long)n, (long long)cap);
    const int32_t* ring_meta = S->B->h_ring_meta;      // [cap][4]: k, sup, seg, round
    const uint16_t* ring_items = S->B->h_ring_items;   // [cap][SG_MAX_DEPTH]
    const int64_t r0 = S->ring_read % cap;
    auto slot = [&](int64_t i) { const int64_t s = r0 + i; return s < cap ? s : s - cap; };
    // Every entry is checked before anything is sized or copied from it: an itemset length
    // outside 1..SG_MAX_DEPTH, an item outside the universe or a segment outside the search
    // would make the copies below read or write out of bounds. Such an entry can only come
    // from a device-side fault, so it stops the run with SGFIM_E_DEVICE.
    const int32_t d = S->U.d;
    const int32_t nseg_search = (int32_t)S->vsegs.size();
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = slot(i);
        if (s < 0 || s >= cap) return fail(c, SGFIM_E_DEVICE, "MFI ring slot %lld outside [0, %lld)", (long long)s, (long long)cap);
        const int32_t* m = ring_meta + s * 4;
        if (m[0] < 1 || m[0] > SG_MAX_DEPTH || m[2] < 0 || m[2] >= nseg_search)
            return fail(c, SGFIM_E_DEVICE, "corrupt MFI ring entry %lld (length %d, segment %d)", (long long)(S->ring_read + i),
                        m[0], m[2]);
        const uint16_t* it = ring_items + s * SG_MAX_DEPTH;
        for (int j = 0; j < m[0]; j++)
            if ((int32_t)it[j] >= d)
                return fail(c, SGFIM_E_DEVICE, "corrupt MFI ring entry %lld (item %d of %d)", (long long)(S->ring_read + i),
                            (int)it[j], d);
    }
    sgfim_result& R = S->pending;
    const size_t n0 = R.sup.size();
    R.off.resize(n0 + 1 + (size_t)n);
    for (int64_t i = 0; i < n; i++) R.off[n0 + 1 + i] = R.off[n0 + i] + ring_meta[slot(i) * 4];
    R.items.resize((size_t)R.off[n0 + n]);
    R.sup.resize(n0 + (size_t)n);
    R.round.resize(n0 + (size_t)n);
    R.time_s.resize(n0 + (size_t)n);
    R.seg.resize(n0 + (size_t)n);
    const int32_t* ids = S->U.ids;
    auto convert = [&](int64_t lo, int64_t hi) {
        uint16_t x[SG_MAX_DEPTH];
        for (int64_t i = lo; i < hi; i++) {
            const int64_t s = slot(i);
            const int32_t* m = ring_meta + s * 4;
            const int k = m[0];
            std::copy(ring_items + s * SG_MAX_DEPTH, ring_items + s * SG_MAX_DEPTH + k, x);
            // descending item ids: the order in which the reference's search adds items (the
            // root is the largest id, every child a smaller one), hence its Results lines
            std::sort(x, x + k, [](uint16_t a, uint16_t b) { return a > b; });
            int32_t* out = R.items.data() + R.off[n0 + i];
            for (int j = 0; j < k; j++) out[j] = ids[x[j]];
            R.sup[n0 + i] = m[1];
            R.seg[n0 + i] = m[2];
            R.round[n0 + i] = m[3];
            R.time_s[n0 + i] = t;
        }
    };
    const int64_t per_thread = 1 << 15;
    const int nt = (int)std::min<int64_t>({(n + per_thread - 1) / per_thread,
                                           (int64_t)std::max(1u, std::thread::hardware_concurrency()), 16});
    if (nt <= 1) {
        convert(0, n);
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < nt; k++) pool.emplace_back(convert, n * k / nt, n * (k + 1) / nt);
        for (auto& th : pool) th.join();
    }
    S->ring_read = pos;
    return SGFIM_OK;
}

int read_ctl(sgfim_ctx* c, SearchState* S)
{
    HIPCHK(c, hipMemcpyAsync(S->B->h_ctl, S->B->D.ctl, CTL_SLOTS * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SGFIM_OK;
}

// Records waiting in the priority buckets of a control-block snapshot.
int64_t frontier_size(const int64_t* h)
{
    int64_t f = 0;
    for (int b = 0; b < SG_NBUCKETS; b++) f += h[CTL_BUCKET_TOP0 + b];
    return f;
}

void fill_stats(sgfim_ctx* c, SearchState* S)
{
    sgfim_round_stats& st = S->stats;
    st = sgfim_round_stats{};
    if (S->B && S->B->h_ctl) {  // a run that found nothing frequent has no buffers
        const int64_t* h = S->B->h_ctl;
        st.nodes = h[CTL_NODES];
        st.counts = h[CTL_COUNTS];
        st.mfis = h[CTL_RING_POS];
        st.frontier = frontier_size(h);
        st.words = h[CTL_WORDS];
        st.projections = h[CTL_PROJS];
        st.proj_words = h[CTL_PROJ_WORDS];
        st.proj_from_rows = h[CTL_PROJ_ROWS];
        st.proj_block_words = h[CTL_PROJ_ARENA];
    }
    st.round = S->rounds;
    st.pair_counts = S->pair_counts;
    st.count_launches = S->count_launches;
    st.count_ms = S->count_ms;
    st.event_ms = S->event_ms;
    st.elapsed_s = now_s(S);
    st.done = S->done ? 1 : 0;
    st.status = S->status;
    c->last = st;
}

int check_err(sgfim_ctx* c, SearchState* S)
{
    const int64_t e = S->B->h_ctl[CTL_ERR];
    if (e & 4) return fail(c, SGFIM_E_DEVICE, "projected block inconsistent (transaction list or item outside its U)");
    if (e & 2) return fail(c, SGFIM_E_DEPTH, "an itemset longer than %d items would be needed", SG_MAX_DEPTH);
    if (e & 1)
        return fail(c, SGFIM_E_CAPACITY,
                    "frontier capacity exhausted (cap %lld nodes, newborn cap %lld); raise opts.frontier_cap",
                    (long long)S->B->D.cap, (long long)S->B->D.newborn_cap);
    return SGFIM_OK;
}

int launch_round(sgfim_ctx* c, SearchState* S)
{
    const int32_t uniform = c->nseg == 1 ? 1 : 0;
    SearchBufs* B = S->B;
    (void)hipGetLastError();
    if (!S->use_graph) {
        HIPCHK(c, sg_search_round(B->D, S->U.rows, c->d_vseg, S->U.pair, S->U.ok, S->U.sup, c->d_vseg_ibase,
                                  uniform, S->order_mode, &c->scorer, B->ev, c->stream));
        return SGFIM_OK;
    }
    const int g = S->order_mode;
    if (B->exec[g] && (B->graph_feat[g] != B->D.feat_on || B->graph_ck[g] != B->D.count_kind ||
                       B->graph_gen[g] != S->U.gen))
        graph_free(B, g);
    if (!B->exec[g]) {
        B->graph_feat[g] = B->D.feat_on;
        B->graph_ck[g] = B->D.count_kind;
        B->graph_gen[g] = S->U.gen;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = sg_search_round(B->D, S->U.rows, c->d_vseg, S->U.pair, S->U.ok, S->U.sup,
                                       c->d_vseg_ibase, uniform, S->order_mode, &c->scorer, nullptr,
                                       c->stream);
        hipError_t e2 = hipStreamEndCapture(c->stream, &B->graph[g]);
        if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "round capture failed: %s", hipGetErrorString(e));
        if (e2 != hipSuccess) return fail(c, SGFIM_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e2));
        HIPCHK(c, hipGraphInstantiate(&B->exec[g], B->graph[g], nullptr, nullptr, 0));
    }
    HIPCHK(c, hipGraphLaunch(B->exec[g], c->stream));
    return SGFIM_OK;
}

// (Re)use the context's search buffers for a run needing the given sizes.
struct ProjSizes {
    int32_t factor = 0, dense = 0, mintail = 0;
    int64_t arena = 0, pdesc = 0, ptid = 0, pent = 0;
};

int ensure_bufs(sgfim_ctx* c, int64_t cap, int64_t newborn_cap, int64_t ring_cap, int32_t d, int32_t dw,
                int32_t bmax, int32_t chunk, int32_t range, int64_t trace_cap, int64_t n_roots, int64_t cell_cap,
                const ProjSizes& ps)
{
    // sibling ring: four rows per frontier record, 8 MB .. 2 GB, unless set
    const int64_t sib_cap = c->sib_words > 0 ? c->sib_words
                                             : std::min<int64_t>(std::max<int64_t>(cap * dw * 4, (int64_t)1 << 20), (int64_t)1 << 28);
    SearchBufs* B = c->SB;
    if (B) {
        const SgSearchDev& D = B->D;
        const bool fits = D.sib_cap == sib_cap && D.cap >= cap && D.newborn_cap >= newborn_cap && D.ring_cap >= ring_cap && D.d_max >= d &&
                          D.dw_max >= dw && D.bmax == bmax && D.chunk_words == chunk && D.chunk_range == range && D.trace_cap == trace_cap &&
                          B->roots_cap >= n_roots && D.cell_cap >= cell_cap && D.ring_cap >= n_roots + 2 * D.newborn_cap &&
                          D.proj_factor == ps.factor && D.proj_dense == ps.dense && D.proj_mintail == ps.mintail &&
                          D.parena_cap == ps.arena && D.pdesc_cap == ps.pdesc &&
                          D.ptid_cap == ps.ptid && D.pent_cap == ps.pent;
        if (fits) return SGFIM_OK;
        bufs_free(c);
    }
    B = new SearchBufs();
    c->SB = B;
    SgSearchDev& D = B->D;
    D.cap = cap;
    D.newborn_cap = newborn_cap;
    D.ring_cap = ring_cap;
    D.d_max = d;
    D.dw_max = dw;
    D.bmax = bmax;
    D.chunk_words = chunk;
    D.chunk_range = range;
    D.count_kind = c->count_kind;
    D.count_grid = sg_count_group_blocks_per_cu(chunk, D.count_kind) * c->n_cu;   // one persistent wave of workgroups
    D.trace_cap = trace_cap;
    D.cell_cap = cell_cap;
    D.sib_cap = sib_cap;
    D.proj_factor = ps.factor;
    D.proj_dense = ps.dense;
    D.proj_mintail = ps.mintail;
    D.parena_cap = ps.arena;
    D.proj_round_words = ps.arena / 4;
    D.proj_max_block = D.proj_round_words;
    D.pdesc_cap = ps.pdesc;
    D.ptid_cap = ps.ptid;
    D.pent_cap = ps.pent;
    B->roots_cap = std::max<int64_t>(n_roots, 1);
    auto& V = B->bufs;
    int rc;
    if ((rc = dev_alloc_t(c, V, &D.items, (size_t)cap * SG_MAX_DEPTH)) || (rc = dev_alloc_t(c, V, &D.meta, (size_t)cap)) ||
        (rc = dev_alloc_t(c, V, &D.cand, (size_t)cap * dw)) || (rc = dev_alloc_t(c, V, &D.key, (size_t)cap)) ||
        (rc = dev_alloc_t(c, V, &D.pkey, (size_t)cap)) || (rc = dev_alloc_t(c, V, &D.free_idx, (size_t)cap)) ||
        (rc = dev_alloc_t(c, V, &D.bucket, (size_t)cap * SG_NBUCKETS)) ||
        (rc = dev_alloc_t(c, V, &D.batch, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.coff, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.unit_off, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.choff, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.celloff, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.cells, (size_t)cell_cap)) ||
        (rc = dev_alloc_t(c, V, &D.chrec, (size_t)newborn_cap)) ||
        (rc = dev_alloc_t(c, V, &D.blkrec, (size_t)(newborn_cap / 16 + bmax + 1))) ||
        (rc = dev_alloc_t(c, V, &D.clist, (size_t)bmax * d)) || (rc = dev_alloc_t(c, V, &D.ccount, (size_t)bmax * d)) ||
        (rc = dev_alloc_t(c, V, &D.newborn, (size_t)newborn_cap)) ||
        (rc = dev_alloc_t(c, V, &D.feat, (size_t)newborn_cap * SG_FEAT)) ||
        (rc = dev_alloc_t(c, V, &D.scores, (size_t)newborn_cap)) ||
        (rc = dev_alloc_t(c, V, &D.ctl, CTL_SLOTS)) ||
        (rc = dev_alloc_t(c, V, &D.sib, (size_t)sib_cap)) || (rc = dev_alloc_t(c, V, &D.sib_off, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.need, (size_t)newborn_cap * dw)) ||
        (rc = dev_alloc_t(c, V, &B->d_roots, (size_t)B->roots_cap * 2)) ||
        (rc = dev_alloc_t(c, V, &D.bview, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.bsrc, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.bpnew, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.bnr, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.bunitB, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.bunitT, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.bunitG, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.bpa, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.bpd, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.ptoff, (size_t)bmax + 1)) ||
        (rc = dev_alloc_t(c, V, &D.pfill, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.crow, (size_t)bmax * d)) ||
        (rc = dev_alloc_t(c, V, &D.bunitC, (size_t)bmax + 1)) || (rc = dev_alloc_t(c, V, &D.peoff, (size_t)bmax + 1))) {
        bufs_free(c);
        return rc;
    }
    if (ps.factor > 0) {   // the block arena holds a ring of arena words plus one largest block past its end
        if ((rc = dev_alloc_t(c, V, &D.parena, (size_t)(ps.arena + D.proj_max_block))) ||
            (rc = dev_alloc_t(c, V, &D.pdesc, (size_t)ps.pdesc)) ||
            (rc = dev_alloc_t(c, V, &D.pmask, (size_t)ps.pdesc * dw)) ||
            (rc = dev_alloc_t(c, V, &D.prank, (size_t)ps.pdesc * dw)) || (rc = dev_alloc_t(c, V, &D.ptid, (size_t)ps.ptid)) ||
            (rc = dev_alloc_t(c, V, &D.pent, (size_t)ps.pent))) {
            bufs_free(c);
            return rc;
        }
    }
    if (trace_cap) {
        if ((rc = dev_alloc_t(c, V, &D.trace_key, (size_t)trace_cap * 2)) ||
            (rc = dev_alloc_t(c, V, &D.trace_feat, (size_t)trace_cap * SG_FEAT)) ||
            (rc = dev_alloc_t(c, V, &D.trace_mfi, (size_t)trace_cap))) {
            bufs_free(c);
            return rc;
        }
    }
    if (hipHostMalloc((void**)&B->h_ctl, 3 * CTL_SLOTS * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        bufs_free(c);
        return fail(c, SGFIM_E_NOMEM, "pinned control block allocation failed");
    }
    memset(B->h_ctl, 0, 3 * CTL_SLOTS * sizeof(int64_t));
    if (hipHostMalloc((void**)&B->h_ring_items, (size_t)ring_cap * SG_MAX_DEPTH * 2, hipHostMallocMapped) != hipSuccess ||
        hipHostMalloc((void**)&B->h_ring_meta, (size_t)ring_cap * 16, hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        bufs_free(c);
        return fail(c, SGFIM_E_NOMEM, "pinned MFI ring allocation failed");
    }
    HIPCHK(c, hipHostGetDevicePointer((void**)&D.ring_items, B->h_ring_items, 0));
    HIPCHK(c, hipHostGetDevicePointer((void**)&D.ring_meta, B->h_ring_meta, 0));
    for (int i = 0; i <= SG_STAGES; i++) HIPCHK(c, hipEventCreate(&B->ev[i]));
    for (int i = 0; i < 2; i++) HIPCHK(c, hipEventCreateWithFlags(&B->snap_ev[i], hipEventDisableTiming));
    return SGFIM_OK;
}

}  // namespace

// =====================================================================================
extern "C" {

const char* sgfim_version(void) { return SGFIM_VERSION; }

const char* sgfim_last_error(const sgfim_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int sgfim_create(const sgfim_opts* opts, sgfim_ctx** out)
{
    if (!out) return SGFIM_E_INVAL;
    *out = nullptr;
    sgfim_ctx* c = new sgfim_ctx();
    if (opts) c->opts = *opts;
    c->device = c->opts.device;
    c->mem_cap = c->opts.mem_cap_bytes;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        delete c;
        return SGFIM_E_DEVICE;
    }
    if (c->device < 0 || c->device >= n) {
        delete c;
        return SGFIM_E_INVAL;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess || strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete c;
        return SGFIM_E_DEVICE;
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SGFIM_E_DEVICE;
    }
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess) {
        (void)hipGetLastError();
        c->wall_khz = 0;
    }
    if (prop.multiProcessorCount > 0) c->n_cu = prop.multiProcessorCount;
    c->dev_name = prop.gcnArchName;
    for (int i = 0; i < 2; i++)
        if (hipEventCreate(&c->pair_ev[i]) != hipSuccess) {
            (void)hipGetLastError();
            c->pair_ev[i] = nullptr;
        }
    // the device address of this device's allocation-free scratch, looked up once per device
    if (sg_count_scratch(c->device, &c->scratch) != hipSuccess || !c->scratch) {
        (void)hipGetLastError();
        if (c->stream) (void)hipStreamDestroy(c->stream);
        for (int i = 0; i < 2; i++)
            if (c->pair_ev[i]) (void)hipEventDestroy(c->pair_ev[i]);
        delete c;
        return SGFIM_E_DEVICE;
    }
    *out = c;
    return SGFIM_OK;
}

void sgfim_destroy(sgfim_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    pipe_settle(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    data_free(c);
    bufs_free(c);
    dev_free_all(c, c->scorer_bufs);
    if (c->h_pair_pin) (void)hipHostFree(c->h_pair_pin);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    for (int i = 0; i < 2; i++)
        if (c->pair_ev[i]) (void)hipEventDestroy(c->pair_ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Load windows: window w holds transactions [win_off[w], win_off[w+1]) of the CSR.
// All windows share one item universe (the distinct ids of the whole CSR); each has
// its own bitmaps (dpad rows x Wp_w words), supports, pair matrix and threshold.
// The bitmaps of a loaded set from a CSR in device memory: dense ids from idmap (original
// id -> dense, -1 = absent; c->ids already filled), one segment per window, the GPU
// transpose per window, supports, segment tables. d_off indexes d_items globally.
static int build_from_device_csr(sgfim_ctx* c, int32_t n_windows, const int64_t* win_off, int64_t n_txn,
                                 const int64_t* d_off, const int32_t* d_items, int64_t nnz, int32_t idmax,
                                 const std::vector<int32_t>& idmap)
{
    c->n_txn = n_txn;
    c->d = (int32_t)c->ids.size();
    if (c->d > 65535) return fail(c, SGFIM_E_INVAL, "at most 65535 distinct items are supported (got %d)", c->d);
    c->dpad = std::max(128, (c->d + 127) / 128 * 128);
    c->nseg = n_windows;
    c->segs.assign(n_windows, SgSegment{});
    int64_t base = 0;
    for (int32_t w = 0; w < n_windows; w++) {
        SgSegment& g = c->segs[w];
        const int64_t n = win_off[w + 1] - win_off[w];
        g.n = (int32_t)n;
        g.words = (n + 63) / 64;
        g.stride = std::max<int64_t>(64, (g.words + 63) / 64 * 64);
        g.row_base = base;
        base += (int64_t)c->dpad * g.stride;
        g.d = c->d;
        g.dpad = c->dpad;
        g.dw = (c->d + 63) / 64;
        g.pair_base = (int64_t)w * c->dpad * c->dpad;
        g.ok_base = (int64_t)w * c->d * g.dw;
        g.thr = 1;
    }
    c->W = c->segs[0].words;
    c->Wp = c->segs[0].stride;
    c->rows_words = base;
    int rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_rows, (size_t)base))) return rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_sup, (size_t)c->dpad * n_windows))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_rows, 0, (size_t)base * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_sup, 0, (size_t)c->dpad * n_windows * 4, c->stream));
    if (nnz > 0) {
        std::vector<DevBuf> tmp;
        int32_t* d_map = nullptr;
        if ((rc = dev_alloc_t(c, tmp, &d_map, (size_t)idmax + 1))) return rc;
        hipError_t e = hipMemcpyAsync(d_map, idmap.data(), ((size_t)idmax + 1) * 4, hipMemcpyHostToDevice, c->stream);
        for (int32_t w = 0; w < n_windows && e == hipSuccess; w++) {
            const SgSegment& g = c->segs[w];
            if (g.n == 0) continue;
            e = sg_launch_transpose(d_off + win_off[w], d_items, d_map, g.n, idmax, c->d_rows + g.row_base, c->d,
                                    c->dpad, g.stride, c->d_sup + (int64_t)w * c->dpad, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dev_free_all(c, tmp);
        if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "bitmap build failed: %s", hipGetErrorString(e));
    }
    std::vector<int32_t> sup32((size_t)c->dpad * n_windows);
    HIPCHK(c, hipMemcpyAsync(sup32.data(), c->d_sup, sup32.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_sup_seg.assign((size_t)c->d * n_windows, 0);
    for (int32_t w = 0; w < n_windows; w++)
        for (int32_t i = 0; i < c->d; i++) c->h_sup_seg[(size_t)w * c->d + i] = sup32[(size_t)w * c->dpad + i];
    c->h_sup.assign(c->h_sup_seg.begin(), c->h_sup_seg.begin() + c->d);
    // segment tables: [0, nseg) search, [nseg] count_batch
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_seg, (size_t)n_windows + 1))) return rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_seg_item_base, (size_t)n_windows))) return rc;
    std::vector<int64_t> ib(n_windows);
    for (int32_t w = 0; w < n_windows; w++) ib[w] = (int64_t)w * c->dpad;
    HIPCHK(c, hipMemcpyAsync(c->d_seg_item_base, ib.data(), ib.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_seg, c->segs.data(), c->segs.size() * sizeof(SgSegment), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = alloc_vsegs(c))) return rc;
    c->loaded = true;
    return SGFIM_OK;
}


int sgfim_load_windows(sgfim_ctx* c, int32_t n_windows, const int64_t* win_off, const int64_t* off, int64_t n_txn,
                       const int32_t* items, int64_t nnz)
{
    (void)hipGetLastError();
    if (!c) return SGFIM_E_INVAL;
    pipe_settle(c);
    if (n_windows < 1 || !win_off || n_txn < 0 || nnz < 0 || (!off && n_txn > 0) || (!items && nnz > 0))
        return fail(c, SGFIM_E_INVAL, "bad CSR or window arguments");
    if (win_off[0] != 0 || win_off[n_windows] != n_txn) return fail(c, SGFIM_E_INVAL, "win_off must span 0..n_txn");
    for (int32_t w = 0; w < n_windows; w++) {
        if (win_off[w + 1] < win_off[w]) return fail(c, SGFIM_E_INVAL, "win_off is not non-decreasing at %d", w);
        if (win_off[w + 1] - win_off[w] >= (int64_t)1 << 31)
            return fail(c, SGFIM_E_INVAL, "at most 2^31-1 transactions per window are supported");
    }
    if (n_txn > 0 && (off[0] != 0 || off[n_txn] != nnz))
        return fail(c, SGFIM_E_INVAL, "txn_off must start at 0 and end at nnz");
    for (int64_t t = 0; t < n_txn; t++)
        if (off[t + 1] < off[t]) return fail(c, SGFIM_E_INVAL, "txn_off is not non-decreasing at row %lld", (long long)t);
    int32_t idmax = 0;
    for (int64_t p = 0; p < nnz; p++) {
        const int32_t x = items[p];
        if (x < 1) return fail(c, SGFIM_E_INVAL, "item id %d at position %lld is not >= 1", x, (long long)p);
        if (x > idmax) idmax = x;
    }
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    data_free(c);
    std::vector<uint8_t> present((size_t)idmax + 1, 0);
    for (int64_t p = 0; p < nnz; p++) present[items[p]] = 1;
    c->ids.clear();
    std::vector<int32_t> idmap((size_t)idmax + 1, -1);
    for (int32_t x = 1; x <= idmax; x++)
        if (present[x]) {
            idmap[x] = (int32_t)c->ids.size();
            c->ids.push_back(x);
        }
    if (nnz == 0) return build_from_device_csr(c, n_windows, win_off, n_txn, nullptr, nullptr, 0, idmax, idmap);
    std::vector<DevBuf> tmp;
    int64_t* d_off = nullptr;
    int32_t* d_items = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, tmp, &d_off, (size_t)n_txn + 1)) || (rc = dev_alloc_t(c, tmp, &d_items, (size_t)nnz))) {
        dev_free_all(c, tmp);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_off, off, ((size_t)n_txn + 1) * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_items, items, (size_t)nnz * 4, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
        dev_free_all(c, tmp);
        return fail(c, SGFIM_E_DEVICE, "CSR upload failed: %s", hipGetErrorString(e));
    }
    rc = build_from_device_csr(c, n_windows, win_off, n_txn, d_off, d_items, nnz, idmax, idmap);
    (void)hipStreamSynchronize(c->stream);
    dev_free_all(c, tmp);
    return rc;
}

// Quantile edges of one sorted column with NumPy's default (linear) method, then
// unique, then the constant-column rule of fit_bins (skab_spike.py:37-50). `val`
// holds the order statistics at the positions sensor_quantile_positions chose.
static void sensor_edges(int64_t n, int32_t B, const double* val, std::vector<double>& e)
{
    e.clear();
    const double step = 1.0 / (double)B;   // linspace(0, 1, B + 1): i * step, the last one exactly 1
    for (int32_t i = 0; i <= B; i++) {
        const double q = i == B ? 1.0 : (double)i * step;
        const double v = (double)(n - 1) * q;
        const double a = val[2 * i], b = val[2 * i + 1];
        double gamma;
        if (v >= (double)(n - 1)) gamma = v + 1.0;   // both neighbours are the last value
        else gamma = v - std::floor(v);
        const double diff = b - a;
        double r = a + diff * gamma;
        if (gamma >= 0.5) r = b - diff * (1.0 - gamma);
        e.push_back(r);
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    if (e.size() < 2) {
        const double x = e[0];
        e.assign({x - 0.5, x + 0.5});
    }
}

namespace {
int finish_sensor_load(sgfim_ctx* c, std::vector<DevBuf>& tmp, hipError_t e, int32_t S, int32_t B, int64_t n_win,
                       const int32_t* d_items, const uint32_t* d_present, const int64_t* d_off);
}

int sgfim_load_sensors(sgfim_ctx* c, const double* x, int64_t n_rows, int32_t n_sensors, int32_t n_bins,
                       int32_t window, int32_t stride)
{
    (void)hipGetLastError();
    if (!c) return SGFIM_E_INVAL;
    pipe_settle(c);
    const int32_t S = n_sensors, B = n_bins;
    if (!x || n_rows < 1 || S < 1 || B < 1 || window < 1 || window > 64 || stride < 1 || n_rows < window)
        return fail(c, SGFIM_E_INVAL, "bad sensor matrix arguments (1 <= window <= 64, window <= rows)");
    if ((int64_t)S * B > 65535) return fail(c, SGFIM_E_INVAL, "sensors x bins must be <= 65535");
    if (n_rows > (int64_t)UINT32_MAX) return fail(c, SGFIM_E_INVAL, "at most 2^32 - 1 sensor rows");
    const int64_t n_win = (n_rows - window) / stride + 1;
    if (n_win >= ((int64_t)1 << 31)) return fail(c, SGFIM_E_INVAL, "too many windows");
    for (int64_t i = 0; i < n_rows * S; i++)
        if (!std::isfinite(x[i])) return fail(c, SGFIM_E_INVAL, "sensor value %lld is not finite", (long long)i);
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    data_free(c);
    const int64_t nv = n_rows * S;
    // order statistics for the quantiles: positions floor(v) and floor(v) + 1 of
    // v = (n - 1) q for q = i / B (the last one 1), the same in every column; the select
    // takes them once each, in ascending order
    const int32_t J = 2 * (B + 1);
    std::vector<int64_t> pos((size_t)J);
    {
        const double step = 1.0 / (double)B;
        for (int32_t i = 0; i <= B; i++) {
            const double q = i == B ? 1.0 : (double)i * step;
            const double v = (double)(n_rows - 1) * q;
            int64_t lo, hi;
            if (v >= (double)(n_rows - 1)) lo = hi = n_rows - 1;
            else { lo = (int64_t)std::floor(v); hi = lo + 1; }
            pos[(size_t)i * 2] = lo;
            pos[(size_t)i * 2 + 1] = hi;
        }
    }
    std::vector<int64_t> uranks(pos);
    std::sort(uranks.begin(), uranks.end());
    uranks.erase(std::unique(uranks.begin(), uranks.end()), uranks.end());
    const int32_t R = (int32_t)uranks.size();
    std::vector<int32_t> umap((size_t)J);
    for (int32_t j = 0; j < J; j++)
        umap[(size_t)j] = (int32_t)(std::lower_bound(uranks.begin(), uranks.end(), pos[(size_t)j]) - uranks.begin());
    std::vector<int64_t> rank0((size_t)S * R);
    for (int32_t s = 0; s < S; s++) std::copy(uranks.begin(), uranks.end(), rank0.begin() + (size_t)s * R);
    const std::vector<int32_t> ngrp0((size_t)S, 1);
    std::vector<DevBuf> tmp;
    double *d_x = nullptr, *d_xc = nullptr, *d_val = nullptr, *d_edges = nullptr;
    int64_t *d_rank = nullptr, *d_off = nullptr;
    uint64_t *d_pref = nullptr, *d_gpref = nullptr;
    int32_t *d_grp = nullptr, *d_ngrp = nullptr, *d_umap = nullptr, *d_ne = nullptr, *d_items = nullptr;
    uint32_t *d_hist = nullptr, *d_present = nullptr;
    const int64_t m = (int64_t)S * J;
    const size_t sr = (size_t)S * R;
    int rc;
    if ((rc = dev_alloc_t(c, tmp, &d_x, (size_t)nv)) || (rc = dev_alloc_t(c, tmp, &d_xc, (size_t)nv)) ||
        (rc = dev_alloc_t(c, tmp, &d_rank, sr)) || (rc = dev_alloc_t(c, tmp, &d_pref, sr)) ||
        (rc = dev_alloc_t(c, tmp, &d_grp, sr)) || (rc = dev_alloc_t(c, tmp, &d_gpref, sr)) ||
        (rc = dev_alloc_t(c, tmp, &d_ngrp, (size_t)S)) || (rc = dev_alloc_t(c, tmp, &d_hist, sr * 256)) ||
        (rc = dev_alloc_t(c, tmp, &d_umap, (size_t)J)) || (rc = dev_alloc_t(c, tmp, &d_val, (size_t)m)) ||
        (rc = dev_alloc_t(c, tmp, &d_edges, (size_t)S * (B + 1))) || (rc = dev_alloc_t(c, tmp, &d_ne, (size_t)S)) ||
        (rc = dev_alloc_t(c, tmp, &d_items, (size_t)(n_win * S))) ||
        (rc = dev_alloc_t(c, tmp, &d_present, (size_t)S * B)) || (rc = dev_alloc_t(c, tmp, &d_off, (size_t)n_win + 1))) {
        dev_free_all(c, tmp);
        return rc;
    }
    auto bail = [&](const char* what, hipError_t e) {
        dev_free_all(c, tmp);
        return fail(c, SGFIM_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    // columns, then their order statistics
    hipError_t e = hipMemcpyAsync(d_x, x, (size_t)nv * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_rank, rank0.data(), sr * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ngrp, ngrp0.data(), (size_t)S * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_umap, umap.data(), (size_t)J * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_pref, 0, sr * 8, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_grp, 0, sr * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_gpref, 0, sr * 8, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_hist, 0, sr * 256 * 4, c->stream);
    if (e == hipSuccess) e = sg_sensor_columns(d_x, n_rows, S, d_xc, c->stream);
    if (e == hipSuccess)
        e = sg_sensor_select(d_xc, n_rows, S, R, d_rank, d_pref, d_grp, d_ngrp, d_gpref, d_hist, d_umap, J, d_val,
                             c->stream);
    std::vector<double> val((size_t)m);
    if (e == hipSuccess) e = hipMemcpyAsync(val.data(), d_val, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail("sensor quantiles", e);
    std::vector<double> edges((size_t)S * (B + 1), 0.0);
    std::vector<int32_t> ne(S);
    std::vector<double> es;
    for (int32_t s = 0; s < S; s++) {
        sensor_edges(n_rows, B, val.data() + (size_t)s * (B + 1) * 2, es);
        ne[s] = (int32_t)es.size();
        std::copy(es.begin(), es.end(), edges.begin() + (size_t)s * (B + 1));
    }
    // medians, bins, items, presence
    e = hipMemcpyAsync(d_edges, edges.data(), edges.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ne, ne.data(), (size_t)S * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_present, 0, (size_t)S * B * 4, c->stream);
    if (e == hipSuccess)
        e = sg_sensor_bins(d_x, S, B, window, stride, n_win, d_edges, d_ne, d_items, d_present, d_off, c->stream);
    return finish_sensor_load(c, tmp, e, S, B, n_win, d_items, d_present, d_off);
}

namespace {

// The items [n_win][S] are on the device with their presence flags and the window offsets:
// keep a host copy of the items, number the present ids densely and build the bitmaps.
int finish_sensor_load(sgfim_ctx* c, std::vector<DevBuf>& tmp, hipError_t e, int32_t S, int32_t B, int64_t n_win,
                       const int32_t* d_items, const uint32_t* d_present, const int64_t* d_off)
{
    std::vector<uint32_t> present((size_t)S * B);
    std::vector<int32_t> items((size_t)(n_win * S));
    if (e == hipSuccess) e = hipMemcpyAsync(present.data(), d_present, present.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpyAsync(items.data(), d_items, items.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        dev_free_all(c, tmp);
        return fail(c, SGFIM_E_DEVICE, "sensor bins: %s", hipGetErrorString(e));
    }
    const int32_t idmax = S * B;
    std::vector<int32_t> idmap((size_t)idmax + 1, -1);
    c->ids.clear();
    for (int32_t id = 1; id <= idmax; id++)
        if (present[id - 1]) {
            idmap[id] = (int32_t)c->ids.size();
            c->ids.push_back(id);
        }
    const int64_t win_off[2] = {0, n_win};
    const int rc = build_from_device_csr(c, 1, win_off, n_win, d_off, d_items, n_win * S, idmax, idmap);
    (void)hipStreamSynchronize(c->stream);
    dev_free_all(c, tmp);
    if (rc == SGFIM_OK) c->sensor_items.swap(items);
    return rc;
}

}  // namespace

int sgfim_load_sensors_online(sgfim_ctx* c, const double* x, int64_t n_rows, int32_t n_sensors, int32_t n_bins,
                              int32_t window, int32_t stride, double* edges_out)
{
    (void)hipGetLastError();
    if (!c) return SGFIM_E_INVAL;
    pipe_settle(c);
    const int32_t S = n_sensors, B = n_bins;
    if (!x || n_rows < 1 || S < 1 || B < 2 || window < 1 || window > 64 || stride < 1 || n_rows < window)
        return fail(c, SGFIM_E_INVAL, "bad sensor matrix arguments (n_bins >= 2, 1 <= window <= 64, window <= rows)");
    if ((int64_t)S * B > 65535) return fail(c, SGFIM_E_INVAL, "sensors x bins must be <= 65535");
    const int64_t n_win = (n_rows - window) / stride + 1;
    if (n_win >= ((int64_t)1 << 31)) return fail(c, SGFIM_E_INVAL, "too many windows");
    for (int64_t i = 0; i < n_rows * S; i++)
        if (!std::isfinite(x[i])) return fail(c, SGFIM_E_INVAL, "sensor value %lld is not finite", (long long)i);
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    data_free(c);
    const int64_t nv = n_rows * S, ne = n_win * S * (B + 1);
    std::vector<DevBuf> tmp;
    double *d_x = nullptr, *d_E = nullptr;
    int64_t* d_off = nullptr;
    int32_t* d_items = nullptr;
    uint32_t* d_present = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, tmp, &d_x, (size_t)nv)) || (rc = dev_alloc_t(c, tmp, &d_E, (size_t)ne)) ||
        (rc = dev_alloc_t(c, tmp, &d_items, (size_t)(n_win * S))) || (rc = dev_alloc_t(c, tmp, &d_present, (size_t)S * B)) ||
        (rc = dev_alloc_t(c, tmp, &d_off, (size_t)n_win + 1))) {
        dev_free_all(c, tmp);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_x, x, (size_t)nv * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_present, 0, (size_t)S * B * 4, c->stream);
    if (e == hipSuccess) e = sg_online_bins(d_x, S, B, window, stride, n_win, d_E, d_items, d_present, d_off, c->stream);
    if (e == hipSuccess && edges_out)
        e = hipMemcpyAsync(edges_out, d_E, (size_t)ne * 8, hipMemcpyDeviceToHost, c->stream);
    return finish_sensor_load(c, tmp, e, S, B, n_win, d_items, d_present, d_off);
}

int sgfim_set_tile(sgfim_ctx* c, int32_t chunk_words, int32_t chunk_range)
{
    if (!c) return SGFIM_E_INVAL;
    if (!(chunk_words == 0 || chunk_words == 128 || chunk_words == 256 || chunk_words == 512) || chunk_range < 0 ||
        chunk_range > 64)
        return fail(c, SGFIM_E_INVAL, "tile: chunk_words must be 0, 128, 256 or 512 and chunk_range 0..64");
    if (c->S && c->S->active && !c->S->done) return fail(c, SGFIM_E_STATE, "tile: an anytime search is in progress");
    c->tile_cw = chunk_words;
    c->tile_cr = chunk_range;
    return SGFIM_OK;
}

int sgfim_set_sib_ring(sgfim_ctx* c, int64_t words)
{
    if (!c) return SGFIM_E_INVAL;
    if (words < 0) return fail(c, SGFIM_E_INVAL, "sibling ring: words must be >= 0");
    if (c->S && c->S->active && !c->S->done) return fail(c, SGFIM_E_STATE, "sibling ring: an anytime search is in progress");
    c->sib_words = words;
    return SGFIM_OK;
}

int sgfim_set_projection(sgfim_ctx* c, int32_t factor, int64_t arena_words)
{
    if (!c) return SGFIM_E_INVAL;
    if (factor < 0 || factor == 1 || factor > 1024 || arena_words < 0 || (arena_words > 0 && arena_words < 1024))
        return fail(c, SGFIM_E_INVAL, "projection: factor must be 0 or 2..1024, arena_words 0 or >= 1024");
    if (c->S && c->S->active && !c->S->done) return fail(c, SGFIM_E_STATE, "projection: an anytime search is in progress");
    c->proj_factor = factor;
    c->proj_arena = arena_words;
    return SGFIM_OK;
}

int sgfim_get_projection(const sgfim_ctx* c, int32_t* factor, int64_t* arena_words)
{
    if (!c || !factor || !arena_words) return SGFIM_E_INVAL;
    *factor = c->proj_factor;
    *arena_words = c->proj_arena;
    return SGFIM_OK;
}

int sgfim_set_block_rule(sgfim_ctx* c, int32_t dense, int32_t min_children)
{
    if (!c) return SGFIM_E_INVAL;
    if (dense < 0 || dense == 1 || dense > 1024 || min_children < 2 || min_children > SG_DMAX)
        return fail(c, SGFIM_E_INVAL, "block rule: dense must be 0 or 2..1024, min_children 2..%d", SG_DMAX);
    if (c->S && c->S->active && !c->S->done) return fail(c, SGFIM_E_STATE, "block rule: an anytime search is in progress");
    c->proj_dense = dense;
    c->proj_mintail = min_children;
    return SGFIM_OK;
}

int sgfim_get_block_rule(const sgfim_ctx* c, int32_t* dense, int32_t* min_children)
{
    if (!c || !dense || !min_children) return SGFIM_E_INVAL;
    *dense = c->proj_dense;
    *min_children = c->proj_mintail;
    return SGFIM_OK;
}

int sgfim_get_tile(const sgfim_ctx* c, int32_t* chunk_words, int32_t* chunk_range)
{
    if (!c || !chunk_words || !chunk_range) return SGFIM_E_INVAL;
    *chunk_words = search_chunk(c);
    *chunk_range = search_range(c);
    return SGFIM_OK;
}

int sgfim_device_info(const sgfim_ctx* c, char* name, int32_t len, int32_t* n_cu)
{
    if (!c) return SGFIM_E_INVAL;
    if (name && len > 0) {
        const size_t n = std::min<size_t>((size_t)len - 1, c->dev_name.size());
        memcpy(name, c->dev_name.data(), n);
        name[n] = 0;
    }
    if (n_cu) *n_cu = c->n_cu;
    return SGFIM_OK;
}

int sgfim_sensor_items(sgfim_ctx* c, int32_t* items, int64_t cap, int64_t* n)
{
    if (!c || !n) return SGFIM_E_INVAL;
    *n = (int64_t)c->sensor_items.size();
    if (items) std::copy(c->sensor_items.begin(), c->sensor_items.begin() + std::min<int64_t>(cap, *n), items);
    return SGFIM_OK;
}

namespace {
uint64_t zmix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

// Item probabilities q_i = min(1, c i^-s), sum = mean_len (bisection on c), and the
// planted itemsets (sizes, items, rates) of the synthetic Zipf set, from a splitmix64
// stream seeded with seed ^ 0x243F6A8885A308D3.
int sgfim_zipf_spec(int32_t n_items, double zipf_s, double mean_len, int32_t n_planted, int32_t size_lo,
                    int32_t size_hi, double rate_lo, double rate_hi, uint64_t seed, double* q, int32_t* planted_off,
                    int32_t* planted_items, double* planted_rate)
{
    if (n_items < 1 || mean_len <= 0 || mean_len > n_items || n_planted < 0 || size_lo < 1 || size_hi < size_lo ||
        size_hi > n_items || rate_lo < 0 || rate_hi < rate_lo || rate_hi > 1)
        return SGFIM_E_INVAL;
    double lo = 0.0, hi = 1.0;
    auto total = [&](double cc) {
        double t = 0.0;
        for (int32_t i = 1; i <= n_items; i++) t += std::min(1.0, cc * std::pow((double)i, -zipf_s));
        return t;
    };
    while (total(hi) < mean_len) hi *= 2.0;
    for (int it = 0; it < 200; it++) {
        const double m = 0.5 * (lo + hi);
        if (total(m) < mean_len) lo = m; else hi = m;
    }
    for (int32_t i = 1; i <= n_items; i++) q[i - 1] = std::min(1.0, hi * std::pow((double)i, -zipf_s));
    uint64_t st = seed ^ 0x243F6A8885A308D3ull;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        return zmix64(st);
    };
    planted_off[0] = 0;
    for (int32_t j = 0; j < n_planted; j++) {
        const int32_t sz = size_lo + (int32_t)(next() % (uint64_t)(size_hi - size_lo + 1));
        int32_t k = planted_off[j];
        while (k - planted_off[j] < sz) {
            const int32_t x = 1 + (int32_t)(next() % (uint64_t)n_items);
            bool dup = false;
            for (int32_t a = planted_off[j]; a < k; a++) dup = dup || planted_items[a] == x;
            if (!dup) planted_items[k++] = x;
        }
        planted_off[j + 1] = k;
        planted_rate[j] = rate_lo + (rate_hi - rate_lo) * ((double)(next() >> 11) * (1.0 / 9007199254740992.0));
    }
    return SGFIM_OK;
}

// Generate the synthetic Zipf set straight into the bitmaps (ids 1..n_items, all kept).
int sgfim_load_zipf(sgfim_ctx* c, int64_t n_txn, int32_t n_items, double zipf_s, double mean_len, int32_t n_planted,
                    int32_t size_lo, int32_t size_hi, double rate_lo, double rate_hi, uint64_t seed)
{
    (void)hipGetLastError();
    if (!c) return SGFIM_E_INVAL;
    pipe_settle(c);
    if (n_txn < 1 || n_txn >= ((int64_t)1 << 31) || n_items < 1 || n_items > 65535 || n_planted > 4096)
        return fail(c, SGFIM_E_INVAL, "bad synthetic set size");
    std::vector<double> q((size_t)n_items), rate((size_t)std::max(n_planted, 1));
    std::vector<int32_t> poff((size_t)n_planted + 1), pit((size_t)std::max(n_planted * size_hi, 1));
    if (sgfim_zipf_spec(n_items, zipf_s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed, q.data(),
                        poff.data(), pit.data(), rate.data()) != SGFIM_OK)
        return fail(c, SGFIM_E_INVAL, "bad synthetic set parameters");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    data_free(c);
    const double two53 = 9007199254740992.0;
    std::vector<uint64_t> q53((size_t)n_items), r53((size_t)std::max(n_planted, 1));
    for (int32_t i = 0; i < n_items; i++) q53[i] = (uint64_t)(q[i] * two53);
    for (int32_t j = 0; j < n_planted; j++) r53[j] = (uint64_t)(rate[j] * two53);
    // per item: the planted sets holding it
    std::vector<std::vector<int32_t>> by_item((size_t)n_items);
    for (int32_t j = 0; j < n_planted; j++)
        for (int32_t a = poff[j]; a < poff[j + 1]; a++) by_item[pit[a] - 1].push_back(j);
    std::vector<int32_t> isoff((size_t)n_items + 1, 0), isl;
    for (int32_t i = 0; i < n_items; i++) {
        isl.insert(isl.end(), by_item[i].begin(), by_item[i].end());
        isoff[i + 1] = (int32_t)isl.size();
    }
    if (isl.empty()) isl.push_back(0);
    c->n_txn = n_txn;
    c->d = n_items;
    c->ids.resize(n_items);
    for (int32_t i = 0; i < n_items; i++) c->ids[i] = i + 1;
    c->dpad = std::max(128, (c->d + 127) / 128 * 128);
    c->nseg = 1;
    c->segs.assign(1, SgSegment{});
    SgSegment& g = c->segs[0];
    g.n = (int32_t)n_txn;
    g.words = (n_txn + 63) / 64;
    g.stride = std::max<int64_t>(64, (g.words + 63) / 64 * 64);
    g.d = c->d;
    g.dpad = c->dpad;
    g.dw = (c->d + 63) / 64;
    g.thr = 1;
    c->W = g.words;
    c->Wp = g.stride;
    c->rows_words = (int64_t)c->dpad * g.stride;
    int rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_rows, (size_t)c->rows_words))) return rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_sup, (size_t)c->dpad))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_rows, 0, (size_t)c->rows_words * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_sup, 0, (size_t)c->dpad * 4, c->stream));
    std::vector<DevBuf> tmp;
    uint64_t *d_q = nullptr, *d_r = nullptr;
    int32_t *d_isoff = nullptr, *d_isl = nullptr;
    if ((rc = dev_alloc_t(c, tmp, &d_q, q53.size())) || (rc = dev_alloc_t(c, tmp, &d_r, r53.size())) ||
        (rc = dev_alloc_t(c, tmp, &d_isoff, isoff.size())) || (rc = dev_alloc_t(c, tmp, &d_isl, isl.size()))) {
        dev_free_all(c, tmp);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_q, q53.data(), q53.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_r, r53.data(), r53.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_isoff, isoff.data(), isoff.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_isl, isl.data(), isl.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = sg_launch_gen_zipf(c->d_rows, g.stride, n_txn, c->d, d_q, d_r, d_isoff, d_isl, seed, c->d_sup, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free_all(c, tmp);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "synthetic set generation failed: %s", hipGetErrorString(e));
    std::vector<int32_t> sup32((size_t)c->dpad);
    HIPCHK(c, hipMemcpyAsync(sup32.data(), c->d_sup, sup32.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_sup_seg.assign(sup32.begin(), sup32.begin() + c->d);
    c->h_sup = c->h_sup_seg;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_seg, 2))) return rc;
    if ((rc = dev_alloc_t(c, c->data_bufs, &c->d_seg_item_base, 1))) return rc;
    const int64_t zero = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_seg_item_base, &zero, 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_seg, c->segs.data(), sizeof(SgSegment), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = alloc_vsegs(c))) return rc;
    c->loaded = true;
    return SGFIM_OK;
}

int sgfim_load_csr(sgfim_ctx* c, const int64_t* off, int64_t n_txn, const int32_t* items, int64_t nnz)
{
    const int64_t w[2] = {0, n_txn};
    return sgfim_load_windows(c, 1, w, off, n_txn, items, nnz);
}

int sgfim_window_supports(sgfim_ctx* c, int32_t window, int64_t* out)
{
    if (!c || !out) return SGFIM_E_INVAL;
    if (!c->loaded) return fail(c, SGFIM_E_STATE, "no transactions loaded");
    if (window < 0 || window >= c->nseg) return fail(c, SGFIM_E_INVAL, "window %d out of range", window);
    std::copy(c->h_sup_seg.begin() + (size_t)window * c->d, c->h_sup_seg.begin() + (size_t)(window + 1) * c->d, out);
    return SGFIM_OK;
}

int sgfim_num_windows(const sgfim_ctx* c, int32_t* n)
{
    if (!c || !n) return SGFIM_E_INVAL;
    *n = c->nseg;
    return SGFIM_OK;
}

int sgfim_info(const sgfim_ctx* c, int64_t* n_txn, int32_t* n_items, int64_t* words)
{
    if (!c) return SGFIM_E_INVAL;
    if (n_txn) *n_txn = c->n_txn;
    if (n_items) *n_items = c->d;
    if (words) *words = c->W;
    return SGFIM_OK;
}

int sgfim_item_ids(const sgfim_ctx* c, int32_t* ids)
{
    if (!c || !ids) return SGFIM_E_INVAL;
    std::copy(c->ids.begin(), c->ids.end(), ids);
    return SGFIM_OK;
}

int sgfim_item_supports(sgfim_ctx* c, int64_t* out)
{
    if (!c || !out) return SGFIM_E_INVAL;
    if (!c->loaded) return fail(c, SGFIM_E_STATE, "no transactions loaded");
    std::copy(c->h_sup.begin(), c->h_sup.end(), out);
    return SGFIM_OK;
}

static int pair_out(sgfim_ctx* c, int32_t* out, bool valu)
{
    (void)hipGetLastError();
    if (!c || !out) return SGFIM_E_INVAL;
    pipe_settle(c);
    if (!c->loaded) return fail(c, SGFIM_E_STATE, "no transactions loaded");
    if (c->nseg != 1) return fail(c, SGFIM_E_STATE, "pair_matrix needs a single loaded transaction set");
    (void)hipSetDevice(c->device);
    int rc;
    if (valu) {
        std::vector<DevBuf> tmp;
        int32_t* d = nullptr;
        if ((rc = dev_alloc_t(c, tmp, &d, (size_t)c->dpad * c->dpad))) return rc;
        hipError_t e = sg_launch_pair_valu(c->d_rows, c->dpad, c->Wp, c->W, d, c->stream);
        std::vector<int32_t> full((size_t)c->dpad * c->dpad);
        if (e == hipSuccess) e = hipMemcpyAsync(full.data(), d, full.size() * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dev_free_all(c, tmp);
        if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "VALU pair matrix failed: %s", hipGetErrorString(e));
        for (int i = 0; i < c->d; i++) memcpy(out + (size_t)i * c->d, &full[(size_t)i * c->dpad], (size_t)c->d * 4);
        return SGFIM_OK;
    }
    c->pair_valid = false;  // recompute: this entry point measures the MFMA path
    if ((rc = ensure_pair(c))) return rc;
    const int64_t n = (int64_t)c->dpad * c->dpad;
    if (c->h_pair_pin_n < n) {
        if (c->h_pair_pin) (void)hipHostFree(c->h_pair_pin);
        c->h_pair_pin = nullptr;
        c->h_pair_pin_n = 0;
        if (hipHostMalloc((void**)&c->h_pair_pin, (size_t)n * 4, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            c->h_pair_pin = nullptr;
            return fail(c, SGFIM_E_NOMEM, "pinned pair-matrix buffer allocation failed");
        }
        c->h_pair_pin_n = n;
    }
    HIPCHK(c, hipMemcpyAsync(c->h_pair_pin, c->d_pair, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->d; i++) memcpy(out + (size_t)i * c->d, c->h_pair_pin + (size_t)i * c->dpad, (size_t)c->d * 4);
    return SGFIM_OK;
}

int sgfim_pair_matrix(sgfim_ctx* c, int32_t* out) { return pair_out(c, out, false); }

int sgfim_pair_time(sgfim_ctx* c, double* ms)
{
    if (!c || !ms) return SGFIM_E_INVAL;
    *ms = 0.0;
    if (!c->pair_timed) return SGFIM_OK;
    float f = 0.0f;
    hipError_t e = hipEventSynchronize(c->pair_ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&f, c->pair_ev[0], c->pair_ev[1]);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "pair timing: %s", hipGetErrorString(e));
    *ms = (double)f;
    return SGFIM_OK;
}
int sgfim_pair_matrix_valu(sgfim_ctx* c, int32_t* out) { return pair_out(c, out, true); }

static_assert(SG_PAIR_FP4 == SGFIM_PAIR_FP4 && SG_PAIR_INT8 == SGFIM_PAIR_INT8, "pair kernel ids");

int sgfim_set_pair_kernel(sgfim_ctx* c, int32_t kind, int32_t max_splits)
{
    if (!c) return SGFIM_E_INVAL;
    if (kind != SGFIM_PAIR_FP4 && kind != SGFIM_PAIR_INT8)
        return fail(c, SGFIM_E_INVAL, "pair kernel %d: expected SGFIM_PAIR_FP4 or SGFIM_PAIR_INT8", kind);
    if (max_splits < 0) return fail(c, SGFIM_E_INVAL, "pair kernel: max_splits %d < 0", max_splits);
    if (kind != c->pair_kind || max_splits != c->pair_split_cap) {
        c->pair_kind = kind;
        c->pair_split_cap = max_splits;
        c->pair_valid = false;
        c->cu.pair_valid = false;
    }
    return SGFIM_OK;
}

int sgfim_set_count_kernel(sgfim_ctx* c, int32_t kind)
{
    static_assert(SG_COUNT_VALU == SGFIM_COUNT_VALU && SG_COUNT_F4 == SGFIM_COUNT_F4, "kernel ids");
    if (!c) return SGFIM_E_INVAL;
    if (kind != SGFIM_COUNT_F4 && kind != SGFIM_COUNT_VALU)
        return fail(c, SGFIM_E_INVAL, "count kernel %d: expected SGFIM_COUNT_F4 or SGFIM_COUNT_VALU", kind);
    if (c->S && c->S->active && !c->S->done) return fail(c, SGFIM_E_STATE, "count kernel: an anytime search is in progress");
    c->count_kind = kind;
    return SGFIM_OK;
}

// The counting behind sgfim_count_batch and sgfim_count_batch_noalloc: the device scratch
// (node descriptors, compacted candidates, counts, work counter) is allocated per call, or
// taken from the device's static scratch (noalloc: no allocation at all, sizes bounded).
static int count_batch_impl(sgfim_ctx* c, const int32_t* prefix_ids, const int64_t* prefix_off, int32_t n_prefix,
                            const int32_t* cand, const int64_t* cand_off, int32_t* out, bool noalloc)
{
    (void)hipGetLastError();
    if (!c || n_prefix < 0 || (n_prefix > 0 && (!prefix_off || !cand_off))) return SGFIM_E_INVAL;
    if (!c->loaded) return fail(c, SGFIM_E_STATE, "no transactions loaded");
    if (c->nseg != 1) return fail(c, SGFIM_E_STATE, "count_batch needs a single loaded transaction set");
    if (n_prefix == 0) return SGFIM_OK;
    if (prefix_off[0] != 0 || cand_off[0] != 0) return fail(c, SGFIM_E_INVAL, "offsets must start at 0");
    if (noalloc && n_prefix > SG_NOALLOC_NODES)
        return fail(c, SGFIM_E_INVAL, "count_batch_noalloc: %d prefixes, at most %d", n_prefix, SG_NOALLOC_NODES);
    // original id -> dense id
    auto dense = [&](int32_t id) -> int32_t {
        auto it = std::lower_bound(c->ids.begin(), c->ids.end(), id);
        return (it != c->ids.end() && *it == id) ? (int32_t)(it - c->ids.begin()) : -1;
    };
    const int32_t cw = pick_chunk(c, c->W);
    const int64_t nchunk = std::max<int64_t>(1, (c->W + cw - 1) / cw);
    std::vector<SgCountNode> nodes((size_t)n_prefix);
    const int64_t ncand_total = cand_off[n_prefix];
    // candidates that occur in the set, compacted; src[j] = output position of compact j
    std::vector<int32_t> dc;
    std::vector<int64_t> src;
    dc.reserve((size_t)ncand_total);
    src.reserve((size_t)ncand_total);
    int64_t units = 0;
    for (int32_t p = 0; p < n_prefix; p++) {
        const int64_t a = prefix_off[p], b = prefix_off[p + 1];
        const int64_t ca = cand_off[p], cb = cand_off[p + 1];
        if (b < a || cb < ca) return fail(c, SGFIM_E_INVAL, "offsets not non-decreasing at prefix %d", p);
        if (b - a > SG_MAX_DEPTH) return fail(c, SGFIM_E_DEPTH, "prefix %d has %lld > %d items", p, (long long)(b - a), SG_MAX_DEPTH);
        SgCountNode& nd = nodes[p];
        memset(&nd, 0, sizeof(nd));
        nd.k = (int32_t)(b - a);
        bool absent = false;  // an absent prefix item makes every count of the prefix 0
        for (int64_t i = a; i < b; i++) {
            const int32_t x = dense(prefix_ids[i]);
            if (x < 0) absent = true;
            nd.items[i - a] = x < 0 ? 0 : x;
        }
        nd.cand_off = (int64_t)dc.size();
        if (!absent && c->W > 0)
            for (int64_t j = ca; j < cb; j++) {
                const int32_t x = dense(cand[j]);
                if (x < 0) continue;  // an absent candidate counts 0
                dc.push_back(x);
                src.push_back(j);
            }
        nd.ncand = (int32_t)((int64_t)dc.size() - nd.cand_off);
        nd.seg = 0;
        nd.nchunk = nd.ncand > 0 ? (int32_t)nchunk : 0;
        nd.unit_off = units;
        units += nd.nchunk;
    }
    if (dc.empty()) {
        std::fill(out, out + ncand_total, 0);
        return SGFIM_OK;
    }
    if (noalloc && (int64_t)dc.size() > SG_NOALLOC_CANDS)
        return fail(c, SGFIM_E_INVAL, "count_batch_noalloc: %lld candidates present in the set, at most %d",
                    (long long)dc.size(), SG_NOALLOC_CANDS);
    (void)hipSetDevice(c->device);
    SgSegment seg{};
    seg.row_base = 0;
    seg.stride = c->Wp;
    seg.words = c->W;
    seg.d = c->d;
    seg.dpad = c->dpad;
    seg.dw = (c->d + 63) / 64;
    seg.n = (int32_t)c->n_txn;
    HIPCHK(c, hipMemcpyAsync(c->d_seg + c->nseg, &seg, sizeof(seg), hipMemcpyHostToDevice, c->stream));
    std::vector<DevBuf> tmp;
    SgCountNode* d_nodes = nullptr;
    int32_t *d_cand = nullptr, *d_out = nullptr, *d_ctr = nullptr;
    int rc;
    if (noalloc) {   // the device's static scratch: [nodes | candidate ids | counts | counter]
        unsigned char* b = (unsigned char*)c->scratch;
        d_nodes = (SgCountNode*)b;
        d_cand = (int32_t*)(b + SG_NOALLOC_NODE_BYTES);
        d_out = d_cand + SG_NOALLOC_CANDS;
        d_ctr = d_out + SG_NOALLOC_CANDS;
    } else if ((rc = dev_alloc_t(c, tmp, &d_nodes, nodes.size())) || (rc = dev_alloc_t(c, tmp, &d_cand, dc.size())) ||
               (rc = dev_alloc_t(c, tmp, &d_out, dc.size())) || (rc = dev_alloc_t(c, tmp, &d_ctr, 1))) {
        dev_free_all(c, tmp);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_nodes, nodes.data(), nodes.size() * sizeof(SgCountNode), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_cand, dc.data(), dc.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_out, 0, dc.size() * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_ctr, 0, 4, c->stream);
    if (e == hipSuccess)
        e = sg_launch_count_nodes(c->d_rows, c->d_seg + c->nseg, d_nodes, n_prefix, units, d_cand, d_out, cw, d_ctr, 2048, c->stream);
    std::vector<int32_t> res(dc.size());
    if (e == hipSuccess) e = hipMemcpyAsync(res.data(), d_out, res.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free_all(c, tmp);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "count batch failed: %s", hipGetErrorString(e));
    std::fill(out, out + ncand_total, 0);
    for (size_t j = 0; j < src.size(); j++) out[src[j]] = res[j];
    return SGFIM_OK;
}

int sgfim_count_batch(sgfim_ctx* c, const int32_t* prefix_ids, const int64_t* prefix_off, int32_t n_prefix,
                      const int32_t* cand, const int64_t* cand_off, int32_t* out)
{
    return count_batch_impl(c, prefix_ids, prefix_off, n_prefix, cand, cand_off, out, false);
}

int sgfim_count_batch_noalloc(sgfim_ctx* c, const int32_t* prefix_ids, const int64_t* prefix_off, int32_t n_prefix,
                              const int32_t* cand, const int64_t* cand_off, int32_t* out)
{
    return count_batch_impl(c, prefix_ids, prefix_off, n_prefix, cand, cand_off, out, true);
}

// ------------------------------------------------------------------ search
static int search_setup(sgfim_ctx* c, const sgfim_mine_args* a)
{
    (void)hipGetLastError();
    if (!c || !a) return SGFIM_E_INVAL;
    if (!c->loaded) return fail(c, SGFIM_E_STATE, "no transactions loaded");
    if (a->min_count < 1) return fail(c, SGFIM_E_INVAL, "min_count must be >= 1");
    if ((a->flags & SGFIM_ORDER_LEARNED) && !c->scorer_set)
        return fail(c, SGFIM_E_STATE, "SGFIM_ORDER_LEARNED needs sgfim_set_scorer first");
    if ((a->flags & SGFIM_ROOT_STRICT) && (a->flags & SGFIM_ROOT_TOPK))
        return fail(c, SGFIM_E_INVAL, "SGFIM_ROOT_STRICT and SGFIM_ROOT_TOPK are exclusive");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    search_free(c);
    SearchState* S = new SearchState();
    c->S = S;
    S->B = c->SB;
    S->args = *a;
    S->order_mode = (a->flags & SGFIM_ORDER_LEARNED) ? 1 : 0;
    S->use_graph = !(a->flags & SGFIM_NO_GRAPH);
    S->t0 = std::chrono::steady_clock::now();
    const int32_t nseg = c->nseg;
    int rc;
    // search segments: one per window (its own threshold), or one per threshold of a
    // single set (one batched search over several thresholds; roots and results per
    // threshold, rows and pair matrix shared)
    const bool multi = !c->multi_thr.empty();
    if (multi && nseg != 1)
        return fail(c, SGFIM_E_INVAL, "a threshold list needs a single loaded set (windows take per-window thresholds)");
    const int32_t nv = multi ? (int32_t)c->multi_thr.size() : nseg;
    if (nv > c->vseg_cap) return fail(c, SGFIM_E_INVAL, "at most %d thresholds per search", c->vseg_cap);
    std::vector<int64_t> thr(nv, a->min_count);
    if (multi) thr = c->multi_thr;
    else if (!c->win_thr.empty()) {
        if ((int32_t)c->win_thr.size() != nseg) return fail(c, SGFIM_E_INVAL, "window threshold count mismatch");
        thr = c->win_thr;
    }
    for (int64_t t : thr)
        if (t < 1 || t > (int64_t)INT32_MAX) return fail(c, SGFIM_E_INVAL, "min_count out of range");
    auto data_seg = [&](int32_t v) { return multi ? 0 : v; };
    // the universe: the loaded items, or -- past SG_DMAX of them -- a compact copy of the
    // items frequent at some search segment's threshold (no other item can be in an MFI,
    // a frontier record or a candidate list)
    if (c->d <= SG_DMAX) {
        S->U = loaded_universe(c);
    } else {
        std::vector<int32_t> keep;
        for (int32_t x = 0; x < c->d; x++)
            for (int32_t v = 0; v < nv; v++)
                if (c->h_sup_seg[(size_t)data_seg(v) * c->d + x] >= thr[v]) {
                    keep.push_back(x);
                    break;
                }
        if ((int64_t)keep.size() > SG_DMAX)
            return fail(c, SGFIM_E_INVAL, "mining supports at most %d items frequent at the search's thresholds (got %lld of %d)",
                        SG_DMAX, (long long)keep.size(), c->d);
        if (keep.empty()) {
            S->B = nullptr;   // nothing frequent: no search buffers are touched
            S->active = true;
            S->done = true;
            fill_stats(c, S);
            return SGFIM_OK;
        }
        if ((rc = compact_build(c, keep))) return rc;
        S->U = compact_universe(c);
    }
    Universe& U = S->U;
    const int32_t d = U.d;
    const int32_t dw = (d + 63) / 64;
    // roots (search segment, dense id): frequent items passing the root rule
    std::vector<int32_t> rp;
    for (int32_t v = 0; v < nv; v++) {
        const int32_t w = data_seg(v);
        const int64_t* sup = U.h_sup_seg + (size_t)w * d;
        const int64_t n = U.segs[w].n;
        if (a->flags & SGFIM_ROOT_TOPK) {
            const int32_t topk = (multi && !c->multi_topk.empty()) ? c->multi_topk[v] : a->root_topk;
            std::vector<int32_t> order(d);
            for (int32_t i = 0; i < d; i++) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return sup[x] > sup[y]; });
            std::vector<int32_t> r;
            for (int32_t i = 0; i < std::min<int32_t>(topk, d); i++)
                if (sup[order[i]] >= thr[v]) r.push_back(order[i]);
            std::sort(r.begin(), r.end());
            for (int32_t x : r) { rp.push_back(v); rp.push_back(x); }
            continue;
        }
        for (int32_t x = 0; x < d; x++) {
            if (sup[x] < thr[v]) continue;
            if ((a->flags & SGFIM_ROOT_STRICT) && !(((double)sup[x] / (double)n) > a->rel_thr)) continue;
            rp.push_back(v);
            rp.push_back(x);
        }
    }
    const int64_t n_roots = (int64_t)rp.size() / 2;
    if (d == 0 || n_roots == 0 || c->n_txn == 0) {
        S->B = nullptr;  // nothing frequent: no search buffers are touched
        S->active = true;
        S->done = true;
        fill_stats(c, S);
        return SGFIM_OK;
    }
    if (a->flags & SGFIM_FRESH_PAIR) *U.pair_valid = false;
    if (U.gen == c->loaded_gen) {   // the loaded items' pair matrix and pair-ok masks are allocated on first use
        if ((rc = ensure_pair(c))) return rc;
        if (!c->d_ok && (rc = dev_alloc_t(c, c->data_bufs, &c->d_ok, (size_t)d * dw * c->vseg_cap))) return rc;
        U.pair = c->d_pair;
        U.ok = c->d_ok;
    } else if ((rc = ensure_pair_u(c, U))) {
        return rc;
    }
    S->pair_counts = n_roots * d;
    // search-segment table (thresholds, pair-ok blocks) + pair_ok masks
    std::vector<SgSegment>& vs = S->vsegs;
    std::vector<int64_t>& vib = S->vibase;
    vs.resize(nv);
    vib.resize(nv);
    for (int32_t v = 0; v < nv; v++) {
        vs[v] = U.segs[data_seg(v)];
        vs[v].thr = (int32_t)thr[v];
        vs[v].ok_base = (int64_t)v * d * dw;
        vib[v] = (int64_t)data_seg(v) * U.dpad;
    }
    // uploads go through pinned staging: [segments | item bases | roots], 256-byte aligned
    const size_t st_seg = 0, st_ib = (st_seg + (size_t)nv * sizeof(SgSegment) + 255) / 256 * 256;
    const size_t st_roots = (st_ib + (size_t)nv * 8 + 255) / 256 * 256;
    if ((rc = stage_reserve(c, st_roots + rp.size() * 4))) return rc;
    memcpy(c->h_stage + st_seg, vs.data(), (size_t)nv * sizeof(SgSegment));
    memcpy(c->h_stage + st_ib, vib.data(), (size_t)nv * 8);
    HIPCHK(c, hipMemcpyAsync(c->d_vseg, c->h_stage + st_seg, (size_t)nv * sizeof(SgSegment), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_vseg_ibase, c->h_stage + st_ib, (size_t)nv * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, sg_launch_pair_ok(U.pair, c->d_vseg, nv, d, U.ok, c->stream));
    // the seed turns every root into at most one frontier record
    const int64_t seed_records = n_roots;
    // search buffers (reused across runs when they fit)
    const int32_t bmax = std::min(SG_BMAX, c->opts.batch_nodes > 0 ? c->opts.batch_nodes : SG_BMAX);
    int64_t cap = c->opts.frontier_cap > 0 ? c->opts.frontier_cap
                                           : std::max<int64_t>({(int64_t)1 << 20, 4 * seed_records + 4 * bmax + 4 * (int64_t)d,
                                                                c->cap_hint});
    if (cap > ((int64_t)1 << 31) - 1) cap = ((int64_t)1 << 31) - 1;
    if (seed_records > cap) return fail(c, SGFIM_E_CAPACITY, "frontier_cap %lld < %lld root records",
                                        (long long)cap, (long long)seed_records);
    // newborn_cap bounds the children of one round (each makes at most one record and emits
    // at most one MFI), so the MFI ring holds the seed's MFIs plus one round's
    int64_t newborn_cap = std::max<int64_t>(seed_records, std::min<int64_t>(cap, (int64_t)1 << 19));
    // count cells of one round; at least one record's full tile (d x d) so every round advances
    const int64_t cell_cap = std::max<int64_t>((int64_t)1 << 26, (int64_t)d * d);
    // two rounds may be in flight before the host drains the ring
    int64_t ring_cap = n_roots + 2 * newborn_cap;
    if (c->opts.ring_cap > 0) {
        const int64_t need = n_roots + 2 * std::max<int64_t>(n_roots, d);
        if (c->opts.ring_cap < need)
            return fail(c, SGFIM_E_CAPACITY, "ring_cap must hold at least %lld MFIs", (long long)need);
        ring_cap = c->opts.ring_cap;
        newborn_cap = std::min(newborn_cap, (ring_cap - n_roots) / 2);
    }
    // projected blocks: arena ring (default 32x the loaded rows, at least 2^28 words, at most a
    // quarter of the device memory), descriptors for every frontier record up to 2^20,
    // transaction lists of one round (at least 2^27 entries, and two blocks projected from
    // the longest window)
    ProjSizes ps;
    if (c->proj_factor > 0) {
        ps.factor = c->proj_factor;
        ps.dense = c->proj_dense;
        ps.mintail = c->proj_mintail;
        int64_t a = c->proj_arena;
        if (a <= 0) {
            a = std::max<int64_t>((int64_t)1 << 28, 32 * c->rows_words);
            size_t fr = 0, tt = 0;   // the total (not the free) memory keeps the size stable across runs
            if (hipMemGetInfo(&fr, &tt) == hipSuccess)
                a = std::max<int64_t>((int64_t)1 << 22, std::min<int64_t>(a, (int64_t)(tt / 4 / 8 * 15 / 16)));
            else
                (void)hipGetLastError();
        }
        ps.arena = a & ~(int64_t)7;   // blocks start at multiples of 8 words (16-byte row loads)
        ps.pdesc = std::min<int64_t>(std::max<int64_t>(cap, (int64_t)1 << 16), (int64_t)1 << 20);
        int64_t nmax = 0;
        for (int32_t v = 0; v < nv; v++) nmax = std::max<int64_t>(nmax, vs[v].n);
        ps.ptid = std::max<int64_t>((int64_t)1 << 27, 2 * nmax / ps.factor + 1);
        ps.pent = ps.ptid / 4;   // 16-byte word entries of compacted blocks
    }
    if ((rc = ensure_bufs(c, cap, newborn_cap, ring_cap, d, dw, bmax, search_chunk(c), search_range(c),
                          c->trace ? c->trace_cap : 0, n_roots, cell_cap, ps)))
        return rc;
    S->B = c->SB;
    SgSearchDev& D = S->B->D;
    D.feat_on = (S->order_mode == 1 || D.trace_cap > 0) ? 1 : 0;   // a round captured with another value is captured again
    if (D.count_kind != c->count_kind) {   // likewise
        D.count_kind = c->count_kind;
        D.count_grid = sg_count_group_blocks_per_cu(D.chunk_words, D.count_kind) * c->n_cu;
    }
    HIPCHK(c, sg_search_init(D, c->stream));
    S->n_roots = (int32_t)n_roots;
    S->roots = std::move(rp);
    memcpy(c->h_stage + st_roots, S->roots.data(), S->roots.size() * 4);
    HIPCHK(c, hipMemcpyAsync(S->B->d_roots, c->h_stage + st_roots, S->roots.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, sg_search_seed(D, c->d_vseg, U.pair, U.ok, S->B->d_roots, S->n_roots, U.sup, c->d_vseg_ibase,
                             S->order_mode, &c->scorer, c->stream));
    // no synchronisation here: the seed's MFIs, errors and frontier are read after the
    // first round (a round over an empty frontier does nothing)
    memset(S->B->h_ctl, 0, CTL_SLOTS * sizeof(int64_t));
    S->active = true;
    S->done = false;
    fill_stats(c, S);
    return SGFIM_OK;
}

int sgfim_anytime_begin(sgfim_ctx* c, const sgfim_mine_args* a)
{
    int rc = search_setup(c, a);
    if (rc && c && c->S) {
        c->S->status = rc;
        c->S->done = true;
    }
    return rc;
}

// Queue one round and the copy of its control snapshot behind it.
static int pipe_push(sgfim_ctx* c, SearchState* S)
{
    SearchBufs* B = S->B;
    const int slot = (S->pipe_head + S->pipe_n) & 1;
    int64_t* snap = B->h_ctl + (1 + slot) * CTL_SLOTS;
    int rc;
    if ((rc = launch_round(c, S))) return rc;
    hipError_t e = hipMemcpyAsync(snap, B->D.ctl, CTL_SLOTS * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipEventRecord(B->snap_ev[slot], c->stream);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "round snapshot: %s", hipGetErrorString(e));
    S->pipe_n++;
    return SGFIM_OK;
}

// Wait for the oldest queued round and take in its snapshot; with `drain`, check its
// errors and move its MFIs to the pending result. Sets S->done when the frontier is empty.
static int pipe_pop(sgfim_ctx* c, SearchState* S, bool drain)
{
    SearchBufs* B = S->B;
    const int slot = S->pipe_head;
    S->pipe_head ^= 1;
    S->pipe_n--;
    hipError_t e = hipEventSynchronize(B->snap_ev[slot]);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "round: %s", hipGetErrorString(e));
    memcpy(B->h_ctl, B->h_ctl + (1 + slot) * CTL_SLOTS, CTL_SLOTS * sizeof(int64_t));
    if (B->h_ctl[CTL_BATCH_N] > 0) S->rounds++;   // a round queued after the frontier emptied is not counted
    // every queued round launches the counting kernel once (an empty one returns at once)
    if (S->pipe_counted > 0) S->pipe_counted--;
    else S->count_launches++;
    if (B->h_ctl[CTL_UNITS] > 0 && B->h_ctl[CTL_BATCH_N] > 0) {
        // device-side duration of the counting kernel: first workgroup start to last
        // workgroup end, from the constant-rate wall clock
        const int64_t t0 = B->h_ctl[CTL_T0], t1 = B->h_ctl[CTL_T1];
        if (t1 > t0 && c->wall_khz > 0) S->count_ms += (double)(t1 - t0) / (double)c->wall_khz;
        if (!S->use_graph) {   // one round in flight: its stage events are this round's
            for (int i = 0; i < SG_STAGES; i++) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, B->ev[i], B->ev[i + 1]) == hipSuccess) S->stage_ms[i] += ms;
                else (void)hipGetLastError();
                if (i == 5) S->event_ms += ms;
            }
        }
    }
    if (!drain) return SGFIM_OK;
    int rc;
    if ((rc = check_err(c, S)) || (rc = drain_ring(c, S, B->h_ctl[CTL_RING_POS], now_s(S)))) return rc;
    if (frontier_size(B->h_ctl) == 0) S->done = true;
    return SGFIM_OK;
}

// Process every round still queued (a call that changes what the rounds read, or needs
// the stream, first lets them finish); their MFIs go to the pending result.
static void pipe_settle(sgfim_ctx* c)
{
    SearchState* S = c ? c->S : nullptr;
    if (!S) return;
    while (S->pipe_n > 0) {
        const int rc = pipe_pop(c, S, S->status == SGFIM_OK);
        if (rc != SGFIM_OK && S->status == SGFIM_OK) {
            S->status = rc;
            S->done = true;
        }
    }
    if (S->active) fill_stats(c, S);
}

int sgfim_anytime_step(sgfim_ctx* c, int32_t max_rounds, sgfim_round_stats* st)
{
    if (!c) return SGFIM_E_INVAL;
    SearchState* S = c->S;
    if (!S || !S->active) return fail(c, SGFIM_E_STATE, "no anytime search in progress");
    (void)hipSetDevice(c->device);
    int rc = S->status;
    if (rc != SGFIM_OK || (S->done && S->pipe_n == 0)) {
        fill_stats(c, S);
        if (st) *st = S->stats;
        return rc;
    }
    // Rounds are pipelined two deep with captured graphs: round r+1 is queued before the
    // host reads round r's control snapshot and drains its MFIs, so the host work and the
    // launch latency hide behind the device's work. The queue stays full when the call
    // returns after max_rounds rounds, so the device keeps working while the caller
    // handles the MFIs. A round queued after the frontier emptied does nothing. Direct
    // launches (event timing) keep one round queued.
    const int depth = S->use_graph ? 2 : 1;
    int64_t processed = 0;
    bool stop = S->done;
    while (true) {
        while (!stop && S->pipe_n < depth) {
            if (S->cancel) { rc = fail(c, SGFIM_E_CANCELLED, "search cancelled"); stop = true; break; }
            if (S->args.deadline_s > 0 && now_s(S) > S->args.deadline_s) {
                rc = fail(c, SGFIM_E_TIMEOUT, "deadline of %.3f s passed", S->args.deadline_s);
                stop = true;
                break;
            }
            if ((rc = pipe_push(c, S))) { stop = true; break; }
        }
        if (S->pipe_n == 0) break;
        if (!stop && max_rounds > 0 && processed >= max_rounds) break;   // queued rounds carry over
        // the search ended: a round still queued runs after the one whose snapshot showed an
        // empty frontier, so it takes no record, counts nothing and finds no MFI. It is not
        // waited for here (the next call that needs the stream processes it), and its
        // counting-kernel launch is counted now.
        if (S->done && rc == SGFIM_OK) {
            S->count_launches += S->pipe_n - S->pipe_counted;
            S->pipe_counted = S->pipe_n;
            break;
        }
        // process the oldest round queued (after an earlier failure, only wait for it)
        const int r2 = pipe_pop(c, S, rc == SGFIM_OK);
        processed++;
        if (r2 != SGFIM_OK) {
            if (rc == SGFIM_OK) rc = r2;
            stop = true;
        }
        if (S->done) stop = true;
    }
    if (rc != SGFIM_OK) {
        S->status = rc;
        S->done = true;
    }
    fill_stats(c, S);
    if (st) *st = S->stats;
    return rc;
}

int sgfim_anytime_poll(sgfim_ctx* c, sgfim_round_stats* st)
{
    if (!c || !st) return SGFIM_E_INVAL;
    if (!c->S) return fail(c, SGFIM_E_STATE, "no anytime search");
    *st = c->S->stats;
    return SGFIM_OK;
}

int sgfim_anytime_cancel(sgfim_ctx* c)
{
    if (!c) return SGFIM_E_INVAL;
    if (!c->S) return fail(c, SGFIM_E_STATE, "no anytime search");
    c->S->cancel = true;
    return SGFIM_OK;
}

int sgfim_anytime_take(sgfim_ctx* c, sgfim_result** out)
{
    if (!c || !out) return SGFIM_E_INVAL;
    if (!c->S) return fail(c, SGFIM_E_STATE, "no anytime search");
    sgfim_result* r = new sgfim_result();
    std::swap(*r, c->S->pending);
    c->S->pending = sgfim_result();
    *out = r;
    return SGFIM_OK;
}

int sgfim_mine(sgfim_ctx* c, const sgfim_mine_args* a, sgfim_result** out)
{
    if (!c || !a || !out) return SGFIM_E_INVAL;
    *out = nullptr;
    for (;;) {
        int rc = sgfim_anytime_begin(c, a);
        if (rc == SGFIM_OK) {
            sgfim_round_stats st;
            rc = sgfim_anytime_step(c, 0, &st);
        }
        // with the default frontier capacity, an overflow restarts the search with 4x the room,
        // while the larger frontier fits the device memory (and the context's cap) left
        if (rc == SGFIM_E_CAPACITY && c->opts.frontier_cap == 0 && c->SB && c->SB->D.cap < ((int64_t)1 << 26)) {
            const SgSearchDev& D = c->SB->D;
            // bytes per frontier record: itemset, meta, extension mask, keys, free list and
            // its slot in each of the 64 priority stacks
            const int64_t per = (int64_t)(SG_MAX_DEPTH * sizeof(*D.items) + sizeof(*D.meta) + D.dw_max * sizeof(*D.cand) +
                                          sizeof(*D.key) + sizeof(*D.pkey) + sizeof(*D.free_idx) +
                                          SG_NBUCKETS * sizeof(*D.bucket));
            const int64_t cap2 = 4 * D.cap;
            const int64_t need = (cap2 - D.cap) * per;   // the current frontier's buffers are reused
            size_t fr = 0, tt = 0;
            int64_t avail = hipMemGetInfo(&fr, &tt) == hipSuccess ? (int64_t)fr : 0;
            if (c->mem_cap > 0) avail = std::min(avail, c->mem_cap - c->mem_used);
            if (need > avail)
                return fail(c, SGFIM_E_CAPACITY,
                            "frontier of %lld records is full; %lld records would need %.2f GB more device "
                            "memory, %.2f GB available (set opts.frontier_cap, or a higher threshold)",
                            (long long)D.cap, (long long)cap2, need / 1e9, avail / 1e9);
            c->cap_hint = cap2;
            continue;
        }
        if (rc) return rc;
        return sgfim_anytime_take(c, out);
    }
}

int sgfim_result_get(const sgfim_result* r, const int32_t** items, const int64_t** off, const int64_t** support,
                     const int32_t** round, int64_t* n)
{
    if (!r) return SGFIM_E_INVAL;
    if (items) *items = r->items.data();
    if (off) *off = r->off.data();
    if (support) *support = r->sup.data();
    if (round) *round = r->round.data();
    if (n) *n = (int64_t)r->sup.size();
    return SGFIM_OK;
}

int sgfim_result_times(const sgfim_result* r, const double** t)
{
    if (!r || !t) return SGFIM_E_INVAL;
    *t = r->time_s.data();
    return SGFIM_OK;
}

int sgfim_result_windows(const sgfim_result* r, const int32_t** win)
{
    if (!r || !win) return SGFIM_E_INVAL;
    *win = r->seg.data();
    return SGFIM_OK;
}

void sgfim_result_free(sgfim_result* r) { delete r; }

int sgfim_set_window_thresholds(sgfim_ctx* c, const int64_t* min_counts, int32_t n)
{
    if (!c || n < 0 || (n > 0 && !min_counts)) return SGFIM_E_INVAL;
    if (n == 0) {
        c->win_thr.clear();
        return SGFIM_OK;
    }
    if (!c->loaded || n != c->nseg) return fail(c, SGFIM_E_INVAL, "need one threshold per loaded window (%d)", c->nseg);
    for (int32_t w = 0; w < n; w++)
        if (min_counts[w] < 1) return fail(c, SGFIM_E_INVAL, "min_count of window %d must be >= 1", w);
    c->win_thr.assign(min_counts, min_counts + n);
    return SGFIM_OK;
}

int sgfim_set_thresholds(sgfim_ctx* c, const int64_t* min_counts, const int32_t* root_topk, int32_t n)
{
    if (!c || n < 0 || (n > 0 && !min_counts)) return SGFIM_E_INVAL;
    c->multi_thr.clear();
    c->multi_topk.clear();
    if (n == 0) return SGFIM_OK;
    if (!c->loaded || c->nseg != 1) return fail(c, SGFIM_E_STATE, "a threshold list needs a single loaded set");
    if (n > c->vseg_cap) return fail(c, SGFIM_E_INVAL, "at most %d thresholds per search", c->vseg_cap);
    for (int32_t i = 0; i < n; i++) {
        if (min_counts[i] < 1) return fail(c, SGFIM_E_INVAL, "min_count %d must be >= 1", i);
        if (root_topk && root_topk[i] < 0) return fail(c, SGFIM_E_INVAL, "root_topk %d must be >= 0", i);
    }
    c->multi_thr.assign(min_counts, min_counts + n);
    if (root_topk) c->multi_topk.assign(root_topk, root_topk + n);
    return SGFIM_OK;
}

int sgfim_stage_times(const sgfim_ctx* c, double* ms, int32_t n)
{
    if (!c || !ms || n < 0) return SGFIM_E_INVAL;
    for (int32_t i = 0; i < n; i++) ms[i] = (c->S && i < SG_STAGES) ? c->S->stage_ms[i] : 0.0;
    return SGFIM_OK;
}

int sgfim_last_stats(const sgfim_ctx* c, sgfim_round_stats* st)
{
    if (!c || !st) return SGFIM_E_INVAL;
    *st = c->last;
    return SGFIM_OK;
}

int sgfim_set_scorer(sgfim_ctx* c, const float* w1, const float* b1, const float* w2, float b2)
{
    if (!c || !w1 || !b1 || !w2) return SGFIM_E_INVAL;
    pipe_settle(c);
    (void)hipSetDevice(c->device);
    c->scorer_set = false;
    auto bf16 = [](float f) -> uint16_t {
        uint32_t u;
        memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    };
    std::vector<uint16_t> w1t((size_t)SG_HID * 32, 0), w2b(SG_HID);
    for (int col = 0; col < SG_HID; col++)
        for (int k = 0; k < SG_FEAT; k++) w1t[(size_t)col * 32 + k] = bf16(w1[(size_t)k * SG_HID + col]);
    for (int i = 0; i < SG_HID; i++) w2b[i] = bf16(w2[i]);
    int rc;
    SgScorer& s = c->scorer;
    // the weights live in buffers allocated once per context and overwritten in stream
    // order: a captured learned-order round names these addresses
    if (!s.w1 && ((rc = dev_alloc_t(c, c->scorer_bufs, &s.w1, w1t.size())) ||
                  (rc = dev_alloc_t(c, c->scorer_bufs, &s.b1, SG_HID)) ||
                  (rc = dev_alloc_t(c, c->scorer_bufs, &s.w2, SG_HID)) ||
                  (rc = dev_alloc_t(c, c->scorer_bufs, &s.b2, 1)))) {
        dev_free_all(c, c->scorer_bufs);
        s = SgScorer{};
        return rc;
    }
    HIPCHK(c, hipMemcpyAsync(s.w1, w1t.data(), w1t.size() * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.b1, b1, SG_HID * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.w2, w2b.data(), SG_HID * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.b2, &b2, 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->scorer_set = true;
    return SGFIM_OK;
}

int sgfim_copy_probe(sgfim_ctx* c, int64_t bytes, int32_t iters, double* gbps)
{
    (void)hipGetLastError();
    if (!c || bytes < 16 || iters < 1 || !gbps) return SGFIM_E_INVAL;
    (void)hipSetDevice(c->device);
    bytes = bytes / 16 * 16;
    std::vector<DevBuf> tmp;
    void *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = dev_alloc(c, tmp, &a, (size_t)bytes)) || (rc = dev_alloc(c, tmp, &b, (size_t)bytes))) {
        dev_free_all(c, tmp);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemsetAsync(a, 1, (size_t)bytes, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = sg_launch_copy_probe(a, b, bytes, c->stream);   // warm
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int32_t i = 0; i < iters && e == hipSuccess; i++) e = sg_launch_copy_probe(i & 1 ? b : a, i & 1 ? a : b, bytes, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free_all(c, tmp);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "copy probe: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;  // bytes read + bytes written
    return SGFIM_OK;
}

int sgfim_mfma_probe(sgfim_ctx* c, int32_t kind, int32_t iters, double* tops)
{
    (void)hipGetLastError();
    if (!c || iters < 1 || !tops) return SGFIM_E_INVAL;
    if (kind != SGFIM_PAIR_FP4 && kind != SGFIM_PAIR_INT8)
        return fail(c, SGFIM_E_INVAL, "mfma probe %d: expected SGFIM_PAIR_FP4 or SGFIM_PAIR_INT8", kind);
    (void)hipSetDevice(c->device);
    const int32_t blocks = 4 * c->n_cu;   // four workgroups of four waves per CU: four waves per SIMD
    std::vector<DevBuf> tmp;
    float* d = nullptr;
    int rc;
    if ((rc = dev_alloc_t(c, tmp, &d, (size_t)blocks))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = sg_launch_mfma_probe(d, blocks, iters, kind, c->stream);   // warm
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    if (e == hipSuccess) e = sg_launch_mfma_probe(d, blocks, iters, kind, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free_all(c, tmp);
    if (e != hipSuccess) return fail(c, SGFIM_E_DEVICE, "mfma probe: %s", hipGetErrorString(e));
    const double k = kind == SGFIM_PAIR_FP4 ? 64.0 : 32.0;   // K of one instruction
    *tops = 2.0 * 32.0 * 32.0 * k * 8.0 * iters * 4.0 * blocks / (ms * 1e-3) / 1e12;
    return SGFIM_OK;
}

int sgfim_set_trace(sgfim_ctx* c, int32_t enable, int64_t cap)
{
    if (!c || cap < 0) return SGFIM_E_INVAL;
    pipe_settle(c);
    c->trace = enable != 0;
    c->trace_cap = enable ? (cap > 0 ? cap : (int64_t)1 << 22) : 0;
    return SGFIM_OK;
}

// Trace of the last run: births (key, parent key, 24 features) and MFI keys.
int sgfim_trace_get(sgfim_ctx* c, int64_t* n_birth, int64_t* n_mfi, int64_t* keys, float* feat, int64_t* mfi_keys,
                    int64_t cap)
{
    if (!c || !n_birth || !n_mfi) return SGFIM_E_INVAL;
    pipe_settle(c);
    SearchState* S = c->S;
    if (!S || !S->B || !S->B->D.trace_cap) return fail(c, SGFIM_E_STATE, "no traced run");
    (void)hipSetDevice(c->device);
    int rc;
    if ((rc = read_ctl(c, S))) return rc;
    const int64_t nb = std::min(S->B->h_ctl[CTL_TRACE_BIRTH], S->B->D.trace_cap);
    const int64_t nm = std::min(S->B->h_ctl[CTL_TRACE_MFI], S->B->D.trace_cap);
    *n_birth = nb;
    *n_mfi = nm;
    if (S->B->h_ctl[CTL_TRACE_BIRTH] > S->B->D.trace_cap || S->B->h_ctl[CTL_TRACE_MFI] > S->B->D.trace_cap)
        return fail(c, SGFIM_E_CAPACITY, "trace capacity %lld exceeded", (long long)S->B->D.trace_cap);
    if (!keys && !feat && !mfi_keys) return SGFIM_OK;
    if (cap < std::max(nb, nm)) return fail(c, SGFIM_E_INVAL, "trace output capacity too small");
    if (keys) HIPCHK(c, hipMemcpyAsync(keys, S->B->D.trace_key, nb * 16, hipMemcpyDeviceToHost, c->stream));
    if (feat) HIPCHK(c, hipMemcpyAsync(feat, S->B->D.trace_feat, nb * SG_FEAT * 4, hipMemcpyDeviceToHost, c->stream));
    if (mfi_keys) HIPCHK(c, hipMemcpyAsync(mfi_keys, S->B->D.trace_mfi, nm * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SGFIM_OK;
}

}  // extern "C"
