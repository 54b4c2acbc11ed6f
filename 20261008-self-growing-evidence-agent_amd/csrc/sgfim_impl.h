// sgfim_impl.h -- internal declarations shared by the host driver and the HIP kernels.
//
// Data layout in HBM (one context = one transaction set, or one set of windows):
//   rows      uint64 [d_pad][Wp]   vertical tidset bitmaps; bit (t & 63) of word (t >> 6)
//                                  of row x is set iff transaction t holds item x.
//                                  Wp = W rounded up to 64 words (512 B), pad words and
//                                  rows d..d_pad-1 are zero.
//   supports  int32  [d]           row popcounts (item supports)
//   pair      int32  [d_pad][d_pad] pair co-occurrence matrix (int8 MFMA)
//   pair_ok   uint64 [d][DW]       bit j of row i set iff pair(i, j) >= min_count
// A "segment" is one transaction set (one window of the batched-window workload);
// every kernel addresses rows through the segment table, so one launch can mix
// windows with different lengths.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SG_MAX_DEPTH 64
#define SG_NBUCKETS 64
#define SG_WAVE 64
#define SG_BMAX 16384   // largest batch of frontier nodes per round
#define SG_DMAX 4096    // most items one search handles (a larger loaded set is compacted to
                        // its frequent items first)

struct SgSegment {
    int64_t row_base;   // word offset of row 0 of this segment in the rows arena
    int64_t stride;     // words per row (Wp of the segment)
    int64_t words;      // W of the segment (ceil(n/64))
    int64_t pair_base;  // int32 offset of the segment's pair matrix
    int64_t ok_base;    // uint64 offset of the segment's pair_ok mask
    int32_t d;          // items in the segment
    int32_t dpad;       // rows allocated (multiple of 128)
    int32_t dw;         // 64-bit words per candidate mask: ceil(d / 64)
    int32_t thr;        // min_count
    int32_t n;          // transactions
    int32_t pad_;
};

// One frontier record: an itemset X whose frequent one-item extensions U = F(X) are
// known (the candidate mask) and whose children X u {e}, e in U below `last`, are still
// to be counted. Items are dense ids of the record's segment.
struct SgNodeMeta {
    int32_t k;          // |X|
    int32_t sup;        // support of X
    int32_t last;       // the enumeration tail bound: children take items below it
    int32_t seg;        // segment id
    int32_t parent_sup; // support of the parent (0 for roots)
    int32_t parent_nf;  // |F(parent)|
    int32_t ncand;      // |U|
    int32_t ntail;      // children: |U below last|
    int32_t nblk;       // blocks of children in the record's counting tiles
    int32_t nrange;     // row-chunk ranges of the record's window
    float score;        // priority score
    int64_t sib;        // the parent's sibling block in the sibling ring (-1: none, a root)
    int64_t proj;       // projected block the record's subtree counts on (-1: none; see SgProj)
    int32_t pk;         // leading items of X that the projected block has already applied
    int32_t pad_;
};

// A projected block: the rows of U(Y) restricted to the transactions of T(Y) for an
// itemset Y of the search, bit-compacted (bit j of a row = the j-th transaction of T(Y)).
// Every descendant X of Y counts on it: T(X) = T(Y) & AND of X's later items, and a
// candidate's row is |T(Y)| / 64 words instead of the window's. Blocks live in a ring of
// words (the arena) and a ring of descriptors, with absolute positions that only grow; a
// block is used while neither ring has wrapped onto it, otherwise its records fall back to
// the window's rows. Row r of a block is the r-th item of U(Y) (ascending).
struct SgProj {
    int64_t off;        // absolute word position of row 0 in the arena ring
    int64_t words;      // words per row: ceil(n / 64)
    int64_t stride;     // words between rows (words rounded up to 8)
    int32_t n;          // transactions: |T(Y)|
    int32_t nrows;      // rows: |U(Y)|
};

// One nonzero source word of a new block's transaction set: T(Y) restricted to source word
// w is the bit mask m, and its transactions are the block's bits p .. p + popcount(m) - 1.
struct SgPWord {
    int32_t w;
    int32_t p;
    uint64_t m;
};

// Launchers (kernels_*.hip); all launch on the given stream and return hipError_t.
hipError_t sg_launch_transpose(const int64_t* off, const int32_t* items, const int32_t* idmap,
                               int64_t n_txn, int32_t id_max, uint64_t* rows, int32_t d,
                               int32_t dpad, int64_t Wp, int32_t* supports, hipStream_t s);
// Row keep[i] of every segment of (src, src_sup) -> row i of (dst, dst_sup), for i < dc;
// the segment tables give both arenas' row bases (equal strides).
hipError_t sg_launch_gather_rows(const uint64_t* src, const int32_t* src_sup, const SgSegment* src_segs,
                                 int32_t src_dpad, uint64_t* dst, int32_t* dst_sup, const SgSegment* dst_segs,
                                 int32_t dst_dpad, const int32_t* keep, int32_t dc, int32_t nseg, hipStream_t s);

hipError_t sg_launch_gen_zipf(uint64_t* rows, int64_t Wp, int64_t n_txn, int32_t d, const uint64_t* q53,
                              const uint64_t* r53, const int32_t* item_sets_off, const int32_t* item_sets,
                              uint64_t seed, int32_t* sup, hipStream_t s);
hipError_t sg_launch_copy_probe(const void* src, void* dst, int64_t bytes, hipStream_t s);
// MFMA throughput probe: `blocks` workgroups of 4 waves, iters x 8 MFMAs per wave; out [blocks].
hipError_t sg_launch_mfma_probe(float* out, int32_t blocks, int32_t iters, int32_t kind, hipStream_t s);

// Pair-matrix kernels: fp4 (e2m1) MFMA with fp32 sums, or int8 MFMA with int32 sums.
#define SG_PAIR_FP4 0
#define SG_PAIR_INT8 1
hipError_t sg_launch_pair_mfma(const uint64_t* rows, int32_t dpad, int64_t Wp, int64_t W,
                               int32_t* pair, int32_t kind, int32_t split_cap, hipStream_t s);
// p[0 .. n) = 0 with a kernel (16-byte stores).
hipError_t sg_launch_zero32(int32_t* p, int64_t n, hipStream_t s);
hipError_t sg_launch_pair_valu(const uint64_t* rows, int32_t dpad, int64_t Wp, int64_t W,
                               int32_t* pair, hipStream_t s);
// Batched pair matrices over segments (one launch for all windows).
hipError_t sg_launch_pair_mfma_segs(const uint64_t* rows, const SgSegment* segs, int32_t nseg,
                                    int32_t max_dpad, int64_t max_words, int32_t* pair, int32_t kind,
                                    int32_t split_cap, hipStream_t s);
hipError_t sg_launch_pair_ok(const int32_t* pair, const SgSegment* segs, int32_t nseg,
                             int32_t max_d, uint64_t* ok, hipStream_t s);

// Generic batched AND+popcount over explicit node descriptors (used by sgfim_count_batch).
struct SgCountNode {
    int32_t k;          // prefix length
    int32_t ncand;      // candidates
    int64_t cand_off;   // offset into cand / out
    int32_t seg;        // segment
    int32_t nchunk;     // chunks of this node
    int64_t unit_off;   // first work unit of this node
    int32_t items[SG_MAX_DEPTH];
};
// Allocation-free count batches: at most SG_NOALLOC_NODES prefixes and SG_NOALLOC_CANDS
// candidates per call, in a static device scratch of SG_NOALLOC_BYTES per device laid out as
// [nodes | candidate ids | counts | work counter].
#define SG_NOALLOC_NODES 16384
#define SG_NOALLOC_CANDS (1 << 20)
#define SG_NOALLOC_NODE_BYTES ((size_t)SG_NOALLOC_NODES * sizeof(SgCountNode))
#define SG_NOALLOC_BYTES (SG_NOALLOC_NODE_BYTES + (size_t)SG_NOALLOC_CANDS * 8 + 256)
#define SG_NOALLOC_MAX_DEVICES 64
// The device address of `device`'s scratch (looked up once per device, then cached).
hipError_t sg_count_scratch(int device, void** ptr);
hipError_t sg_launch_count_nodes(const uint64_t* rows, const SgSegment* segs,
                                 const SgCountNode* nodes, int32_t n_nodes, int64_t n_units,
                                 const int32_t* cand, int32_t* out, int32_t chunk_words,
                                 int32_t* unit_counter, int32_t grid, hipStream_t s);

// Search engine device state (kernels_search.hip).
struct SgSearchDev {
    // arena
    uint16_t* items;     // [cap][SG_MAX_DEPTH]
    SgNodeMeta* meta;    // [cap]
    uint64_t* cand;      // [cap][dw_max] candidate masks
    int64_t* key;        // [cap] node key (unique per run)
    int64_t* pkey;       // [cap] parent key
    int32_t* free_idx;   // [cap]
    int32_t* bucket;     // [SG_NBUCKETS][cap]
    // batch
    int32_t* batch;      // [bmax + 1]
    int32_t* coff;       // [bmax + 1] candidate offsets
    int64_t* unit_off;   // [bmax + 1] child-block offsets of the batch's records
    int32_t* clist;      // [bmax * d_max] U lists of the batch (ascending ids)
    int32_t* ccount;     // [bmax * d_max] child supports |T(X) & T(e)|
    int32_t* choff;      // [bmax + 1] child offsets
    int64_t* celloff;    // [bmax + 1] count-matrix offsets
    int32_t* cells;      // [cell_cap] child x candidate counts
    int32_t* chrec;      // [newborn_cap] batch record of each child of the round
    int32_t* blkrec;     // [newborn_cap / SG_CB + bmax] batch record of each child block
    int64_t cell_cap;
    // Sibling ring: a record X expanded in a round writes a block of (1 + ntail) rows of
    // dw_max words: its U, then F(X u {e}) for each child e. A child record Y of X keeps the
    // block's offset; when Y is counted, child e of Y needs candidate f only if f is in
    // F(X u {e}) (|T(Y u {e, f})| <= |T(X u {e, f})|), a tighter test than the pair matrix.
    // Offsets are absolute positions in an endless stream of words, stored modulo sib_cap;
    // a block is readable while it lies in the latest sib_cap words written.
    uint64_t* sib;       // [sib_cap]
    int64_t* sib_off;    // [bmax + 1] block of each batch record (-1: none)
    int64_t sib_cap;
    uint64_t* need;      // [newborn_cap][dw_max] need row of each child of the round
    // children born in the current round
    int32_t* newborn;    // [newborn_cap]
    float* feat;         // [newborn_cap][SG_FEAT]
    float* scores;       // [newborn_cap]
    // MFI ring
    uint16_t* ring_items;   // [ring_cap][SG_MAX_DEPTH]
    int32_t* ring_meta;     // [ring_cap][4]: k, sup, seg, round
    // optional search trace (training data for the proposal network)
    int64_t* trace_key;     // [trace_cap][2]: key, parent key
    float* trace_feat;      // [trace_cap][SG_FEAT]
    int64_t* trace_mfi;     // [trace_cap] keys of emitted MFIs
    int64_t trace_cap;
    // control block (device int64s, see SgCtl)
    int64_t* ctl;
    // parameters
    int64_t cap;
    int64_t newborn_cap;
    int64_t ring_cap;
    int32_t dw_max;
    int32_t d_max;
    int32_t bmax;
    int32_t chunk_words;  // row words per counting chunk (128, 256 or 512)
    int32_t chunk_range;  // chunks per counting unit
    int32_t count_grid;
    int32_t feat_on;     // compute the 24 features of newborn nodes (learned order or tracing)
    // projected blocks (SgProj); proj_factor 0 turns them off
    int32_t proj_factor;  // project a record when its support is at most 1/proj_factor of its view's transactions
    int32_t proj_dense;   // a new block whose support is at least 1/proj_dense of its source's is compacted
                          // word by word (k_proj_compress), a sparser one gathered bit by bit
    int32_t proj_mintail; // children a record needs for a compacted block (2 for a gathered one)
    int32_t count_kind;   // counting kernel: SG_COUNT_F4 (fp4 MFMA) or SG_COUNT_VALU (AND + popcount)
    uint64_t* parena;     // [parena_cap + proj_max_block] words
    int64_t parena_cap;
    int64_t proj_max_block;  // largest block (words)
    int64_t proj_round_words;  // arena words one round may allocate (blocks used in a round are
                               // those the ring cannot reach within it)
    SgProj* pdesc;        // [pdesc_cap]
    uint64_t* pmask;      // [pdesc_cap][dw_max] U(Y) of each block
    int32_t* prank;       // [pdesc_cap][dw_max] items of U(Y) in the words before each word
    int64_t pdesc_cap;
    int32_t* ptid;        // [ptid_cap] transaction lists T(Y) of the round's gathered blocks, and
                          // the first word entry of each compress unit of the compacted ones
    int64_t ptid_cap;
    SgPWord* pent;        // [pent_cap] nonzero source words of the round's compacted blocks
    int64_t pent_cap;
    // per batch record (bmax + 1 entries)
    int64_t* bview;       // block the record counts on this round (-1: the window's rows)
    int64_t* bsrc;        // a new block's source: the block it is projected from (-1: the window)
    int32_t* bpnew;       // 1: the record's block is projected this round
    int32_t* bnr;         // row-chunk ranges of the record's view
    int64_t* bunitB;      // counting units of records on blocks (prefix; those records run record-major)
    int64_t* bunitT;      // transaction-list units of new blocks (prefix)
    int64_t* bunitG;      // gather units of new blocks (prefix)
    int64_t* bunitC;      // compress units of new blocks (prefix)
    int64_t* peoff;       // a compacted block's word entries in pent
    int64_t* bpa;         // arena position of the record's new block (absolute; the prefix at the cut)
    int64_t* bpd;         // descriptor id of the record's new block (absolute; the prefix at the cut)
    int64_t* ptoff;       // a new block's transaction list in ptid
    int64_t* pfill;       // transactions written into that list (compacted blocks: word entries
                          // << 40 | transactions)
    int32_t* crow;        // [bmax * d_max] row of each U-list item in its record's view
};

// Control block layout (int64 slots of SgSearchDev::ctl).
enum SgCtl {
    CTL_FREE_TOP = 0,     // free-list size
    CTL_BATCH_N,          // nodes in the current batch
    CTL_UNITS,            // counting work units of the current batch
    CTL_UNIT_NEXT,        // persistent-kernel work counter
    CTL_RING_POS,         // MFIs written so far (ring position = pos % ring_cap)
    CTL_NODES,            // nodes (children of records, plus roots) processed (cumulative)
    CTL_COUNTS,           // needed support counts computed by AND+popcount (cumulative)
    CTL_WORDS,            // algorithmic words: (needed counts + 1) * W per node (cumulative)
    CTL_ERR,              // error flags: bit 0 capacity, bit 1 depth
    CTL_ROUND,            // rounds done
    CTL_NEWBORN,          // records born this round
    CTL_BLOCKS,           // child blocks of the current batch (units = blocks x chunk ranges)
    CTL_KEY,              // next record key
    CTL_TRACE_BIRTH,      // trace records written (births)
    CTL_TRACE_MFI,        // trace records written (MFIs)
    CTL_T0,               // earliest wall_clock64() at which a counting workgroup started (this round)
    CTL_BUCKET_TOP0 = 16, // SG_NBUCKETS slots
    CTL_T1 = 16 + SG_NBUCKETS, // latest wall_clock64() at which a counting workgroup finished
    CTL_CHILDREN,         // children to expand in the current batch
    CTL_CELLS,            // count-matrix cells of the current batch
    CTL_SIB_NEXT,         // sibling-ring words written so far (end of the current round's blocks)
    CTL_PROJ_NEXT,        // projected-block descriptors allocated so far (absolute)
    CTL_ARENA_NEXT,       // projected-block arena words allocated so far (absolute)
    CTL_UNITS_A,          // counting units of the records on window rows (range-major, first)
    CTL_PROJ_TUNITS,      // transaction-list units of the round's new blocks
    CTL_PROJ_GUNITS,      // gather units of the round's new blocks
    CTL_PROJS,            // projected blocks built (cumulative)
    CTL_PROJ_WORDS,       // words read and written building them (cumulative)
    CTL_PROJ_ROWS,        // of them, projected from window rows (cumulative)
    CTL_PROJ_ARENA,       // arena words of the blocks built (cumulative)
    CTL_PROJ_CUNITS,      // compress units of the round's new blocks
    CTL_SLOTS = 32 + SG_NBUCKETS
};

struct SgScorer;  // proposal network weights on the device (kernels_mlp.hip)

hipError_t sg_sensor_columns(const double* x, int64_t n, int32_t S, double* xc, hipStream_t s);
// Order statistics of each column of xc[S][n]: R ascending ranks shared by every column
// (rank[S][R] initialised to them, pref/grp/gpref zero, ngrp 1, hist zero); out[s][j] is the
// value at rank umap[j] of column s.
hipError_t sg_sensor_select(const double* xc, int64_t n, int32_t S, int32_t R, int64_t* rank, uint64_t* pref,
                            int32_t* grp, int32_t* ngrp, uint64_t* gpref, uint32_t* hist, const int32_t* umap,
                            int32_t J, double* out, hipStream_t s);
hipError_t sg_sensor_bins(const double* x, int32_t S, int32_t B, int32_t window, int32_t stride, int64_t n_win,
                          const double* edges, const int32_t* ne, int32_t* items, uint32_t* present, int64_t* off,
                          hipStream_t s);
// Online (P^2) variant: E [n_win][S][B + 1] receives the edges used for each window.
hipError_t sg_online_bins(const double* x, int32_t S, int32_t B, int32_t window, int32_t stride, int64_t n_win,
                          double* E, int32_t* items, uint32_t* present, int64_t* off, hipStream_t s);
#define SG_COUNT_VALU 0
#define SG_COUNT_F4 1
int sg_count_group_blocks_per_cu(int32_t chunk_words, int32_t kind);
hipError_t sg_search_init(const SgSearchDev& S, hipStream_t s);
hipError_t sg_search_seed(const SgSearchDev& S, const SgSegment* segs, const int32_t* pair,
                          const uint64_t* ok, const int32_t* roots, int32_t n_roots,
                          const int32_t* sup_all, const int64_t* seg_item_base, int32_t order_mode,
                          const SgScorer* scorer, hipStream_t s);
// Words per transaction-list unit and transactions per gather unit of the projection kernels.
#define SG_PT_WORDS 1024
#define SG_PG_TIDS 1024
#define SG_PG_ROWS 64
// Compacted blocks (dense enough, SgSearchDev::proj_dense) are built word by word
// (k_proj_compress) instead of bit by bit (k_proj_gather); a compress unit is SG_PC_WORDS
// block words of SG_PC_ROWS rows.
#ifndef SG_PC_WORDS
#define SG_PC_WORDS 32
#endif
#ifndef SG_PC_ROWS
#define SG_PC_ROWS 16
#endif
// Stages of a search round, timed with events on direct launches: selection + U lists,
// the new blocks' transaction lists, their gathered rows, their compacted rows, need rows,
// counting, expansion, scoring + commit.
#define SG_STAGES 8
hipError_t sg_search_round(const SgSearchDev& S, const uint64_t* rows, const SgSegment* segs,
                           const int32_t* pair, const uint64_t* ok, const int32_t* sup_all,
                           const int64_t* seg_item_base, int32_t uniform, int32_t order_mode,
                           const SgScorer* scorer, const hipEvent_t* ev, hipStream_t s);

// Proposal network (kernels_mlp.hip).
#define SG_FEAT 24
#define SG_HID 256
struct SgScorer {
    uint16_t* w1;   // bf16 [SG_HID][32] (hidden-major, features padded to 32)
    float* b1;      // [SG_HID]
    uint16_t* w2;   // bf16 [SG_HID]
    float* b2;      // [1]
};
// Scores out[i] = sigmoid(MLP(feat[i])) for i < *n_dev (n_dev on the device), i < n_max.
hipError_t sg_launch_score(const SgScorer* sc, const float* feat, int32_t n_max,
                           const int64_t* n_dev, float* out, hipStream_t s);

// Training state of the proposal network (kernels_mlp.hip).
// Master weights (FP32): W1 [SG_FEAT][SG_HID], b1 [SG_HID], w2 [SG_HID], b2 [1]; the
// packed parameter/gradient/moment order is W1, b1, w2, b2.
struct SgTrainState {
    float* W1; float* b1; float* w2; float* b2;   // master weights
    float* m;  float* v;                          // Adam moments (packed)
    uint16_t* w1t; uint16_t* w2b;                 // BF16 copies: w1t [SG_HID][32], w2b [SG_HID]
    float* grad;                                  // packed gradient of the last step
    float* loss;                                  // [1]
};
hipError_t sg_mlp_pack(const SgTrainState& T, hipStream_t s);
hipError_t sg_mlp_train_step(const SgTrainState& T, const float* X, const float* Y, const int32_t* idx, int32_t n,
                             float lr, float beta1, float beta2, float eps, int32_t step, hipStream_t s);
#define SG_NPARAM (SG_FEAT * SG_HID + SG_HID + SG_HID + 1)
hipError_t sg_mlp_forward(const uint16_t* w1t, const float* b1, const uint16_t* w2b, const float* b2,
                          const float* X, int32_t n, float* out, int32_t logits, hipStream_t s);
