/*
 * sgfim.h -- C ABI of the MI355X (gfx950) maximal-frequent-itemset support-counting library.
 *
 * The reference drives its mining engines as subprocesses over FIMI files
 * (code/pipeline/engine_adapter.py:70-131). This ABI replaces that process
 * boundary with a shared library: integer transactions go in as CSR arrays,
 * maximal frequent itemsets (MFIs) with exact supports come out. Every entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status: SGFIM_OK (0) or a negative
 *     SGFIM_E_* code; sgfim_last_error() holds a message for the last failure.
 *   - Item ids are the FIMI ids (>= 1). Negative ids are rejected (the
 *     reference's parser drops them, kernel_bitmap.cu:225-229; the Python
 *     adapter never writes them, engine_adapter.py:45-48). Id 0 is rejected
 *     too: the reference counts it in masks but never as a candidate
 *     (kernel_bitmap.cu:356-357), which no caller relies on.
 *   - Counts are exact integers; a transaction set must have fewer than 2^31
 *     transactions (the reference stores counts in int, kernel_bitmap.cu:412).
 *   - Caller-owned input arrays are copied before the call returns; results
 *     are library-owned until sgfim_result_free().
 *   - One context per host thread; a context owns one HIP stream on one
 *     device and all of its device memory. Contexts share nothing.
 */
#ifndef SGFIM_H
#define SGFIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sgfim_status {
    SGFIM_OK = 0,
    SGFIM_E_INVAL = -1,      /* invalid argument                                  */
    SGFIM_E_NOMEM = -2,      /* device or host allocation failed / memory cap hit */
    SGFIM_E_DEVICE = -3,     /* HIP runtime error, or no usable gfx950 device     */
    SGFIM_E_DEPTH = -4,      /* an itemset longer than SGFIM_MAX_DEPTH would be needed */
    SGFIM_E_CANCELLED = -5,  /* sgfim_anytime_cancel() stopped the run            */
    SGFIM_E_TIMEOUT = -6,    /* the deadline passed before the search finished    */
    SGFIM_E_STATE = -7,      /* call out of order (e.g. mining before loading)    */
    SGFIM_E_CAPACITY = -8    /* a device table (frontier, MFI ring) is too small  */
};

#define SGFIM_MAX_DEPTH 64   /* longest itemset; the reference's BMMA packer stops at 31 (kernel_bmma.cu:798),
                                its itemsets are vectors (kernel_bitmap.cu:38-43) */

/* Mining flags (sgfim_mine / sgfim_anytime_begin). */
#define SGFIM_ROOT_STRICT  0x1u  /* TensorFIM root rule: only items with freq/n > rel_thr root a
                                    subtree (kernel_bitmap.cu:816,841-850); needs opts.rel_thr */
#define SGFIM_ROOT_TOPK    0x2u  /* AnyFIM stage rule: only the first root_topk items of the
                                    stable descending-frequency order root a subtree (kernel.cu:850-859) */
#define SGFIM_ORDER_LEARNED 0x4u /* order the frontier by the proposal network's score (64
                                    score buckets, highest first); never changes the result */
#define SGFIM_NO_GRAPH     0x8u  /* launch each round's kernels directly instead of replaying
                                    the round's captured hipGraph */
#define SGFIM_FRESH_PAIR   0x10u /* recompute the pair matrix that seeds the search (on the
                                    device, in the run's stream order) instead of reusing the
                                    one kept from an earlier run; sgfim_pair_time times it */

typedef struct sgfim_ctx sgfim_ctx;
typedef struct sgfim_result sgfim_result;

typedef struct sgfim_opts {
    int32_t device;          /* HIP device ordinal                                         */
    int64_t mem_cap_bytes;   /* device-memory cap for this context; 0 = no cap             */
    int32_t batch_nodes;     /* frontier nodes per search round; 0 = default               */
    int32_t chunk_words;     /* 64-bit words per counting work unit; 0 = autotuned/default */
    int64_t frontier_cap;    /* frontier node capacity; 0 = default                        */
    int64_t ring_cap;        /* MFI output ring capacity (itemsets); 0 = default           */
} sgfim_opts;

/* Search flags and thresholds for one mining run. */
typedef struct sgfim_mine_args {
    int64_t min_count;       /* an itemset is frequent iff support >= min_count (>= 1)     */
    uint32_t flags;          /* SGFIM_* flags above                                        */
    double rel_thr;          /* SGFIM_ROOT_STRICT: relative threshold s                    */
    int32_t root_topk;       /* SGFIM_ROOT_TOPK: number of roots                           */
    double deadline_s;       /* wall-clock budget in seconds; 0 = none                     */
} sgfim_mine_args;

/* Round statistics reported by the anytime driver. */
typedef struct sgfim_round_stats {
    int64_t round;           /* rounds completed                                           */
    int64_t nodes;           /* search nodes processed so far (roots, then every child
                                X u {e} of a frontier record that a round counted)         */
    int64_t counts;          /* support counts (itemset, item) computed by AND+popcount    */
    int64_t mfis;            /* MFIs confirmed so far                                      */
    int64_t frontier;        /* records (itemset X with its frequent extensions U) waiting
                                in the frontier; a round counts all children of a record   */
    int64_t words;           /* algorithmic 64-bit words of the counting kernel so far:
                                (needed counts + 1) * W per child, W = the child's window
                                words; a needed count is a candidate f of U \ {e} whose
                                pair (e, f) is frequent and that is not a later child of
                                the same record (that support is counted once, in the
                                later child's row)                                        */
    int64_t pair_counts;     /* counts taken from the pair matrix (roots x items)          */
    int64_t count_launches;  /* launches of the counting kernel (one per round queued,
                                including a last one queued after the frontier emptied,
                                which returns at once)                                     */
    double count_ms;         /* summed duration of the counting kernel, from the device's
                                constant-rate wall clock (first workgroup start to last end) */
    double event_ms;         /* the same duration from HIP events around the kernel; only
                                with SGFIM_NO_GRAPH (0 otherwise)                          */
    double elapsed_s;        /* wall time since sgfim_anytime_begin                        */
    int32_t done;            /* 1 when the frontier is empty: every MFI is confirmed       */
    int32_t status;          /* SGFIM_OK or the error that ended the run                   */
    int64_t projections;     /* projected blocks built so far (sgfim_set_projection)       */
    int64_t proj_words;      /* 64-bit words the projections read and wrote: each new
                                block's unapplied prefix rows and the U rows' words at its
                                transactions in its source view, plus its own rows         */
    int64_t proj_from_rows;  /* of the projected blocks, those built from window rows (the
                                others from an ancestor's block)                           */
    int64_t proj_block_words;/* 64-bit words of the blocks built (their rows x strides)    */
} sgfim_round_stats;

/* ---- context (replaces the engine process: kernel_bitmap.cu:796 main) ---- */
int sgfim_create(const sgfim_opts* opts, sgfim_ctx** out);
void sgfim_destroy(sgfim_ctx* ctx);
const char* sgfim_last_error(const sgfim_ctx* ctx);
const char* sgfim_version(void);

/* Load integer transactions and build the vertical tidset bitmaps on the GPU.
 * Replaces readTransSetFile + getItemsFrequenc + buildBitmap + uploadBitmapToGPU
 * (kernel_bitmap.cu:205-298, 342-388). txn_off has n_txn + 1 entries; row t is
 * items[txn_off[t] .. txn_off[t+1]); duplicates inside a row are allowed (a row
 * is a set). An empty row is an empty transaction and counts in n_txn
 * (kernel_bitmap.cu:219). Every load replaces the loaded data; the search buffers of
 * earlier searches stay allocated for the next search when they fit it. */
int sgfim_load_csr(sgfim_ctx* ctx, const int64_t* txn_off, int64_t n_txn,
                   const int32_t* items, int64_t nnz);

/* Load many independent windows in one context: window w holds transactions
 * [win_off[w], win_off[w+1]) of the CSR (win_off has n_windows + 1 entries). Each
 * window gets its own bitmaps, supports, pair matrix and threshold, and one search
 * run mines all windows together: every counting launch mixes candidates of
 * windows of different lengths. No reference counterpart as one call: the
 * reference mines each window's FIMI file with its own engine process
 * (engine_adapter.py:70-131). */
int sgfim_load_windows(sgfim_ctx* ctx, int32_t n_windows, const int64_t* win_off, const int64_t* txn_off,
                       int64_t n_txn, const int32_t* items, int64_t nnz);
int sgfim_num_windows(const sgfim_ctx* ctx, int32_t* n);
/* Item supports of one window (sgfim_item_ids order). */
int sgfim_window_supports(sgfim_ctx* ctx, int32_t window, int64_t* out);
/* Per-window min_count for the next runs (n = number of windows; n = 0 clears, and
 * sgfim_mine_args.min_count then applies to every window). */
int sgfim_set_window_thresholds(sgfim_ctx* ctx, const int64_t* min_counts, int32_t n);

/* Sensor matrix -> transactions on the GPU (replaces the reference's transactionisation,
 * skab_spike.py:37-80 fit_bins / bin_index / window_transactions, and the FIMI text
 * round trip): x is row-major [n_rows][n_sensors] (finite values); per sensor the bin
 * edges are unique(quantile(column, linspace(0, 1, n_bins + 1))) over the whole column
 * (NumPy's linear interpolation; a constant column gets [v - 0.5, v + 0.5]); windows
 * [s, s + window) for s = 0, stride, ... give one transaction each, holding for every
 * sensor the item s * n_bins + b + 1 of the bin b of the window's median (the mean of the
 * two middle values for an even window). window <= 64; window = stride = 1 bins rows. */
int sgfim_load_sensors(sgfim_ctx* ctx, const double* x, int64_t n_rows, int32_t n_sensors, int32_t n_bins,
                       int32_t window, int32_t stride);

/* Online (streaming) variant for the self-growing pattern library: per sensor, n_bins - 1
 * P^2 quantile estimators at i / n_bins follow the stream (replaces online_discretize.py:
 * 22-108 P2Quantile / OnlineBinner / bin_index, driven as run_file :111-131 and the
 * ablation scripts drive them): for each window [s, s + window), s = 0, stride, ..., every
 * sample of the window is added to the sensor's estimators (overlapping windows add a
 * sample again), then the window median is binned with the edges of that moment --
 * sorted estimates, minimum gap 0.05 of their span, outer edges one span outside. The
 * items and the loaded bitmaps are as for sgfim_load_sensors. When edges is not NULL it
 * receives the edges used, [windows][n_sensors][n_bins + 1] doubles. n_bins >= 2,
 * window <= 64. */
int sgfim_load_sensors_online(sgfim_ctx* ctx, const double* x, int64_t n_rows, int32_t n_sensors, int32_t n_bins,
                              int32_t window, int32_t stride, double* edges);

/* Counting-tile shape of the search: chunk_words row words per chunk (the 16 child
 * prefixes of one chunk sit in LDS: 128, 256 or 512; 0 = the default 256) and chunk_range
 * chunks per work unit (1..64; 0 = the default 4). Results do not depend on it; it trades
 * LDS and registers (occupancy) against set-up work per unit. sgfim/autotune.py picks it
 * per device and input shape. No counterpart in the reference, whose engines take such
 * knobs from environment variables read at start-up (kernel_bmma.cu:351-368). */
int sgfim_set_tile(sgfim_ctx* ctx, int32_t chunk_words, int32_t chunk_range);

/* Words of the search's sibling ring (0 = default: four rows of candidate-mask words per
 * frontier record, 8 MB to 2 GB). A record expanded in a round writes a block of its
 * children's frequent-extension masks F(X u {e}) to the ring; when a child record is
 * counted, child e of it needs candidate f only if f is in F(X u {e}), a tighter test than
 * the pair matrix, exact since |T(Y u {e, f})| <= |T(X u {e, f})| for Y >= X. A block that
 * later writes have wrapped onto falls back to the pair test. Results do not depend on it;
 * a larger ring prunes more counts on deep searches (1 = none). No reference counterpart. */
int sgfim_set_sib_ring(sgfim_ctx* ctx, int64_t words);
int sgfim_get_tile(const sgfim_ctx* ctx, int32_t* chunk_words, int32_t* chunk_range);

/* Projected blocks of later searches. With factor >= 2, a frontier record X whose support
 * is at most 1/factor of the transactions its rows span gets the rows of its frequent
 * extensions U(X) restricted to T(X) and bit-compacted (|T(X)| / 64 words per row instead
 * of the window's), built on the device in the round that counts X; X and its whole
 * subtree then count on them, and a deeper record can be projected again from them. The
 * reference has no counterpart: its counter always ANDs the full-length rows of every item
 * of X for every node (kernel_bitmap.cu:390-442), which is what a record without a block
 * does here. Blocks live in a ring of arena_words 64-bit words (0 = default: 32x the
 * loaded rows and at least 2^28 words, at most a quarter of the device memory); a block that later blocks have overwritten is not
 * used, its records count on the window's rows. factor 0 turns projection off. Results do
 * not depend on either value. */
int sgfim_set_projection(sgfim_ctx* ctx, int32_t factor, int64_t arena_words);
int sgfim_get_projection(const sgfim_ctx* ctx, int32_t* factor, int64_t* arena_words);

/* How projected blocks of later searches are built. A block whose support is at least
 * 1/dense of its source's transactions (dense = 0: none) is compacted word by word: each
 * nonzero word of T(X) in the source view is bit-compressed in every row, about 40 lane
 * operations per source word and row, which a record repays only with min_children or more
 * children (2..4096, default 16); sparser blocks are gathered bit by bit and need 2
 * children. Default dense = 32. No reference counterpart (see sgfim_set_projection).
 * Results do not depend on either value. */
int sgfim_set_block_rule(sgfim_ctx* ctx, int32_t dense, int32_t min_children);
int sgfim_get_block_rule(const sgfim_ctx* ctx, int32_t* dense, int32_t* min_children);
/* The context's device: its gcnArchName (NUL-terminated, at most len - 1 characters) and
 * compute units; the key of the tile autotuner's cache. No counterpart in the reference
 * (its Results header prints the device, kernel_bitmap.cu:608-690). */
int sgfim_device_info(const sgfim_ctx* ctx, char* name, int32_t len, int32_t* n_cu);

/* Items [windows][n_sensors] of the latest sensor load (either variant): *n receives their
 * number, and when items is not NULL the first min(cap, *n) are copied to it. */
int sgfim_sensor_items(sgfim_ctx* ctx, int32_t* items, int64_t cap, int64_t* n);

/* Several thresholds of one loaded set in one batched search: the next runs mine the
 * exact MFI family at each min_counts[i] (with SGFIM_ROOT_TOPK, root_topk[i] roots for
 * threshold i when root_topk is given) and tag every MFI with i (sgfim_result_windows).
 * Replaces the AnyFIM stage loop, which re-mines every stage from scratch
 * (kernel.cu:844-907): the stages share the bitmaps, the pair matrix and every round.
 * n = 0 clears; at most 256 thresholds. */
int sgfim_set_thresholds(sgfim_ctx* ctx, const int64_t* min_counts, const int32_t* root_topk, int32_t n);

/* Synthetic dense-sparse set generated on the GPU straight into the bitmaps: item i
 * (1..n_items) is present independently with probability q_i = min(1, c i^-zipf_s),
 * c set so that sum q_i = mean_len; n_planted itemsets of size_lo..size_hi items are
 * ORed in at rates drawn in [rate_lo, rate_hi]. All draws are counter-based hashes of
 * (seed, transaction, item or set), so any transaction can be regenerated on the CPU.
 * No reference counterpart (the reference's inputs are files). */
int sgfim_load_zipf(sgfim_ctx* ctx, int64_t n_txn, int32_t n_items, double zipf_s, double mean_len,
                    int32_t n_planted, int32_t size_lo, int32_t size_hi, double rate_lo, double rate_hi,
                    uint64_t seed);
/* The parameters it derives: q [n_items], planted_off [n_planted + 1], planted_items
 * [n_planted * size_hi], planted_rate [n_planted]. */
int sgfim_zipf_spec(int32_t n_items, double zipf_s, double mean_len, int32_t n_planted, int32_t size_lo,
                    int32_t size_hi, double rate_lo, double rate_hi, uint64_t seed, double* q,
                    int32_t* planted_off, int32_t* planted_items, double* planted_rate);

/* Number of transactions and the distinct item ids of the loaded set
 * (ascending; the library's dense id of ids[k] is k). */
int sgfim_info(const sgfim_ctx* ctx, int64_t* n_txn, int32_t* n_items, int64_t* words);
int sgfim_item_ids(const sgfim_ctx* ctx, int32_t* ids /* n_items */);

/* Item supports (row popcounts), one per loaded item id in sgfim_item_ids order.
 * Replaces the frequency histogram of getItemsFrequenc (kernel_bitmap.cu:269-298). */
int sgfim_item_supports(sgfim_ctx* ctx, int64_t* out);

/* Pair co-occurrence matrix over all loaded items, out[i*n + j] = |T_i & T_j|
 * (n = n_items, dense order), computed with MFMA on 0/1 tiles (fp4 by default, see
 * sgfim_set_pair_kernel). Replaces the first BMMA wave over the single-item roots
 * (kernel_bmma.cu:772-872,1224-1236). */
int sgfim_pair_matrix(sgfim_ctx* ctx, int32_t* out);

/* Device duration in ms of the most recent MFMA pair-matrix computation of the
 * context (HIP events around its kernels; 0 if none yet). No reference counterpart:
 * the measurement of the pair path. */
int sgfim_pair_time(sgfim_ctx* ctx, double* ms);

/* The pair-matrix kernel of the context: SGFIM_PAIR_FP4 (default; e2m1 MFMA
 * v_mfma_scale_f32_32x32x64_f8f6f4 on 0/1 nibbles, fp32 partial sums of at most 2^20
 * transactions per split, so exact) or SGFIM_PAIR_INT8 (v_mfma_i32_32x32x32_i8 on 0/1
 * bytes, int32 sums). max_splits > 0 caps the transaction splits of the grid (never
 * below what the fp4 bound needs; 0 = sized to whole dispatch rounds). The matrix does
 * not depend on either; a change recomputes it on next use. No reference counterpart:
 * the reference has one BMMA kernel (kernel_bmma.cu:636-768). */
#define SGFIM_PAIR_FP4 0
#define SGFIM_PAIR_INT8 1
int sgfim_set_pair_kernel(sgfim_ctx* ctx, int32_t kind, int32_t max_splits);

/* The counting kernel of the context's searches: SGFIM_COUNT_VALU (default; 64-bit AND +
 * popcount per needed cell) or SGFIM_COUNT_F4 (the cells of 16 children x 32 listed
 * candidates as two v_mfma_scale_f32_16x16x128_f8f6f4 products of bit-phase fp4 operands
 * per 512 transactions, exact fp32 sums). Both count the same units with the same need
 * rules and give identical cell matrices, so results never depend on it. E_STATE while an
 * anytime search runs. Replaces the reference's per-node counter
 * (getResFrequencyCudaKernel, kernel_bitmap.cu:412-442) and its BMMA waves
 * (kernel_bmma.cu:636-768). */
#define SGFIM_COUNT_VALU 0
#define SGFIM_COUNT_F4 1
int sgfim_set_count_kernel(sgfim_ctx* ctx, int32_t kind);

/* Same matrix by the VALU AND+popcount kernel (cross-check path). */
int sgfim_pair_matrix_valu(sgfim_ctx* ctx, int32_t* out);

/* Batched AND+popcount. For prefix p (item ids prefix_ids[prefix_off[p] ..
 * prefix_off[p+1])) and each candidate id c in cand[cand_off[p] .. cand_off[p+1]),
 * out[j] = |T(prefix) & T(c)|. An empty prefix stands for all transactions.
 * Replaces getResFrequencyCuda (kernel_bitmap.cu:451-536) and getResFrequency
 * (kernel_bitmap.cu:390-408), batched over many prefixes. */
int sgfim_count_batch(sgfim_ctx* ctx, const int32_t* prefix_ids, const int64_t* prefix_off,
                      int32_t n_prefix, const int32_t* cand, const int64_t* cand_off,
                      int32_t* out);

/* sgfim_count_batch without any allocation: no counterpart in the reference. It is for
 * callers that cannot allocate device memory (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, results and errors as
 * sgfim_count_batch; the device scratch (node descriptors, candidate ids, counts, work
 * counter) is a static array of each device, 13.1 MB, set up when the context is created.
 * At most 16384 prefixes and 2^20 candidates present in the set per call (more:
 * SGFIM_E_INVAL). Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several contexts or threads) race and must be serialised by the caller. */
int sgfim_count_batch_noalloc(sgfim_ctx* ctx, const int32_t* prefix_ids, const int64_t* prefix_off,
                              int32_t n_prefix, const int32_t* cand, const int64_t* cand_off,
                              int32_t* out);

/* Exact MFI mining at one threshold; blocks until the search ends.
 * Replaces the engine run CoParaCG_<baseline|bmma>.exe <fimi> <rel_thr>
 * (kernel_bitmap.cu:796-910, kernel_bmma.cu:1179-1295) and one AnyFIM stage
 * (AnyTime kernel.cu:844-907). Without root flags the result is the exact MFI
 * family {X : sup(X) >= min_count, no single-item extension frequent}, the
 * family code/pipeline/skab_spike.py:84-135 returns. One search handles at most
 * 4096 items: with more loaded, it runs on a device copy of the items frequent at
 * its thresholds (the reference compacts to those items, kernel_bitmap.cu:692-703),
 * and more than 4096 frequent items is SGFIM_E_INVAL. */
int sgfim_mine(sgfim_ctx* ctx, const sgfim_mine_args* args, sgfim_result** out);

/* Anytime search: begin, advance by rounds, poll, cancel. MFIs are confirmed
 * final when they are reported. Replaces the AnyFIM stage loop
 * (AnyTime kernel.cu:844-915) with a threshold-free, priority-ordered frontier
 * that runs down to args->min_count (the support floor). sgfim_anytime_step
 * processes up to max_rounds rounds (0: to the end); with captured graphs it returns
 * with the next rounds already queued, so the device keeps searching while the caller
 * takes and handles the MFIs. Calls that reload data, change the scorer or the trace,
 * or read the pair matrix first process the queued rounds (their MFIs become pending);
 * after a cancel the next step waits for them without reporting their MFIs. When a step
 * finds the search done, a round still queued (queued before the frontier was seen empty)
 * takes nothing and is not waited for; the next call that needs the stream processes it. */
int sgfim_anytime_begin(sgfim_ctx* ctx, const sgfim_mine_args* args);
int sgfim_anytime_step(sgfim_ctx* ctx, int32_t max_rounds, sgfim_round_stats* stats);
int sgfim_anytime_poll(sgfim_ctx* ctx, sgfim_round_stats* stats);
int sgfim_anytime_cancel(sgfim_ctx* ctx);
/* Take the MFIs confirmed since the last call (or all, after the run ended). */
int sgfim_anytime_take(sgfim_ctx* ctx, sgfim_result** out);

/* Result access: itemsets as original ids (the root first, then the order the
 * search added them, as in the reference's Results lines, kernel_bitmap.cu:681-687),
 * off has n + 1 entries, support has n entries; round[i] is the anytime round in
 * which set i was confirmed (0 for sgfim_mine). */
int sgfim_result_get(const sgfim_result* r, const int32_t** items, const int64_t** off,
                     const int64_t** support, const int32_t** round, int64_t* n);
/* Window of each result set (0 for a single loaded set). */
int sgfim_result_windows(const sgfim_result* r, const int32_t** win);
/* Wall time (seconds since the run began) at which each set was confirmed. */
int sgfim_result_times(const sgfim_result* r, const double** t);
void sgfim_result_free(sgfim_result* r);

/* On-device copy bandwidth: `iters` copies of `bytes` between two buffers with a plain
 * 16-byte-load/store kernel, timed with HIP events; gbps = (read + written bytes) / s.
 * The reference number for the HBM roofline on this box, and the known-size kernel for
 * calibrating hardware byte counters. No reference counterpart. */
int sgfim_copy_probe(sgfim_ctx* ctx, int64_t bytes, int32_t iters, double* gbps);

/* Matrix-core throughput of this device in TOP/s (fp4 flops or int8 ops, 2 per
 * multiply-add): iters x 8 independent 32 x 32 MFMAs per wave on register operands, four
 * waves per SIMD, no memory traffic, timed with HIP events. kind: SGFIM_PAIR_FP4 or
 * SGFIM_PAIR_INT8. No reference counterpart: the sustained rate the pair kernels are
 * compared with beside the published peak. */
int sgfim_mfma_probe(sgfim_ctx* ctx, int32_t kind, int32_t iters, double* tops);

/* Cumulative search statistics of the last sgfim_mine / anytime run. */
int sgfim_last_stats(const sgfim_ctx* ctx, sgfim_round_stats* stats);

/* Device time (ms, HIP events) of each stage of the current or latest search, summed over
 * its rounds, for searches run with SGFIM_NO_GRAPH (one round at a time, every launch
 * bracketed); 0 otherwise. ms receives min(n, 8) values: selection and U lists, the
 * projected blocks' transaction lists, their rows built bit by bit, their rows built word by
 * word, need rows, counting, expansion, scoring and commit. A profile of the round's
 * kernels without an external profiler; no reference counterpart (the engines print phase
 * totals, kernel_bmma.cu:985-986). */
int sgfim_stage_times(const sgfim_ctx* ctx, double* ms, int32_t n);

/* Proposal network (24 -> 256 ReLU -> 1): load BF16 weights used by
 * SGFIM_ORDER_LEARNED. w1 is [24][256] row-major, b1 [256], w2 [256], b2 [1], all
 * float32 here and rounded to BF16 on the device. A later call overwrites the weights in
 * stream order: rounds already enqueued use the old weights, later ones the new. No
 * reference counterpart: the reference has no learned component. */
int sgfim_set_scorer(sgfim_ctx* ctx, const float* w1, const float* b1, const float* w2, float b2);

/* Search trace for training the proposal network. With tracing on, the next
 * runs record every frontier record born (its key, its parent record's key and its
 * 24 features) and, for every confirmed MFI, the key of the record whose round
 * confirmed it (-1 for MFIs found by the seed), up to cap records each. sgfim_trace_get
 * returns the counts, and copies the records when the output arrays are given
 * (keys: 2 per birth; feat: 24 per birth; cap = capacity of the outputs). */
int sgfim_set_trace(sgfim_ctx* ctx, int32_t enable, int64_t cap);
int sgfim_trace_get(sgfim_ctx* ctx, int64_t* n_birth, int64_t* n_mfi, int64_t* keys, float* feat,
                    int64_t* mfi_keys, int64_t cap);

/* ---- proposal-network trainer (no reference counterpart) ----
 * FP32 master weights and Adam state on one device; forward and backward run as
 * BF16-MFMA HIP kernels with FP32 accumulation. Parameter order of the packed
 * gradient: W1 [24][256] row-major, b1 [256], w2 [256], b2. */
typedef struct sgfim_mlp sgfim_mlp;
int sgfim_mlp_create(int32_t device, const float* W1, const float* b1, const float* w2, float b2,
                     sgfim_mlp** out);
void sgfim_mlp_destroy(sgfim_mlp* m);
const char* sgfim_mlp_last_error(const sgfim_mlp* m);
/* One step on host rows X [n][24], labels Y [n] (0/1): mean BCE loss and packed gradient
 * out (either may be null); an Adam update (beta 0.9/0.999, eps 1e-8) when lr > 0. */
int sgfim_mlp_step(sgfim_mlp* m, const float* X, const float* Y, int32_t n, float lr, float* loss,
                   float* grad);
/* Resident training set and minibatch Adam epochs over it. */
int sgfim_mlp_set_data(sgfim_mlp* m, const float* X, const float* Y, int64_t n);
int sgfim_mlp_train(sgfim_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* losses);
int sgfim_mlp_get(sgfim_mlp* m, float* W1, float* b1, float* w2, float* b2);
/* Logits (probs = 0) or sigmoid scores (probs = 1) of n rows with the current weights. */
int sgfim_mlp_forward(sgfim_mlp* m, const float* X, int32_t n, int32_t probs, float* out);

#ifdef __cplusplus
}
#endif
#endif /* SGFIM_H */
