"""Benchmark of the MFI support-counting hot path on one MI355X.

Workload (workload 2 of the benchmark set): 1,000,000 transactions from a seeded
64-sensor x 8-bin regime-switching AR(1) stream (512 items, one item per sensor),
held as vertical uint64 bitmaps in HBM. One step = the int8-MFMA pair matrix plus
exact MFI mining at min supports 10%, 5% and 2% (every support counted exactly).

Prints one JSON line: whole-job support counts/s (value), the counting kernel's
roofline against MI355X HBM, and a CPU baseline (the C restatement of the
reference engine's CPU search, timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "20261008-self-growing-evidence-agent_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA
# 32-bit VALU lane-ops/s: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (a wave64 instruction
# issues over four cycles)
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12
# HBM bytes per k_count_group dispatch for this command, from the two rocprofv3 --pmc
# passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, mean over the run), scaled by
# what the same counters read for the 2 GiB copy probe (FETCH x2.0, WRITE x1.0);
# profiles/k_count_group_workload2_pmc.md holds the commands and the table.
PMC_COUNT_KERNEL = {"fetch_kib": 116982.0, "write_kib": 61325.3, "fetch_scale": 2.0, "write_scale": 1.0}
SUPPORTS = (0.10, 0.05, 0.02)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(T, min_s: float = 10.0):
    """The oracle's restatement of the reference CPU search (count every survivor per
    node, kernel_bitmap.cu:390-408 / :853-894) on all host cores, over the step's own
    workload: workload 2 at min supports 10%, 5% and 2%, repeated whole until at least
    min_s seconds have passed (about 2.5 s per repetition on 16 cores)."""
    import numpy as np
    from oracle import cref
    from sgfim import gen
    cores = len(os.sched_getaffinity(0))
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    plan = []
    for s in SUPPORTS:
        thr = gen.min_count(s, T.n_txn)
        plan.append((thr, rows[np.where(sup >= thr)[0]]))
    counts = nodes = mfis = reps = 0
    t0 = time.perf_counter()
    while reps == 0 or time.perf_counter() - t0 < min_s:
        for thr, r in plan:
            out, st = cref.mine(r, thr, count_mode=cref.COUNT_ALL, threads=cores)
            counts += st["counts"]
            nodes += st["nodes"]
            mfis += len(out)
        reps += 1
    dt = time.perf_counter() - t0
    return {"value": counts / dt, "unit": "counts/s", "cores": cores, "kind": "port",
            "sample": f"workload 2 ({T.n_txn} txns, {T.n_items} items) at min supports 10%/5%/2%, "
                      f"{reps} repetitions: {nodes} nodes, {counts} counts in {dt:.2f} s "
                      f"(count-all reference search, {mfis // reps} MFIs per repetition)"}


def workload3_run(Context, seed=0, workers=16):
    """512 windows, lengths log-uniform in [2,000, 200,000], 32 sensors x 6 bins, 3% per window:
    one search over all windows (every counting launch mixes windows)."""
    from sgfim import gen
    t0 = time.perf_counter()
    win_off, off, items, lens = gen.workload3_csr(seed, 512, workers)
    log(f"workload 3 generated: {len(lens)} windows, {int(win_off[-1])} txns in {time.perf_counter() - t0:.1f} s")
    ctx = Context(device=0)
    ctx.load_windows(win_off, off, items)
    from sgfim import autotune
    tile = autotune.apply(ctx, len(lens))
    thr = [gen.min_count(0.03, int(n)) for n in lens]
    ctx.mine_windows(thr, convert=False)                   # warm
    t0 = time.perf_counter()
    r = ctx.mine_windows(thr, convert=False)
    dt = time.perf_counter() - t0
    st = r.stats
    ctx.close()
    counts = st["counts"] + st["pair_counts"]
    return {"workload": "workload 3: 512 windows, 2k-200k txns, 32 sensors x 6 bins, 3% per window",
            "seconds": dt, "mfis": len(r), "counts": counts, "counts_per_s": counts / dt,
            "count_kernel_ms": st["count_ms"], "launches": st["count_launches"],
            "alg_gbs": 8.0 * st["words"] / (st["count_ms"] * 1e-3) / 1e9 if st["count_ms"] > 0 else None,
            "tile": {"chunk_words": tile[0], "chunk_range": tile[1]}}


def workload5_run(Context, budget_s=8.0, seed=0, learned_order=False):
    """10M txns x 2,048 Zipf items (2.56 GB of bitmaps, 10x the MALL): a budgeted anytime run
    down to the 0.5% floor; reports the counting kernel against the HBM roofline. With
    learned_order the proposal network orders the frontier."""
    from sgfim import gen, learned
    from sgfim._lib import ORDER_LEARNED
    ctx = Context(device=0, frontier_cap=1 << 23)   # 8M frontier records (~5 GB at 2,048 items)
    t0 = time.perf_counter()
    ctx.load_zipf(seed=seed, **gen.WORKLOAD5)
    log(f"workload 5 generated on the GPU in {time.perf_counter() - t0:.2f} s (W = {ctx.words})")
    if learned_order:
        ctx.set_scorer(*learned.load_weights())
    thr = gen.min_count(0.005, ctx.n_txn)
    t0 = time.perf_counter()
    ctx.anytime_begin(thr, ORDER_LEARNED if learned_order else 0)
    mfis = 0
    tag = "learned" if learned_order else "plain"
    while True:
        st = ctx.anytime_step(1)
        mfis += len(ctx.anytime_take(convert=False))
        log(f"workload 5 ({tag}) round {st['round']}: {st['counts']} counts, {mfis} MFIs, frontier {st['frontier']}")
        if st["done"] or time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    ctx.close()
    alg = 8.0 * st["words"] / (st["count_ms"] * 1e-3) / 1e9 if st["count_ms"] > 0 else None
    return {"workload": "workload 5: 10M txns x 2048 Zipf(1.1) items, mean length 60, 40 planted sets, "
                        f"floor 0.5%, budgeted anytime run, {tag} frontier order",
            "seconds": dt, "rounds": st["round"], "done": bool(st["done"]), "mfis_confirmed": mfis,
            "counts": st["counts"] + st["pair_counts"], "counts_per_s": (st["counts"] + st["pair_counts"]) / dt,
            "count_kernel_ms": st["count_ms"], "launches": st["count_launches"],
            "alg_gbs": alg, "alg_frac_of_hbm_peak": alg / HBM_PEAK_GBS if alg else None}


def workload4_run(Context, n_seeds=4, batch_nodes=4096):
    """Anytime coverage on workload 4 (500,000 txns x 256 items, 2% floor): time until 25%,
    50% and 95% of the final MFI set is confirmed, learned frontier order against the plain
    support order, on the first held-out seeds of the proposal network's evaluation (the
    committed weights). Each stream gets an untimed run first, and the timed runs alternate
    which order goes first."""
    from sgfim import learned
    from sgfim.train_proposal import anytime_curve, compare_orders, streams, t_at
    from sgfim import gen
    from sgfim import autotune
    ctx = Context(device=0, batch_nodes=batch_nodes)
    ctx.set_scorer(*learned.load_weights())
    tile = None
    seeds = list(range(1_000_000, 1_000_000 + n_seeds))
    rows, warm = [], True
    for k, (s, (off, items)) in enumerate(streams(seeds, 500_000, min(16, len(seeds)))):
        ctx.load_csr(off, items)
        tile = tile or autotune.apply(ctx, 1)
        thr = gen.min_count(0.02, ctx.n_txn)
        if warm:   # warm the captured rounds of the learned order
            anytime_curve(ctx, thr, True)
            warm = False
        rp = fp = rl = fl = None
        rp, fp, rl, fl = compare_orders(ctx, thr, plain_first=k % 2 == 0)
        tot = len(fp)
        rows.append({"seed": s, "mfis": tot, "identical": fp == fl, "plain_first": k % 2 == 0,
                     "t95_plain_s": t_at(rp, 0.95, tot, 0), "t95_learned_s": t_at(rl, 0.95, tot, 0),
                     "counts95_plain": t_at(rp, 0.95, tot, 1), "counts95_learned": t_at(rl, 0.95, tot, 1),
                     "t25_plain_s": t_at(rp, 0.25, tot, 0), "t25_learned_s": t_at(rl, 0.25, tot, 0),
                     "t50_plain_s": t_at(rp, 0.5, tot, 0), "t50_learned_s": t_at(rl, 0.5, tot, 0),
                     "counts50_plain": t_at(rp, 0.5, tot, 1), "counts50_learned": t_at(rl, 0.5, tot, 1)})
        log(f"workload 4 seed {s}: t95 plain {rows[-1]['t95_plain_s']:.4f} s, learned {rows[-1]['t95_learned_s']:.4f} s")
    ctx.close()
    med = lambda v: float(np.median(v))
    return {"workload": f"workload 4: 500,000 txns x 256 items, 2% floor, held-out seeds {seeds[0]}..{seeds[-1]}, "
                        f"anytime rounds of up to {batch_nodes} records",
            "metric": "time to 25%, 50% and 95% MFI coverage, learned / plain",
            "median_t95_ratio": med([r["t95_learned_s"] / r["t95_plain_s"] for r in rows]),
            "median_counts95_ratio": med([r["counts95_learned"] / max(r["counts95_plain"], 1) for r in rows]),
            "median_t25_ratio": med([r["t25_learned_s"] / r["t25_plain_s"] for r in rows]),
            "median_t50_ratio": med([r["t50_learned_s"] / r["t50_plain_s"] for r in rows]),
            "median_counts50_ratio": med([r["counts50_learned"] / max(r["counts50_plain"], 1) for r in rows]),
            "median_t95_plain_s": med([r["t95_plain_s"] for r in rows]),
            "median_t95_learned_s": med([r["t95_learned_s"] for r in rows]),
            "all_identical": all(r["identical"] for r in rows),
            "tile": {"chunk_words": tile[0], "chunk_range": tile[1]}, "seeds": rows}


def cpu_extra(T2=None, budget_s=15.0):
    """Single-thread restatements of the reference's Python miners: the set-based miner
    (skab_spike.py:84-135) and the NumPy packed-bitmap miner (exp_e6_cpuvec.py:43-107) on
    workload 1 (10k txns, 2%), and the NumPy miner on workload 2 at 10% under a budget."""
    from oracle import cpuvec as V
    from oracle import mfi_oracle as O
    from sgfim import gen
    T = gen.workload1(0)
    rows = T.rows()
    t0 = time.perf_counter()
    out, _ = O.mine_mfi(rows, min_count=gen.min_count(0.02, T.n_txn), node_cap=0)
    res = {"pure_python_mine_mfi_workload1_2pct_s": time.perf_counter() - t0, "mfis": len(out), "cores": 1}
    bm, n = V.bitmap_from_rows(rows, T.n_items)
    mf, dt = V.mine_mfi_cpuvec(bm, n, 0.02)
    res["numpy_cpuvec_workload1_2pct_s"] = dt
    res["numpy_cpuvec_workload1_mfis"] = len(mf)
    if T2 is not None:
        bm2, n2 = V.bitmap_from_rows(T2.rows(), T2.n_items)
        try:
            mf2, dt2 = V.mine_mfi_cpuvec(bm2, n2, 0.10, budget_s=budget_s)
            res["numpy_cpuvec_workload2_10pct_s"] = dt2
            res["numpy_cpuvec_workload2_10pct_mfis"] = len(mf2)
        except TimeoutError:
            res["numpy_cpuvec_workload2_10pct_s"] = f">{budget_s} (budget)"
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-txn", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the workload 3, 4 and 5 measurements")
    ap.add_argument("--w5-budget", type=float, default=8.0)
    ap.add_argument("--sequential", action="store_true",
                    help="mine the three supports as three searches instead of one batched search")
    ap.add_argument("--extra-only", choices=("workload3", "workload4", "workload5", "workload5_learned"), default=None,
                    help="run one extra workload alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              f"described, not built, in this repository", file=sys.stderr)
        sys.exit(2)

    from sgfim import gen
    from sgfim.core import Context

    if args.extra_only:
        res = {"workload3": lambda: workload3_run(Context), "workload4": lambda: workload4_run(Context),
               "workload5": lambda: workload5_run(Context, args.w5_budget),
               "workload5_learned": lambda: workload5_run(Context, args.w5_budget, learned_order=True)}[args.extra_only]()
        print(json.dumps(res))
        return

    t0 = time.perf_counter()
    T = gen.workload2(args.seed, args.n_txn)
    log(f"workload 2 generated: {T.n_txn} txns, {T.n_items} items in {time.perf_counter() - t0:.1f} s")
    ctx = Context(device=0)
    t0 = time.perf_counter()
    ctx.load_csr(T.offsets, T.items)
    log(f"loaded + transposed on GPU in {time.perf_counter() - t0:.2f} s (W = {ctx.words} words)")
    thrs = [gen.min_count(s, T.n_txn) for s in SUPPORTS]
    from sgfim import autotune
    tile_key = autotune.context_key(ctx, 1 if args.sequential else len(thrs))
    tile = autotune.apply(ctx, 1 if args.sequential else len(thrs))
    tile_src = "tile_cache.json" if autotune.lookup(tile_key) else "default"
    log(f"counting tile {tile[0]} words x {tile[1]} chunks ({tile_src}, {tile_key})")

    pair_ms = []

    from sgfim._lib import FRESH_PAIR

    def step():
        # the int8 MFMA pair matrix is recomputed on the device at the start of every step,
        # in the search's stream order (it seeds the search; no host read-back)
        counts = ctx.n_items * ctx.n_items
        words = 0
        cms = 0.0
        launches = 0
        mfis = []
        if args.sequential:                    # one search per support
            for k, thr in enumerate(thrs):
                r = ctx.mine(thr, flags=FRESH_PAIR if k == 0 else 0, convert=False)
                st = r.stats
                counts += st["counts"]
                words += st["words"]
                cms += st["count_ms"]
                launches += st["count_launches"]
                mfis.append(len(r))
        else:                                  # the three supports in one batched search
            r = ctx.mine_thresholds(thrs, flags=FRESH_PAIR, convert=False)
            st = r.stats
            counts += st["counts"]
            words += st["words"]
            cms += st["count_ms"]
            launches += st["count_launches"]
            tags = r.windows if len(r) else np.zeros(0, dtype=np.int64)
            mfis = np.bincount(np.asarray(tags, dtype=np.int64), minlength=len(thrs)).tolist()
        pair_ms.append(ctx.pair_time())
        return counts, words, cms, launches, mfis

    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    pair_ms.clear()
    tot_counts = tot_words = tot_launch = 0
    tot_cms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        c, w, cms, nl, mfis = step()
        tot_counts += c
        tot_words += w
        tot_cms += cms
        tot_launch += nl
        log(f"step {i + 1}/{args.steps}: {c} counts, MFIs per support {mfis}")
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    # cross-check of the counting kernel's duration with HIP events on its stream
    # (one direct-launch run at the lowest support, outside the timed steps)
    from sgfim._lib import NO_GRAPH
    rx = ctx.mine(thrs[-1], flags=NO_GRAPH).stats
    event_check = {"device_clock_ms": rx["count_ms"], "hip_event_ms": rx["event_ms"],
                   "launches": rx["count_launches"]}
    copy_gbs = ctx.copy_probe(1 << 31, 5)   # on-box HBM copy rate, 2 GiB buffers (8x the MALL)
    value = tot_counts / dt
    pair_avg = sum(pair_ms) / len(pair_ms) if pair_ms else 0.0
    pair_ops = 2.0 * ctx.n_items * ctx.n_items * T.n_txn
    pair_tops = pair_ops / (pair_avg * 1e-3) / 1e12 if pair_avg > 0 else None
    avg_ms = tot_cms / max(tot_launch, 1)
    bytes_per_launch = 8.0 * tot_words / max(tot_launch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # the AND + popcount a count needs: per 64-bit word two 32-bit ANDs and two popcounts
    valu_achieved = 4.0 * tot_words / (tot_cms * 1e-3) / 1e12 if tot_cms > 0 else 0.0
    out = {
        "metric": "support counts/s (exact MFI mining, workload 2 at 10%/5%/2%)",
        "value": value,
        "unit": "counts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded regime-switching AR(1) sensor stream)",
        "config": {"workload": f"workload 2: {T.n_txn} txns, 64 sensors x 8 bins = {T.n_items} items, "
                               f"min supports 10%/5%/2%", "mfis_per_support": mfis,
                   "search": "three searches, one per support" if args.sequential
                             else "one batched search over the three supports",
                   "tile": {"chunk_words": tile[0], "chunk_range": tile[1], "source": tile_src}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": 1024.0 * (PMC_COUNT_KERNEL["fetch_scale"] * PMC_COUNT_KERNEL["fetch_kib"] +
                                          PMC_COUNT_KERNEL["write_scale"] * PMC_COUNT_KERNEL["write_kib"]),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this command, "
                                       "calibrated by the copy probe (profiles/k_count_group_workload2_pmc.md)",
                     "kernel": "k_count_group", "avg_launch_ms": avg_ms,
                     "alg_bytes_per_launch": bytes_per_launch, "timing_check": event_check,
                     "copy_probe_gbs": copy_gbs,
                     # the bound that holds: candidate rows are reread from L2 and the MALL, so
                     # the algorithmic bytes exceed the HBM peak; the kernel is bound by VALU
                     # issue, and this is the share of it spent on the AND + popcount itself
                     "valu": {"achieved": valu_achieved, "peak": VALU_PEAK_TLANE, "unit": "T lane-op/s",
                              "frac": valu_achieved / VALU_PEAK_TLANE,
                              "counted": "4 lane-ops (2 AND, 2 popcount) per 64-bit word of every count"}},
        # the pair matrix (part of every step): 2 d^2 N int8 ops (the full square) per launch
        "pair_roofline": {"bound": "mfma", "kernel": "k_pair_tile", "avg_launch_ms": pair_avg,
                          "achieved": pair_tops, "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                          "frac": pair_tops / INT8_PEAK_TOPS if pair_tops else None},
    }
    ctx.close()
    if not args.no_extra:
        out["extra"] = {"workload5": workload5_run(Context, args.w5_budget),
                        "workload5_learned": workload5_run(Context, args.w5_budget, learned_order=True),
                        "workload3": workload3_run(Context),
                        "workload4": workload4_run(Context)}
    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(T)
        out.setdefault("extra", {})["cpu_python"] = cpu_extra(T)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
