"""Autotuner of the search's counting tile.

The counting kernel stages the 16 child prefixes of one row chunk in LDS and streams the
candidate rows through registers. Two knobs shape it and neither changes a result:
  chunk_words  row words per chunk: words per lane = chunk_words / 64 (candidate words a
               lane holds per step), LDS prefix size = 16 x 8 x chunk_words bytes (128 ->
               16 KB and four workgroups per CU, 256 -> 32 KB and three, 512 -> 64 KB and one);
  chunk_range  chunks per work unit: the unit's set-up (need masks, the sorted candidate
               list -- the candidates a workgroup takes) is paid once per range, and a
               range of every record is counted before the next (its rows stay in L2).
A third knob is the projection factor (sgfim_set_projection): a record whose support is at
most 1/factor of its rows' transactions gets its U rows compacted to its transactions, and
its subtree counts on them. Deep lattices over long rows gain (workload 5), shallow ones
pay the projection without reusing it (workload 2), so it is tuned per input shape too, and
so is the block rule (sgfim_set_block_rule): which blocks are compacted word by word and
how many children such a block needs to repay its build (deep subtrees reuse a block
many levels down, shallow ones only in the record's own round).
tune() times a caller's search for every projection factor on the loaded data, then every
block rule with the chosen factor, then every candidate shape, checks that every setting
returns the same MFIs, keeps the fastest, and records it in a JSON cache keyed by device
and input shape; apply() sets a cached setting. `python -m sgfim.autotune` tunes the benchmark workloads and writes
tile_cache.json next to this file.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

CANDIDATES = ((128, 2), (128, 4), (128, 8), (256, 2), (256, 4), (256, 8), (512, 1), (512, 2), (512, 4))
DEFAULT = (256, 4)
PROJ_CANDIDATES = (0, 4, 8, 16, 32)
PROJ_DEFAULT = 4   # the library's default factor
RULE_CANDIDATES = ((32, 16), (32, 4), (32, 2), (8, 32), (8, 24), (8, 16), (8, 2), (0, 2))   # (dense, min_children)
RULE_DEFAULT = (32, 16)   # the library's default block rule
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_cache.json")


def _bucket(x: int) -> int:
    """Next power of two (shapes within a factor of two share an entry)."""
    return 1 << max(0, math.ceil(math.log2(max(1, int(x)))))


def shape_key(device: str, n_cu: int, n_items: int, words: int, n_sets: int) -> str:
    return f"{device}:{n_cu}cu|items:{_bucket(n_items)}|words:{_bucket(words)}|sets:{_bucket(n_sets)}"


def context_key(ctx, n_sets: int = 1) -> str:
    dev, ncu = ctx.device_info()
    return shape_key(dev, ncu, ctx.n_items, ctx.words, n_sets)


def load_cache(path: str = CACHE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_entry(key: str, entry: dict, path: str = CACHE):
    cache = load_cache(path)
    cache[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def lookup(key: str, path: str = CACHE):
    """(chunk_words, chunk_range) cached for key, or None."""
    e = load_cache(path).get(key)
    return (int(e["chunk_words"]), int(e["chunk_range"])) if e else None


def lookup_projection(key: str, path: str = CACHE):
    """Projection factor cached for key, or None."""
    e = load_cache(path).get(key)
    return int(e["proj_factor"]) if e and "proj_factor" in e else None


def lookup_rule(key: str, path: str = CACHE):
    """(dense, min_children) cached for key, or None."""
    e = load_cache(path).get(key)
    return (int(e["proj_dense"]), int(e["proj_mintail"])) if e and "proj_dense" in e else None


def apply(ctx, n_sets: int = 1, path: str = CACHE):
    """Set the cached tile shape, projection factor and block rule for this context's
    device and data; returns the shape (or the default). Without a cached value the
    library default stays."""
    key = context_key(ctx, n_sets)
    t = lookup(key, path)
    ctx.set_tile(*(t or (0, 0)))
    f = lookup_projection(key, path)
    ctx.set_projection(PROJ_DEFAULT if f is None else f)
    ctx.set_block_rule(*(lookup_rule(key, path) or RULE_DEFAULT))
    return t or DEFAULT


def choose(timings: dict):
    """The fastest shape; ties go to the earlier candidate."""
    best = None
    for shape, ms in timings.items():
        if best is None or ms < timings[best]:
            best = shape
    return best


def _mix(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 (wrapping arithmetic)."""
    z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def signature(res) -> tuple:
    """Order-free digest of a result: the number of MFIs and a wrapping sum over MFIs of a
    hash of (set of items, support, set index)."""
    if res.arrays:
        items, off, sup = (np.asarray(a) for a in res.arrays)
    else:
        lens = [len(x) for x in res.itemsets]
        items = np.asarray([i for x in res.itemsets for i in x], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        sup = np.asarray(res.supports, dtype=np.int64)
    n = len(sup)
    if n == 0:
        return (0, 0)
    win = np.asarray(res.windows, dtype=np.int64) if len(res.windows) else np.zeros(n, dtype=np.int64)
    with np.errstate(over="ignore"):
        per_item = _mix(items.astype(np.int64))
        csum = np.concatenate([[np.uint64(0)], np.cumsum(per_item, dtype=np.uint64)])
        set_h = csum[off[1:]] - csum[off[:-1]]
        h = _mix(set_h ^ _mix(sup.astype(np.uint64) * np.uint64(0x100000001B3) + win.astype(np.uint64)))
        return (n, int(np.sum(h, dtype=np.uint64)))


def _time(run, reps: int) -> float:
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    return 1e3 * min(ts)


def tune(ctx, run, candidates=CANDIDATES, reps: int = 3, key: str | None = None, path: str | None = CACHE,
         verify: bool = True, log=None, proj_candidates=PROJ_CANDIDATES, rule_candidates=RULE_CANDIDATES) -> dict:
    """Time run() (a search on ctx returning a MineResult) under every projection factor with
    the default tile shape and block rule, keep the fastest factor, then every block rule
    with that factor (when blocks are on), then every candidate tile shape, keeping the
    fastest each time (the best of reps wall times after one warm-up run, which also
    captures the round graphs). With verify, every setting must give
    the same MFIs. The chosen setting is left on ctx and, when key and path are given, stored
    in the cache."""
    ref = None

    def warm_check(what):
        nonlocal ref
        r = run()
        if verify:
            sig = signature(r)
            if ref is None:
                ref = sig
            elif sig != ref:
                raise RuntimeError(f"{what} changed the result")

    ctx.set_tile(*DEFAULT)
    ctx.set_block_rule(*RULE_DEFAULT)
    ptimes = {}
    for f in proj_candidates:
        ctx.set_projection(f)
        warm_check(f"projection factor {f}")
        ptimes[f] = _time(run, reps)
        if log:
            log(f"projection factor {f}: {ptimes[f]:.3f} ms")
    pf = min(ptimes, key=lambda f: (ptimes[f], f != PROJ_DEFAULT))
    ctx.set_projection(pf)
    rtimes = {RULE_DEFAULT: ptimes[pf]}
    if pf > 0:
        for rule in rule_candidates:
            if rule == RULE_DEFAULT:
                continue
            ctx.set_block_rule(*rule)
            warm_check(f"block rule {rule}")
            rtimes[rule] = _time(run, reps)
            if log:
                log(f"block rule dense={rule[0]} min_children={rule[1]}: {rtimes[rule]:.3f} ms")
    rule = min(rtimes, key=lambda r: (rtimes[r], r != RULE_DEFAULT))
    ctx.set_block_rule(*rule)
    timings = {DEFAULT: rtimes[rule]}
    for cw, cr in candidates:
        if (cw, cr) == DEFAULT:
            continue
        ctx.set_tile(cw, cr)
        warm_check(f"tile {cw}x{cr}")
        timings[(cw, cr)] = _time(run, reps)
        if log:
            log(f"tile chunk_words={cw} chunk_range={cr}: {timings[(cw, cr)]:.3f} ms")
    best = choose({sh: timings[sh] for sh in candidates if sh in timings})
    ctx.set_tile(*best)
    entry = {"chunk_words": best[0], "chunk_range": best[1], "proj_factor": pf,
             "proj_dense": rule[0], "proj_mintail": rule[1], "ms": round(timings[best], 4),
             "default_ms": round(ptimes.get(PROJ_DEFAULT, float("nan")), 4),
             "all": {f"{cw}x{cr}": round(ms, 4) for (cw, cr), ms in timings.items()},
             "proj_all": {str(f): round(ms, 4) for f, ms in sorted(ptimes.items())},
             "rule_all": {f"{d}/{m}": round(ms, 4) for (d, m), ms in sorted(rtimes.items())}}
    if key and path:
        save_entry(key, entry, path)
    return entry


def main(argv=None):
    """Tune the benchmark's workload 2 (three supports in one search), workload 3 (512
    windows), workload 4 (rounds of up to 4096 records) and workload 5 (the whole search to
    the 0.5% floor in the learned order) on device 0 and write the cache."""
    import argparse
    from . import gen
    from .core import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="2,3")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cache", default=CACHE)
    a = ap.parse_args(argv)
    say = lambda m: print(m, flush=True)
    for w in a.workloads.split(","):
        ctx = Context(device=0, batch_nodes=4096 if w == "4" else 0, frontier_cap=(1 << 23) if w == "5" else 0)
        if w == "2":
            T = gen.workload2(0)
            ctx.load_csr(T.offsets, T.items)
            thrs = [gen.min_count(s, T.n_txn) for s in (0.10, 0.05, 0.02)]
            run = lambda: ctx.mine_thresholds(thrs, convert=False)
            n_sets = len(thrs)
        elif w == "3":
            win_off, off, items, lens = gen.workload3_csr(0)
            ctx.load_windows(win_off, off, items)
            thrs = [gen.min_count(0.03, int(n)) for n in lens]
            run = lambda: ctx.mine_windows(thrs, convert=False)
            n_sets = len(lens)
        elif w == "4":   # anytime rounds of up to 4096 records, as the coverage evaluation runs them
            T = gen.workload4(1_000_000)
            ctx.load_csr(T.offsets, T.items)
            thr = gen.min_count(0.02, T.n_txn)
            run = lambda: ctx.mine(thr, convert=False)
            n_sets = 1
        elif w == "5":   # the dense-sparse Zipf set to its floor, learned frontier order
            from . import learned
            from ._lib import ORDER_LEARNED
            ctx.load_zipf(seed=0, **gen.WORKLOAD5)
            ctx.set_scorer(*learned.load_weights())
            thr = gen.min_count(0.005, ctx.n_txn)
            run = lambda: ctx.mine(thr, flags=ORDER_LEARNED, convert=False)
            n_sets = 1
        else:
            raise SystemExit(f"unknown workload {w}")
        key = context_key(ctx, n_sets)
        say(f"workload {w}: {key}")
        e = tune(ctx, run, reps=a.reps, key=key, path=a.cache, log=say)
        say(f"workload {w}: chose {e['chunk_words']}x{e['chunk_range']} ({e['ms']} ms; default {e['default_ms']} ms)")
        ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
