"""Drop-in engine adapter: the reference pipeline's engine interface, backed by the GPU library.

Mirrors code/pipeline/engine_adapter.py (same function names, arguments, return
shapes and error convention) but runs the MFI search in-process through the
sgfim C ABI instead of spawning CoParaCG_<engine>.exe / AnytimeMining.exe:

  write_fimi / load_items_map / decode_itemset     engine_adapter.py:37-66
  run_tensorfim(engine, fimi, rel_thr, reps, timeout) engine_adapter.py:70-100
  parse_tensorfim_results                          engine_adapter.py:103-111
  run_anyfim(fimi, upto_stage, timeout)            engine_adapter.py:115-131
  parse_anyfim_results                             engine_adapter.py:134-166
  cumulative_mfis / support_confidence_lift        engine_adapter.py:169-191

Both run_* functions also write the engine's Results file next to the dataset,
in the engine's line grammar (kernel_bitmap.cu:651-690, AnyTime kernel.cu:615-659),
so scripts that read those files keep working. Engine semantics are kept:
TensorFIM roots need freq/n > rel_thr (kernel_bitmap.cu:816); an AnyFIM stage k
mines at the k-th highest item count with the first k items of the stable
frequency order as roots (AnyTime kernel.cu:844-882). Pass exact=True to get
the exact MFI family instead (the Python miner's semantics, skab_spike.py:84-135).
"""
from __future__ import annotations

import itertools
import json
import math
import os
import re
import time

import numpy as np

from . import _lib as L
from .core import Context

TF_LINE = re.compile(r"^\d+th:\s+\d+\s+\{\s*([^}]*)\}\s+support:\s+\S+\s+frequency:\s+(\d+)")
TF_COMP = re.compile(r"computation_time\(s\):\s*([\d.eE+-]+)")
TF_TOT = re.compile(r"total_time\(s\):\s*([\d.eE+-]+)")
AF_STAGE = re.compile(r"^=+ (\d+)th stage MFIs Number: (\d+) =+")

_CTX = None


def _ctx() -> Context:
    global _CTX
    if _CTX is None:
        _CTX = Context(device=int(os.environ.get("SGFIM_DEVICE", "0")))
    return _CTX


# ---------------------------------------------------------------------------- FIMI I/O
def write_fimi(txns, fimi_path, items_map_path=None):
    """Encode 'name=state' transactions as FIMI lines of 1-based ids (first-seen order),
    each line the sorted distinct ids; returns {item_str: id}. The id map is written as
    JSON (unescaped, indent 1) when a path is given, the form the pipeline reads back."""
    mapping = {}
    # setdefault sizes the map before inserting, so a new item gets the next id
    rows = [sorted({mapping.setdefault(it, len(mapping) + 1) for it in txn}) for txn in txns]
    text = "\n".join(" ".join(map(str, r)) for r in rows) + "\n"
    with open(fimi_path, "w", encoding="utf-8") as f:
        f.write(text)
    if items_map_path:
        with open(items_map_path, "w", encoding="utf-8") as f:
            json.dump(mapping, f, ensure_ascii=False, indent=1)
    return mapping


def load_items_map(path):
    with open(path, encoding="utf-8") as f:
        return json.load(f)


def decode_itemset(itemset_ids, inv_map):
    def one(i):
        return inv_map[i] if i in inv_map else inv_map[str(i)]
    return frozenset(one(i) for i in itemset_ids)


def read_fimi(path):
    """CSR (int64 offsets, int32 ids) of a FIMI file with the engines' parse rules
    (kernel_bitmap.cu:205-232): whitespace-separated integers, one transaction per
    line, a negative id and its digits are dropped, an empty line is an empty
    transaction, a last line without newline counts only if it holds an item."""
    data = open(path, "rb").read()
    off = [0]
    items: list[int] = []
    row: list[int] = []
    i, n = 0, len(data)
    while i < n:
        c = data[i]
        if c == 0x0A:
            items.extend(row)
            off.append(len(items))
            row = []
            i += 1
        elif 0x30 <= c <= 0x39:
            v = 0
            while i < n and 0x30 <= data[i] <= 0x39:
                v = v * 10 + (data[i] - 0x30)
                i += 1
            row.append(v)
        elif c == 0x2D:
            i += 1
            while i < n and 0x30 <= data[i] <= 0x39:
                i += 1
        else:
            i += 1
    if row:
        items.extend(row)
        off.append(len(items))
    return np.asarray(off, dtype=np.int64), np.asarray(items, dtype=np.int32)


def _csr_without_zero(off, items):
    """The engines count id 0 in prefix masks but never as a candidate (kernel_bitmap.cu:344-366,
    AnyTime kernel.cu:355-358: survivors are ids >= 1), so for every root other than 0 their
    output is the MFI family of the transactions without item 0; the ABI takes ids >= 1, and
    id 0 is dropped here. Returns (off, items, support of id 0)."""
    if items.size and (items == 0).any():
        keep = items != 0
        cs = np.concatenate([[0], np.cumsum(keep)])
        return cs[off], items[keep], int((~keep).sum())
    return off, items, 0


def _csr_shift(off, items):
    """Ids + 1, so id 0 is an ordinary item (the exact family: the Python miner treats 0 like
    any other id, skab_spike.py:84-135); results are shifted back."""
    return off, items + 1


# ---------------------------------------------------------------------------- Results files
def _fmt(x: float) -> str:
    """C++ iostream default formatting (6 significant digits, %g rules)."""
    s = "%.6g" % x
    return s


def _mfi_lines(sets, n_txn):
    order = sorted(range(len(sets)), key=lambda i: (-sets[i][1], sorted(sets[i][0])))
    lines = []
    for r, i in enumerate(order):
        items, freq = sets[i]
        ids = "".join(f"{x} " for x in items)
        lines.append(f"{r + 1}th: {len(items)} {{ {ids}}}   support: {_fmt(freq / n_txn)}  frequency: {freq}")
    return lines


def write_tensorfim_results(path, fimi_path, sets, n_txn, rel_thr, comp_s, pre_s, extra=None):
    """Results file of CoParaCG_<engine>.exe (kernel_bitmap.cu:651-690)."""
    bar = "=" * 66
    with open(path, "w", encoding="utf-8") as f:
        f.write(bar + "\n")
        f.write("GPU 0: AMD Instinct MI355X (gfx950)\n")
        f.write(bar + "\n")
        f.write(f"transSetFile: {fimi_path}\n")
        for k, v in (extra or {}).items():
            f.write(f"{k}: {v}\n")
        f.write(f"\nsupportThreshold: {_fmt(rel_thr)}\n\n")
        f.write(bar + "\n")
        f.write(f"preprocessing_time(s): {_fmt(pre_s)}\n")
        f.write(f"computation_time(s): {_fmt(comp_s)}\n")
        f.write(f"total_time(s): {_fmt(pre_s + comp_s)}\n")
        f.write(f"\nMFIsNumber: {len(sets)}\n")
        f.write("============================List of MFIs:=============================\n")
        for ln in _mfi_lines(sets, n_txn):
            f.write(ln + "\n")
        f.write(bar + "\n")


def write_anyfim_results(path, fimi_path, stages, thresholds, runtimes, n_txn, pre_s):
    """Staged Results file of AnytimeMining.exe (AnyTime kernel.cu:615-659)."""
    bar = "=" * 65
    K = len(stages)
    with open(path, "w", encoding="utf-8") as f:
        f.write(bar + "\n")
        f.write("GPU 0: AMD Instinct MI355X (gfx950)\n")
        f.write(bar + "\n")
        f.write(f"transSetFile: {fimi_path}\n")
        f.write(f"transNum: {n_txn}\n")
        f.write(f"\nUptoStage: {K}\n")
        f.write(f"preprocessing_time(s): {_fmt(pre_s)}\n")
        f.write(bar + "\n")
        f.write("runtimePerStage(s):" + "".join(f" {_fmt(v)}," for v in runtimes) + "\n")
        f.write("\nfrequencyThrePerStage:" + "".join(f" {t}," for t in thresholds))
        f.write("\nsupportThrePerStage:" + "".join(f" {_fmt(t / n_txn)}," for t in thresholds) + "\n")
        f.write("\ndimensionPerStage:" + "".join(f" {k}," for k in range(1, K + 1)) + "\n")
        nums = [len(stages[k]) for k in range(1, K + 1)]
        its = [sum(len(x) for x, _ in stages[k]) for k in range(1, K + 1)]
        f.write("\nMFIsNumPerStage:" + "".join(f" {v}," for v in nums))
        f.write("\nitemsNumPerStage:" + "".join(f" {v}," for v in its) + "\n")
        f.write("\nMFIsNumPerStage_withRemoveDuplicate:" + "".join(f" {v}," for v in nums))
        f.write("\nitemsNumPerStage_withRemoveDuplicate:" + "".join(f" {v}," for v in its) + "\n")
        for k in range(1, K + 1):
            f.write(f"====================== {k}th stage MFIs Number: {len(stages[k])} ======================\n")
            for ln in _mfi_lines(stages[k], n_txn):
                f.write(ln + "\n")
            f.write(f"\nfrequencyThreshold: {thresholds[k - 1]}\n")
        f.write(bar + "\n")


def _mfi_line(line):
    """(frozenset(ids), frequency) of one MFI line of a Results file, or None."""
    hit = TF_LINE.match(line)
    return (frozenset(map(int, hit.group(1).split())), int(hit.group(2))) if hit else None


def _lines(path):
    with open(path, encoding="utf-8", errors="replace") as f:
        for raw in f:
            yield raw.strip()


def parse_tensorfim_results(path):
    """{frozenset(item ids): frequency} of a TensorFIM Results file (a repeated set keeps
    its last frequency)."""
    return dict(filter(None, map(_mfi_line, _lines(path))))


# Per-stage lists in an AnyFIM Results file: the label before the first ':' and the type of
# the comma-separated values after it.
_AF_META = {"runtimePerStage(s)": float, "supportThrePerStage": float, "frequencyThrePerStage": float,
            "MFIsNumPerStage": int, "itemsNumPerStage": int, "dimensionPerStage": int,
            "preprocessing_time(s)": float}


def parse_anyfim_results(path):
    """(stages, meta): stages[k] = [(frozenset(ids), frequency)] in file order, meta[label]
    = the label's per-stage list. MFI lines before the first stage header are ignored."""
    stages, meta, cur = {}, {}, None
    for line in _lines(path):
        head = AF_STAGE.match(line)
        if head:
            cur = stages.setdefault(int(head.group(1)), [])
            cur.clear()
            continue
        label, colon, rest = line.partition(":")
        if colon and label in _AF_META:
            meta[label] = [_AF_META[label](v) for v in rest.replace(",", " ").split()]
        mfi = _mfi_line(line) if cur is not None else None
        if mfi:
            cur.append(mfi)
    return stages, meta


# ---------------------------------------------------------------------------- engines
def run_tensorfim(engine, fimi_path, rel_thr, reps=1, timeout=1800, exact=False):
    """engine: 'baseline' | 'bmma' (both map to the same exact GPU search). Returns
    {'ok', 'mfis': {frozenset: freq}, 'comp_times': [...], 'tot_times': [...]} and writes
    <fimi>-<rel_thr:.6f>=Results.txt.

    Failures follow the reference adapter (engine_adapter.py:82-99): a run that times out
    or produces no Results file sets ok=False and is skipped; nothing is raised. Here a
    library error of any kind (time-out, more frequent items or a longer itemset than the
    library supports, capacity, device) is such a run; its message goes under 'log'."""
    if engine not in ("baseline", "bmma"):
        raise ValueError(f"unknown engine {engine!r}")
    rp = f"{fimi_path}-{rel_thr:.6f}=Results.txt"
    comp, tot = [], []
    ok, mfis = True, {}
    log = []
    ctx = _ctx()   # no usable device or library: raised, as a missing engine executable is
    t_pre0 = time.perf_counter()
    try:
        off, items = read_fimi(fimi_path)
        zero = 0
        if exact:
            off, items = _csr_shift(off, items)
        else:
            off, items, zero = _csr_without_zero(off, items)
        ctx.load_csr(off, items)
    except (OSError, L.SgfimError) as e:
        # the engine cannot read its input and writes no Results file (kernel_bitmap.cu:801)
        if os.path.exists(rp):
            os.remove(rp)
        return {"ok": False, "mfis": {}, "comp_times": [], "tot_times": [], "log": str(e)}
    pre = time.perf_counter() - t_pre0
    n = ctx.n_txn
    thr = max(1, math.ceil(n * rel_thr)) if n else 1
    if zero and n and zero / n > rel_thr:
        # id 0 passes the root rule (kernel_bitmap.cu:816): the engine also mines a subtree
        # under root 0 whose output (kernel_bitmap.cu:866-889 with id 0 never a candidate)
        # is not reproduced here
        if os.path.exists(rp):
            os.remove(rp)
        return {"ok": False, "mfis": {}, "comp_times": [], "tot_times": [],
                "log": "item id 0 is a root of the engine's search; its subtree is not reproduced"}
    for _ in range(reps):
        if os.path.exists(rp):
            os.remove(rp)
        t0 = time.perf_counter()
        try:
            res = ctx.mine(thr, flags=0 if exact else L.ROOT_STRICT, rel_thr=rel_thr, deadline_s=float(timeout))
        except L.SgfimError as e:
            ok = False
            log.append(str(e))
            continue
        dt = time.perf_counter() - t0
        comp.append(dt)
        tot.append(dt + pre)
        sets = [tuple(x - 1 for x in it) for it in res.itemsets] if exact else res.itemsets
        mfis = {frozenset(x): s for x, s in zip(sets, res.supports)}
        write_tensorfim_results(rp, fimi_path, list(zip(sets, res.supports)), n, rel_thr, dt, pre,
                                {"transNum": n, "dimension": ctx.n_items})
    out = {"ok": ok, "mfis": mfis, "comp_times": comp, "tot_times": tot}
    if log:
        out["log"] = "\n".join(log)
    return out


def anyfim_thresholds(supports: np.ndarray, upto_stage: int):
    """Stage thresholds: the k-th highest item count, k = 1..upto_stage (kernel.cu:846)."""
    return [int(x) for x in np.sort(supports)[::-1][:upto_stage]]


def _anyfim_failed(rp, why):
    """The reference adapter's answer when the engine leaves no Results file
    (engine_adapter.py:123-129): ok=False, no stages, the engine's output under 'log'."""
    if os.path.exists(rp):
        os.remove(rp)
    return {"ok": False, "stages": {}, "meta": {}, "log": why, "results_path": rp}


def run_anyfim(fimi_path, upto_stage, timeout=3600, exact=False, batched=False):
    """Run stages 1..upto_stage; returns {'ok', 'stages': {k: [(frozenset, freq)]}, 'meta', 'results_path'}
    and writes <fimi>-<upto_stage>-stages=Results.txt.

    batched=True mines every stage in one search (sgfim_set_thresholds) instead of
    re-mining stage by stage as the engine does (kernel.cu:844-907); the stage sets are
    the same, and runtimePerStage then holds the time at which each stage's last MFI was
    confirmed in that shared run.

    Failures follow the reference adapter (engine_adapter.py:115-131): a time-out returns
    ok=False with empty stages, and so does any library error (its message under 'log',
    as the reference keeps the engine's stdout when the Results file is missing)."""
    rp = f"{fimi_path}-{upto_stage}-stages=Results.txt"
    if os.path.exists(rp):
        os.remove(rp)
    ctx = _ctx()   # no usable device or library: raised, as a missing engine executable is
    t_pre0 = time.perf_counter()
    try:
        off, items = read_fimi(fimi_path)
        zero = 0
        if exact:
            off, items = _csr_shift(off, items)
        else:
            off, items, zero = _csr_without_zero(off, items)
        ctx.load_csr(off, items)
    except (OSError, L.SgfimError) as e:
        return _anyfim_failed(rp, str(e))
    pre = time.perf_counter() - t_pre0
    if zero:
        # id 0 ranks in the engine's stable frequency order (kernel.cu:269-300): with it among
        # the first upto_stage items its stage thresholds and root sets differ from the ones
        # computed without it, and the engine's root-0 output is not reproduced
        s0 = ctx.item_supports()
        rank0 = int((s0 > zero).sum())          # items ranked before id 0 (ties: lower id first)
        if rank0 < upto_stage:
            return _anyfim_failed(rp, "item id 0 is among the engine's first stage roots; not reproduced")
    if upto_stage > ctx.n_items:
        # the engine refuses (kernel.cu:804-806) and writes no file
        out = _anyfim_failed(rp, "Up to the stage number being greater than the dimension !")
        out["refused"] = True
        return out
    sup = ctx.item_supports()
    thrs = anyfim_thresholds(sup, upto_stage)
    stages, runtimes = {}, []
    t_start = time.perf_counter()
    if batched:
        try:
            res = ctx.mine_thresholds(thrs, flags=0 if exact else L.ROOT_TOPK,
                                      root_topk=list(range(1, upto_stage + 1)))
        except L.SgfimError as e:
            return _anyfim_failed(rp, str(e))
        total = time.perf_counter() - t_start
        if total > timeout:
            return _anyfim_failed(rp, f"time-out after {total:.3f} s")
        for k in range(1, upto_stage + 1):
            stages[k] = []
        done_at = [0.0] * upto_stage
        for x, sp, w, t in zip(res.itemsets, res.supports, res.windows, res.times):
            stages[w + 1].append((tuple(i - 1 for i in x) if exact else x, sp))
            done_at[w] = max(done_at[w], t)
        runtimes = done_at
        write_anyfim_results(rp, fimi_path, stages, thrs, runtimes, ctx.n_txn, pre)
        st, meta = parse_anyfim_results(rp)
        return {"ok": True, "stages": st, "meta": meta, "results_path": rp}
    for k in range(1, upto_stage + 1):
        left = timeout - (time.perf_counter() - t_start)
        if left <= 0:
            return _anyfim_failed(rp, f"time-out before stage {k}")
        t0 = time.perf_counter()
        try:
            res = ctx.mine(thrs[k - 1], flags=0 if exact else L.ROOT_TOPK, root_topk=k, deadline_s=left)
        except L.SgfimError as e:
            return _anyfim_failed(rp, str(e))
        runtimes.append(time.perf_counter() - t0)
        sets = [tuple(i - 1 for i in x) for x in res.itemsets] if exact else res.itemsets
        stages[k] = list(zip(sets, res.supports))
    write_anyfim_results(rp, fimi_path, stages, thrs, runtimes, ctx.n_txn, pre)
    st, meta = parse_anyfim_results(rp)
    return {"ok": True, "stages": st, "meta": meta, "results_path": rp}


def cumulative_mfis(stages):
    """{k: set of every itemset seen in stages <= k}."""
    order = sorted(stages)
    per_stage = (frozenset(x for x, _ in stages[k]) for k in order)
    return {k: set(u) for k, u in zip(order, itertools.accumulate(per_stage, frozenset.union))}


def support_confidence_lift(itemset, txns, labels):
    """(support, P(X | label 1), lift = P(X | 1) / P(X)) over integer transactions."""
    n = len(txns)
    x = frozenset(itemset)
    covers = [x.issubset(t) for t in txns]   # X within transaction t
    cnt = sum(covers)
    npos = sum(lab == 1 for lab in labels)
    pos = sum(c and lab == 1 for c, lab in zip(covers, labels))
    sup = cnt / n if n else 0.0
    conf = pos / npos if npos else 0.0
    lift = (pos / npos) / (cnt / n) if cnt and npos else 0.0
    return sup, conf, lift


def support_confidence_lift_batch(itemsets, txns, labels):
    """support_confidence_lift for many patterns at once on the GPU: the label becomes one
    more item (present in the label-1 transactions) and every pattern takes two AND+popcount
    counts, |T(X)| (prefix X minus its last item, candidate that item) and |T(X) & T(label)|
    (prefix X, candidate the label item), in one count_batch call. Same formula and float
    results as support_confidence_lift."""
    n = len(txns)
    npos = sum(1 for lab in labels if lab == 1)
    label_id = 1 + max((max(t) for t in txns if len(t)), default=0)
    rows = [tuple(sorted(set(t))) + ((label_id,) if lab == 1 else ()) for t, lab in zip(txns, labels)]
    from .core import csr_from_rows
    off, items = csr_from_rows(rows)
    ctx = _ctx()
    ctx.load_csr(off, items)
    prefixes, cands = [], []
    for x in itemsets:
        xs = sorted(set(x))
        prefixes.append(xs[:-1] if xs else [])
        cands.append(xs[-1:] if xs else [])
        prefixes.append(xs)
        cands.append([label_id])
    got = ctx.count_batch(prefixes, cands) if itemsets else []
    out = []
    for i, x in enumerate(itemsets):
        cnt = int(got[2 * i][0]) if len(x) else n
        pos = int(got[2 * i + 1][0])
        sup = cnt / n if n else 0.0
        conf = pos / npos if npos else 0.0
        lift = (pos / npos) / (cnt / n) if cnt and npos else 0.0
        out.append((sup, conf, lift))
    return out

