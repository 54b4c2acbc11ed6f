"""NumPy restatement of the reference's CPU-vectorised MFI miner -- TEST INFRASTRUCTURE ONLY.

Only tests and the benchmark's CPU-baseline leg import this module; the product path
never does.

Restates code/pipeline/exp_e6_cpuvec.py:43-107 (single thread, NumPy):
  * load_fimi_bitmap: one packed vertical bitmap per item id, np.packbits along the
    transactions (MSB-first within a byte); blank lines are skipped, ids are 1-based
    and the bitmap row of id x is x - 1                         (exp_e6_cpuvec.py:46-56)
  * popcount by a 256-entry lookup table over the packed bytes  (:43, :63, :81)
  * frequent items (count >= rel * n, compared as floats) sorted by descending
    support, ties in ascending id (stable)                      (:62-67)
  * depth-first search: a node ANDs its bitmap with every frequent item at once; it
    is a maximal itemset when no item outside it is frequent with it, else it
    recurses into the frequent items after its own position; a node whose itemset
    plus every later item is covered by an MFI found earlier is skipped (:70-101)
It returns (mfis, seconds) with mfis a list of frozensets of 0-based item rows; with a
budget it raises TimeoutError once the budget has passed (checked at every node).
"""
from __future__ import annotations

import time

import numpy as np

LUT = np.array([bin(i).count("1") for i in range(256)], dtype=np.int32)


def load_fimi_bitmap(lines):
    """Packed bitmaps from FIMI text lines: (bm uint8[n_items][ceil(n/8)], n_txns)."""
    rows = [ln.split() for ln in lines if ln.strip()]
    n = len(rows)
    n_items = max((int(x) for r in rows for x in r), default=0)
    cols = np.zeros((n_items, n), dtype=np.uint8)
    for t, r in enumerate(rows):
        for x in r:
            cols[int(x) - 1, t] = 1
    return np.packbits(cols, axis=1), n


def bitmap_from_rows(rows, n_items: int):
    """The same packed layout from in-memory transactions (tuples of 1-based ids)."""
    n = len(rows)
    cols = np.zeros((n_items, n), dtype=np.uint8)
    for t, r in enumerate(rows):
        for x in r:
            cols[x - 1, t] = 1
    return np.packbits(cols, axis=1), n


def mine_mfi_cpuvec(bm, n_txns: int, rel_thr: float, budget_s=None):
    t0 = time.perf_counter()
    abs_sup = rel_thr * n_txns
    sup1 = LUT[bm].sum(axis=1)
    keep = np.nonzero(sup1 >= abs_sup)[0]
    order = keep[np.argsort(-sup1[keep], kind="stable")]
    packed = bm[order]
    m = len(order)
    tail_items = [frozenset(int(x) for x in order[k:]) for k in range(m + 1)]
    found: list[frozenset] = []
    visits = [0]

    def visit(node_bits, items: frozenset, nxt: int):
        visits[0] += 1
        if budget_s is not None and time.perf_counter() - t0 > budget_s:   # checked at every node
            raise TimeoutError
        counts = LUT[packed & node_bits].sum(axis=1)
        ext = [p for p in np.nonzero(counts >= abs_sup)[0].tolist() if int(order[p]) not in items]
        if not ext:
            if items and not any(items <= f for f in found):
                found.append(items)
            return
        cover = next((f for f in found if items <= f), None)
        if cover is not None and (items | tail_items[nxt]) <= cover:
            return
        for p in ext:
            if p >= nxt:
                visit(packed[p] & node_bits, items | {int(order[p])}, p + 1)

    for k in range(m):
        visit(packed[k], frozenset({int(order[k])}), k + 1)
    return found, time.perf_counter() - t0
