"""Pure-Python oracle of the exact MFI family -- TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module; the product path never does.

mine_mfi restates the reference Python miner code/pipeline/skab_spike.py:84-135:
  * threshold: max(1, int(min_count)) when min_count is given, else
    max(1, ceil(min_sup * n))                                  (skab_spike.py:98-100)
  * tidsets as Python sets, frequent items in ascending order  (:101-107)
  * depth-first set enumeration over forward extensions; a node with no frequent
    forward extension is a candidate; the node cap raises RuntimeError (:108-126)
  * subsumption filter: sorted by decreasing length, a candidate properly contained
    in an already kept one is dropped                           (:128-135)
It returns ([(tuple(items), count)], n) like the reference.

brute_force_mfi enumerates every subset of the frequent items (small inputs only);
it pins mine_mfi independently of any search order.

tensorfim_family / anyfim_stage_family restate the two engines' emitted families
(kernel_bitmap.cu:816,841-894 and AnyTime kernel.cu:844-882): the engines search
only the subtrees rooted at selected items, children having smaller ids than their
parent's last item, so they emit exactly the MFIs whose largest item id is a root.
"""
from __future__ import annotations

import math
from collections import defaultdict
from itertools import combinations


def threshold(n: int, min_sup: float = 0.05, min_count=None) -> int:
    if min_count is not None:
        return max(1, int(min_count))
    return max(1, int(math.ceil(min_sup * n)))


def mine_mfi(transactions, min_sup=0.05, min_count=None, node_cap=200000):
    n = len(transactions)
    thresh = threshold(n, min_sup, min_count)
    tidsets = defaultdict(set)
    for i, t in enumerate(transactions):
        for item in t:
            tidsets[item].add(i)
    items = sorted(it for it, ts in tidsets.items() if len(ts) >= thresh)
    found = []
    nodes = [0]

    def dfs(prefix, cand, cur):
        nodes[0] += 1
        if node_cap and nodes[0] > node_cap:
            raise RuntimeError(f"search tree exceeds the node cap {node_cap}")
        ext = []
        for c in cand:
            ts = cur & tidsets[c]
            if len(ts) >= thresh:
                ext.append((c, ts))
        if not ext:
            if prefix:
                found.append((tuple(prefix), len(cur)))
            return
        for i, (c, ts) in enumerate(ext):
            dfs(prefix + [c], [x for x, _ in ext[i + 1:]], ts)

    dfs([], items, set(range(n)))
    kept = []
    for pat, cnt in sorted(found, key=lambda pc: -len(pc[0])):
        s = set(pat)
        if any(s < set(k) for k, _ in kept):
            continue
        kept.append((pat, cnt))
    return kept, n


def support(transactions, itemset) -> int:
    s = set(itemset)
    return sum(1 for t in transactions if s <= set(t))


def brute_force_mfi(transactions, thresh: int):
    """Every frequent itemset by exhaustive enumeration, then the maximal ones."""
    sets = [frozenset(t) for t in transactions]
    items = sorted({x for t in sets for x in t})
    items = [x for x in items if sum(1 for t in sets if x in t) >= thresh]
    freq = {}
    for k in range(1, len(items) + 1):
        level = False
        for comb in combinations(items, k):
            c = frozenset(comb)
            if k > 1 and any((c - {x}) not in freq for x in c):
                continue
            s = sum(1 for t in sets if c <= t)
            if s >= thresh:
                freq[c] = s
                level = True
        if not level:
            break
    return {x: s for x, s in freq.items() if not any(x < y for y in freq if len(y) == len(x) + 1)}


def item_counts(transactions):
    cnt = defaultdict(int)
    for t in transactions:
        for x in set(t):
            cnt[x] += 1
    return cnt


def tensorfim_family(transactions, rel_thr: float, n_total=None):
    """MFIs TensorFIM emits at relative threshold s: survivors have count >= n*s
    (double compare, kernel_bitmap.cu:350), roots need count/n > s (kernel_bitmap.cu:816)."""
    n = len(transactions) if n_total is None else n_total
    thr = math.ceil(n * rel_thr)
    cnt = item_counts(transactions)
    exact = brute_force_mfi(transactions, max(1, thr)) if len(cnt) <= 20 else \
        {frozenset(p): c for p, c in mine_mfi(transactions, min_count=max(1, thr), node_cap=0)[0]}
    return {x: s for x, s in exact.items() if cnt[max(x)] / n > rel_thr}


def anyfim_stage_family(transactions, k: int):
    """Stage k of AnyFIM: thr = k-th highest item count, roots = the first k items of
    the stable descending-count order with ties by ascending id (kernel.cu:276-285,846-859)."""
    cnt = item_counts(transactions)
    order = sorted(cnt, key=lambda x: (-cnt[x], x))
    thr = cnt[order[k - 1]]
    roots = set(order[:k])
    exact = {frozenset(p): c for p, c in mine_mfi(transactions, min_count=thr, node_cap=0)[0]}
    return {x: s for x, s in exact.items() if max(x) in roots}, thr
