"""GPU parity on data built to trigger parent-equivalence pruning: items present in every
transaction, duplicated item columns, nested tidsets and dense random sets. The oracle
searches without any pruning (the reference's node step), so equality pins the pruning."""
import random

import numpy as np
import pytest

from oracle import cref
from oracle import mfi_oracle as O
from sgfim._lib import ROOT_STRICT, ROOT_TOPK
from sgfim.core import Context, csr_from_rows

pytestmark = pytest.mark.gpu


def _dense_rows(seed, n, d, p, always=(), copies=(), nested=()):
    rng = random.Random(seed)
    rows = []
    for _ in range(n):
        t = {i for i in range(1, d + 1) if rng.random() < p}
        t.update(always)
        for a, b in copies:          # item b is a copy of item a
            t.discard(b)
            if a in t:
                t.add(b)
        for a, b in nested:          # T(a) is contained in T(b)
            if a in t:
                t.add(b)
        rows.append(tuple(sorted(t)))
    return rows


def _oracle(rows, d, thr):
    off, items = csr_from_rows(rows)
    R = cref.bitmap(off, items, d)
    sup = cref.row_supports(R)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(R[keep], thr, count_mode=cref.COUNT_ALL)
    return {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}


CASES = [
    dict(seed=1, n=400, d=14, p=0.5, always=(2, 9)),
    dict(seed=2, n=600, d=16, p=0.45, copies=((1, 5), (3, 12), (12, 13))),
    dict(seed=3, n=500, d=15, p=0.4, nested=((2, 7), (4, 1), (9, 15)), always=(11,)),
    dict(seed=4, n=3000, d=24, p=0.35, always=(24,), copies=((20, 21),), nested=((6, 3),)),
    dict(seed=5, n=200, d=12, p=0.7),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pep_dense_cases_match_unpruned_oracle(case):
    kw = CASES[case]
    print(f"case: pep {kw}", flush=True)
    d = kw["d"]
    rows = _dense_rows(kw["seed"], kw["n"], d, kw["p"], kw.get("always", ()), kw.get("copies", ()),
                       kw.get("nested", ()))
    off, items = csr_from_rows(rows)
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        for rel in (0.05, 0.15, 0.3):
            thr = max(1, int(np.ceil(rel * len(rows))))
            got = ctx.mine(thr)
            ref = _oracle(rows, d, thr)
            assert got.as_dict() == ref, (case, rel)
            # items come back in descending id order (the reference's Results order)
            assert all(list(x) == sorted(x, reverse=True) for x in got.itemsets)
            if len(rows) <= 600:
                py = {frozenset(p): c for p, c in O.mine_mfi(rows, min_count=thr, node_cap=0)[0]}
                assert py == ref
        # engine root rules on the same data
        s = 0.2
        thr = int(np.ceil(len(rows) * s))
        assert ctx.mine(thr, flags=ROOT_STRICT, rel_thr=s).as_dict() == O.tensorfim_family(rows, s)
        for k in (1, 3, d // 2):
            fam, t = O.anyfim_stage_family(rows, k)
            assert ctx.mine(t, flags=ROOT_TOPK, root_topk=k).as_dict() == fam
