"""GPU known-answer test: the reference's seeded batch-prototype dataset (N = 2^23, 512 items).

The published per-level frequent-itemset counts at minsup 12582
(code/engine_tensorfim/code/batch_proto/run_dispatch.txt:5-15) must follow from the
MFIs the GPU mines: the downward closure of the MFI family is the frequent family.
"""
import sys
from itertools import combinations

import numpy as np
import pytest

from oracle import cref
from sgfim.core import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kat():
    rows, pats = cref.kat_generate()
    off, items = cref.kat_csr(rows)
    return rows, pats, off, items


def test_kat_mfis_give_published_levels(kat):
    print("case: batch_proto KAT mine at 12582", flush=True)
    rows, pats, off, items = kat
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        assert ctx.n_txn == cref.KAT_N
        res = ctx.mine(cref.KAT_MINSUP)
        # supports of a sample of MFIs against the oracle counter
        for x, s in list(zip(res.itemsets, res.supports))[:50]:
            ref = cref.count(rows, [i - 1 for i in x[1:]], [x[0] - 1])[0]
            assert ref == s
    closure = set()
    for x in res.itemsets:
        xs = sorted(x)
        for k in range(1, len(xs) + 1):
            closure.update(combinations(xs, k))
    levels = [sum(1 for c in closure if len(c) == k) for k in range(1, 8)]
    assert levels == cref.KAT_LEVELS
    assert len(closure) == cref.KAT_TOTAL
    recalled = sum(1 for p, _ in pats if tuple(i + 1 for i in p) in closure)
    assert recalled == 80


def test_kat_pair_matrix_and_counts(kat):
    print("case: batch_proto KAT pair matrix + level-2 counts", flush=True)
    rows, pats, off, items = kat
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        ids = ctx.item_ids()
        P = ctx.pair_matrix()
        sub = np.ascontiguousarray(rows[ids - 1][:64])
        R = cref.pair_matrix(sub)
        assert (P[:64, :64] == R).all()
        sup = ctx.item_supports()
        assert (sup == cref.row_supports(rows)[ids - 1]).all()
        # level-2 frequent count from the pair matrix (F2 = 585)
        f1 = np.where(sup >= cref.KAT_MINSUP)[0]
        sub_p = P[np.ix_(f1, f1)]
        assert len(f1) == 248
        assert int((np.triu(sub_p, 1) >= cref.KAT_MINSUP).sum()) == 585
