"""GPU tests of the sibling ring. A record's counting tile asks, for child e and candidate f,
whether f is in F(P u {e}) of the parent P's expansion. That test is tighter than the pair
matrix and exact, since |T(Y u {e, f})| <= |T(P u {e, f})|.

Every ring size keeps the exact MFI family:
- the default;
- a ring small enough to wrap many times per search (its old blocks fall back to the pair
  test);
- one word (no block fits, so every tile uses the pair test).

This holds in both frontier orders and at two round sizes. On this deep search the default
ring counts fewer supports than no ring."""
import numpy as np
import pytest

from oracle import cref
from sgfim import gen, learned
from sgfim._lib import ORDER_LEARNED
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle(T, thr):
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): int(s) for x, s in out}


@pytest.mark.parametrize("batch", [64, 4096])
def test_sibling_ring_sizes_keep_the_mfi_family(batch):
    print(f"case: sibling ring sizes, workload 4 seed 1000004 at 50,000 txns, 2%, rounds of {batch}", flush=True)
    T = gen.workload4(1_000_004, 50_000)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle(T, thr)
    counts = {}
    with Context(device=0, batch_nodes=batch) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_scorer(*learned.load_weights())
        for flags in (0, ORDER_LEARNED):
            for words in (0, 1 << 14, 1):
                ctx.set_sib_ring(words)
                r = ctx.mine(thr, flags=flags)
                assert r.as_dict() == ref, (flags, words)
                counts[(flags, words)] = r.stats["counts"]
    print(f"support counts by (order, ring words): {counts}", flush=True)
    for flags in (0, ORDER_LEARNED):
        assert counts[(flags, 0)] < counts[(flags, 1)]
        assert counts[(flags, 1 << 14)] <= counts[(flags, 1)]
