"""GPU parity at the widest item sets.

- 4,096 loaded items are searched in place: 64-word extension masks and a 32 x 32 grid of
  128 x 128 pair tiles.
- More than 4,096 loaded items are searched on a compact copy of the items frequent at the
  search's thresholds. This covers a single set, new thresholds (a new copy), both
  frontier orders, windows and several thresholds in one search. The query functions
  keep the loaded items.
- A search with more than 4,096 frequent items is refused with E_INVAL, and the context
  stays usable.

The reference drops infrequent items before it builds its bitmap (kernel_bitmap.cu:692-703),
so any number of loaded items with few enough frequent ones is a valid input there."""
import numpy as np
import pytest

from oracle import cref
from sgfim import _lib as L
from sgfim import learned
from sgfim.core import Context

pytestmark = pytest.mark.gpu

WIDE = (20_000, 6000, 0.3, 120.0, 10, 3, 6, 0.02, 0.05, 5)   # 6,000 items, flat Zipf(0.3), 10 planted sets


def _oracle(off, items, d, thr):
    rows = cref.bitmap(off, items, d)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): int(s) for x, s in out}, len(keep)


def test_4096_loaded_items_searched_in_place():
    print("case: 4,096 loaded items, zipf(0.3) n=20000 mean 120, min_count 300", flush=True)
    off, items = cref.zipf_csr(20_000, 4096, 0.3, 120.0, 10, 3, 6, 0.02, 0.05, 5)
    ref, nf = _oracle(off, items, 4096, 300)
    assert nf == 4096
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        assert ctx.n_items == 4096
        assert ctx.mine(300).as_dict() == ref


def test_compact_universe_single_set_orders_and_rebuilds():
    print("case: 6,000 loaded items, min_count 318 (4,030 frequent) and 400, both orders", flush=True)
    off, items = cref.zipf_csr(*WIDE)
    ref_a, na = _oracle(off, items, 6000, 318)
    ref_b, nb = _oracle(off, items, 6000, 400)
    assert 4000 < na <= 4096 and nb < na
    with Context(device=0, batch_nodes=1024) as ctx:
        ctx.load_csr(off, items)
        ctx.set_scorer(*learned.load_weights())
        assert ctx.n_items == 6000
        assert ctx.mine(318).as_dict() == ref_a
        assert ctx.mine(400).as_dict() == ref_b                          # a new compact copy
        assert ctx.mine(318, flags=L.ORDER_LEARNED).as_dict() == ref_a   # and back, learned order
        assert ctx.mine(318).as_dict() == ref_a                          # the copy kept
        # the query functions still see all loaded items
        rows = cref.bitmap(off, items, 6000)
        sup = cref.row_supports(rows)
        ids = ctx.item_ids()
        assert len(ids) == 6000
        assert (ctx.item_supports() == sup[ids - 1]).all()
        pm = ctx.pair_matrix()
        assert pm.shape == (6000, 6000)
        assert (np.diag(pm) == sup[ids - 1]).all()
        rng = np.random.default_rng(0)
        for a, b in rng.integers(0, 6000, size=(50, 2)):
            ra, rb = rows[ids[a] - 1], rows[ids[b] - 1]
            assert pm[a, b] == int(np.unpackbits((ra & rb).view(np.uint8)).sum())


def test_compact_universe_windows_and_thresholds():
    print("case: 6,000 loaded items over two windows (12,000 + 9,000 txns), and thresholds 318/400 in one search",
          flush=True)
    o1, i1 = cref.zipf_csr(12_000, 6000, 0.3, 120.0, 6, 3, 6, 0.02, 0.05, 7)
    o2, i2 = cref.zipf_csr(9_000, 6000, 0.3, 120.0, 6, 3, 6, 0.02, 0.05, 8)
    win_off = np.array([0, 12_000, 21_000], dtype=np.int64)
    off = np.concatenate([o1, o2[1:] + o1[-1]]).astype(np.int64)
    items = np.concatenate([i1, i2]).astype(np.int32)
    thr = [260, 200]
    refs = [_oracle(o1, i1, 6000, thr[0]), _oracle(o2, i2, 6000, thr[1])]
    assert all(0 < n <= 4096 for _, n in refs)
    with Context(device=0) as ctx:
        ctx.load_windows(win_off, off, items)
        res = ctx.mine_windows(thr)
        got = [dict(), dict()]
        for x, s, w in zip(res.itemsets, res.supports, res.windows):
            got[w][frozenset(x)] = s
        assert got[0] == refs[0][0]
        assert got[1] == refs[1][0]
    off, items = cref.zipf_csr(*WIDE)
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        res = ctx.mine_thresholds([318, 400])
        got = [dict(), dict()]
        for x, s, w in zip(res.itemsets, res.supports, res.windows):
            got[w][frozenset(x)] = s
    assert got[0] == _oracle(off, items, 6000, 318)[0]
    assert got[1] == _oracle(off, items, 6000, 400)[0]


def test_more_than_4096_frequent_items_refused():
    print("case: 6,000 loaded items at min_count 250 (over 4,096 frequent), then 400", flush=True)
    off, items = cref.zipf_csr(*WIDE)
    assert int((cref.row_supports(cref.bitmap(off, items, 6000)) >= 250).sum()) > 4096
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(250)
        assert e.value.code == L.E_INVAL
        assert ctx.mine(400).as_dict() == _oracle(off, items, 6000, 400)[0]
