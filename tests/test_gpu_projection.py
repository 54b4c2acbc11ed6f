"""GPU parity of projected blocks: every search setting that changes where supports are
counted (projection factor, arena size, round size, frontier order, windows, several
thresholds) gives the oracle's exact MFI family.

A projected block holds the rows of U(Y) restricted and bit-compacted to T(Y); the
records under Y count on it. Dense blocks are compacted word by word (bit compress of each
nonzero source word), sparse ones gathered bit by bit; the block rule says which and how
many children a dense block needs. The cases cover:
- factor 0 (window rows only), 2, 4 and 16;
- block rules that build every block word by word (dense 2), every block bit by bit
  (dense 0), and mixes, with 2 to 16 children required;
- arenas so small that blocks are refused or overwritten before their records are
  counted (the records then fall back to window rows or project again);
- rounds of one record (each block built alone) up to whole batches;
- the support and learned frontier orders (different block lifetimes);
- windows of different lengths and three thresholds in one search.
"""
import numpy as np
import pytest

from oracle import cref
from sgfim import gen, learned
from sgfim import _lib as L
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle(off, items, d, thr):
    rows = cref.bitmap(off, items, d)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): int(s) for x, s in out}


def _by_tag(res, n):
    got = [dict() for _ in range(n)]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    return got


CASES = [
    # kind, n, rel, factor, arena words, batch, order, block rule (dense, min children)
    ("w4", 30_000, 0.02, 4, 0, 0, 0, (32, 2)),
    ("w4", 30_000, 0.02, 4, 0, 0, 0, (32, 16)),
    ("w4", 30_000, 0.02, 4, 0, 0, L.ORDER_LEARNED, (2, 2)),    # every block word by word
    ("w4", 30_000, 0.02, 4, 0, 0, 0, (0, 2)),                  # every block bit by bit
    ("w4", 30_000, 0.02, 2, 0, 64, L.ORDER_LEARNED, (8, 2)),
    ("w4", 30_000, 0.02, 16, 1 << 12, 256, 0, (32, 2)),        # tiny arena: most blocks refused
    ("w4", 30_000, 0.02, 4, 1 << 16, 7, 0, (32, 4)),           # small arena: blocks overwritten
    ("w4", 30_000, 0.02, 4, 1 << 16, 1, L.ORDER_LEARNED, (2, 2)),
    ("w1", 20_000, 0.005, 2, 0, 0, 0, (32, 2)),
    ("w1", 20_000, 0.005, 4, 1 << 14, 33, L.ORDER_LEARNED, (2, 2)),
    ("zipf", 8_000, 0.03, 4, 0, 0, 0, (2, 2)),
    ("zipf", 8_000, 0.03, 2, 1 << 15, 128, L.ORDER_LEARNED, (8, 16)),
    ("windows", 0, 0.03, 4, 0, 0, 0, (2, 2)),
    ("windows", 0, 0.02, 2, 1 << 15, 16, L.ORDER_LEARNED, (32, 2)),
    ("multi", 25_000, 0.02, 4, 0, 0, 0, (2, 2)),
    ("multi", 25_000, 0.02, 4, 1 << 16, 100, L.ORDER_LEARNED, (32, 16)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_projected_search_matches_oracle(case):
    kind, n, rel, factor, arena, batch, flags, rule = CASES[case]
    print(f"case {case}: kind={kind} n={n} rel={rel} factor={factor} arena={arena} batch={batch} flags={flags} "
          f"rule={rule}", flush=True)
    with Context(device=0, batch_nodes=batch) as ctx:
        ctx.set_scorer(*learned.load_weights())
        ctx.set_projection(factor, arena)
        ctx.set_block_rule(*rule)
        assert ctx.projection() == (factor, arena)
        assert ctx.block_rule() == rule
        if kind == "windows":
            wins = gen.workload3(case, n_windows=6, lo=500, hi=20_000)
            win_off, off, items, _ = gen.concat_windows(wins)
            thr = [gen.min_count(rel, T.n_txn) for T in wins]
            ctx.load_windows(win_off, off, items)
            got = _by_tag(ctx.mine_windows(thr, flags=flags), len(wins))
            for w, T in enumerate(wins):
                assert got[w] == _oracle(T.offsets, T.items, T.n_items, thr[w]), w
            return
        if kind == "zipf":
            off, items = cref.zipf_csr(n, 400, 0.6, 25.0, 6, 3, 7, 0.02, 0.06, case)
            d = 400
        elif kind == "w1":
            T = gen.workload1(case, n)
            off, items, d = T.offsets, T.items, T.n_items
        else:
            T = gen.workload4(case, n)
            off, items, d = T.offsets, T.items, T.n_items
        thr = gen.min_count(rel, len(off) - 1)
        ctx.load_csr(off, items)
        if kind == "multi":
            thrs = [thr, 2 * thr, 4 * thr]
            got = _by_tag(ctx.mine_thresholds(thrs, flags=flags), 3)
            for i, t in enumerate(thrs):
                assert got[i] == _oracle(off, items, d, t), t
            return
        res = ctx.mine(thr, flags=flags)
        st = ctx.last_stats()
        print(f"  {len(res)} MFIs, {st['projections']} blocks ({st['proj_from_rows']} from window rows)", flush=True)
        assert res.as_dict() == _oracle(off, items, d, thr)
        if factor and not arena and kind == "w4" and rule[1] == 2:
            assert st["projections"] > 0   # the case exercises blocks, and blocks of blocks
            assert st["projections"] > st["proj_from_rows"]


def test_projection_settings_are_validated():
    with Context(device=0) as ctx:
        for f, a in ((1, 0), (-1, 0), (4, 100), (2000, 0)):
            with pytest.raises(L.SgfimError) as e:
                ctx.set_projection(f, a)
            assert e.value.code == L.E_INVAL
        ctx.set_projection(0)
        assert ctx.projection() == (0, 0)
        for d, m in ((1, 2), (-1, 2), (2000, 2), (32, 1), (32, 0), (32, 5000)):
            with pytest.raises(L.SgfimError) as e:
                ctx.set_block_rule(d, m)
            assert e.value.code == L.E_INVAL
        ctx.set_block_rule(0, 2)
        assert ctx.block_rule() == (0, 2)
