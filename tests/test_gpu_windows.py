"""GPU tests of the batched-window search (many uneven windows mined in one run)."""
import numpy as np
import pytest

from oracle import cref
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle(T, thr):
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}


@pytest.mark.parametrize("seed,nwin,lo,hi", [(0, 24, 300, 20000), (1, 64, 2000, 6000)])
def test_windows_match_per_window_oracle(seed, nwin, lo, hi):
    print(f"case: windows seed={seed} n={nwin} lens [{lo},{hi}]", flush=True)
    wins = gen.workload3(seed, n_windows=nwin, lo=lo, hi=hi)
    win_off, off, items, d = gen.concat_windows(wins)
    thr = [gen.min_count(0.03, T.n_txn) for T in wins]
    with Context(device=0) as ctx:
        ctx.load_windows(win_off, off, items)
        assert ctx.n_windows == nwin
        for w in (0, nwin - 1):
            rows = cref.bitmap(wins[w].offsets, wins[w].items, wins[w].n_items)
            assert (ctx.window_supports(w) == cref.row_supports(rows)[ctx.item_ids() - 1]).all()
        res = ctx.mine_windows(thr)
        got = [dict() for _ in range(nwin)]
        for x, s, w in zip(res.itemsets, res.supports, res.windows):
            got[w][frozenset(x)] = s
        for w, T in enumerate(wins):
            assert got[w] == _oracle(T, thr[w]), w
        # the same windows one at a time through the single-set entry point
        for w in (0, nwin // 2):
            ctx2 = Context(device=0)
            ctx2.load_csr(wins[w].offsets, wins[w].items)
            assert ctx2.mine(thr[w]).as_dict() == got[w]
            ctx2.close()
