"""GPU parity over seeded random configurations, each against the oracle's exact MFI family.

Every case draws one of five data shapes:
- a sparse Zipf set (skew 0.3 or 0.6, at most a tenth of the items per transaction),
  whose loaded item count may pass 4,096 (searched on a compact copy of the frequent
  items);
- a workload-1 or workload-4 stream;
- 2-11 uneven windows;
- a stream with three thresholds in one search.

It also draws the search options: records per round (1 to 16,384), counting-tile shape,
sibling-ring size (one word: no blocks; small: blocks wrap and fall back to the pair
test; default) and frontier order. The draws are fixed by the seed, so a failure names its
case and repeats exactly."""
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


CASES = 30


@pytest.mark.parametrize("case", range(CASES))
def test_random_configuration_matches_oracle(case):
    rng = np.random.default_rng([20261204, case])
    kind = str(rng.choice(["zipf", "w4", "w1", "windows", "multi"]))
    batch = int(rng.choice([1, 7, 64, 256, 1024, 4096, 0]))
    tile = (int(rng.choice([128, 256, 512])), int(rng.choice([1, 2, 4, 8])))
    ring = int(rng.choice([0, 0, 1, 64, 4096, 1 << 16]))
    flags = int(rng.choice([0, L.ORDER_LEARNED]))
    seed = int(rng.integers(0, 10_000))
    print(f"case {case}: kind={kind} batch={batch} tile={tile} ring={ring} flags={flags} seed={seed}", flush=True)
    with Context(device=0, batch_nodes=batch) as ctx:
        ctx.set_scorer(*learned.load_weights())
        ctx.set_tile(*tile)
        ctx.set_sib_ring(ring)
        if kind == "windows":
            nw = int(rng.integers(2, 12))
            wins = gen.workload3(seed, n_windows=nw, lo=200, hi=6000)
            win_off, off, items, _ = gen.concat_windows(wins)
            rel = float(rng.choice([0.02, 0.03, 0.05]))
            thr = [gen.min_count(rel, T.n_txn) for T in wins]
            ctx.load_windows(win_off, off, items)
            got = _by_tag(ctx.mine_windows(thr, flags=flags), nw)
            for w, T in enumerate(wins):
                assert got[w] == _oracle(T.offsets, T.items, T.n_items, thr[w]), w
            return
        if kind == "zipf":
            n = int(rng.integers(2_000, 30_000))
            d = int(rng.choice([60, 300, 700, 5000]))
            zs = float(rng.choice([0.3, 0.6]))                          # sparse enough for the oracle
            mean = min(float(rng.choice([10.0, 30.0, 80.0])), d / 10.0)
            off, items = cref.zipf_csr(n, d, zs, mean, 6, 3, 6, 0.02, 0.05, seed)
            srt = np.sort(cref.row_supports(cref.bitmap(off, items, d)))[::-1]
            thr = int(max(2, srt[min(len(srt) - 1, int(rng.integers(20, 3000)))]))
        elif kind == "w1":
            T = gen.workload1(seed, int(rng.integers(500, 20_000)))
            off, items, d = T.offsets, T.items, T.n_items
            thr = gen.min_count(float(rng.choice([0.01, 0.02, 0.05])), T.n_txn)
        else:
            T = gen.workload4(seed, int(rng.integers(5_000, 40_000)))
            off, items, d = T.offsets, T.items, T.n_items
            thr = gen.min_count(float(rng.choice([0.02, 0.03, 0.05])), T.n_txn)
        ctx.load_csr(off, items)
        if kind == "multi":
            thrs = sorted({thr, int(thr * 1.5) + 1, int(thr * 3) + 1})
            got = _by_tag(ctx.mine_thresholds(thrs, flags=flags), len(thrs))
            for i, t in enumerate(thrs):
                assert got[i] == _oracle(off, items, d, t), t
        else:
            assert ctx.mine(thr, flags=flags).as_dict() == _oracle(off, items, d, thr)
