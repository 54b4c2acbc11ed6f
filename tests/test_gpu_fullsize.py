"""GPU parity at the benchmark's full sizes against the C oracle's exact MFI families:
workload 2 (1,000,000 txns x 512 items at 10%, 5% and 2%, one batched search), workload 3
(512 windows, 24.7M txns, 3% per window, one search) and workload 4 (500,000 txns x 256
items at 2%, plain and learned frontier order)."""
import numpy as np
import pytest

from oracle import cref
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle(off, items, n_items, thr):
    """{(sorted original ids): support} of the exact MFI family."""
    rows = cref.bitmap(off, items, n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    if len(keep) == 0:
        return {}
    out, _ = cref.mine(rows[keep], thr)
    return {tuple(sorted(int(keep[i]) + 1 for i in x)): int(s) for x, s in out}


def _by_set(res, n_sets):
    """Per set index: {(sorted ids): support} from a flat-array result."""
    items, off, sup = res.arrays
    win = np.asarray(res.windows)
    got = [dict() for _ in range(n_sets)]
    for i in range(len(sup)):
        got[int(win[i])][tuple(sorted(items[off[i]:off[i + 1]].tolist()))] = int(sup[i])
    return got


def test_workload2_full_size_matches_oracle():
    print("case: workload 2 full size (1M txns x 512 items) at 10%/5%/2%", flush=True)
    T = gen.workload2(0)
    thrs = [gen.min_count(s, T.n_txn) for s in (0.10, 0.05, 0.02)]
    with Context(device=0) as ctx:
        ctx.load_csr(T.offsets, T.items)
        got = _by_set(ctx.mine_thresholds(thrs, convert=False), len(thrs))
    for thr, g in zip(thrs, got):
        ref = _oracle(T.offsets, T.items, T.n_items, thr)
        assert len(ref) > 0
        assert g == ref, thr
    assert [len(g) for g in got] == [512, 530, 9174]


def test_workload3_full_size_matches_oracle():
    print("case: workload 3 full size (512 windows, 24.7M txns) at 3% per window", flush=True)
    win_off, off, items, lens = gen.workload3_csr(0)
    thrs = [gen.min_count(0.03, int(n)) for n in lens]
    with Context(device=0) as ctx:
        ctx.load_windows(win_off, off, items)
        got = _by_set(ctx.mine_windows(thrs, convert=False), len(lens))
    total = 0
    for w in range(len(lens)):
        a, b = int(win_off[w]), int(win_off[w + 1])
        o = off[a:b + 1] - off[a]
        it = items[off[a]:off[b]]
        ref = _oracle(o, it, 192, thrs[w])
        assert got[w] == ref, w
        total += len(ref)
    assert total == 2_702_719


def test_workload4_full_size_both_orders_match_oracle():
    print("case: workload 4 full size (500,000 txns x 256 items, seed 1000004) at 2%, plain and learned order",
          flush=True)
    from sgfim import learned
    from sgfim._lib import ORDER_LEARNED
    T = gen.workload4(1_000_004)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle(T.offsets, T.items, T.n_items, thr)
    assert len(ref) == 544_863
    with Context(device=0, batch_nodes=1024) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_scorer(*learned.load_weights())
        for flags in (0, ORDER_LEARNED):
            got = _by_set(ctx.mine(thr, flags=flags, convert=False), 1)[0]
            assert got == ref, flags
