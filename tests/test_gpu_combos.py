"""GPU parity for combinations of features, each against the oracle's exact MFI families:
- windows with the learned frontier order;
- several thresholds in one search with the learned order;
- an anytime run cancelled partway: every MFI it reported is already final (in the
  oracle's family, with its support), and a full run on the same context afterwards is
  exact;
- anytime rounds over more than 4,096 loaded items (a compact copy of the frequent ones)
  in learned order, taken round by round."""
import numpy as np
import pytest

from oracle import cref
from sgfim import _lib as L
from sgfim import gen, learned
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle_rows(rows, thr):
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): int(s) for x, s in out}


def _oracle(T, thr):
    return _oracle_rows(cref.bitmap(T.offsets, T.items, T.n_items), thr)


def _by_tag(res, n):
    got = [dict() for _ in range(n)]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    return got


def test_windows_in_learned_order():
    print("case: 16 windows (seed 3, 500-8000 txns) at 3%, learned order", flush=True)
    wins = gen.workload3(3, n_windows=16, lo=500, hi=8000)
    win_off, off, items, _ = gen.concat_windows(wins)
    thr = [gen.min_count(0.03, T.n_txn) for T in wins]
    with Context(device=0, batch_nodes=256) as ctx:
        ctx.load_windows(win_off, off, items)
        ctx.set_scorer(*learned.load_weights())
        got = _by_tag(ctx.mine_windows(thr, flags=L.ORDER_LEARNED), len(wins))
    for w, T in enumerate(wins):
        assert got[w] == _oracle(T, thr[w]), w


def test_thresholds_in_learned_order():
    print("case: workload 4 seed 1000007 at 60,000 txns, thresholds 2%/4%/8% in one search, learned order",
          flush=True)
    T = gen.workload4(1_000_007, 60_000)
    thrs = [gen.min_count(s, T.n_txn) for s in (0.02, 0.04, 0.08)]
    with Context(device=0, batch_nodes=1024) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_scorer(*learned.load_weights())
        got = _by_tag(ctx.mine_thresholds(thrs, flags=L.ORDER_LEARNED), len(thrs))
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    for i, thr in enumerate(thrs):
        assert got[i] == _oracle_rows(rows, thr), thr


def test_cancelled_anytime_run_reports_only_final_mfis():
    print("case: workload 4 seed 1000008 at 60,000 txns, 2%, rounds of 64, cancelled after 5 rounds", flush=True)
    T = gen.workload4(1_000_008, 60_000)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle(T, thr)
    with Context(device=0, batch_nodes=64) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.anytime_begin(thr, 0)
        part = {}
        for _ in range(5):
            st = ctx.anytime_step(1)
            part.update(ctx.anytime_take().as_dict())
            if st["done"]:
                break
        assert not st["done"]
        ctx.anytime_cancel()
        with pytest.raises(L.SgfimError) as e:
            while not ctx.anytime_step(1)["done"]:
                pass
        assert e.value.code == L.E_CANCELLED
        assert 0 < len(part) < len(ref)
        assert all(ref.get(x) == s for x, s in part.items())
        assert ctx.mine(thr).as_dict() == ref


@pytest.mark.parametrize("flags", [0, L.ORDER_LEARNED])
def test_direct_launch_anytime_rounds_of_64_match_oracle(flags):
    """Anytime stepping with direct launches (no captured graphs, stage events on every
    launch) in rounds of 64 records, the host path that drains the MFI ring and reads the
    control snapshot and stage events every round; the stage times add up to device time."""
    print(f"case: workload 4 seed 1000009 at 60,000 txns, 2%, NO_GRAPH anytime, rounds of 64, flags={flags}",
          flush=True)
    T = gen.workload4(1_000_009, 60_000)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle(T, thr)
    with Context(device=0, batch_nodes=64) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_scorer(*learned.load_weights())
        ctx.anytime_begin(thr, flags | L.NO_GRAPH)
        got = {}
        rounds = 0
        while True:
            st = ctx.anytime_step(1)
            got.update(ctx.anytime_take().as_dict())
            rounds += 1
            if st["done"]:
                break
        stages = ctx.stage_times()
    assert rounds > 20
    assert got == ref
    assert all(v >= 0.0 for v in stages.values()) and stages["count"] > 0.0


def test_anytime_rounds_on_a_compact_universe_in_learned_order():
    print("case: 6,000 loaded items (zipf 0.3, n=20000), min_count 400, anytime learned order, rounds of 256",
          flush=True)
    off, items = cref.zipf_csr(20_000, 6000, 0.3, 120.0, 10, 3, 6, 0.02, 0.05, 5)
    rows = cref.bitmap(off, items, 6000)
    ref = _oracle_rows(rows, 400)
    with Context(device=0, batch_nodes=256) as ctx:
        ctx.load_csr(off, items)
        ctx.set_scorer(*learned.load_weights())
        ctx.anytime_begin(400, L.ORDER_LEARNED)
        got = {}
        rounds = 0
        while True:
            st = ctx.anytime_step(1)
            got.update(ctx.anytime_take().as_dict())
            rounds += 1
            if st["done"]:
                break
    assert rounds > 1
    assert got == ref


def test_anytime_queue_carries_over_steps_and_settles_for_other_calls():
    """Captured rounds of 64 records taken one or three at a time: each step returns with
    the next rounds still queued. Every few steps a pair-matrix read first processes the
    queued rounds (their MFIs become pending) before it runs; the search then goes on.
    Every MFI is reported exactly once and the union is the oracle's family."""
    print("case: workload 4 seed 1000010 at 60,000 txns, 2%, graph anytime, rounds of 64, settle by pair reads",
          flush=True)
    T = gen.workload4(1_000_010, 60_000)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle(T, thr)
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    pair_ref = cref.pair_matrix(rows)
    with Context(device=0, batch_nodes=64) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.anytime_begin(thr, 0)
        got = {}
        steps = 0
        while True:
            st = ctx.anytime_step(1 if steps % 2 else 3)
            r = ctx.anytime_take()
            for x, s in zip(r.itemsets, r.supports):
                k = frozenset(x)
                assert k not in got
                got[k] = int(s)
            steps += 1
            if st["done"]:
                break
            if steps % 5 == 0:
                assert (ctx.pair_matrix() == pair_ref).all()
        assert steps > 10
        r = ctx.anytime_take()
        assert r.count == 0
    assert got == ref
