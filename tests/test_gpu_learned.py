"""GPU test of what the trained proposal network gains: on a held-out workload-4 stream
(500,000 txns x 256 items, 2% floor), the learned frontier order confirms half of the
final MFI set with clearly fewer support counts than the plain support order, and both
orders end with the same exact MFI set."""
import pytest

from sgfim import gen, learned
from sgfim.core import Context
from sgfim.train_proposal import anytime_curve, t_at

pytestmark = pytest.mark.gpu


def test_learned_order_reaches_half_coverage_with_fewer_counts():
    print("case: learned vs plain order, workload 4 seed 1000002, rounds of 64 records", flush=True)
    T = gen.workload4(1_000_002)
    thr = gen.min_count(0.02, T.n_txn)
    with Context(device=0, batch_nodes=64) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_scorer(*learned.load_weights())
        rp, fp = anytime_curve(ctx, thr, False)
        rl, fl = anytime_curve(ctx, thr, True)
    assert fp == fl                       # the order never changes the result
    tot = len(fp)
    assert tot == 630_358
    assert rp[-1][1] == rl[-1][1]         # nor the total work
    c50_plain, c50_learned = t_at(rp, 0.5, tot, 1), t_at(rl, 0.5, tot, 1)
    print(f"counts to 50% coverage: plain {c50_plain}, learned {c50_learned}", flush=True)
    assert c50_learned < 0.85 * c50_plain
