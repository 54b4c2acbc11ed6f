"""GPU test of what the trained proposal network gains: on a held-out workload-4 stream
(500,000 txns x 256 items, 2% floor), the learned frontier order confirms half of the
final MFI set with clearly fewer support counts than the plain support order, and both
orders end with the same exact MFI set."""
import numpy as np
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


def _oracle_mfis(T, thr):
    from oracle import cref
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}


def test_scorer_swap_and_reloads_on_one_context():
    """A second set_scorer overwrites the weights that a captured learned-order round reads
    (the negated network, score 1 - s, swaps the two halves of the frontier and loses its
    head start; the original weights set again win it back), and new loads on the same
    context -- smaller data on the kept search buffers, then larger data that needs new
    ones -- re-capture the rounds and keep the exact MFI families."""
    from sgfim._lib import ORDER_LEARNED
    print("case: scorer swap and reloads, workload 1 seed 0 at 2% and workload 4 seed 1000002, "
          "rounds of 64 records", flush=True)
    w1, b1, w2, b2 = learned.load_weights()
    T1 = gen.workload1(0)
    thr1 = gen.min_count(0.02, T1.n_txn)
    ref1 = _oracle_mfis(T1, thr1)
    T4 = gen.workload4(1_000_002)
    thr4 = gen.min_count(0.02, T4.n_txn)
    with Context(device=0, batch_nodes=64) as ctx:
        ctx.set_scorer(w1, b1, w2, b2)
        ctx.load_csr(T1.offsets, T1.items)
        for flags in (ORDER_LEARNED, 0):
            assert ctx.mine(thr1, flags=flags).as_dict() == ref1, flags
        ctx.load_csr(T4.offsets, T4.items)
        ra, fa = anytime_curve(ctx, thr4, True)
        ctx.set_scorer(w1, b1, -w2, -b2)
        rn, fn = anytime_curve(ctx, thr4, True)
        ctx.set_scorer(w1, b1, w2, b2)
        rb, fb = anytime_curve(ctx, thr4, True)
        ctx.load_csr(T1.offsets, T1.items)
        for flags in (0, ORDER_LEARNED):
            assert ctx.mine(thr1, flags=flags).as_dict() == ref1, flags
    assert fa == fn == fb
    tot = len(fa)
    assert tot == 630_358
    c50 = [t_at(r, 0.5, tot, 1) for r in (ra, rn, rb)]
    print(f"counts to 50% coverage: trained {c50[0]}, negated {c50[1]}, trained again {c50[2]}", flush=True)
    assert c50[1] > 1.3 * c50[0]
    assert abs(c50[2] - c50[0]) < 0.1 * c50[0]
