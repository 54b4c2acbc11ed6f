"""Pair co-occurrence matrix kernels on the GPU: fp4 MFMA (default) and int8 MFMA against
the VALU kernel, the C oracle and integer closed forms, including a transaction split that
fills the fp4 kernel's exact fp32 range."""
import sys

import numpy as np
import pytest

from oracle import cref
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu

KINDS = [(Context.PAIR_FP4, "fp4"), (Context.PAIR_INT8, "int8")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(device=0)
    yield c
    c.set_pair_kernel(Context.PAIR_FP4, 0)
    c.close()


def _csr(rows_l):
    off = np.zeros(len(rows_l) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(r) for r in rows_l])
    items = np.concatenate(rows_l).astype(np.int32) if len(rows_l) else np.zeros(0, np.int32)
    return off, items


@pytest.mark.parametrize("kind,name", KINDS)
@pytest.mark.parametrize("seed,n", [(0, 4000), (3, 70001)])
def test_pair_kernels_match_oracle(ctx, kind, name, seed, n):
    say(f"case: pair matrix {name} workload1-shape seed={seed} n={n}")
    T = gen.workload1(seed, n)
    ctx.load_csr(T.offsets, T.items)
    ctx.set_pair_kernel(kind, 0)
    P = ctx.pair_matrix()
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    assert (P == cref.pair_matrix(rows)).all()
    assert (P == ctx.pair_matrix(valu=True)).all()


@pytest.mark.parametrize("kind,name", KINDS)
@pytest.mark.parametrize("cap", [0, 1, 3])
def test_pair_kernels_ragged_tiles_any_split(ctx, kind, name, cap):
    say(f"case: pair matrix {name} 700 items (6 x 6 tiles, the last ragged), 20011 txns, max_splits={cap}")
    rng = np.random.default_rng(11)
    n, d = 20011, 700
    dense = rng.random((n, d)) < 0.05
    off, items = _csr([np.nonzero(r)[0] + 1 for r in dense])
    ctx.load_csr(off, items)
    ctx.set_pair_kernel(kind, cap)
    ids = ctx.item_ids()
    P = ctx.pair_matrix()
    D = dense[:, ids - 1].astype(np.float64)   # exact: counts < 2^53
    assert (P == (D.T @ D).astype(np.int64)).all()
    assert ctx.pair_time() > 0.0


def test_pair_fp4_full_split_is_exact(ctx):
    """One split of the fp4 kernel spans its maximum, 2^20 transactions, with rows set in
    every one of them: its fp32 partial sums reach 2^20 and must still be exact. Closed
    forms: item 1 is in every transaction, item 2 in all but the last, item 3 in the even
    ones, item 4 in every third; the rest are seeded random."""
    n = (1 << 21) - 3 * 64 - 5   # two splits at max_splits=1: 2^20 and the ragged rest
    say(f"case: pair matrix fp4 full split, n={n}, 24 items, max_splits=1")
    rng = np.random.default_rng(7)
    t = np.arange(n)
    member = np.zeros((n, 24), dtype=bool)
    member[:, 0] = True
    member[:-1, 1] = True
    member[:, 2] = (t % 2) == 0
    member[:, 3] = (t % 3) == 0
    member[:, 4:] = rng.random((n, 20)) < 0.3
    cnt = member.sum(axis=1)
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(cnt)
    items = (np.nonzero(member)[1] + 1).astype(np.int32)   # row-major: per transaction, ascending
    ctx.load_csr(off, items)
    assert list(ctx.item_ids()) == list(range(1, 25))
    ref = None
    for kind, name in KINDS:
        ctx.set_pair_kernel(kind, 1)
        P = ctx.pair_matrix().astype(np.int64)
        assert P[0, 0] == n and P[1, 1] == n - 1 and P[0, 1] == n - 1
        assert P[2, 2] == (n + 1) // 2 and P[3, 3] == (n + 2) // 3
        assert P[2, 3] == (n + 5) // 6   # multiples of 6 below n
        if ref is None:
            M = member.astype(np.float64)
            ref = (M.T @ M).astype(np.int64)
        assert (P == ref).all(), name
    ctx.set_pair_kernel(Context.PAIR_FP4, 0)
    assert (ctx.pair_matrix() == ref).all()


def test_pair_kernel_choice_keeps_mining_results(ctx):
    say("case: mining workload1 seed 2 at 2% with each pair kernel")
    T = gen.workload1(2)
    thr = gen.min_count(0.02, T.n_txn)
    ctx.load_csr(T.offsets, T.items)
    got = []
    for kind, _ in KINDS:
        ctx.set_pair_kernel(kind, 0)
        got.append(ctx.mine(thr).as_dict())
    ctx.set_pair_kernel(Context.PAIR_FP4, 0)
    assert got[0] == got[1] and len(got[0]) > 0


def test_pair_kernel_rejects_unknown_kind(ctx):
    with pytest.raises(Exception):
        ctx.set_pair_kernel(7, 0)
    with pytest.raises(Exception):
        ctx.set_pair_kernel(Context.PAIR_FP4, -1)


def test_mfma_probe_reports_a_plausible_rate(ctx):
    say("case: MFMA probe fp4 and int8")
    f4 = ctx.mfma_probe(Context.PAIR_FP4, 200)
    i8 = ctx.mfma_probe(Context.PAIR_INT8, 200)
    # dense peaks 10.1 PFLOP/s (fp4) and 5.0 POP/s (int8); the sustained rate is below them
    assert 500.0 < f4 < 10100.0 * 1.05
    assert 250.0 < i8 < 5000.0 * 1.05
    with pytest.raises(Exception):
        ctx.mfma_probe(7, 10)
