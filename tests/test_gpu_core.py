"""GPU parity tests of the HIP path against the oracle (bit-exact integer results)."""
import random
import sys

import numpy as np
import pytest

from oracle import cref
from oracle import mfi_oracle as O
from sgfim import gen
from sgfim._lib import NO_GRAPH, ROOT_STRICT, ROOT_TOPK
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = Context(device=0)
    yield c
    c.close()


def _oracle_mfis(T, thr, **kw):
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, st = cref.mine(rows[keep], thr, **kw)
    return {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}


def test_item_supports_and_bitmap(ctx):
    say("case: item supports, workload1 seed 1 n=10000")
    T = gen.workload1(1)
    ctx.load_csr(T.offsets, T.items)
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    assert ctx.n_txn == T.n_txn and ctx.n_items == 40
    assert (ctx.item_ids() == np.arange(1, 41)).all()
    assert (ctx.item_supports() == cref.row_supports(rows)).all()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 70001])
def test_count_batch_random(ctx, n):
    say(f"case: count_batch random n={n}")
    rng = random.Random(n)
    d = 37
    txns = [tuple(sorted(rng.sample(range(1, d + 1), rng.randint(0, 9)))) for _ in range(n)]
    off = np.zeros(n + 1, dtype=np.int64)
    flat = []
    for i, t in enumerate(txns):
        flat.extend(t)
        off[i + 1] = len(flat)
    items = np.asarray(flat, dtype=np.int32)
    ctx.load_csr(off, items)
    rows = cref.bitmap(off, items, d)
    ids = list(ctx.item_ids())
    prefixes, cands = [], []
    for _ in range(50):
        k = rng.randint(0, 4)
        prefixes.append(rng.sample(ids, k) if len(ids) >= k else [])
        cands.append(rng.sample(ids, min(len(ids), rng.randint(0, 20))))
    got = ctx.count_batch(prefixes, cands)
    for p, c, g in zip(prefixes, cands, got):
        ref = cref.count(rows, [x - 1 for x in p], [x - 1 for x in c])
        assert list(g) == list(ref)


def test_count_batch_absent_ids(ctx):
    say("case: count_batch absent ids")
    off = np.asarray([0, 2, 3, 5], dtype=np.int64)
    items = np.asarray([1, 3, 3, 1, 3], dtype=np.int32)
    ctx.load_csr(off, items)
    got = ctx.count_batch([[1], [2], []], [[3, 2, 1], [1, 3], [1, 3, 7]])
    assert [list(g) for g in got] == [[2, 0, 2], [0, 0], [2, 3, 0]]


@pytest.mark.parametrize("seed,n", [(0, 10_000), (1, 200_000)])
def test_pair_matrix_mfma_vs_valu_vs_oracle(ctx, seed, n):
    say(f"case: pair matrix workload1-shape seed={seed} n={n}")
    T = gen.workload1(seed, n)
    ctx.load_csr(T.offsets, T.items)
    P = ctx.pair_matrix()
    V = ctx.pair_matrix(valu=True)
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    R = cref.pair_matrix(rows)
    assert (V == R).all()
    assert (P == R).all()


def test_pair_matrix_ragged_tiles_and_timing(ctx):
    say("case: pair matrix 700 items (6 x 6 tiles of 128, partial last tile), 20011 txns")
    rng = np.random.default_rng(11)
    n, d = 20011, 700
    dense = rng.random((n, d)) < 0.05
    rows_l = [np.nonzero(r)[0] + 1 for r in dense]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(r) for r in rows_l])
    items = np.concatenate(rows_l).astype(np.int32)
    ctx.load_csr(off, items)
    ids = ctx.item_ids()
    P = ctx.pair_matrix()
    assert ctx.pair_time() > 0.0
    D = dense[:, ids - 1].astype(np.float64)   # exact: counts < 2^53
    assert (P == (D.T @ D).astype(np.int64)).all()
    assert (P == P.T).all()


def test_pair_matrix_sparse_random(ctx):
    say("case: pair matrix sparse random 300 items")
    rng = np.random.default_rng(5)
    n, d = 5000, 300
    dense = rng.random((n, d)) < 0.03
    rows_l = [np.nonzero(r)[0] + 1 for r in dense]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(r) for r in rows_l])
    items = np.concatenate(rows_l).astype(np.int32)
    ctx.load_csr(off, items)
    ids = ctx.item_ids()
    P = ctx.pair_matrix()
    ref = dense.T.astype(np.int64) @ dense.astype(np.int64)
    assert (P == ref[np.ix_(ids - 1, ids - 1)]).all()


@pytest.mark.parametrize("rel", [0.02, 0.05, 0.1])
def test_mine_workload1_matches_oracles(ctx, rel):
    say(f"case: mine workload1 seed 0 rel={rel}")
    T = gen.workload1(0)
    thr = gen.min_count(rel, T.n_txn)
    ctx.load_csr(T.offsets, T.items)
    got = ctx.mine(thr)
    ref_c = _oracle_mfis(T, thr)
    ref_py = {frozenset(p): c for p, c in O.mine_mfi(T.rows(), min_count=thr, node_cap=0)[0]}
    assert ref_c == ref_py
    assert got.as_dict() == ref_c
    assert len(got.itemsets) == len(set(map(frozenset, got.itemsets)))


def test_mine_graph_and_direct_launch_agree(ctx):
    say("case: mine graph vs direct")
    T = gen.workload1(2, 20_000)
    thr = gen.min_count(0.02, T.n_txn)
    ctx.load_csr(T.offsets, T.items)
    a = ctx.mine(thr).as_dict()
    b = ctx.mine(thr, flags=NO_GRAPH).as_dict()
    assert a == b == _oracle_mfis(T, thr)


@pytest.mark.parametrize("batch", [1, 7, 64])
def test_mine_small_batches(batch):
    say(f"case: mine batch_nodes={batch}")
    T = gen.workload1(4, 5000)
    thr = gen.min_count(0.02, T.n_txn)
    with Context(device=0, batch_nodes=batch) as c:
        c.load_csr(T.offsets, T.items)
        assert c.mine(thr).as_dict() == _oracle_mfis(T, thr)


def test_engine_root_rules(ctx):
    say("case: engine root rules workload1 seed 5")
    T = gen.workload1(5, 3000)
    ctx.load_csr(T.offsets, T.items)
    txns = T.rows()
    for s in (0.03, 0.05, 0.1):
        thr = int(np.ceil(T.n_txn * s))
        got = ctx.mine(thr, flags=ROOT_STRICT, rel_thr=s).as_dict()
        assert got == O.tensorfim_family(txns, s)
    for k in (1, 5, 17, 40):
        fam, thr = O.anyfim_stage_family(txns, k)
        got = ctx.mine(thr, flags=ROOT_TOPK, root_topk=k).as_dict()
        assert got == fam


def test_mine_thresholds_batched_equals_separate_runs(ctx):
    say("case: batched thresholds vs separate runs, workload1 seed 6 n=8000")
    T = gen.workload1(6, 8000)
    ctx.load_csr(T.offsets, T.items)
    thrs = [gen.min_count(s, T.n_txn) for s in (0.1, 0.05, 0.02, 0.2)]
    res = ctx.mine_thresholds(thrs)
    got = [dict() for _ in thrs]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    for i, thr in enumerate(thrs):
        assert got[i] == ctx.mine(thr).as_dict() == _oracle_mfis(T, thr), i
    # AnyFIM stages: threshold k = k-th item count, top-k roots
    txns = T.rows()
    fams = [O.anyfim_stage_family(txns, k) for k in range(1, 13)]
    res = ctx.mine_thresholds([t for _, t in fams], flags=ROOT_TOPK, root_topk=list(range(1, 13)))
    got = [dict() for _ in fams]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    assert got == [f for f, _ in fams]


def test_empty_and_degenerate_inputs(ctx):
    say("case: degenerate inputs")
    ctx.load_csr(np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int32))
    assert len(ctx.mine(1)) == 0
    ctx.load_csr(np.asarray([0, 0, 0], dtype=np.int64), np.zeros(0, dtype=np.int32))
    assert len(ctx.mine(1)) == 0
    ctx.load_csr(np.asarray([0, 1, 2], dtype=np.int64), np.asarray([4, 4], dtype=np.int32))
    assert ctx.mine(1).as_dict() == {frozenset({4}): 2}
    assert ctx.mine(3).as_dict() == {}


def test_fresh_pair_flag_recomputes_in_stream_order(ctx):
    say("case: FRESH_PAIR, workload1 seed 7 n=30000")
    from sgfim._lib import FRESH_PAIR
    T = gen.workload1(7, 30_000)
    ctx.load_csr(T.offsets, T.items)
    thr = gen.min_count(0.02, T.n_txn)
    ref = _oracle_mfis(T, thr)
    assert ctx.mine(thr).as_dict() == ref
    assert ctx.pair_time() > 0.0                   # the first run computed the pair matrix
    for _ in range(2):
        assert ctx.mine(thr, flags=FRESH_PAIR).as_dict() == ref
        assert ctx.pair_time() > 0.0
    res = ctx.mine_thresholds([thr, 2 * thr], flags=FRESH_PAIR)
    got = [dict(), dict()]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    assert got[0] == ref and got[1] == _oracle_mfis(T, 2 * thr)
