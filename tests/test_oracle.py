"""CPU tests: the oracle against brute force, closed forms and the reference's semantics."""
import random

import numpy as np
import pytest

from oracle import cref
from oracle import mfi_oracle as O
from sgfim import gen


def _rand_txns(seed, n, d, p):
    rng = random.Random(seed)
    return [tuple(sorted(i for i in range(1, d + 1) if rng.random() < p)) for _ in range(n)]


def _dense(T_rows, n_items):
    off = np.zeros(len(T_rows) + 1, dtype=np.int64)
    items = []
    for i, r in enumerate(T_rows):
        items.extend(r)
        off[i + 1] = len(items)
    return off, np.asarray(items or [1], dtype=np.int32)[: len(items)]


@pytest.mark.parametrize("seed", range(8))
def test_mine_mfi_equals_brute_force(seed):
    txns = _rand_txns(seed, 60, 9, 0.45)
    for thr in (3, 8, 15):
        got = {frozenset(p): c for p, c in O.mine_mfi(txns, min_count=thr)[0]}
        assert got == O.brute_force_mfi(txns, thr)


def test_threshold_rule():
    # max(1, ceil(min_sup * n)) and max(1, int(min_count)) (skab_spike.py:99-100)
    assert O.threshold(1000, 0.02) == 20
    assert O.threshold(999, 0.02) == 20      # 19.98 -> 20
    assert O.threshold(10, 0.0) == 1
    assert O.threshold(10, min_count=0) == 1
    assert O.threshold(10, min_count=7.9) == 7


def test_node_cap_raises():
    txns = [tuple(range(1, 15))] * 4
    with pytest.raises(RuntimeError):
        O.mine_mfi(txns, min_count=1, node_cap=100)


def test_subsumption_filter_drops_non_maximal_leaves():
    # {1,3} is a leaf of the forward DFS (no forward extension) but {1,2,3} contains it.
    txns = [(1, 2, 3), (1, 2, 3), (1, 3), (2, 3)]
    got = {frozenset(p): c for p, c in O.mine_mfi(txns, min_count=2)[0]}
    assert got == {frozenset({1, 2, 3}): 2}


def test_empty_and_trivial():
    assert O.mine_mfi([], min_count=1)[0] == []
    assert O.mine_mfi([(), ()], min_count=1)[0] == []
    assert O.mine_mfi([(5,)], min_count=1)[0] == [((5,), 1)]


@pytest.mark.parametrize("seed", range(4))
def test_c_oracle_matches_python(seed):
    txns = _rand_txns(100 + seed, 200, 14, 0.35)
    off, items = _dense(txns, 14)
    rows = cref.bitmap(off, items, 14)
    sup = cref.row_supports(rows)
    for thr in (10, 30, 60):
        keep = np.where(sup >= thr)[0]
        for mode in (cref.COUNT_ALL, cref.COUNT_INHERIT):
            out, _ = cref.mine(rows[keep], thr, count_mode=mode, threads=2)
            got = {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}
            ref = {frozenset(p): c for p, c in O.mine_mfi(txns, min_count=thr)[0]}
            assert got == ref


def test_vectorised_counter_equals_scalar():
    """The AVX-512 VPOPCNTDQ counter path and the scalar popcnt path give the same counts
    and the same search (row lengths around the 8- and 32-word vector steps, dense and
    sparse rows); on a host without VPOPCNTDQ both runs take the scalar path."""
    rng = np.random.default_rng(11)
    best = cref.set_simd(True)
    try:
        for W in (1, 7, 8, 9, 31, 32, 33, 95, 257):
            rows = rng.integers(0, 2**63, size=(6, W), dtype=np.uint64)
            rows[3] &= rng.integers(0, 2**63, size=W, dtype=np.uint64)   # sparser rows
            rows[4] = 0
            rows[5] = np.uint64(2**64 - 1)
            prefix, cands = [1, 3], [0, 2, 4, 5]
            cref.set_simd(True)
            v = cref.count(rows, prefix, cands)
            cref.set_simd(False)
            s = cref.count(rows, prefix, cands)
            assert (v == s).all()
            m = np.bitwise_and(rows[1], rows[3])
            ref = [sum(bin(int(x)).count("1") for x in np.bitwise_and(m, rows[c])) for c in cands]
            assert s.tolist() == ref
        T = gen.workload1(5, 5000)
        rows = cref.bitmap(T.offsets, T.items, T.n_items)
        thr = gen.min_count(0.02, T.n_txn)
        cref.set_simd(True)
        a, sa = cref.mine(rows, thr, count_mode=cref.COUNT_ALL, threads=2)
        cref.set_simd(False)
        b, sb = cref.mine(rows, thr, count_mode=cref.COUNT_ALL, threads=2)
        assert sorted(a) == sorted(b) and sa == sb
    finally:
        cref.set_simd(True)
    assert cref.simd_path() == best


def test_c_counter_closed_forms():
    T = gen.workload1(3, 3000)
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    # one item per sensor: each sensor's bin rows partition the transactions
    for s in range(8):
        assert sup[s * 5:(s + 1) * 5].sum() == T.n_txn
    P = cref.pair_matrix(rows)
    assert (P == P.T).all() and (np.diag(P) == sup).all()
    for s in range(8):
        blk = P[s * 5:(s + 1) * 5, s * 5:(s + 1) * 5]
        assert (blk[~np.eye(5, dtype=bool)] == 0).all()
        # sum over the bins of sensor s of C[i][j] equals C[i][i] for i outside s
        others = [i for i in range(40) if i // 5 != s]
        assert (P[others][:, s * 5:(s + 1) * 5].sum(axis=1) == sup[others]).all()
    # the counter against explicit set arithmetic
    txs = [set(r) for r in T.rows()]
    got = cref.count(rows, [0, 7], [12, 20, 33])
    for j, c in enumerate([12, 20, 33]):
        assert got[j] == sum(1 for t in txs if {1, 8, c + 1} <= t)


def test_engine_root_rules_on_tie_cases():
    # TensorFIM: an MFI whose largest item has count exactly n*s is not emitted
    # (root rule freq/n > s, kernel_bitmap.cu:816); at s = 1.0 nothing is emitted
    # (data/fimi/skab_valve2.txt-1.000000=Results.txt reports MFIsNumber: 0).
    txns = [(1, 2)] * 4 + [(1,)] * 2 + [(3,)] * 4
    fam = O.tensorfim_family(txns, 0.4)          # thr = 4; counts 1:6, 2:4, 3:4
    assert fam == {}                             # {1,2} largest item 2 has 4/10 == 0.4, {3} likewise
    assert O.tensorfim_family([(1,), (1,)], 1.0) == {}
    exact = {frozenset(p): c for p, c in O.mine_mfi(txns, min_count=4)[0]}
    assert exact == {frozenset({1, 2}): 4, frozenset({3}): 4}
    # AnyFIM: items tied with the k-th count but ranked after k are not roots
    fam2, thr = O.anyfim_stage_family(txns, 2)   # order 1 (6), 2 (4), 3 (4): roots {1, 2}
    assert thr == 4 and fam2 == {frozenset({1, 2}): 4}


def test_batch_proto_kat_oracle_reproduces_published_levels():
    # code/engine_tensorfim/code/batch_proto/run_dispatch.txt:5-15
    rows, pats = cref.kat_generate()
    levels, total, sets = cref.kat_levelwise(rows)
    assert levels == cref.KAT_LEVELS and total == cref.KAT_TOTAL
    fs = set(sets)
    assert sum(tuple(p) in fs for p, _ in pats) == 80
    # downward closure
    for s in sets:
        if len(s) > 2:
            for i in range(len(s)):
                assert s[:i] + s[i + 1:] in fs


def test_oracle_on_equivalence_heavy_data():
    """Items present everywhere, copied columns and nested tidsets: the DFS oracle, the C
    oracle and brute force agree (these are the cases the device search prunes)."""
    rng = random.Random(9)
    rows = []
    for _ in range(150):
        t = {i for i in range(1, 11) if rng.random() < 0.5}
        t.add(4)
        t.discard(7)
        if 2 in t:
            t.add(7)
        if 5 in t:
            t.add(8)
        rows.append(tuple(sorted(t)))
    off, items = _dense(rows, 10)
    R = cref.bitmap(off, items, 10)
    for thr in (8, 30, 70):
        bf = O.brute_force_mfi(rows, thr)
        assert {frozenset(p): c for p, c in O.mine_mfi(rows, min_count=thr)[0]} == bf
        sup = cref.row_supports(R)
        keep = np.where(sup >= thr)[0]
        out, _ = cref.mine(R[keep], thr, count_mode=cref.COUNT_ALL)
        assert {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out} == bf


def test_zipf_spec_library_matches_oracle():
    """The synthetic Zipf set's parameters: the library's host code and the oracle's
    restatement agree exactly, and q sums to the mean length."""
    for args in [(2048, 1.1, 60.0, 40, 4, 12, 0.005, 0.05, 0), (300, 1.1, 20.0, 7, 3, 6, 0.01, 0.1, 9)]:
        q, poff, pit, pr = gen.zipf_spec(*args)
        q2, poff2, pit2, pr2 = cref.zipf_spec(*args)
        assert np.array_equal(q, q2) and np.array_equal(poff, poff2) and np.array_equal(pit, pit2)
        assert np.array_equal(pr, pr2)
        assert abs(q.sum() - args[2]) < 1e-9 * args[2]
        assert (q <= 1.0).all() and (np.diff(q) <= 0).all()
        assert ((pr >= args[6]) & (pr <= args[7])).all()
        sizes = np.diff(poff)
        assert ((sizes >= args[4]) & (sizes <= args[5])).all()
        for j in range(args[3]):
            s = pit[poff[j]:poff[j + 1]]
            assert len(set(s.tolist())) == len(s) and s.min() >= 1 and s.max() <= args[0]
    # the CSR twin: mean length = target + what the planted sets add (less overlaps)
    off, items = cref.zipf_csr(4000, 300, 1.1, 20.0, 7, 3, 6, 0.01, 0.1, 9)
    q, poff, pit, pr = cref.zipf_spec(300, 1.1, 20.0, 7, 3, 6, 0.01, 0.1, 9)
    extra = float((np.diff(poff) * pr).sum())
    assert 20.0 - 0.5 < np.diff(off).mean() < 20.0 + extra + 0.5


def test_zipf_bitmap_equals_csr_draws():
    """The multithreaded bitmap regenerator draws exactly the CSR generator's set."""
    for n, d, mean, npl, seed in ((1000, 70, 9.0, 3, 1), (4097, 300, 30.0, 8, 5)):
        args = (n, d, 1.1, mean, npl, 3, 8, 0.01, 0.08, seed)
        off, items = cref.zipf_csr(*args)
        assert np.array_equal(cref.zipf_bitmap(*args), cref.bitmap(off, items, d))


def test_antichain_check_against_brute_force():
    rng = np.random.default_rng(3)
    for case in range(40):
        d = int(rng.integers(3, 9))
        fam = {tuple(sorted(rng.choice(d, size=int(rng.integers(1, d + 1)), replace=False).tolist()))
               for _ in range(int(rng.integers(1, 12)))}
        fam = [list(f) for f in fam]
        for f in fam:
            rng.shuffle(f)
        it = np.array([x for f in fam for x in f], np.int32)
        off = np.cumsum([0] + [len(f) for f in fam]).astype(np.int64)
        want = sum(1 for a in fam if any(set(a) < set(b) for b in fam))
        assert cref.antichain_violations(it, off, d) == want, fam
