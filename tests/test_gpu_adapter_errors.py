"""GPU tests: the drop-in engine adapter against the engines' semantics, and the C ABI's
error paths (every error is forced from the host; no kernel is made to fail)."""
import os

import numpy as np
import pytest

from oracle import mfi_oracle as O
from sgfim import _lib as L
from sgfim import engine_adapter as EA
from sgfim import gen
from sgfim.core import Context, csr_from_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fimi(tmp_path_factory):
    T = gen.workload1(7, 3000)
    rows = T.rows()
    p = tmp_path_factory.mktemp("fimi") / "w1.txt"
    with open(p, "w") as f:
        f.write("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    return str(p), rows


def test_run_tensorfim_matches_engine_family(fimi):
    print("case: adapter run_tensorfim", flush=True)
    path, rows = fimi
    for s in (0.03, 0.1):
        r = EA.run_tensorfim("bmma", path, s, reps=2)
        assert r["ok"] and len(r["comp_times"]) == 2
        assert r["mfis"] == O.tensorfim_family(rows, s)
        rp = f"{path}-{s:.6f}=Results.txt"
        assert EA.parse_tensorfim_results(rp) == r["mfis"]
        ex = EA.run_tensorfim("baseline", path, s, exact=True)
        n = len(rows)
        assert ex["mfis"] == {frozenset(p): c for p, c in O.mine_mfi(rows, min_count=int(np.ceil(n * s)),
                                                                    node_cap=0)[0]}


def test_run_anyfim_matches_stage_families(fimi):
    print("case: adapter run_anyfim", flush=True)
    path, rows = fimi
    r = EA.run_anyfim(path, 12)
    assert r["ok"] and os.path.exists(r["results_path"])
    for k in range(1, 13):
        fam, thr = O.anyfim_stage_family(rows, k)
        assert dict(r["stages"][k]) == fam
        assert r["meta"]["frequencyThrePerStage"][k - 1] == thr
    assert not EA.run_anyfim(path, 41)["ok"]        # more stages than items: the engine refuses
    b = EA.run_anyfim(path, 12, batched=True)        # all stages in one search: same stage sets
    assert b["ok"] and b["stages"] == r["stages"]
    assert b["meta"]["frequencyThrePerStage"] == r["meta"]["frequencyThrePerStage"]


def test_error_paths():
    print("case: error paths", flush=True)
    off, items = csr_from_rows([(1, 2), (2, 3), (1, 2, 3)])
    with Context(device=0) as ctx:
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(1)
        assert e.value.code == L.E_STATE                       # nothing loaded
        with pytest.raises(L.SgfimError) as e:
            ctx.load_csr(np.asarray([0, 2]), np.asarray([1, 0]))
        assert e.value.code == L.E_INVAL                       # id 0
        with pytest.raises(L.SgfimError) as e:
            ctx.load_csr(np.asarray([0, 2, 1]), np.asarray([1, 2]))
        assert e.value.code == L.E_INVAL                       # offsets decrease
        ctx.load_csr(off, items)
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(0)
        assert e.value.code == L.E_INVAL
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(1, flags=L.ORDER_LEARNED)
        assert e.value.code == L.E_STATE                       # no scorer loaded
        with pytest.raises(L.SgfimError) as e:
            ctx.count_batch([list(range(1, 70))], [[1]])
        assert e.value.code == L.E_DEPTH                       # prefix longer than 64 items
    # device-memory cap
    with Context(device=0, mem_cap_bytes=1 << 16) as ctx:
        T = gen.workload1(1, 50_000)
        with pytest.raises(L.SgfimError) as e:
            ctx.load_csr(T.offsets, T.items)
        assert e.value.code == L.E_NOMEM
    # explicit frontier capacity too small for the level-2 nodes
    T = gen.workload1(1, 10_000)
    with Context(device=0, frontier_cap=16) as ctx:
        ctx.load_csr(T.offsets, T.items)
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(gen.min_count(0.02, T.n_txn))
        assert e.value.code == L.E_CAPACITY
    # default capacity: an overflow restarts the search with more room, same result
    with Context(device=0, frontier_cap=0) as a:
        a.load_csr(T.offsets, T.items)
        ref = a.mine(gen.min_count(0.02, T.n_txn)).as_dict()
    assert len(ref) > 0
    # an explicit MFI ring bounds the children per round: smaller rounds, same result;
    # a ring that cannot hold the roots plus one record's children is refused
    with Context(device=0, ring_cap=1024) as a:
        a.load_csr(T.offsets, T.items)
        assert a.mine(gen.min_count(0.02, T.n_txn)).as_dict() == ref
    with Context(device=0, ring_cap=8) as a:
        a.load_csr(T.offsets, T.items)
        with pytest.raises(L.SgfimError) as e:
            a.mine(gen.min_count(0.02, T.n_txn))
        assert e.value.code == L.E_CAPACITY


def test_depth_overflow_and_deadline_and_cancel():
    print("case: depth, deadline, cancel", flush=True)
    # 70 items that always occur together with independent noise items: one MFI of length 70
    rng = np.random.default_rng(0)
    rows = [tuple(range(1, 71)) + tuple(int(x) for x in rng.choice(np.arange(71, 91), 3, replace=False))
            for _ in range(200)]
    off, items = csr_from_rows(rows)
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(150)
        assert e.value.code == L.E_DEPTH
    T = gen.workload1(2, 10_000)
    with Context(device=0, batch_nodes=1) as ctx:
        ctx.load_csr(T.offsets, T.items)
        thr = gen.min_count(0.02, T.n_txn)
        with pytest.raises(L.SgfimError) as e:
            ctx.mine(thr, deadline_s=1e-9)
        assert e.value.code == L.E_TIMEOUT
        ctx.anytime_begin(thr)
        st = ctx.anytime_step(2)
        assert not st["done"]
        got = ctx.anytime_take()
        ctx.anytime_cancel()
        with pytest.raises(L.SgfimError) as e:
            ctx.anytime_step(1)
        assert e.value.code == L.E_CANCELLED
        # every MFI reported before the cancel is final
        full = ctx.mine(thr).as_dict()
        for x, s in zip(got.itemsets, got.supports):
            assert full[frozenset(x)] == s


def test_scoring_batch_matches_cpu(fimi):
    print("case: support/confidence/lift on the GPU", flush=True)
    _, rows = fimi
    rng = np.random.default_rng(4)
    labels = (rng.random(len(rows)) < 0.3).astype(int).tolist()
    items = sorted({x for r in rows for x in r})
    pats = [frozenset(rng.choice(items, size=k, replace=False).tolist()) for k in (1, 2, 2, 3, 4, 5) for _ in range(5)]
    pats += [frozenset(), frozenset({items[0], 999})]            # empty pattern, absent id
    got = EA.support_confidence_lift_batch(pats, rows, labels)
    assert got == [EA.support_confidence_lift(p, rows, labels) for p in pats]


def test_engine_argv_shims(fimi):
    import subprocess
    import sys
    print("case: engine argv shims", flush=True)
    path, rows = fimi
    bindir = os.path.join(os.path.dirname(EA.__file__), "bin")
    rp = f"{path}-{0.05:.6f}=Results.txt"
    if os.path.exists(rp):
        os.remove(rp)
    p = subprocess.run([sys.executable, os.path.join(bindir, "CoParaCG_bmma.exe"), path, "0.05"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 1, p.stderr                            # the engines exit 1 after a run
    assert "computation_time(s):" in p.stdout and "total_time(s):" in p.stdout
    assert EA.parse_tensorfim_results(rp) == O.tensorfim_family(rows, 0.05)
    p = subprocess.run([sys.executable, os.path.join(bindir, "AnytimeMining.exe"), path, "5"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 1, p.stderr
    st, meta = EA.parse_anyfim_results(f"{path}-5-stages=Results.txt")
    for k in range(1, 6):
        assert dict(st[k]) == O.anyfim_stage_family(rows, k)[0]
    p = subprocess.run([sys.executable, os.path.join(bindir, "AnytimeMining.exe"), path, "41"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0                                      # more stages than items: refused, exit 0
    assert not os.path.exists(f"{path}-41-stages=Results.txt")
    p = subprocess.run([sys.executable, os.path.join(bindir, "AnytimeMining.exe"), path + ".missing", "5"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 1                                      # unreadable input: exit 1, no file
    assert not os.path.exists(f"{path}.missing-5-stages=Results.txt")


def test_adapter_failures_return_not_ok(tmp_path):
    """Library failures come back as ok=False with empty results and a 'log', as the
    reference adapter reports an engine that left no Results file (engine_adapter.py:82-99,
    123-129); nothing is raised."""
    print("case: adapter failures -> ok=False", flush=True)
    # 70 items that always occur together: at 150 of 200 one MFI of 70 items, past the
    # longest itemset the search supports (E_DEPTH)
    rng = np.random.default_rng(0)
    rows = [tuple(range(1, 71)) + tuple(int(x) for x in rng.choice(np.arange(71, 91), 3, replace=False))
            for _ in range(200)]
    deep = tmp_path / "deep.txt"
    deep.write_text("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    r = EA.run_tensorfim("bmma", str(deep), 0.75, reps=2)
    assert r["ok"] is False and r["mfis"] == {} and r["comp_times"] == [] and "longer than" in r["log"]
    assert not os.path.exists(f"{deep}-{0.75:.6f}=Results.txt")
    # the exact family at each stage threshold: the 70-item set is an MFI there (E_DEPTH); the
    # engine's stage rule (roots = the first k items, extensions below them) never reaches it
    a = EA.run_anyfim(str(deep), 3, exact=True)
    assert a["ok"] is False and a["stages"] == {} and "longer than" in a["log"]
    assert not os.path.exists(a["results_path"])
    # 4,200 distinct items, every one frequent: more than the 4,096 a search supports (E_INVAL)
    wide = tmp_path / "wide.txt"
    wide.write_text("\n".join(" ".join(str(x) for x in range(1 + (t % 2), 4201, 2)) for t in range(4)) + "\n")
    r = EA.run_tensorfim("baseline", str(wide), 0.25, exact=True)
    assert r["ok"] is False and r["mfis"] == {} and "4096" in r["log"]
    r = EA.run_tensorfim("bmma", str(tmp_path / "absent.txt"), 0.1)
    assert r["ok"] is False and r["mfis"] == {}
    p = subprocess_run_shim("CoParaCG_bmma.exe", str(deep), "0.75")
    assert p.returncode == 1 and not os.path.exists(f"{deep}-{0.75:.6f}=Results.txt")


def subprocess_run_shim(name, *args):
    import subprocess
    import sys
    bindir = os.path.join(os.path.dirname(EA.__file__), "bin")
    return subprocess.run([sys.executable, os.path.join(bindir, name), *args], capture_output=True, text=True,
                          timeout=300)


def test_item_id_zero(tmp_path):
    """Id 0: an ordinary item of the exact family (the Python miner's semantics), and in
    engine mode refused (ok=False) only when it would be a root of the engine's search --
    otherwise the engines' output without item 0 is reproduced (id 0 is never a candidate)."""
    print("case: item id 0", flush=True)
    rng = np.random.default_rng(8)
    # items 1..11 at random, and id 0 in most transactions (the most frequent item)
    rows = [tuple(sorted(set(int(x) for x in rng.choice(np.arange(1, 12), size=4, replace=False)) |
                         ({0} if rng.random() < 0.8 else set()))) for _ in range(300)]
    p = tmp_path / "zero.txt"
    p.write_text("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    n = len(rows)
    ex = EA.run_tensorfim("bmma", str(p), 0.1, exact=True)
    assert ex["ok"]
    assert ex["mfis"] == {frozenset(q): c for q, c in O.mine_mfi(rows, min_count=int(np.ceil(0.1 * n)), node_cap=0)[0]}
    assert any(0 in k for k in ex["mfis"])
    eng = EA.run_tensorfim("bmma", str(p), 0.1)           # id 0 is frequent: a root of the engine
    assert eng["ok"] is False and "id 0" in eng["log"]
    rare = [r if 0 not in r or i % 20 == 0 else tuple(x for x in r if x) for i, r in enumerate(rows)]
    q = tmp_path / "rare0.txt"
    q.write_text("\n".join(" ".join(map(str, r)) for r in rare) + "\n")
    eng = EA.run_tensorfim("bmma", str(q), 0.1)           # id 0 infrequent: reproduced
    without = [tuple(x for x in r if x) for r in rare]
    assert eng["ok"] and eng["mfis"] == O.tensorfim_family(without, 0.1)
    af = EA.run_anyfim(str(p), 3)
    assert af["ok"] is False and "id 0" in af["log"]


@pytest.mark.parametrize("k", [40, 64])
def test_long_itemsets_up_to_the_limit(k):
    """Itemsets of up to 64 items. Closed forms (the oracle's search has no equivalence
    pruning, so it cannot enumerate the 2^k subsets): k items in every one of 200 rows,
    each with 3 of 20 noise items (at most 30 rows each), give the single MFI {1..k} at
    support 200 for any threshold above 30; two overlapping patterns in 100 rows each give
    exactly the two patterns."""
    print(f"case: one {k}-item pattern in every row, and two overlapping patterns", flush=True)
    rng = np.random.default_rng(k)
    rows = [tuple(range(1, k + 1)) + tuple(int(x) for x in rng.choice(np.arange(k + 1, k + 21), 3, replace=False))
            for _ in range(200)]
    off, items = csr_from_rows(rows)
    a = tuple(range(1, 51))
    b = tuple(range(21, 61))
    rows2 = [a + (61 + t % 7,) for t in range(100)] + [b + (68 + t % 7,) for t in range(100)]
    off2, items2 = csr_from_rows(rows2)
    with Context(device=0) as ctx:
        ctx.load_csr(off, items)
        assert ctx.mine(150).as_dict() == {frozenset(range(1, k + 1)): 200}
        ctx.load_csr(off2, items2)
        assert ctx.mine(60).as_dict() == {frozenset(a): 100, frozenset(b): 100}
        # one more item in every row: 65 items, past the limit
        if k == 64:
            rows3 = [tuple(range(1, 66)) for _ in range(50)]
            off3, items3 = csr_from_rows(rows3)
            ctx.load_csr(off3, items3)
            with pytest.raises(L.SgfimError) as e:
                ctx.mine(40)
            assert e.value.code == L.E_DEPTH
