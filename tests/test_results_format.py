"""CPU tests: Results-file grammar, FIMI parsing and the adapter's helpers, against the
reference engines' own output files kept as fixtures under tests/golden/results/."""
import glob
import os
import re

import numpy as np

from sgfim import engine_adapter as EA

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "results")


def _header_int(path, key):
    for ln in open(path, encoding="utf-8", errors="replace"):
        if ln.startswith(key + ":"):
            return int(ln.split(":")[1])
    raise KeyError(key)


def test_tensorfim_files_parse_to_declared_counts():
    files = [f for f in glob.glob(os.path.join(GOLD, "*=Results.txt")) if "stages" not in f]
    assert files
    for f in files:
        m = EA.parse_tensorfim_results(f)
        assert len(m) == _header_int(f, "MFIsNumber"), f
        # an MFI family is an antichain
        sets = list(m)
        for a in sets:
            assert not any(a < b for b in sets)
    # TensorFIM at s = 1.0 emits nothing (the strict root rule, kernel_bitmap.cu:816)
    assert EA.parse_tensorfim_results(os.path.join(GOLD, "skab_valve2.txt-1.000000=Results.txt")) == {}


def test_anyfim_stage_files_are_consistent():
    for f in glob.glob(os.path.join(GOLD, "*-stages=Results.txt")):
        stages, meta = EA.parse_anyfim_results(f)
        K = len(stages)
        assert meta["MFIsNumPerStage"] == [len(stages[k]) for k in range(1, K + 1)]
        assert meta["itemsNumPerStage"] == [sum(len(x) for x, _ in stages[k]) for k in range(1, K + 1)]
        thr = meta["frequencyThrePerStage"]
        assert all(thr[i] >= thr[i + 1] for i in range(K - 1))
        for k in range(1, K + 1):
            sets = [x for x, _ in stages[k]]
            assert len(set(sets)) == len(sets)
            for x, s in stages[k]:
                assert s >= thr[k - 1]
                assert not any(x < y for y in sets)
            # roots of stage k: an emitted set's largest... every set holds at least one item
            assert all(len(x) >= 1 for x in sets)


def test_writer_reproduces_reference_lines_byte_for_byte():
    """Our line writer, fed each stage pool's (items in file order, frequency), yields the
    reference's exact lines (up to the rank prefix and the order of equal frequencies)."""
    for f in glob.glob(os.path.join(GOLD, "*-stages=Results.txt")):
        n = _header_int(f, "transNum")
        cur = None
        pools = {}
        pat = re.compile(r"^(\d+)th: (\d+) \{ ((?:\d+ )*)\}")
        for ln in open(f, encoding="utf-8"):
            ln = ln.rstrip("\n")
            m = EA.AF_STAGE.match(ln)
            if m:
                cur = int(m.group(1))
                pools[cur] = []
                continue
            m = pat.match(ln)
            if m and cur is not None:
                items = tuple(int(x) for x in m.group(3).split())
                freq = int(ln.rsplit("frequency: ", 1)[1])
                pools[cur].append((items, freq, ln))
        for k, pool in pools.items():
            mine = EA._mfi_lines([(it, fr) for it, fr, _ in pool], n)
            strip = lambda s: s.split("th: ", 1)[1]
            assert sorted(map(strip, mine)) == sorted(strip(ln) for _, _, ln in pool), (f, k)


def test_tensorfim_writer_roundtrip(tmp_path):
    sets = [((4, 7, 13), 863), ((2,), 699), ((10, 4), 577)]
    p = tmp_path / "x.txt-0.358888=Results.txt"
    EA.write_tensorfim_results(str(p), "x.txt", sets, 1187, 0.358888, 0.01, 0.0)
    got = EA.parse_tensorfim_results(str(p))
    assert got == {frozenset(a): b for a, b in sets}
    assert _header_int(str(p), "MFIsNumber") == 3
    txt = p.read_text()
    assert "computation_time(s): 0.01" in txt
    assert EA.TF_COMP.findall(txt) == ["0.01"]


def test_anyfim_writer_roundtrip(tmp_path):
    stages = {1: [((4,), 1186)], 2: [((4,), 1186), ((7,), 1121)], 3: [((7, 4), 1120), ((13,), 914)]}
    p = tmp_path / "x.txt-3-stages=Results.txt"
    EA.write_anyfim_results(str(p), "x.txt", stages, [1186, 1121, 914], [1e-4, 2e-4, 3e-4], 1187, 0.0)
    st, meta = EA.parse_anyfim_results(str(p))
    assert {k: sorted(v, key=lambda t: -t[1]) for k, v in st.items()} == \
        {k: sorted([(frozenset(a), b) for a, b in v], key=lambda t: -t[1]) for k, v in stages.items()}
    assert meta["MFIsNumPerStage"] == [1, 2, 2]
    assert meta["frequencyThrePerStage"] == [1186.0, 1121.0, 914.0]
    assert EA.cumulative_mfis(st)[3] == {frozenset({4}), frozenset({7}), frozenset({4, 7}), frozenset({13})}


def test_fimi_reader_follows_engine_parse_rules(tmp_path):
    p = tmp_path / "t.txt"
    # blank line = empty transaction; negative ids dropped; CRLF tolerated; last line without newline
    p.write_bytes(b"1 2 3\n\n4 -5 6\r\n7\n2 9")
    off, items = EA.read_fimi(str(p))
    assert list(off) == [0, 3, 3, 5, 6, 8]
    assert list(items) == [1, 2, 3, 4, 6, 7, 2, 9]
    q = tmp_path / "u.txt"
    q.write_bytes(b"1\n2\n")
    off, items = EA.read_fimi(str(q))
    assert list(off) == [0, 1, 2]


def test_write_fimi_matches_adapter_encoding(tmp_path):
    txns = [["A=1", "B=0"], ["B=0", "C=2", "A=1"], ["C=2", "C=2"]]
    p = tmp_path / "f.txt"
    mp = EA.write_fimi(txns, str(p), str(tmp_path / "m.json"))
    assert mp == {"A=1": 1, "B=0": 2, "C=2": 3}
    assert p.read_text() == "1 2\n1 2 3\n3\n"
    inv = {v: k for k, v in EA.load_items_map(str(tmp_path / "m.json")).items()}
    assert EA.decode_itemset([1, 3], inv) == frozenset({"A=1", "C=2"})


def test_support_confidence_lift():
    txns = [(1, 2), (1, 2, 3), (2,), (1, 2)]
    labels = [1, 0, 1, 1]
    sup, conf, lift = EA.support_confidence_lift({1, 2}, txns, labels)
    assert sup == 0.75 and abs(conf - 2 / 3) < 1e-12 and abs(lift - (2 / 3) / 0.75) < 1e-12


def test_anyfim_thresholds():
    assert EA.anyfim_thresholds(np.array([5, 9, 9, 2]), 3) == [9, 9, 5]
