"""CPU tests of the counting-tile autotuner's host logic: cache keys, choice, JSON cache,
and the order-free result digest it uses to check that a tile shape changes nothing."""
import numpy as np

from sgfim import autotune as A
from sgfim.core import MineResult


def test_shape_key_buckets_by_powers_of_two():
    k = A.shape_key("gfx950", 256, 512, 15625, 3)
    assert k == "gfx950:256cu|items:512|words:16384|sets:4"
    assert A.shape_key("gfx950", 256, 300, 9000, 1) == "gfx950:256cu|items:512|words:16384|sets:1"
    assert A.shape_key("gfx950", 256, 513, 15625, 3) != k


def test_choose_fastest_and_ties_to_first():
    assert A.choose({(256, 4): 2.0, (128, 2): 1.5, (512, 1): 1.7}) == (128, 2)
    assert A.choose({(256, 4): 1.0, (128, 2): 1.0}) == (256, 4)


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "tiles.json")
    assert A.lookup("k", p) is None
    A.save_entry("k", {"chunk_words": 128, "chunk_range": 8, "ms": 1.0}, p)
    A.save_entry("j", {"chunk_words": 512, "chunk_range": 1, "ms": 2.0}, p)
    assert A.lookup("k", p) == (128, 8)
    assert A.lookup("j", p) == (512, 1)
    assert A.lookup("other-device", p) is None


def test_committed_cache_entries_are_valid_shapes():
    for key, e in A.load_cache().items():
        assert (e["chunk_words"], e["chunk_range"]) in A.CANDIDATES, key
        assert e["ms"] <= e["default_ms"] + 1e-9, key      # the choice is the fastest measured
        if "proj_factor" in e:
            assert e["proj_factor"] in A.PROJ_CANDIDATES, key
            assert e["ms"] <= min(e["proj_all"].values()) + 1e-9, key   # tile tuned with the best factor
            base = e["proj_all"][str(e["proj_factor"])]
            if "rule_all" in e:   # rule tuned with the best factor, tile with the best rule
                rule = (e["proj_dense"], e["proj_mintail"])
                assert rule in A.RULE_CANDIDATES, key
                assert e["rule_all"][f"{A.RULE_DEFAULT[0]}/{A.RULE_DEFAULT[1]}"] == base, key
                assert e["ms"] <= min(e["rule_all"].values()) + 1e-9, key
                base = e["rule_all"][f"{rule[0]}/{rule[1]}"]
            assert e["all"][f"{A.DEFAULT[0]}x{A.DEFAULT[1]}"] == base, key


def test_projection_lookup(tmp_path):
    p = str(tmp_path / "tiles.json")
    A.save_entry("k", {"chunk_words": 256, "chunk_range": 4, "proj_factor": 16, "ms": 1.0}, p)
    A.save_entry("old", {"chunk_words": 256, "chunk_range": 4, "ms": 1.0}, p)
    assert A.lookup_projection("k", p) == 16
    assert A.lookup_projection("old", p) is None
    assert A.lookup_projection("absent", p) is None


def test_rule_lookup(tmp_path):
    p = str(tmp_path / "tiles.json")
    A.save_entry("k", {"chunk_words": 256, "chunk_range": 4, "proj_factor": 16, "proj_dense": 8,
                       "proj_mintail": 2, "ms": 1.0}, p)
    A.save_entry("old", {"chunk_words": 256, "chunk_range": 4, "proj_factor": 16, "ms": 1.0}, p)
    assert A.lookup_rule("k", p) == (8, 2)
    assert A.lookup_rule("old", p) is None
    assert A.RULE_DEFAULT in A.RULE_CANDIDATES


def test_signature_is_order_free_and_sensitive():
    a = MineResult(itemsets=[(3, 1, 2), (5,)], supports=[7, 9], windows=[0, 1], count=2)
    b = MineResult(itemsets=[(5,), (1, 2, 3)], supports=[9, 7], windows=[1, 0], count=2)
    assert A.signature(a) == A.signature(b)
    for changed in (MineResult(itemsets=[(5,), (1, 2, 4)], supports=[9, 7], windows=[1, 0], count=2),
                    MineResult(itemsets=[(5,), (1, 2, 3)], supports=[9, 8], windows=[1, 0], count=2),
                    MineResult(itemsets=[(5,), (1, 2, 3)], supports=[9, 7], windows=[0, 0], count=2),
                    MineResult(itemsets=[(5,)], supports=[9], windows=[1], count=1)):
        assert A.signature(changed) != A.signature(a)
    flat = MineResult(arrays=(np.array([1, 2, 3, 5], dtype=np.int32), np.array([0, 3, 4]), np.array([7, 9])),
                      windows=np.array([0, 1]), count=2)
    assert A.signature(flat) == A.signature(a)
