"""The NumPy restatement of the reference's CPU-vectorised miner (A14) against the exact
MFI oracles (CPU only)."""
import math

import numpy as np
import pytest

from oracle import cpuvec as V
from oracle import mfi_oracle as O
from sgfim import gen


def test_packed_layout_and_blank_lines():
    lines = ["1 3\n", "\n", "3\n", "2 3 1\n"]
    bm, n = V.load_fimi_bitmap(lines)
    assert n == 3                                   # the blank line is skipped
    assert bm.shape == (3, 1)
    # MSB-first bits of transactions 0, 1, 2
    assert bm[:, 0].tolist() == [0b10100000, 0b00100000, 0b11100000]


@pytest.mark.parametrize("seed,rel", [(0, 0.05), (1, 0.1), (2, 0.2)])
def test_cpuvec_equals_exact_family(seed, rel):
    T = gen.workload1(seed, 2000)
    rows = T.rows()
    bm, n = V.bitmap_from_rows(rows, T.n_items)
    mfis, _ = V.mine_mfi_cpuvec(bm, n, rel)
    got = {frozenset(x + 1 for x in m) for m in mfis}
    thr = max(1, math.ceil(rel * n))
    ref = {frozenset(p) for p, _ in O.mine_mfi(rows, min_count=thr, node_cap=0)[0]}
    assert got == ref
    assert len(got) == len(mfis)                    # no duplicates


def test_cpuvec_brute_force_small():
    rng = np.random.default_rng(3)
    rows = [tuple(sorted(int(x) for x in np.nonzero(rng.random(9) < 0.45)[0] + 1)) for _ in range(60)]
    bm, n = V.bitmap_from_rows(rows, 9)
    mfis, _ = V.mine_mfi_cpuvec(bm, n, 0.15)
    got = {frozenset(x + 1 for x in m) for m in mfis}
    ref = set(O.brute_force_mfi(rows, math.ceil(0.15 * n)))
    assert got == ref
