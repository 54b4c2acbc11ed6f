"""CPU tests: the online P^2 discretiser restatement (oracle/online_p2.py) against the
reference's own input/output vectors (tests/golden/online_p2.npz, written by
tests/golden/make_online_p2.py) and against what the estimator's definition fixes."""
import os

import numpy as np
import pytest

from oracle import online_p2 as P2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "online_p2.npz"))
CASES = sorted(int(k[2:]) for k in GOLD.files if k.startswith("x_"))


@pytest.mark.parametrize("k", CASES)
def test_oracle_matches_reference_vectors(k):
    x = GOLD[f"x_{k}"]
    B, w, st = (int(v) for v in GOLD[f"p_{k}"])
    bins, edges = P2.online_bins(x, B, w, st)
    assert (bins == GOLD[f"bins_{k}"]).all()
    assert edges.tobytes() == GOLD[f"edges_{k}"].tobytes()


def test_fixture_covers_the_edge_cases():
    params = [tuple(int(v) for v in GOLD[f"p_{k}"]) for k in CASES]
    assert {B for B, _, _ in params} >= {2, 3, 4, 5, 7}         # both edge rules, B = 2
    assert any(w < 5 for _, w, _ in params)                     # estimates before five samples
    assert any(w > st for _, w, st in params)                   # overlapping windows re-add samples
    assert any(w == 64 for _, w, _ in params)                   # largest window
    assert any(len(GOLD[f"bins_{k}"]) == 0 for k in CASES)      # no window fits


def test_fewer_than_five_samples_give_the_exact_median():
    rng = np.random.default_rng(0)
    for m in range(0, 5):
        xs = rng.normal(size=m)
        e = P2.P2Quantile(0.25)
        for v in xs:
            e.add(float(v))
        assert e.value() == (float(np.median(xs)) if m else 0.0)


def test_marker_invariants():
    """After n >= 5 samples: heights sorted, h[0] and h[4] the running minimum and maximum,
    n[0] = n - 5 and n[4] = n - 1 (every sample moves both end markers by one)."""
    rng = np.random.default_rng(1)
    for q in (1 / 3, 0.5, 2 / 3, 0.125):
        e = P2.P2Quantile(q)
        xs = rng.standard_cauchy(400)
        for t, v in enumerate(xs, 1):
            e.add(float(v))
            if t >= 5:
                assert e.h == sorted(e.h)
                assert e.h[0] == xs[:t].min() and e.h[4] == xs[:t].max()
                assert e.n[0] == t - 5 and e.n[4] == t - 1


def test_constant_stream_edges_closed_form():
    """A constant sensor: every estimate is c, span = 1e-9; for three bins the two inner
    edges are re-centred to c -/+ 0.025e-9 and the median lands in the middle bin."""
    c = 2.5
    x = np.full((60, 1), c)
    bins, edges = P2.online_bins(x, 3, 30, 15)
    span, gap = 1e-9, 0.05 * 1e-9
    lo, hi = c - gap / 2, c + gap / 2
    assert (edges[:, 0] == np.array([lo - span, lo, hi, hi + span])).all()
    assert (bins == 1).all()
    # more bins: each edge pushed one gap above the previous one
    bins, edges = P2.online_bins(x, 5, 30, 15)
    q = [c, c + gap, c + gap + gap, c + gap + gap + gap]
    assert (edges[:, 0] == np.array([q[0] - span] + q + [q[-1] + span])).all()
    assert (bins == 1).all()                         # c equals the lowest inner edge: side right puts it in bin 1


def test_bins_are_searchsorted_of_the_median():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(200, 3))
    bins, edges = P2.online_bins(x, 4, 20, 10)
    for w in range(len(bins)):
        for s in range(3):
            med = np.median(x[w * 10:w * 10 + 20, s])
            b = int(np.searchsorted(edges[w, s], med, side="right") - 1)
            assert bins[w, s] == min(max(b, 0), 3)
            assert (np.diff(edges[w, s]) >= 0).all()
