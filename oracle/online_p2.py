"""Python restatement of the reference's online P^2 discretiser -- TEST INFRASTRUCTURE ONLY.

Only tests import this module; the product path (the C++ routine behind
sgfim_online_items) never does.

Restates code/pipeline/online_discretize.py:
  P2Quantile (:22-71)      five markers; the first five samples initialise them (sorted
                           heights, positions 0, 1 + 2q, 1 + 4q, 3 + 2q, 4); each later
                           sample finds its cell k, bumps the extremes, adds 1 to every
                           position and to positions 1..k, then moves markers 1..3 by
                           qp = (d + 1)(h[i+1] - h[i-1]) / span when d >= 1, span > 1 and
                           the move stays strictly between the neighbours; the estimate is
                           the middle marker (the median of the samples before five)
  OnlineBinner (:74-103)   estimators at i / B, i = 1..B-1; edges: sorted estimates q,
                           span = max(q[-1] - q[0], 1e-9), minimum gap 0.05 span (B = 3:
                           the first two edges re-centred on their midpoint; else pushed up
                           one after another), then [q0 - span] + q + [q[-1] + span]
  bin_index (:106-108)     searchsorted right - 1, clamped to [0, B - 1]
  run_file (:111-131)      per window [s, s + window), s = 0, stride, ...: every sample of
                           the window is added (overlapping windows add samples again), then
                           the window median is binned with the current edges
Items are numbered s * B + b + 1 for sensor s, bin b.
"""
from __future__ import annotations

import numpy as np


class P2Quantile:
    def __init__(self, q):
        self.q = q
        self.init = []
        self.n = None
        self.h = None

    def add(self, x):
        if self.n is None:
            self.init.append(x)
            if len(self.init) == 5:
                h = sorted(self.init)
                self.h = h[:5]
                self.n = [0.0, 1 + 2 * self.q, 1 + 4 * self.q, 3 + 2 * self.q, 4.0]
            return
        h, n = self.h, self.n
        if x < h[0]:
            h[0] = x
            k = 0
        elif x < h[1]:
            k = 0
        elif x < h[2]:
            k = 1
        elif x < h[3]:
            k = 2
        elif x <= h[4]:
            k = 3
        else:
            h[4] = x
            k = 3
        for i in range(5):
            n[i] = n[i] + 1
        for i in range(1, 4):
            n[i] += 1 if i <= k else 0
        for i in (1, 2, 3):
            d = n[i] - n[i - 1]
            span = n[i + 1] - n[i - 1]
            if d >= 1 and span > 1:
                qp = (d + 1) * (h[i + 1] - h[i - 1]) / span
                if h[i - 1] < h[i] + qp < h[i + 1]:
                    h[i] += qp

    def value(self):
        if self.n is None:
            return float(np.median(self.init)) if self.init else 0.0
        return float(self.h[2])


def edges_of(est, n_bins):
    q = sorted(e.value() for e in est)
    span = max(q[-1] - q[0], 1e-9)
    gap = 0.05 * span
    if n_bins == 3:
        if q[1] - q[0] < gap:
            mid = (q[0] + q[1]) / 2
            q[0], q[1] = mid - gap / 2, mid + gap / 2
    else:
        for i in range(1, len(q)):
            if q[i] < q[i - 1] + gap:
                q[i] = q[i - 1] + gap
    return np.array([q[0] - span] + q + [q[-1] + span])


def online_bins(x: np.ndarray, n_bins: int = 3, window: int = 30, stride: int = 15):
    """(bins int32 [windows][sensors], edges float64 [windows][sensors][n_bins + 1]) of the
    online discretisation of x[rows][sensors]; item ids are s * n_bins + bins[w, s] + 1."""
    n, S = x.shape
    starts = list(range(0, n - window + 1, stride))
    bins = np.empty((len(starts), S), dtype=np.int32)
    edges = np.empty((len(starts), S, n_bins + 1), dtype=np.float64)
    for s in range(S):
        est = [P2Quantile(i / n_bins) for i in range(1, n_bins)]
        for w, a in enumerate(starts):
            for t in range(a, a + window):
                v = float(x[t, s])
                for e in est:
                    e.add(v)
            med = float(np.median(x[a:a + window, s]))
            ed = edges_of(est, n_bins)
            b = int(np.searchsorted(ed, med, side="right") - 1)
            bins[w, s] = min(max(b, 0), len(ed) - 2)
            edges[w, s] = ed
    return bins, edges


def online_items(x: np.ndarray, n_bins: int = 3, window: int = 30, stride: int = 15) -> np.ndarray:
    """Item ids [windows][sensors]: sensor s, bin b -> s * n_bins + b + 1."""
    bins, _ = online_bins(x, n_bins, window, stride)
    return bins + (np.arange(x.shape[1], dtype=np.int32) * n_bins + 1)[None, :]
