"""NumPy restatement of the reference's sensor transactionisation -- TEST INFRASTRUCTURE ONLY.

Only tests import this module; the product path never does.

Restates code/pipeline/skab_spike.py:
  fit_bins (:37-50)             per sensor, edges = unique(quantile(column,
                                linspace(0, 1, B + 1))); fewer than two edges (a
                                constant column) become [v - 0.5, v + 0.5]
  bin_index (:53-56)            searchsorted(edges, v, side="right") - 1, clamped to
                                [0, len(edges) - 2]
  window_transactions (:60-80)  windows [s, s + window) for s in range(0, n - window + 1,
                                stride); per sensor the bin of the window's median
with the item of sensor s in bin b numbered s * B + b + 1 (the reference names it
"sensor=bin" and numbers names in first-seen order when writing FIMI files).
"""
from __future__ import annotations

import numpy as np


def fit_edges(col: np.ndarray, n_bins: int) -> np.ndarray:
    e = np.unique(np.quantile(col, np.linspace(0, 1, n_bins + 1)))
    if len(e) < 2:
        e = np.array([e[0] - 0.5, e[0] + 0.5])
    return e


def bin_of(v, edges: np.ndarray):
    return np.clip(np.searchsorted(edges, v, side="right") - 1, 0, len(edges) - 2)


def transactions(x: np.ndarray, n_bins: int, window: int = 1, stride: int = 1):
    """CSR (int64 offsets, int32 ids) of the windowed, binned sensor matrix x[rows][sensors]."""
    n, S = x.shape
    starts = np.arange(0, n - window + 1, stride)
    ids = np.empty((len(starts), S), dtype=np.int32)
    for s in range(S):
        e = fit_edges(x[:, s], n_bins)
        med = np.array([np.median(x[a:a + window, s]) for a in starts])
        ids[:, s] = bin_of(med, e) + s * n_bins + 1
    off = np.arange(len(starts) + 1, dtype=np.int64) * S
    return off, ids.reshape(-1)
