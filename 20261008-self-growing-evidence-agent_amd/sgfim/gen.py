"""Seeded synthetic transaction generators for the five benchmark workloads.

Every generator returns transactions in CSR form: ``offsets`` (int64, length
n_txn + 1) and ``items`` (int32, 1-based item ids, each row sorted ascending and
free of duplicates), which is the in-memory equivalent of one FIMI file
(one transaction per line, whitespace-separated integer ids), the input format
of code/pipeline/engine_adapter.py:37-54 (write_fimi maps items to 1-based ids).

Sensor streams (workloads 1-4) follow the reference pipeline's transactionisation:
a sensor value is discretised into quantile bins with the semantics of
code/pipeline/skab_spike.py:37-57 (np.quantile on linspace(0, 1, B+1) edges,
np.unique, searchsorted(side="right") - 1, clamped), and each transaction holds
exactly one item per sensor (skab_spike.py:61-80), item id = s*B + b + 1.

Choices the workload description leaves open, fixed here once:
  * regime means mu[r, s] ~ N(0, 1), drawn per (regime, sensor);
  * AR(1) noise e_t = phi*e_{t-1} + eps_t with eps_t ~ N(0, 1 - phi^2) and a
    stationary start e_0 ~ N(0, 1), so the noise has unit variance;
  * the regime chain starts uniform, stays with probability ``stay`` and
    otherwise jumps uniformly to one of the other regimes; workloads that do
    not state ``stay`` use 0.98 like workload 1.
All draws come from numpy's PCG64 with the given seed, in a fixed order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Transactions:
    """CSR transactions: row t is items[offsets[t]:offsets[t+1]] (1-based ids)."""
    offsets: np.ndarray  # int64 [n_txn + 1]
    items: np.ndarray    # int32 [nnz]
    n_items: int         # largest item id (ids are 1..n_items)

    @property
    def n_txn(self) -> int:
        return int(self.offsets.shape[0] - 1)

    def rows(self):
        o, it = self.offsets, self.items
        return [tuple(int(x) for x in it[o[t]:o[t + 1]]) for t in range(self.n_txn)]


def _regime_chain(rng: np.random.Generator, n: int, n_regimes: int, stay: float) -> np.ndarray:
    r0 = int(rng.integers(n_regimes))
    if n_regimes == 1:
        return np.zeros(n, dtype=np.int64)
    switch = rng.random(n) >= stay
    switch[0] = False
    off = rng.integers(1, n_regimes, size=n)
    steps = np.where(switch, off, 0)
    return (r0 + np.cumsum(steps)) % n_regimes


def _ar1(rng: np.random.Generator, n: int, n_sensors: int, phi: float) -> np.ndarray:
    from scipy.signal import lfilter
    eps = rng.standard_normal((n, n_sensors)) * math.sqrt(1.0 - phi * phi)
    e0 = rng.standard_normal(n_sensors)
    # e_t = phi*e_{t-1} + eps_t, with e_{-1} chosen so that e_0 = phi*e_{-1}+eps_0
    # equals the stationary draw e0 (zi = phi*e_{-1} = e0 - eps_0).
    zi = (e0 - eps[0])[None, :]
    e, _ = lfilter([1.0], [1.0, -phi], eps, axis=0, zi=zi)
    return e


def quantile_bins(x: np.ndarray, n_bins: int) -> np.ndarray:
    """Bin index per value with fit_bins/bin_index semantics (skab_spike.py:37-57)."""
    qs = np.linspace(0, 1, n_bins + 1)
    e = np.unique(np.quantile(x, qs))
    if len(e) < 2:
        e = np.array([e[0] - 0.5, e[0] + 0.5])
    idx = np.searchsorted(e, x, side="right") - 1
    return np.clip(idx, 0, len(e) - 2)


def sensor_stream(seed: int, n_txn: int, n_sensors: int, n_bins: int, n_regimes: int,
                  phi: float = 0.9, stay: float = 0.98) -> Transactions:
    """One regime-switching AR(1) sensor stream, discretised to one item per sensor."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mu = rng.standard_normal((n_regimes, n_sensors))
    reg = _regime_chain(rng, n_txn, n_regimes, stay)
    x = mu[reg] + _ar1(rng, n_txn, n_sensors, phi)
    ids = np.empty((n_txn, n_sensors), dtype=np.int32)
    for s in range(n_sensors):
        ids[:, s] = quantile_bins(x[:, s], n_bins) + s * n_bins + 1
    offsets = np.arange(n_txn + 1, dtype=np.int64) * n_sensors
    return Transactions(offsets, ids.reshape(-1), n_sensors * n_bins)


# --------------------------------------------------------------------------- workloads
def workload1(seed: int = 0, n_txn: int = 10_000) -> Transactions:
    """10,000 txns, 8 sensors x 5 bins = 40 items, 3 regimes, stay 0.98."""
    return sensor_stream(seed, n_txn, 8, 5, 3)


def workload2(seed: int = 0, n_txn: int = 1_000_000) -> Transactions:
    """1,000,000 txns, 64 sensors x 8 bins = 512 items, 6 regimes."""
    return sensor_stream(seed, n_txn, 64, 8, 6)


def window_lengths(seed: int, n_windows: int = 512, lo: int = 2_000, hi: int = 200_000) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=n_windows))).astype(np.int64)


def workload3(seed: int = 0, n_windows: int = 512, lo: int = 2_000, hi: int = 200_000):
    """Independent windows, lengths log-uniform in [lo, hi]; 32 sensors x 6 bins; 4 regimes.

    Window w uses stream seed ``seed * 100_003 + w + 1``.
    """
    lens = window_lengths(seed, n_windows, lo, hi)
    return [sensor_stream(seed * 100_003 + w + 1, int(n), 32, 6, 4) for w, n in enumerate(lens)]


def _workload3_window(args):
    seed, w, n = args
    T = sensor_stream(seed * 100_003 + w + 1, int(n), 32, 6, 4)
    return T.offsets, T.items


def workload3_csr(seed: int = 0, n_windows: int = 512, workers: int = 16):
    """Workload 3 as one CSR: (win_off, txn_off, items, lengths), windows generated in a
    process pool (window w as in workload3)."""
    from concurrent.futures import ProcessPoolExecutor
    lens = window_lengths(seed, n_windows)
    args = [(seed, w, n) for w, n in enumerate(lens)]
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            parts = list(ex.map(_workload3_window, args))
    else:
        parts = [_workload3_window(a) for a in args]
    win_off = np.zeros(len(parts) + 1, dtype=np.int64)
    offs, its, base = [np.zeros(1, dtype=np.int64)], [], 0
    for w, (o, it) in enumerate(parts):
        win_off[w + 1] = win_off[w] + o.shape[0] - 1
        offs.append(o[1:] + base)
        its.append(it)
        base += it.shape[0]
    return win_off, np.concatenate(offs), np.concatenate(its), lens


def workload4(seed: int, n_txn: int = 500_000) -> Transactions:
    """500,000 txns, 256 items shaped like workload 1 (3 regimes, N(0,1) means, phi 0.9,
    stay 0.98): 64 sensors x 4 bins. (32 x 8 bins would also give 256 items, but at 2%
    its frequent lattice is two levels deep -- pairs -- so no search order differs.)"""
    return sensor_stream(seed, n_txn, 64, 4, 3)


def min_count(rel: float, n: int) -> int:
    """Absolute threshold of the reference Python miner: max(1, ceil(rel*n)) (skab_spike.py:99-100)."""
    return max(1, int(np.ceil(rel * n)))


def concat_windows(windows):
    """(win_off, offsets, items, n_items) of a list of Transactions laid end to end."""
    win_off = np.zeros(len(windows) + 1, dtype=np.int64)
    offs, its = [np.zeros(1, dtype=np.int64)], []
    base = 0
    n_items = 0
    for w, T in enumerate(windows):
        win_off[w + 1] = win_off[w] + T.n_txn
        offs.append(T.offsets[1:] + base)
        its.append(T.items)
        base += int(T.items.shape[0])
        n_items = max(n_items, T.n_items)
    items = np.concatenate(its) if its else np.zeros(0, dtype=np.int32)
    return win_off, np.concatenate(offs), items.astype(np.int32), n_items


WORKLOAD5 = dict(n_txn=10_000_000, n_items=2048, zipf_s=1.1, mean_len=60.0, n_planted=40, size_lo=4,
                 size_hi=12, rate_lo=0.005, rate_hi=0.05)


def zipf_spec(n_items=2048, zipf_s=1.1, mean_len=60.0, n_planted=40, size_lo=4, size_hi=12,
              rate_lo=0.005, rate_hi=0.05, seed=0):
    """(q, planted_off, planted_items, planted_rate) of the synthetic Zipf set, from the library
    (host code: item probabilities and planted itemsets; the bitmaps are drawn on the GPU)."""
    import ctypes as C
    from . import _lib
    q = np.zeros(n_items, dtype=np.float64)
    poff = np.zeros(n_planted + 1, dtype=np.int32)
    pit = np.zeros(max(n_planted * size_hi, 1), dtype=np.int32)
    pr = np.zeros(max(n_planted, 1), dtype=np.float64)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = _lib.lib().sgfim_zipf_spec(n_items, zipf_s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi,
                                    seed, P(q), P(poff), P(pit), P(pr))
    if rc != 0:
        raise ValueError("bad synthetic set parameters")
    return q, poff, pit[: poff[-1]], pr[:n_planted]
