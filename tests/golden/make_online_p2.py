"""Writes tests/golden/online_p2.npz: input/output vectors of the reference's online P^2
discretiser (code/pipeline/online_discretize.py:22-131), produced by importing that module
from the reference checkout. Run in a container that has the reference at REF (it is not
needed afterwards: the fixture holds only data).

Each case k stores
  x_k       float64 [rows][sensors]   seeded synthetic sensor matrix
  p_k       int64   [n_bins, window, stride]
  bins_k    int32   [windows][sensors] bin of each window median under the edges of that moment
  edges_k   float64 [windows][sensors][n_bins + 1] the edges used for it
The windowing follows run_file (:111-131) and the ablation scripts that drive OnlineBinner
with other bin counts: every sample of window [s, s + window) is added, then the window
median is binned. For n_bins = 3 the bins are also checked against run_file itself.
"""
import importlib.util
import os
import sys

import numpy as np
import pandas as pd

REF = "/root/reference/code/pipeline/online_discretize.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "online_p2.npz")


def load_ref():
    spec = importlib.util.spec_from_file_location("ref_online_discretize", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def sensors(rng, n, kinds):
    cols = []
    for kind in kinds:
        if kind == "walk":
            c = np.cumsum(rng.normal(0, 1, n))
        elif kind == "uniform":
            c = rng.uniform(-3, 7, n)
        elif kind == "heavy":
            c = rng.standard_cauchy(n)
        elif kind == "const":
            c = np.full(n, 2.5)
        elif kind == "levels":
            c = rng.integers(0, 4, n).astype(np.float64) * 0.5
        elif kind == "step":
            c = np.where(np.arange(n) < n // 2, 1.0, 9.0) + rng.normal(0, 0.05, n)
        elif kind == "spiky":
            c = rng.normal(10, 0.1, n)
            c[rng.random(n) < 0.03] += 50.0
        elif kind == "trend":
            c = np.linspace(-1, 1, n) ** 3 + rng.normal(0, 1e-3, n)
        else:
            raise ValueError(kind)
        cols.append(c)
    return np.stack(cols, axis=1)


CASES = [
    # n_bins, window, stride, rows, sensor kinds, seed
    (3, 30, 15, 600, ("walk", "uniform", "heavy", "const", "levels", "step", "spiky", "trend"), 1),
    (5, 10, 3, 300, ("walk", "heavy", "levels", "trend"), 2),
    (7, 4, 1, 60, ("uniform", "levels", "const"), 3),
    (2, 7, 7, 100, ("walk", "step"), 4),
    (3, 1, 2, 40, ("uniform", "levels", "const", "heavy"), 5),
    (3, 2, 1, 30, ("walk", "levels"), 6),
    (4, 64, 5, 400, ("walk", "spiky", "trend"), 7),
    (3, 30, 15, 29, ("walk",), 8),          # fewer rows than one window: no windows
]


def run_case(ref, x, n_bins, window, stride):
    n, S = x.shape
    df = pd.DataFrame(x, columns=[f"c{s}" for s in range(S)])
    binners = [ref.OnlineBinner(n_bins) for _ in range(S)]
    bins, edges = [], []
    for s in range(0, n - window + 1, stride):
        e = s + window
        for t in range(s, e):
            for c in range(S):
                binners[c].add(float(df[f"c{c}"].iloc[t]))
        brow, erow = [], []
        for c in range(S):
            ed = binners[c].edges()
            brow.append(ref.bin_index(float(df[f"c{c}"].iloc[s:e].median()), ed))
            erow.append(np.asarray(ed, dtype=np.float64))
        bins.append(brow)
        edges.append(erow)
    bins = np.asarray(bins, dtype=np.int32).reshape(-1, S)
    edges = np.asarray(edges, dtype=np.float64).reshape(-1, S, n_bins + 1)
    if n_bins == 3:
        res = ref.run_file(df, list(df.columns), window, stride)
        rb = np.asarray([[items[c] for c in df.columns] for _, _, items in res], dtype=np.int32).reshape(-1, S)
        assert (rb == bins).all()
    return bins, edges


def main():
    ref = load_ref()
    out = {}
    for k, (B, w, st, n, kinds, seed) in enumerate(CASES):
        x = sensors(np.random.default_rng(seed), n, kinds)
        bins, edges = run_case(ref, x, B, w, st)
        out[f"x_{k}"] = x
        out[f"p_{k}"] = np.asarray([B, w, st], dtype=np.int64)
        out[f"bins_{k}"] = bins
        out[f"edges_{k}"] = edges
        print(f"case {k}: n_bins={B} window={w} stride={st} rows={n} sensors={len(kinds)} windows={len(bins)}")
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
