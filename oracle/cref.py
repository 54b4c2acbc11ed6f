"""ctypes wrapper of the C oracle (fim_ref.c) -- test infrastructure only.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfim_oracle.so")
SRC = os.path.join(HERE, "fim_ref.c")

ROOTS_ALL, ROOTS_STRICT, ROOTS_TOPK = 0, 1, 2
COUNT_ALL, COUNT_INHERIT = 0, 1


def build(force: bool = True) -> str:
    """Compile fim_ref.c with gcc -O3 -fopenmp into libfim_oracle.so."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "libfim_oracle.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(force=False)
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.ref_bitmap.argtypes = [P, P, C.c_int64, C.c_int32, C.c_int64, P]
        L.ref_count.argtypes = [P, C.c_int64, P, C.c_int32, P, C.c_int32, P]
        L.ref_row_supports.argtypes = [P, C.c_int32, C.c_int64, P]
        L.ref_pair_matrix.argtypes = [P, C.c_int32, C.c_int64, P]
        L.ref_mine.argtypes = [P, C.c_int32, C.c_int64, C.c_double, C.c_int, C.c_int64,
                               C.c_double, C.c_int32, C.c_int, C.c_int, C.c_int64, P]
        L.ref_mine.restype = P
        L.ref_result_n.argtypes = [P]; L.ref_result_n.restype = C.c_int64
        L.ref_result_nitems.argtypes = [P]; L.ref_result_nitems.restype = C.c_int64
        L.ref_result_copy.argtypes = [P, P, P, P]
        L.ref_result_free.argtypes = [P]
        L.ref_trace_set.argtypes = [P, C.c_int64]
        L.ref_trace_count.restype = C.c_int64
        L.zipf_spec_ref.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_double, C.c_double, C.c_uint64, P, P, P, P]
        L.zipf_csr_ref.argtypes = [C.c_int64, C.c_int32, P, C.c_int32, P, P, P, C.c_uint64, P, P]
        L.zipf_csr_ref.restype = C.c_int64
        L.kat_generate.argtypes = [C.c_int64, P, P, P, P]
        L.kat_nnz.argtypes = [P, C.c_int64]; L.kat_nnz.restype = C.c_int64
        L.kat_csr.argtypes = [P, C.c_int64, P, P]
        L.kat_levelwise.argtypes = [P, C.c_int64, C.c_int64, P, P, C.c_int64]
        L.kat_levelwise.restype = C.c_int64
        L.zipf_bitmap_ref.argtypes = [C.c_int64, C.c_int32, P, C.c_int32, P, P, P, C.c_uint64, C.c_int64, P]
        L.ref_antichain_violations.argtypes = [P, P, C.c_int64, C.c_int32]
        L.ref_antichain_violations.restype = C.c_int64
        L.ref_simd_path.restype = C.c_int
        L.ref_set_simd.argtypes = [C.c_int]; L.ref_set_simd.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def simd_path() -> str:
    """The counter's inner loop in use: "avx512-vpopcntdq" or "scalar-popcnt"."""
    return "avx512-vpopcntdq" if lib().ref_simd_path() else "scalar-popcnt"


def set_simd(best: bool) -> str:
    """Use the host's best counter path (best=True) or force the scalar popcnt loop."""
    lib().ref_set_simd(1 if best else 0)
    return simd_path()


def bitmap(offsets: np.ndarray, items: np.ndarray, n_items: int) -> np.ndarray:
    """Vertical bitmap rows[id-1, w] (uint64), W = ceil(n_txn / 64)."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    it = np.ascontiguousarray(items, dtype=np.int32)
    n = off.shape[0] - 1
    W = (n + 63) // 64
    rows = np.zeros((n_items, max(W, 0)), dtype=np.uint64)
    lib().ref_bitmap(_p(off), _p(it), n, n_items, W, _p(rows))
    return rows


def count(rows: np.ndarray, prefix, cands) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    pre = np.ascontiguousarray(prefix, dtype=np.int32)
    ca = np.ascontiguousarray(cands, dtype=np.int32)
    out = np.zeros(ca.shape[0], dtype=np.int64)
    lib().ref_count(_p(rows), rows.shape[1], _p(pre), pre.shape[0], _p(ca), ca.shape[0], _p(out))
    return out


def mine_trace(rows: np.ndarray, thr: float, cap: int = 1 << 22, **kw):
    """ref_mine with a node trace: (results, stats, trace [n, 6] = parent, depth, support,
    candidates counted, is MFI, last item); single-threaded."""
    tr = np.zeros((cap, 6), dtype=np.int64)
    lib().ref_trace_set(_p(tr), cap)
    out, st = mine(rows, thr, threads=1, **kw)
    n = lib().ref_trace_count()
    return out, st, tr[:n]


def row_supports(rows: np.ndarray) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    out = np.zeros(rows.shape[0], dtype=np.int64)
    lib().ref_row_supports(_p(rows), rows.shape[0], rows.shape[1], _p(out))
    return out


def pair_matrix(rows: np.ndarray) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    d = rows.shape[0]
    out = np.zeros((d, d), dtype=np.int64)
    lib().ref_pair_matrix(_p(rows), d, rows.shape[1], _p(out))
    return out


def mine(rows: np.ndarray, thr: float, root_mode: int = ROOTS_ALL, n_total: int = 0,
         s: float = 0.0, topk: int = 0, count_mode: int = COUNT_INHERIT,
         threads: int = 0, max_nodes: int = 0):
    """Run the restated engine search over dense rows.

    Returns (list of (tuple(dense ids in insertion order), support), stats dict).
    """
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    d, W = rows.shape if rows.ndim == 2 else (0, 0)
    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    stats = np.zeros(3, dtype=np.int64)
    L = lib()
    h = L.ref_mine(_p(rows), d, W, float(thr), root_mode, n_total, float(s), topk,
                   count_mode, threads, max_nodes, _p(stats))
    try:
        n = L.ref_result_n(h)
        ni = L.ref_result_nitems(h)
        items = np.zeros(max(ni, 1), dtype=np.int32)
        off = np.zeros(n + 1, dtype=np.int64)
        sup = np.zeros(max(n, 1), dtype=np.int64)
        L.ref_result_copy(h, _p(items), _p(off), _p(sup))
    finally:
        L.ref_result_free(h)
    out = [(tuple(int(x) for x in items[off[i]:off[i + 1]]), int(sup[i])) for i in range(n)]
    return out, {"nodes": int(stats[0]), "counts": int(stats[1]), "overflow": int(stats[2])}


# --------------------------------------------------------------------------- batch-prototype KAT
KAT_N = 1 << 23
KAT_MINSUP = int(KAT_N * 0.15 / 100.0)        # 12582 (batch_proto.cu:240)
KAT_LEVELS = [248, 585, 587, 350, 117, 17, 0]  # run_dispatch.txt:5-11
KAT_TOTAL = 1904


def kat_generate(n: int = KAT_N):
    """(rows [512, n/64] uint64, patterns [(sorted item list, freq)])."""
    rows = np.zeros((512, n // 64), dtype=np.uint64)
    pi = np.zeros((80, 6), dtype=np.int32)
    ps = np.zeros(80, dtype=np.int32)
    pf = np.zeros(80, dtype=np.float64)
    lib().kat_generate(n, _p(rows), _p(pi), _p(ps), _p(pf))
    pats = [([int(x) for x in pi[p, :ps[p]]], float(pf[p])) for p in range(80)]
    return rows, pats


def kat_csr(rows: np.ndarray):
    n = rows.shape[1] * 64
    nnz = lib().kat_nnz(_p(rows), n)
    off = np.zeros(n + 1, dtype=np.int64)
    items = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().kat_csr(_p(rows), n, _p(off), _p(items))
    return off, items[:nnz]


def kat_levelwise(rows: np.ndarray, minsup: int = KAT_MINSUP, out_cap: int = 4096):
    """(level counts [1..7], total, list of frequent itemsets as tuples of 0-based items)."""
    n = rows.shape[1] * 64
    lc = np.zeros(8, dtype=np.int64)
    out = np.zeros((out_cap, 8), dtype=np.int32)
    tot = lib().kat_levelwise(_p(rows), n, minsup, _p(lc), _p(out), out_cap)
    sets = [tuple(int(x) for x in out[i, 1:1 + out[i, 0]]) for i in range(min(tot, out_cap))]
    return [int(x) for x in lc[1:8]], int(tot), sets


# --------------------------------------------------------------------------- synthetic Zipf set
def zipf_spec(d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed):
    q = np.zeros(d, dtype=np.float64)
    poff = np.zeros(n_planted + 1, dtype=np.int32)
    pit = np.zeros(max(n_planted * size_hi, 1), dtype=np.int32)
    pr = np.zeros(max(n_planted, 1), dtype=np.float64)
    lib().zipf_spec_ref(d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed,
                        _p(q), _p(poff), _p(pit), _p(pr))
    return q, poff, pit[: poff[-1]], pr[:n_planted]


def zipf_csr(n, d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed):
    q, poff, pit, pr = zipf_spec(d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed)
    pit_c = np.ascontiguousarray(pit if pit.size else np.zeros(1, np.int32))
    pr_c = np.ascontiguousarray(pr if pr.size else np.zeros(1))
    nnz = lib().zipf_csr_ref(n, d, _p(q), n_planted, _p(poff), _p(pit_c), _p(pr_c), seed, None, None)
    off = np.zeros(n + 1, dtype=np.int64)
    items = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().zipf_csr_ref(n, d, _p(q), n_planted, _p(poff), _p(pit_c), _p(pr_c), seed, _p(off), _p(items))
    return off, items[:nnz]


def zipf_bitmap(n, d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed):
    """Vertical bitmap rows [d, ceil(n / 64)] of the synthetic Zipf set (the draws of
    zipf_csr), built with all host threads."""
    q, poff, pit, pr = zipf_spec(d, s, mean_len, n_planted, size_lo, size_hi, rate_lo, rate_hi, seed)
    pit_c = np.ascontiguousarray(pit if pit.size else np.zeros(1, np.int32))
    pr_c = np.ascontiguousarray(pr if pr.size else np.zeros(1))
    W = (n + 63) // 64
    rows = np.empty((d, W), dtype=np.uint64)
    lib().zipf_bitmap_ref(n, d, _p(q), n_planted, _p(poff), _p(pit_c), _p(pr_c), seed, W, _p(rows))
    return rows


def antichain_violations(items: np.ndarray, off: np.ndarray, d: int) -> int:
    """Members of the family (items[off[i]:off[i+1]], ids 0..d-1) that are proper subsets of
    another member; 0 for an antichain."""
    it = np.ascontiguousarray(items, dtype=np.int32)
    of = np.ascontiguousarray(off, dtype=np.int64)
    if it.size == 0:
        it = np.zeros(1, dtype=np.int32)
    return int(lib().ref_antichain_violations(_p(it), _p(of), int(of.shape[0]) - 1, int(d)))
