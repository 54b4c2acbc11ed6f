"""Python host API over the C ABI: one Context per transaction set on one GPU."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _lib as L


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def csr_from_rows(rows: Iterable[Iterable[int]]):
    """CSR arrays (int64 offsets, int32 ids) from an iterable of transactions."""
    off = [0]
    items: list[int] = []
    for r in rows:
        items.extend(int(x) for x in r)
        off.append(len(items))
    return np.asarray(off, dtype=np.int64), np.asarray(items, dtype=np.int32)


@dataclass
class MineResult:
    """MFIs as tuples of original ids in search insertion order, with exact supports."""
    itemsets: list = field(default_factory=list)
    supports: list = field(default_factory=list)
    rounds: list = field(default_factory=list)
    times: list = field(default_factory=list)
    windows: list = field(default_factory=list)
    stats: dict = field(default_factory=dict)
    count: int = 0
    arrays: tuple = ()   # (items int32, off int64, support int64) without per-set tuples

    def as_dict(self) -> dict:
        """{frozenset(ids): support} -- the adapter's comparison form (engine_adapter.py:103-111)."""
        return {frozenset(x): s for x, s in zip(self.itemsets, self.supports)}

    def __len__(self):
        return self.count


def _copy(ptr: C.c_void_p, ctype, n: int, dtype) -> np.ndarray:
    """A new array holding the n values of ctype at ptr (frombuffer on a ctypes array type
    costs about half of np.ctypeslib.as_array per call)."""
    if n <= 0:
        return np.zeros(0, dtype=dtype)
    return np.frombuffer((ctype * n).from_address(ptr.value), dtype=dtype).copy()


def _take_result(h: C.c_void_p, stats: dict | None = None, convert: bool = True) -> MineResult:
    Lb = L.lib()
    items = C.c_void_p()
    off = C.c_void_p()
    sup = C.c_void_p()
    rnd = C.c_void_p()
    n = C.c_int64()
    tms = C.c_void_p()
    win = C.c_void_p()
    try:
        Lb.sgfim_result_get(h, C.byref(items), C.byref(off), C.byref(sup), C.byref(rnd), C.byref(n))
        Lb.sgfim_result_times(h, C.byref(tms))
        Lb.sgfim_result_windows(h, C.byref(win))
        nn = n.value
        res = MineResult(stats=stats or {}, count=nn)
        if nn and not convert:
            o = _copy(off, C.c_int64, nn + 1, np.int64)
            it = _copy(items, C.c_int32, int(o[-1]), np.int32)
            s = _copy(sup, C.c_int64, nn, np.int64)
            res.arrays = (it, o, s)
            res.windows = _copy(win, C.c_int32, nn, np.int32)
            res.times = _copy(tms, C.c_double, nn, np.float64)
        elif nn:
            o = _copy(off, C.c_int64, nn + 1, np.int64)
            it = _copy(items, C.c_int32, int(o[-1]), np.int32)
            s = _copy(sup, C.c_int64, nn, np.int64)
            r = _copy(rnd, C.c_int32, nn, np.int32)
            t = _copy(tms, C.c_double, nn, np.float64)
            res.itemsets = [tuple(int(x) for x in it[o[i]:o[i + 1]]) for i in range(nn)]
            res.supports = [int(x) for x in s]
            res.rounds = [int(x) for x in r]
            res.times = [float(x) for x in t]
            wv = _copy(win, C.c_int32, nn, np.int32)
            res.windows = [int(x) for x in wv]
        return res
    finally:
        Lb.sgfim_result_free(h)


class Context:
    """A device context holding one loaded transaction set."""

    def __init__(self, device: int = 0, mem_cap_bytes: int = 0, batch_nodes: int = 0, chunk_words: int = 0,
                 frontier_cap: int = 0, ring_cap: int = 0):
        self._L = L.lib()
        o = L.Opts(device, mem_cap_bytes, batch_nodes, chunk_words, frontier_cap, ring_cap)
        h = C.c_void_p()
        rc = self._L.sgfim_create(C.byref(o), C.byref(h))
        if rc != 0:
            raise L.SgfimError(rc, f"sgfim_create(device={device}) failed (no gfx950 device?)")
        self._h = h
        self.n_txn = 0
        self.n_items = 0
        self.words = 0

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None):
            self._L.sgfim_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.sgfim_last_error(self._h)
            raise L.SgfimError(rc, f"{what}: {msg.decode() if msg else ''}")

    # ------------------------------------------------------------------ data
    def load_csr(self, offsets: np.ndarray, items: np.ndarray):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        it = np.ascontiguousarray(items, dtype=np.int32)
        n = int(off.shape[0]) - 1
        self._check(self._L.sgfim_load_csr(self._h, _p(off), n, _p(it), int(it.shape[0])), "load_csr")
        nt = C.c_int64()
        ni = C.c_int32()
        w = C.c_int64()
        self._L.sgfim_info(self._h, C.byref(nt), C.byref(ni), C.byref(w))
        self.n_txn, self.n_items, self.words = nt.value, ni.value, w.value

    def load_windows(self, win_off: np.ndarray, offsets: np.ndarray, items: np.ndarray):
        """Load windows: window w = transactions [win_off[w], win_off[w+1]) of the CSR."""
        wo = np.ascontiguousarray(win_off, dtype=np.int64)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        it = np.ascontiguousarray(items, dtype=np.int32)
        n = int(off.shape[0]) - 1
        self._check(self._L.sgfim_load_windows(self._h, int(wo.shape[0]) - 1, _p(wo), _p(off), n, _p(it),
                                               int(it.shape[0])), "load_windows")
        nt = C.c_int64()
        ni = C.c_int32()
        w = C.c_int64()
        self._L.sgfim_info(self._h, C.byref(nt), C.byref(ni), C.byref(w))
        self.n_txn, self.n_items, self.words = nt.value, ni.value, w.value
        self.n_windows = int(wo.shape[0]) - 1

    def load_zipf(self, n_txn: int, n_items: int = 2048, zipf_s: float = 1.1, mean_len: float = 60.0,
                  n_planted: int = 40, size_lo: int = 4, size_hi: int = 12, rate_lo: float = 0.005,
                  rate_hi: float = 0.05, seed: int = 0):
        """Generate the synthetic dense-sparse set on the GPU (see sgfim_load_zipf)."""
        self._check(self._L.sgfim_load_zipf(self._h, int(n_txn), int(n_items), float(zipf_s), float(mean_len),
                                            int(n_planted), int(size_lo), int(size_hi), float(rate_lo),
                                            float(rate_hi), int(seed)), "load_zipf")
        nt = C.c_int64()
        ni = C.c_int32()
        w = C.c_int64()
        self._L.sgfim_info(self._h, C.byref(nt), C.byref(ni), C.byref(w))
        self.n_txn, self.n_items, self.words = nt.value, ni.value, w.value

    def window_supports(self, w: int) -> np.ndarray:
        out = np.zeros(max(self.n_items, 1), dtype=np.int64)
        self._check(self._L.sgfim_window_supports(self._h, int(w), _p(out)), "window_supports")
        return out[: self.n_items]

    def set_window_thresholds(self, min_counts):
        mc = np.ascontiguousarray(min_counts, dtype=np.int64)
        self._check(self._L.sgfim_set_window_thresholds(self._h, _p(mc), int(mc.shape[0])), "set_window_thresholds")

    def load_sensors(self, x: np.ndarray, n_bins: int, window: int = 1, stride: int = 1):
        """Transactionise a sensor matrix [rows][sensors] on the GPU: quantile bins per
        sensor over the whole column, one transaction per window of its medians' bins
        (item of sensor s, bin b = s * n_bins + b + 1)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        if x.ndim != 2:
            raise ValueError("x must be [rows][sensors]")
        self._check(self._L.sgfim_load_sensors(self._h, _p(x), x.shape[0], x.shape[1], int(n_bins), int(window),
                                               int(stride)), "load_sensors")
        self._refresh_info()

    def load_sensors_online(self, x: np.ndarray, n_bins: int = 3, window: int = 30, stride: int = 15,
                            return_edges: bool = False):
        """Transactionise a sensor matrix [rows][sensors] on the GPU with streaming P^2
        quantile bins (each window's median binned with the edges of that moment); the
        items and bitmaps are as for load_sensors. With return_edges, returns the edges used,
        float64 [windows][sensors][n_bins + 1]."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        if x.ndim != 2:
            raise ValueError("x must be [rows][sensors]")
        n, S = x.shape
        edges = None
        if return_edges and n >= window >= 1 and stride >= 1:
            edges = np.empty(((n - window) // stride + 1, S, int(n_bins) + 1), dtype=np.float64)
        self._check(self._L.sgfim_load_sensors_online(self._h, _p(x), n, S, int(n_bins), int(window), int(stride),
                                                      _p(edges) if edges is not None else None),
                    "load_sensors_online")
        self._refresh_info()
        return edges

    def sensor_items(self) -> np.ndarray:
        """Items [windows][sensors] of the latest sensor load (sensor s, bin b -> s * n_bins + b + 1)."""
        n = C.c_int64()
        self._check(self._L.sgfim_sensor_items(self._h, None, 0, C.byref(n)), "sensor_items")
        out = np.empty(n.value, dtype=np.int32)
        self._check(self._L.sgfim_sensor_items(self._h, _p(out), n.value, C.byref(n)), "sensor_items")
        return out

    def _refresh_info(self):
        nt = C.c_int64()
        ni = C.c_int32()
        w = C.c_int64()
        self._L.sgfim_info(self._h, C.byref(nt), C.byref(ni), C.byref(w))
        self.n_txn, self.n_items, self.words = nt.value, ni.value, w.value

    def mine_thresholds(self, min_counts, flags: int = 0, root_topk=None, convert: bool = True) -> MineResult:
        """Exact MFIs of one loaded set at several thresholds in one batched search
        (result.windows tags each set with its threshold's index). With ROOT_TOPK,
        root_topk[i] gives threshold i's number of roots (AnyFIM stages)."""
        mc = np.ascontiguousarray(np.asarray(min_counts, dtype=np.int64))
        tk = None if root_topk is None else np.ascontiguousarray(np.asarray(root_topk, dtype=np.int32))
        self._check(self._L.sgfim_set_thresholds(self._h, _p(mc), _p(tk) if tk is not None else None, len(mc)),
                    "set_thresholds")
        try:
            return self.mine(int(mc.max()) if len(mc) else 1, flags=flags, convert=convert)
        finally:
            self._L.sgfim_set_thresholds(self._h, None, None, 0)

    def mine_windows(self, min_counts, flags: int = 0, convert: bool = True) -> MineResult:
        """Exact MFIs of every loaded window in one search (result.windows tags each set;
        convert=False keeps them as flat arrays, result.arrays)."""
        self.set_window_thresholds(min_counts)
        try:
            return self.mine(1, flags=flags, convert=convert)
        finally:
            self._L.sgfim_set_window_thresholds(self._h, None, 0)

    def load_rows(self, rows: Iterable[Iterable[int]]):
        off, it = csr_from_rows(rows)
        self.load_csr(off, it)

    def item_ids(self) -> np.ndarray:
        out = np.zeros(max(self.n_items, 1), dtype=np.int32)
        self._check(self._L.sgfim_item_ids(self._h, _p(out)), "item_ids")
        return out[: self.n_items]

    def item_supports(self) -> np.ndarray:
        out = np.zeros(max(self.n_items, 1), dtype=np.int64)
        self._check(self._L.sgfim_item_supports(self._h, _p(out)), "item_supports")
        return out[: self.n_items]

    def pair_matrix(self, valu: bool = False) -> np.ndarray:
        d = self.n_items
        out = np.zeros((max(d, 1), max(d, 1)), dtype=np.int32)
        fn = self._L.sgfim_pair_matrix_valu if valu else self._L.sgfim_pair_matrix
        self._check(fn(self._h, _p(out)), "pair_matrix")
        return out[:d, :d]

    def count_batch(self, prefixes: Sequence[Sequence[int]], cands: Sequence[Sequence[int]]) -> list:
        """counts[p][j] = |T(prefixes[p]) & T(cands[p][j])| (original ids)."""
        assert len(prefixes) == len(cands)
        poff = np.zeros(len(prefixes) + 1, dtype=np.int64)
        coff = np.zeros(len(prefixes) + 1, dtype=np.int64)
        for i, (p, c) in enumerate(zip(prefixes, cands)):
            poff[i + 1] = poff[i] + len(p)
            coff[i + 1] = coff[i] + len(c)
        pid = np.asarray([x for p in prefixes for x in p] or [0], dtype=np.int32)
        cid = np.asarray([x for c in cands for x in c] or [0], dtype=np.int32)
        out = np.zeros(max(int(coff[-1]), 1), dtype=np.int32)
        self._check(self._L.sgfim_count_batch(self._h, _p(pid), _p(poff), len(prefixes), _p(cid), _p(coff),
                                              _p(out)), "count_batch")
        return [out[coff[i]:coff[i + 1]].copy() for i in range(len(prefixes))]

    def count_batch_flat(self, prefix_ids: np.ndarray, prefix_off: np.ndarray, cand: np.ndarray,
                         cand_off: np.ndarray) -> np.ndarray:
        pid = np.ascontiguousarray(prefix_ids, dtype=np.int32)
        poff = np.ascontiguousarray(prefix_off, dtype=np.int64)
        cid = np.ascontiguousarray(cand, dtype=np.int32)
        coff = np.ascontiguousarray(cand_off, dtype=np.int64)
        out = np.zeros(max(int(coff[-1]), 1), dtype=np.int32)
        self._check(self._L.sgfim_count_batch(self._h, _p(pid), _p(poff), int(poff.shape[0]) - 1, _p(cid),
                                              _p(coff), _p(out)), "count_batch")
        return out[: int(coff[-1])]

    # ------------------------------------------------------------------ mining
    @staticmethod
    def _args(min_count: int, flags: int = 0, rel_thr: float = 0.0, root_topk: int = 0,
              deadline_s: float = 0.0) -> L.MineArgs:
        return L.MineArgs(int(min_count), int(flags), float(rel_thr), int(root_topk), float(deadline_s))

    def mine(self, min_count: int, flags: int = 0, rel_thr: float = 0.0, root_topk: int = 0,
             deadline_s: float = 0.0, convert: bool = True) -> MineResult:
        """Exact MFIs; convert=False keeps them as flat arrays (result.arrays)."""
        a = self._args(min_count, flags, rel_thr, root_topk, deadline_s)
        h = C.c_void_p()
        self._check(self._L.sgfim_mine(self._h, C.byref(a), C.byref(h)), "mine")
        return _take_result(h, self.last_stats(), convert)

    def pair_time(self) -> float:
        """Device milliseconds of the most recent pair-matrix computation (0 if none)."""
        ms = C.c_double()
        self._check(self._L.sgfim_pair_time(self._h, C.byref(ms)), "pair_time")
        return ms.value

    PAIR_FP4 = 0
    PAIR_INT8 = 1
    COUNT_VALU = 0
    COUNT_F4 = 1

    def set_count_kernel(self, kind: int = 0):
        """Counting kernel of the searches: COUNT_VALU (default, AND + popcount per needed
        cell) or COUNT_F4 (fp4 MFMA 16 x 16 products of children and candidates, exact); the
        results do not depend on it."""
        self._check(self._L.sgfim_set_count_kernel(self._h, int(kind)), "set_count_kernel")

    def set_pair_kernel(self, kind: int = 0, max_splits: int = 0):
        """Pair-matrix kernel: PAIR_FP4 (default, e2m1 MFMA with exact fp32 sums) or PAIR_INT8
        (int8 MFMA), and a cap on its transaction splits (0 = automatic); the matrix does
        not depend on either."""
        self._check(self._L.sgfim_set_pair_kernel(self._h, int(kind), int(max_splits)), "set_pair_kernel")

    def set_tile(self, chunk_words: int = 0, chunk_range: int = 0):
        """Counting-tile shape of later searches (0 = default); results do not depend on it."""
        self._check(self._L.sgfim_set_tile(self._h, int(chunk_words), int(chunk_range)), "set_tile")

    def set_sib_ring(self, words: int = 0):
        """Words of the sibling ring of later searches (0 = default); results do not depend
        on it, a larger ring prunes more counts on deep searches."""
        self._check(self._L.sgfim_set_sib_ring(self._h, int(words)), "set_sib_ring")

    def set_projection(self, factor: int = 4, arena_words: int = 0):
        """Projected blocks of later searches: a record whose support is at most 1/factor of
        its rows' transactions counts its subtree on rows compacted to its transactions
        (factor 0 = off; arena_words 0 = default ring size). Results do not depend on it."""
        self._check(self._L.sgfim_set_projection(self._h, int(factor), int(arena_words)), "set_projection")

    def projection(self) -> tuple:
        """(factor, arena_words) the next search uses (arena_words 0 = default)."""
        f, a = C.c_int32(), C.c_int64()
        self._check(self._L.sgfim_get_projection(self._h, C.byref(f), C.byref(a)), "get_projection")
        return f.value, a.value

    def set_block_rule(self, dense: int = 32, min_children: int = 16):
        """How later searches build projected blocks: one at least 1/dense of its source's
        transactions (dense 0: none) is compacted word by word and needs min_children
        children; sparser ones are gathered bit by bit. Results do not depend on it."""
        self._check(self._L.sgfim_set_block_rule(self._h, int(dense), int(min_children)), "set_block_rule")

    def block_rule(self) -> tuple:
        """(dense, min_children) the next search uses."""
        d, m = C.c_int32(), C.c_int32()
        self._check(self._L.sgfim_get_block_rule(self._h, C.byref(d), C.byref(m)), "get_block_rule")
        return d.value, m.value

    def tile(self) -> tuple:
        """(chunk_words, chunk_range) the next search uses."""
        cw, cr = C.c_int32(), C.c_int32()
        self._check(self._L.sgfim_get_tile(self._h, C.byref(cw), C.byref(cr)), "get_tile")
        return cw.value, cr.value

    def device_info(self) -> tuple:
        """(gcnArchName, compute units) of the context's GPU."""
        buf = C.create_string_buffer(64)
        n = C.c_int32()
        self._check(self._L.sgfim_device_info(self._h, buf, 64, C.byref(n)), "device_info")
        return buf.value.decode(), n.value

    def mfma_probe(self, kind: int = 0, iters: int = 4000) -> float:
        """Sustained MFMA rate in TOP/s with no memory traffic (kind PAIR_FP4 or PAIR_INT8)."""
        g = C.c_double()
        self._check(self._L.sgfim_mfma_probe(self._h, int(kind), int(iters), C.byref(g)), "mfma_probe")
        return g.value

    def copy_probe(self, nbytes: int = 1 << 31, iters: int = 5) -> float:
        """On-device copy bandwidth in GB/s (bytes read + written)."""
        g = C.c_double()
        self._check(self._L.sgfim_copy_probe(self._h, int(nbytes), int(iters), C.byref(g)), "copy_probe")
        return g.value

    STAGES = ("select", "proj_tids", "proj_gather", "proj_compress", "need_rows", "count", "expand", "commit")

    def stage_times(self) -> dict:
        """Device ms of each round stage of the latest search run with NO_GRAPH (zeros otherwise)."""
        out = np.zeros(len(self.STAGES), dtype=np.float64)
        self._check(self._L.sgfim_stage_times(self._h, _p(out), len(out)), "stage_times")
        return dict(zip(self.STAGES, out.tolist()))

    def last_stats(self) -> dict:
        st = L.RoundStats()
        self._L.sgfim_last_stats(self._h, C.byref(st))
        return st.as_dict()

    def anytime_begin(self, min_count: int, flags: int = 0, deadline_s: float = 0.0):
        a = self._args(min_count, flags, 0.0, 0, deadline_s)
        self._check(self._L.sgfim_anytime_begin(self._h, C.byref(a)), "anytime_begin")

    def anytime_step(self, max_rounds: int = 1) -> dict:
        st = L.RoundStats()
        self._check(self._L.sgfim_anytime_step(self._h, int(max_rounds), C.byref(st)), "anytime_step")
        return st.as_dict()

    def anytime_cancel(self):
        self._check(self._L.sgfim_anytime_cancel(self._h), "anytime_cancel")

    def anytime_take(self, convert: bool = True) -> MineResult:
        """MFIs confirmed since the last take (convert=False: flat arrays, result.arrays)."""
        h = C.c_void_p()
        self._check(self._L.sgfim_anytime_take(self._h, C.byref(h)), "anytime_take")
        return _take_result(h, self.last_stats(), convert)

    # ------------------------------------------------------------------ proposal network
    def set_scorer(self, w1: np.ndarray, b1: np.ndarray, w2: np.ndarray, b2: float):
        w1 = np.ascontiguousarray(w1, dtype=np.float32)
        b1 = np.ascontiguousarray(b1, dtype=np.float32)
        w2 = np.ascontiguousarray(w2, dtype=np.float32)
        assert w1.shape == (24, 256) and b1.shape == (256,) and w2.shape == (256,)
        self._check(self._L.sgfim_set_scorer(self._h, _p(w1), _p(b1), _p(w2), float(b2)), "set_scorer")

    def set_trace(self, enable: bool, cap: int = 0):
        self._check(self._L.sgfim_set_trace(self._h, 1 if enable else 0, int(cap)), "set_trace")

    def trace(self):
        """(keys [n,2] = (key, parent key), features [n,24], mfi_keys [m]) of the last traced run."""
        nb = C.c_int64()
        nm = C.c_int64()
        self._check(self._L.sgfim_trace_get(self._h, C.byref(nb), C.byref(nm), None, None, None, 0), "trace_get")
        cap = max(nb.value, nm.value, 1)
        keys = np.zeros((cap, 2), dtype=np.int64)
        feat = np.zeros((cap, 24), dtype=np.float32)
        mk = np.zeros(cap, dtype=np.int64)
        self._check(self._L.sgfim_trace_get(self._h, C.byref(nb), C.byref(nm), _p(keys), _p(feat), _p(mk), cap),
                    "trace_get")
        return keys[: nb.value], feat[: nb.value], mk[: nm.value]
