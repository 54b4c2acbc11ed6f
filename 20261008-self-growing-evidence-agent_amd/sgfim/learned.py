"""Learned proposal ordering for the anytime frontier.

The proposal network (24 features -> 256 ReLU -> 1 logit, BF16 weights, FP32
accumulation) is trained on the GPU by the library's own forward/backward HIP
kernels (sgfim_mlp_*), on traces the exact miner records: every node born during
a search, with its features, labelled 1 when it is a confirmed MFI or has a
confirmed MFI among its descendants. At search time the network's score orders
the frontier; counts stay exact and the final MFI set is unchanged.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .core import Context

F, HID = 24, 256
NPARAM = F * HID + HID + HID + 1
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_weights.npz")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def init_params(seed: int = 0):
    """He-normal W1, zero b1, scaled-normal w2, zero b2 (float32)."""
    rng = np.random.default_rng(seed)
    W1 = (rng.standard_normal((F, HID)) * np.sqrt(2.0 / F)).astype(np.float32)
    b1 = np.zeros(HID, dtype=np.float32)
    w2 = (rng.standard_normal(HID) * np.sqrt(1.0 / HID)).astype(np.float32)
    return W1, b1, w2, 0.0


class Trainer:
    """Device-resident trainer (sgfim_mlp) for the proposal network."""

    def __init__(self, params=None, device: int = 0, seed: int = 0):
        self._L = L.lib()
        W1, b1, w2, b2 = params if params is not None else init_params(seed)
        W1 = np.ascontiguousarray(W1, np.float32)
        b1 = np.ascontiguousarray(b1, np.float32)
        w2 = np.ascontiguousarray(w2, np.float32)
        h = C.c_void_p()
        rc = self._L.sgfim_mlp_create(device, _p(W1), _p(b1), _p(w2), float(b2), C.byref(h))
        if rc != 0:
            raise L.SgfimError(rc, "sgfim_mlp_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.sgfim_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.sgfim_mlp_last_error(self._h)
            raise L.SgfimError(rc, f"{what}: {msg.decode() if msg else ''}")

    def step(self, X, Y, lr: float = 0.0):
        """(loss, packed gradient float32[NPARAM]) of one fused step; Adam update if lr > 0."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(NPARAM, np.float32)
        self._check(self._L.sgfim_mlp_step(self._h, _p(X), _p(Y), X.shape[0], float(lr), _p(loss), _p(grad)), "step")
        return float(loss[0]), grad

    def set_data(self, X, Y):
        self._X = np.ascontiguousarray(X, np.float32)
        self._Y = np.ascontiguousarray(Y, np.float32)
        self._check(self._L.sgfim_mlp_set_data(self._h, _p(self._X), _p(self._Y), self._X.shape[0]), "set_data")

    def train(self, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0):
        losses = np.zeros(max(epochs, 1), np.float32)
        self._check(self._L.sgfim_mlp_train(self._h, epochs, batch, float(lr), seed, _p(losses)), "train")
        return losses[:epochs]

    def params(self):
        W1 = np.zeros((F, HID), np.float32)
        b1 = np.zeros(HID, np.float32)
        w2 = np.zeros(HID, np.float32)
        b2 = np.zeros(1, np.float32)
        self._check(self._L.sgfim_mlp_get(self._h, _p(W1), _p(b1), _p(w2), _p(b2)), "get")
        return W1, b1, w2, float(b2[0])

    def forward(self, X, probs: bool = False):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(max(X.shape[0], 1), np.float32)
        self._check(self._L.sgfim_mlp_forward(self._h, _p(X), X.shape[0], 1 if probs else 0, _p(out)), "forward")
        return out[: X.shape[0]]


# ---------------------------------------------------------------------------- traces
def trace_labels(keys: np.ndarray, mfi_keys: np.ndarray, mode: str = "ancestor") -> np.ndarray:
    """Label of every born frontier record: 1 iff the round that counted its children
    confirmed an MFI (mfi_keys holds that record's key per MFI), or it is an ancestor
    of such a record.

    keys[i] = (key, parent key) of birth i; keys are unique per run and a record is
    always born after its parent, so one reverse pass propagates labels upward.
    mode="direct" keeps only the first part: 1 iff the record's own round confirmed an
    MFI (what a coverage-first order wants to reach early)."""
    n = keys.shape[0]
    pos = {int(k): i for i, k in enumerate(keys[:, 0])}
    lab = np.zeros(n, dtype=np.float32)
    for k in mfi_keys:
        i = pos.get(int(k))
        if i is not None:
            lab[i] = 1.0
    if mode == "direct":
        return lab
    for i in range(n - 1, -1, -1):
        if lab[i]:
            j = pos.get(int(keys[i, 1]))
            if j is not None:
                lab[j] = 1.0
    return lab


def collect_trace(ctx: Context, min_count: int, cap: int = 1 << 22, mode: str = "ancestor"):
    """Run one traced exact search; return (features, labels, result)."""
    ctx.set_trace(True, cap)
    try:
        res = ctx.mine(min_count)
        keys, feat, mk = ctx.trace()
    finally:
        ctx.set_trace(False)
    if mode == "yield":
        return feat, yield_labels(keys, mk, feat), res
    return feat, trace_labels(keys, mk, mode), res


def yield_labels(keys: np.ndarray, mfi_keys: np.ndarray, feat: np.ndarray) -> np.ndarray:
    """Soft label of every born record: the fraction of its children that its round
    confirms as MFIs, min(1, MFIs / children). Feature 6 is log2(1 + children)."""
    uk, cnt = np.unique(mfi_keys, return_counts=True)
    per = dict(zip(uk.tolist(), cnt.tolist()))
    direct = np.asarray([per.get(int(k), 0) for k in keys[:, 0]], dtype=np.float64)
    children = np.maximum(1.0, np.rint(np.exp2(feat[:, 6].astype(np.float64)) - 1.0))
    return np.minimum(1.0, direct / children).astype(np.float32)


def save_weights(params, path: str = WEIGHTS_PATH, meta: dict | None = None):
    W1, b1, w2, b2 = params
    np.savez(path, W1=W1, b1=b1, w2=w2, b2=np.float32(b2), meta=np.array(str(meta or {})))


def load_weights(path: str = WEIGHTS_PATH):
    z = np.load(path, allow_pickle=False)
    return z["W1"], z["b1"], z["w2"], float(z["b2"])


# ---------------------------------------------------------------------------- coverage
def coverage_curve(times, final_count: int, grid=None):
    """(t, fraction of the final MFI set confirmed by t) from confirmation times."""
    t = np.sort(np.asarray(times, dtype=np.float64))
    if grid is None:
        grid = t
    frac = np.searchsorted(t, grid, side="right") / max(final_count, 1)
    return np.asarray(grid), frac


def time_to_fraction(times, final_count: int, frac: float = 0.95) -> float:
    t = np.sort(np.asarray(times, dtype=np.float64))
    if final_count == 0:
        return 0.0
    k = int(np.ceil(frac * final_count)) - 1
    return float(t[min(max(k, 0), len(t) - 1)])
