"""GPU test of the allocation-free count batch: through the C ABI, sgfim_count_batch_noalloc
gives exactly what sgfim_count_batch gives, on two calls in a row (the second smaller, so
stale counts left in the shared scratch would show), and refuses sizes past its scratch."""
import ctypes as C

import numpy as np
import pytest

from oracle import cref
from sgfim import _lib as L
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _batch(rng, n_prefix, n_items, max_k, max_c):
    k = rng.integers(0, max_k + 1, size=n_prefix)
    nc = rng.integers(0, max_c + 1, size=n_prefix)
    pid = np.concatenate([rng.choice(np.arange(1, n_items + 1), size=int(x), replace=False) for x in k] +
                         [np.zeros(0, np.int64)]).astype(np.int32)
    cid = rng.integers(1, n_items + 3, size=int(nc.sum())).astype(np.int32)   # ids past the set count 0
    poff = np.concatenate([[0], np.cumsum(k)]).astype(np.int64)
    coff = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
    return pid, poff, cid, coff


def _call(Lb, ctx, fn, pid, poff, cid, coff):
    out = np.full(max(int(coff[-1]), 1), -7, dtype=np.int32)
    rc = getattr(Lb, fn)(ctx._h, _p(pid) if pid.size else None, _p(poff), int(poff.shape[0]) - 1,
                         _p(cid) if cid.size else None, _p(coff), _p(out))
    return rc, out[: int(coff[-1])]


def test_noalloc_count_batch_matches_allocating_entry_point():
    print("case: count_batch_noalloc vs count_batch (workload 1, 30,000 txns)", flush=True)
    T = gen.workload1(6, 30_000)
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    rng = np.random.default_rng(42)
    Lb = L.lib()
    with Context(device=0) as ctx:
        ctx.load_csr(T.offsets, T.items)
        for n_prefix in (3000, 17):                  # the second call is smaller
            pid, poff, cid, coff = _batch(rng, n_prefix, T.n_items, 5, 40)
            rc_a, a = _call(Lb, ctx, "sgfim_count_batch", pid, poff, cid, coff)
            rc_b, b = _call(Lb, ctx, "sgfim_count_batch_noalloc", pid, poff, cid, coff)
            assert rc_a == 0 and rc_b == 0
            assert np.array_equal(a, b)
            for p in range(0, n_prefix, max(1, n_prefix // 50)):   # a sample against the oracle
                pre = pid[poff[p]:poff[p + 1]] - 1
                cs = cid[coff[p]:coff[p + 1]]
                ok = cs <= T.n_items
                want = np.zeros(cs.shape[0], dtype=np.int64)
                if ok.any():
                    want[ok] = cref.count(rows, pre, cs[ok] - 1)
                assert np.array_equal(b[coff[p]:coff[p + 1]], want), p
        # past the scratch: refused with E_INVAL, nothing written
        pid, poff, cid, coff = _batch(rng, 16385, T.n_items, 1, 1)
        rc, _ = _call(Lb, ctx, "sgfim_count_batch_noalloc", pid, poff, cid, coff)
        assert rc == L.E_INVAL
        n = (1 << 20) + 5
        pid = np.array([1], np.int32)
        poff = np.array([0, 1], np.int64)
        cid = np.full(n, 2, np.int32)
        coff = np.array([0, n], np.int64)
        rc, _ = _call(Lb, ctx, "sgfim_count_batch_noalloc", pid, poff, cid, coff)
        assert rc == L.E_INVAL
        # and a call after the refusals still counts
        pid, poff, cid, coff = _batch(rng, 5, T.n_items, 3, 8)
        rc_a, a = _call(Lb, ctx, "sgfim_count_batch", pid, poff, cid, coff)
        rc_b, b = _call(Lb, ctx, "sgfim_count_batch_noalloc", pid, poff, cid, coff)
        assert rc_a == 0 and rc_b == 0 and np.array_equal(a, b)
