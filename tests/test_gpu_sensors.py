"""GPU tests of the sensor transactionisation front-ends: global quantile bins against their
NumPy restatement, online P^2 bins against the reference's own vectors and the oracle."""
import os

import numpy as np
import pytest

from oracle import online_p2 as P2
from oracle import transact as X
from sgfim import _lib as L
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _matrix(seed, n, S):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal((n, S)), axis=0)       # random walks
    x[:, 1] = np.round(x[:, 1])                                # heavy ties
    x[:, 2] = 7.25                                             # constant column
    x[::97, 3] = -0.0                                          # signed zeros
    return x


@pytest.mark.parametrize("seed,n,S,B,window,stride", [(0, 5000, 6, 5, 30, 15), (1, 4099, 5, 3, 1, 1),
                                                      (2, 3000, 8, 8, 7, 3), (3, 640, 4, 6, 64, 64),
                                                      # one window per row bins every value, so
                                                      # each order statistic behind an edge shows:
                                                      # 302 ranks (more key groups than LDS holds),
                                                      # many workgroups per column, a single row
                                                      (4, 20000, 4, 150, 1, 1), (5, 300000, 4, 8, 1, 1),
                                                      (6, 1, 4, 4, 1, 1)])
def test_sensor_front_end_matches_numpy(seed, n, S, B, window, stride):
    print(f"case: sensors seed={seed} n={n} S={S} B={B} window={window} stride={stride}", flush=True)
    x = _matrix(seed, n, S)
    off, ids = X.transactions(x, B, window, stride)
    with Context(device=0) as a, Context(device=0) as b:
        a.load_sensors(x, B, window, stride)
        b.load_csr(off, ids)
        assert a.n_txn == b.n_txn == len(off) - 1
        assert (a.item_ids() == b.item_ids()).all()
        assert (a.item_supports() == b.item_supports()).all()
        thr = gen.min_count(0.05, a.n_txn)
        assert a.mine(thr).as_dict() == b.mine(thr).as_dict()
        assert (a.sensor_items() == np.asarray(ids, dtype=np.int32)).all()


GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "online_p2.npz"))
GOLD_CASES = sorted(int(k[2:]) for k in GOLD.files if k.startswith("x_"))


@pytest.mark.parametrize("k", GOLD_CASES)
def test_online_p2_matches_reference_vectors(k):
    x = GOLD[f"x_{k}"]
    B, w, st = (int(v) for v in GOLD[f"p_{k}"])
    print(f"case: online P2 fixture {k} rows={len(x)} sensors={x.shape[1]} n_bins={B} window={w} stride={st}",
          flush=True)
    with Context(device=0) as c:
        if len(x) < w:                                     # no window fits: refused
            with pytest.raises(L.SgfimError) as e:
                c.load_sensors_online(x, B, w, st)
            assert e.value.code == L.E_INVAL
            return
        edges = c.load_sensors_online(x, B, w, st, return_edges=True)
        assert edges.tobytes() == GOLD[f"edges_{k}"].tobytes()     # bit for bit
        bins = GOLD[f"bins_{k}"]
        items = bins + (np.arange(x.shape[1], dtype=np.int32) * B + 1)[None, :]
        assert (c.sensor_items() == items.ravel()).all()
        assert c.n_txn == len(bins)


@pytest.mark.parametrize("seed,n,S,B,window,stride", [(10, 3000, 12, 3, 30, 15), (11, 2000, 7, 6, 20, 4)])
def test_online_p2_matches_oracle_and_mines(seed, n, S, B, window, stride):
    print(f"case: online P2 seed={seed} n={n} S={S} B={B} window={window} stride={stride}", flush=True)
    x = _matrix(seed, n, S)
    bins, edges = P2.online_bins(x, B, window, stride)
    items = bins + (np.arange(S, dtype=np.int32) * B + 1)[None, :]
    off = np.arange(len(items) + 1, dtype=np.int64) * S
    with Context(device=0) as a, Context(device=0) as b:
        got = a.load_sensors_online(x, B, window, stride, return_edges=True)
        assert got.tobytes() == edges.tobytes()
        assert (a.sensor_items() == items.ravel()).all()
        b.load_csr(off, items.ravel())
        assert (a.item_ids() == b.item_ids()).all()
        assert (a.item_supports() == b.item_supports()).all()
        thr = gen.min_count(0.05, a.n_txn)
        assert a.mine(thr).as_dict() == b.mine(thr).as_dict()
        # a later offline load replaces the kept items
        a.load_sensors(x, B, window, stride)
        assert (a.sensor_items() == np.asarray(X.transactions(x, B, window, stride)[1], dtype=np.int32)).all()
