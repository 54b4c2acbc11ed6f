"""GPU tests of the sensor transactionisation front-end against its NumPy restatement."""
import numpy as np
import pytest

from oracle import transact as X
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
                                                      (2, 3000, 8, 8, 7, 3), (3, 640, 4, 6, 64, 64)])
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
