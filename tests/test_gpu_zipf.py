"""GPU test: the synthetic Zipf set drawn on the GPU equals the oracle's CPU draws, and the
MFIs mined from it equal the oracle's."""
import numpy as np
import pytest

from oracle import cref
from sgfim.core import Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,mean,np_,seed", [(20_000, 300, 12.0, 6, 3), (70_001, 512, 20.0, 10, 11)])
def test_zipf_set_matches_cpu_draws(n, d, mean, np_, seed):
    print(f"case: zipf n={n} d={d}", flush=True)
    args = (n, d, 1.1, mean, np_, 3, 8, 0.01, 0.08, seed)
    off, items = cref.zipf_csr(*args)
    with Context(device=0) as a, Context(device=0) as b:
        a.load_zipf(n, d, 1.1, mean, np_, 3, 8, 0.01, 0.08, seed)
        b.load_csr(off, items)
        sa = dict(zip(a.item_ids().tolist(), a.item_supports().tolist()))
        sb = dict(zip(b.item_ids().tolist(), b.item_supports().tolist()))
        assert {k: v for k, v in sa.items() if v} == sb
        thr = int(0.03 * n)
        ra = a.mine(thr).as_dict()
        rb = b.mine(thr).as_dict()
        assert ra == rb
        rows = cref.bitmap(off, items, d)
        sup = cref.row_supports(rows)
        keep = np.where(sup >= thr)[0]
        out, _ = cref.mine(rows[keep], thr)
        assert ra == {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}
