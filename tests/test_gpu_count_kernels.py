"""Both counting kernels of the search (AND + popcount on the VALU, and 16 x 16 fp4 MFMA
products of children and candidates) give the oracle's exact MFI families: plain rows,
projected blocks (gathered and compacted), every tile width, several thresholds in one
search and many windows; and the switch is refused while an anytime search runs."""
import numpy as np
import pytest

from oracle import cref
from sgfim import _lib as L
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu

KINDS = [(Context.COUNT_VALU, "valu"), (Context.COUNT_F4, "f4")]


def _oracle(T, thr):
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): int(s) for x, s in out}


def _by_tag(res, n):
    got = [dict() for _ in range(n)]
    for x, s, w in zip(res.itemsets, res.supports, res.windows):
        got[w][frozenset(x)] = s
    return got


@pytest.mark.parametrize("cw", [128, 256, 512])
@pytest.mark.parametrize("proj,dense", [(0, 0), (4, 0), (4, 6)])
def test_count_kernels_match_oracle(cw, proj, dense):
    print(f"case: workload1 seed 5 at 2%, tile {cw} x 3, projection {proj} dense {dense}", flush=True)
    T = gen.workload1(5, 20011)
    thr = gen.min_count(0.02, T.n_txn)
    want = _oracle(T, thr)
    with Context(device=0, batch_nodes=512) as ctx:
        ctx.load_csr(T.offsets, T.items)
        ctx.set_tile(cw, 3)
        ctx.set_projection(proj)
        if proj:
            ctx.set_block_rule(dense, 2)
        for kind, name in KINDS:
            ctx.set_count_kernel(kind)
            got = ctx.mine(thr).as_dict()
            assert got == want, name
    assert len(want) > 0


def test_count_kernels_thresholds_and_windows():
    print("case: workload 4 seed 1000011 at 40,000 txns, 2%/4%/8%; 12 windows at 3%", flush=True)
    T = gen.workload4(1_000_011, 40_000)
    thrs = [gen.min_count(s, T.n_txn) for s in (0.02, 0.04, 0.08)]
    wins = gen.workload3(9, n_windows=12, lo=500, hi=6000)
    win_off, off, items, _ = gen.concat_windows(wins)
    wthr = [gen.min_count(0.03, W.n_txn) for W in wins]
    for kind, name in KINDS:
        with Context(device=0, batch_nodes=1024) as ctx:
            ctx.set_count_kernel(kind)
            ctx.load_csr(T.offsets, T.items)
            got = _by_tag(ctx.mine_thresholds(thrs), len(thrs))
            for i, t in enumerate(thrs):
                assert got[i] == _oracle(T, t), (name, i)
            ctx.load_windows(win_off, off, items)
            gw = _by_tag(ctx.mine_windows(wthr), len(wins))
            for w, W in enumerate(wins):
                assert gw[w] == _oracle(W, wthr[w]), (name, w)


def test_count_kernel_switch_is_checked():
    T = gen.workload1(1, 3000)
    with Context(device=0) as ctx:
        ctx.load_csr(T.offsets, T.items)
        with pytest.raises(L.SgfimError):
            ctx.set_count_kernel(5)
        ctx.anytime_begin(gen.min_count(0.05, T.n_txn))
        ctx.anytime_step(1)
        st = ctx.last_stats()
        if not st["done"]:
            with pytest.raises(L.SgfimError) as e:
                ctx.set_count_kernel(Context.COUNT_F4)
            assert e.value.code == L.E_STATE
        ctx.anytime_step(0)
        ctx.set_count_kernel(Context.COUNT_F4)
        ctx.set_count_kernel(Context.COUNT_VALU)
