"""GPU tests: every counting-tile shape gives the same MFIs, and the autotuner keeps one."""
import numpy as np
import pytest

from oracle import cref
from sgfim import _lib as L
from sgfim import autotune as A
from sgfim import gen
from sgfim.core import Context

pytestmark = pytest.mark.gpu


def _oracle(T, thr):
    rows = cref.bitmap(T.offsets, T.items, T.n_items)
    sup = cref.row_supports(rows)
    keep = np.where(sup >= thr)[0]
    out, _ = cref.mine(rows[keep], thr)
    return {frozenset(int(keep[i]) + 1 for i in x): s for x, s in out}


def test_every_tile_shape_matches_the_oracle():
    print("case: tile shapes, workload1 seed 3 n=40000 at 2% and 5%", flush=True)
    T = gen.workload1(3, 40_000)
    thrs = [gen.min_count(0.02, T.n_txn), gen.min_count(0.05, T.n_txn)]
    refs = [_oracle(T, t) for t in thrs]
    with Context(device=0) as ctx:
        ctx.load_csr(T.offsets, T.items)
        assert ctx.tile() == A.DEFAULT
        for cw, cr in A.CANDIDATES:
            ctx.set_tile(cw, cr)
            assert ctx.tile() == (cw, cr)
            for thr, ref in zip(thrs, refs):
                assert ctx.mine(thr).as_dict() == ref, (cw, cr, thr)
        with pytest.raises(L.SgfimError) as e:
            ctx.set_tile(384, 4)
        assert e.value.code == L.E_INVAL
        with pytest.raises(L.SgfimError) as e:
            ctx.set_tile(256, 65)
        assert e.value.code == L.E_INVAL


def test_tile_shapes_on_windows():
    print("case: tile shapes over 24 windows", flush=True)
    win_off, off, items, lens = gen.workload3_csr(5, 24, 1)
    thrs = [gen.min_count(0.03, int(n)) for n in lens]
    with Context(device=0) as ctx:
        ctx.load_windows(win_off, off, items)
        ref = None
        for cw, cr in ((256, 4), (128, 8), (512, 1)):
            ctx.set_tile(cw, cr)
            sig = A.signature(ctx.mine_windows(thrs, convert=False))
            ref = ref or sig
            assert sig == ref, (cw, cr)


def test_tune_picks_a_candidate(tmp_path):
    print("case: autotune on workload1 seed 4", flush=True)
    T = gen.workload1(4, 20_000)
    thr = gen.min_count(0.02, T.n_txn)
    p = str(tmp_path / "tiles.json")
    with Context(device=0) as ctx:
        ctx.load_csr(T.offsets, T.items)
        key = A.context_key(ctx)
        assert key.startswith("gfx950:")
        e = A.tune(ctx, lambda: ctx.mine(thr, convert=False), candidates=((256, 4), (128, 4), (512, 2)), reps=1,
                   key=key, path=p)
        assert (e["chunk_words"], e["chunk_range"]) in ((256, 4), (128, 4), (512, 2))
        assert ctx.tile() == (e["chunk_words"], e["chunk_range"])
        assert A.lookup(key, p) == ctx.tile()
        ctx.set_tile(0, 0)
        assert A.apply(ctx, path=p) == ctx.tile() == (e["chunk_words"], e["chunk_range"])
        assert ctx.mine(thr).as_dict() == _oracle(T, thr)
