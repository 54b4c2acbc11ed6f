"""GPU tests of the proposal network's HIP forward/backward kernels against the float64 oracle."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from sgfim import learned

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, M.F)) * 0.8).astype(np.float32)
    Y = (rng.random(n) < 0.3).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("n", [1, 31, 33, 200, 4096])
def test_gradient_matches_oracle(n):
    print(f"case: mlp gradient n={n}", flush=True)
    X, Y = _data(n, n)
    params = M.init_params(7)
    params = (params[0], params[1] + 0.01, params[2], 0.1)
    tr = learned.Trainer(params)
    loss, g = tr.step(X, Y, lr=0.0)
    lref, gref = M.loss_grad(X, Y, *params, bf16_ops=True)
    # FP32 accumulation on the device vs float64 here; dH is rounded to BF16 on both
    # sides from slightly different FP32/float64 values, so a rounding may differ by one
    # BF16 step: tolerance 2e-2 of the largest |gradient| per block, 1e-4 relative for the loss.
    assert abs(loss - lref) <= 1e-4 * max(1.0, abs(lref))
    blocks = [(0, M.F * M.HID), (M.F * M.HID, M.F * M.HID + M.HID),
              (M.F * M.HID + M.HID, M.F * M.HID + 2 * M.HID), (M.NPARAM - 1, M.NPARAM)]
    for a, b in blocks:
        scale = np.abs(gref[a:b]).max() + 1e-12
        assert np.abs(g[a:b] - gref[a:b]).max() <= 2e-2 * scale + 1e-7, (a, b)
    tr.close()


def test_forward_matches_oracle():
    print("case: mlp forward", flush=True)
    X, _ = _data(1000, 3)
    params = M.init_params(11)
    tr = learned.Trainer(params)
    z = tr.forward(X)
    zref, _, _, _ = M.forward(X, *params, bf16_ops=True)
    assert np.abs(z - zref).max() <= 1e-4 * (np.abs(zref).max() + 1.0)
    p = tr.forward(X, probs=True)
    assert np.abs(p - 1 / (1 + np.exp(-zref))).max() <= 1e-4
    tr.close()


def test_adam_steps_match_oracle_and_are_deterministic():
    print("case: mlp adam determinism", flush=True)
    X, Y = _data(512, 5)
    params = M.init_params(2)
    a = learned.Trainer(params)
    b = learned.Trainer(params)
    for _ in range(5):
        a.step(X, Y, lr=1e-2)
        b.step(X, Y, lr=1e-2)
    pa, pb = a.params(), b.params()
    for x, y in zip(pa, pb):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    # one oracle Adam step from the same start agrees with one device step
    c = learned.Trainer(params)
    c.step(X, Y, lr=1e-2)
    _, g = M.loss_grad(X, Y, *params, bf16_ops=True)
    p1, _, _ = M.adam_step(M.pack(*params), g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, 1e-2)
    pc = M.pack(*c.params())
    # the first Adam step moves each parameter by ~lr*sign(g); compare where |g| is not tiny
    mask = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.abs(pc[mask] - p1[mask]).max() <= 1e-3
    for t in (a, b, c):
        t.close()


def test_training_learns_a_separable_rule():
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((20000, M.F)).astype(np.float32)
    Y = ((X[:, 0] + 0.5 * X[:, 3] - X[:, 7]) > 0.2).astype(np.float32)
    tr = learned.Trainer(seed=1)
    tr.set_data(X, Y)
    losses = tr.train(epochs=8, batch=1024, lr=3e-3, seed=0)
    assert losses[-1] < 0.25 * losses[0]
    p = tr.forward(X, probs=True)
    acc = ((p > 0.5) == (Y > 0.5)).mean()
    assert acc > 0.95
    tr.close()
