"""CPU tests pinning the float64 MLP oracle by finite differences and closed forms."""
import numpy as np

from oracle import mlp_oracle as M


def _data(n=37, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.F)).astype(np.float32)
    Y = (rng.random(n) < 0.4).astype(np.float32)
    return X, Y


def test_analytic_gradient_matches_finite_differences():
    X, Y = _data()
    W1, b1, w2, b2 = M.init_params(3)
    b1 = b1 + 0.05
    p0 = M.pack(W1, b1, w2, b2)
    _, g = M.loss_grad(X, Y, *M.unpack(p0), bf16_ops=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.F * M.HID, 40, replace=False),
                          M.F * M.HID + rng.choice(M.HID, 10, replace=False),
                          M.F * M.HID + M.HID + rng.choice(M.HID, 10, replace=False), [M.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = M.loss_grad(X, Y, *M.unpack(pp), bf16_ops=False)
        lm, _ = M.loss_grad(X, Y, *M.unpack(pm), bf16_ops=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_closed_forms():
    # zero weights: z = b2 for every row, loss = softplus(b2) - mean(y) b2, dz = (sigmoid(b2) - y) / n
    X, Y = _data(20, 2)
    Y = Y.astype(np.float64)
    W1 = np.zeros((M.F, M.HID), np.float32)
    loss, g = M.loss_grad(X, Y, W1, np.zeros(M.HID), np.zeros(M.HID), 0.7, bf16_ops=False)
    sp = np.log1p(np.exp(0.7))
    assert abs(loss - (sp - Y.mean() * 0.7)) < 1e-12
    assert abs(g[-1] - (1 / (1 + np.exp(-0.7)) - Y.mean())) < 1e-12
    assert np.all(g[:-1] == 0)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # halfway rounds to even
    assert r[2] == 1.0078125                    # above halfway rounds up
    assert abs(r[3] - (-3.0e-3)) < 2e-5
    assert r[4] == 65536.0


def test_adam_matches_textbook_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = M.adam_step(p, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first bias-corrected step moves each parameter by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)


def test_trace_labels_direct_and_ancestor():
    from sgfim.learned import trace_labels
    # births (key, parent): 0 root record, 1 and 2 children of 0, 3 child of 1
    keys = np.asarray([[0, -1], [1, 0], [2, 0], [3, 1]], dtype=np.int64)
    mfi_keys = np.asarray([3, -1, 3], dtype=np.int64)   # record 3 confirmed two MFIs; -1 = the seed
    assert trace_labels(keys, mfi_keys, "direct").tolist() == [0, 0, 0, 1]
    assert trace_labels(keys, mfi_keys, "ancestor").tolist() == [1, 1, 0, 1]


def test_yield_labels():
    from sgfim.learned import yield_labels
    keys = np.asarray([[0, -1], [1, 0], [2, 0]], dtype=np.int64)
    mfi_keys = np.asarray([0, 0, 2, -1], dtype=np.int64)
    feat = np.zeros((3, 24), dtype=np.float32)
    feat[:, 6] = np.log2(1.0 + np.asarray([4.0, 3.0, 1.0]))   # children 4, 3, 1
    assert np.allclose(yield_labels(keys, mfi_keys, feat), [0.5, 0.0, 1.0])
