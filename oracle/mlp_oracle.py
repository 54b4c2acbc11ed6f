"""Float64 oracle of the proposal network (24 -> 256 ReLU -> 1) -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component; this oracle is written from the
mathematics of a two-layer perceptron with a binary cross-entropy loss on logits:
  H = relu(X W1 + b1), z = H w2 + b2, L = mean(softplus(z) - y z)
  dz = (sigmoid(z) - y) / n, dw2 = H^T dz, db2 = sum dz,
  dH = dz w2^T * [H > 0], db1 = sum_rows dH, dW1 = X^T dH.
bf16=True applies the device kernels' roundings (round-to-nearest-even to BF16 of
X, W1, w2 and of dH before the X^T dH product) so device gradients can be checked
tightly; bf16=False is the exact-arithmetic reference that finite differences pin.
"""
from __future__ import annotations

import numpy as np

F, HID = 24, 256
NPARAM = F * HID + HID + HID + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def init_params(seed: int = 0):
    rng = np.random.default_rng(seed)
    W1 = (rng.standard_normal((F, HID)) * np.sqrt(2.0 / F)).astype(np.float32)
    b1 = np.zeros(HID, dtype=np.float32)
    w2 = (rng.standard_normal(HID) * np.sqrt(1.0 / HID)).astype(np.float32)
    return W1, b1, w2, np.float32(0.0)


def forward(X, W1, b1, w2, b2, bf16_ops=True):
    Xc = bf16(X) if bf16_ops else np.asarray(X, np.float64)
    W1c = bf16(W1) if bf16_ops else np.asarray(W1, np.float64)
    w2c = bf16(w2) if bf16_ops else np.asarray(w2, np.float64)
    H = np.maximum(Xc @ W1c + np.asarray(b1, np.float64), 0.0)
    z = H @ w2c + float(b2)
    return z, H, Xc, w2c


def loss_grad(X, Y, W1, b1, w2, b2, bf16_ops=True):
    """(loss, packed gradient float64 [NPARAM]) in the order W1, b1, w2, b2."""
    n = X.shape[0]
    z, H, Xc, w2c = forward(X, W1, b1, w2, b2, bf16_ops)
    Y = np.asarray(Y, np.float64)
    loss = np.mean(np.maximum(z, 0) - Y * z + np.log1p(np.exp(-np.abs(z))))
    dz = (1.0 / (1.0 + np.exp(-z)) - Y) / n
    dw2 = H.T @ dz
    db2 = dz.sum()
    dH = np.outer(dz, w2c) * (H > 0)
    db1 = dH.sum(axis=0)
    dHc = bf16(dH.astype(np.float32)) if bf16_ops else dH
    dW1 = Xc.T @ dHc
    return float(loss), np.concatenate([dW1.reshape(-1), db1, dw2, [db2]])


def adam_step(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def pack(W1, b1, w2, b2):
    return np.concatenate([np.asarray(W1, np.float64).reshape(-1), b1, w2, [b2]])


def unpack(p):
    W1 = p[:F * HID].reshape(F, HID)
    b1 = p[F * HID:F * HID + HID]
    w2 = p[F * HID + HID:F * HID + 2 * HID]
    return W1, b1, w2, p[-1]
