"""CPU reference of the surrogate MLP (test infrastructure only).

float64 restatement of the network in csrc/mlp_kernels.hip with the same bf16
round-to-nearest-even points: features, weights W1/W2/w3, the hidden
activations H1/H2 and the back-propagated errors dH2/dH1 are rounded to bf16
where the GPU feeds them to the matrix cores; every sum is exact float64 here
(fp32 on the GPU), so GPU and CPU gradients agree to bf16 precision, not bit
for bit.  The tolerance lives in tests/test_gpu_mlp.py.

There is no reference counterpart (the reference has no learned component);
"parity" here is the gradient check of the hand-written backward pass.
"""

from __future__ import annotations

import numpy as np

IN, HID = 128, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
N_PARAMS = OFF_B3 + 1


def bf16(x) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(masks: np.ndarray) -> np.ndarray:
    masks = np.asarray(masks, dtype=np.uint64)
    n = len(masks)
    X = np.zeros((n, IN), dtype=np.float32)
    for r, m in enumerate(masks.tolist()):
        pts = [p for p in range(49) if (m >> p) & 1]
        k = len(pts)
        cnt = np.zeros(49, dtype=np.int64)
        for a in pts:
            for b in pts:
                if a != b:
                    cnt[((a // 7 - b // 7) % 7) * 7 + (a % 7 - b % 7) % 7] += 1
        X[r, [p for p in pts]] = 1.0
        X[r, 49] = np.float32(k) / np.float32(16.0)
        invk = np.float32(1.0) / np.float32(k) if k else np.float32(0.0)
        X[r, 50:98] = cnt[1:].astype(np.float32) * invk
    return bf16(X)


def unpack(theta: np.ndarray):
    t = np.asarray(theta, dtype=np.float64)
    return (t[OFF_W1:OFF_B1].reshape(HID, IN), t[OFF_B1:OFF_W2],
            t[OFF_W2:OFF_B2].reshape(HID, HID), t[OFF_B2:OFF_W3], t[OFF_W3:OFF_B3], t[OFF_B3])


def forward(theta, X):
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    W1b, W2b, w3b = bf16(W1), bf16(W2), bf16(w3)
    H1 = bf16(np.maximum(X @ W1b.T + b1, 0.0))
    H2 = bf16(np.maximum(H1 @ W2b.T + b2, 0.0))
    y = H2 @ w3b + b3
    return y, (W1b, W2b, w3b, H1, H2)


def loss_and_grads(theta, masks, target):
    X = features(masks)
    t = np.asarray(target, dtype=np.float64)
    n = len(t)
    y, (W1b, W2b, w3b, H1, H2) = forward(theta, X)
    diff = y - t
    loss = 0.5 * np.mean(diff ** 2)
    dy = diff / n
    D2 = bf16(np.where(H2 != 0, dy[:, None] * w3b[None, :], 0.0))
    D1 = bf16(np.where(H1 != 0, D2 @ W2b, 0.0))
    g = np.zeros(N_PARAMS)
    g[OFF_W1:OFF_B1] = (D1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    g[OFF_W2:OFF_B2] = (D2.T @ H1).ravel()
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    g[OFF_W3:OFF_B3] = dy @ H2
    g[OFF_B3] = dy.sum()
    return loss, g


def finite_difference(theta, masks, target, idx, eps=1e-3):
    """Central difference of the float64 loss without bf16 roundings of the
    activations (a check of the analytic backward pass itself)."""
    X = features(masks)
    t = np.asarray(target, dtype=np.float64)

    def L(th):
        W1, b1, W2, b2, w3, b3 = unpack(th)
        H1 = np.maximum(X @ W1.T + b1, 0.0)
        H2 = np.maximum(H1 @ W2.T + b2, 0.0)
        return 0.5 * np.mean((H2 @ w3 + b3 - t) ** 2)

    out = []
    for i in idx:
        a = np.array(theta, dtype=np.float64)
        b = a.copy()
        a[i] += eps
        b[i] -= eps
        out.append((L(a) - L(b)) / (2 * eps))
    return np.array(out)


def grads_exact(theta, masks, target):
    """Analytic gradient of the same float64 loss (no roundings)."""
    X = features(masks)
    t = np.asarray(target, dtype=np.float64)
    n = len(t)
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    H1 = np.maximum(X @ W1.T + b1, 0.0)
    H2 = np.maximum(H1 @ W2.T + b2, 0.0)
    dy = (H2 @ w3 + b3 - t) / n
    D2 = np.where(H2 > 0, dy[:, None] * w3[None, :], 0.0)
    D1 = np.where(H1 > 0, D2 @ W2, 0.0)
    g = np.zeros(N_PARAMS)
    g[OFF_W1:OFF_B1] = (D1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    g[OFF_W2:OFF_B2] = (D2.T @ H1).ravel()
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    g[OFF_W3:OFF_B3] = dy @ H2
    g[OFF_B3] = dy.sum()
    return g
