"""Surrogate MLP that ranks exponent-set proposals before exact evaluation.

The network (csrc/mlp_kernels.hip) maps a set S to a score trained to predict
its standardised exact minimum distance.  Callers pass canonical masks (the
AGL2 canonical image of S): distance is a class invariant, and on raw masks
the position features do not generalise.  It only orders proposals: every distance the
search reports comes from the exact kernel.  With no trained network the
search falls back to unranked proposals (:meth:`Surrogate.rank` of an
untrained model is never used for that; callers pass ``surrogate=None``).
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._native import check, lib, ptr

N_PARAMS = 33153


class Surrogate:
    def __init__(self, device: int = 0, seed: int = 0):
        h = ctypes.c_void_p()
        check(lib().gt_mlp_create(int(device), ctypes.c_uint64(int(seed)), ctypes.byref(h)))
        self.handle = h
        self.steps = 0
        self.mean, self.std = 0.0, 1.0

    # ---- parameters --------------------------------------------------------
    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, dtype=np.float32)
        check(lib().gt_mlp_get_params(self.handle, ptr(out)))
        return out

    def set_params(self, theta) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == N_PARAMS
        check(lib().gt_mlp_set_params(self.handle, ptr(t)))

    # ---- compute -------------------------------------------------------------
    def features(self, masks) -> np.ndarray:
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        out = np.zeros((len(m), 128), dtype=np.uint16)
        if len(m):
            check(lib().gt_mlp_features(self.handle, ptr(m), len(m), ptr(out)))
        return out

    def grads(self, masks, target) -> tuple[np.ndarray, float]:
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().gt_mlp_grads(self.handle, ptr(m), ptr(t), len(m), ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, masks, target, lr: float = 1e-3) -> float:
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().gt_mlp_train_step(self.handle, ptr(m), ptr(t), len(m), float(lr),
                                      ctypes.byref(loss)))
        self.steps += 1
        return loss.value

    def fit(self, masks, distances, epochs: int = 20, batch: int = 1024, lr: float = 2e-3,
            seed: int = 0) -> list[float]:
        """Minibatch Adam on (mask, exact distance) pairs.  Targets are the
        distances standardised over the training set (the spread of d among
        sets of one size is a few units, so raw d / 49 gives gradients too
        small to move the hidden layers); the score keeps the order of d."""
        masks = np.asarray(masks, dtype=np.uint64)
        d = np.asarray(distances, dtype=np.float64)
        self.mean = float(d.mean()) if len(d) else 0.0
        self.std = float(d.std()) if len(d) and d.std() > 0 else 1.0
        target = ((d - self.mean) / self.std).astype(np.float32)
        rng = np.random.default_rng(seed)
        losses = []
        for _ in range(epochs):
            order = rng.permutation(len(masks))
            m = np.ascontiguousarray(masks[order])
            t = np.ascontiguousarray(target[order])
            ep = np.zeros((len(order) + batch - 1) // batch, dtype=np.float32)
            if len(order):
                check(lib().gt_mlp_train_batches(self.handle, ptr(m), ptr(t), len(m), int(batch),
                                                 float(lr), ptr(ep)))
                self.steps += len(ep)
            losses.append(float(np.mean(ep)) if len(ep) else float("nan"))
        return losses

    def predict(self, masks) -> np.ndarray:
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        out = np.zeros(len(m), dtype=np.float32)
        if len(m):
            check(lib().gt_mlp_predict(self.handle, ptr(m), len(m), ptr(out)))
        return out

    def predict_distance(self, masks) -> np.ndarray:
        """Predicted distance in the units of the last fit (for reporting only)."""
        return self.predict(masks) * self.std + self.mean

    def rank(self, masks) -> np.ndarray:
        """Indices of `masks` from most to least promising (stable)."""
        return np.argsort(-self.predict(masks), kind="stable")

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            lib().gt_mlp_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bf16_gemm(A: np.ndarray, B: np.ndarray, device: int = 0) -> np.ndarray:
    """C = A B through the MLP's bf16 MFMA tile routine (A, B given as uint16
    bf16 bit patterns, row-major)."""
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = np.zeros((M, N), dtype=np.float32)
    check(lib().gt_bf16_gemm(int(device), ptr(A), ptr(B), ptr(C), M, N, K))
    return C
