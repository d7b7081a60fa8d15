"""CPU checks of the MLP reference itself (no GPU)."""

import numpy as np

from oracle import mlp_ref as R


def test_analytic_backward_matches_finite_differences():
    import sys
    rng = np.random.default_rng(0)
    th = rng.standard_normal(R.N_PARAMS) * 0.1
    th[R.OFF_W1:R.OFF_B1].reshape(R.HID, R.IN)[:, 98:] = 0.0
    masks = np.array([sum(1 << int(p) for p in rng.choice(49, int(k), replace=False))
                      for k in rng.integers(4, 15, 48)], dtype=np.uint64)
    t = rng.random(48)
    g = R.grads_exact(th, masks, t)
    idx = rng.choice(R.N_PARAMS, 40, replace=False)
    idx = np.concatenate([idx, [R.OFF_B3, R.OFF_W3 + 5, R.OFF_B1 + 7]])
    fd = R.finite_difference(th, masks, t, idx, eps=1e-5)
    assert np.max(np.abs(fd - g[idx])) <= 1e-8 * max(1.0, np.max(np.abs(g)))


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(1)
    th = rng.standard_normal(R.N_PARAMS) * 0.1
    masks = np.array([sum(1 << int(p) for p in rng.choice(49, 9, replace=False))
                      for _ in range(64)], dtype=np.uint64)
    t = rng.random(64)
    _, gb = R.loss_and_grads(th, masks, t)
    g = R.grads_exact(th, masks, t)
    assert gb @ g / (np.linalg.norm(gb) * np.linalg.norm(g)) > 0.999


def test_features_definition():
    # S = {0, 1, 7}: differences +-(0,1), +-(1,0), +-(1,-1) each once.
    X = R.features(np.array([(1 << 0) | (1 << 1) | (1 << 7)], dtype=np.uint64))[0]
    assert X[0] == X[1] == X[7] == 1.0 and X[:49].sum() == 3
    assert X[49] == 3 / 16
    bins = {1: 1, 6: 1, 7: 1, 42: 1, 8: 0, 13: 1, 43: 1}
    # (0,1)->1, (0,-1)->6, (1,0)->7, (-1,0)->42, (1,-1)->13, (-1,1)->43
    for d, c in bins.items():
        assert X[49 + d] == R.bf16(np.float32(c) * np.float32(1.0 / 3.0)), d
    assert np.count_nonzero(X[50:98]) == 6 and np.all(X[98:] == 0)
