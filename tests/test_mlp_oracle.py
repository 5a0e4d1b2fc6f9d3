"""Pins the predictor's CPU oracle (oracle/mlp_oracle.py): its float64 gradients against
central finite differences of its own loss, and its bfloat16 rounding against known bit
patterns."""
from __future__ import annotations

import numpy as np

import mlp_oracle as M


def _model(seed):
    rng = np.random.default_rng(seed)
    flat = rng.normal(scale=0.3, size=M.N_PARAMS)
    mean = rng.normal(size=M.IN)
    inv_std = rng.uniform(0.5, 2.0, size=M.IN)
    X = rng.normal(size=(37, M.IN)) / inv_std + mean
    T = rng.normal(size=(37, M.OUT))
    return flat, mean, inv_std, X, T


def test_gradients_match_finite_differences():
    flat, mean, inv_std, X, T = _model(0)
    g, loss = M.grads_f64(flat, mean, inv_std, X, T)
    assert abs(loss - M.loss_f64(flat, mean, inv_std, X, T)) < 1e-12
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 120, replace=False),
                          [0, 1023, 1024, 1087, 1088, 5183, 5184, 5247, 5248, 5375, 5376, 5377]])
    eps = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        num = (M.loss_f64(fp, mean, inv_std, X, T) - M.loss_f64(fm, mean, inv_std, X, T)) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_layout_sizes():
    assert M.N_PARAMS == 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2 == 5378
    parts = M.unpack(np.arange(M.N_PARAMS, dtype=np.float64))
    assert parts["W1"][0, 1] == 1 and parts["b1"][0] == 1024 and parts["W3"][0, 1] == 5249
    assert np.array_equal(M.pack(parts), np.arange(M.N_PARAMS))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e-3, 0.0], np.float32)
    got = M.bf16(v).view(np.uint32) >> 16
    # 1 + 2^-8 is a tie between 0x3F80 and 0x3F81: to even (0x3F80); 1 + 3*2^-8 ties to 0x3F82
    assert list(got) == [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x3B45, 0x0000]


def test_bf16_forward_close_to_float64():
    flat, mean, inv_std, X, T = _model(2)
    y64, _ = M.forward_f64(flat, mean, inv_std, X)
    yb = M.forward_bf16(flat, mean, inv_std, X)
    assert np.max(np.abs(yb - y64)) < 0.05 * (1 + np.max(np.abs(y64)))
