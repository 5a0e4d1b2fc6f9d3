"""CPU restatement of the per-path count predictor (a 16-64-64-2 ReLU MLP).

TEST INFRASTRUCTURE ONLY: the checker for the predictor's HIP kernels (FP32 training
forward/backward in vkscanlinepr_amd/csrc/k_learned.hip, BF16 matrix-core inference in
k_mlp_infer).  No reference counterpart exists; the network is new.  The float64 gradients are
pinned by central finite differences in tests/test_mlp_oracle.py.

Parameter layout (flat, float32): W1[16][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2];
inputs are normalised as (x - mean) * inv_std before the first layer.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 16, 64, 2
SIZES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)),
         ("b3", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(flat):
    flat = np.asarray(flat)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name]).reshape(-1) for name, _ in SIZES])


def forward_f64(flat, mean, inv_std, X):
    p = {k: v.astype(np.float64) for k, v in unpack(flat).items()}
    x = (np.asarray(X, np.float64) - np.asarray(mean, np.float64)) * np.asarray(inv_std, np.float64)
    h1 = np.maximum(x @ p["W1"] + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["W2"] + p["b2"], 0.0)
    y = h2 @ p["W3"] + p["b3"]
    return y, (x, h1, h2)


def loss_f64(flat, mean, inv_std, X, T):
    y, _ = forward_f64(flat, mean, inv_std, X)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def grads_f64(flat, mean, inv_std, X, T):
    """Gradient of the mean squared error over all B x 2 outputs, in the flat layout."""
    p = {k: v.astype(np.float64) for k, v in unpack(flat).items()}
    y, (x, h1, h2) = forward_f64(flat, mean, inv_std, X)
    e = y - np.asarray(T, np.float64)
    dy = 2.0 * e / e.size
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ p["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ dh2, dh2.sum(0)
    dh1 = (dh2 @ p["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ dh1, dh1.sum(0)
    return pack(g), float(np.mean(e ** 2))


def bf16(a) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest even) -> float32, for finite inputs."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(flat, mean, inv_std, X):
    """The inference kernel's arithmetic: inputs, weights and hidden activations rounded to
    bfloat16, products accumulated in (here: float64, there: float32) precision, biases in
    float32."""
    p = unpack(np.asarray(flat, np.float32))
    x = ((np.asarray(X, np.float32) - np.asarray(mean, np.float32)) * np.asarray(inv_std, np.float32))
    a = bf16(x).astype(np.float64)
    h1 = np.maximum(a @ bf16(p["W1"]).astype(np.float64) + p["b1"], 0.0).astype(np.float32)
    h2 = np.maximum(bf16(h1).astype(np.float64) @ bf16(p["W2"]).astype(np.float64) + p["b2"], 0.0)
    h2 = h2.astype(np.float32)
    return bf16(h2).astype(np.float64) @ bf16(p["W3"]).astype(np.float64) + p["b3"]
