"""CPU oracle for the learned CG warm start -- TEST INFRASTRUCTURE ONLY.

The per-particle network has no counterpart in the reference; it is pinned
by the mathematics instead: this float64 numpy forward/backward is checked
against central finite differences (tests/test_mlp_oracle.py), and the GPU's
BF16-MFMA kernels are checked against it.

Network: 16 features -> 64 -> 64 -> 64 -> 1, SiLU after the three hidden
layers, loss = mean((y - t)^2).  Parameters in the flat order
W1[16x64], b1[64], W2[64x64], b2[64], W3[64x64], b3[64], W4[64], b4[1]
(9,473 values; W stored [in][out] row-major).  With ``bf16=True`` every
matrix-product input (weights and activations, forward and backward) is
rounded to bfloat16 (round to nearest even) as the device's MFMA operands
are; the last layer's backward (vector units) rounds only W4; sums, biases
and SiLU stay in float64.
"""
from __future__ import annotations

import numpy as np

IN, HID = 16, 64
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, k = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = np.asarray(theta[k:k + n], np.float64).reshape(s)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def to_bf16(a):
    """Round float64 values to the nearest bfloat16 (ties to even), via float32."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, X, bf16=False):
    P = unpack(theta)
    q = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    H = np.asarray(X, np.float64)
    cache = []
    for l in (1, 2, 3):
        Z = q(H) @ q(P[f"W{l}"]) + P[f"b{l}"]
        cache.append((H, Z))
        H = silu(Z)
    y = q(H) @ q(P["W4"]) + P["b4"][0]
    cache.append((H, None))
    return y, cache


def loss_and_grad(theta, X, t, bf16=False):
    P = unpack(theta)
    q = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    y, cache = forward(theta, X, bf16)
    n = len(t)
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {}
    H3 = cache[3][0]
    # the last layer's backward runs on the vector units: only W4 is rounded
    g["W4"] = H3.T @ dy
    g["b4"] = np.array([dy.sum()])
    dH = np.outer(dy, q(P["W4"]))
    for l in (3, 2, 1):
        Hin, Z = cache[l - 1]
        dZ = dH * dsilu(Z)
        g[f"W{l}"] = q(Hin).T @ q(dZ)
        g[f"b{l}"] = dZ.sum(0)
        if l > 1:
            dH = q(dZ) @ q(P[f"W{l}"]).T
    return loss, pack(g), y


def init(seed=0):
    """Scaled-normal initialisation (tests pass explicit parameters)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, s in SHAPES:
        if name.startswith("W"):
            fan_in = s[0]
            d[name] = rng.standard_normal(s) * np.sqrt(1.0 / fan_in)
        else:
            d[name] = 0.01 * rng.standard_normal(s)
    return pack(d)
