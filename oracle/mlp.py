"""CPU reference of the surrogate MLP (ReLU hidden layers, linear output,
mean-squared-error loss) in NumPy -- TEST INFRASTRUCTURE ONLY.

Parameters use the device library's flat layout: per layer W [rows][cols]
(cols = output width rounded up to 32) then b [cols].  `bf16=True` rounds the
GEMM operands to bfloat16 (round to nearest even) exactly where the device
kernels do, so the two agree to FP32-accumulation noise; `bf16=False` is the
plain FP32 network.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params, layout):
    rows, cols, w_off, b_off = layout
    out = []
    for r, c, wo, bo in zip(rows, cols, w_off, b_off):
        W = params[wo:wo + r * c].reshape(r, c)
        b = params[bo:bo + c]
        out.append((W, b))
    return out


def forward(params, layout, x, bf16=False, dtype=np.float32):
    layers = unpack(np.asarray(params, dtype=dtype), layout)
    h = np.asarray(x, dtype=dtype)
    cache = []
    for i, (W, b) in enumerate(layers):
        hin = bf16_round(h) if bf16 else h
        Wq = bf16_round(W) if bf16 else W
        if hin.shape[1] < W.shape[0]:  # inputs narrower than the padded rows
            hin = np.pad(hin, ((0, 0), (0, W.shape[0] - hin.shape[1])))
        z = (hin.astype(dtype) @ Wq.astype(dtype)) + b
        last = i == len(layers) - 1
        h = z if last else np.maximum(z, 0)
        cache.append((hin, z))
    return h[:, 0], cache


def loss_and_grad(params, layout, x, t, bf16=False, dtype=np.float32):
    """MSE loss and the analytic gradient w.r.t. the flat parameters."""
    params = np.asarray(params, dtype=dtype)
    layers = unpack(params, layout)
    y, cache = forward(params, layout, x, bf16, dtype)
    m = len(t)
    e = y - np.asarray(t, dtype=dtype)
    loss = float(np.mean(e * e))
    grad = np.zeros_like(params)
    rows, cols, w_off, b_off = layout
    dz = np.zeros((m, cols[-1]), dtype=dtype)
    dz[:, 0] = 2 * e / m
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        hin, _ = cache[i]
        dzq = bf16_round(dz) if bf16 else dz
        grad[w_off[i]:w_off[i] + rows[i] * cols[i]] = (hin.T @ dzq).ravel()
        grad[b_off[i]:b_off[i] + cols[i]] = dz.sum(axis=0)
        if i == 0:
            break
        Wq = bf16_round(W) if bf16 else W
        dh = dzq @ Wq.T
        z_prev = cache[i - 1][1]
        dz = np.where(z_prev > 0, dh, 0)
    return loss, grad
