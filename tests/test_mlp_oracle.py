"""Pins the CPU reference of the surrogate MLP (oracle/mlp.py): its analytic
gradient against central finite differences in float64, and its bfloat16
rounding against the IEEE definition."""
import numpy as np

from oracle import mlp as M


def small_layout(dims):
    rows, cols, w_off, b_off = [], [], [], []
    off = 0
    prev = dims[0]
    for o in dims[1:]:
        c = (o + 31) // 32 * 32
        rows.append(prev)
        cols.append(c)
        w_off.append(off)
        off += prev * c
        b_off.append(off)
        off += c
        prev = c
    return (rows, cols, w_off, b_off), off


def test_finite_differences_float64():
    rng = np.random.default_rng(0)
    layout, n = small_layout((32, 32, 32, 1))
    p = rng.normal(0, 0.3, n)
    # zero the padded output columns (as the device initialises them)
    rows, cols, w_off, b_off = layout
    W = p[w_off[-1]:w_off[-1] + rows[-1] * cols[-1]].reshape(rows[-1], cols[-1])
    W[:, 1:] = 0
    p[b_off[-1] + 1:b_off[-1] + cols[-1]] = 0
    x = rng.normal(size=(7, 32))
    t = rng.normal(size=7)
    loss, g = M.loss_and_grad(p, layout, x, t, dtype=np.float64)
    idx = rng.choice(n, 60, replace=False)
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = M.loss_and_grad(q, layout, x, t, dtype=np.float64)
        q[i] -= 2 * eps
        lm, _ = M.loss_and_grad(q, layout, x, t, dtype=np.float64)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_round():
    vals = np.array([1.0, 1.00390625, 1.0078125, 3.14159, -2.5e-3, 65504.0], dtype=np.float32)
    r = M.bf16_round(vals)
    # exactly representable values are unchanged; others are within half a bf16 ulp
    assert r[0] == 1.0 and r[2] == 1.0078125
    assert r[1] == 1.0  # tie 1 + 2^-8 rounds to even (1.0)
    ulp = 2.0 ** (np.floor(np.log2(np.abs(vals))) - 7)
    assert np.all(np.abs(r - vals) <= ulp / 2)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
