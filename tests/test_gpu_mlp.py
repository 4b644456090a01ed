"""The surrogate MLP's BF16 MFMA forward/backward kernels against the CPU
reference (oracle/mlp.py), and on-device training."""
import numpy as np
import pytest

from oracle import mlp as OM
from graphgolf_amd.surrogate import Mlp

pytestmark = pytest.mark.gpu

DIMS = (64, 256, 256, 256, 1)


@pytest.fixture(scope="module")
def net():
    m = Mlp(DIMS, seed=3)
    yield m
    m.close()


@pytest.mark.parametrize("rows", [1, 31, 100, 4096])
def test_gpu_mlp_forward_matches_reference(net, rows):
    rng = np.random.default_rng(rows)
    x = rng.normal(size=(rows, 64)).astype(np.float32)
    p = net.params()
    got = net.forward(x)
    want_bf, _ = OM.forward(p, net.layout, x, bf16=True)
    want_32, _ = OM.forward(p, net.layout, x, bf16=False)
    scale = np.abs(want_32).max() + 1e-6
    # BF16 operands, FP32 accumulation: equal to the bf16-emulating reference up
    # to accumulation order; within bf16 precision of the FP32 network.
    assert np.max(np.abs(got - want_bf)) <= 2e-3 * scale
    assert np.max(np.abs(got - want_32)) <= 3e-2 * scale


def test_gpu_mlp_gradient_matches_reference(net):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(200, 64)).astype(np.float32)
    t = rng.normal(size=200).astype(np.float32)
    p = net.params()
    loss, g = net.gradient(x, t)
    l_bf, g_bf = OM.loss_and_grad(p, net.layout, x, t, bf16=True)
    l_32, g_32 = OM.loss_and_grad(p, net.layout, x, t, bf16=False)
    assert abs(loss - l_bf) <= 1e-3 * abs(l_bf)
    gs = np.abs(g_bf).max()
    assert np.max(np.abs(g - g_bf)) <= 2e-2 * gs
    def cos(a, b):
        return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
    # Tolerances: against the reference that rounds the same operands to BF16
    # the gradient agrees to FP32 accumulation noise; against the plain FP32
    # network it agrees to BF16 precision (gradients themselves are BF16 in
    # the backward GEMMs).
    assert cos(g, g_bf) > 0.9999
    assert cos(g, g_32) > 0.995
    # padded output columns and rows carry no gradient
    rows, cols, w_off, b_off = net.layout
    W_last = g[w_off[-1]:w_off[-1] + rows[-1] * cols[-1]].reshape(rows[-1], cols[-1])
    assert np.all(W_last[:, 1:] == 0)


def test_gpu_mlp_params_roundtrip():
    m = Mlp(DIMS, seed=11)
    p = m.params()
    q = p * 0.5
    m.set_params(q)
    assert np.array_equal(m.params(), q)
    m.close()


def test_gpu_mlp_trains():
    rng = np.random.default_rng(0)
    m = Mlp(DIMS, seed=5)
    w = rng.normal(size=64).astype(np.float32)
    x = rng.normal(size=(1024, 64)).astype(np.float32)
    t = np.tanh(x @ w / 8).astype(np.float32) + 0.1 * (x[:, 0] ** 2)
    first = m.train_step(x, t, lr=1e-3)
    for _ in range(300):
        last = m.train_step(x, t, lr=1e-3)
    assert last < 0.1 * first, (first, last)
    m.close()


@pytest.mark.parametrize("dims", [(20, 64, 33, 1), (8, 40, 1), (64, 96, 1)])
def test_gpu_mlp_odd_shapes_match_reference(dims):
    """Input widths that are not multiples of 8 (padded to 16-byte BF16 rows
    inside), hidden widths that are not multiples of 32 or 64 (padded columns,
    partial 64x64 GEMM tiles): forward and gradient against the reference."""
    m = Mlp(dims, seed=11)
    try:
        rng = np.random.default_rng(5)
        x = rng.normal(size=(77, dims[0])).astype(np.float32)
        t = rng.normal(size=77).astype(np.float32)
        p = m.params()
        got = m.forward(x)
        want, _ = OM.forward(p, m.layout, x, bf16=True)
        assert np.max(np.abs(got - want)) <= 2e-3 * (np.abs(want).max() + 1e-6)
        loss, g = m.gradient(x, t)
        l_bf, g_bf = OM.loss_and_grad(p, m.layout, x, t, bf16=True)
        assert abs(loss - l_bf) <= 1e-3 * abs(l_bf)
        assert np.max(np.abs(g - g_bf)) <= 2e-2 * np.abs(g_bf).max()
    finally:
        m.close()
