"""Pins the CPU restatement of the radius-proposal MLP (oracle/mlp_ref.py) on the CPU.

- FP64 back-propagation against central finite differences, on every parameter block;
- bfloat16 rounding against hand-computed bit patterns (ties to even, carries into the exponent);
- the first Adam step against its closed form, and a later step against a direct evaluation;
- the bf16-rounded gradient stays close to the FP64 one (rounding, not a different formula).
"""
import numpy as np
import pytest

from oracle import mlp_ref as M


def _fd_check(params, X, T, idx, h=1e-6):
    _, _, g = M.forward_backward(params, X, T, rounding=None)
    for i in idx:
        p1 = params.copy()
        p2 = params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (M.loss(p1, X, T) - M.loss(p2, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_backprop_matches_finite_differences():
    params = M.init_params(3).astype(np.float64)
    X, T = M.synthetic_batch(50, 4)
    rng = np.random.default_rng(5)
    blocks = [(M.OFF_W1, M.OFF_W1 + M.NF * M.NH), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
              (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)]
    idx = np.concatenate([rng.choice(np.arange(a, b), min(12, b - a), replace=False) for a, b in blocks])
    _fd_check(params, X, T, idx)


def test_padding_rows_have_zero_gradient():
    params = M.init_params(7)
    X, T = M.synthetic_batch(64, 8)
    for rounding in (None, M.bf16):
        _, _, g = M.forward_backward(params, X, T, rounding=rounding)
        assert np.all(g[M.OFF_W1 + M.NF * M.NH:M.OFF_B1] == 0.0)


def test_bf16_rounding_bit_patterns():
    cases = {
        0x3F800000: 0x3F800000,  # 1.0 exact
        0x3F808000: 0x3F800000,  # halfway, even stays
        0x3F818000: 0x3F820000,  # halfway, odd rounds up
        0x3F808001: 0x3F810000,  # above halfway
        0x3F807FFF: 0x3F800000,  # below halfway
        0x3FFFFFFF: 0x40000000,  # carry into the exponent
        0xBF818000: 0xBF820000,  # sign preserved
    }
    src = np.array(list(cases.keys()), np.uint32).view(np.float32)
    out = M.bf16(src).view(np.uint32)
    assert list(out) == list(cases.values())


def test_bf16_gradient_close_to_fp64():
    params = M.init_params(11)
    X, T = M.synthetic_batch(500, 12)
    _, l64, g64 = M.forward_backward(params, X, T, rounding=None)
    _, l16, g16 = M.forward_backward(params, X, T, rounding=M.bf16)
    assert abs(l64 - l16) <= 0.02 * l64
    assert np.linalg.norm(g16 - g64) <= 0.02 * np.linalg.norm(g64)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(1)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    g[:5] = 0.0
    lr = 1e-3
    p1, m, v = M.adam(p, np.zeros(100), np.zeros(100), g, 1, lr)
    # step 1: m_hat = g, v_hat = g^2, update = -lr * g / (|g| + eps)
    expect = p - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=0, atol=1e-7)
    # step 2 against a direct float64 evaluation
    g2 = rng.normal(size=100).astype(np.float32)
    p2, _, _ = M.adam(p1, m, v, g2, 2, lr)
    m2 = 0.9 * (0.1 * g.astype(np.float64)) + 0.1 * g2
    v2 = 0.999 * (0.001 * g.astype(np.float64) ** 2) + 0.001 * g2.astype(np.float64) ** 2
    ref = p1 - lr * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + 1e-8)
    np.testing.assert_allclose(p2, ref, rtol=0, atol=2e-6)


def test_gradient_descent_reduces_loss():
    params = M.init_params(21).astype(np.float64)
    X, T = M.synthetic_batch(256, 22)
    l0 = M.loss(params, X, T)
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    for step in range(1, 31):
        _, _, g = M.forward_backward(params, X, T, rounding=M.bf16)
        params, m, v = M.adam(params, m, v, g, step, 3e-3)
        params = params.astype(np.float64)
    assert M.loss(params, X, T) < 0.5 * l0


@pytest.mark.parametrize("n", [1, 31, 33])
def test_shapes(n):
    params = M.init_params(2)
    X, T = M.synthetic_batch(n, 3)
    y, loss, g = M.forward_backward(params, X, T, rounding=M.bf16)
    assert y.shape == (n,) and g.shape == (M.NPARAM,) and np.isfinite(loss)
