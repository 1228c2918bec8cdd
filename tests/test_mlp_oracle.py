"""Pins of the adaptive-grid network oracle (oracle/mlp_oracle.py), CPU only."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_round_rules():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.4e38, 0.0],
                 dtype=np.float32)
    want = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, np.inf, 0.0]
    np.testing.assert_array_equal(M.bf16_round(x), want)
    assert np.isnan(M.bf16_round(np.array([np.nan], dtype=np.float32)))[0]


def test_bf16_round_is_nearest_even_brute_force():
    """Against the two bfloat16 neighbours of each value: nearest, ties to the even one."""
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000)).astype(np.float32)
    u = x.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)                 # toward zero
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - lo), np.abs(hi - xd)
    lo_even = ((u >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo_even, lo, hi)))
    np.testing.assert_array_equal(M.bf16_round(x), want)


def test_gelu_closed_form_and_derivative():
    z = np.linspace(-6, 6, 61)
    g = M.gelu(z)
    assert g[30] == 0.0
    np.testing.assert_allclose(g[-1], 6.0, atol=1e-7)
    np.testing.assert_allclose(g[0], 0.0, atol=1e-7)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)


def small_net(rng, sizes):
    return [(rng.standard_normal((a, b)) / math.sqrt(a), 0.1 * rng.standard_normal(b))
            for a, b in zip(sizes[:-1], sizes[1:])]


def test_gradients_match_finite_differences():
    """Exact arithmetic (no bf16 rounding): every dW and db of the Gaussian NLL agrees with
    central differences, so a dropped GELU' factor, a transposed product or a wrong NLL
    gradient fails."""
    rng = np.random.default_rng(1)
    P = 3
    params = small_net(rng, [5, 7, 6, 2 * P])
    X = rng.standard_normal((4, 5))
    T = rng.random((4, P))
    loss, grads = M.loss_and_grads(params, X, T, rounding=False)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = M.loss_and_grads(params, X, T, rounding=False)
                arr[idx] = old - h
                lm, _ = M.loss_and_grads(params, X, T, rounding=False)
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def test_nll_clamp_and_minimum():
    T = np.array([[0.3, 0.7]])
    O = np.array([[0.3, 0.7, 25.0, -25.0]])
    loss, G = M.nll(O, T)
    assert G[0, 2] == 0.0 and G[0, 3] == 0.0 and G[0, 0] == 0.0 and G[0, 1] == 0.0
    # at mu = T the loss is minimised in s at s -> -inf; at |r| = 1 the optimum is s = 0
    O = np.array([[T[0, 0] + 1.0, T[0, 1] - 1.0, 0.0, 0.0]])
    _, G = M.nll(O, T)
    np.testing.assert_allclose(G[0, 2:], 0.0, atol=1e-15)


def test_adam_first_step_is_sign_step():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -7.0, 1e-3])
    w1, m, v = M.adam(w, g, np.zeros(3), np.zeros(3), 1e-2, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(w1, w - 1e-2 * np.sign(g), rtol=1e-9)
