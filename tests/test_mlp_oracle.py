"""The FP64 warm-start-network oracle (oracle/mlp_ref.py) pinned by
mathematics: central finite differences of the loss for every parameter
block, the GELU derivative, the BF16 rounding rule, and closed forms of a
network reduced to a linear map."""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

import mlp_ref as M  # noqa: E402


def _data(rng, n):
    x = rng.standard_normal((n, 16)) * 2.0 + 0.5
    t = rng.standard_normal((n, 2))
    return x, t


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = M.random_params(rng)
    x, t = _data(rng, 12)
    _, g = M.loss_and_grad(p, x, t)
    base = M.unpack(p)
    h = 1e-4
    for k in M.WEIGHTS:
        idx = rng.choice(base[k].size, size=min(6, base[k].size), replace=False)
        for i in idx:
            plus = {kk: v.copy() for kk, v in base.items()}
            minus = {kk: v.copy() for kk, v in base.items()}
            plus[k].reshape(-1)[i] += h
            minus[k].reshape(-1)[i] -= h
            fd = (M.loss(plus, x, t) - M.loss(minus, x, t)) / (2 * h)
            an = g[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(an), (k, i, fd, an)


def test_gelu_derivative():
    x = np.linspace(-6, 6, 241)
    fd = (M.gelu(x + 1e-6) - M.gelu(x - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(x), fd, atol=1e-8)
    # gelu(0) = 0, gelu(x) -> x for large x, -> 0 for very negative x
    assert M.gelu(np.array(0.0)) == 0.0
    assert abs(M.gelu(np.array(10.0)) - 10.0) < 1e-12
    assert abs(M.gelu(np.array(-10.0))) < 1e-12


def test_bf16_rounding_rule():
    one = np.float32(1.0)
    half_ulp = np.float32(2.0 ** -8)  # BF16 has 7 stored mantissa bits
    assert M.bf16(one) == one
    assert M.bf16(one + half_ulp) == one                           # tie -> even (1.0)
    assert M.bf16(one + 3 * half_ulp) == one + np.float32(2 ** -6)  # tie -> even (mantissa 2)
    assert M.bf16(one + np.float32(1.25) * half_ulp) == one + np.float32(2 ** -7)
    assert M.bf16(np.float32(-3.0)) == np.float32(-3.0)
    r = np.random.default_rng(1).standard_normal(10000).astype(np.float32)
    b = M.bf16(r)
    assert np.all(b.view(np.uint32) & 0xFFFF == 0)
    assert np.all(np.abs(b - r) <= np.abs(r) * 2.0 ** -8 + 1e-30)


def test_linear_network_closed_form():
    """With W2 = W3 = 0 the network is y' = W4 gelu(b3) + b4: constant in x,
    so its loss gradient with respect to b4 is mean(y' - t') exactly."""
    rng = np.random.default_rng(3)
    p = M.random_params(rng, norm=False)
    p["W2"] = 0.0
    p["W3"] = 0.0
    x, t = _data(rng, 20)
    y, c = M.forward(p, x)
    assert np.allclose(y, y[0])
    L, g = M.loss_and_grad(p, x, t)
    assert np.allclose(g["b4"], np.mean(c["ys"] - t, axis=0))
    assert np.allclose(g["W1"], 0.0) and np.allclose(g["W2"], 0.0)
    assert np.isclose(L, 0.5 * np.mean(np.sum((c["ys"] - t) ** 2, axis=1)))


def test_normalisation_is_applied():
    """Standardising the inputs and de-standardising the outputs equals a
    network with identity normalisation fed standardised data."""
    rng = np.random.default_rng(5)
    p = M.random_params(rng)
    q = p.copy()
    q["xmean"] = 0.0
    q["xstd"] = 1.0
    q["ymean"] = 0.0
    q["ystd"] = 1.0
    x, _ = _data(rng, 9)
    xs = (x - p["xmean"].astype(np.float64)) * np.where(p["xstd"] > 0, 1.0 / np.maximum(p["xstd"], 1e-30), 0.0)
    y1, _ = M.forward(p, x)
    y2, _ = M.forward(q, xs)
    assert np.allclose(y1, y2 * p["ystd"] + p["ymean"])
