"""ORACLE TEST INFRASTRUCTURE -- not part of the product; only tests/ may use it.

FP64 restatement of the warm-start network of the learned warm start
(16-64-64-64-2 perceptron, GELU with the Gaussian CDF, mean squared error on
standardised targets) and of its gradient, written from the definitions:

    x' = (x - xmean) * xinv          (xinv = 1/xstd, 0 where xstd == 0)
    z1 = W1 x' + b1, a1 = gelu(z1), z2 = W2 a1 + b2, a2 = gelu(z2),
    z3 = W3 a2 + b3, a3 = gelu(z3), y' = W4 a3 + b4, y = y' * ystd + ymean
    L  = 1/(2N) sum_i |y'_i - (t_i - ymean) / ystd|^2

The reference has no counterpart (its solver starts new contact points at
zero impulse, REF/Dynamics/Contacts/b2ContactSolver.cpp:100-109).

With emulate=True the BF16 roundings of the matrix-core path are applied at
the same points as the device: weights, standardised inputs, activations,
and the output-gradient tiles fed to the products (round to nearest even).
The FP64 gradient itself is pinned by central finite differences in the
tests.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, HIDDEN, OUT = 16, 64, 2

PARAMS_DTYPE = np.dtype([
    ("W1", np.float32, (64 * 16,)), ("b1", np.float32, (64,)),
    ("W2", np.float32, (64 * 64,)), ("b2", np.float32, (64,)),
    ("W3", np.float32, (64 * 64,)), ("b3", np.float32, (64,)),
    ("W4", np.float32, (2 * 64,)), ("b4", np.float32, (2,)),
    ("xmean", np.float32, (16,)), ("xstd", np.float32, (16,)),
    ("ymean", np.float32, (2,)), ("ystd", np.float32, (2,)),
])
WEIGHTS = ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")
SHAPES = {"W1": (64, 16), "b1": (64,), "W2": (64, 64), "b2": (64,), "W3": (64, 64), "b3": (64,),
          "W4": (2, 64), "b4": (2,)}


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(p) -> dict:
    """b2g_mlp_params record (or an already unpacked dict) -> dict of float64
    arrays in matrix shapes."""
    if isinstance(p, dict):
        return p
    return {k: np.asarray(p[k], dtype=np.float64).reshape(SHAPES[k]) for k in WEIGHTS} | {
        k: np.asarray(p[k], dtype=np.float64) for k in ("xmean", "xstd", "ymean", "ystd")}


def random_params(rng: np.random.Generator, norm: bool = True) -> np.ndarray:
    p = np.zeros((), PARAMS_DTYPE)
    for k in WEIGHTS:
        shape = SHAPES[k]
        fan = shape[-1] if len(shape) == 2 else 64
        p[k] = (rng.standard_normal(int(np.prod(shape))) / np.sqrt(fan)).astype(np.float32)
    if norm:
        p["xmean"] = rng.standard_normal(16).astype(np.float32)
        p["xstd"] = rng.uniform(0.5, 2.0, 16).astype(np.float32)
        p["xstd"][15] = 0.0  # an ignored feature
        p["ymean"] = rng.standard_normal(2).astype(np.float32)
        p["ystd"] = rng.uniform(0.5, 2.0, 2).astype(np.float32)
    else:
        p["xstd"] = 1.0
        p["ystd"] = 1.0
    return p


def _standardise(P, x, emulate):
    xinv = np.where(P["xstd"] > 0, 1.0 / np.where(P["xstd"] > 0, P["xstd"], 1.0), 0.0)
    if emulate:
        xinv32 = np.where(P["xstd"] > 0, (np.float32(1.0) / np.where(P["xstd"] > 0, P["xstd"], 1.0).astype(np.float32)),
                          np.float32(0.0)).astype(np.float32)
        xs = (x.astype(np.float32) - P["xmean"].astype(np.float32)) * xinv32
        return bf16(xs).astype(np.float64)
    return (x - P["xmean"]) * xinv


def forward(p, x, emulate=False):
    """Outputs (N x 2, raw units) and the cache of the backward pass."""
    P = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    r = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: a)
    W = {k: r(P[k]) if k.startswith("W") else P[k] for k in WEIGHTS}
    xs = _standardise(P, x, emulate)
    z1 = xs @ W["W1"].T + W["b1"]
    a1 = r(gelu(z1))
    z2 = a1 @ W["W2"].T + W["b2"]
    a2 = r(gelu(z2))
    z3 = a2 @ W["W3"].T + W["b3"]
    a3 = r(gelu(z3))
    ys = a3 @ W["W4"].T + W["b4"]
    y = ys * P["ystd"] + P["ymean"]
    return y, dict(P=P, W=W, xs=xs, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, ys=ys)


def loss_and_grad(p, x, t, emulate=False):
    """L and dL/d(W1, b1, ..., W4, b4) as a dict of float64 arrays."""
    y, c = forward(p, x, emulate)
    P, W = c["P"], c["W"]
    r = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: a)
    n = len(x)
    ts = (np.asarray(t, dtype=np.float64) - P["ymean"]) / P["ystd"]
    e = c["ys"] - ts
    L = 0.5 * np.sum(e * e) / n
    d4 = e / n
    g = {"W4": r(d4).T @ c["a3"], "b4": d4.sum(0)}
    d3 = (r(d4) @ W["W4"]) * gelu_grad(c["z3"])
    g["W3"], g["b3"] = r(d3).T @ c["a2"], d3.sum(0)
    d2 = (r(d3) @ W["W3"]) * gelu_grad(c["z2"])
    g["W2"], g["b2"] = r(d2).T @ c["a1"], d2.sum(0)
    d1 = (r(d2) @ W["W2"]) * gelu_grad(c["z1"])
    g["W1"], g["b1"] = r(d1).T @ c["xs"], d1.sum(0)
    return L, g


def loss(p, x, t):
    return loss_and_grad(p, x, t)[0]
