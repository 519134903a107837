"""Learned warm start on the GPU: the BF16 matrix-core network against the
FP64 oracle (oracle/mlp_ref.py, pinned by finite differences), the fresh-point
features and the FP64 residual trace against closed forms, the fallback, and
what a trained network gains over the plain warm start."""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np
import pytest

import engines as E

sys.path.insert(0, os.path.join(E.ROOT, "oracle"))
import mlp_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return E.b2g_lib()


@pytest.fixture
def ctx(lib):
    c = ctypes.c_void_p()
    assert lib.b2g_create(ctypes.byref(c), 0) == 0
    yield c
    lib.b2g_destroy(c)


def _set(lib, ctx, p):
    E.b2g_check(lib, ctx, lib.b2g_mlp_set_params(ctx, E.ptr(np.atleast_1d(p))))


def _inputs(rng, n):
    x = (rng.standard_normal((n, 16)) * 1.5 + rng.standard_normal(16)).astype(np.float32)
    t = rng.standard_normal((n, 2)).astype(np.float32)
    return x, t


def test_forward_matches_oracle(lib, ctx):
    rng = np.random.default_rng(11)
    p = M.random_params(rng)
    _set(lib, ctx, p)
    for n in (1, 15, 16, 1000, 4099):
        x, _ = _inputs(rng, n)
        y = np.zeros((n, 2), np.float32)
        E.b2g_check(lib, ctx, lib.b2g_mlp_forward(ctx, E.ptr(x), n, E.ptr(y)))
        ye, _ = M.forward(p, x, emulate=True)
        yx, _ = M.forward(p, x)
        # BF16 roundings emulated: FP32 accumulation order is what differs,
        # except where an activation lies on a BF16 rounding boundary and
        # rounds the other way (one BF16 step of one activation)
        err = np.max(np.abs(y - ye) / (1.0 + np.abs(ye)), axis=1)
        assert np.sum(err > 2e-6) <= max(1, 0.02 * n), n
        assert err.max() < 5e-3, n
        # against exact FP64: the BF16 error of a 3-hidden-layer network
        assert np.sqrt(np.mean((y - yx) ** 2)) < 3e-2 * np.sqrt(np.mean(yx ** 2)) + 1e-3, n


def test_gradient_matches_oracle(lib, ctx):
    """The training kernel's gradient against the oracle with the BF16
    roundings emulated (tight) and against exact FP64 (the BF16 bound)."""
    rng = np.random.default_rng(12)
    p = M.random_params(rng)
    _set(lib, ctx, p)
    n = 777
    x, t = _inputs(rng, n)
    g = np.zeros(1, M.PARAMS_DTYPE)
    loss = ctypes.c_double()
    E.b2g_check(lib, ctx, lib.b2g_mlp_gradient(ctx, E.ptr(x), E.ptr(t), n, E.ptr(g), ctypes.byref(loss)))
    Le, ge = M.loss_and_grad(p, x, t, emulate=True)
    Lx, gx = M.loss_and_grad(p, x, t)
    assert abs(loss.value - Le) <= 1e-4 * Le
    assert abs(loss.value - Lx) <= 2e-2 * Lx
    for k in M.WEIGHTS:
        gk = g[0][k].astype(np.float64).reshape(M.SHAPES[k])
        ek = ge[k]
        xk = gx[k]
        rel_e = np.linalg.norm(gk - ek) / np.linalg.norm(ek)
        rel_x = np.linalg.norm(gk - xk) / np.linalg.norm(xk)
        cos = float(np.sum(gk * xk) / (np.linalg.norm(gk) * np.linalg.norm(xk)))
        assert rel_e < 3e-3, (k, rel_e)
        assert rel_x < 5e-2 and cos > 0.998, (k, rel_x, cos)


def test_gradient_is_deterministic(lib, ctx):
    """Workgroup sums are added up in a fixed order and the wavefronts of a
    workgroup in a fixed order: the same batch gives the same gradient bits
    (enough samples for many workgroups)."""
    rng = np.random.default_rng(21)
    _set(lib, ctx, M.random_params(rng))
    n = 40000
    x, t = _inputs(rng, n)
    out = []
    for _ in range(3):
        g = np.zeros(1, M.PARAMS_DTYPE)
        loss = ctypes.c_double()
        E.b2g_check(lib, ctx, lib.b2g_mlp_gradient(ctx, E.ptr(x), E.ptr(t), n, E.ptr(g), ctypes.byref(loss)))
        out.append((g.view(np.uint8).copy(), loss.value))
    assert all(np.array_equal(o[0], out[0][0]) and o[1] == out[0][1] for o in out)


def test_adam_training_fits_a_smooth_map(lib, gpu):
    """Log fresh points from a pile, then train: the loss falls well below
    the loss of predicting the mean."""
    w = gpu.world("world", 2000, 4, sleep=False)
    c = w.device_context()
    E.b2g_check(lib, c, lib.b2g_log_fresh(c, 400000))
    w.step(150)
    n = ctypes.c_int()
    lib.b2g_log_size(c, ctypes.byref(n))
    assert n.value > 5000
    feat = np.zeros((n.value, 16), np.float32)
    tgt = np.zeros((n.value, 2), np.float32)
    E.b2g_check(lib, c, lib.b2g_get_log(c, 0, n.value, E.ptr(feat), E.ptr(tgt)))
    assert np.all(np.isfinite(feat)) and np.all(np.isfinite(tgt))
    assert set(np.unique(feat[:, 9])) <= {1.0, 2.0}          # point count
    assert np.all(tgt[:, 0] >= 0.0)                            # normal impulse
    E.b2g_check(lib, c, lib.b2g_mlp_init(c, 1))
    losses = np.zeros(12, np.float64)
    E.b2g_check(lib, c, lib.b2g_mlp_train(c, 12, 4096, ctypes.c_float(2e-3), 7, E.ptr(losses)))
    assert np.all(np.isfinite(losses))
    # normalised targets have unit variance: predicting the mean scores
    # L = 1/2 (1 + 1) = 1; the network explains a good part of the variance
    # and keeps improving over the epochs
    assert losses[-1] < 0.7 and losses[-1] < 0.8 * losses[0], losses
    w.close()


def test_fresh_features_closed_form(lib, gpu):
    """HelloWorld's box landing on the ground box: at the first touching step
    both points are fresh; their features are the closed-form masses, the
    ground normal and the box's velocity after gravity integration."""
    w = gpu.world("hello", sleep=False)
    c = w.device_context()
    E.b2g_check(lib, c, lib.b2g_log_fresh(c, 1000))
    n = ctypes.c_int()
    prev = None
    for s in range(200):
        prev = w.bodies()
        w.step(1, vel=6, pos=2)
        lib.b2g_log_size(c, ctypes.byref(n))
        if n.value > 0:
            break
    assert n.value == 2, n.value
    f = np.zeros((2, 16), np.float32)
    t = np.zeros((2, 2), np.float32)
    E.b2g_check(lib, c, lib.b2g_get_log(c, 0, 2, E.ptr(f), E.ptr(t)))
    vy = np.float32(prev[1, E.VY]) + np.float32(1.0 / 60.0) * np.float32(-10.0)
    for j in range(2):
        assert f[j, 0] == pytest.approx(vy, abs=1e-6)           # relative normal velocity
        assert f[j, 2] == 0.0 and f[j, 3] == 1.0                 # ground normal
        assert f[j, 5] == 0.0 and f[j, 7] == 0.0                 # static ground
        assert f[j, 6] == pytest.approx(0.25)                    # 1 / (density * 2 * 2)
        assert f[j, 8] == pytest.approx(1.0 / (4.0 * 8.0 / 12.0))  # 1 / (m (w^2 + h^2) / 12)
        assert f[j, 9] == 2.0
        assert f[j, 12] == pytest.approx(np.sqrt(np.float32(0.3) * np.float32(0.2)))
        assert t[j, 0] > 0.0
    w.close()


def _trace(lib, c):
    cur = np.zeros(3 * 257, np.float64)
    it = ctypes.c_int()
    E.b2g_check(lib, c, lib.b2g_get_trace(c, E.ptr(cur), len(cur), ctypes.byref(it)))
    T = it.value
    return cur[:3 * (T + 1)].reshape(3, T + 1)


def test_trace_closed_form_single_box(lib, gpu):
    """At the landing step of HelloWorld's box, before any iteration, both
    points approach at |vn| with zero impulse, so the cold and warm
    residuals equal |vn|; the solver then drives the residual to zero."""
    w = gpu.world("hello", sleep=False)
    c = w.device_context()
    E.b2g_check(lib, c, lib.b2g_log_fresh(c, 1000))
    n = ctypes.c_int()
    for s in range(200):
        prev = w.bodies()
        E.b2g_check(lib, c, lib.b2g_set_trace(c, 32))
        w.step(1, vel=6, pos=2)
        lib.b2g_log_size(c, ctypes.byref(n))
        if n.value > 0:
            break
    cur = _trace(lib, c)
    vy = float(np.float32(prev[1, E.VY]) + np.float32(1.0 / 60.0) * np.float32(-10.0))
    assert cur[0, 0] == pytest.approx(abs(vy), rel=1e-5)
    assert cur[1, 0] == pytest.approx(abs(vy), rel=1e-5)
    assert np.isnan(cur[2]).all()            # no network loaded
    assert cur[0, 32] < 1e-5 and cur[1, 32] < 1e-5
    lib.b2g_set_trace(c, 0)
    w.close()


def test_traced_step_leaves_the_step_unchanged(lib, gpu):
    a = gpu.world("world", 1500, 6, sleep=False)
    b = gpu.world("world", 1500, 6, sleep=False)
    a.step(80)
    b.step(80)
    ba, bb = a.bodies(), b.bodies()
    diff = np.where(~np.all(ba.view(np.uint32) == bb.view(np.uint32), axis=1))[0]
    assert len(diff) == 0, f"two identical worlds differ before tracing: {len(diff)} bodies, " \
                           f"max {np.abs(ba - bb).max()}"
    ca = a.device_context()
    E.b2g_check(lib, ca, lib.b2g_set_trace(ca, 12))
    a.step(3)
    b.step(3)
    cur = _trace(lib, ca)
    assert np.all(np.isfinite(cur[:2]))
    ba, bb = a.bodies(), b.bodies()
    assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32))
    a.close(); b.close()


def test_fallback_restores_the_plain_step(lib, gpu):
    """A residual threshold no start can meet makes every learned step fall
    back to the plain start: the trajectory equals the plain solver's bit
    for bit. A threshold that always passes keeps the learned start."""
    rng = np.random.default_rng(3)
    p = M.random_params(rng, norm=False)
    plain = gpu.world("world", 1200, 8, sleep=False)
    strict = gpu.world("world", 1200, 8, sleep=False)
    loose = gpu.world("world", 1200, 8, sleep=False)
    for w, thr in ((strict, 1e-20), (loose, 1e9)):
        c = w.device_context()
        _set(lib, c, p)
        w.set_learned(True, thr)
    fb = 0
    for s in range(60):
        plain.step(1)
        strict.step(1)
        loose.step(1)
        st = strict.stats()
        if st["freshPoints"] > 0:
            assert st["fallback"] == 1 and st["residual"] >= 0.0
            fb += 1
        assert loose.stats()["fallback"] == 0
    assert fb > 10
    assert np.array_equal(plain.bodies().view(np.uint32), strict.bodies().view(np.uint32))
    assert not np.array_equal(plain.bodies().view(np.uint32), loose.bodies().view(np.uint32))
    plain.close(); strict.close(); loose.close()


def test_learned_start_lowers_early_residual(lib, gpu):
    """Trained on impulses logged from plain-solver piles, the network's
    start for fresh points leaves a lower residual than the plain warm start
    (fresh points at zero) after the first iterations, averaged over
    traced steps of an unseen pile."""
    lw = gpu.world("world", 3000, 21, sleep=False)
    c0 = lw.device_context()
    E.b2g_check(lib, c0, lib.b2g_log_fresh(c0, 2_000_000))
    lw.step(240)
    E.b2g_check(lib, c0, lib.b2g_mlp_init(c0, 5))
    losses = np.zeros(30, np.float64)
    E.b2g_check(lib, c0, lib.b2g_mlp_train(c0, 30, 4096, ctypes.c_float(1e-3), 9, E.ptr(losses)))
    params = np.zeros(1, M.PARAMS_DTYPE)
    E.b2g_check(lib, c0, lib.b2g_mlp_get_params(c0, E.ptr(params)))
    lw.close()
    ev = gpu.world("world", 3000, 99, sleep=False)
    c = ev.device_context()
    _set(lib, c, params[0])
    curves = []
    for s in range(200):
        if s >= 40 and s % 8 == 0:
            E.b2g_check(lib, c, lib.b2g_set_trace(c, 8))
            ev.step(1)
            E.b2g_check(lib, c, lib.b2g_set_trace(c, 0))
            if ev.stats()["freshPoints"] > 20:
                curves.append(_trace(lib, c))
        else:
            ev.step(1)
    ev.close()
    cur = np.array(curves)           # steps x 3 x 9
    assert len(cur) >= 10
    warm = cur[:, 1, :].mean(axis=0)
    learned = cur[:, 2, :].mean(axis=0)
    cold = cur[:, 0, :].mean(axis=0)
    print("cold", cold, "\nwarm", warm, "\nlearned", learned)
    # After the first iteration the learned start stays below the plain warm
    # start (before any iteration it can be above it: predicted impulses on
    # points that then separate count in the residual).
    assert np.all(learned[1:] < warm[1:]), (warm, learned)


def test_every_point_scope_features_and_targets(lib, gpu):
    """Scope 1 serves every solved point. HelloWorld's box at rest: both
    persistent points are candidates; feature 9 is the carried normal
    impulse over the point's normal mass (k_n = 1/m + (r x n)^2 / I with
    r x n = -1, +1 at the corners: 0.25 + 1 / (4 * 8 / 12) = 0.625) and the
    target is the solved impulse in the same units minus that."""
    w = gpu.world("hello", sleep=False)
    c = w.device_context()
    w.step(150, vel=6, pos=2)
    _, man = w.contacts()
    carried = np.array([man[0]["points"][j]["ni"] for j in range(2)], np.float64)
    assert np.all(carried > 0)
    E.b2g_check(lib, c, lib.b2g_set_learned_scope(c, 1, 1))
    E.b2g_check(lib, c, lib.b2g_log_fresh(c, 100))
    w.step(1, vel=6, pos=2)
    n = ctypes.c_int()
    lib.b2g_log_size(c, ctypes.byref(n))
    assert n.value == 2
    f = np.zeros((2, 16), np.float32)
    t = np.zeros((2, 2), np.float32)
    E.b2g_check(lib, c, lib.b2g_get_log(c, 0, 2, E.ptr(f), E.ptr(t)))
    _, man = w.contacts()
    solved = np.array([man[0]["points"][j]["ni"] for j in range(2)], np.float64)
    kn = 0.25 + 1.0 / (4.0 * 8.0 / 12.0)
    for j in range(2):
        jj = j if abs(f[j, 9] - carried[j] * kn) < abs(f[j, 9] - carried[1 - j] * kn) else 1 - j
        assert f[j, 9] == pytest.approx(carried[jj] * kn, rel=1e-5)
        assert f[j, 10] == pytest.approx(1.0 / kn, rel=1e-5)
        assert t[j, 0] == pytest.approx(solved[jj] * kn - f[j, 9], abs=1e-5)
    assert lib.b2g_set_learned_scope(c, 2, 1) != 0
    assert lib.b2g_set_learned_scope(c, 1, 0) != 0
    w.close()


def test_every_point_zero_network_is_the_warm_start(lib, gpu):
    """In scope 1 the network's outputs are added to the carried impulse, so
    a network that outputs zero starts every point from its carried impulse
    (up to the rounding of dividing and multiplying by the point's mass and
    the friction-cone clamp): its residual curve is the warm start's."""
    p = np.zeros(1, M.PARAMS_DTYPE)
    p["xstd"] = 1.0
    p["ystd"] = 1.0
    w = gpu.world("world", 1500, 4, sleep=False)
    c = w.device_context()
    w.step(120)
    E.b2g_check(lib, c, lib.b2g_set_learned_scope(c, 1, 1))
    _set(lib, c, p[0])
    E.b2g_check(lib, c, lib.b2g_set_trace(c, 6))
    w.step(1)
    E.b2g_check(lib, c, lib.b2g_set_trace(c, 0))
    cur = _trace(lib, c)
    assert np.all(np.isfinite(cur))
    # the clamp moves carried tangent impulses that left the cone when the
    # normal impulse dropped after the tangent update; the iterations then
    # take both starts to the same residual
    assert abs(cur[2, 0] - cur[1, 0]) <= 0.05 * cur[1, 0], cur
    assert np.allclose(cur[2, 1:], cur[1, 1:], rtol=2e-2, atol=1e-4), cur
    w.close()


@pytest.mark.parametrize("threshold", [1e-20, 1e9])
def test_persistent_solve_with_learned_start(lib, gpu, threshold):
    """The one-launch solve splits at the residual check and the impulse log
    (velocity passes, host check, position passes): with the learned start,
    falling back (threshold 1e-20) or not (1e9) and logging fresh points, it
    steps as the per-colour launches do, bit for bit, and logs the same."""
    rng = np.random.default_rng(5)
    p = M.random_params(rng, norm=False)
    ws = [gpu.world("world", 1200, 8, sleep=False) for _ in range(2)]
    for w, persistent in zip(ws, (0, 2)):
        s = np.array([(0, 0, 256, 256, 256, 1024, 0, persistent)], E.SCHEDULE)
        assert w.e.lib.b2pp_world_set_schedule(w.h, E.ptr(s)) == 0, w.error()
        c = w.device_context()
        _set(lib, c, p)
        E.b2g_check(lib, c, lib.b2g_log_fresh(c, 100000))
        w.set_learned(True, threshold)
    for _ in range(40):
        for w in ws:
            w.step(1)
    a, b = ws
    assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32))
    na, nb = ctypes.c_int(), ctypes.c_int()
    lib.b2g_log_size(a.device_context(), ctypes.byref(na))
    lib.b2g_log_size(b.device_context(), ctypes.byref(nb))
    assert na.value == nb.value > 0
    a.close(); b.close()
