"""Per-function parity of the gfx950 kernels against the CPU oracle, called
through the C ABI (include/b2g.h) on host arrays.

  narrowphase   b2g_collide_batch        vs orc_collide          bit-exact (ids, counts, points, normals)
  world manif.  b2g_world_manifold_batch vs orc_world_manifold   bit-exact
  pair set      b2g_pairs_batch          vs orc_pairs_bruteforce identical sorted (min, max) lists
  fat AABB rule b2g_fat_rule_batch       vs orc_fat_rule         bit-exact box and move flag
  trig          b2g_sincos_batch         vs rounded double sin/cos bit-exact; vs glibc <= 1 ulp
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import corpus
import engines as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b2g():
    lib = ctypes.CDLL(E.GPU_B2G_LIB)
    lib.b2g_last_error.restype = ctypes.c_char_p
    ctx = ctypes.c_void_p()
    rc = lib.b2g_create(ctypes.byref(ctx), 0)
    assert rc == 0, f"b2g_create failed rc={rc}"
    yield lib, ctx
    lib.b2g_destroy(ctx)


def _check(lib, ctx, rc):
    assert rc == 0, lib.b2g_last_error(ctx).decode()


def _gpu_collide(lib, ctx, A, XA, B, XB):
    out = np.zeros(len(A), E.MANIFOLD)
    _check(lib, ctx, lib.b2g_collide_batch(ctx, E.ptr(A), E.ptr(XA), E.ptr(B), E.ptr(XB), len(A), E.ptr(out)))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_collide_batch_bitexact(b2g, orclib, seed):
    lib, ctx = b2g
    A, XA, B, XB = corpus.random_corpus(orclib, "orc_", 600, seed)
    want = corpus.collide_all(orclib, "orc_", A, XA, B, XB)
    got = _gpu_collide(lib, ctx, A, XA, B, XB)
    bad = np.where(~np.all(E.geometry_bits(got) == E.geometry_bits(want), axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(A)} manifolds differ; first {bad[:8]} types {A['type'][bad[:8]]}/{B['type'][bad[:8]]}"
    # every contact type produced touching manifolds
    for t in range(5):
        sl = slice(t * 600, (t + 1) * 600)
        assert (want["count"][sl] > 0).sum() > 20, t


def test_collide_special_cases(b2g, orclib):
    lib, ctx = b2g
    A, XA, B, XB = corpus.special_corpus(orclib, "orc_")
    want = corpus.collide_all(orclib, "orc_", A, XA, B, XB)
    got = _gpu_collide(lib, ctx, A, XA, B, XB)
    assert np.array_equal(E.geometry_bits(got), E.geometry_bits(want))


def test_world_manifold_batch(b2g, orclib):
    lib, ctx = b2g
    A, XA, B, XB = corpus.random_corpus(orclib, "orc_", 300, 9)
    m = corpus.collide_all(orclib, "orc_", A, XA, B, XB)
    keep = m["count"] > 0
    m, XA, XB, A, B = m[keep], XA[keep], XB[keep], A[keep], B[keep]
    ra = A["radius"].astype(np.float32).copy()
    rb = B["radius"].astype(np.float32).copy()
    got = np.zeros((len(m), 8), np.float32)
    _check(lib, ctx, lib.b2g_world_manifold_batch(ctx, E.ptr(m), E.ptr(XA), E.ptr(ra), E.ptr(XB), E.ptr(rb),
                                                  len(m), E.ptr(got)))
    want = np.zeros((len(m), 8), np.float32)
    for i in range(len(m)):
        orclib.orc_world_manifold(E.ptr(m[i:i + 1]), E.ptr(XA[i:i + 1]), ctypes.c_float(ra[i]),
                                  E.ptr(XB[i:i + 1]), ctypes.c_float(rb[i]), E.ptr(want[i]))
    one = m["count"] == 1
    want[one, 4:6] = 0
    want[one, 7] = 0
    got[one, 4:6] = 0
    got[one, 7] = 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _boxes(rng, n, spread, lo_size, hi_size, wide_every=0):
    lo = rng.uniform(-spread, spread, (n, 2)).astype(np.float32)
    size = rng.uniform(lo_size, hi_size, (n, 2)).astype(np.float32)
    if wide_every:
        size[::wide_every] *= 40
    a = np.zeros(n, E.AABB)
    a["lx"], a["ly"] = lo[:, 0], lo[:, 1]
    a["ux"], a["uy"] = lo[:, 0] + size[:, 0], lo[:, 1] + size[:, 1]
    return a


@pytest.mark.parametrize("case", ["uniform", "dense", "wide", "touching", "sparse_moves", "large"])
def test_pairs_batch_matches_bruteforce(b2g, orclib, case):
    lib, ctx = b2g
    rng = np.random.default_rng(hash(case) % 1000)
    n = 4000 if case == "large" else 1500
    if case == "uniform":
        fat = _boxes(rng, n, 60, 0.3, 2.5)
    elif case == "dense":
        fat = _boxes(rng, n, 12, 0.3, 2.0)
    elif case == "wide":
        fat = _boxes(rng, n, 40, 0.2, 2.0, wide_every=97)
    elif case == "touching":
        g = np.arange(n)
        fat = np.zeros(n, E.AABB)
        fat["lx"] = (g % 40).astype(np.float32)
        fat["ly"] = (g // 40).astype(np.float32)
        fat["ux"] = fat["lx"] + 1.0
        fat["uy"] = fat["ly"] + 1.0
    else:
        fat = _boxes(rng, n, 50, 0.3, 2.5)
    moved = (rng.random(n) < (0.1 if case == "sparse_moves" else 1.0)).astype(np.uint8)
    cap = 1 << 21
    want = np.zeros(2 * cap, np.int32)
    nw = orclib.orc_pairs_bruteforce(E.ptr(fat), E.ptr(moved), n, E.ptr(want), cap)
    got = np.zeros(2 * cap, np.int32)
    cnt = ctypes.c_int(0)
    _check(lib, ctx, lib.b2g_pairs_batch(ctx, E.ptr(fat), E.ptr(moved), n, E.ptr(got), cap, ctypes.byref(cnt)))
    assert cnt.value == nw
    assert np.array_equal(got[:2 * nw], want[:2 * nw])


def test_fat_rule_batch(b2g, orclib):
    lib, ctx = b2g
    rng = np.random.default_rng(3)
    n = 20000
    tight = _boxes(rng, n, 50, 0.1, 2.0)
    fat = tight.copy()
    grow = rng.uniform(0.0, 0.3, (n, 4)).astype(np.float32)
    fat["lx"] -= grow[:, 0]; fat["ly"] -= grow[:, 1]; fat["ux"] += grow[:, 2]; fat["uy"] += grow[:, 3]
    d = rng.uniform(-0.25, 0.25, (n, 2)).astype(np.float32)
    tight["lx"] += d[:, 0]; tight["ux"] += d[:, 0]; tight["ly"] += d[:, 1]; tight["uy"] += d[:, 1]
    got = np.zeros(n, E.AABB)
    mv = np.zeros(n, np.uint8)
    _check(lib, ctx, lib.b2g_fat_rule_batch(ctx, E.ptr(fat), E.ptr(tight), E.ptr(d), n, E.ptr(got), E.ptr(mv)))
    want = np.zeros(n, E.AABB)
    wm = np.zeros(n, np.uint8)
    for i in range(n):
        wm[i] = orclib.orc_fat_rule(E.ptr(fat[i:i + 1]), E.ptr(tight[i:i + 1]), ctypes.c_float(d[i, 0]),
                                    ctypes.c_float(d[i, 1]), E.ptr(want[i:i + 1]))
    assert 0.05 < wm.mean() < 0.95
    assert np.array_equal(mv, wm)
    assert got.tobytes() == want.tobytes()


def test_sincos_batch(b2g, orclib):
    lib, ctx = b2g
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.uniform(-7, 7, 400000), rng.uniform(-200, 200, 100000),
                        np.array([0.0, -0.0, np.pi / 2, np.pi, 1e-8, -1e-8])]).astype(np.float32)
    s = np.zeros_like(a)
    c = np.zeros_like(a)
    _check(lib, ctx, lib.b2g_sincos_batch(ctx, E.ptr(a), len(a), E.ptr(s), E.ptr(c)))
    ws = np.sin(a.astype(np.float64)).astype(np.float32)
    wc = np.cos(a.astype(np.float64)).astype(np.float32)
    ulp = lambda x, y: np.abs(x.view(np.int32).astype(np.int64) - y.view(np.int32).astype(np.int64))
    assert ulp(s, ws).max() == 0 and ulp(c, wc).max() == 0
    gs, gc = np.zeros_like(a), np.zeros_like(a)
    orclib.orc_sincos(E.ptr(a), len(a), E.ptr(gs), E.ptr(gc))
    assert ulp(s, gs).max() <= 1 and ulp(c, gc).max() <= 1
