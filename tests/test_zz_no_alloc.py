"""b2g_pairs_batch_noalloc, the pair set without device allocation, against
b2g_pairs_batch through the C ABI: identical sorted (min, max) pair lists
and counts on two calls in a row, the second with a smaller input (the
scratch array keeps the first call's contents, so anything that must start
at zero has to be cleared by the call), plus its size limits and a pair
count above the caller's capacity.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

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


def _boxes(rng, n, spread, lo_size, hi_size, wide_every=0):
    lo = rng.uniform(-spread, spread, (n, 2)).astype(np.float32)
    size = rng.uniform(lo_size, hi_size, (n, 2)).astype(np.float32)
    if wide_every:
        size[::wide_every] *= 40
    a = np.zeros(n, E.AABB)
    a["lx"], a["ly"] = lo[:, 0], lo[:, 1]
    a["ux"], a["uy"] = lo[:, 0] + size[:, 0], lo[:, 1] + size[:, 1]
    return a


def _pairs(lib, ctx, fn, fat, moved, cap):
    out = np.full(2 * max(cap, 1), -1, np.int32)
    cnt = ctypes.c_int(-1)
    rc = getattr(lib, fn)(ctx, E.ptr(fat), E.ptr(moved), len(fat), E.ptr(out), cap, ctypes.byref(cnt))
    assert rc == 0, lib.b2g_last_error(ctx).decode()
    return cnt.value, out


def test_noalloc_matches_allocating_entry_point(b2g):
    lib, ctx = b2g
    rng = np.random.default_rng(7)
    cases = [
        (_boxes(rng, 20000, 150, 0.3, 2.5, wide_every=211), 1.0),  # the larger call first
        (_boxes(rng, 3000, 40, 0.3, 2.0), 0.3),                    # then a smaller one
    ]
    for fat, share in cases:
        moved = (rng.random(len(fat)) < share).astype(np.uint8)
        print(f"case: pairs, {len(fat)} proxies, {share:.0%} moved", flush=True)
        cap = 1 << 20
        nw, want = _pairs(lib, ctx, "b2g_pairs_batch", fat, moved, cap)
        ng, got = _pairs(lib, ctx, "b2g_pairs_batch_noalloc", fat, moved, cap)
        assert nw > 1000
        assert ng == nw
        assert np.array_equal(got[:2 * ng], want[:2 * nw])
        assert np.all(got[2 * ng:2 * ng + 8] == -1)  # nothing written past the pairs
        # fewer slots than pairs: the count is the full count, the first cap pairs are written
        small = nw // 3
        ns, part = _pairs(lib, ctx, "b2g_pairs_batch_noalloc", fat, moved, small)
        assert ns == nw
        assert np.array_equal(part[:2 * small], want[:2 * small])
        assert np.all(part[2 * small:] == -1)


def test_noalloc_limits(b2g):
    lib, ctx = b2g
    fat = _boxes(np.random.default_rng(1), 65537, 500, 0.3, 1.0)
    moved = np.ones(len(fat), np.uint8)
    out = np.zeros(2, np.int32)
    cnt = ctypes.c_int(0)
    rc = lib.b2g_pairs_batch_noalloc(ctx, E.ptr(fat), E.ptr(moved), len(fat), E.ptr(out), 1, ctypes.byref(cnt))
    assert rc != 0 and b"65536" in lib.b2g_last_error(ctx)
    # empty input: no pairs, no device work
    _, empty = _pairs(lib, ctx, "b2g_pairs_batch_noalloc", fat[:0], moved[:0], 4)
    assert np.all(empty == -1)
