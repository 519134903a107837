"""The device scan and sorts under the step (contact compaction, grid build,
colour order, new-pair order, island edges), called through the C ABI
(b2g_prim_*) on host arrays and compared with numpy:

  exclusive sum      b2g_prim_exclusive_sum               == shifted np.cumsum
  radix pair sort    b2g_prim_sort_pairs method 0         == stable argsort of key & (2^bits - 1)
  7-bit pair sort    b2g_prim_sort_pairs method 1         == stable argsort of key & 127, totals += bincount
  64-bit key sort    b2g_prim_sort_keys_u64 (bitonic <= 4096 keys, 4096-key chunks merged
                     pairwise up to 2^17, radix above)   == np.sort

All bit-exact. The sizes straddle the scan's 4096-item tiles, the sorts'
256-item tiles and the small-sort limit; repeated keys check stability.
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


def _check(lib, ctx, rc):
    assert rc == 0, lib.b2g_last_error(ctx).decode()


@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 100003, (1 << 20) + 3])
def test_exclusive_sum(b2g, n):
    lib, ctx = b2g
    rng = np.random.default_rng(n)
    x = rng.integers(-50, 100, n).astype(np.int32)
    out = np.full(n, -7, np.int32)
    _check(lib, ctx, lib.b2g_prim_exclusive_sum(ctx, E.ptr(x), n, E.ptr(out)))
    want = np.concatenate([[0], np.cumsum(x.astype(np.int64))[:-1]]).astype(np.int32)
    assert np.array_equal(out, want)


def _sort_pairs(lib, ctx, keys, bits, method, totals=None):
    n = len(keys)
    vals = np.arange(n, dtype=np.int32)
    ko = np.zeros(n, np.uint32)
    vo = np.full(n, -1, np.int32)
    tk = 0 if totals is None else len(totals)
    tp = None if totals is None else E.ptr(totals)
    _check(lib, ctx, lib.b2g_prim_sort_pairs(ctx, E.ptr(keys), E.ptr(vals), n, bits, method, E.ptr(ko), E.ptr(vo),
                                             tp, tk))
    return ko, vo


@pytest.mark.parametrize("n", [1, 255, 256, 257, 5000, 300001])
@pytest.mark.parametrize("bits,span", [(32, 1 << 32), (13, 1 << 32), (31, 1 << 31), (8, 5)])
def test_radix_sort_pairs(b2g, n, bits, span):
    """Keys above `bits` carry random high bits that the sort must ignore;
    span 5 keys repeat thousands of times (stability)."""
    lib, ctx = b2g
    rng = np.random.default_rng(n * 131 + bits)
    keys = rng.integers(0, span, n, dtype=np.uint64).astype(np.uint32)
    ko, vo = _sort_pairs(lib, ctx, keys, bits, 0)
    order = np.argsort(keys & np.uint32((1 << bits) - 1 if bits < 32 else 0xFFFFFFFF), kind="stable")
    assert np.array_equal(vo, order)
    assert np.array_equal(ko, keys[order])


@pytest.mark.parametrize("n", [1, 300, 70000, 271387])
def test_count_sort_pairs(b2g, n):
    """The colour order: 7-bit keys (65 colour slots and the key of
    constraints outside the set), per-key totals added to the caller's."""
    lib, ctx = b2g
    rng = np.random.default_rng(n)
    keys = rng.choice(np.array([0, 1, 2, 3, 5, 17, 64, 65, 127], np.uint32), n, p=[.3, .2, .1, .1, .1, .1, .05, .04, .01])
    keys |= (rng.integers(0, 4, n).astype(np.uint32) << np.uint32(7))  # ignored bits
    totals = rng.integers(0, 10, 66).astype(np.int32)
    before = totals.copy()
    ko, vo = _sort_pairs(lib, ctx, keys, 7, 1, totals)
    low = keys & np.uint32(127)
    order = np.argsort(low, kind="stable")
    assert np.array_equal(vo, order)
    assert np.array_equal(ko, keys[order])
    assert np.array_equal(totals, before + np.bincount(low, minlength=128)[:66].astype(np.int32))


@pytest.mark.parametrize("n", [1, 100, 4096, 4097, 8192, 12289, 65536, 100003, 131072, 131073, 200001])
@pytest.mark.parametrize("end_bit", [64, 40])
def test_sort_keys_u64(b2g, n, end_bit):
    lib, ctx = b2g
    rng = np.random.default_rng(n + end_bit)
    if end_bit == 64:
        keys = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    else:  # equal high bits, as the pair keys a * n + b below n * n
        keys = (rng.integers(0, 1 << end_bit, n, dtype=np.uint64) | (np.uint64(0x5A5) << np.uint64(end_bit)))
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # repeated keys
    out = np.zeros(n, np.uint64)
    _check(lib, ctx, lib.b2g_prim_sort_keys_u64(ctx, E.ptr(keys), n, end_bit, E.ptr(out)))
    assert np.array_equal(out, np.sort(keys))


def test_bad_arguments(b2g):
    lib, ctx = b2g
    k = np.zeros(4, np.uint32)
    v = np.zeros(4, np.int32)
    assert lib.b2g_prim_sort_pairs(ctx, E.ptr(k), E.ptr(v), 4, 33, 0, E.ptr(k), E.ptr(v), None, 0) != 0
    assert lib.b2g_prim_sort_pairs(ctx, E.ptr(k), E.ptr(v), 4, 7, 2, E.ptr(k), E.ptr(v), None, 0) != 0
    assert lib.b2g_prim_sort_pairs(ctx, E.ptr(k), E.ptr(v), 4, 7, 1, E.ptr(k), E.ptr(v), None, 3) != 0
    u = np.zeros(4, np.uint64)
    assert lib.b2g_prim_sort_keys_u64(ctx, E.ptr(u), 4, 0, E.ptr(u)) != 0
    assert lib.b2g_prim_exclusive_sum(ctx, None, 4, None) != 0
