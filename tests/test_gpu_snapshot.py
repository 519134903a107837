"""b2g_snapshot / b2g_restore: a world restored into a new context steps
exactly as the world it was taken from.

Each case runs a scene for a while through the box2d:: world, snapshots its
context, restores the bytes into a fresh b2g_context and then steps both
contexts through the C ABI: bodies, the contact table (fixture pairs, flags,
colours, manifolds with their impulses) and the joint impulses must agree
bit for bit after every step. The scenes cover sleeping islands, joints,
sub-worlds, a removed joint and a removed fixture.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu
BODY_WORDS = 26  # b2g_body_state: 2 int32 + 24 float


def _params():
    p = np.zeros(1, E.STEP_PARAMS)
    p[0] = (1.0 / 60.0, 8, 3, 0, 0, 0.0)
    return p


def _state(lib, ctx):
    """(contact meta, manifold bits) of a context's contact table."""
    count = ctypes.c_int(0)
    E.b2g_check(lib, ctx, lib.b2g_get_contacts(ctx, None, None, 0, ctypes.byref(count)))
    nc = count.value
    meta = np.zeros(4 * max(nc, 1), np.int32)
    man = np.zeros(max(nc, 1), E.MANIFOLD)
    E.b2g_check(lib, ctx, lib.b2g_get_contacts(ctx, E.ptr(meta), E.ptr(man), nc, ctypes.byref(count)))
    return meta[:4 * nc], man[:nc].view(np.uint32)


def _bodies(lib, ctx, n):
    out = np.zeros(BODY_WORDS * n, np.uint32)
    E.b2g_check(lib, ctx, lib.b2g_get_bodies(ctx, 0, n, E.ptr(out)))
    return out.reshape(n, BODY_WORDS)


def _joints(lib, ctx, n):
    out = np.zeros(4 * max(n, 1), np.uint32)
    if n:
        E.b2g_check(lib, ctx, lib.b2g_get_joints(ctx, 0, n, E.ptr(out)))
    return out[:4 * n]


def _snapshot(lib, ctx):
    size = ctypes.c_size_t(0)
    E.b2g_check(lib, ctx, lib.b2g_snapshot(ctx, None, ctypes.c_size_t(0), ctypes.byref(size)))
    buf = np.zeros(size.value, np.uint8)
    E.b2g_check(lib, ctx, lib.b2g_snapshot(ctx, E.ptr(buf), ctypes.c_size_t(buf.size), ctypes.byref(size)))
    assert size.value == buf.size
    return buf


def _fresh(lib):
    c = ctypes.c_void_p()
    assert lib.b2g_create(ctypes.byref(c), 0) == 0
    return c


def _edit_world(lib, w, ctx, kind, p):
    """A removed joint and a removed fixture before the snapshot, so the
    restored tables carry holes."""
    if kind == "chains":
        assert w.e.fn("world_destroy_joint")(w.h, 5) == 0
    w.step(60)
    gone = np.array([7], np.int32)
    E.b2g_check(lib, ctx, lib.b2g_remove_fixtures(ctx, E.ptr(gone), 1))
    for _ in range(5):
        E.b2g_check(lib, ctx, lib.b2g_step(ctx, E.ptr(p)))


@pytest.mark.parametrize("kind,n,arg,sleep,worlds", [
    ("pyramid", 20, 0, True, 0),
    ("chains", 4, 14, True, 0),
    ("stress", 600, 11, False, 2),
    ("pulley", 8, 0, True, 0),      # pulley constants live in the joint rows
    ("gears", 4, 0, False, 0),      # gear, prismatic and revolute joints: axis and direction rows
    ("wheels", 6, 0, False, 0),     # wheel springs keep their axis rows across steps
])
def test_restored_world_steps_identically(gpu, kind, n, arg, sleep, worlds):
    lib = E.b2g_lib()
    if worlds:
        w = E.World.worlds(gpu, kind, worlds, n, arg, sleep=sleep)
    else:
        w = gpu.world(kind, n, arg, sleep=sleep)
    w.step(60)
    a = w.device_context()
    p = _params()
    _edit_world(lib, w, a, kind, p)
    nb = w.e.fn("world_body_count")(w.h)
    nj = w.e.fn("world_joint_count")(w.h)
    snap = _snapshot(lib, a)
    b = _fresh(lib)
    try:
        E.b2g_check(lib, b, lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size)))
        # the restored context snapshots to the same bytes
        assert np.array_equal(_snapshot(lib, b), snap)
        for s in range(90):
            E.b2g_check(lib, a, lib.b2g_step(a, E.ptr(p)))
            E.b2g_check(lib, b, lib.b2g_step(b, E.ptr(p)))
            if s % 15 == 14 or s < 2:
                ba, bb = _bodies(lib, a, nb), _bodies(lib, b, nb)
                bad = np.where(~np.all(ba == bb, axis=1))[0]
                assert len(bad) == 0, f"step {s + 1}: bodies {bad[:8]} differ"
                ma, na = _state(lib, a)
                mb, nbm = _state(lib, b)
                assert np.array_equal(ma, mb), f"step {s + 1}: contact tables differ"
                assert np.array_equal(na, nbm), f"step {s + 1}: manifolds differ"
                assert np.array_equal(_joints(lib, a, nj), _joints(lib, b, nj)), f"step {s + 1}: joints differ"
        assert len(ma) > 0 or kind == "gears"  # the gear trains touch nothing
    finally:
        lib.b2g_destroy(b)
        w.close()


def test_snapshot_errors(gpu):
    lib = E.b2g_lib()
    w = gpu.world("pyramid", 6, 0)
    w.step(10)
    a = w.device_context()
    snap = _snapshot(lib, a)
    size = ctypes.c_size_t(0)
    small = np.zeros(16, np.uint8)
    assert lib.b2g_snapshot(a, E.ptr(small), ctypes.c_size_t(16), ctypes.byref(size)) != 0
    assert size.value == snap.size
    b = _fresh(lib)
    try:
        assert lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size - 4)) != 0   # truncated
        bad = snap.copy()
        bad[0] ^= 0xFF
        assert lib.b2g_restore(b, E.ptr(bad), ctypes.c_size_t(bad.size)) != 0      # not a snapshot
        E.b2g_check(lib, b, lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size)))
        assert lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size)) != 0    # not a new context
        assert b"new context" in lib.b2g_last_error(b)
    finally:
        lib.b2g_destroy(b)
        w.close()


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


def _rows(state):
    """Oracle body rows (px py s c cx cy a vx vy w awake) from b2g_body_state
    words: xf at 2-5, sweep c at 11-12, a at 13, v at 14-15, w at 16, the
    awake flag in bit 1 of the flags word."""
    awake = ((state[:, 1] & 2) != 0).astype(np.float32).view(np.uint32)
    return np.concatenate([state[:, 2:6], state[:, 11:14], state[:, 14:17], awake[:, None]], axis=1)


@pytest.mark.parametrize("kind,n,seed,sleep,ccd,k,m", [
    ("pyramid", 20, 0, True, False, 150, 250),   # sleeps after the resume (near step 280)
    ("pile", 1200, 9, False, False, 100, 100),
    ("bullets", 24, 1, True, True, 40, 80),      # time-of-impact events after the resume
])
def test_resumed_reference_mode_matches_oracle(gpu, orc, wide_trig, kind, n, seed, sleep, ccd, k, m):
    """Reference-order mode, the counterpart of b2World::Dump
    (Box2D/Dynamics/b2World.cpp:1271-1328): a snapshot taken after k steps
    and restored into a new context continues exactly as the oracle that
    stepped k + m times without a break (bodies after every step, the
    contact table and manifolds at the end)."""
    lib = E.b2g_lib()
    g = gpu.world(kind, n, seed, sleep=sleep, ccd=ccd)
    g.set_mode(1)
    o = orc.world(kind, n, seed, sleep=sleep, ccd=ccd)
    g.step(k)
    o.step(k)
    nb = g.e.fn("world_body_count")(g.h)
    snap = _snapshot(lib, g.device_context())
    g.close()
    b = _fresh(lib)
    p = _params()
    p[0]["mode"] = 1
    events = 0
    try:
        E.b2g_check(lib, b, lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size)))
        st = np.zeros(1, E.STATS)
        for s in range(m):
            E.b2g_check(lib, b, lib.b2g_step(b, E.ptr(p)))
            o.step(1)
            E.b2g_check(lib, b, lib.b2g_get_stats(b, E.ptr(st)))
            events += int(st[0]["toiEvents"])
            got = _rows(_bodies(lib, b, nb))
            want = (o.bodies()[:, :11] + np.float32(0)).view(np.uint32)
            bad = np.where(~np.all(got == want, axis=1))[0]
            assert len(bad) == 0, f"step {k + s + 1}: {len(bad)} bodies differ, first {bad[:5]}"
        meta, man = _state(lib, b)
        gm, gman = E.contact_table(meta.reshape(-1, 4), man.view(E.MANIFOLD).reshape(-1))
        om, oman = E.contact_table(*o.contacts())
        assert np.array_equal(gm[:, :3], om[:, :3])
        assert np.array_equal(E.manifold_bits(gman), E.manifold_bits(oman))
        if ccd:
            assert events > 5, events
        if sleep and kind == "pyramid":
            assert not got[:, 10].any()  # the resumed pyramid fell asleep as the oracle's did
    finally:
        lib.b2g_destroy(b)
        o.close()


def test_resumed_after_destruction_matches_oracle(gpu, orc, wide_trig):
    """A snapshot taken after bodies were destroyed and others added (the
    first two phases of the destroy plan) holds the proxies' tree ids: the
    restored context creates its new contacts in the reference's order and
    stays equal to the oracle, bit for bit, over 150 more steps."""
    from test_oracle import DESTROY_PLAN, destroy_plan_run
    lib = E.b2g_lib()
    g = gpu.world("world", 150, 5, sleep=True)
    g.set_mode(1)
    o = orc.world("world", 150, 5, sleep=True)
    alive = destroy_plan_run(g, o, lambda tag, alive: None, plan=DESTROY_PLAN[:2])
    g.step(1)  # uploads the added bodies; their pairs are found with the reused ids
    o.step(1)
    nb = g.e.fn("world_body_count")(g.h)
    snap = _snapshot(lib, g.device_context())
    g.close()
    b = _fresh(lib)
    p = _params()
    p[0]["mode"] = 1
    try:
        E.b2g_check(lib, b, lib.b2g_restore(b, E.ptr(snap), ctypes.c_size_t(snap.size)))
        for s in range(150):
            E.b2g_check(lib, b, lib.b2g_step(b, E.ptr(p)))
            o.step(1)
            got = _rows(_bodies(lib, b, nb))[alive]
            want = (o.bodies()[alive, :11] + np.float32(0)).view(np.uint32)
            bad = np.where(~np.all(got == want, axis=1))[0]
            assert len(bad) == 0, f"step {s + 1} after the resume: {len(bad)} bodies differ, first {bad[:5]}"
        meta, man = _state(lib, b)
        gm = meta.reshape(-1, 4)
        om, _ = o.contacts()
        assert np.array_equal(gm[::-1, :3], om[:, :3])  # creation order against the newest-first list
    finally:
        lib.b2g_destroy(b)
        o.close()
