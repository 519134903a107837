"""Reference-order mode (B2G_MODE_REFERENCE): islands found in Box2D's order
(newest-first bodies and contact edges, depth-first), one wavefront per
island, constraints solved sequentially in island order with the per-island
early exit and sleep of b2Island::Solve.

With full solver iterations, its state equals the CPU oracle's bit for bit
when the oracle evaluates in-step trig as the GPU does (rounded double
sin/cos, trig mode 1), and stays within 1e-4 of the reference build's poses
(glibc sinf/cosf) over 60 steps.
"""
from __future__ import annotations

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu

REFERENCE = 1


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


def _canon(b):
    return (b + np.float32(0.0)).view(np.uint32)


@pytest.mark.parametrize("kind,n,seed,steps,sleep,vel,pos", [
    ("hello", 0, 0, 60, True, 6, 2),
    ("pyramid", 20, 0, 400, True, 8, 3),     # the whole pyramid falls asleep near step 280
    ("pyramid", 12, 0, 200, False, 8, 3),
    ("pile", 400, 3, 300, True, 8, 3),
    ("pile", 1200, 9, 200, False, 8, 3),
    ("terrain", 60, 1, 300, True, 8, 3),     # chain shapes (children as edges without skin in their boxes)
    ("sensors", 30, 1, 300, True, 8, 3),     # sensor fixtures: a slab and halos, overlap-only contacts
    ("sensors", 40, 2, 200, False, 8, 3),
])
def test_reference_mode_bitexact(gpu, orc, wide_trig, kind, n, seed, steps, sleep, vel, pos):
    g = gpu.world(kind, n, seed, sleep=sleep)
    g.set_mode(REFERENCE)
    o = orc.world(kind, n, seed, sleep=sleep)
    for s in range(steps):
        g.step(1, vel=vel, pos=pos)
        o.step(1, vel=vel, pos=pos)
        if s % 20 == 19 or s == steps - 1:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_canon(gb[:, :11]) == _canon(ob[:, :11]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: {len(bad)} bodies differ, first {bad[:5]}"
    gm, gman = E.contact_table(*g.contacts())
    om, oman = E.contact_table(*o.contacts())
    assert np.array_equal(gm[:, :3], om[:, :3])
    assert np.array_equal(E.manifold_bits(gman), E.manifold_bits(oman))
    g.close(); o.close()


def test_reference_mode_matches_reference_build(gpu, ref):
    """Against upstream Box2D itself (glibc trig): poses within 1e-4 over 60 steps."""
    g = gpu.world("pyramid", 20, 0, sleep=True)
    g.set_mode(REFERENCE)
    r = ref.world("pyramid", 20, 0, sleep=True)
    worst = 0.0
    for s in range(60):
        g.step(1)
        r.step(1)
        d = np.abs(g.bodies()[:, [E.CX, E.CY, E.ANG]] - r.bodies()[:, [E.CX, E.CY, E.ANG]]).max()
        worst = max(worst, float(d))
    assert worst <= 1e-4, worst


def test_coloured_mode_pyramid_sleeps(gpu, orc):
    """Coloured mode with sleeping on: the pyramid comes to rest and falls
    asleep as one island, like the reference's."""
    g = gpu.world("pyramid", 20, 0, sleep=True)
    o = orc.world("pyramid", 20, 0, sleep=True)
    g.step(600)
    o.step(600)
    gb, ob = g.bodies(), o.bodies()
    assert np.all(ob[1:, E.AWAKE] == 0)
    assert np.all(gb[1:, E.AWAKE] == 0)
    d = np.hypot(gb[1:, E.CX] - ob[1:, E.CX], gb[1:, E.CY] - ob[1:, E.CY])
    assert d.max() < 0.1
    g.close(); o.close()
