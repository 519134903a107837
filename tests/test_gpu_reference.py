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


@pytest.mark.parametrize("ccd", [False, True])
def test_destroy_body_reference_order_bitexact(gpu, orc, wide_trig, ccd):
    """Bodies destroyed mid-run and others added after them (the plan of
    test_oracle.destroy_plan_run, which pins the oracle against the reference
    build): the host models the reference tree's proxy ids and the device
    orders new pairs by them, so contacts are created in the reference's
    order with its A/B roles. Reference-order mode stays equal to the oracle
    bit for bit, the contact list in order, with and without continuous
    collision (whose events find new pairs too)."""
    from test_oracle import destroy_plan_run, proxy_ids
    g = gpu.world("world", 150, 5, sleep=True, ccd=ccd)
    g.set_mode(REFERENCE)
    o = orc.world("world", 150, 5, sleep=True, ccd=ccd)
    print(f"case: destroy plan, reference order, ccd={ccd}", flush=True)

    def compare(tag, alive):
        assert np.array_equal(proxy_ids(g), proxy_ids(o)), f"{tag}: proxy ids differ"
        if "step" not in tag:
            return  # the device drops a destroyed body's contacts at the next step
        gb, ob = g.bodies()[alive], o.bodies()[alive]
        bad = np.where(~np.all(_canon(gb[:, :11]) == _canon(ob[:, :11]), axis=1))[0]
        assert len(bad) == 0, f"{tag}: {len(bad)} bodies differ, first {bad[:5]}"
        gm, gman = g.contacts()
        om, oman = o.contacts()
        # the device table is in creation order, the reference's list newest first
        assert np.array_equal(gm[::-1, :3], om[:, :3]), f"{tag}: contact lists differ"
        assert np.array_equal(E.manifold_bits(gman[::-1]), E.manifold_bits(oman)), f"{tag}: manifolds differ"

    destroy_plan_run(g, o, compare)
    ids = proxy_ids(g)
    assert np.any(np.diff(ids[ids >= 0]) < 0)  # freed ids were taken again
    g.close(); o.close()


def test_destroy_body_coloured(gpu, orc):
    """The coloured solver through the same plan: the proxy ids equal the
    oracle's, every contact's fixture A is the one with the smaller proxy id
    unless the shape pair's registry swaps them, no contact survives its
    body's next step, and the penetration at the end stays within the
    oracle's on the same plan plus 0.5 cm (mean) and 5 cm (max)."""
    from test_oracle import destroy_plan_run, proxy_ids
    g = gpu.world("world", 150, 5, sleep=True)
    o = orc.world("world", 150, 5, sleep=True)
    def compare(tag, alive):
        ids = proxy_ids(g)
        assert np.array_equal(ids, proxy_ids(o)), f"{tag}: proxy ids differ"
        if "step" not in tag:
            return
        meta, _ = g.contacts()
        body, _ = g.fixture_info()
        assert np.all(alive[body[meta[:, 0]]]) and np.all(alive[body[meta[:, 1]]]), f"{tag}: a destroyed body's contact"
        fa, fb = meta[:, 0], meta[:, 1]
        assert np.all(ids[fa] >= 0) and np.all(ids[fb] >= 0)
        # the same fixtures in the oracle: their shape types
        nf = len(ids)
        sh = np.zeros(nf, E.SHAPE)
        o.e.fn("world_get_shapes")(o.h, E.ptr(sh))
        ta, tb = sh["type"][fa], sh["type"][fb]
        # A has the larger id only where the registry swapped a pair whose
        # smaller-id fixture comes second in b2Contact's table: (circle,
        # polygon), (circle, edge), (polygon, edge) become (polygon, circle),
        # (edge, circle), (edge, polygon) (types 0 circle, 1 edge, 2 polygon)
        swapped = ids[fa] > ids[fb]
        pair = ta * 3 + tb
        assert np.all(np.isin(pair[swapped], [2 * 3 + 0, 1 * 3 + 0, 1 * 3 + 2])), tag
        assert np.all(np.isfinite(g.bodies()[alive]))

    destroy_plan_run(g, o, compare)
    # matched penetration: the oracle ran the same plan (2.1 cm mean, 19.5 cm max)
    mean, mx = g.penetration()[:2]
    omean, omx = o.penetration()[:2]
    assert mean <= omean + 0.005 and mx <= omx + 0.05, (mean, mx, omean, omx)
    g.close(); o.close()


def test_set_type_reference_order_bitexact(gpu, orc, wide_trig):
    """b2Body::SetType and SetActive mid-run (the plan of
    test_oracle.type_plan_run, which pins the oracle against the reference
    build): the body's contacts end at the next step, its neighbours wake at
    once, its proxies are queried again (or destroyed, or created with new
    tree ids) by the end of the next step; reference-order mode stays equal
    to the oracle bit for bit, the contact list in order."""
    from test_oracle import type_plan_run
    g = gpu.world("world", 150, 5, sleep=True)
    g.set_mode(REFERENCE)
    o = orc.world("world", 150, 5, sleep=True)
    print("case: type plan, reference order", flush=True)

    def compare(tag):
        if "step" not in tag:
            return  # the neighbours' wake-ups reach the host copy with the next step
        gb, ob = g.bodies(), o.bodies()
        bad = np.where(~np.all(_canon(gb[:, :11]) == _canon(ob[:, :11]), axis=1))[0]
        assert len(bad) == 0, f"{tag}: {len(bad)} bodies differ, first {bad[:5]}"
        gm, _ = g.contacts()
        om, _ = o.contacts()
        assert np.array_equal(gm[::-1, :3], om[:, :3]), f"{tag}: contact lists differ"

    type_plan_run(g, o, compare)
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
