"""Revolute joints on the GPU step (coloured solver) against the oracle.

* Pendulums (one joint per island, no contacts): the coloured solver's
  order inside such an island is the reference's, so trajectories and joint
  impulses equal the oracle's bit for bit (trig as the device evaluates it),
  also across mid-run edits through the box2d:: setters (limits switched and
  moved, motors retuned, a joint destroyed).
* Chains (joints and contacts in one island, colliding chains): the
  coloured order differs from the reference's, so the checks are the
  constraint's own: every hinge holds within the slop the oracle itself
  leaves, no contact joins two bodies that a joint keeps apart, and the
  bodies stay finite.
* Reference-order mode (island search through joint edges, joints solved
  before the contacts in every velocity iteration and after them in every
  position iteration, as b2Island::Solve does): every joint type, chains
  with contacts in one island and the stress world equal the oracle bit for
  bit, and chains stay within 1e-4 of the reference build over 60 steps.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu
F = ctypes.c_float


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


def _joints(w):
    nj = w.e.fn("world_joint_count")(w.h)
    out = np.zeros(4 * nj, np.float32)
    w.e.fn("world_get_joints")(w.h, E.ptr(out))
    return out.reshape(-1, 4)


def _bits(a):
    return (a + np.float32(0)).view(np.uint32)  # -0 == +0


def test_pendulums_bitexact(gpu, orc, wide_trig):
    g = gpu.world("pendulums", 24, 0, sleep=True)
    o = orc.world("pendulums", 24, 0, sleep=True)
    for s in range(400):
        g.step(1)
        o.step(1)
        if s % 50 == 49 or s < 3:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_bits(gb[:, :10]) == _bits(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: bodies {bad[:8]} differ"
    assert np.array_equal(_bits(_joints(g)), _bits(_joints(o)))
    st = g.stats()
    assert st["contacts"] == 0
    g.close()
    o.close()


def _edit(w):
    w.e.fn("world_set_joint")(w.h, 0, 1, F(-0.2), F(0.25), 0, F(0.0), F(0.0))
    w.e.fn("world_set_joint")(w.h, 1, 1, F(-0.1), F(0.1), 0, F(0.0), F(0.0))
    w.e.fn("world_set_joint")(w.h, 2, 0, F(0.0), F(0.0), 1, F(-2.0), F(5.0))
    w.e.fn("world_set_joint")(w.h, 4, 0, F(-0.4), F(0.3), 0, F(0.0), F(0.0))
    w.e.fn("world_set_joint")(w.h, 6, 1, F(0.0), F(0.0), 1, F(0.5), F(10.0))
    assert w.e.fn("world_destroy_joint")(w.h, 3) == 0


def test_joint_edits_bitexact(gpu, orc, wide_trig):
    g = gpu.world("pendulums", 24, 0, sleep=True)
    o = orc.world("pendulums", 24, 0, sleep=True)
    g.step(90)
    o.step(90)
    _edit(g)
    _edit(o)
    for s in range(240):
        g.step(1)
        o.step(1)
        if s % 40 == 39 or s < 3:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_bits(gb[:, :10]) == _bits(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 91}: bodies {bad[:8]} differ"
    jg, jo = _joints(g), _joints(o)
    assert np.array_equal(_bits(jg), _bits(jo))
    assert np.all(jg[3] == 0.0)
    g.close()
    o.close()


def _chain_gaps(b, n, links):
    """Hinge gaps |anchor on A - anchor on B| of the chains scene: per chain a
    static anchor at the chain's first hinge, then `links` boxes 0.5 m long
    joined end to end."""
    px, py, qs, qc = b[:, E.PX], b[:, E.PY], b[:, E.QS], b[:, E.QC]

    def world_point(i, lx):
        return np.array([px[i] + qc[i] * lx, py[i] + qs[i] * lx])

    gaps = []
    for c in range(n):
        anchor = 1 + c * (links + 1)
        for i in range(links):
            cur = anchor + 1 + i
            pa = world_point(anchor, 0.0) if i == 0 else world_point(cur - 1, 0.25)
            pb = world_point(cur, -0.25)
            gaps.append(np.hypot(*(pa - pb)))
    return np.array(gaps)


@pytest.mark.parametrize("n,links,steps,sleep,vel,pos", [
    (3, 12, 400, False, 8, 3),
    (6, 16, 300, True, 4, 2),
])
def test_chains_hold(gpu, orc, n, links, steps, sleep, vel, pos):
    g = gpu.world("chains", n, links, sleep=sleep)
    o = orc.world("chains", n, links, sleep=sleep)
    worst_g = worst_o = 0.0
    # fixture -> body: fixture 0 is the ground edge, then one box per link
    fb = np.concatenate([[0], [1 + c * (links + 1) + 1 + i for c in range(n) for i in range(links)]])
    joined = set()
    for c in range(n):
        anchor = 1 + c * (links + 1)
        for i in range(links):
            a, b = (anchor if i == 0 else anchor + i), anchor + 1 + i
            joined.add((min(a, b), max(a, b)))
    touched = 0
    for s in range(steps):
        g.step(1, vel=vel, pos=pos)
        o.step(1, vel=vel, pos=pos)
        if s % 20 == 19:
            gb = g.bodies()
            assert np.all(np.isfinite(gb)), f"step {s + 1}"
            worst_g = max(worst_g, _chain_gaps(gb, n, links).max())
            worst_o = max(worst_o, _chain_gaps(o.bodies(), n, links).max())
            meta, _ = g.contacts()
            touched += len(meta)
            for fa, fbb in meta[:, :2]:
                ba, bb = fb[fa], fb[fbb]
                assert (min(ba, bb), max(ba, bb)) not in joined, f"contact between jointed bodies {ba} {bb}"
    assert touched > 0  # the chains do hit each other or the ground
    assert worst_g <= max(0.02, 3.0 * worst_o), (worst_g, worst_o)
    j = _joints(g)
    assert np.all(np.isfinite(j))
    g.close()
    o.close()


@pytest.mark.parametrize("kind,n,seed,steps,sleep", [
    ("pendulums", 24, 0, 300, True),
    ("chains", 3, 12, 300, False),
    ("chains", 6, 16, 200, True),
    ("distance", 16, 1, 200, True),
    ("ropes", 16, 1, 200, True),
    ("welds", 12, 2, 200, True),
    ("friction", 16, 0, 200, True),
    ("motor", 12, 1, 200, True),
    ("pulley", 8, 0, 200, False),
    ("mouse", 8, 0, 200, True),
    ("prismatic", 12, 0, 200, True),
    ("wheels", 6, 1, 200, True),
    ("gears", 4, 0, 200, False),
    ("stress", 1500, 11, 150, False),
])
def test_reference_mode_joints_bitexact(gpu, orc, wide_trig, kind, n, seed, steps, sleep):
    g = gpu.world(kind, n, seed, sleep=sleep)
    g.set_mode(1)
    o = orc.world(kind, n, seed, sleep=sleep)
    print(f"case: reference-order mode, {kind} n={n} seed={seed}", flush=True)
    for s in range(steps):
        g.step(1)
        o.step(1)
        if s % 25 == 24 or s == steps - 1:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_bits(gb[:, :10]) == _bits(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: bodies {bad[:8]} differ"
    assert np.array_equal(_bits(_joints(g)), _bits(_joints(o)))
    if kind == "stress":
        assert g.stats()["constraints"] > 1000
    g.close()
    o.close()


def test_reference_mode_chains_match_reference_build(gpu, ref):
    """Against upstream Box2D itself (glibc trig): chains colliding with the
    ground, poses within 1e-4 over 60 steps."""
    g = gpu.world("chains", 3, 12, sleep=False)
    g.set_mode(1)
    r = ref.world("chains", 3, 12, sleep=False)
    worst = 0.0
    for _ in range(60):
        g.step(1)
        r.step(1)
        d = np.abs(g.bodies()[:, [E.CX, E.CY, E.ANG]] - r.bodies()[:, [E.CX, E.CY, E.ANG]]).max()
        worst = max(worst, float(d))
    assert worst <= 1e-4, worst
    g.close()
    r.close()


def _b2g(gpu):
    lib = ctypes.CDLL(E.GPU_B2G_LIB)
    return lib


def test_energy_and_drift_closed_forms(gpu):
    """b2g_energy at load against the pendulums' closed form (boxes 0.5 x 0.5
    m, density 1: m = 0.25, I = m (0.5^2 + 0.5^2) / 12; v0 = 0.5 + 0.25 (i mod
    13); height 20 - length), and b2g_joint_drift against the hinge gaps
    computed from the body poses."""
    lib = _b2g(gpu)
    g = gpu.world("pendulums", 24, 0, sleep=False)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    assert lib.b2g_energy(ctx, E.ptr(out)) == 0
    m, inertia = 0.25, 0.25 * 0.5 / 12.0
    i = np.arange(24)
    ke = np.sum(0.5 * m * np.float32(0.5 + 0.25 * (i % 13)).astype(np.float64) ** 2)
    length = np.float32(1.0 + 0.05 * (i % 20))
    pe = np.sum(m * 10.0 * np.float32(20.0 - length).astype(np.float64))
    assert abs(out[0] - ke) <= 1e-6 * ke and abs(out[1] - pe) <= 1e-6 * pe, (out, ke, pe)
    assert inertia > 0
    g.step(120)
    assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
    b = g.bodies()[1:].astype(np.float64)
    x = 5.0 * i
    ax = b[:, E.PX] - b[:, E.QS] * length
    ay = b[:, E.PY] + b[:, E.QC] * length
    gaps = np.hypot(ax - x, ay - 20.0)
    assert out[2] == 24
    assert abs(out[0] - gaps.mean()) < 2e-5 and abs(out[1] - gaps.max()) < 2e-5, (out, gaps.mean(), gaps.max())
    g.close()


def _stress_gaps(b, base, n, links):
    """Hinge gaps of the stress scene's chains, whose links follow the
    container and the n pile bodies from body `base` on, end to end."""
    nc = n // 50
    idx = base + 1 + n + np.arange(nc * links).reshape(nc, links)
    a, c = idx[:, :-1].ravel(), idx[:, 1:].ravel()
    b = b.astype(np.float64)
    pa = np.stack([b[a, E.PX] + b[a, E.QC] * 0.25, b[a, E.PY] + b[a, E.QS] * 0.25], 1)
    pb = np.stack([b[c, E.PX] - b[c, E.QC] * 0.25, b[c, E.PY] - b[c, E.QS] * 0.25], 1)
    return np.hypot(*(pa - pb).T)


def test_stress_worlds_small(gpu, orc):
    """Four stress worlds (a pile of 1500 bodies dropped on a mat of 30
    chains of 32 links each) in sub-worlds of one context: no contact joins
    two worlds or two jointed links, b2g_joint_drift agrees with the gaps
    computed from the poses, and world 0's hinge gaps stay within a small
    factor of the oracle's on the same world (fast bodies tear links in
    both; the solver order differs)."""
    lib = _b2g(gpu)
    n, worlds, links, seed = 1500, 4, 32, 11
    g = E.World.worlds(gpu, "stress", worlds, n, seed, sleep=False)
    o = orc.world("stress", n, seed, sleep=False)
    per_world = 1 + n + (n // 50) * links
    nb = g.e.fn("world_body_count")(g.h)
    assert nb == worlds * per_world
    assert g.e.fn("world_joint_count")(g.h) == worlds * (n // 50) * (links - 1)
    # fixture -> body: the container's 3 edges on body 0 of each world, then one per body
    fb = []
    for w in range(worlds):
        base = w * per_world
        fb += [base] * 3 + list(range(base + 1, base + per_world))
    fb = np.array(fb)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    for s in range(3):
        g.step(100)
        o.step(100)
        meta, _ = g.contacts()
        ba, bb = fb[meta[:, 0]], fb[meta[:, 1]]
        assert np.all(ba // per_world == bb // per_world), "a contact joins two worlds"
        local_a, local_b = ba % per_world, bb % per_world
        lo, hi = np.minimum(local_a, local_b), np.maximum(local_a, local_b)
        chain_lo = lo - (1 + n)
        jointed = (chain_lo >= 0) & (hi == lo + 1) & (chain_lo % links != links - 1)
        assert not np.any(jointed), "contact between jointed links"
    gb = g.bodies()
    assert np.all(np.isfinite(gb))
    gaps = np.concatenate([_stress_gaps(gb, w * per_world, n, links) for w in range(worlds)])
    assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
    assert out[2] == len(gaps)
    assert abs(out[0] - gaps.mean()) < 1e-5 and abs(out[1] - gaps.max()) < 1e-5, (out, gaps.mean(), gaps.max())
    g0 = _stress_gaps(gb, 0, n, links)
    o0 = _stress_gaps(o.bodies(), 0, n, links)
    assert g0.mean() <= 2.0 * o0.mean() + 0.002, (g0.mean(), o0.mean())
    assert np.percentile(g0, 99) <= 3.0 * np.percentile(o0, 99) + 0.01, (np.percentile(g0, 99), np.percentile(o0, 99))
    pen = g.penetration()
    assert pen[2] > 0 and pen[0] < 0.1
    g.close()
    o.close()


def test_distance_joints_bitexact(gpu, orc, wide_trig):
    """Distance joints, one per island (rods and springs from the ground):
    the coloured solver's order is the reference's there, so bodies and the
    joints' reaction forces equal the oracle's bit for bit, also after
    SetLength / SetFrequency / SetDampingRatio through the box2d:: API."""
    g = gpu.world("distance", 16, 1, sleep=True)
    o = orc.world("distance", 16, 1, sleep=True)
    for s in range(300):
        if s == 150:
            for w in (g, o):
                assert w.e.fn("world_set_distance")(w.h, 0, F(5.0), F(1.0), F(0.5)) == 0
                assert w.e.fn("world_set_distance")(w.h, 1, F(6.0), F(0.0), F(0.0)) == 0
                assert w.e.fn("world_set_distance")(w.h, 2, F(3.0), F(0.0), F(0.0)) == 0
        g.step(1)
        o.step(1)
        if s % 30 == 29 or s < 3:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_bits(gb[:, :10]) == _bits(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: bodies {bad[:8]} differ"
    assert np.array_equal(_bits(_joints(g)), _bits(_joints(o)))
    assert np.abs(_joints(g)[:, :2]).max() > 0.05  # the joints carry load
    g.close()
    o.close()


def test_distance_rods_hold(gpu):
    """Boxes on rods hanging from rods (two joints per island, solved by
    colour): every rod keeps its length within 2 cm, as b2g_joint_drift
    measures it (springs are left out of that measure)."""
    lib = _b2g(gpu)
    g = gpu.world("distance", 16, 0, sleep=False)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    worst = 0.0
    for _ in range(10):
        g.step(30)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 8 + 16  # rods from the ground (even i) and the hanging rods
        worst = max(worst, out[1])
    assert np.all(np.isfinite(g.bodies()))
    assert worst < 0.02, worst
    g.close()


@pytest.mark.parametrize("kind,n,seed,edit", [
    ("ropes", 16, 1, ("world_set_rope", 0, F(2.0))),
    ("welds", 12, 2, ("world_set_weld", 1, F(0.0), F(0.0))),
    ("friction", 16, 0, ("world_set_friction", 2, F(0.0), F(0.0))),
    ("motor", 12, 1, ("world_set_motor", 0, F(-8.0), F(6.0), F(2.0), F(30.0), F(8.0), F(0.4))),
    ("pulley", 8, 0, ("world_joint_count",)),
    ("mouse", 8, 0, ("world_set_mouse", 0, F(-9.0), F(5.0), F(1000.0), F(8.0), F(0.3))),
    ("prismatic", 12, 0, ("world_set_prismatic", 2, 1, F(-0.5), F(0.5), 1, F(-1.0), F(50.0))),
    ("wheels", 6, 1, ("world_set_wheel", 0, 1, F(5.0), F(40.0), F(2.0), F(0.2))),
])
def test_rope_weld_friction_motor_joints_bitexact(gpu, orc, wide_trig, kind, n, seed, edit):
    """Rope, weld, friction and motor joints, one per island (or joints on
    disjoint bodies where boxes collide): boxes on ropes from the ground
    (slack, taut, one of length zero), single-link beams on rigid and soft
    welds (one with a fixed-rotation link), boxes thrown against friction
    joints and boxes driven to offsets by motor joints: bodies and the
    joints' reaction forces and torques equal the oracle's bit for bit, also
    after the setters through the box2d:: API."""
    g = gpu.world(kind, n, seed, sleep=True)
    o = orc.world(kind, n, seed, sleep=True)
    for s in range(300):
        if s == 150:
            for w in (g, o):
                if edit[0] != "world_joint_count":  # the pulley has no setters
                    assert w.e.fn(edit[0])(w.h, *edit[1:]) == 0
            # reactions read right after a setter show its side effects (a
            # prismatic limit change drops the limit impulse at once)
            assert np.array_equal(_bits(_joints(g)), _bits(_joints(o)))
        g.step(1)
        o.step(1)
        if s % 30 == 29 or s < 3:
            gb, ob = g.bodies(), o.bodies()
            bad = np.where(~np.all(_bits(gb[:, :10]) == _bits(ob[:, :10]), axis=1))[0]
            assert len(bad) == 0, f"step {s + 1}: bodies {bad[:8]} differ"
    jg, jo = _joints(g), _joints(o)
    assert np.array_equal(_bits(jg), _bits(jo))
    assert np.abs(jg[:, :2]).max() > 0.03  # the joints carry load
    g.close()
    o.close()


def _weld_gaps(bodies, n):
    """Anchor gap |cB + rB - cA - rA| of every weld of the welds scene (n
    beams of 4 + i % 4 links from x = 10 i - 5 n, a dropped box after each)."""
    def world_point(i, local):
        p = bodies[i, [E.PX, E.PY]].astype(np.float64)
        s, c = float(bodies[i, E.QS]), float(bodies[i, E.QC])
        return p + np.array([c * local[0] - s * local[1], s * local[0] + c * local[1]])
    gaps, body = [], 1
    for i in range(n):
        x, prev = 10.0 * i - 5.0 * n, 0
        for _ in range(4 + i % 4):
            anchor = (x, 10.0) if prev == 0 else (0.5, 0.0)
            gaps.append(np.hypot(*(world_point(body, (-0.5, 0.0)) - world_point(prev, anchor))))
            prev, body = body, body + 1
        body += 1
    return np.array(gaps)


def test_ropes_hold(gpu):
    """Ropes with a second box on a rope (two joints per island, solved by
    colour) hold: every rope stays within 2 cm of its maximum length, as
    b2g_joint_drift measures the stretch."""
    lib = _b2g(gpu)
    g = gpu.world("ropes", 16, 0, sleep=False)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    worst = 0.0
    for _ in range(10):
        g.step(30)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 32
        worst = max(worst, out[1])
    assert np.all(np.isfinite(g.bodies()))
    assert worst < 0.02, worst
    g.close()


def test_welds_hold(gpu, orc):
    """Welded beams of 4-7 links under dropped boxes (solved by colour):
    b2g_joint_drift's anchor gap equals the gap computed from the bodies, and
    the worst gap over the run stays within 1.5x (+1 cm) of the oracle's on
    the same scene, which sags too (soft welds, long beams)."""
    lib = _b2g(gpu)
    g = gpu.world("welds", 12, 0, sleep=False)
    o = orc.world("welds", 12, 0, sleep=False)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    worst_g = worst_o = 0.0
    for _ in range(10):
        g.step(30)
        o.step(30)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 66
        gaps = _weld_gaps(g.bodies(), 12)
        assert abs(out[1] - gaps.max()) < 1e-4 and abs(out[0] - gaps.mean()) < 1e-4, (out, gaps.max())
        worst_g = max(worst_g, out[1])
        worst_o = max(worst_o, _weld_gaps(o.bodies(), 12).max())
    assert np.all(np.isfinite(g.bodies()))
    assert worst_g <= 1.5 * worst_o + 0.01, (worst_g, worst_o)
    g.close()
    o.close()


def test_pulleys_hold_after_the_joint_table_grows(gpu):
    """Pulleys keep lengthA + ratio lengthB (their constant, kept in the
    joint rows) when a second scene's joints grow the joint table past its
    first capacity: b2g_joint_drift's pulley error stays under 1 cm."""
    lib = _b2g(gpu)
    g = gpu.world("pulley", 8, 0, sleep=False)
    g.step(30)
    assert g.e.fn("world_load_scene")(g.h, b"friction", 100, ctypes.c_uint(0)) == 0
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    for _ in range(4):
        g.step(15)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 8 and out[1] < 0.01, out
    g.close()


def test_wheels_hold(gpu):
    """Carts on two wheel joints each (two joints per island, solved by
    colour) over bumps: every wheel stays within 1 cm of its axis line, as
    b2g_joint_drift measures it, and the driven carts move forwards."""
    lib = _b2g(gpu)
    g = gpu.world("wheels", 6, 0, sleep=False)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    x0 = g.bodies()[:, E.PX].copy()
    for _ in range(8):
        g.step(30)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 12 and out[1] < 0.01, out
    b = g.bodies()
    chassis = [1 + 3 * i for i in range(6)]
    assert all(b[c, E.PX] - x0[c] > 2.0 for c in chassis[::2]), [b[c, E.PX] - x0[c] for c in chassis]
    g.close()


def test_gear_trains_coloured(gpu):
    """Gear trains (a gear joint writes four bodies; it is coloured with all
    of them): the GPU keeps the gear relations and the rack's closed-form
    fall (2%), and b2g_joint_drift stays under 1 cm / 0.01 rad over every
    revolute, prismatic and gear joint of the four trains."""
    from test_oracle import gear_train_check
    lib = _b2g(gpu)
    g = gpu.world("gears", 4, 0, sleep=False)
    gear_train_check(g, 60, 0.02)
    ctx = g.device_context()
    out = np.zeros(3, np.float64)
    for _ in range(4):
        g.step(30)
        assert lib.b2g_joint_drift(ctx, E.ptr(out)) == 0
        assert out[2] == 20 and out[1] < 0.01, out
    g.close()
