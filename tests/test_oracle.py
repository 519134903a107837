"""Pins the CPU oracle (oracle/cpu_ref.cpp) against the upstream Box2D 2.3.2
build of the reference's own sources (oracle/_ref), against the reference's
known-answer collision cases (Regression/tests/helloworld.cpp:268-418), and
against closed forms. Everything here is bit-exact unless stated."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import corpus
import engines as E

F = ctypes.c_float


# ------------------------------------------------------------ whole step
@pytest.mark.parametrize("kind,n,seed,steps,sleep,vel,pos", [
    ("hello", 0, 0, 60, True, 6, 2),           # helloworld.cpp:420-432 (6/2 iterations)
    ("pyramid", 20, 0, 600, True, 8, 3),       # sleeps at step ~284
    ("pyramid", 20, 0, 300, False, 8, 3),
    ("pyramid", 6, 0, 200, True, 4, 1),
    ("pile", 300, 1, 400, False, 8, 3),
    ("pile", 1500, 7, 150, True, 8, 3),
    ("pendulums", 24, 0, 400, True, 8, 3),     # revolute joints: free, limited, motorised
    ("chains", 3, 12, 400, False, 8, 3),       # 3 chains of 12 links hitting the ground
    ("chains", 2, 32, 250, True, 4, 2),
    ("stress", 1000, 3, 200, False, 8, 3),     # pile dropped on a mat of 20 chains of 32 links
    ("plank", 100, 0, 200, True, 8, 3),        # one body in 101 contacts
    ("terrain", 60, 1, 400, True, 8, 3),       # chain shapes: open chains with and without ghosts, a loop
    ("distance", 16, 0, 400, True, 8, 3),      # distance joints: rods and springs, a second box on a rod
    ("distance", 12, 1, 300, False, 8, 3),
    ("terrain", 100, 2, 300, False, 8, 3),
    ("ropes", 16, 0, 400, True, 8, 3),         # rope joints: slack and taut, zero length, a second box on a rope
    ("ropes", 12, 1, 300, False, 8, 3),
    ("welds", 12, 0, 400, True, 8, 3),         # weld joints: rigid and soft beams, a fixed-rotation link
    ("welds", 10, 0, 250, False, 4, 2),
    ("welds", 12, 2, 300, True, 8, 3),         # single-link beams: soft welds swing, a fixed-rotation rigid weld
    ("friction", 16, 0, 300, True, 8, 3),      # friction joints: boxes thrown along the ground, some colliding
    ("motor", 12, 0, 400, True, 8, 3),         # motor joints: boxes driven to offsets from the ground and each other
    ("motor", 9, 1, 300, False, 4, 2),
    ("pulley", 8, 0, 300, True, 8, 3),         # pulley joints: ratios 1-2, anchors at centres and corners
    ("mouse", 8, 0, 300, True, 8, 3),          # mouse joints: boxes pulled up by soft springs
    ("prismatic", 12, 0, 300, True, 8, 3),     # prismatic joints: free, limited (one equal limit), motorised
    ("wheels", 6, 0, 400, True, 8, 3),         # wheel joints: carts on suspension springs over bumps, driven
    ("wheels", 6, 1, 300, False, 4, 2),
    ("gears", 4, 0, 300, True, 8, 3),          # gear joints: revolute-revolute and revolute-prismatic (rack) trains
    ("sensors", 30, 1, 300, True, 8, 3),       # sensors: a slab bodies fall through, sensor halos on bodies
    ("sensors", 40, 2, 200, False, 8, 3),
    ("sensors", 8, 0, 200, True, 8, 3),
])
def test_world_trajectory_bitexact(ref, orc, kind, n, seed, steps, sleep, vel, pos):
    a = ref.world(kind, n, seed, sleep=sleep)
    b = orc.world(kind, n, seed, sleep=sleep)
    for s in range(steps):
        a.step(1, vel=vel, pos=pos)
        b.step(1, vel=vel, pos=pos)
        ba, bb = a.bodies(), b.bodies()
        assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32)), f"{kind} bodies differ at step {s + 1}"
    ma, na = a.contacts()
    mb, nb = b.contacts()
    assert np.array_equal(ma, mb), "contact lists differ (order or fixtures or flags)"
    assert np.array_equal(E.manifold_bits(na), E.manifold_bits(nb)), "manifolds differ"
    assert a.penetration() == b.penetration()
    nj = a.e.fn("world_joint_count")(a.h)
    assert nj == b.e.fn("world_joint_count")(b.h)
    if nj:
        ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
        a.e.fn("world_get_joints")(a.h, E.ptr(ja))
        b.e.fn("world_get_joints")(b.h, E.ptr(jb))
        assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32)), "joint impulses differ"
    a.close(); b.close()


def _edit_joints(w, kind):
    """Mid-run joint edits through each engine's setters: limits switched
    and moved, motors switched and retuned, joints destroyed."""
    if kind == "pendulums":
        w.e.fn("world_set_joint")(w.h, 0, 1, F(-0.2), F(0.25), 0, F(0.0), F(0.0))     # limit on
        w.e.fn("world_set_joint")(w.h, 1, 1, F(-0.1), F(0.1), 0, F(0.0), F(0.0))      # limits moved
        w.e.fn("world_set_joint")(w.h, 2, 0, F(0.0), F(0.0), 1, F(-2.0), F(5.0))      # motor retuned
        w.e.fn("world_set_joint")(w.h, 4, 0, F(-0.4), F(0.3), 0, F(0.0), F(0.0))      # limit off
        w.e.fn("world_set_joint")(w.h, 6, 1, F(0.0), F(0.0), 1, F(0.5), F(10.0))      # equal limits + motor
        assert w.e.fn("world_destroy_joint")(w.h, 3) == 0
    else:
        assert w.e.fn("world_destroy_joint")(w.h, 5) == 0                             # chain torn apart
        assert w.e.fn("world_destroy_joint")(w.h, 17) == 0


@pytest.mark.parametrize("kind,n,seed,sleep", [
    ("pendulums", 24, 0, True),
    ("chains", 2, 12, False),
])
def test_joint_edits_bitexact(ref, orc, kind, n, seed, sleep):
    a = ref.world(kind, n, seed, sleep=sleep)
    b = orc.world(kind, n, seed, sleep=sleep)
    a.step(90)
    b.step(90)
    _edit_joints(a, kind)
    _edit_joints(b, kind)
    for s in range(240):
        a.step(1)
        b.step(1)
        assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{kind} step {s + 91}"
    nj = a.e.fn("world_joint_count")(a.h)
    ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
    a.e.fn("world_get_joints")(a.h, E.ptr(ja))
    b.e.fn("world_get_joints")(b.h, E.ptr(jb))
    assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
    ma, _ = a.contacts()
    mb, _ = b.contacts()
    assert np.array_equal(ma, mb)
    c = orc.world(kind, n, seed, sleep=sleep)  # the edits change the motion
    c.step(330)
    assert not np.array_equal(c.bodies(), b.bodies())
    if kind == "pendulums":
        jb = jb.reshape(-1, 4)
        assert np.all(jb[3] == 0.0)
        cb, bb = c.bodies(), b.bodies()
        assert cb[4, E.CY] > bb[4, E.CY] + 10.0  # the box of the destroyed joint falls
    a.close(); b.close(); c.close()


def _destroy_bodies(w, ids):
    for i in ids:
        assert w.e.fn("world_destroy_body")(w.h, int(i)) == 0, i


def proxy_ids(w):
    """Per fixture proxy in creation order: its broad-phase tree id (-1 once
    destroyed)."""
    n = w.e.fn("world_fixture_count")(w.h)
    out = np.zeros(max(n, 1), np.int32)
    assert w.e.fn("world_proxy_ids")(w.h, E.ptr(out), n) == n
    return out[:n]


# (steps before, bodies destroyed, bodies then added by the "rain" scene, its seed)
DESTROY_PLAN = [(60, 12, 30, 7), (90, 8, 20, 8), (45, 25, 0, 0), (60, 0, 40, 9), (120, 0, 0, 0)]


def destroy_plan_run(a, b, compare, seed=3, n=150, plan=DESTROY_PLAN, fixtures=True):
    """Drives two worlds loaded with the same 150-body "world" scene through
    `plan` (the same bodies destroyed in both, picked by `seed`, and with
    `fixtures` three more bodies' fixtures), calling compare(tag, alive)
    after every step and after every destroy or load; returns which bodies
    are alive."""
    rng = np.random.default_rng(seed)
    gone = set()
    alive = np.ones(n + 1, bool)  # body 0 is the container
    for k, (steps, nd, na, s) in enumerate(plan):
        for i in range(steps):
            a.step(1)
            b.step(1)
            compare(f"phase {k} step {i + 1}", alive)
        if nd:
            ids = rng.choice(np.nonzero(alive[1:])[0] + 1, nd, replace=False)
            _destroy_bodies(a, ids)
            _destroy_bodies(b, ids)
            alive[ids] = False
            compare(f"phase {k} destroyed {sorted(ids.tolist())}", alive)
            if fixtures:
                # b2Body::DestroyFixture on a few live bodies' only fixture: they
                # fall on as shapeless bodies of unit mass
                body, _ = b.fixture_info()
                live = [f for f in range(len(body)) if body[f] > 0 and alive[body[f]] and f not in gone]
                for f in rng.choice(live, min(3, len(live)), replace=False):
                    for w in (a, b):
                        assert w.e.fn("world_destroy_fixture")(w.h, int(f)) == 0, f
                    gone.add(int(f))
                compare(f"phase {k} destroyed fixtures", alive)
        if na:
            for w in (a, b):
                assert w.e.fn("world_load_scene")(w.h, b"rain", na, ctypes.c_uint(s)) == 0
            alive = np.concatenate([alive, np.ones(na, bool)])
            compare(f"phase {k} added {na}", alive)
    return alive


def test_destroy_body_reuses_proxy_ids_bitexact(ref, orc):
    """b2World::DestroyBody frees the body's tree proxies (REF
    b2World.cpp:133-215, b2DynamicTree.cpp:56-131); fixtures created later take
    the freed ids, so new contacts come in another order and with other A/B
    roles than creation order gives. The oracle's free-list model must give
    the reference build's proxy ids exactly, and with them the same contact
    list and trajectories, bit for bit."""
    a = ref.world("world", 150, 5, sleep=True)
    b = orc.world("world", 150, 5, sleep=True)

    def compare(tag, alive):
        pa, pb = proxy_ids(a), proxy_ids(b)
        assert np.array_equal(pa, pb), f"{tag}: proxy ids differ at {np.nonzero(pa != pb)[0][:8]}"
        ba, bb = a.bodies()[alive], b.bodies()[alive]
        assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32)), f"{tag}: bodies differ"
        ma, _ = a.contacts()
        mb, _ = b.contacts()
        assert np.array_equal(ma, mb), f"{tag}: contact lists differ"

    destroy_plan_run(a, b, compare)
    ids = proxy_ids(a)
    live = ids[ids >= 0]
    # freed ids were taken again: the ids no longer grow with creation order
    assert np.any(np.diff(live) < 0)
    assert len(np.unique(live)) == len(live)
    a.close(); b.close()


# (steps before, [(body, change)]): change 0 static, 1 kinematic, 2 dynamic,
# 3 deactivate, 4 activate
TYPE_PLAN = [(60, [(5, 0), (17, 0), (40, 1), (0, 2), (60, 3), (61, 3)]),
             (90, [(5, 2), (0, 0), (33, 0), (40, 2), (60, 4), (90, 3)]),
             (60, [(17, 2), (80, 1), (81, 0), (61, 4)]), (120, [(90, 4)]), (60, [])]


def type_plan_run(a, b, compare):
    """b2Body::SetType and SetActive mid-run on two worlds loaded with the
    150-body "world" scene: bodies made static and kinematic and back, the
    container made dynamic and static again, bodies deactivated and
    reactivated; compare(tag) after every step and change."""
    for k, (steps, changes) in enumerate(TYPE_PLAN):
        for i in range(steps):
            a.step(1)
            b.step(1)
            compare(f"phase {k} step {i + 1}")
        for body, t in changes:
            for w in (a, b):
                if t < 3:
                    assert w.e.fn("world_set_type")(w.h, body, t) == 0
                else:
                    assert w.e.fn("world_set_active")(w.h, body, t - 3) == 0
        if changes:
            compare(f"phase {k} types changed")


def test_set_type_bitexact(ref, orc):
    """The oracle's SetType and SetActive (REF b2Body.cpp:115-172, 456-496)
    against the reference build: bodies and contact lists bit for bit
    through the type plan."""
    a = ref.world("world", 150, 5, sleep=True)
    b = orc.world("world", 150, 5, sleep=True)

    def compare(tag):
        assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{tag}: bodies differ"
        ma, _ = a.contacts()
        mb, _ = b.contacts()
        assert np.array_equal(ma, mb), f"{tag}: contact lists differ"

    type_plan_run(a, b, compare)
    a.close(); b.close()


def test_pendulum_closed_forms(orc):
    """Free pendulums: the hinge holds each box at its length from the
    anchor (position error below the linear slop) and the motorised joints
    spin their box at the motor speed while the torque suffices."""
    w = orc.world("pendulums", 24, 0, sleep=False)
    b0 = w.bodies()
    w.step(240)
    b = w.bodies()
    anchor_y = 20.0
    for i in range(24):
        x = 5.0 * i
        length = 1.0 + 0.05 * (i % 20)
        r = np.hypot(b[1 + i, E.CX] - x, b[1 + i, E.CY] - anchor_y)
        assert abs(r - length) < 0.005, (i, r, length)
    w.close()


def test_pyramid_reference_facts(orc):
    """Pyramid with sleeping on: 590 contacts, about 800 manifold points (799
    with continuous collision on, 800 without), asleep by the end, mean
    penetration about 0.7 cm (the reference's values)."""
    w = orc.world("pyramid", 20, 0, sleep=True)
    w.step(600)
    bodies = w.bodies()
    mean, mx, npts = w.penetration()
    assert w.e.fn("world_contact_count")(w.h) == 590
    assert npts == 800
    assert np.all(bodies[1:, E.AWAKE] == 0)
    assert 0.004 < mean < 0.01 and mx < 0.02


def test_hello_world_closed_form(orc):
    """HelloWorld: the box falls freely for the first steps (v = -g t in FP32)
    and comes to rest on the ground at y = 1 within the polygon skin."""
    w = orc.world("hello", sleep=True)
    h = np.float32(1.0 / 60.0)
    v = np.float32(0.0)
    y = np.float32(4.0)
    for s in range(10):
        w.step(1, vel=6, pos=2)
        v = np.float32(v + h * np.float32(np.float32(1.0) * np.float32(-10.0)))
        y = np.float32(y + h * v)
        b = w.bodies()[1]
        assert b[E.VY] == v and b[E.CY] == y
    w.step(50, vel=6, pos=2)
    b = w.bodies()[1]
    assert abs(b[E.PY] - 1.0) < 0.02 and abs(b[E.ANG]) < 1e-3


# ------------------------------------------------------------ narrowphase
def _compare_collide(reflib, orclib, A, XA, B, XB):
    # b2Collide* leave the impulse fields unwritten, so compare geometry and ids.
    mr = corpus.collide_all(reflib, "ref_", A, XA, B, XB)
    mo = corpus.collide_all(orclib, "orc_", A, XA, B, XB)
    bad = np.where(~np.all(E.geometry_bits(mr) == E.geometry_bits(mo), axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} manifolds differ, first case {bad[:5]}"
    return mr


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_collide_corpus_bitexact(reflib, orclib, seed):
    A, XA, B, XB = corpus.random_corpus(reflib, "ref_", 400, seed)
    m = _compare_collide(reflib, orclib, A, XA, B, XB)
    counts = np.bincount(m["count"], minlength=3)
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0, counts


def test_collide_special_cases_bitexact(reflib, orclib):
    A, XA, B, XB = corpus.special_corpus(reflib, "ref_")
    _compare_collide(reflib, orclib, A, XA, B, XB)


def _xf(lib, pre, x, y, a):
    t = np.zeros(1, E.XF)
    getattr(lib, pre + "transform")(F(x), F(y), F(a), E.ptr(t))
    return t


def test_reference_known_answers(reflib, orclib):
    """The reference's collision tests (Regression/tests/helloworld.cpp:268-418),
    with closed forms for what they compute."""
    for lib, pre in ((reflib, "ref_"), (orclib, "orc_")):
        box = np.array([E.box(lib, pre, 1.0, 1.0)], E.SHAPE)
        # Testb2CollideCirclesTouching (REF placement: centres 1 apart, radii 1)
        c = np.array([E.circle(1.0)], E.SHAPE)
        m = np.zeros(1, E.MANIFOLD)
        getattr(lib, pre + "collide")(E.ptr(c), E.ptr(_xf(lib, pre, 0, 0, 1)), E.ptr(c), E.ptr(_xf(lib, pre, -1, 0, 1)), E.ptr(m))
        assert m[0]["count"] == 1 and m[0]["type"] == 0 and m[0]["points"][0]["id"] == 0
        # Testb2FindMaxSeparation: unit boxes 3 apart, both turned 1 rad:
        # separation = 3 sin(1) - 2 along the face normals.
        edge = ctypes.c_int(0)
        s = getattr(lib, pre + "find_max_separation")(E.ptr(box), E.ptr(_xf(lib, pre, 2, 0, 1)), E.ptr(box),
                                                      E.ptr(_xf(lib, pre, -1, 0, 1)), ctypes.byref(edge))
        assert abs(s - (3 * np.sin(1.0) - 2)) < 1e-5
        # overlap variant: centres 2 apart -> 2 sin(1) - 2 < 0
        s2 = getattr(lib, pre + "find_max_separation")(E.ptr(box), E.ptr(_xf(lib, pre, 1, 0, 1)), E.ptr(box),
                                                       E.ptr(_xf(lib, pre, -1, 0, 1)), ctypes.byref(edge))
        assert abs(s2 - (2 * np.sin(1.0) - 2)) < 1e-5
        # Testb2FindIncidentEdge: axis-aligned boxes touching at x = 0: the
        # incident edge of box 2 is its right face x = 0, from y = -1 to 1.
        s3 = getattr(lib, pre + "find_max_separation")(E.ptr(box), E.ptr(_xf(lib, pre, 1, 0, 0)), E.ptr(box),
                                                       E.ptr(_xf(lib, pre, -1, 0, 0)), ctypes.byref(edge))
        assert s3 == 0.0 and edge.value == 3
        out = np.zeros(4, np.float32)
        ids = np.zeros(2, np.uint32)
        getattr(lib, pre + "find_incident_edge")(E.ptr(box), E.ptr(_xf(lib, pre, 1, 0, 0)), edge.value, E.ptr(box),
                                                 E.ptr(_xf(lib, pre, -1, 0, 0)), E.ptr(out), E.ptr(ids))
        assert np.allclose(out, [0, -1, 0, 1]) or np.allclose(out, [0, 1, 0, -1])
        # Testb2CollidePolygonsTouching: two points, normal along the face.
        getattr(lib, pre + "collide")(E.ptr(box), E.ptr(_xf(lib, pre, 1, 0, 1)), E.ptr(box),
                                      E.ptr(_xf(lib, pre, -1, 0, 1)), E.ptr(m))
        assert m[0]["count"] == 2
    # and both checkers agree bit for bit on all of them (compared in the corpus tests too)


def test_box_on_box_closed_form(orclib):
    """Box resting on a box: two points at the clipped corners, face normal
    (0, 1); the side planes are pushed out by the two skins (0.02), so the
    corner beyond the reference face clips at x = 0.5 + 0.02."""
    box = np.array([E.box(orclib, "orc_", 0.5, 0.5)], E.SHAPE)
    xa = np.array([E.transform(0, 0, 0)], E.XF)
    xb = np.array([E.transform(0.25, 0.995, 0)], E.XF)
    m = np.zeros(1, E.MANIFOLD)
    orclib.orc_collide(E.ptr(box), E.ptr(xa), E.ptr(box), E.ptr(xb), E.ptr(m))
    assert m[0]["count"] == 2 and m[0]["type"] in (1, 2)
    wm = np.zeros(8, np.float32)
    ra = np.float32(0.01)
    orclib.orc_world_manifold(E.ptr(m), E.ptr(xa), F(ra), E.ptr(xb), F(ra), E.ptr(wm))
    assert np.allclose(wm[:2], [0, 1], atol=1e-6)
    xs = sorted([wm[2], wm[4]])
    assert abs(xs[0] - (-0.25)) < 1e-5 and abs(xs[1] - 0.52) < 1e-5
    assert np.allclose(wm[6:8], 0.995 - 1.0 - 0.02, atol=1e-6)


# ------------------------------------------------------------ shapes
def test_polygon_hull_and_mass(reflib, orclib):
    rng = np.random.default_rng(5)
    for i in range(300):
        k = int(rng.integers(3, 9))
        pts = rng.uniform(-2, 2, (k, 2)).astype(np.float32)
        if i % 5 == 0:
            pts[1] = pts[0] + 1e-3  # welded duplicate
        a = E.polygon_from_points(reflib, "ref_", pts)
        b = E.polygon_from_points(orclib, "orc_", pts)
        assert a.tobytes() == b.tobytes(), i
        ma, mb = np.zeros(4, np.float32), np.zeros(4, np.float32)
        sa = np.array([a], E.SHAPE)
        reflib.ref_shape_mass(E.ptr(sa), F(1.3), E.ptr(ma))
        orclib.orc_shape_mass(E.ptr(sa), F(1.3), E.ptr(mb))
        assert ma.tobytes() == mb.tobytes(), i
    for (hx, hy, cx, cy, ang, rot) in ((0.5, 0.5, 0, 0, 0, 0), (2, 0.3, 0.4, -0.2, 0.7, 1), (1, 1, 0, 0, 1.2, 1)):
        a = E.box(reflib, "ref_", hx, hy, cx, cy, ang, rot)
        b = E.box(orclib, "orc_", hx, hy, cx, cy, ang, rot)
        assert a.tobytes() == b.tobytes()
        mb = np.zeros(4, np.float32)
        orclib.orc_shape_mass(E.ptr(np.array([b], E.SHAPE)), F(2.0), E.ptr(mb))
        mass = 2.0 * 4 * hx * hy
        inertia_c = mass * (hx * hx + hy * hy) / 3.0
        assert abs(mb[0] - mass) < 1e-5 * mass
        assert abs(mb[3] - (inertia_c + mass * (cx * cx + cy * cy))) < 1e-4 * mb[3]
    # circle mass: pi r^2 rho, I = m (r^2 / 2 + |p|^2)
    c = np.array([E.circle(0.7, 0.1, -0.2)], E.SHAPE)
    mc = np.zeros(4, np.float32)
    orclib.orc_shape_mass(E.ptr(c), F(1.0), E.ptr(mc))
    m = np.pi * 0.49
    assert abs(mc[0] - m) < 1e-6 and abs(mc[3] - m * (0.245 + 0.05)) < 1e-6


# ------------------------------------------------------------ broad-phase
def _ref_pairs(reflib, bp, cap=1 << 20):
    out = np.zeros(2 * cap, np.int32)
    n = reflib.ref_bp_update_pairs(ctypes.c_void_p(bp), E.ptr(out), cap)
    return out[:2 * n].reshape(-1, 2)


@pytest.mark.parametrize("seed", [11, 12])
def test_update_pairs_vs_bruteforce(reflib, orclib, seed):
    """b2BroadPhase::UpdatePairs (tree queries) equals the O(n^2) pair set on
    the same fat AABBs and move set, in the same sorted order."""
    rng = np.random.default_rng(seed)
    n = 600
    lo = rng.uniform(-40, 40, (n, 2)).astype(np.float32)
    size = rng.uniform(0.2, 3.0, (n, 2)).astype(np.float32)
    size[::97] *= 20  # a few wide proxies
    tight = np.zeros(n, E.AABB)
    tight["lx"], tight["ly"] = lo[:, 0], lo[:, 1]
    tight["ux"], tight["uy"] = lo[:, 0] + size[:, 0], lo[:, 1] + size[:, 1]
    bp = reflib.ref_bp_create()
    for i in range(n):
        reflib.ref_bp_create_proxy(ctypes.c_void_p(bp), E.ptr(tight[i:i + 1]), i)
    ids = np.arange(n)  # proxy k gets tree node 0, 1, 3, 5, ...: order preserved, mapped by user index
    got = _ref_pairs(reflib, bp)
    fat = np.zeros(n, E.AABB)
    node = np.where(ids == 0, 0, 2 * ids - 1)
    for i in range(n):
        reflib.ref_bp_get_fat(ctypes.c_void_p(bp), int(node[i]), E.ptr(fat[i:i + 1]))
    moved = np.ones(n, np.uint8)
    want = np.zeros(2 * (1 << 20), np.int32)
    cnt = orclib.orc_pairs_bruteforce(E.ptr(fat), E.ptr(moved), n, E.ptr(want), 1 << 20)
    assert np.array_equal(got, want[:2 * cnt].reshape(-1, 2))
    # second round: move a subset (fat-box rule decides which re-enter the move buffer)
    moved = np.zeros(n, np.uint8)
    for i in rng.choice(n, 150, replace=False):
        d = rng.uniform(-0.4, 0.4, 2).astype(np.float32)
        nt = tight[i:i + 1].copy()
        nt["lx"] += d[0]; nt["ux"] += d[0]; nt["ly"] += d[1]; nt["uy"] += d[1]
        fat_before = fat[i:i + 1].copy()
        exp = np.zeros(1, E.AABB)
        em = orclib.orc_fat_rule(E.ptr(fat_before), E.ptr(nt), F(d[0]), F(d[1]), E.ptr(exp))
        m = reflib.ref_bp_move_proxy(ctypes.c_void_p(bp), int(node[i]), E.ptr(nt), F(d[0]), F(d[1]))
        reflib.ref_bp_get_fat(ctypes.c_void_p(bp), int(node[i]), E.ptr(fat[i:i + 1]))
        assert m == em and fat[i:i + 1].tobytes() == exp.tobytes()
        moved[i] = max(moved[i], m)
    got = _ref_pairs(reflib, bp)
    cnt = orclib.orc_pairs_bruteforce(E.ptr(fat), E.ptr(moved), n, E.ptr(want), 1 << 20)
    assert np.array_equal(got, want[:2 * cnt].reshape(-1, 2))
    reflib.ref_bp_destroy(ctypes.c_void_p(bp))


def test_trig_modes(orclib):
    """Mode 0 is glibc sinf/cosf; mode 1 the rounded double sin/cos; they agree
    to within one ulp and differ for about 1% of angles."""
    a = np.linspace(-60, 60, 200001).astype(np.float32)
    s0, c0 = np.zeros_like(a), np.zeros_like(a)
    orclib.orc_sincos(E.ptr(a), len(a), E.ptr(s0), E.ptr(c0))
    s1 = np.sin(a.astype(np.float64)).astype(np.float32)
    c1 = np.cos(a.astype(np.float64)).astype(np.float32)
    ulp = lambda x, y: np.abs(x.view(np.int32).astype(np.int64) - y.view(np.int32).astype(np.int64))
    assert ulp(s0, s1).max() <= 1 and ulp(c0, c1).max() <= 1
    frac = (ulp(s0, s1) > 0).mean()
    assert 0.0 < frac < 0.05


def _events(w):
    n = w.e.fn("world_events")(w.h, None, 0)
    ev = np.zeros(5 * max(n, 1), np.int32)
    w.e.fn("world_events")(w.h, E.ptr(ev), n)
    return ev[:5 * n].reshape(-1, 5)


def test_one_sided_platform_listener(ref):
    """The listener the GPU's PreSolve path is checked with, on the
    reference build: boxes launched at 10.5 m/s from y = 1 pass up through
    the platform (top y = 5.25) while PreSolve disables its contacts, peak
    near 1 + v^2 / 2g = 6.51, land on it (origin 5.25 + 0.25 above, less the
    skin) and fall asleep. Each box begins touching the platform twice (rising
    through, landing) and ends once; every PreSolve that disabled a contact
    came while the box was below the clearance."""
    w = ref.world("oneway", 8, 0, sleep=True)
    w.e.fn("world_set_recorder")(w.h, 1, F(5.25), F(0.1))
    peak = np.full(8, -1e9)
    for _ in range(240):
        w.step(1)
        peak = np.maximum(peak, w.bodies()[2:, E.CY])
    b = w.bodies()[2:]
    assert np.all(np.abs(peak - (1.0 + 10.5 ** 2 / 20.0)) < 0.05), peak
    assert np.all(np.abs(b[:, E.CY] - 5.5) < 0.02) and np.all(b[:, E.AWAKE] == 0)
    ev = _events(w)
    plat = (ev[:, 2] == 1) | (ev[:, 3] == 1)
    for k, per_box in ((1, 2), (2, 1)):
        assert ((ev[:, 1] == k) & plat).sum() == 8 * per_box, k
    assert ((ev[:, 1] == 3) & (ev[:, 4] == 0)).sum() > 8
    w.close()


def test_sensor_facts(ref, orc):
    """What b2Contact::Update's sensor branch fixes, on the reference build
    (listener record) and the oracle (contact table): the row of bodies falls
    through the sensor slab (fixture 1) to the ground, each solid fixture
    and each halo (the sensor circle after every third body's solid)
    beginning and ending touching the slab once; no PreSolve or PostSolve names a
    sensor; every sensor contact keeps no manifold points; the bodies rest
    on the ground."""
    n = 8
    w = ref.world("sensors", n, 0, sleep=True)
    w.e.fn("world_set_recorder")(w.h, -1, F(0.0), F(0.0))
    o = orc.world("sensors", n, 0, sleep=True)
    for _ in range(200):
        w.step(1)
        o.step(1)
    solid, halo = [], []
    f = 2
    for i in range(n):
        solid.append(f)
        f += 1
        if i % 3 == 0:
            halo.append(f)
            f += 1
    ev = _events(w)
    with_slab = (ev[:, 2] == 1) | (ev[:, 3] == 1)
    for k in (1, 2):
        other = np.where(ev[:, 2] == 1, ev[:, 3], ev[:, 2])[(ev[:, 1] == k) & with_slab]
        assert sorted(other.tolist()) == sorted(solid + halo), (k, other)
    sensors = set(halo) | {1}
    for kind in (3, 4):
        named = ev[ev[:, 1] == kind][:, 2:4]
        assert not (set(named.ravel().tolist()) & sensors), kind
    meta, man = o.contacts()
    sens = np.isin(meta[:, 0], list(sensors)) | np.isin(meta[:, 1], list(sensors))
    assert sens.any() and np.all(man[sens]["count"] == 0)
    b = w.bodies()[2:]
    assert np.all(b[:, E.CY] < 1.0), b[:, E.CY]
    w.close()
    o.close()


def test_conveyor_belt_facts(ref):
    """b2Contact::SetTangentSpeed on the reference build (a listener giving
    the belt's contacts 5 m/s from PreSolve): a box that lands on the belt
    slides until it moves with the belt, so while on the belt it reaches
    v_x = 5 m/s (friction 0.8: 0.6 s at mu g); the boxes leave the belt's
    right end (x = 10) and come to rest beyond it, below its top."""
    n = 6
    w = ref.world("conveyor", n, 0, sleep=True)
    w.e.fn("world_set_conveyor")(w.h, 1, F(5.0))
    top = np.full(n, -1e9)
    for _ in range(600):
        w.step(1)
        b = w.bodies()[2:]
        on = np.abs(b[:, E.CY] - 5.65) < 0.02
        top = np.where(on, np.maximum(top, b[:, E.VX]), top)
    assert np.all(np.abs(top - 5.0) < 1e-3), top
    b = w.bodies()[2:]
    assert np.all(b[:, E.CX] > 10.0) and np.all(b[:, E.CY] < 5.0), b[:, [E.CX, E.CY]]
    assert np.all(np.abs(b[:, [E.VX, E.VY]]) < 0.05), b[:, [E.VX, E.VY]]
    w.close()


# ------------------------------------------------------------ continuous collision
@pytest.mark.parametrize("kind,n,seed,steps,sleep,substep", [
    ("bullets", 24, 1, 120, True, False),      # bullets and fast bodies into a column and a thin wall
    ("bullets", 48, 2, 200, False, False),
    ("bullets", 24, 3, 120, True, True),       # sub-stepping: one TOI event per step
    ("pyramid", 20, 0, 300, True, False),      # the reference defaults (sleep and CCD on)
    ("pile", 400, 3, 300, False, False),       # fast bodies against the static container
    ("terrain", 60, 3, 300, True, False),      # chain shapes
])
def test_ccd_trajectory_bitexact(ref, orc, kind, n, seed, steps, sleep, substep):
    a = ref.world(kind, n, seed, sleep=sleep, ccd=True, substep=substep)
    b = orc.world(kind, n, seed, sleep=sleep, ccd=True, substep=substep)
    for s in range(steps):
        a.step(1)
        b.step(1)
        ba, bb = a.bodies(), b.bodies()
        assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32)), f"{kind} bodies differ at step {s + 1}"
    ma, na = a.contacts()
    mb, nb = b.contacts()
    assert np.array_equal(ma, mb), "contact lists differ"
    assert np.array_equal(E.manifold_bits(na), E.manifold_bits(nb)), "manifolds differ"
    a.close(); b.close()


def test_ccd_stops_tunnelling(orc):
    """The bullets scene's wall is 0.2 m thick and bodies travel 1.3-3.3 m
    per step: without TOI many end up behind it, with TOI none does."""
    past = {}
    for ccd in (False, True):
        w = orc.world("bullets", 48, 2, sleep=False, ccd=ccd)
        worst = 0
        for _ in range(200):
            w.step(1)
            worst = max(worst, int((w.bodies()[7:, E.PX] > 12.0).sum()))
        past[ccd] = worst
        w.close()
    assert past[False] >= 10 and past[True] == 0, past


def _seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / max(np.dot(ab, ab), 1e-300), 0.0, 1.0)
    return np.hypot(*(p - (a + t * ab)))


def _world_verts(s, xf):
    n = 1 if s["type"] == 0 else (2 if s["type"] == 1 else int(s["count"]))
    if s["type"] == 0:
        v = np.array([[s["centroid"]["x"], s["centroid"]["y"]]], np.float64)
    else:
        v = np.array([[s["verts"][i]["x"], s["verts"][i]["y"]] for i in range(n)], np.float64)
    c, sn = float(xf["c"]), float(xf["s"])
    return np.stack([c * v[:, 0] - sn * v[:, 1] + float(xf["px"]), sn * v[:, 0] + c * v[:, 1] + float(xf["py"])], 1)


def _cross2(u, v):
    return u[0] * v[1] - u[1] * v[0]


def _edges(P):
    return [(P[i], P[(i + 1) % len(P)]) for i in range(len(P))] if len(P) > 1 else [(P[0], P[0])]


def _brute_distance(va, vb):
    """Distance between two convex point sets' hulls (segments, vertices):
    0 when a vertex of one lies inside the other or two edges cross, else
    the smallest vertex-to-edge distance either way."""
    def inside(p, poly):
        if len(poly) < 3:
            return False
        cr = [_cross2(poly[(i + 1) % len(poly)] - poly[i], p - poly[i]) for i in range(len(poly))]
        return all(c >= 0 for c in cr) or all(c <= 0 for c in cr)

    def crosses(a, b, c, d):
        d1, d2 = _cross2(b - a, c - a), _cross2(b - a, d - a)
        d3, d4 = _cross2(d - c, a - c), _cross2(d - c, b - c)
        return d1 * d2 < 0 and d3 * d4 < 0
    if any(inside(p, vb) for p in va) or any(inside(p, va) for p in vb):
        return 0.0
    if any(crosses(a, b, c, d) for a, b in _edges(va) for c, d in _edges(vb)):
        return 0.0
    best = np.inf
    for P, Q in ((va, vb), (vb, va)):
        for p in P:
            for a, b in _edges(Q):
                best = min(best, _seg_dist(p, a, b))
    return best


@pytest.mark.parametrize("seed", [5, 6])
def test_distance_matches_reference(reflib, orclib, seed):
    """GJK distance (b2Distance, no radii, empty cache) equals the reference
    build bit for bit, and the brute-force hull distance in FP64."""
    A, XA, B, XB = corpus.random_corpus(orclib, "orc_", 60, seed)
    out_o, out_r = np.zeros(7, np.float32), np.zeros(7, np.float32)
    checked = 0
    for i in range(len(A)):
        orclib.orc_distance(E.ptr(A[i:i + 1]), E.ptr(XA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(XB[i:i + 1]), E.ptr(out_o))
        reflib.ref_distance(E.ptr(A[i:i + 1]), E.ptr(XA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(XB[i:i + 1]), E.ptr(out_r))
        keep = [0, 1, 2, 3, 4, 6]
        assert np.array_equal(out_o[keep].view(np.uint32), out_r[keep].view(np.uint32)), f"case {i}"
        d = _brute_distance(_world_verts(A[i], XA[i]), _world_verts(B[i], XB[i]))
        scale = 1.0 + np.abs(_world_verts(A[i], XA[i])).max()
        if d > 1e-3:
            assert abs(out_o[4] - d) <= 2e-5 * scale, (i, out_o[4], d)
            checked += 1
        else:
            assert out_o[4] <= 1e-3 * scale, (i, out_o[4], d)
    assert checked > 100


@pytest.mark.parametrize("seed", [7, 8])
def test_toi_matches_reference(reflib, orclib, seed):
    """b2TimeOfImpact state and time equal the reference build's bit for
    bit over swept pairs of every shape combination."""
    A, SA, B, SB = corpus.sweep_corpus(orclib, "orc_", 60, seed)
    t_o, t_r = np.zeros(1, np.float32), np.zeros(1, np.float32)
    states = []
    for i in range(len(A)):
        so = orclib.orc_time_of_impact(E.ptr(A[i:i + 1]), E.ptr(SA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(SB[i:i + 1]),
                                       F(1.0), E.ptr(t_o))
        sr = reflib.ref_time_of_impact(E.ptr(A[i:i + 1]), E.ptr(SA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(SB[i:i + 1]),
                                       F(1.0), E.ptr(t_r))
        assert so == sr and t_o.view(np.uint32)[0] == t_r.view(np.uint32)[0], (i, so, sr, t_o[0], t_r[0])
        states.append(so)
    states = np.array(states)
    assert (states == 3).sum() > 50 and (states == 4).sum() > 20  # touching and separated both occur


def test_toi_closed_form(orclib):
    """Circle B translating head-on towards circle A: the time returned is
    where the centre distance reaches max(linearSlop, rA + rB - 3 linearSlop)
    within the root finder's tolerance of linearSlop / 4."""
    slop = 0.005
    t = np.zeros(1, np.float32)
    for rA, rB, d0, L in [(0.5, 0.5, 5.0, 8.0), (0.1, 0.3, 3.0, 40.0), (1.0, 0.05, 2.5, 2.0), (0.02, 0.02, 1.0, 3.0)]:
        A, B = np.array([E.circle(rA)], E.SHAPE), np.array([E.circle(rB)], E.SHAPE)
        sa = np.zeros(1, E.SWEEP)
        sb = np.zeros(1, E.SWEEP)
        sb["c0"] = (d0, 0.0)
        sb["c"] = (d0 - L, 0.0)
        state = orclib.orc_time_of_impact(E.ptr(A), E.ptr(sa), E.ptr(B), E.ptr(sb), F(1.0), E.ptr(t))
        target = max(slop, rA + rB - 3 * slop)
        assert state == 3, state
        dist = d0 - float(t[0]) * L
        assert abs(dist - target) < 0.25 * slop + 1e-6, (rA, rB, d0, L, dist, target)


def test_chain_children(ref, orc):
    """A chain fixture is one proxy per child edge (b2ChainShape): the child
    edges' ghost vertices (open chain with and without SetPrev/NextVertex, a
    loop closed across its seam) equal the reference's GetChildEdge, and each
    child's fat box is its segment's box without the skin, widened by the
    AABB extension (0.1 m), as the reference's broad-phase holds it."""
    a = ref.world("terrain", 10, 1)
    b = orc.world("terrain", 10, 1)
    n = a.e.fn("world_fixture_count")(a.h)
    assert n == b.e.fn("world_fixture_count")(b.h) == 7 + 3 + 6 + 10
    sa, sb = np.zeros(n, E.SHAPE), np.zeros(n, E.SHAPE)
    a.e.fn("world_get_shapes")(a.h, E.ptr(sa))
    b.e.fn("world_get_shapes")(b.h, E.ptr(sb))
    assert np.array_equal(sa[:16].view(np.uint8), sb[:16].view(np.uint8))
    flags = sa["flags"][:16]
    assert np.all(flags & 4)
    assert flags[0] & 1 and flags[6] & 2                 # ghosts given beyond the floor's ends
    assert not (flags[7] & 1) and not (flags[9] & 2)     # the ramp has free ends
    assert np.all(flags[10:16] & 3 == 3)                 # a loop is smooth all round
    fa, fb = a.fat_aabbs(), b.fat_aabbs()
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    for i in range(16):
        v = np.array([[sa[i]["verts"][k]["x"], sa[i]["verts"][k]["y"]] for k in range(2)], np.float32)
        lo, hi = v.min(0) - np.float32(0.1), v.max(0) + np.float32(0.1)
        assert np.array_equal(np.array([fa[i]["lx"], fa[i]["ly"]], np.float32), lo), i
        assert np.array_equal(np.array([fa[i]["ux"], fa[i]["uy"]], np.float32), hi), i
    a.step(100); b.step(100)
    assert np.array_equal(a.fat_aabbs().view(np.uint32), b.fat_aabbs().view(np.uint32))
    meta, _ = b.contacts()
    assert np.any((meta[:, 0] < 16) & (meta[:, 2] & 1 == 1))  # bodies rest on chain children
    a.close(); b.close()


def test_distance_joint_edits_bitexact(ref, orc):
    """b2DistanceJoint::SetLength, SetFrequency and SetDampingRatio mid-run
    (a rod made a spring, a spring made a rod, a rod shortened) against the
    reference's own setters, every step, with the joints' reaction forces."""
    a = ref.world("distance", 8, 1, sleep=True)
    b = orc.world("distance", 8, 1, sleep=True)
    a.step(80); b.step(80)
    for w in (a, b):
        assert w.e.fn("world_set_distance")(w.h, 0, F(5.0), F(1.0), F(0.5)) == 0
        assert w.e.fn("world_set_distance")(w.h, 1, F(6.0), F(0.0), F(0.0)) == 0
        assert w.e.fn("world_set_distance")(w.h, 2, F(3.0), F(0.0), F(0.0)) == 0
    for s in range(200):
        a.step(1); b.step(1)
        assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"step {s + 81}"
    ja, jb = np.zeros(32, np.float32), np.zeros(32, np.float32)
    a.e.fn("world_get_joints")(a.h, E.ptr(ja))
    b.e.fn("world_get_joints")(b.h, E.ptr(jb))
    assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
    a.close(); b.close()


@pytest.mark.parametrize("hz", [1.0, 2.0])
def test_distance_spring_closed_form(orc, hz):
    """A box on an undamped distance spring without gravity oscillates along
    the spring at the frequency it was given: the period between upward
    crossings of the rest position is 1 / hz within 3%; with damping ratio 1
    (critical) it creeps back without crossing."""
    w = orc.world("spring", int(round(10 * hz)), 0, gravity=(0.0, 0.0))
    ys, t = [], []
    for s in range(300):
        w.step(1)
        ys.append(float(w.bodies()[1, E.CY]))
        t.append((s + 1) / 60.0)
    y = np.array(ys) - 7.0  # rest: 3 m below the anchor at y = 10
    up = [t[i] + (t[i + 1] - t[i]) * (-y[i]) / (y[i + 1] - y[i]) for i in range(len(y) - 1) if y[i] < 0 <= y[i + 1]]
    assert len(up) >= 3
    period = np.diff(up).mean()
    assert abs(period - 1.0 / hz) < 0.03 / hz, (period, hz)
    w.close()
    c = orc.world("spring", int(round(10 * hz)), 10, gravity=(0.0, 0.0))
    yc = []
    for _ in range(300):
        c.step(1)
        yc.append(float(c.bodies()[1, E.CY]) - 7.0)
    assert max(yc) < 1e-3 and yc[-1] < 0.0 and yc[-1] > -0.05, (max(yc), yc[-1])
    c.close()


def test_rope_and_weld_edits_bitexact(ref, orc):
    """b2RopeJoint::SetMaxLength and b2WeldJoint::SetFrequency /
    SetDampingRatio mid-run (a rope shortened and lengthened, a rigid weld
    made soft, a soft one made rigid) against the reference's setters, with
    the joints' reaction forces and torques."""
    for kind, n, edits in (("ropes", 8, lambda w: (w.e.fn("world_set_rope")(w.h, 0, F(2.0)),
                                                    w.e.fn("world_set_rope")(w.h, 3, F(6.0)))),
                           ("welds", 6, lambda w: (w.e.fn("world_set_weld")(w.h, 0, F(3.0), F(0.2)),
                                                   w.e.fn("world_set_weld")(w.h, 5, F(0.0), F(0.0))))):
        a = ref.world(kind, n, 0, sleep=True)
        b = orc.world(kind, n, 0, sleep=True)
        a.step(80); b.step(80)
        for w in (a, b):
            assert edits(w) == (0, 0)
        assert b.e.fn("world_set_rope" if kind == "welds" else "world_set_weld")(b.h, 0, F(1.0), F(0.0)) != 0
        for s in range(200):
            a.step(1); b.step(1)
            assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{kind} step {s + 81}"
        nj = a.e.fn("world_joint_count")(a.h)
        ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
        a.e.fn("world_get_joints")(a.h, E.ptr(ja))
        b.e.fn("world_get_joints")(b.h, E.ptr(jb))
        assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
        a.close(); b.close()


def test_rope_slack_free_fall_and_bound(orc):
    """A box that starts on its rope's anchor (rope of 4 m, box 3 of the
    unchained rope scene, moving at -3 m/s) falls freely while the rope is
    slack: no impulse, and the positions of symplectic Euler under gravity,
    x_k = x0 - 3 k h and y_k = 16 - g h^2 k (k + 1) / 2. Afterwards the
    rope holds it: the centre never gets further than 4 m + 1 cm from the
    anchor."""
    w = orc.world("ropes", 4, 1, sleep=False)
    b0 = w.bodies()
    x0 = float(b0[4, E.CX])
    h = 1.0 / 60.0
    anchor = np.array([6.0 * 3 - 12.0, 16.0])
    worst = 0.0
    for k in range(1, 301):
        w.step(1)
        c = w.bodies()[4, [E.CX, E.CY]].astype(np.float64)
        if k <= 30:
            assert abs(c[0] - (x0 - 3.0 * k * h)) < 1e-4, (k, c[0])
            assert abs(c[1] - (16.0 - 10.0 * h * h * k * (k + 1) / 2)) < 1e-4, (k, c[1])
            assert w.joints()[3, 0] == 0.0
        worst = max(worst, float(np.hypot(*(c - anchor))))
    assert 3.99 < worst < 4.01, worst
    w.close()


def test_weld_cantilever_statics(orc):
    """A rigid cantilever of four 1 x 0.25 m links (density 1: 0.25 kg
    each) welded to the ground at its left end comes to rest, and the
    wall weld then carries the statics of the beam: linear impulse
    (0, W h) with W = 1 kg x 10 m/s^2 and angular impulse M h with the
    moment M = 0.25 x 10 x (0.5 + 1.5 + 2.5 + 3.5) = 20 N m about the wall,
    within 2%. The welds in between carry the load of the links beyond them."""
    w = orc.world("welds", 1, 1, sleep=False)
    w.step(600)
    j = w.joints() * 60.0  # reaction force and torque at inv_dt = 60
    assert abs(j[0, 0]) < 0.05 and abs(j[0, 1] - 10.0) < 0.2 and abs(j[0, 2] - 20.0) < 0.4, j[0]
    for k in range(1, 4):  # weld k holds links k..3: force 2.5 (4 - k), moment about its anchor
        arm = sum(0.5 + q for q in range(4 - k))
        assert abs(j[k, 1] - 2.5 * (4 - k)) < 0.05 * 2.5 * (4 - k), (k, j[k])
        assert abs(j[k, 2] - 2.5 * arm) < 0.05 * 2.5 * arm, (k, j[k], arm)
    w.close()


def test_friction_and_motor_edits_bitexact(ref, orc):
    """b2FrictionJoint::SetMaxForce / SetMaxTorque and b2MotorJoint's
    offset, force, torque and correction setters mid-run (a box set free, a
    box held harder, a target moved and turned) against the reference's
    setters, with the joints' reaction forces and torques."""
    for kind, n, edits in (
            ("friction", 8, lambda w: (w.e.fn("world_set_friction")(w.h, 0, F(0.0), F(0.0)),
                                       w.e.fn("world_set_friction")(w.h, 3, F(20.0), F(4.0)))),
            ("motor", 6, lambda w: (w.e.fn("world_set_motor")(w.h, 0, F(-8.0), F(6.0), F(2.0), F(30.0), F(8.0),
                                                               F(0.4)),
                                    w.e.fn("world_set_motor")(w.h, 2, F(6.0), F(0.5), F(1.5), F(5.0), F(1.0),
                                                               F(0.2))))):
        a = ref.world(kind, n, 0, sleep=True)
        b = orc.world(kind, n, 0, sleep=True)
        a.step(90); b.step(90)
        for w in (a, b):
            assert edits(w) == (0, 0)
        for s in range(200):
            a.step(1); b.step(1)
            assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{kind} step {s + 91}"
        nj = a.e.fn("world_joint_count")(a.h)
        ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
        a.e.fn("world_get_joints")(a.h, E.ptr(ja))
        b.e.fn("world_get_joints")(b.h, E.ptr(jb))
        assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
        a.close(); b.close()


def test_friction_joint_closed_form(orc):
    """Without gravity, box 0 of the friction scene (1 kg, 1 x 1 m, I = 1/6,
    thrown at 2 m/s with 1 rad/s, maxForce 2 N, maxTorque 0.5 N m) slows at
    exactly F / m and T / I: speed 2 - 2 k h and spin 1 - 3 k h after k steps,
    until it stops (at k = 60 and k = 20) and stays stopped."""
    w = orc.world("friction", 16, 0, sleep=False, gravity=(0.0, 0.0))
    h = 1.0 / 60.0
    for k in range(1, 91):
        w.step(1)
        b = w.bodies()[1]
        speed, spin = float(np.hypot(b[E.VX], b[E.VY])), float(b[E.W])
        want_v, want_w = max(2.0 - 2.0 * k * h, 0.0), max(1.0 - 3.0 * k * h, 0.0)
        assert abs(speed - want_v) < 1e-4 and abs(spin - want_w) < 1e-4, (k, speed, want_v, spin, want_w)
    w.close()


def test_motor_joint_reaches_offsets(orc):
    """Without gravity every box of the motor scene ends at its target: the
    linear offset from its body A (in A's frame) within 1 mm and the angular
    offset within 1e-3 rad, whether driven from the ground or from the
    previous box."""
    n = 12
    w = orc.world("motor", n, 0, sleep=False, gravity=(0.0, 0.0))
    w.step(900)
    bodies = w.bodies()
    for i in range(n):
        b = bodies[1 + i]
        from_prev = i % 3 == 2
        x = 6.0 * i - 3.0 * n
        if from_prev:
            a = bodies[i]
            s, c = float(a[E.QS]), float(a[E.QC])
            off = (6.0, 0.5)
            target = np.array([a[E.PX] + c * off[0] - s * off[1], a[E.PY] + s * off[0] + c * off[1]])
            base_angle = float(a[E.ANG])
        else:
            target = np.array([x + 1.5 * (1 + i % 2), 4.0 + 1.0 + 0.5 * (i % 3)])
            base_angle = 0.0
        angle = base_angle + 0.5 + 0.5 * (i % 3)
        assert np.hypot(*(b[[E.PX, E.PY]] - target)) < 1e-3, (i, b[[E.PX, E.PY]], target)
        assert abs(float(b[E.ANG]) - angle) < 1e-3, (i, b[E.ANG], angle)
    w.close()


def test_pulley_and_mouse_edits_bitexact(ref, orc):
    """b2MouseJoint::SetTarget, SetMaxForce, SetFrequency and SetDampingRatio
    mid-run (a target moved sideways, a spring stiffened, a force limit cut)
    on the mouse scene, and the pulley scene run long enough for a box to
    reach its pulley, against the reference, with the reaction forces."""
    for kind, n, edits in (
            ("mouse", 6, lambda w: (w.e.fn("world_set_mouse")(w.h, 0, F(-9.0), F(5.0), F(1000.0), F(8.0), F(0.3)),
                                    w.e.fn("world_set_mouse")(w.h, 3, F(2.5), F(1.0), F(3.0), F(2.0), F(0.7)))),
            ("pulley", 6, lambda w: (0, 0))):
        a = ref.world(kind, n, 0, sleep=True)
        b = orc.world(kind, n, 0, sleep=True)
        a.step(90); b.step(90)
        for w in (a, b):
            assert edits(w) == (0, 0)
        for s in range(300):
            a.step(1); b.step(1)
            assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{kind} step {s + 91}"
        nj = a.e.fn("world_joint_count")(a.h)
        ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
        a.e.fn("world_get_joints")(a.h, E.ptr(ja))
        b.e.fn("world_get_joints")(b.h, E.ptr(jb))
        assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
        a.close(); b.close()


def test_pulley_atwood_closed_form(orc):
    """Boxes hung by their centres on a pulley of ratio r accelerate as an
    Atwood machine: body B at a_B = g (r m_A - m_B) / (m_B + r^2 m_A) and body
    A at -r a_B, so y_k = y_0 + a h^2 k (k + 1) / 2 after k steps (1 mm over
    30 steps), with lengthA + r lengthB held. Pulley 0: r = 1, m_A = 1,
    m_B = 0.64; pulley 2: r = 2, m_A = 1, m_B = 1.96."""
    w = orc.world("pulley", 8, 0, sleep=False)
    h, g = 1.0 / 60.0, 10.0
    y0 = w.bodies()[:, E.CY].copy()
    for k in range(1, 31):
        w.step(1)
        y = w.bodies()[:, E.CY]
        for i, r, mb in ((0, 1.0, 0.64), (2, 2.0, 1.96)):
            ab = g * (r * 1.0 - mb) / (mb + r * r * 1.0)
            for body, a in ((1 + 2 * i, -r * ab), (2 + 2 * i, ab)):
                want = y0[body] + a * h * h * k * (k + 1) / 2
                assert abs(y[body] - want) < 1e-3, (i, body, k, y[body], want)
    w.close()


def test_mouse_joint_hooke_closed_form(orc):
    """A box hanging at rest from a mouse joint of frequency f sags below the
    target by g / (2 pi f)^2 (Hooke's law with k = m (2 pi f)^2), and the
    joint carries its weight: reaction force (0, m g) within 1%."""
    w = orc.world("mouse", 6, 0, sleep=False)
    w.step(2400)  # the box also swings about the grab point, damped slowly
    bodies, j = w.bodies(), w.joints() * 60.0
    for i in range(3):
        hz = 2.0 + 2.0 * i
        b = bodies[1 + i]
        s, c = float(b[E.QS]), float(b[E.QC])
        # grab point (0.4, 0.4) on the box, at its local anchor
        anchor = np.array([b[E.PX] + c * 0.4 - s * 0.4, b[E.PY] + s * 0.4 + c * 0.4])
        target = np.array([4.0 * i - 12.0 + 0.4, 3.5 + 0.5 * i])
        sag = 10.0 / (2.0 * np.pi * hz) ** 2
        assert abs(anchor[0] - target[0]) < 2e-3 and abs(target[1] - anchor[1] - sag) < 2e-3, (i, anchor, target, sag)
        assert abs(j[i, 0]) < 0.02 and abs(j[i, 1] - 10.0) < 0.1, (i, j[i])
    w.close()


def test_prismatic_and_wheel_edits_bitexact(ref, orc):
    """b2PrismaticJoint's EnableLimit, SetLimits, EnableMotor, SetMotorSpeed,
    SetMaxMotorForce and b2WheelJoint's EnableMotor, SetMotorSpeed,
    SetMaxMotorTorque, SetSpringFrequencyHz, SetSpringDampingRatio mid-run,
    against the reference's setters, with the reaction forces and torques."""
    for kind, n, edits in (
            ("prismatic", 8, lambda w: (w.e.fn("world_set_prismatic")(w.h, 2, 1, F(-0.5), F(0.5), 1, F(-1.0), F(50.0)),
                                        w.e.fn("world_set_prismatic")(w.h, 3, 0, F(-1.0), F(2.0), 0, F(0.0), F(0.0)))),
            ("wheels", 4, lambda w: (w.e.fn("world_set_wheel")(w.h, 0, 1, F(5.0), F(40.0), F(2.0), F(0.2)),
                                     w.e.fn("world_set_wheel")(w.h, 3, 1, F(-20.0), F(10.0), F(8.0), F(1.0))))):
        a = ref.world(kind, n, 0, sleep=True)
        b = orc.world(kind, n, 0, sleep=True)
        a.step(60); b.step(60)
        for w in (a, b):
            assert edits(w) == (0, 0)
        for s in range(240):
            a.step(1); b.step(1)
            assert np.array_equal(a.bodies().view(np.uint32), b.bodies().view(np.uint32)), f"{kind} step {s + 61}"
        nj = a.e.fn("world_joint_count")(a.h)
        ja, jb = np.zeros(4 * nj, np.float32), np.zeros(4 * nj, np.float32)
        a.e.fn("world_get_joints")(a.h, E.ptr(ja))
        b.e.fn("world_get_joints")(b.h, E.ptr(jb))
        assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))
        a.close(); b.close()


def test_prismatic_closed_forms(orc):
    """Prismatic joints of the prismatic scene (no contact with the ground,
    which is body A): box 2 (axis at 30 degrees, free, started at 1 m/s
    along x) slides as a frictionless incline, s_k = 0.866 k h - 5 h^2
    k (k + 1) / 2 along the axis with no drift across it; box 3 (45 degrees,
    limit [-1, 2]) stops at -1 m; box 1 (15 degrees, motor 2 m/s, 30 N >
    2.6 N of gravity along the axis) runs at 2 m/s."""
    w = orc.world("prismatic", 12, 0, sleep=False)
    h = 1.0 / 60.0
    p0 = w.bodies()[:, [E.PX, E.PY]].astype(np.float64)
    axis = lambda i: np.array([np.cos(0.2618 * (i % 6)), np.sin(0.2618 * (i % 6))])
    lowest = 0.0
    for k in range(1, 181):
        w.step(1)
        p = w.bodies()[:, [E.PX, E.PY]].astype(np.float64)
        d2 = p[3] - p0[3]
        s2 = float(np.dot(d2, axis(2)))
        if k <= 60:
            want = np.cos(0.5236) * k * h - 5.0 * h * h * k * (k + 1) / 2
            assert abs(s2 - want) < 1e-3, (k, s2, want)
            ax2 = axis(2)
            assert abs(float(ax2[0] * d2[1] - ax2[1] * d2[0])) < 1e-4
        lowest = min(lowest, float(np.dot(p[4] - p0[4], axis(3))))
    assert -1.0 - 0.01 < lowest < -0.99, lowest
    v1 = w.bodies()[2, [E.VX, E.VY]].astype(np.float64)
    assert abs(float(np.dot(v1, axis(1))) - 2.0) < 1e-3, v1
    w.close()


def gear_train_check(w, steps, tol):
    """Shared by the oracle and GPU gear tests: train 0 of the gears scene
    (disc 1: r 1, density 5; disc 2: r 2, density 5; rack 1 x 6 m, density
    1) keeps its gear relations, angle1 = -angle2 / 2 and angle2 = y / 2 for
    the rack's drop y, and while free the rack falls at
    a = -m g / (m + I2 / 4 + I1 / 16) (one degree of freedom, the energy of
    the train), so y_k = a h^2 k (k + 1) / 2 within `tol` of its value."""
    h = 1.0 / 60.0
    m, I1, I2 = 6.0, 0.5 * 5.0 * np.pi * 1.0, 0.5 * 20.0 * np.pi * 4.0
    a = -m * 10.0 / (m + I2 / 4.0 + I1 / 16.0)
    y0 = float(w.bodies()[3, E.CY])
    for k in range(1, steps + 1):
        w.step(1)
        b = w.bodies()
        y = float(b[3, E.CY]) - y0
        a1, a2 = float(b[1, E.ANG]), float(b[2, E.ANG])
        assert abs(a1 + 0.5 * a2) < 1e-3 and abs(a2 - 0.5 * y) < 1e-3, (k, a1, a2, y)
        want = a * h * h * k * (k + 1) / 2
        if k >= 10:
            assert abs(y - want) < tol * abs(want), (k, y, want)


def test_gear_train_closed_form(orc):
    w = orc.world("gears", 4, 0, sleep=False)
    gear_train_check(w, 60, 0.02)
    w.close()
