"""b2ContactListener on the GPU path against the reference build.

With a listener set, b2World::Step splits at the narrowphase
(b2g_step_collide / b2g_step_solve), with either solver, and makes the
reference's calls inside the step: over the contact list, newest first,
EndContact for destroyed touching contacts, BeginContact / EndContact on
touching changes and PreSolve with the old manifold; PreSolve's
SetEnabled(false) keeps the contact out of the step's solve. PostSolve
follows for the island contacts, in the order of b2World::Solve's island
search (island after island, each in the order the search adds them), the
recorded flag being the bits of the first normal impulse.

The one-sided platform scene has one contact per island and bodies of fixed
rotation, so the coloured solver's order is the reference's and no in-step
trig rounds differently: poses and the whole sequence of listener calls must
equal the reference's exactly. The reference-order solver keeps them equal
with several contacts per island too (the conveyor's boxes landing on each
other).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu
F = ctypes.c_float


def _events(w):
    n = w.e.fn("world_events")(w.h, None, 0)
    ev = np.zeros(5 * max(n, 1), np.int32)
    w.e.fn("world_events")(w.h, E.ptr(ev), n)
    return ev[:5 * n].reshape(-1, 5)


def _bits(a):
    return (a + np.float32(0)).view(np.uint32)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("sleep", [True, False])
def test_one_sided_platform_matches_reference(gpu, ref, sleep, mode):
    g = gpu.world("oneway", 8, 0, sleep=sleep)
    g.set_mode(mode)
    r = ref.world("oneway", 8, 0, sleep=sleep)
    for w in (g, r):
        w.e.fn("world_set_recorder")(w.h, 1, F(5.25), F(0.1))
    for s in range(240):
        g.step(1)
        r.step(1)
        gb, rb = g.bodies(), r.bodies()
        assert np.array_equal(_bits(gb[:, :11]), _bits(rb[:, :11])), f"step {s + 1}"
    ge, re_ = _events(g), _events(r)
    assert np.array_equal(ge, re_)                          # every call, in order, with PostSolve's impulses
    assert ((ge[:, 1] == 3) & (ge[:, 4] == 0)).sum() > 8  # PreSolve did disable contacts
    assert (ge[:, 1] == 4).sum() > 100
    assert np.all(np.abs(gb[2:, E.CY] - 5.5) < 0.02)       # the boxes landed on the platform
    g.close()
    r.close()


def test_recording_listener_without_rule_matches_plain_step(gpu):
    """A listener that only records changes nothing: the split step gives
    the same poses as the unsplit one, and every touching contact reports
    PostSolve while it is solved."""
    a = gpu.world("pyramid", 10, 0, sleep=True)
    b = gpu.world("pyramid", 10, 0, sleep=True)
    a.e.fn("world_set_recorder")(a.h, -1, F(0.0), F(0.0))
    for s in range(300):
        a.step(1)
        b.step(1)
    assert np.array_equal(_bits(a.bodies()), _bits(b.bodies()))
    ev = _events(a)
    assert (ev[:, 1] == 1).sum() >= 100 and (ev[:, 1] == 3).sum() > 1000 and (ev[:, 1] == 4).sum() > 1000
    assert np.all(ev[ev[:, 1] == 3][:, 4] == 1)
    a.close()
    b.close()


def test_split_step_state_errors(gpu):
    lib = E.b2g_lib()
    w = gpu.world("pyramid", 4, 0)
    c = w.device_context()
    p = np.zeros(1, E.STEP_PARAMS)
    p[0] = (1.0 / 60.0, 8, 3, 0, 0, 0.0)
    assert lib.b2g_step_solve(c) != 0                      # nothing open
    E.b2g_check(lib, c, lib.b2g_step_collide(c, E.ptr(p)))
    assert lib.b2g_step_collide(c, E.ptr(p)) != 0          # already open
    assert lib.b2g_step(c, E.ptr(p)) != 0                  # must finish first
    E.b2g_check(lib, c, lib.b2g_step_solve(c))
    E.b2g_check(lib, c, lib.b2g_step(c, E.ptr(p)))
    p[0]["mode"] = 1                                       # the reference-order solver splits too
    E.b2g_check(lib, c, lib.b2g_step_collide(c, E.ptr(p)))
    E.b2g_check(lib, c, lib.b2g_step_solve(c))
    p[0]["mode"] = 2
    assert lib.b2g_step_collide(c, E.ptr(p)) != 0          # no such solver
    w.close()


@pytest.mark.parametrize("ccd", [False, True])
def test_begin_end_balanced_with_ccd(gpu, ccd):
    """With continuous collision the TOI events change touching states after
    the narrowphase; those changes reach the listener after the step. Per
    fixture pair, BeginContact and EndContact still alternate starting with
    Begin, and the pairs left open are exactly the touching contacts."""
    g = gpu.world("bullets", 48, 2, sleep=False, ccd=ccd)
    g.e.fn("world_set_recorder")(g.h, -1, F(0.0), F(0.0))
    toi = 0
    for _ in range(160):
        g.step(1)
        toi += int(g.stats()["toiEvents"])
    ev = _events(g)
    open_pairs = {}
    for kind, a, b in ev[np.isin(ev[:, 1], (1, 2))][:, 1:4].tolist():
        key = (min(a, b), max(a, b))
        if kind == 1:
            assert not open_pairs.get(key, False), f"Begin twice for {key}"
            open_pairs[key] = True
        else:
            assert open_pairs.get(key, False), f"End without Begin for {key}"
            open_pairs[key] = False
    meta, _ = g.contacts()
    touching = {(min(a, b), max(a, b)) for a, b, f in meta[:, :3].tolist() if f & 1}
    assert {k for k, v in open_pairs.items() if v} == touching
    assert (toi > 20) == ccd
    g.close()


def test_sensors_match_reference(gpu, ref):
    """Sensor fixtures through the listener: a row of bodies of fixed
    rotation falling through a sensor slab, every third with a sensor halo,
    landing apart (one solid contact per island). Poses equal the reference
    build's bit for bit, and so does the sequence of Begin / End / PreSolve
    calls (sensor contacts report Begin and End, never PreSolve, and never
    join an island's PostSolve)."""
    g = gpu.world("sensors", 8, 0, sleep=True)
    r = ref.world("sensors", 8, 0, sleep=True)
    for w in (g, r):
        w.e.fn("world_set_recorder")(w.h, -1, F(0.0), F(0.0))
    for s in range(200):
        g.step(1)
        r.step(1)
        assert np.array_equal(_bits(g.bodies()[:, :11]), _bits(r.bodies()[:, :11])), f"step {s + 1}"
    ge, re_ = _events(g), _events(r)
    assert np.array_equal(ge, re_)
    slab = ((ge[:, 2] == 1) | (ge[:, 3] == 1))
    assert ((ge[:, 1] == 1) & slab).sum() == 8 + 3 and not ((ge[:, 1] == 3) & slab).any()
    g.close()
    r.close()


def test_conveyor_matches_reference(gpu, ref):
    """b2Contact::SetTangentSpeed from PreSolve (a conveyor belt carrying
    boxes of fixed rotation off its end). While each box has one solid
    contact (its island's only one: the first 180 steps, until boxes that
    fell off the end land on each other) poses equal the reference build's
    bit for bit, and so do the listener calls; all boxes end up beyond the
    belt's end."""
    g = gpu.world("conveyor", 6, 0, sleep=True)
    r = ref.world("conveyor", 6, 0, sleep=True)
    for w in (g, r):
        w.e.fn("world_set_conveyor")(w.h, 1, F(5.0))
    exact = 180
    for s in range(exact):
        g.step(1)
        r.step(1)
        assert np.array_equal(_bits(g.bodies()[:, :11]), _bits(r.bodies()[:, :11])), f"step {s + 1}"
    ge, re_ = _events(g), _events(r)
    assert np.array_equal(ge, re_)
    g.step(220)
    assert np.all(g.bodies()[2:, E.CX] > 10.0)
    g.close()
    r.close()


def test_conveyor_reference_order_islands(gpu, ref):
    """The conveyor with the reference-order solver over 400 steps, past the
    boxes landing on each other beyond the belt's end: islands of several
    contacts, solved in Box2D's order. Poses equal the reference build's bit
    for bit every step, and so does the whole sequence of listener calls,
    PostSolve island by island in the island search's order with its
    impulses."""
    g = gpu.world("conveyor", 6, 0, sleep=True)
    g.set_mode(1)
    r = ref.world("conveyor", 6, 0, sleep=True)
    for w in (g, r):
        w.e.fn("world_set_conveyor")(w.h, 1, F(5.0))
    for s in range(400):
        g.step(1)
        r.step(1)
        assert np.array_equal(_bits(g.bodies()[:, :11]), _bits(r.bodies()[:, :11])), f"step {s + 1}"
    ge, re_ = _events(g), _events(r)
    assert np.array_equal(ge, re_)
    post = ge[ge[:, 1] == 4]
    steps, counts = np.unique(post[:, 0], return_counts=True)
    assert counts.max() >= 3  # a step reported islands with several contacts
    g.close()
    r.close()
