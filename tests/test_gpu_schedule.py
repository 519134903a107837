"""The step schedule (island labelling method, collide order, workgroup
sizes, trailing-colour fusion) changes how the coloured step runs, not what
it computes: worlds stepped under different schedules stay identical bit for
bit, and the autotuner's trial steps are ordinary steps.

The clustered solve (schedule field clusterBodies) changes the order of a
pass: constraints inside a body cluster first, by one workgroup per cluster
with the bodies in LDS, then those between clusters. With one cluster holding
every awake body the order is the unclustered one, and the steps are
identical bit for bit; with many clusters the steps are reproducible and stay
within bounds of the unclustered solve."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu

DEFAULT = (0, 0, 256, 256, 256, 0, 0, 0)
SCHEDULES = [(1, 1, 64, 128, 64, 1024, 0, 0), (2, 1, 512, 64, 128, 8192, 0, 0), (2, 0, 128, 1024, 256, 512, 0, 0),
             (0, 1, 192, 64, 512, 2048, 0, 0), (1, 2, 256, 256, 64, 2048, 0, 0),
             # one-launch solve: grid barriers between colours, workgroup sizes, trailing colours
             (0, 0, 256, 256, 256, 0, 0, 1), (2, 1, 1024, 256, 64, 2048, 0, 2), (0, 0, 64, 256, 256, 512, 0, 16)]


def _world(gpu):
    # three stress worlds: piles, edges and jointed chains in one context
    return E.World.worlds(gpu, "stress", 3, 1500, 5, sleep=True)


def _state(w):
    meta, man = w.contacts()
    return (w.bodies() + np.float32(0)).view(np.uint32), meta, E.manifold_bits(man), (w.joints() + np.float32(0)).view(np.uint32)


def _set(w, sched):
    s = np.array([sched], E.SCHEDULE)
    rc = w.e.lib.b2pp_world_set_schedule(w.h, E.ptr(s))
    assert rc == 0, w.error()


def _same(a, b):
    sa, sb = _state(a), _state(b)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("sched", SCHEDULES)
def test_schedules_give_identical_steps(gpu, sched):
    a, b = _world(gpu), _world(gpu)
    _set(a, DEFAULT)
    _set(b, sched)
    for _ in range(4):
        a.step(50)
        b.step(50)
        _same(a, b)
    assert a.stats()["contacts"] > 0
    a.close(); b.close()


def test_autotune_steps_are_ordinary_steps(gpu):
    a, b = _world(gpu), _world(gpu)
    a.step(100)
    b.step(100)
    rounds = 2
    out = np.zeros(1, E.SCHEDULE)
    rc = b.e.lib.b2pp_world_autotune(b.h, ctypes.c_float(1.0 / 60.0), 8, 3, rounds, E.ptr(out))
    assert rc == 0, b.error()
    a.step(20 * rounds)  # one step per candidate value: 3 + 3 + 4 + 3 + 3 + 4 per round
    _same(a, b)
    o = out[0]
    assert o["islands"] in (0, 1, 2) and o["collideOrder"] in (0, 1, 2)
    assert o["solverBlock"] in (64, 128, 256, 512) and o["collideBlock"] in (64, 128, 256)
    assert o["queryBlock"] in (64, 128, 256) and o["tailColours"] in (0, 512, 1024, 2048)
    assert o["persistentSolve"] == 0 and o["clusterBodies"] == 0  # opt-in only
    a.step(60)
    b.step(60)
    _same(a, b)
    a.close(); b.close()


def test_set_schedule_rejects_bad_values(gpu):
    w = gpu.world("pyramid", 4, 0)
    for bad in [(3, 0, 256, 256, 256, 0, 0, 0), (0, 3, 256, 256, 256, 0, 0, 0), (0, 0, 100, 256, 256, 0, 0, 0),
                (0, 0, 256, 2048, 256, 0, 0, 0), (0, 0, 256, 256, 0, 0, 0, 0), (0, 0, 256, 256, 256, -1, 0, 0),
                (0, 0, 256, 256, 256, 9000, 0, 0), (0, 0, 256, 256, 256, 0, 32, 0), (0, 0, 256, 256, 256, 0, 100, 0),
                (0, 0, 256, 256, 256, 0, 4096, 0), (0, 0, 256, 256, 256, 0, 0, -1), (0, 0, 256, 256, 256, 0, 0, 17)]:
        s = np.array([bad], E.SCHEDULE)
        assert w.e.lib.b2pp_world_set_schedule(w.h, E.ptr(s)) != 0
    w.close()


@pytest.mark.parametrize("scene", ["stress", "plank", "pendulums"])
def test_graph_solve_matches_direct(gpu, scene):
    """The solve captured as a hipGraph (replayed while its layout repeats)
    and the directly launched solve give the same steps bit for bit: a
    stress context (piles, edges, jointed chains, sleep), a plank whose
    contacts overflow into the serial colour, and pendulums (joints only)."""
    lib = E.b2g_lib()

    def make():
        if scene == "stress":
            return _world(gpu)
        if scene == "plank":
            return gpu.world("plank", 100, 0, sleep=True)
        return gpu.world("pendulums", 24, 0, sleep=True)

    a, b = make(), make()
    ca, cb = a.device_context(), b.device_context()
    E.b2g_check(lib, ca, lib.b2g_set_graphs(ca, 1))
    E.b2g_check(lib, cb, lib.b2g_set_graphs(cb, 0))
    for _ in range(5):
        a.step(40)
        b.step(40)
        _same(a, b)
    launches, captures = ctypes.c_int64(), ctypes.c_int64()
    E.b2g_check(lib, ca, lib.b2g_get_graph_counts(ca, ctypes.byref(launches), ctypes.byref(captures)))
    assert launches.value > 50, (launches.value, captures.value)
    E.b2g_check(lib, cb, lib.b2g_get_graph_counts(cb, ctypes.byref(launches), None))
    assert launches.value == 0
    a.close(); b.close()


def _pen_energy(w):
    m = w.measures()
    return m["mean_penetration_m"], m["kinetic_energy_J"]


def test_single_cluster_is_the_unclustered_solve(gpu):
    a = gpu.world("pile", 1200, 9, sleep=False)
    b = gpu.world("pile", 1200, 9, sleep=False)
    _set(a, DEFAULT)
    _set(b, (0, 0, 256, 256, 256, 1024, 2048, 0))
    for _ in range(6):
        a.step(50)
        b.step(50)
        _same(a, b)
    assert a.stats()["constraints"] > 100
    a.close(); b.close()


@pytest.mark.parametrize("size", [64, 256])
def test_many_clusters(gpu, size):
    lib = E.b2g_lib()
    plain, a, b = _world(gpu), _world(gpu), _world(gpu)
    _set(plain, DEFAULT)
    _set(a, (0, 0, 256, 256, 256, 1024, size, 0))
    _set(b, (0, 0, 128, 256, 256, 0, size, 0))   # other workgroup size and no fusion: same order
    conflicts, mask = ctypes.c_int(-1), ctypes.c_int(-1)
    for _ in range(4):
        plain.step(50)
        a.step(50)
        b.step(50)
        _same(a, b)
        E.b2g_check(lib, a.device_context(), lib.b2g_check_colouring(a.device_context(), ctypes.byref(conflicts),
                                                                     ctypes.byref(mask)))
        assert conflicts.value == 0 and mask.value == 0
    pp, kp = _pen_energy(plain)
    pa, ka = _pen_energy(a)
    assert np.isfinite(a.bodies()).all()
    assert pa <= pp + 0.005, (pa, pp)
    assert abs(ka / kp - 1.0) <= 0.2, (ka, kp)
    plain.close(); a.close(); b.close()


FUSED = (0, 0, 256, 256, 256, 1024, 0, 0)
PERSIST = (0, 0, 256, 256, 256, 1024, 0, 2)


@pytest.mark.parametrize("scene,n,seed,sleep", [("plank", 100, 0, True), ("pendulums", 24, 0, True),
                                                ("chains", 6, 16, True), ("pile", 3000, 4, False)])
def test_persistent_solve_matches_per_colour(gpu, scene, n, seed, sleep):
    """The one-launch solve (grid barriers between colours) against a launch
    per colour, bit for bit: a plank whose contacts overflow into the serial
    colour, pendulums (joints only), chains on the ground (joints and
    contacts) and a pile."""
    a = gpu.world(scene, n, seed, sleep=sleep)
    b = gpu.world(scene, n, seed, sleep=sleep)
    _set(a, FUSED)
    _set(b, PERSIST)
    for _ in range(4):
        a.step(40)
        b.step(40)
        _same(a, b)
    a.close(); b.close()


def _colour_counts(lib, w):
    counts = np.zeros(65, np.int32)
    n = ctypes.c_int()
    E.b2g_check(lib, w.device_context(), lib.b2g_get_colour_counts(w.device_context(), E.ptr(counts), 65,
                                                                   ctypes.byref(n)))
    assert n.value == 65
    return counts


def test_colour_demotion(gpu):
    """Demotion rounds (b2g_set_colour_demotion) move constraints down to
    colours freed since they were coloured: on a settling pile the pass has
    fewer colours than without them, the colouring invariant (no dynamic body
    twice in a colour, masks equal to the constraints' colours) holds after
    every step, the colour counts add up to the constraints, and the state
    stays within the bounds of the solve without demotion (the visit order
    differs)."""
    lib = E.b2g_lib()
    plain = gpu.world("pile", 3000, 4, sleep=False)
    dem = gpu.world("pile", 3000, 4, sleep=False)
    E.b2g_check(lib, plain.device_context(), lib.b2g_set_colour_demotion(plain.device_context(), 0))
    E.b2g_check(lib, dem.device_context(), lib.b2g_set_colour_demotion(dem.device_context(), 1))
    assert lib.b2g_set_colour_demotion(dem.device_context(), -1) != 0
    conf, mask = ctypes.c_int(-1), ctypes.c_int(-1)
    for s in range(500):
        plain.step(1)
        dem.step(1)
        if s % 5 == 4:
            E.b2g_check(lib, dem.device_context(), lib.b2g_check_colouring(dem.device_context(), ctypes.byref(conf),
                                                                           ctypes.byref(mask)))
            assert conf.value == 0 and mask.value == 0, (s, conf.value, mask.value)
    cp, cd = _colour_counts(lib, plain), _colour_counts(lib, dem)
    assert cd.sum() == dem.stats()["constraints"] and cp.sum() == plain.stats()["constraints"]
    assert cd.sum() > 1000
    # constraints sit in lower colours: the mean colour drops, the count does not grow
    k = np.arange(65)
    assert (cd * k).sum() / cd.sum() < (cp * k).sum() / cp.sum(), (cp[:16], cd[:16])
    assert np.count_nonzero(cd) <= np.count_nonzero(cp), (cp[:16], cd[:16])
    pp, kp = _pen_energy(plain)
    pd, kd = _pen_energy(dem)
    assert np.isfinite(dem.bodies()).all()
    assert pd <= pp + 0.005, (pd, pp)
    assert abs(kd / kp - 1.0) <= 0.25, (kd, kp)
    plain.close(); dem.close()
