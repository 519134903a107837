"""Continuous collision on the GPU step against the oracle.

* b2g_distance_batch and b2g_toi_batch (GJK and the time of impact, one lane
  per pair) equal the oracle's restatement bit for bit over seeded corpora,
  with the oracle evaluating rotations as the device does (trig mode 1).
* The reference-order mode with continuous physics on equals the oracle bit
  for bit step by step: bullets and fast bodies through a column of boxes
  into a thin wall (sleep on and off, and with sub-stepping), the pyramid at
  the reference's defaults, and a pile falling into its container. The TOI
  events themselves are then the reference's, in its order.
* The coloured mode with continuous physics on stops the same bodies at the
  wall that tunnel through it without.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import corpus
import engines as E

pytestmark = pytest.mark.gpu
REFERENCE = 1
F = ctypes.c_float


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


@pytest.fixture(scope="module")
def ctx():
    lib = E.b2g_lib()
    c = ctypes.c_void_p()
    assert lib.b2g_create(ctypes.byref(c), 0) == 0
    yield lib, c
    lib.b2g_destroy(c)


def _canon(b):
    return (b + np.float32(0.0)).view(np.uint32)


@pytest.mark.parametrize("seed", [5, 6])
def test_distance_batch_matches_oracle(ctx, orclib, seed):
    lib, c = ctx
    A, XA, B, XB = corpus.random_corpus(orclib, "orc_", 200, seed)
    n = len(A)
    out = np.zeros(2 * n, np.float32)
    E.b2g_check(lib, c, lib.b2g_distance_batch(c, E.ptr(A), E.ptr(XA), E.ptr(B), E.ptr(XB), n, E.ptr(out)))
    ref = np.zeros(7, np.float32)
    for i in range(n):
        orclib.orc_distance(E.ptr(A[i:i + 1]), E.ptr(XA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(XB[i:i + 1]), E.ptr(ref))
        assert out[2 * i].view(np.uint32) == ref[4].view(np.uint32) and out[2 * i + 1] == ref[6], (i, out[2 * i], ref)


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_toi_batch_matches_oracle(ctx, orclib, wide_trig, seed):
    lib, c = ctx
    A, SA, B, SB = corpus.sweep_corpus(orclib, "orc_", 200, seed)
    n = len(A)
    state = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    E.b2g_check(lib, c, lib.b2g_toi_batch(c, E.ptr(A), E.ptr(SA), E.ptr(B), E.ptr(SB), n, F(1.0), E.ptr(state),
                                          E.ptr(t)))
    to = np.zeros(1, np.float32)
    for i in range(n):
        so = orclib.orc_time_of_impact(E.ptr(A[i:i + 1]), E.ptr(SA[i:i + 1]), E.ptr(B[i:i + 1]), E.ptr(SB[i:i + 1]),
                                       F(1.0), E.ptr(to))
        assert state[i] == so and t[i].view(np.uint32) == to.view(np.uint32)[0], (i, state[i], so, t[i], to[0])
    assert (state == 3).sum() > 100 and (state == 4).sum() > 50


@pytest.mark.parametrize("kind,n,seed,steps,sleep,substep", [
    ("bullets", 24, 1, 120, True, False),
    ("bullets", 48, 2, 160, False, False),
    ("bullets", 24, 3, 120, True, True),
    ("pyramid", 20, 0, 300, True, False),
    ("pile", 400, 3, 250, False, False),
    ("terrain", 60, 3, 200, True, False),
])
def test_reference_mode_ccd_bitexact(gpu, orc, wide_trig, kind, n, seed, steps, sleep, substep):
    g = gpu.world(kind, n, seed, sleep=sleep, ccd=True, substep=substep)
    g.set_mode(REFERENCE)
    o = orc.world(kind, n, seed, sleep=sleep, ccd=True, substep=substep)
    events = 0
    for s in range(steps):
        g.step(1)
        o.step(1)
        events += g.stats()["toiEvents"]
        gb, ob = g.bodies(), o.bodies()
        bad = np.where(~np.all(_canon(gb[:, :11]) == _canon(ob[:, :11]), axis=1))[0]
        assert len(bad) == 0, f"step {s + 1}: {len(bad)} bodies differ, first {bad[:5]}"
    gm, gman = E.contact_table(*g.contacts())
    om, oman = E.contact_table(*o.contacts())
    assert np.array_equal(gm[:, :3], om[:, :3])
    assert np.array_equal(E.manifold_bits(gman), E.manifold_bits(oman))
    if kind == "bullets":
        assert events > 20, events
    g.close(); o.close()


def test_coloured_mode_ccd_stops_tunnelling(gpu):
    past = {}
    for ccd in (False, True):
        g = gpu.world("bullets", 48, 2, sleep=False, ccd=ccd)
        worst, events = 0, 0
        for _ in range(160):
            g.step(1)
            events += g.stats()["toiEvents"]
            worst = max(worst, int((g.bodies()[7:, E.PX] > 12.0).sum()))
        past[ccd] = (worst, events)
        g.close()
    assert past[False][0] >= 10 and past[False][1] == 0, past
    assert past[True][0] == 0 and past[True][1] > 20, past
