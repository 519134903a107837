"""Batched worlds: many independent worlds in one b2World's sub-worlds,
advanced together by one Step on one device context.

* With zero solver iterations every sub-world's trajectory equals, bit for
  bit, the oracle stepping that world alone (trig as the device evaluates
  it): the broad-phase keeps worlds apart and the per-body stages do not
  depend on the batch.
* With full iterations no contact joins two sub-worlds, the contact count of
  every sub-world equals that of the same world stepped alone on the GPU at
  the first step, and the batch stays finite and settles.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu


@pytest.fixture
def wide_trig(orclib):
    orclib.orc_set_trig_mode(1)
    yield
    orclib.orc_set_trig_mode(0)


def _batch(gpu, n_worlds, seed, lo, hi, sleep=False):
    return E.World.batch(gpu, n_worlds, seed, lo, hi, sleep)


def _slices(sizes):
    starts = np.concatenate([[0], np.cumsum(sizes + 1)])
    return [(int(starts[i]), int(starts[i + 1])) for i in range(len(sizes))]


def test_batch_zero_iterations_bitexact_per_world(gpu, orc, wide_trig):
    seed = 3
    h, sizes = _batch(gpu, 24, seed, 16, 256)
    sl = _slices(sizes)
    worlds = [orc.world("world", int(n), seed + i, sleep=False) for i, n in enumerate(sizes)]
    for s in range(90):
        h.step(1, vel=0, pos=0)
        for o in worlds:
            o.step(1, vel=0, pos=0)
        if s % 30 == 29:
            gb = h.bodies()
            for i, o in enumerate(worlds):
                ob = o.bodies()
                a, b = sl[i]
                g = (gb[a:b, :10] + np.float32(0)).view(np.uint32)
                r = (ob[:, :10] + np.float32(0)).view(np.uint32)
                bad = np.where(~np.all(g == r, axis=1))[0]
                assert len(bad) == 0, f"step {s + 1} world {i}: {len(bad)} bodies differ"
    for o in worlds:
        o.close()
    h.close()


def test_batch_keeps_worlds_apart(gpu):
    seed = 5
    h, sizes = _batch(gpu, 40, seed, 16, 512)
    sl = _slices(sizes)
    # fixture -> world: one fixture per body here except each container (3 edges)
    fw = []
    for i, n in enumerate(sizes):
        fw += [i] * (3 + int(n))
    fw = np.array(fw)
    h.step(1)
    meta, _ = h.contacts()
    assert len(meta) > 0
    assert np.all(fw[meta[:, 0]] == fw[meta[:, 1]]), "a contact joins two sub-worlds"
    per_world = np.bincount(fw[meta[:, 0]], minlength=len(sizes))
    for i in (0, 7, 19, 39):
        alone = gpu.world("world", int(sizes[i]), seed + i, sleep=False)
        alone.step(1)
        assert per_world[i] == len(alone.contacts()[0]), i
        alone.close()
    h.step(299)
    b = h.bodies()
    assert np.all(np.isfinite(b))
    meta, _ = h.contacts()
    assert np.all(fw[meta[:, 0]] == fw[meta[:, 1]])
    pen = h.penetration()
    assert pen[0] < 0.1, pen
    h.close()
