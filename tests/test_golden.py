"""Committed fixtures from the reference build (tests/golden/ref_vectors.npz,
written by tests/golden/make_golden.py): they pin the oracle and the GPU
path where oracle/_ref is not built.

* CPU: the oracle (glibc trig) reproduces every fixture bit for bit.
* GPU: the narrowphase batch reproduces the corpus manifolds bit for bit
  (transforms are given, so no trig is involved); whole steps differ from
  the reference by the trig rounding (the device rounds double sin/cos), so
  they are held to pose bands.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

REFERENCE = 1


def _bits(a):
    return (np.asarray(a, np.float32) + np.float32(0)).view(np.uint32)


def test_oracle_collide_corpus(orclib, gold):
    import corpus
    got = corpus.collide_all(orclib, "orc_", gold["corpus_A"], gold["corpus_XA"], gold["corpus_B"], gold["corpus_XB"])
    assert np.array_equal(E.geometry_bits(got), E.geometry_bits(gold["corpus_manifold"]))
    assert (gold["corpus_manifold"]["count"] > 0).sum() > 80


def test_oracle_hello_world(orc, gold):
    w = orc.world("hello", sleep=True)
    for s in range(60):
        w.step(1, vel=6, pos=2)
        assert np.array_equal(_bits(w.bodies()[1]), _bits(gold["hello_bodies"][s])), s
    w.close()


def test_oracle_pyramid(orc, gold):
    w = orc.world("pyramid", 20, 0, sleep=True)
    w.step(600)
    assert np.array_equal(_bits(w.bodies()), _bits(gold["pyramid_bodies"]))
    assert w.e.fn("world_contact_count")(w.h) == int(gold["pyramid_contacts"][0])
    assert np.array_equal(np.array(w.penetration()), gold["pyramid_penetration"])
    w.close()


def test_oracle_pendulums(orc, gold):
    w = orc.world("pendulums", 24, 0, sleep=True)
    w.step(200)
    assert np.array_equal(_bits(w.bodies()), _bits(gold["pendulum_bodies"]))
    assert np.array_equal(_bits(w.joints()), _bits(gold["pendulum_joints"]))
    w.close()


@pytest.mark.gpu
def test_gpu_collide_corpus(gold):
    lib = E.b2g_lib()
    ctx = ctypes.c_void_p()
    assert lib.b2g_create(ctypes.byref(ctx), 0) == 0
    A, XA, B, XB = gold["corpus_A"], gold["corpus_XA"], gold["corpus_B"], gold["corpus_XB"]
    out = np.zeros(len(A), E.MANIFOLD)
    E.b2g_check(lib, ctx, lib.b2g_collide_batch(ctx, E.ptr(A), E.ptr(XA), E.ptr(B), E.ptr(XB), len(A), E.ptr(out)))
    lib.b2g_destroy(ctx)
    assert np.array_equal(E.geometry_bits(out), E.geometry_bits(gold["corpus_manifold"]))


@pytest.mark.gpu
def test_gpu_reference_mode_hello_world(gpu, gold):
    g = gpu.world("hello", sleep=True)
    g.set_mode(REFERENCE)
    for s in range(60):
        g.step(1, vel=6, pos=2)
        d = np.abs(g.bodies()[1, :10] - gold["hello_bodies"][s, :10]).max()
        assert d <= 1e-5, (s, d)
    g.close()


@pytest.mark.gpu
def test_gpu_reference_mode_pyramid(gpu, gold):
    """600 steps with sleep: the same contact count, all boxes asleep, poses
    within 1 cm and the mean penetration within 1 mm of the reference's."""
    g = gpu.world("pyramid", 20, 0, sleep=True)
    g.set_mode(REFERENCE)
    g.step(600)
    b, want = g.bodies(), gold["pyramid_bodies"]
    assert g.e.fn("world_contact_count")(g.h) == int(gold["pyramid_contacts"][0])
    assert np.all(b[1:, E.AWAKE] == 0)
    assert np.hypot(b[:, E.CX] - want[:, E.CX], b[:, E.CY] - want[:, E.CY]).max() < 0.01
    assert abs(g.penetration()[0] - gold["pyramid_penetration"][0]) < 1e-3
    g.close()


@pytest.mark.gpu
def test_gpu_pendulums(gpu, gold):
    """Coloured mode, one joint per island: poses within 1e-4 and joint
    impulses within 1e-4 relative of the reference's after 200 steps."""
    g = gpu.world("pendulums", 24, 0, sleep=True)
    g.step(200)
    b, want = g.bodies(), gold["pendulum_bodies"]
    assert np.abs(b[:, [E.CX, E.CY, E.ANG]] - want[:, [E.CX, E.CY, E.ANG]]).max() < 1e-4
    j, jw = g.joints(), gold["pendulum_joints"]
    assert np.all(np.abs(j - jw) <= 1e-4 * (1.0 + np.abs(jw)))
    g.close()
