"""Writes tests/golden/ref_vectors.npz from the reference build
(oracle/_ref/libb2ref.so, upstream Box2D 2.3.2 compiled from the
reference's own sources by oracle/Makefile):

  corpus_*        narrowphase cases (shapes and transforms built by the
                  reference's own b2PolygonShape::Set / SetAsBox) and the
                  manifolds its b2Collide* functions return
  hello_bodies    HelloWorld (6/2 iterations), the box's row after each of 60 steps
  hello_ccd_bodies  the same with continuous collision on, the reference's
                  default (HelloWorld/HelloWorld.cpp sets no flag)
  pyramid_*       the 20-row pyramid, sleep on, 8/3 iterations, after 600 steps:
                  bodies, contact count, penetration (mean, max, points)
  pendulum_*      24 pendulums, sleep on, 8/3 iterations, after 200 steps:
                  bodies and joint impulses

The reference's b2Collide* functions leave the points' normal and tangent
impulses unset, so those bytes of corpus_manifold vary between runs; the
tests compare the manifold geometry only (tests/engines.py geometry_bits).
Body rows are those of *_world_get_bodies (tests/engines.py). Run from the
repository root after building oracle/_ref:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import corpus  # noqa: E402
import engines as E  # noqa: E402


def main() -> None:
    reflib = E.reference()
    ref = E.Engine("ref")
    out = {}
    A, XA, B, XB = corpus.random_corpus(reflib, "ref_", 40, 2024)
    SA, SXA, SB, SXB = corpus.special_corpus(reflib, "ref_")
    A, XA = np.concatenate([A, SA]), np.concatenate([XA, SXA])
    B, XB = np.concatenate([B, SB]), np.concatenate([XB, SXB])
    out.update(corpus_A=A, corpus_XA=XA, corpus_B=B, corpus_XB=XB,
               corpus_manifold=corpus.collide_all(reflib, "ref_", A, XA, B, XB))

    w = ref.world("hello", sleep=True)
    rows = []
    for _ in range(60):
        w.step(1, vel=6, pos=2)
        rows.append(w.bodies()[1])
    out["hello_bodies"] = np.array(rows, np.float32)
    w.close()

    w = ref.world("hello", sleep=True, ccd=True)
    rows = []
    for _ in range(60):
        w.step(1, vel=6, pos=2)
        rows.append(w.bodies()[1])
    out["hello_ccd_bodies"] = np.array(rows, np.float32)
    w.close()

    w = ref.world("pyramid", 20, 0, sleep=True)
    w.step(600)
    out["pyramid_bodies"] = w.bodies()
    out["pyramid_contacts"] = np.array([w.e.fn("world_contact_count")(w.h)], np.int32)
    out["pyramid_penetration"] = np.array(w.penetration(), np.float64)
    w.close()

    w = ref.world("pendulums", 24, 0, sleep=True)
    w.step(200)
    out["pendulum_bodies"] = w.bodies()
    out["pendulum_joints"] = w.joints()
    w.close()

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
