"""b2World::QueryAABB and b2World::RayCast (Box2D/Dynamics/b2World.h:115-123)
through the box2d:: host API, against the reference build on the same
freshly loaded scenes (identical fat boxes: the shapes' AABBs at the load
transforms plus the AABB extension):

  QueryAABB   the reported fixtures, as a sorted list (the reference reports in
              its tree's order, the host in proxy order), equal
  RayCast     the closest hit (the callback clips at each hit's fraction): the
              same fixture, fraction, normal and point within 1e-5

Scenes: the batched-world pile (circles and convex polygons in an edge
container) and the terrain (chain shapes: every child a proxy).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu
F = ctypes.c_float


def _query(w, box):
    out = np.zeros(4096, np.int32)
    b = np.asarray(box, np.float32)
    n = w.e.fn("world_query_aabb")(w.h, E.ptr(b), E.ptr(out), len(out))
    return np.sort(out[:n])


def _ray(w, p1, p2):
    out = np.zeros(5, np.float32)
    hit = w.e.fn("world_raycast_closest")(w.h, F(p1[0]), F(p1[1]), F(p2[0]), F(p2[1]), E.ptr(out))
    return hit, out


@pytest.mark.parametrize("kind,n,seed", [("world", 150, 5), ("terrain", 60, 1)])
def test_queries_match_reference_build(gpu, ref, kind, n, seed):
    g = gpu.world(kind, n, seed, sleep=True)
    r = ref.world(kind, n, seed, sleep=True)
    print(f"case: queries on {kind} n={n}", flush=True)
    fat = r.fat_aabbs()
    names = fat.dtype.names
    lo = np.array([fat[names[0]].min(), fat[names[1]].min()])
    hi = np.array([fat[names[2]].max(), fat[names[3]].max()])
    rng = np.random.default_rng(seed)
    reports = 0
    for _ in range(200):
        c = lo + (hi - lo) * rng.random(2)
        half = rng.uniform(0.05, 4.0, 2)
        box = [c[0] - half[0], c[1] - half[1], c[0] + half[0], c[1] + half[1]]
        got, want = _query(g, box), _query(r, box)
        assert np.array_equal(got, want), (box, got, want)
        reports += len(want)
    assert reports > 200
    hits = 0
    for _ in range(200):
        p1 = lo + (hi - lo) * rng.random(2)
        p2 = lo + (hi - lo) * rng.random(2)
        gh, go = _ray(g, p1, p2)
        rh, ro = _ray(r, p1, p2)
        assert gh == rh, (p1, p2, gh, rh)
        if rh >= 0:
            assert np.abs(go - ro).max() <= 1e-5, (p1, p2, go, ro)
            hits += 1
    assert hits > 50
    g.close(); r.close()
