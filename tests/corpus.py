"""Seeded narrowphase corpora: shape pairs with transforms, in contact-registry
order (polygon-polygon, circle-circle, polygon-circle, edge-circle,
edge-polygon), concentrated near contact so every branch of the collide
routines is exercised: separated, touching, shallow and deep overlap,
parallel faces, the reference-face hysteresis boundary, vertex/face
Voronoi regions and edges with and without ghost vertices."""
from __future__ import annotations

import numpy as np

import engines as E


def _rand_polygon(lib, pre, rng, scale=1.0):
    kind = rng.integers(0, 4)
    if kind == 0:
        hx, hy = rng.uniform(0.1, 1.5, 2) * scale
        if rng.random() < 0.5:
            return E.box(lib, pre, hx, hy)
        return E.box(lib, pre, hx, hy, rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-3, 3), True)
    k = int(rng.integers(3, 9))
    R = rng.uniform(0.2, 1.6) * scale
    ph = rng.uniform(0, 2 * np.pi)
    t = ph + 2 * np.pi * np.arange(k) / k + rng.uniform(-0.2, 0.2, k) * (kind == 2)
    pts = np.stack([R * np.cos(t), R * np.sin(t)], 1).astype(np.float32)
    if kind == 3:
        pts += rng.uniform(-0.3, 0.3, 2).astype(np.float32)
    return E.polygon_from_points(lib, pre, pts)


def _extent(s):
    if s["type"] == 0:
        return float(np.hypot(*s["centroid"].tolist()) + s["radius"])
    n = 2 if s["type"] == 1 else int(s["count"])
    v = np.array([[s["verts"][i]["x"], s["verts"][i]["y"]] for i in range(n)])
    return float(np.max(np.hypot(v[:, 0], v[:, 1]))) + float(s["radius"])


def _rand_edge(rng):
    L = rng.uniform(0.5, 6.0)
    a = rng.uniform(0, 2 * np.pi)
    v1 = np.array([-0.5 * L * np.cos(a), -0.5 * L * np.sin(a)])
    v2 = -v1
    v0 = v3 = None
    if rng.random() < 0.6:
        b = a + rng.uniform(-1.2, 1.2)
        v0 = v1 - rng.uniform(0.5, 3) * np.array([np.cos(b), np.sin(b)])
    if rng.random() < 0.6:
        b = a + rng.uniform(-1.2, 1.2)
        v3 = v2 + rng.uniform(0.5, 3) * np.array([np.cos(b), np.sin(b)])
    return E.edge(v1.astype(np.float32), v2.astype(np.float32),
                  None if v0 is None else v0.astype(np.float32), None if v3 is None else v3.astype(np.float32))


def _place(rng, A, B, tight=False):
    xa = E.transform(rng.uniform(-50, 50), rng.uniform(-50, 50), rng.uniform(-6, 6))
    d = (_extent(A) + _extent(B)) * (rng.uniform(0.85, 1.02) if tight else rng.uniform(0.2, 1.15))
    if A["type"] == 1:  # edge: keep the other shape near the segment
        d = _extent(B) * rng.uniform(0.3, 1.2)
        along = rng.uniform(-1.2, 1.2) * _extent(A)
    else:
        along = 0.0
    th = rng.uniform(0, 2 * np.pi)
    ca, sa = float(xa["c"]), float(xa["s"])
    ox, oy = d * np.cos(th), d * np.sin(th)
    if A["type"] == 1:
        ex = float(A["verts"][1]["x"] - A["verts"][0]["x"])
        ey = float(A["verts"][1]["y"] - A["verts"][0]["y"])
        L = np.hypot(ex, ey)
        ex, ey = ex / L, ey / L
        ox, oy = along * ex + d * (-ey) * np.sign(np.sin(th) or 1.0), along * ey + d * ex * np.sign(np.sin(th) or 1.0)
    wx, wy = ca * ox - sa * oy, sa * ox + ca * oy
    xb = E.transform(float(xa["px"]) + wx, float(xa["py"]) + wy, rng.uniform(-6, 6))
    return xa, xb


def random_corpus(lib, pre: str, n_per_type: int, seed: int):
    """Arrays (A, xfA, B, xfB) of 5 * n_per_type cases."""
    rng = np.random.default_rng(seed)
    As, Bs, XA, XB = [], [], [], []
    for t in range(5):
        for i in range(n_per_type):
            if t == 0:
                A, B = _rand_polygon(lib, pre, rng), _rand_polygon(lib, pre, rng)
            elif t == 1:
                A, B = E.circle(rng.uniform(0.1, 1.5)), E.circle(rng.uniform(0.1, 1.5))
            elif t == 2:
                A, B = _rand_polygon(lib, pre, rng), E.circle(rng.uniform(0.1, 1.5))
            elif t == 3:
                A, B = _rand_edge(rng), E.circle(rng.uniform(0.1, 1.5))
            else:
                A, B = _rand_edge(rng), _rand_polygon(lib, pre, rng)
            xa, xb = _place(rng, A, B, tight=(i % 3 == 0))
            As.append(A); Bs.append(B); XA.append(xa); XB.append(xb)
    return (np.array(As, E.SHAPE), np.array(XA, E.XF), np.array(Bs, E.SHAPE), np.array(XB, E.XF))


def special_corpus(lib, pre: str):
    """Hand-placed configurations at branch boundaries."""
    As, Bs, XA, XB = [], [], [], []

    def add(A, xa, B, xb):
        As.append(A); XA.append(xa); Bs.append(B); XB.append(xb)

    box1 = E.box(lib, pre, 0.5, 0.5)
    I = E.transform(0.0, 0.0, 0.0)
    # box resting exactly on box, aligned faces (two-point clip, equal separations)
    for dy in (1.0, 0.99, 1.01, 1.0 + 0.02, 0.5):
        add(box1, I, box1, E.transform(0.0, dy, 0.0))
        add(box1, I, box1, E.transform(0.3, dy, 0.0))
    # hysteresis boundary: separations differ by about 0.1 * linearSlop
    for a in (0.0, 1e-4, 5e-4, 1e-3):
        add(box1, I, E.box(lib, pre, 0.5, 0.5005), E.transform(1.0 - 0.005, 0.0, a))
    # rotated 45 degrees: vertex-face
    add(box1, I, box1, E.transform(0.0, 0.5 + 0.5 * np.sqrt(2.0) - 0.01, np.pi / 4))
    # circles: coincident centres, touching, separated
    c1 = E.circle(1.0)
    add(c1, I, c1, I)
    add(c1, I, c1, E.transform(2.0, 0.0, 0.3))
    add(c1, I, c1, E.transform(2.0001, 0.0, 0.3))
    # polygon-circle: centre inside, vertex regions, face region
    for (x, y) in ((0.0, 0.0), (0.9, 0.9), (-0.9, 0.6), (0.0, 1.2), (0.5, 0.5), (1.5, 0.0)):
        add(box1, I, E.circle(0.5), E.transform(x, y, 0.0))
    # edge-circle: vertex A, vertex B, interior, with ghosts blocking
    e = E.edge((-1.0, 0.0), (1.0, 0.0))
    eg = E.edge((-1.0, 0.0), (1.0, 0.0), (-2.0, 0.0), (2.0, 0.0))
    for (x, y) in ((-1.2, 0.3), (1.2, 0.3), (0.0, 0.4), (0.0, -0.4), (-1.0, 0.5)):
        add(e, I, E.circle(0.5), E.transform(x, y, 0.0))
        add(eg, I, E.circle(0.5), E.transform(x, y, 0.0))
    # edge-polygon: flat, tilted, convex and concave neighbours, back side
    convex = E.edge((-1.0, 0.0), (1.0, 0.0), (-2.0, -0.5), (2.0, -0.5))
    concave = E.edge((-1.0, 0.0), (1.0, 0.0), (-2.0, 0.5), (2.0, 0.5))
    for ed in (e, eg, convex, concave):
        for (x, y, a) in ((0.0, 0.5, 0.0), (0.9, 0.45, 0.2), (-1.1, 0.4, 0.7), (0.0, -0.5, 0.0), (1.4, 0.3, 0.0)):
            add(ed, I, box1, E.transform(x, y, a))
    return (np.array(As, E.SHAPE), np.array(XA, E.XF), np.array(Bs, E.SHAPE), np.array(XB, E.XF))


def collide_all(lib, pre: str, A, XA, B, XB) -> np.ndarray:
    out = np.zeros(len(A), E.MANIFOLD)
    f = getattr(lib, pre + "collide")
    for i in range(len(A)):
        a, b = A[i:i + 1], B[i:i + 1]
        xa, xb = XA[i:i + 1], XB[i:i + 1]
        o = np.zeros(1, E.MANIFOLD)
        f(E.ptr(a), E.ptr(xa), E.ptr(b), E.ptr(xb), E.ptr(o))
        out[i] = o[0]
    return out


def sweep_corpus(lib, pre: str, n_per_type: int, seed: int):
    """Arrays (A, sweepA, B, sweepB) of 5 * n_per_type swept pairs for the
    time of impact: B (and A unless it is an edge) moves up to several
    extents per step with rotations up to 6 rad, towards, past or away from
    the other shape, some sweeps starting at alpha0 > 0."""
    rng = np.random.default_rng(seed)
    As, Bs, SA, SB = [], [], [], []

    def sweep(c0, c, a0, a, lc, alpha0):
        sw = np.zeros((), E.SWEEP)
        sw["localCenter"] = lc
        sw["c0"] = c0
        sw["c"] = c
        sw["a0"], sw["a"], sw["alpha0"] = a0, a, alpha0
        return sw

    for t in range(5):
        for i in range(n_per_type):
            if t == 0:
                A, B = _rand_polygon(lib, pre, rng), _rand_polygon(lib, pre, rng)
            elif t == 1:
                A, B = E.circle(rng.uniform(0.05, 1.0)), E.circle(rng.uniform(0.05, 1.0))
            elif t == 2:
                A, B = _rand_polygon(lib, pre, rng), E.circle(rng.uniform(0.05, 1.0))
            elif t == 3:
                A, B = _rand_edge(rng), E.circle(rng.uniform(0.05, 1.0))
            else:
                A, B = _rand_edge(rng), _rand_polygon(lib, pre, rng)
            reach = _extent(A) + _extent(B)
            pa = rng.uniform(-20, 20, 2)
            aa = rng.uniform(-4, 4)
            if t >= 3:
                sa = sweep(pa, pa, aa, aa, (0.0, 0.0), 0.0)
            else:
                lc = rng.uniform(-0.2, 0.2, 2) * (rng.random() < 0.5)
                sa = sweep(pa, pa + rng.uniform(-1, 1, 2) * reach, aa, aa + rng.uniform(-3, 3), lc, 0.0)
            th = rng.uniform(0, 2 * np.pi)
            start = pa + reach * rng.uniform(1.05, 4.0) * np.array([np.cos(th), np.sin(th)])
            aim = pa + rng.uniform(-0.8, 0.8, 2) * reach
            end = start + (aim - start) * rng.uniform(0.5, 3.0)
            if rng.random() < 0.15:
                end = start + (start - pa) * 0.5  # moving away
            ab = rng.uniform(-4, 4)
            alpha0 = float(rng.uniform(0.0, 0.6)) if rng.random() < 0.2 else 0.0
            sb = sweep(start, end, ab, ab + rng.uniform(-6, 6) * (rng.random() < 0.7), rng.uniform(-0.2, 0.2, 2), alpha0)
            As.append(A); Bs.append(B); SA.append(sa); SB.append(sb)
    return (np.array(As, E.SHAPE), np.array(SA, E.SWEEP), np.array(Bs, E.SHAPE), np.array(SB, E.SWEEP))
