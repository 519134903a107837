"""Post-step measures of a world, in FP64 on the host, one function for every
engine (the device step, the CPU restatement and the reference build all
export the same body rows, contact tables and fixture data).

  penetration   mean and max of max(0, -separation) over the points of
                touching contacts, separations from the world manifold
                (b2WorldManifold::Initialize, Box2D/Collision/b2Collision.cpp:24-87)
  residual      RMS over the points of touching, enabled contacts of the
                complementarity residual of the velocity constraints
                (b2ContactSolver::SolveVelocityConstraints,
                Box2D/Dynamics/Contacts/b2ContactSolver.cpp:280-597):
                r_n = vn where the stored normal impulse is positive, else
                min(vn, 0); r_t = vt where |tangent impulse| < friction *
                normal impulse, else 0. vn and vt are the relative velocity of
                the two bodies at the world-manifold point along the normal and
                the tangent, from the post-step velocities, centres and
                transforms. The restitution bias is not included (it depends on
                the pre-step velocity; the measured scenes have restitution 0).
  kinetic       sum over bodies of m |v|^2 / 2 + I w^2 / 2 (I about the centre)

Inputs are the arrays every engine's *_world_get_bodies, *_world_get_contacts,
*_world_fixture_info and *_world_body_mass write:
  bodies    float32 (n, 12): px py s c cx cy a vx vy w awake sleepTime
  meta      int32 (m, 4): fixture A, fixture B, flags (bit 0 touching, bit 1 enabled)
  man       the 64-byte manifold records (points: lp, ni, ti, id; ln; lpt; type; count)
  fix_body  int32 per fixture proxy: body index
  fix_mat   float32 (proxies, 3): radius, friction, restitution
  mass      float32 (n, 2): mass, rotational inertia about the centre
"""
from __future__ import annotations

import numpy as np

# b2Manifold::Type (Box2D/Collision/b2Collision.h:98-112)
CIRCLES, FACE_A, FACE_B = 0, 1, 2
FLT_EPSILON = np.finfo(np.float32).eps


def _rot(s, c, x, y):
    return c * x - s * y, s * x + c * y


def world_manifold(bodies, meta, man, fix_body, fix_mat):
    """World normal (m, 2), points (m, 2, 2), separations (m, 2) and point
    counts (m,) of every contact, FP64. b2WorldManifold::Initialize
    (Box2D/Collision/b2Collision.cpp:24-87) evaluated on the given transforms."""
    b = bodies.astype(np.float64)
    fa, fb = meta[:, 0], meta[:, 1]
    A, B = fix_body[fa], fix_body[fb]
    rA, rB = fix_mat[fa, 0].astype(np.float64), fix_mat[fb, 0].astype(np.float64)
    pxA, pyA, sA, cA = b[A, 0], b[A, 1], b[A, 2], b[A, 3]
    pxB, pyB, sB, cB = b[B, 0], b[B, 1], b[B, 2], b[B, 3]
    typ = man["type"].astype(np.int64)
    cnt = man["count"].astype(np.int64)
    lnx, lny = man["ln"]["x"].astype(np.float64), man["ln"]["y"].astype(np.float64)
    lpx, lpy = man["lpt"]["x"].astype(np.float64), man["lpt"]["y"].astype(np.float64)
    m = len(meta)
    normal = np.zeros((m, 2))
    points = np.zeros((m, 2, 2))
    sep = np.zeros((m, 2))

    # e_circles: centres of both circles
    ax, ay = _rot(sA, cA, lpx, lpy)
    ax, ay = ax + pxA, ay + pyA
    q0x = man["points"]["lp"]["x"][:, 0].astype(np.float64)
    q0y = man["points"]["lp"]["y"][:, 0].astype(np.float64)
    bx, by = _rot(sB, cB, q0x, q0y)
    bx, by = bx + pxB, by + pyB
    dx, dy = bx - ax, by - ay
    d2 = dx * dx + dy * dy
    far = d2 > FLT_EPSILON * FLT_EPSILON
    dl = np.sqrt(np.where(far, d2, 1.0))
    nx = np.where(far, dx / dl, 1.0)
    ny = np.where(far, dy / dl, 0.0)
    cax, cay = ax + rA * nx, ay + rA * ny
    cbx, cby = bx - rB * nx, by - rB * ny
    circ = typ == CIRCLES
    normal[circ, 0], normal[circ, 1] = nx[circ], ny[circ]
    points[circ, 0, 0] = 0.5 * (cax + cbx)[circ]
    points[circ, 0, 1] = 0.5 * (cay + cby)[circ]
    sep[circ, 0] = ((cbx - cax) * nx + (cby - cay) * ny)[circ]

    # e_faceA / e_faceB: the reference face's plane, points clipped on the other body
    for kind in (FACE_A, FACE_B):
        sel = typ == kind
        if not sel.any():
            continue
        if kind == FACE_A:
            ps, pc, ppx, ppy, qs, qc, qpx, qpy, rr, ri = sA, cA, pxA, pyA, sB, cB, pxB, pyB, rA, rB
        else:
            ps, pc, ppx, ppy, qs, qc, qpx, qpy, rr, ri = sB, cB, pxB, pyB, sA, cA, pxA, pyA, rB, rA
        fnx, fny = _rot(ps, pc, lnx, lny)
        plx, ply = _rot(ps, pc, lpx, lpy)
        plx, ply = plx + ppx, ply + ppy
        for j in range(2):
            lx = man["points"]["lp"]["x"][:, j].astype(np.float64)
            ly = man["points"]["lp"]["y"][:, j].astype(np.float64)
            clx, cly = _rot(qs, qc, lx, ly)
            clx, cly = clx + qpx, cly + qpy
            off = rr - ((clx - plx) * fnx + (cly - ply) * fny)
            c1x, c1y = clx + off * fnx, cly + off * fny   # on the reference face
            c2x, c2y = clx - ri * fnx, cly - ri * fny     # on the incident body
            points[sel, j, 0] = (0.5 * (c1x + c2x))[sel]
            points[sel, j, 1] = (0.5 * (c1y + c2y))[sel]
            sep[sel, j] = ((c2x - c1x) * fnx + (c2y - c1y) * fny)[sel]
        if kind == FACE_A:
            normal[sel, 0], normal[sel, 1] = fnx[sel], fny[sel]
        else:
            normal[sel, 0], normal[sel, 1] = -fnx[sel], -fny[sel]
    return normal, points, sep, cnt


def _point_mask(meta, cnt, enabled_only):
    touching = (meta[:, 2] & 1) != 0
    if enabled_only:
        touching &= (meta[:, 2] & 2) != 0
    return touching[:, None] & (np.arange(2)[None, :] < cnt[:, None])


def penetration(bodies, meta, man, fix_body, fix_mat):
    """(mean, max, points) of max(0, -separation) over touching points."""
    _, _, sep, cnt = world_manifold(bodies, meta, man, fix_body, fix_mat)
    mask = _point_mask(meta, cnt, False)
    p = np.maximum(0.0, -sep[mask])
    return (float(p.mean()) if p.size else 0.0, float(p.max()) if p.size else 0.0, int(p.size))


def residual(bodies, meta, man, fix_body, fix_mat):
    """(rms, points) of the complementarity residual over the points of
    touching, enabled contacts, m/s."""
    normal, points, _, cnt = world_manifold(bodies, meta, man, fix_body, fix_mat)
    b = bodies.astype(np.float64)
    fa, fb = meta[:, 0], meta[:, 1]
    A, B = fix_body[fa], fix_body[fb]
    # b2MixFriction (Box2D/Dynamics/Contacts/b2Contact.h:30-33), FP32 as the solver holds it
    mu = np.sqrt(fix_mat[fa, 1].astype(np.float32) * fix_mat[fb, 1].astype(np.float32)).astype(np.float64)
    nx, ny = normal[:, 0], normal[:, 1]
    tx, ty = ny, -nx
    mask = _point_mask(meta, cnt, True)
    total, npts = 0.0, 0
    for j in range(2):
        sel = mask[:, j]
        if not sel.any():
            continue
        px, py = points[:, j, 0], points[:, j, 1]
        rax, ray = px - b[A, 4], py - b[A, 5]
        rbx, rby = px - b[B, 4], py - b[B, 5]
        dvx = b[B, 7] - b[B, 9] * rby - b[A, 7] + b[A, 9] * ray
        dvy = b[B, 8] + b[B, 9] * rbx - b[A, 8] - b[A, 9] * rax
        vn = dvx * nx + dvy * ny
        vt = dvx * tx + dvy * ty
        ln = man["points"]["ni"][:, j].astype(np.float64)
        lt = man["points"]["ti"][:, j].astype(np.float64)
        rn = np.where(ln > 0.0, vn, np.minimum(vn, 0.0))
        rt = np.where(np.abs(lt) < mu * ln, vt, 0.0)
        total += float(np.sum((rn * rn + rt * rt)[sel]))
        npts += int(sel.sum())
    return (float(np.sqrt(total / npts)) if npts else 0.0, npts)


def kinetic_energy(bodies, mass):
    """Sum of m |v|^2 / 2 + I w^2 / 2, joules."""
    b = bodies.astype(np.float64)
    m = mass.astype(np.float64)
    return float(0.5 * np.sum(m[:, 0] * (b[:, 7] ** 2 + b[:, 8] ** 2) + m[:, 1] * b[:, 9] ** 2))


def all_measures(bodies, meta, man, fix_body, fix_mat, mass):
    """Penetration, residual and kinetic energy of one post-step state."""
    pm, px, pn = penetration(bodies, meta, man, fix_body, fix_mat)
    rr, rn = residual(bodies, meta, man, fix_body, fix_mat)
    return {"mean_penetration_m": pm, "max_penetration_m": px, "touching_points": pn,
            "residual_rms_mps": rr, "residual_points": rn, "kinetic_energy_J": kinetic_energy(bodies, mass)}
