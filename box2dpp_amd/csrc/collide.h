// Narrowphase on gfx950: contact manifolds for the five shape pairs the
// step needs, one call per contact, FP32 in the reference's operation order.
//
//   polygon-polygon   b2CollidePolygons        REF/Collision/b2CollidePolygon.cpp:121-247
//   circle-circle     b2CollideCircles         REF/Collision/b2CollideCircle.cpp:25-50
//   polygon-circle    b2CollidePolygonAndCircle REF/Collision/b2CollideCircle.cpp:53-157
//   edge-circle       b2CollideEdgeAndCircle   REF/Collision/b2CollideEdge.cpp:27-151
//   edge-polygon      b2EPCollider::Collide    REF/Collision/b2CollideEdge.cpp:153-700
//
// The edge-polygon routine follows upstream, not the fork's copy, which
// rotates vertices into normals (Box2D/Collision/b2CollideEdge.cpp:443) and
// takes a normal for a local point (:597).
//
// Shapes arrive as b2g_shape records (verts/normals in body frame). A
// polygon's arrays are read through `const b2g_shape&`; callers may pass a
// record in global memory, LDS or registers.
#pragma once
#include "b2g_device.h"

namespace b2g {

struct ClipV { V2 v; uint32_t id; };

// b2ClipSegmentToLine, REF/Collision/b2Collision.cpp:201-232 (Sutherland-Hodgman).
__device__ inline int clipSegment(ClipV out[2], const ClipV in[2], V2 normal, float offset, int vertexIndexA) {
    int n = 0;
    const float d0 = dot(normal, in[0].v) - offset;
    const float d1 = dot(normal, in[1].v) - offset;
    if (d0 <= 0.0f) out[n++] = in[0];
    if (d1 <= 0.0f) out[n++] = in[1];
    if (d0 * d1 < 0.0f) {
        const float t = d0 / (d0 - d1);
        out[n].v = in[0].v + t * (in[1].v - in[0].v);
        out[n].id = makeId(vertexIndexA, (int)((in[0].id >> 8) & 0xFF), kFeatVertex, kFeatFace);
        ++n;
    }
    return n;
}

// Deepest separation of poly2 along poly1's face normals, in poly2's frame
// (b2FindMaxSeparation, REF/Collision/b2CollidePolygon.cpp:28-66).
__device__ inline float maxSeparation(int* edge, const b2g_shape& p1, const Xf& xf1, const b2g_shape& p2,
                                      const Xf& xf2) {
    const Xf xf = mulT(xf2, xf1);
    int best = 0;
    float maxSep = -FLT_MAX;
    const int n1 = p1.count, n2 = p2.count;
    for (int i = 0; i < n1; ++i) {
        const V2 n = mul(xf.q, ld(p1.normals[i]));
        const V2 v1 = mul(xf, ld(p1.verts[i]));
        float si = FLT_MAX;
        for (int j = 0; j < n2; ++j) {
            const float sij = dot(n, ld(p2.verts[j]) - v1);
            if (sij < si) si = sij;
        }
        if (si > maxSep) { maxSep = si; best = i; }
    }
    *edge = best;
    return maxSep;
}

__device__ inline void collidePolygonsFaces(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                            const Xf& xfB, float sepA, int edgeA, float sepB, int edgeB);

// b2CollidePolygons with b2FindIncidentEdge inlined.
__device__ inline void collidePolygons(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                       const Xf& xfB) {
    m.pointCount = 0;
    const float totalRadius = A.radius + B.radius;
    int edgeA = 0;
    const float sepA = maxSeparation(&edgeA, A, xfA, B, xfB);
    if (sepA > totalRadius) return;
    int edgeB = 0;
    const float sepB = maxSeparation(&edgeB, B, xfB, A, xfA);
    if (sepB > totalRadius) return;
    collidePolygonsFaces(m, A, xfA, B, xfB, sepA, edgeA, sepB, edgeB);
}

// The rest of b2CollidePolygons once both separations are known: reference
// face, incident edge (b2FindIncidentEdge) and the two clips.
__device__ inline void collidePolygonsFaces(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                            const Xf& xfB, float sepA, int edgeA, float sepB, int edgeB) {
    const float totalRadius = A.radius + B.radius;
    const bool flip = sepB > sepA + 0.1f * kLinearSlop;
    const b2g_shape& p1 = flip ? B : A;
    const b2g_shape& p2 = flip ? A : B;
    const Xf xf1 = flip ? xfB : xfA;
    const Xf xf2 = flip ? xfA : xfB;
    const int edge1 = flip ? edgeB : edgeA;
    m.type = flip ? B2G_MANIFOLD_FACE_B : B2G_MANIFOLD_FACE_A;

    // Incident edge: poly2 face most anti-parallel to the reference normal.
    ClipV inc[2];
    {
        const V2 n1 = mulT(xf2.q, mul(xf1.q, ld(p1.normals[edge1])));
        int index = 0;
        float minDot = FLT_MAX;
        for (int i = 0; i < p2.count; ++i) {
            const float d = dot(n1, ld(p2.normals[i]));
            if (d < minDot) { minDot = d; index = i; }
        }
        const int i1 = index, i2 = i1 + 1 < p2.count ? i1 + 1 : 0;
        inc[0].v = mul(xf2, ld(p2.verts[i1]));
        inc[0].id = makeId(edge1, i1, kFeatFace, kFeatVertex);
        inc[1].v = mul(xf2, ld(p2.verts[i2]));
        inc[1].id = makeId(edge1, i2, kFeatFace, kFeatVertex);
    }
    const int iv1 = edge1, iv2 = edge1 + 1 < p1.count ? edge1 + 1 : 0;
    V2 v11 = ld(p1.verts[iv1]), v12 = ld(p1.verts[iv2]);
    V2 localTangent = v12 - v11;
    normalizeV(localTangent);
    const V2 localNormal = cross(localTangent, 1.0f);
    const V2 planePoint = 0.5f * (v11 + v12);
    const V2 tangent = mul(xf1.q, localTangent);
    const V2 normal = cross(tangent, 1.0f);
    v11 = mul(xf1, v11);
    v12 = mul(xf1, v12);
    const float frontOffset = dot(normal, v11);
    const float sideOffset1 = -dot(tangent, v11) + totalRadius;
    const float sideOffset2 = dot(tangent, v12) + totalRadius;
    ClipV c1[2], c2[2];
    if (clipSegment(c1, inc, -tangent, sideOffset1, iv1) < 2) return;
    if (clipSegment(c2, c1, tangent, sideOffset2, iv2) < 2) return;
    m.localNormal = st(localNormal);
    m.localPoint = st(planePoint);
    int pc = 0;
    for (int i = 0; i < 2; ++i) {
        const float separation = dot(normal, c2[i].v) - frontOffset;
        if (separation <= totalRadius) {
            m.points[pc].localPoint = st(mulT(xf2, c2[i].v));
            m.points[pc].id = flip ? swapId(c2[i].id) : c2[i].id;
            ++pc;
        }
    }
    m.pointCount = pc;
}

// ---------------------------------------------------------------- tiles
// b2CollidePolygons by a group of kTileLanes lanes of one wavefront (the
// group's lanes are consecutive and aligned). Every lane of the group calls
// it with the same shapes (staged in LDS) and transforms; lane i of the
// group evaluates face normal i of each b2FindMaxSeparation, and the group
// reduces to the deepest separation, the lowest face index winning a tie,
// which is the face the reference's sequential scan keeps (it replaces the
// running maximum only on a strictly greater value). The rest of the
// routine (incident edge, clipping) is evaluated identically on every lane,
// so all lanes end with the same manifold.
constexpr int kTileLanes = 8;  // polygons have at most 8 vertices

__device__ inline float tileMaxSeparation(int* edge, const b2g_shape& p1, const Xf& xf1, const b2g_shape& p2,
                                          const Xf& xf2, int lane) {
    const Xf xf = mulT(xf2, xf1);
    // this lane's face: the minimum over poly2's vertices, as the scan forms it
    float si = -FLT_MAX;
    int index = kTileLanes;  // no face
    if (lane < p1.count) {
        const V2 n = mul(xf.q, ld(p1.normals[lane]));
        const V2 v1 = mul(xf, ld(p1.verts[lane]));
        float s = FLT_MAX;
        for (int j = 0; j < p2.count; ++j) {
            const float sij = dot(n, ld(p2.verts[j]) - v1);
            if (sij < s) s = sij;
        }
        // faces at or below the scan's starting value (or NaN) never win
        if (s > -FLT_MAX) {
            si = s;
            index = lane;
        }
    }
    for (int mask = 1; mask < kTileLanes; mask <<= 1) {
        const float os = __shfl_xor(si, mask, kTileLanes);
        const int oi = __shfl_xor(index, mask, kTileLanes);
        if (os > si || (os == si && oi < index)) {
            si = os;
            index = oi;
        }
    }
    *edge = index < kTileLanes ? index : 0;
    return si;
}

__device__ inline void collidePolygonsTile(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                           const Xf& xfB, int lane) {
    m.pointCount = 0;
    const float totalRadius = A.radius + B.radius;
    int edgeA = 0;
    const float sepA = tileMaxSeparation(&edgeA, A, xfA, B, xfB, lane);
    if (sepA > totalRadius) return;
    int edgeB = 0;
    const float sepB = tileMaxSeparation(&edgeB, B, xfB, A, xfA, lane);
    if (sepB > totalRadius) return;
    collidePolygonsFaces(m, A, xfA, B, xfB, sepA, edgeA, sepB, edgeB);
}

__device__ inline void collideCircles(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                      const Xf& xfB) {
    m.pointCount = 0;
    const V2 pA = mul(xfA, ld(A.centroid));
    const V2 pB = mul(xfB, ld(B.centroid));
    const V2 d = pB - pA;
    const float distSqr = dot(d, d);
    const float r = A.radius + B.radius;
    if (distSqr > r * r) return;
    m.type = B2G_MANIFOLD_CIRCLES;
    m.localPoint = A.centroid;
    m.localNormal = st(mk(0.0f, 0.0f));
    m.pointCount = 1;
    m.points[0].localPoint = B.centroid;
    m.points[0].id = 0;
}

__device__ inline void onePoint(b2g_manifold& m, V2 normal, V2 point, const b2g_shape& circle) {
    m.pointCount = 1;
    m.type = B2G_MANIFOLD_FACE_A;
    m.localNormal = st(normal);
    m.localPoint = st(point);
    m.points[0].localPoint = circle.centroid;
    m.points[0].id = 0;
}

__device__ inline void collidePolygonCircle(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                            const Xf& xfB) {
    m.pointCount = 0;
    const V2 cLocal = mulT(xfA, mul(xfB, ld(B.centroid)));
    int normalIndex = 0;
    float separation = -FLT_MAX;
    const float radius = A.radius + B.radius;
    const int n = A.count;
    for (int i = 0; i < n; ++i) {
        const float s = dot(ld(A.normals[i]), cLocal - ld(A.verts[i]));
        if (s > radius) return;
        if (s > separation) { separation = s; normalIndex = i; }
    }
    const int vi1 = normalIndex, vi2 = vi1 + 1 < n ? vi1 + 1 : 0;
    const V2 v1 = ld(A.verts[vi1]), v2 = ld(A.verts[vi2]);
    if (separation < FLT_EPSILON) {
        onePoint(m, ld(A.normals[normalIndex]), 0.5f * (v1 + v2), B);
        return;
    }
    const float u1 = dot(cLocal - v1, v2 - v1);
    const float u2 = dot(cLocal - v2, v1 - v2);
    if (u1 <= 0.0f) {
        if (distSq(cLocal, v1) > radius * radius) return;
        V2 nrm = cLocal - v1;
        normalizeV(nrm);
        onePoint(m, nrm, v1, B);
    } else if (u2 <= 0.0f) {
        if (distSq(cLocal, v2) > radius * radius) return;
        V2 nrm = cLocal - v2;
        normalizeV(nrm);
        onePoint(m, nrm, v2, B);
    } else {
        const V2 faceCenter = 0.5f * (v1 + v2);
        const float sep = dot(cLocal - faceCenter, ld(A.normals[vi1]));
        if (sep > radius) return;
        onePoint(m, ld(A.normals[vi1]), faceCenter, B);
    }
}

__device__ inline void collideEdgeCircle(b2g_manifold& m, const b2g_shape& E, const Xf& xfA, const b2g_shape& C,
                                         const Xf& xfB) {
    m.pointCount = 0;
    const V2 Q = mulT(xfA, mul(xfB, ld(C.centroid)));
    const V2 A = ld(E.verts[0]), B = ld(E.verts[1]);
    const V2 e = B - A;
    const float u = dot(e, B - Q);
    const float w = dot(e, Q - A);
    const float radius = E.radius + C.radius;
    int region;  // 0: vertex A, 1: vertex B, 2: interior
    V2 P;
    if (w <= 0.0f) {
        region = 0;
        P = A;
    } else if (u <= 0.0f) {
        region = 1;
        P = B;
    } else {
        region = 2;
        const float den = dot(e, e);
        P = (1.0f / den) * (u * A + w * B);
    }
    const V2 d = Q - P;
    if (dot(d, d) > radius * radius) return;
    if (region == 0 && (E.flags & 1)) {
        const V2 e1 = A - ld(E.v0);
        if (dot(e1, A - Q) > 0.0f) return;  // belongs to the previous edge
    }
    if (region == 1 && (E.flags & 2)) {
        const V2 e2 = ld(E.v3) - B;
        if (dot(e2, Q - B) > 0.0f) return;  // belongs to the next edge
    }
    m.pointCount = 1;
    m.points[0].localPoint = C.centroid;
    if (region < 2) {
        m.type = B2G_MANIFOLD_CIRCLES;
        m.localNormal = st(mk(0.0f, 0.0f));
        m.localPoint = st(P);
        m.points[0].id = makeId(region, 0, kFeatVertex, kFeatVertex);
    } else {
        V2 n = mk(-e.y, e.x);
        if (dot(n, Q - A) < 0.0f) n = mk(-n.x, -n.y);
        normalizeV(n);
        m.type = B2G_MANIFOLD_FACE_A;
        m.localNormal = st(n);
        m.localPoint = st(A);
        m.points[0].id = makeId(0, 0, kFeatFace, kFeatVertex);
    }
}

// Edge (A) against polygon (B) with the edge's neighbour vertices limiting
// the admissible normals (b2EPCollider).
__device__ inline void collideEdgePolygon(b2g_manifold& m, const b2g_shape& E, const Xf& xfA, const b2g_shape& P,
                                          const Xf& xfB) {
    const Xf xf = mulT(xfA, xfB);
    const V2 centroidB = mul(xf, ld(P.centroid));
    const V2 v0 = ld(E.v0), v1 = ld(E.verts[0]), v2 = ld(E.verts[1]), v3 = ld(E.v3);
    const bool has0 = (E.flags & 1) != 0, has3 = (E.flags & 2) != 0;
    V2 edge1 = v2 - v1;
    normalizeV(edge1);
    const V2 n1 = mk(edge1.y, -edge1.x);
    V2 n0 = mk(0.0f, 0.0f), n2 = mk(0.0f, 0.0f);
    const float off1 = dot(n1, centroidB - v1);
    float off0 = 0.0f, off2 = 0.0f;
    bool convex1 = false, convex2 = false;
    if (has0) {
        V2 edge0 = v1 - v0;
        normalizeV(edge0);
        n0 = mk(edge0.y, -edge0.x);
        convex1 = cross(edge0, edge1) >= 0.0f;
        off0 = dot(n0, centroidB - v0);
    }
    if (has3) {
        V2 edge2 = v3 - v2;
        normalizeV(edge2);
        n2 = mk(edge2.y, -edge2.x);
        convex2 = cross(edge1, edge2) > 0.0f;
        off2 = dot(n2, centroidB - v2);
    }
    // Front/back classification and admissible normal range
    // (the decision table of REF :296-447).
    bool front;
    V2 normal, lower, upper;
    if (has0 && has3) {
        if (convex1 && convex2) {
            front = off0 >= 0.0f || off1 >= 0.0f || off2 >= 0.0f;
            lower = front ? n0 : -n1; upper = front ? n2 : -n1;
        } else if (convex1) {
            front = off0 >= 0.0f || (off1 >= 0.0f && off2 >= 0.0f);
            lower = front ? n0 : -n2; upper = front ? n1 : -n1;
        } else if (convex2) {
            front = off2 >= 0.0f || (off0 >= 0.0f && off1 >= 0.0f);
            lower = front ? n1 : -n1; upper = front ? n2 : -n0;
        } else {
            front = off0 >= 0.0f && off1 >= 0.0f && off2 >= 0.0f;
            lower = front ? n1 : -n2; upper = front ? n1 : -n0;
        }
    } else if (has0) {
        if (convex1) {
            front = off0 >= 0.0f || off1 >= 0.0f;
            lower = front ? n0 : n1; upper = -n1;
        } else {
            front = off0 >= 0.0f && off1 >= 0.0f;
            lower = n1; upper = front ? -n1 : -n0;
        }
    } else if (has3) {
        if (convex2) {
            front = off1 >= 0.0f || off2 >= 0.0f;
            lower = -n1; upper = front ? n2 : n1;
        } else {
            front = off1 >= 0.0f && off2 >= 0.0f;
            lower = front ? -n1 : -n2; upper = n1;
        }
    } else {
        front = off1 >= 0.0f;
        lower = front ? -n1 : n1; upper = lower;
    }
    normal = front ? n1 : -n1;

    const int count = P.count;
    V2 pv[8], pn[8];
    for (int i = 0; i < count; ++i) {
        pv[i] = mul(xf, ld(P.verts[i]));
        pn[i] = mul(xf.q, ld(P.normals[i]));
    }
    const float radius = 2.0f * kPolygonRadius;
    m.pointCount = 0;

    // Edge axis: deepest polygon vertex along the edge normal.
    float edgeSep = FLT_MAX;
    for (int i = 0; i < count; ++i) {
        const float s = dot(normal, pv[i] - v1);
        if (s < edgeSep) edgeSep = s;
    }
    if (edgeSep > radius) return;

    // Polygon axes inside the admissible range.
    int polyType = 0, polyIndex = -1;  // 0 unknown, 2 edgeB
    float polySep = -FLT_MAX;
    {
        const V2 perp = mk(-normal.y, normal.x);
        for (int i = 0; i < count; ++i) {
            const V2 n = -pn[i];
            const float s1 = dot(n, pv[i] - v1);
            const float s2 = dot(n, pv[i] - v2);
            const float s = fminB(s1, s2);
            if (s > radius) { polyType = 2; polyIndex = i; polySep = s; break; }
            if (dot(n, perp) >= 0.0f) {
                if (dot(n - upper, normal) < -kAngularSlop) continue;
            } else {
                if (dot(n - lower, normal) < -kAngularSlop) continue;
            }
            if (s > polySep) { polyType = 2; polyIndex = i; polySep = s; }
        }
    }
    if (polyType != 0 && polySep > radius) return;
    const bool usePoly = polyType != 0 && polySep > 0.98f * edgeSep + 0.001f;

    ClipV ie[2];
    int rfi1, rfi2;
    V2 rfv1, rfv2, rfn;
    if (!usePoly) {
        m.type = B2G_MANIFOLD_FACE_A;
        int best = 0;
        float bestValue = dot(normal, pn[0]);
        for (int i = 1; i < count; ++i) {
            const float value = dot(normal, pn[i]);
            if (value < bestValue) { bestValue = value; best = i; }
        }
        const int i1 = best, i2 = i1 + 1 < count ? i1 + 1 : 0;
        ie[0].v = pv[i1]; ie[0].id = makeId(0, i1, kFeatFace, kFeatVertex);
        ie[1].v = pv[i2]; ie[1].id = makeId(0, i2, kFeatFace, kFeatVertex);
        rfi1 = front ? 0 : 1;
        rfi2 = front ? 1 : 0;
        rfv1 = front ? v1 : v2;
        rfv2 = front ? v2 : v1;
        rfn = front ? n1 : -n1;
    } else {
        m.type = B2G_MANIFOLD_FACE_B;
        ie[0].v = v1; ie[0].id = makeId(0, polyIndex, kFeatVertex, kFeatFace);
        ie[1].v = v2; ie[1].id = makeId(0, polyIndex, kFeatVertex, kFeatFace);
        rfi1 = polyIndex;
        rfi2 = rfi1 + 1 < count ? rfi1 + 1 : 0;
        rfv1 = pv[rfi1]; rfv2 = pv[rfi2]; rfn = pn[rfi1];
    }
    const V2 side1 = mk(rfn.y, -rfn.x);
    const V2 side2 = -side1;
    ClipV c1[2], c2[2];
    if (clipSegment(c1, ie, side1, dot(side1, rfv1), rfi1) < 2) return;
    if (clipSegment(c2, c1, side2, dot(side2, rfv2), rfi2) < 2) return;
    if (!usePoly) {
        m.localNormal = st(rfn);
        m.localPoint = st(rfv1);
    } else {
        m.localNormal = P.normals[rfi1];
        m.localPoint = P.verts[rfi1];
    }
    int pc = 0;
    for (int i = 0; i < 2; ++i) {
        const float separation = dot(rfn, c2[i].v - rfv1);
        if (separation <= radius) {
            if (!usePoly) {
                m.points[pc].localPoint = st(mulT(xf, c2[i].v));
                m.points[pc].id = c2[i].id;
            } else {
                m.points[pc].localPoint = st(c2[i].v);
                m.points[pc].id = swapId(c2[i].id);
            }
            ++pc;
        }
    }
    m.pointCount = pc;
}

// Contact-type dispatch for shapes already in registry order
// (REF/Dynamics/Contacts/b2Contact.cpp:34-43). Returns false for pairs
// without a contact type (edge-edge).
__device__ inline bool collideShapes(b2g_manifold& m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                                     const Xf& xfB) {
    const int ta = A.type, tb = B.type;
    if (ta == B2G_POLYGON && tb == B2G_POLYGON) collidePolygons(m, A, xfA, B, xfB);
    else if (ta == B2G_CIRCLE && tb == B2G_CIRCLE) collideCircles(m, A, xfA, B, xfB);
    else if (ta == B2G_POLYGON && tb == B2G_CIRCLE) collidePolygonCircle(m, A, xfA, B, xfB);
    else if (ta == B2G_EDGE && tb == B2G_CIRCLE) collideEdgeCircle(m, A, xfA, B, xfB);
    else if (ta == B2G_EDGE && tb == B2G_POLYGON) collideEdgePolygon(m, A, xfA, B, xfB);
    else { m.pointCount = 0; return false; }
    return true;
}

// True when the contact registry has a creator for (ta, tb) as given
// (b2Contact::s_registers primary entries).
__host__ __device__ inline bool primaryPair(int ta, int tb) {
    return (ta == B2G_CIRCLE && tb == B2G_CIRCLE) || (ta == B2G_POLYGON && tb == B2G_CIRCLE) ||
           (ta == B2G_POLYGON && tb == B2G_POLYGON) || (ta == B2G_EDGE && tb == B2G_CIRCLE) ||
           (ta == B2G_EDGE && tb == B2G_POLYGON);
}

// b2WorldManifold::Initialize, REF/Collision/b2Collision.cpp:22-86.
struct WorldManifold { V2 normal; V2 points[2]; float separations[2]; };
__device__ inline void worldManifold(WorldManifold& wm, const b2g_manifold& m, const Xf& xfA, float radiusA,
                                     const Xf& xfB, float radiusB) {
    if (m.pointCount == 0) return;
    if (m.type == B2G_MANIFOLD_CIRCLES) {
        wm.normal = mk(1.0f, 0.0f);
        const V2 pointA = mul(xfA, ld(m.localPoint));
        const V2 pointB = mul(xfB, ld(m.points[0].localPoint));
        if (distSq(pointA, pointB) > FLT_EPSILON * FLT_EPSILON) {
            wm.normal = pointB - pointA;
            normalizeV(wm.normal);
        }
        const V2 cA = pointA + radiusA * wm.normal;
        const V2 cB = pointB - radiusB * wm.normal;
        wm.points[0] = 0.5f * (cA + cB);
        wm.separations[0] = dot(cB - cA, wm.normal);
    } else if (m.type == B2G_MANIFOLD_FACE_A) {
        wm.normal = mul(xfA.q, ld(m.localNormal));
        const V2 planePoint = mul(xfA, ld(m.localPoint));
        for (int i = 0; i < m.pointCount; ++i) {
            const V2 clipPoint = mul(xfB, ld(m.points[i].localPoint));
            const V2 cA = clipPoint + (radiusA - dot(clipPoint - planePoint, wm.normal)) * wm.normal;
            const V2 cB = clipPoint - radiusB * wm.normal;
            wm.points[i] = 0.5f * (cA + cB);
            wm.separations[i] = dot(cB - cA, wm.normal);
        }
    } else {
        wm.normal = mul(xfB.q, ld(m.localNormal));
        const V2 planePoint = mul(xfB, ld(m.localPoint));
        for (int i = 0; i < m.pointCount; ++i) {
            const V2 clipPoint = mul(xfA, ld(m.points[i].localPoint));
            const V2 cB = clipPoint + (radiusB - dot(clipPoint - planePoint, wm.normal)) * wm.normal;
            const V2 cA = clipPoint - radiusA * wm.normal;
            wm.points[i] = 0.5f * (cA + cB);
            wm.separations[i] = dot(cA - cB, wm.normal);
        }
        wm.normal = -wm.normal;
    }
}

}  // namespace b2g
