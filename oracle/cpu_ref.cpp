// ORACLE TEST INFRASTRUCTURE - not part of the product. See cpu_ref.h.
//
// Each function names the upstream file:line it restates (REF =
// Regression/box2d-ref/Box2DRef). Arithmetic is written out in the same
// order as upstream so the results are bit-identical; the data structures
// are this file's own (index-linked lists instead of pointers, one flat
// pool per entity).
#include "cpu_ref.h"

#include <math.h>
#include <float.h>
#include <string.h>
#include <vector>
#include <algorithm>

#include "scenes/scenes.h"

namespace orc {

// ------------------------------------------------------------ constants
// REF/Common/b2Settings.h:40-150
constexpr float kPi = 3.14159265359f;
constexpr float kLinearSlop = 0.005f;
constexpr float kToiBaumgarte = 0.75f;
constexpr int kMaxSubSteps = 8;
constexpr int kMaxTOIContacts = 32;
constexpr float kAngularSlop = 2.0f / 180.0f * kPi;
constexpr float kMaxAngularCorrection = 8.0f / 180.0f * kPi;
constexpr float kPolygonRadius = 2.0f * kLinearSlop;
constexpr float kAabbExtension = 0.1f;
constexpr float kAabbMultiplier = 2.0f;
constexpr float kVelocityThreshold = 1.0f;
constexpr float kMaxLinearCorrection = 0.2f;
constexpr float kMaxTranslation = 2.0f;
constexpr float kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
constexpr float kMaxRotation = 0.5f * kPi;
constexpr float kMaxRotationSq = kMaxRotation * kMaxRotation;
constexpr float kBaumgarte = 0.2f;
constexpr float kTimeToSleep = 0.5f;
constexpr float kLinearSleepTol = 0.01f;
constexpr float kAngularSleepTol = 2.0f / 180.0f * kPi;

// ------------------------------------------------------------ FP32 math
// REF/Common/b2Math.h: the operators below evaluate in the same order.
struct V2 { float x, y; };
inline V2 mk(float x, float y) { V2 v; v.x = x; v.y = y; return v; }
inline V2 operator+(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }
inline V2 operator-(V2 a) { return mk(-a.x, -a.y); }
inline V2 operator*(float s, V2 a) { return mk(s * a.x, s * a.y); }
inline void operator+=(V2& a, V2 b) { a.x += b.x; a.y += b.y; }
inline void operator-=(V2& a, V2 b) { a.x -= b.x; a.y -= b.y; }
inline void operator*=(V2& a, float s) { a.x *= s; a.y *= s; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline V2 cross(V2 a, float s) { return mk(s * a.y, -s * a.x); }
inline V2 cross(float s, V2 a) { return mk(-s * a.y, s * a.x); }
inline float lengthSq(V2 a) { return a.x * a.x + a.y * a.y; }
inline float length(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
inline float distSq(V2 a, V2 b) { V2 c = a - b; return dot(c, c); }
inline float vmin(float a, float b) { return a < b ? a : b; }
inline float vmax(float a, float b) { return a > b ? a : b; }
inline V2 vmin(V2 a, V2 b) { return mk(vmin(a.x, b.x), vmin(a.y, b.y)); }
inline V2 vmax(V2 a, V2 b) { return mk(vmax(a.x, b.x), vmax(a.y, b.y)); }
inline float clampf(float a, float lo, float hi) { return vmax(lo, vmin(a, hi)); }
inline float absf(float a) { return a > 0.0f ? a : -a; }

// b2Vec2::Normalize (b2Math.h): zero-length vectors stay as they are.
inline float normalize(V2& v) {
    float len = length(v);
    if (len < FLT_EPSILON) return 0.0f;
    float inv = 1.0f / len;
    v.x *= inv;
    v.y *= inv;
    return len;
}

// Trig used inside Step. 0: glibc sinf/cosf, as b2Rot::Set calls them
// (REF/Common/b2Math.h:302-311); the reference build agrees bit for bit.
// 1: the FP32 rounding of the double-precision sine/cosine, which is what the
// GPU step evaluates (glibc sinf is not correctly rounded, so the two differ
// in the last bit for about 1% of angles). Body creation always uses glibc.
static int g_trigMode = 0;
inline float stepSin(float a) { return g_trigMode ? (float)sin((double)a) : sinf(a); }
inline float stepCos(float a) { return g_trigMode ? (float)cos((double)a) : cosf(a); }

struct Rot { float s, c; };
inline Rot rotOf(float angle) { Rot q; q.s = sinf(angle); q.c = cosf(angle); return q; }
inline Rot stepRot(float angle) { Rot q; q.s = stepSin(angle); q.c = stepCos(angle); return q; }
struct Xf { V2 p; Rot q; };
inline V2 mul(Rot q, V2 v) { return mk(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
inline V2 mulT(Rot q, V2 v) { return mk(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
inline V2 mul(const Xf& T, V2 v) {
    float x = (T.q.c * v.x - T.q.s * v.y) + T.p.x;
    float y = (T.q.s * v.x + T.q.c * v.y) + T.p.y;
    return mk(x, y);
}
inline V2 mulT(const Xf& T, V2 v) {
    float px = v.x - T.p.x;
    float py = v.y - T.p.y;
    float x = (T.q.c * px + T.q.s * py);
    float y = (-T.q.s * px + T.q.c * py);
    return mk(x, y);
}
inline Rot mulT(Rot q, Rot r) {
    Rot o;
    o.s = q.c * r.s - q.s * r.c;
    o.c = q.c * r.c + q.s * r.s;
    return o;
}
inline Xf mulT(const Xf& A, const Xf& B) {
    Xf C;
    C.q = mulT(A.q, B.q);
    C.p = mulT(A.q, B.p - A.p);
    return C;
}

struct AABB { V2 lo, hi; };
// b2TestOverlap(b2AABB, b2AABB), REF/Collision/b2Collision.h:247-259
inline bool overlap(const AABB& a, const AABB& b) {
    V2 d1 = b.lo - a.hi;
    V2 d2 = a.lo - b.hi;
    if (d1.x > 0.0f || d1.y > 0.0f) return false;
    if (d2.x > 0.0f || d2.y > 0.0f) return false;
    return true;
}
// b2AABB::Contains (b2Collision.h:200-208)
inline bool contains(const AABB& a, const AABB& b) {
    return a.lo.x <= b.lo.x && a.lo.y <= b.lo.y && b.hi.x <= a.hi.x && b.hi.y <= a.hi.y;
}

inline V2 v(const b2g_vec2& a) { return mk(a.x, a.y); }
inline b2g_vec2 g(V2 a) { b2g_vec2 o; o.x = a.x; o.y = a.y; return o; }
inline Xf xfOf(const b2g_transform* t) { Xf x; x.p = mk(t->px, t->py); x.q.s = t->s; x.q.c = t->c; return x; }

// ------------------------------------------------------------ shapes
// REF/Collision/Shapes/b2PolygonShape.cpp:42-61 (SetAsBox)
void setAsBox(b2g_shape* s, float hx, float hy) {
    memset(s, 0, sizeof(*s));
    s->type = B2G_POLYGON;
    s->radius = kPolygonRadius;
    s->count = 4;
    s->verts[0] = g(mk(-hx, -hy)); s->verts[1] = g(mk(hx, -hy));
    s->verts[2] = g(mk(hx, hy));   s->verts[3] = g(mk(-hx, hy));
    s->normals[0] = g(mk(0.0f, -1.0f)); s->normals[1] = g(mk(1.0f, 0.0f));
    s->normals[2] = g(mk(0.0f, 1.0f));  s->normals[3] = g(mk(-1.0f, 0.0f));
}

// REF/Collision/Shapes/b2PolygonShape.cpp:63-89 (SetAsBox with centre and angle)
void setAsBoxRotated(b2g_shape* s, float hx, float hy, V2 center, float angle) {
    setAsBox(s, hx, hy);
    s->centroid = g(center);
    Xf xf;
    xf.p = center;
    xf.q = rotOf(angle);
    for (int i = 0; i < 4; ++i) {
        s->verts[i] = g(mul(xf, v(s->verts[i])));
        s->normals[i] = g(mul(xf.q, v(s->normals[i])));
    }
}

// REF/Collision/Shapes/b2PolygonShape.cpp:96-142 (ComputeCentroid, origin reference)
V2 polygonCentroid(const V2* vs, int count) {
    V2 c = mk(0.0f, 0.0f);
    float area = 0.0f;
    const V2 pRef = mk(0.0f, 0.0f);
    const float inv3 = 1.0f / 3.0f;
    for (int i = 0; i < count; ++i) {
        V2 p1 = pRef, p2 = vs[i], p3 = i + 1 < count ? vs[i + 1] : vs[0];
        V2 e1 = p2 - p1, e2 = p3 - p1;
        float D = cross(e1, e2);
        float tri = 0.5f * D;
        area += tri;
        c += tri * inv3 * (p1 + p2 + p3);
    }
    c *= 1.0f / area;
    return c;
}

// REF/Collision/Shapes/b2PolygonShape.cpp:144-245 (Set: weld, gift-wrap hull,
// normals, centroid). Degenerate input falls back to a 1x1 box as upstream.
void polygonSet(b2g_shape* s, const V2* input, int count) {
    if (count < 3) { setAsBox(s, 1.0f, 1.0f); return; }
    int n = count < 8 ? count : 8;
    V2 ps[8];
    int tc = 0;
    for (int i = 0; i < n; ++i) {
        V2 p = input[i];
        bool unique = true;
        for (int j = 0; j < tc; ++j) {
            if (distSq(p, ps[j]) < 0.5f * kLinearSlop) { unique = false; break; }
        }
        if (unique) ps[tc++] = p;
    }
    n = tc;
    if (n < 3) { setAsBox(s, 1.0f, 1.0f); return; }
    int i0 = 0;
    float x0 = ps[0].x;
    for (int i = 1; i < n; ++i) {
        float x = ps[i].x;
        if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) { i0 = i; x0 = x; }
    }
    int hull[8];
    int m = 0, ih = i0;
    for (;;) {
        hull[m] = ih;
        int ie = 0;
        for (int j = 1; j < n; ++j) {
            if (ie == ih) { ie = j; continue; }
            V2 r = ps[ie] - ps[hull[m]];
            V2 w = ps[j] - ps[hull[m]];
            float c = cross(r, w);
            if (c < 0.0f) ie = j;
            if (c == 0.0f && lengthSq(w) > lengthSq(r)) ie = j;
        }
        ++m;
        ih = ie;
        if (ie == i0) break;
    }
    if (m < 3) { setAsBox(s, 1.0f, 1.0f); return; }
    memset(s, 0, sizeof(*s));
    s->type = B2G_POLYGON;
    s->radius = kPolygonRadius;
    s->count = m;
    V2 vs[8];
    for (int i = 0; i < m; ++i) { vs[i] = ps[hull[i]]; s->verts[i] = g(vs[i]); }
    for (int i = 0; i < m; ++i) {
        V2 e = vs[i + 1 < m ? i + 1 : 0] - vs[i];
        V2 nrm = cross(e, 1.0f);
        normalize(nrm);
        s->normals[i] = g(nrm);
    }
    s->centroid = g(polygonCentroid(vs, m));
}

void circleShape(b2g_shape* s, V2 p, float r) {
    memset(s, 0, sizeof(*s));
    s->type = B2G_CIRCLE;
    s->radius = r;
    s->centroid = g(p);
}

void edgeShape(b2g_shape* s, V2 v0, V2 v1, V2 v2, V2 v3, bool has0, bool has3) {
    memset(s, 0, sizeof(*s));
    s->type = B2G_EDGE;
    s->radius = kPolygonRadius;
    s->count = 2;
    s->verts[0] = g(v1); s->verts[1] = g(v2);
    s->v0 = g(v0); s->v3 = g(v3);
    s->flags = (has0 ? 1 : 0) | (has3 ? 2 : 0);
}

// ComputeAABB: polygon b2PolygonShape.cpp:335-352, circle b2CircleShape.cpp:85-94,
// edge b2EdgeShape.cpp:118-131, chain child (flags bit2, no skin) b2ChainShape.cpp:171-187.
AABB shapeAABB(const b2g_shape* s, const Xf& xf) {
    AABB a;
    if (s->type == B2G_CIRCLE) {
        V2 p = xf.p + mul(xf.q, v(s->centroid));
        a.lo = mk(p.x - s->radius, p.y - s->radius);
        a.hi = mk(p.x + s->radius, p.y + s->radius);
        return a;
    }
    const int n = s->type == B2G_EDGE ? 2 : s->count;
    V2 lo = mul(xf, v(s->verts[0]));
    V2 hi = lo;
    for (int i = 1; i < n; ++i) {
        V2 w = mul(xf, v(s->verts[i]));
        lo = vmin(lo, w);
        hi = vmax(hi, w);
    }
    if (s->type == B2G_EDGE && (s->flags & 4)) {
        a.lo = lo;
        a.hi = hi;
        return a;
    }
    V2 r = mk(s->radius, s->radius);
    a.lo = lo - r;
    a.hi = hi + r;
    return a;
}

struct MassData { float mass; V2 center; float I; };

// ComputeMass: polygon b2PolygonShape.cpp:354-437, circle b2CircleShape.cpp:96-103,
// edge b2EdgeShape.cpp:133-139.
MassData shapeMass(const b2g_shape* s, float density) {
    MassData md;
    if (s->type == B2G_CIRCLE) {
        md.mass = density * kPi * s->radius * s->radius;
        md.center = v(s->centroid);
        md.I = md.mass * (0.5f * s->radius * s->radius + dot(md.center, md.center));
        return md;
    }
    if (s->type == B2G_EDGE) {
        md.mass = 0.0f;
        md.center = 0.5f * (v(s->verts[0]) + v(s->verts[1]));
        md.I = 0.0f;
        return md;
    }
    const int n = s->count;
    V2 center = mk(0.0f, 0.0f);
    float area = 0.0f, I = 0.0f;
    V2 sref = mk(0.0f, 0.0f);
    for (int i = 0; i < n; ++i) sref += v(s->verts[i]);
    sref *= 1.0f / n;
    const float k_inv3 = 1.0f / 3.0f;
    for (int i = 0; i < n; ++i) {
        V2 e1 = v(s->verts[i]) - sref;
        V2 e2 = i + 1 < n ? v(s->verts[i + 1]) - sref : v(s->verts[0]) - sref;
        float D = cross(e1, e2);
        float tri = 0.5f * D;
        area += tri;
        center += tri * k_inv3 * (e1 + e2);
        float ex1 = e1.x, ey1 = e1.y, ex2 = e2.x, ey2 = e2.y;
        float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
        float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
        I += (0.25f * k_inv3 * D) * (intx2 + inty2);
    }
    md.mass = density * area;
    center *= 1.0f / area;
    md.center = center + sref;
    md.I = density * I;
    md.I += md.mass * (dot(md.center, md.center) - dot(center, center));
    return md;
}

// ------------------------------------------------------------ narrowphase
struct ClipVertex { V2 v; uint32_t id; };

inline uint32_t makeId(int indexA, int indexB, int typeA, int typeB) {
    return (uint32_t)(indexA & 0xFF) | ((uint32_t)(indexB & 0xFF) << 8) |
           ((uint32_t)(typeA & 0xFF) << 16) | ((uint32_t)(typeB & 0xFF) << 24);
}
inline uint32_t swapId(uint32_t id) {
    uint32_t iA = id & 0xFF, iB = (id >> 8) & 0xFF, tA = (id >> 16) & 0xFF, tB = (id >> 24) & 0xFF;
    return iB | (iA << 8) | (tB << 16) | (tA << 24);
}
constexpr int kVertex = 0, kFace = 1;

void clearManifold(b2g_manifold* m) { memset(m, 0, sizeof(*m)); }

// b2ClipSegmentToLine, REF/Collision/b2Collision.cpp:201-232
int clipSegment(ClipVertex out[2], const ClipVertex in[2], V2 normal, float offset, int vertexIndexA) {
    int n = 0;
    float d0 = dot(normal, in[0].v) - offset;
    float d1 = dot(normal, in[1].v) - offset;
    if (d0 <= 0.0f) out[n++] = in[0];
    if (d1 <= 0.0f) out[n++] = in[1];
    if (d0 * d1 < 0.0f) {
        float interp = d0 / (d0 - d1);
        out[n].v = in[0].v + interp * (in[1].v - in[0].v);
        out[n].id = makeId(vertexIndexA, (int)((in[0].id >> 8) & 0xFF), kVertex, kFace);
        ++n;
    }
    return n;
}

// b2FindMaxSeparation, REF/Collision/b2CollidePolygon.cpp:28-66
float findMaxSeparation(int* edge, const b2g_shape* p1, const Xf& xf1, const b2g_shape* p2, const Xf& xf2) {
    Xf xf = mulT(xf2, xf1);
    int best = 0;
    float maxSep = -FLT_MAX;
    for (int i = 0; i < p1->count; ++i) {
        V2 n = mul(xf.q, v(p1->normals[i]));
        V2 v1 = mul(xf, v(p1->verts[i]));
        float si = FLT_MAX;
        for (int j = 0; j < p2->count; ++j) {
            float sij = dot(n, v(p2->verts[j]) - v1);
            if (sij < si) si = sij;
        }
        if (si > maxSep) { maxSep = si; best = i; }
    }
    *edge = best;
    return maxSep;
}

// b2FindIncidentEdge, REF/Collision/b2CollidePolygon.cpp:69-112
void findIncidentEdge(ClipVertex c[2], const b2g_shape* p1, const Xf& xf1, int edge1,
                      const b2g_shape* p2, const Xf& xf2) {
    V2 normal1 = mulT(xf2.q, mul(xf1.q, v(p1->normals[edge1])));
    int index = 0;
    float minDot = FLT_MAX;
    for (int i = 0; i < p2->count; ++i) {
        float d = dot(normal1, v(p2->normals[i]));
        if (d < minDot) { minDot = d; index = i; }
    }
    int i1 = index, i2 = i1 + 1 < p2->count ? i1 + 1 : 0;
    c[0].v = mul(xf2, v(p2->verts[i1]));
    c[0].id = makeId(edge1, i1, kFace, kVertex);
    c[1].v = mul(xf2, v(p2->verts[i2]));
    c[1].id = makeId(edge1, i2, kFace, kVertex);
}

// b2CollidePolygons, REF/Collision/b2CollidePolygon.cpp:121-247
void collidePolygons(b2g_manifold* m, const b2g_shape* A, const Xf& xfA, const b2g_shape* B, const Xf& xfB) {
    m->pointCount = 0;
    float totalRadius = A->radius + B->radius;
    int edgeA = 0;
    float sepA = findMaxSeparation(&edgeA, A, xfA, B, xfB);
    if (sepA > totalRadius) return;
    int edgeB = 0;
    float sepB = findMaxSeparation(&edgeB, B, xfB, A, xfA);
    if (sepB > totalRadius) return;
    const b2g_shape *p1, *p2;
    Xf xf1, xf2;
    int edge1;
    bool flip;
    const float k_tol = 0.1f * kLinearSlop;
    if (sepB > sepA + k_tol) {
        p1 = B; p2 = A; xf1 = xfB; xf2 = xfA; edge1 = edgeB;
        m->type = B2G_MANIFOLD_FACE_B; flip = true;
    } else {
        p1 = A; p2 = B; xf1 = xfA; xf2 = xfB; edge1 = edgeA;
        m->type = B2G_MANIFOLD_FACE_A; flip = false;
    }
    ClipVertex incident[2];
    findIncidentEdge(incident, p1, xf1, edge1, p2, xf2);
    int count1 = p1->count;
    int iv1 = edge1, iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
    V2 v11 = v(p1->verts[iv1]), v12 = v(p1->verts[iv2]);
    V2 localTangent = v12 - v11;
    normalize(localTangent);
    V2 localNormal = cross(localTangent, 1.0f);
    V2 planePoint = 0.5f * (v11 + v12);
    V2 tangent = mul(xf1.q, localTangent);
    V2 normal = cross(tangent, 1.0f);
    v11 = mul(xf1, v11);
    v12 = mul(xf1, v12);
    float frontOffset = dot(normal, v11);
    float sideOffset1 = -dot(tangent, v11) + totalRadius;
    float sideOffset2 = dot(tangent, v12) + totalRadius;
    ClipVertex cp1[2], cp2[2];
    if (clipSegment(cp1, incident, -tangent, sideOffset1, iv1) < 2) return;
    if (clipSegment(cp2, cp1, tangent, sideOffset2, iv2) < 2) return;
    m->localNormal = g(localNormal);
    m->localPoint = g(planePoint);
    int pc = 0;
    for (int i = 0; i < 2; ++i) {
        float separation = dot(normal, cp2[i].v) - frontOffset;
        if (separation <= totalRadius) {
            b2g_manifold_point* mp = m->points + pc;
            mp->localPoint = g(mulT(xf2, cp2[i].v));
            mp->id = flip ? swapId(cp2[i].id) : cp2[i].id;
            ++pc;
        }
    }
    m->pointCount = pc;
}

// b2CollideCircles, REF/Collision/b2CollideCircle.cpp:25-50
void collideCircles(b2g_manifold* m, const b2g_shape* A, const Xf& xfA, const b2g_shape* B, const Xf& xfB) {
    m->pointCount = 0;
    V2 pA = mul(xfA, v(A->centroid));
    V2 pB = mul(xfB, v(B->centroid));
    V2 d = pB - pA;
    float distSqr = dot(d, d);
    float radius = A->radius + B->radius;
    if (distSqr > radius * radius) return;
    m->type = B2G_MANIFOLD_CIRCLES;
    m->localPoint = A->centroid;
    m->localNormal = g(mk(0.0f, 0.0f));
    m->pointCount = 1;
    m->points[0].localPoint = B->centroid;
    m->points[0].id = 0;
}

// b2CollidePolygonAndCircle, REF/Collision/b2CollideCircle.cpp:53-157
void collidePolygonCircle(b2g_manifold* m, const b2g_shape* A, const Xf& xfA, const b2g_shape* B, const Xf& xfB) {
    m->pointCount = 0;
    V2 c = mul(xfB, v(B->centroid));
    V2 cLocal = mulT(xfA, c);
    int normalIndex = 0;
    float separation = -FLT_MAX;
    float radius = A->radius + B->radius;
    const int n = A->count;
    for (int i = 0; i < n; ++i) {
        float s = dot(v(A->normals[i]), cLocal - v(A->verts[i]));
        if (s > radius) return;
        if (s > separation) { separation = s; normalIndex = i; }
    }
    int vi1 = normalIndex, vi2 = vi1 + 1 < n ? vi1 + 1 : 0;
    V2 v1 = v(A->verts[vi1]), v2 = v(A->verts[vi2]);
    if (separation < FLT_EPSILON) {
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_FACE_A;
        m->localNormal = A->normals[normalIndex];
        m->localPoint = g(0.5f * (v1 + v2));
        m->points[0].localPoint = B->centroid;
        m->points[0].id = 0;
        return;
    }
    float u1 = dot(cLocal - v1, v2 - v1);
    float u2 = dot(cLocal - v2, v1 - v2);
    if (u1 <= 0.0f) {
        if (distSq(cLocal, v1) > radius * radius) return;
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_FACE_A;
        V2 nrm = cLocal - v1;
        normalize(nrm);
        m->localNormal = g(nrm);
        m->localPoint = g(v1);
        m->points[0].localPoint = B->centroid;
        m->points[0].id = 0;
    } else if (u2 <= 0.0f) {
        if (distSq(cLocal, v2) > radius * radius) return;
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_FACE_A;
        V2 nrm = cLocal - v2;
        normalize(nrm);
        m->localNormal = g(nrm);
        m->localPoint = g(v2);
        m->points[0].localPoint = B->centroid;
        m->points[0].id = 0;
    } else {
        V2 faceCenter = 0.5f * (v1 + v2);
        float sep = dot(cLocal - faceCenter, v(A->normals[vi1]));
        if (sep > radius) return;
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_FACE_A;
        m->localNormal = A->normals[vi1];
        m->localPoint = g(faceCenter);
        m->points[0].localPoint = B->centroid;
        m->points[0].id = 0;
    }
}

// b2CollideEdgeAndCircle, REF/Collision/b2CollideEdge.cpp:27-151
void collideEdgeCircle(b2g_manifold* m, const b2g_shape* E, const Xf& xfA, const b2g_shape* C, const Xf& xfB) {
    m->pointCount = 0;
    V2 Q = mulT(xfA, mul(xfB, v(C->centroid)));
    V2 A = v(E->verts[0]), B = v(E->verts[1]);
    V2 e = B - A;
    float u = dot(e, B - Q);
    float w = dot(e, Q - A);
    float radius = E->radius + C->radius;
    if (w <= 0.0f) {
        V2 P = A;
        V2 d = Q - P;
        if (dot(d, d) > radius * radius) return;
        if (E->flags & 1) {
            V2 A1 = v(E->v0), B1 = A;
            V2 e1 = B1 - A1;
            float u1 = dot(e1, B1 - Q);
            if (u1 > 0.0f) return;
        }
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_CIRCLES;
        m->localNormal = g(mk(0.0f, 0.0f));
        m->localPoint = g(P);
        m->points[0].id = makeId(0, 0, kVertex, kVertex);
        m->points[0].localPoint = C->centroid;
        return;
    }
    if (u <= 0.0f) {
        V2 P = B;
        V2 d = Q - P;
        if (dot(d, d) > radius * radius) return;
        if (E->flags & 2) {
            V2 B2 = v(E->v3), A2 = B;
            V2 e2 = B2 - A2;
            float v2 = dot(e2, Q - A2);
            if (v2 > 0.0f) return;
        }
        m->pointCount = 1;
        m->type = B2G_MANIFOLD_CIRCLES;
        m->localNormal = g(mk(0.0f, 0.0f));
        m->localPoint = g(P);
        m->points[0].id = makeId(1, 0, kVertex, kVertex);
        m->points[0].localPoint = C->centroid;
        return;
    }
    float den = dot(e, e);
    V2 P = (1.0f / den) * (u * A + w * B);
    V2 d = Q - P;
    if (dot(d, d) > radius * radius) return;
    V2 n = mk(-e.y, e.x);
    if (dot(n, Q - A) < 0.0f) n = mk(-n.x, -n.y);
    normalize(n);
    m->pointCount = 1;
    m->type = B2G_MANIFOLD_FACE_A;
    m->localNormal = g(n);
    m->localPoint = g(A);
    m->points[0].id = makeId(0, 0, kFace, kVertex);
    m->points[0].localPoint = C->centroid;
}

// b2EPCollider (edge versus polygon with adjacency),
// REF/Collision/b2CollideEdge.cpp:153-700. The fork's copy of this function
// has two transcription bugs (Box2D/Collision/b2CollideEdge.cpp:443, :597);
// the upstream arithmetic is the one restated here.
struct EPAxis { int type; int index; float separation; };  // type: 0 unknown, 1 edgeA, 2 edgeB
void collideEdgePolygon(b2g_manifold* m, const b2g_shape* E, const Xf& xfA, const b2g_shape* P, const Xf& xfB) {
    Xf xf = mulT(xfA, xfB);
    V2 centroidB = mul(xf, v(P->centroid));
    V2 v0 = v(E->v0), v1 = v(E->verts[0]), v2 = v(E->verts[1]), v3 = v(E->v3);
    bool has0 = (E->flags & 1) != 0, has3 = (E->flags & 2) != 0;
    V2 edge1 = v2 - v1;
    normalize(edge1);
    V2 normal0 = mk(0.0f, 0.0f), normal1 = mk(edge1.y, -edge1.x), normal2 = mk(0.0f, 0.0f);
    float offset1 = dot(normal1, centroidB - v1);
    float offset0 = 0.0f, offset2 = 0.0f;
    bool convex1 = false, convex2 = false;
    if (has0) {
        V2 edge0 = v1 - v0;
        normalize(edge0);
        normal0 = mk(edge0.y, -edge0.x);
        convex1 = cross(edge0, edge1) >= 0.0f;
        offset0 = dot(normal0, centroidB - v0);
    }
    if (has3) {
        V2 edge2 = v3 - v2;
        normalize(edge2);
        normal2 = mk(edge2.y, -edge2.x);
        convex2 = cross(edge1, edge2) > 0.0f;
        offset2 = dot(normal2, centroidB - v2);
    }
    bool front;
    V2 normal, lower, upper;
    if (has0 && has3) {
        if (convex1 && convex2) {
            front = offset0 >= 0.0f || offset1 >= 0.0f || offset2 >= 0.0f;
            if (front) { normal = normal1; lower = normal0; upper = normal2; }
            else { normal = -normal1; lower = -normal1; upper = -normal1; }
        } else if (convex1) {
            front = offset0 >= 0.0f || (offset1 >= 0.0f && offset2 >= 0.0f);
            if (front) { normal = normal1; lower = normal0; upper = normal1; }
            else { normal = -normal1; lower = -normal2; upper = -normal1; }
        } else if (convex2) {
            front = offset2 >= 0.0f || (offset0 >= 0.0f && offset1 >= 0.0f);
            if (front) { normal = normal1; lower = normal1; upper = normal2; }
            else { normal = -normal1; lower = -normal1; upper = -normal0; }
        } else {
            front = offset0 >= 0.0f && offset1 >= 0.0f && offset2 >= 0.0f;
            if (front) { normal = normal1; lower = normal1; upper = normal1; }
            else { normal = -normal1; lower = -normal2; upper = -normal0; }
        }
    } else if (has0) {
        if (convex1) {
            front = offset0 >= 0.0f || offset1 >= 0.0f;
            if (front) { normal = normal1; lower = normal0; upper = -normal1; }
            else { normal = -normal1; lower = normal1; upper = -normal1; }
        } else {
            front = offset0 >= 0.0f && offset1 >= 0.0f;
            if (front) { normal = normal1; lower = normal1; upper = -normal1; }
            else { normal = -normal1; lower = normal1; upper = -normal0; }
        }
    } else if (has3) {
        if (convex2) {
            front = offset1 >= 0.0f || offset2 >= 0.0f;
            if (front) { normal = normal1; lower = -normal1; upper = normal2; }
            else { normal = -normal1; lower = -normal1; upper = normal1; }
        } else {
            front = offset1 >= 0.0f && offset2 >= 0.0f;
            if (front) { normal = normal1; lower = -normal1; upper = normal1; }
            else { normal = -normal1; lower = -normal2; upper = normal1; }
        }
    } else {
        front = offset1 >= 0.0f;
        if (front) { normal = normal1; lower = -normal1; upper = -normal1; }
        else { normal = -normal1; lower = normal1; upper = normal1; }
    }
    // Polygon B in frame A.
    const int count = P->count;
    V2 pv[8], pn[8];
    for (int i = 0; i < count; ++i) {
        pv[i] = mul(xf, v(P->verts[i]));
        pn[i] = mul(xf.q, v(P->normals[i]));
    }
    const float radius = 2.0f * kPolygonRadius;
    m->pointCount = 0;
    // ComputeEdgeSeparation (REF :606-623)
    EPAxis edgeAxis;
    edgeAxis.type = 1;
    edgeAxis.index = front ? 0 : 1;
    edgeAxis.separation = FLT_MAX;
    for (int i = 0; i < count; ++i) {
        float s = dot(normal, pv[i] - v1);
        if (s < edgeAxis.separation) edgeAxis.separation = s;
    }
    if (edgeAxis.separation > radius) return;
    // ComputePolygonSeparation (REF :625-676)
    EPAxis polyAxis;
    polyAxis.type = 0;
    polyAxis.index = -1;
    polyAxis.separation = -FLT_MAX;
    {
        V2 perp = mk(-normal.y, normal.x);
        for (int i = 0; i < count; ++i) {
            V2 n = -pn[i];
            float s1 = dot(n, pv[i] - v1);
            float s2 = dot(n, pv[i] - v2);
            float s = vmin(s1, s2);
            if (s > radius) {
                polyAxis.type = 2; polyAxis.index = i; polyAxis.separation = s;
                break;
            }
            if (dot(n, perp) >= 0.0f) {
                if (dot(n - upper, normal) < -kAngularSlop) continue;
            } else {
                if (dot(n - lower, normal) < -kAngularSlop) continue;
            }
            if (s > polyAxis.separation) {
                polyAxis.type = 2; polyAxis.index = i; polyAxis.separation = s;
            }
        }
    }
    if (polyAxis.type != 0 && polyAxis.separation > radius) return;
    const float k_relativeTol = 0.98f, k_absoluteTol = 0.001f;
    EPAxis primary;
    if (polyAxis.type == 0) primary = edgeAxis;
    else if (polyAxis.separation > k_relativeTol * edgeAxis.separation + k_absoluteTol) primary = polyAxis;
    else primary = edgeAxis;

    ClipVertex ie[2];
    int rfi1, rfi2;
    V2 rfv1, rfv2, rfn;
    if (primary.type == 1) {
        m->type = B2G_MANIFOLD_FACE_A;
        int best = 0;
        float bestValue = dot(normal, pn[0]);
        for (int i = 1; i < count; ++i) {
            float value = dot(normal, pn[i]);
            if (value < bestValue) { bestValue = value; best = i; }
        }
        int i1 = best, i2 = i1 + 1 < count ? i1 + 1 : 0;
        ie[0].v = pv[i1]; ie[0].id = makeId(0, i1, kFace, kVertex);
        ie[1].v = pv[i2]; ie[1].id = makeId(0, i2, kFace, kVertex);
        if (front) { rfi1 = 0; rfi2 = 1; rfv1 = v1; rfv2 = v2; rfn = normal1; }
        else { rfi1 = 1; rfi2 = 0; rfv1 = v2; rfv2 = v1; rfn = -normal1; }
    } else {
        m->type = B2G_MANIFOLD_FACE_B;
        ie[0].v = v1; ie[0].id = makeId(0, primary.index, kVertex, kFace);
        ie[1].v = v2; ie[1].id = makeId(0, primary.index, kVertex, kFace);
        rfi1 = primary.index;
        rfi2 = rfi1 + 1 < count ? rfi1 + 1 : 0;
        rfv1 = pv[rfi1]; rfv2 = pv[rfi2]; rfn = pn[rfi1];
    }
    V2 side1 = mk(rfn.y, -rfn.x);
    V2 side2 = -side1;
    float sideOffset1 = dot(side1, rfv1);
    float sideOffset2 = dot(side2, rfv2);
    ClipVertex cp1[2], cp2[2];
    if (clipSegment(cp1, ie, side1, sideOffset1, rfi1) < 2) return;
    if (clipSegment(cp2, cp1, side2, sideOffset2, rfi2) < 2) return;
    if (primary.type == 1) {
        m->localNormal = g(rfn);
        m->localPoint = g(rfv1);
    } else {
        m->localNormal = P->normals[rfi1];
        m->localPoint = P->verts[rfi1];
    }
    int pc = 0;
    for (int i = 0; i < 2; ++i) {
        float separation = dot(rfn, cp2[i].v - rfv1);
        if (separation <= radius) {
            b2g_manifold_point* mp = m->points + pc;
            if (primary.type == 1) {
                mp->localPoint = g(mulT(xf, cp2[i].v));
                mp->id = cp2[i].id;
            } else {
                mp->localPoint = g(cp2[i].v);
                mp->id = swapId(cp2[i].id);
            }
            ++pc;
        }
    }
    m->pointCount = pc;
}

// Dispatch for a contact whose shapes are already in registry order
// (REF/Dynamics/Contacts/b2Contact.cpp:34-43 and each Evaluate override).
bool collide(b2g_manifold* m, const b2g_shape* A, const Xf& xfA, const b2g_shape* B, const Xf& xfB) {
    if (A->type == B2G_POLYGON && B->type == B2G_POLYGON) collidePolygons(m, A, xfA, B, xfB);
    else if (A->type == B2G_CIRCLE && B->type == B2G_CIRCLE) collideCircles(m, A, xfA, B, xfB);
    else if (A->type == B2G_POLYGON && B->type == B2G_CIRCLE) collidePolygonCircle(m, A, xfA, B, xfB);
    else if (A->type == B2G_EDGE && B->type == B2G_CIRCLE) collideEdgeCircle(m, A, xfA, B, xfB);
    else if (A->type == B2G_EDGE && B->type == B2G_POLYGON) collideEdgePolygon(m, A, xfA, B, xfB);
    else return false;
    return true;
}

// b2WorldManifold::Initialize, REF/Collision/b2Collision.cpp:22-86
struct WorldManifold { V2 normal; V2 points[2]; float separations[2]; };
void worldManifold(WorldManifold* wm, const b2g_manifold* m, const Xf& xfA, float radiusA, const Xf& xfB, float radiusB) {
    if (m->pointCount == 0) return;
    if (m->type == B2G_MANIFOLD_CIRCLES) {
        wm->normal = mk(1.0f, 0.0f);
        V2 pointA = mul(xfA, v(m->localPoint));
        V2 pointB = mul(xfB, v(m->points[0].localPoint));
        if (distSq(pointA, pointB) > FLT_EPSILON * FLT_EPSILON) {
            wm->normal = pointB - pointA;
            normalize(wm->normal);
        }
        V2 cA = pointA + radiusA * wm->normal;
        V2 cB = pointB - radiusB * wm->normal;
        wm->points[0] = 0.5f * (cA + cB);
        wm->separations[0] = dot(cB - cA, wm->normal);
    } else if (m->type == B2G_MANIFOLD_FACE_A) {
        wm->normal = mul(xfA.q, v(m->localNormal));
        V2 planePoint = mul(xfA, v(m->localPoint));
        for (int i = 0; i < m->pointCount; ++i) {
            V2 clipPoint = mul(xfB, v(m->points[i].localPoint));
            V2 cA = clipPoint + (radiusA - dot(clipPoint - planePoint, wm->normal)) * wm->normal;
            V2 cB = clipPoint - radiusB * wm->normal;
            wm->points[i] = 0.5f * (cA + cB);
            wm->separations[i] = dot(cB - cA, wm->normal);
        }
    } else {
        wm->normal = mul(xfB.q, v(m->localNormal));
        V2 planePoint = mul(xfB, v(m->localPoint));
        for (int i = 0; i < m->pointCount; ++i) {
            V2 clipPoint = mul(xfA, v(m->points[i].localPoint));
            V2 cB = clipPoint + (radiusB - dot(clipPoint - planePoint, wm->normal)) * wm->normal;
            V2 cA = clipPoint - radiusA * wm->normal;
            wm->points[i] = 0.5f * (cA + cB);
            wm->separations[i] = dot(cA - cB, wm->normal);
        }
        wm->normal = -wm->normal;
    }
}

// ------------------------------------------------------------ world
enum BodyType { kStatic = 0, kKinematic = 1, kDynamic = 2 };
enum BodyFlag { kIsland = 1, kAwake = 2, kAutoSleep = 4, kBullet = 8, kFixedRotation = 16, kActive = 32 };
enum ContactFlag { cIsland = 1, cTouching = 2, cEnabled = 4, cFilter = 8, cToi = 0x20 };

// ------------------------------------------------------------ time of impact
// Convex proxy of a shape for GJK (b2DistanceProxy::Set,
// REF/Collision/b2Distance.cpp:31-82): a circle is its centre with its
// radius, a polygon its vertices with the polygon skin, an edge its two ends.
struct Proxy { V2 v[8]; int count; float radius; };
inline Proxy proxyOf(const b2g_shape* s) {
    Proxy p;
    p.radius = s->radius;
    if (s->type == B2G_CIRCLE) {
        p.count = 1;
        p.v[0] = v(s->centroid);
    } else if (s->type == B2G_EDGE) {
        p.count = 2;
        p.v[0] = v(s->verts[0]);
        p.v[1] = v(s->verts[1]);
    } else {
        p.count = s->count;
        for (int i = 0; i < s->count; ++i) p.v[i] = v(s->verts[i]);
    }
    return p;
}
// b2DistanceProxy::GetSupport, REF/Collision/b2Distance.h:109-124: first
// vertex of largest projection.
inline int support(const Proxy& p, V2 d) {
    int best = 0;
    float bestValue = dot(p.v[0], d);
    for (int i = 1; i < p.count; ++i) {
        const float value = dot(p.v[i], d);
        if (value > bestValue) { best = i; bestValue = value; }
    }
    return best;
}

// b2Sweep (REF/Common/b2Math.h:383-405, 695-721)
struct Sweep { V2 localCenter, c0, c; float a0, a, alpha0; };
inline Xf sweepXf(const Sweep& sw, float beta) {
    Xf xf;
    xf.p = (1.0f - beta) * sw.c0 + beta * sw.c;
    const float angle = (1.0f - beta) * sw.a0 + beta * sw.a;
    xf.q = stepRot(angle);
    xf.p -= mul(xf.q, sw.localCenter);
    return xf;
}
inline void sweepAdvance(Sweep& sw, float alpha) {
    const float beta = (alpha - sw.alpha0) / (1.0f - sw.alpha0);
    sw.c0 += beta * (sw.c - sw.c0);
    sw.a0 += beta * (sw.a - sw.a0);
    sw.alpha0 = alpha;
}
inline void sweepNormalize(Sweep& sw) {
    const float twoPi = 2.0f * kPi;
    const float d = twoPi * floorf(sw.a0 / twoPi);
    sw.a0 -= d;
    sw.a -= d;
}

// GJK distance between two proxies (b2Distance with b2Simplex and
// b2SimplexCache, REF/Collision/b2Distance.cpp:84-606), without radii.
struct SimplexCache { float metric; int count; int indexA[3], indexB[3]; };
struct SimplexVertex { V2 wA, wB, w; float a; int indexA, indexB; };
struct Simplex {
    SimplexVertex v[3];
    int count;
    float metric() const {
        if (count == 2) return length(v[0].w - v[1].w);
        if (count == 3) return cross(v[1].w - v[0].w, v[2].w - v[0].w);
        return 0.0f;
    }
};
inline void gjk(V2* pointA, V2* pointB, float* distance, SimplexCache* cache, const Proxy& pA, const Xf& xfA,
                const Proxy& pB, const Xf& xfB) {
    Simplex sx;
    // ReadCache: reuse the cached vertices unless the metric changed a lot.
    sx.count = cache->count;
    for (int i = 0; i < sx.count; ++i) {
        SimplexVertex& vx = sx.v[i];
        vx.indexA = cache->indexA[i];
        vx.indexB = cache->indexB[i];
        vx.wA = mul(xfA, pA.v[vx.indexA]);
        vx.wB = mul(xfB, pB.v[vx.indexB]);
        vx.w = vx.wB - vx.wA;
        vx.a = 0.0f;
    }
    if (sx.count > 1) {
        const float m1 = cache->metric, m2 = sx.metric();
        if (m2 < 0.5f * m1 || 2.0f * m1 < m2 || m2 < FLT_EPSILON) sx.count = 0;
    }
    if (sx.count == 0) {
        SimplexVertex& vx = sx.v[0];
        vx.indexA = 0;
        vx.indexB = 0;
        vx.wA = mul(xfA, pA.v[0]);
        vx.wB = mul(xfB, pB.v[0]);
        vx.w = vx.wB - vx.wA;
        vx.a = 1.0f;
        sx.count = 1;
    }
    int saveA[3], saveB[3];
    for (int iter = 0; iter < 20;) {
        const int saveCount = sx.count;
        for (int i = 0; i < saveCount; ++i) { saveA[i] = sx.v[i].indexA; saveB[i] = sx.v[i].indexB; }
        if (sx.count == 2) {
            // closest feature of a segment (Solve2)
            const V2 w1 = sx.v[0].w, w2 = sx.v[1].w;
            const V2 e12 = w2 - w1;
            const float d12_2 = -dot(w1, e12);
            const float d12_1 = dot(w2, e12);
            if (d12_2 <= 0.0f) {
                sx.v[0].a = 1.0f;
                sx.count = 1;
            } else if (d12_1 <= 0.0f) {
                sx.v[1].a = 1.0f;
                sx.count = 1;
                sx.v[0] = sx.v[1];
            } else {
                const float inv = 1.0f / (d12_1 + d12_2);
                sx.v[0].a = d12_1 * inv;
                sx.v[1].a = d12_2 * inv;
                sx.count = 2;
            }
        } else if (sx.count == 3) {
            // closest feature of a triangle (Solve3)
            const V2 w1 = sx.v[0].w, w2 = sx.v[1].w, w3 = sx.v[2].w;
            const V2 e12 = w2 - w1, e13 = w3 - w1, e23 = w3 - w2;
            const float d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
            const float d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
            const float d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
            const float n123 = cross(e12, e13);
            const float d123_1 = n123 * cross(w2, w3);
            const float d123_2 = n123 * cross(w3, w1);
            const float d123_3 = n123 * cross(w1, w2);
            if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
                sx.v[0].a = 1.0f;
                sx.count = 1;
            } else if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
                const float inv = 1.0f / (d12_1 + d12_2);
                sx.v[0].a = d12_1 * inv;
                sx.v[1].a = d12_2 * inv;
                sx.count = 2;
            } else if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
                const float inv = 1.0f / (d13_1 + d13_2);
                sx.v[0].a = d13_1 * inv;
                sx.v[2].a = d13_2 * inv;
                sx.count = 2;
                sx.v[1] = sx.v[2];
            } else if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
                sx.v[1].a = 1.0f;
                sx.count = 1;
                sx.v[0] = sx.v[1];
            } else if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
                sx.v[2].a = 1.0f;
                sx.count = 1;
                sx.v[0] = sx.v[2];
            } else if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
                const float inv = 1.0f / (d23_1 + d23_2);
                sx.v[1].a = d23_1 * inv;
                sx.v[2].a = d23_2 * inv;
                sx.count = 2;
                sx.v[0] = sx.v[2];
            } else {
                const float inv = 1.0f / (d123_1 + d123_2 + d123_3);
                sx.v[0].a = d123_1 * inv;
                sx.v[1].a = d123_2 * inv;
                sx.v[2].a = d123_3 * inv;
                sx.count = 3;
            }
        }
        if (sx.count == 3) break;  // the origin is inside the triangle
        // search direction towards the origin
        V2 d;
        if (sx.count == 1) {
            d = -sx.v[0].w;
        } else {
            const V2 e12 = sx.v[1].w - sx.v[0].w;
            d = cross(e12, -sx.v[0].w) > 0.0f ? cross(1.0f, e12) : cross(e12, 1.0f);
        }
        if (dot(d, d) < FLT_EPSILON * FLT_EPSILON) break;
        SimplexVertex& nv = sx.v[sx.count];
        nv.indexA = support(pA, mulT(xfA.q, -d));
        nv.wA = mul(xfA, pA.v[nv.indexA]);
        nv.indexB = support(pB, mulT(xfB.q, d));
        nv.wB = mul(xfB, pB.v[nv.indexB]);
        nv.w = nv.wB - nv.wA;
        ++iter;
        bool duplicate = false;
        for (int i = 0; i < saveCount; ++i) {
            if (nv.indexA == saveA[i] && nv.indexB == saveB[i]) { duplicate = true; break; }
        }
        if (duplicate) break;
        ++sx.count;
    }
    // witness points
    if (sx.count == 1) {
        *pointA = sx.v[0].wA;
        *pointB = sx.v[0].wB;
    } else if (sx.count == 2) {
        *pointA = sx.v[0].a * sx.v[0].wA + sx.v[1].a * sx.v[1].wA;
        *pointB = sx.v[0].a * sx.v[0].wB + sx.v[1].a * sx.v[1].wB;
    } else {
        *pointA = sx.v[0].a * sx.v[0].wA + sx.v[1].a * sx.v[1].wA + sx.v[2].a * sx.v[2].wA;
        *pointB = *pointA;
    }
    *distance = length(*pointA - *pointB);
    cache->metric = sx.metric();
    cache->count = sx.count;
    for (int i = 0; i < sx.count; ++i) { cache->indexA[i] = sx.v[i].indexA; cache->indexB[i] = sx.v[i].indexB; }
}

// Separating-axis function of the TOI root finder (b2SeparationFunction,
// REF/Collision/b2TimeOfImpact.cpp:35-253).
struct SepFn {
    enum { kPoints, kFaceA, kFaceB };
    const Proxy* pA;
    const Proxy* pB;
    Sweep sA, sB;
    int type;
    V2 localPoint, axis;

    float init(const SimplexCache& cache, float t1) {
        const Xf xfA = sweepXf(sA, t1), xfB = sweepXf(sB, t1);
        if (cache.count == 1) {
            type = kPoints;
            axis = mul(xfB, pB->v[cache.indexB[0]]) - mul(xfA, pA->v[cache.indexA[0]]);
            return normalize(axis);
        }
        if (cache.indexA[0] == cache.indexA[1]) {
            type = kFaceB;
            const V2 b1 = pB->v[cache.indexB[0]], b2 = pB->v[cache.indexB[1]];
            axis = cross(b2 - b1, 1.0f);
            normalize(axis);
            const V2 normal = mul(xfB.q, axis);
            localPoint = 0.5f * (b1 + b2);
            const V2 pointB = mul(xfB, localPoint);
            const V2 pointA = mul(xfA, pA->v[cache.indexA[0]]);
            float sep = dot(pointA - pointB, normal);
            if (sep < 0.0f) { axis = -axis; sep = -sep; }
            return sep;
        }
        type = kFaceA;
        const V2 a1 = pA->v[cache.indexA[0]], a2 = pA->v[cache.indexA[1]];
        axis = cross(a2 - a1, 1.0f);
        normalize(axis);
        const V2 normal = mul(xfA.q, axis);
        localPoint = 0.5f * (a1 + a2);
        const V2 pointA = mul(xfA, localPoint);
        const V2 pointB = mul(xfB, pB->v[cache.indexB[0]]);
        float sep = dot(pointB - pointA, normal);
        if (sep < 0.0f) { axis = -axis; sep = -sep; }
        return sep;
    }
    float findMin(int* iA, int* iB, float t) const {
        const Xf xfA = sweepXf(sA, t), xfB = sweepXf(sB, t);
        if (type == kPoints) {
            *iA = support(*pA, mulT(xfA.q, axis));
            *iB = support(*pB, mulT(xfB.q, -axis));
            return dot(mul(xfB, pB->v[*iB]) - mul(xfA, pA->v[*iA]), axis);
        }
        if (type == kFaceA) {
            const V2 normal = mul(xfA.q, axis);
            const V2 pointA = mul(xfA, localPoint);
            *iA = -1;
            *iB = support(*pB, mulT(xfB.q, -normal));
            return dot(mul(xfB, pB->v[*iB]) - pointA, normal);
        }
        const V2 normal = mul(xfB.q, axis);
        const V2 pointB = mul(xfB, localPoint);
        *iB = -1;
        *iA = support(*pA, mulT(xfA.q, -normal));
        return dot(mul(xfA, pA->v[*iA]) - pointB, normal);
    }
    float eval(int iA, int iB, float t) const {
        const Xf xfA = sweepXf(sA, t), xfB = sweepXf(sB, t);
        if (type == kPoints) return dot(mul(xfB, pB->v[iB]) - mul(xfA, pA->v[iA]), axis);
        if (type == kFaceA) {
            const V2 normal = mul(xfA.q, axis);
            return dot(mul(xfB, pB->v[iB]) - mul(xfA, localPoint), normal);
        }
        const V2 normal = mul(xfB.q, axis);
        return dot(mul(xfA, pA->v[iA]) - mul(xfB, localPoint), normal);
    }
};

// b2TimeOfImpact, REF/Collision/b2TimeOfImpact.cpp:257-489 (conservative
// advancement on separating axes with a bisection/secant root finder).
enum { kToiUnknown, kToiFailed, kToiOverlapped, kToiTouching, kToiSeparated };
inline int timeOfImpact(float* tOut, const Proxy& pA, const Proxy& pB, Sweep sA, Sweep sB, float tMax) {
    int state = kToiUnknown;
    *tOut = tMax;
    sweepNormalize(sA);
    sweepNormalize(sB);
    const float totalRadius = pA.radius + pB.radius;
    const float target = vmax(kLinearSlop, totalRadius - 3.0f * kLinearSlop);
    const float tolerance = 0.25f * kLinearSlop;
    float t1 = 0.0f;
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    for (int iter = 0;;) {
        const Xf xfA = sweepXf(sA, t1), xfB = sweepXf(sB, t1);
        V2 wa, wb;
        float dist;
        gjk(&wa, &wb, &dist, &cache, pA, xfA, pB, xfB);
        if (dist <= 0.0f) { state = kToiOverlapped; *tOut = 0.0f; break; }
        if (dist < target + tolerance) { state = kToiTouching; *tOut = t1; break; }
        SepFn fcn;
        fcn.pA = &pA; fcn.pB = &pB; fcn.sA = sA; fcn.sB = sB;
        fcn.init(cache, t1);
        bool done = false;
        float t2 = tMax;
        for (int pushBack = 0;;) {
            int iA, iB;
            float s2 = fcn.findMin(&iA, &iB, t2);
            if (s2 > target + tolerance) { state = kToiSeparated; *tOut = tMax; done = true; break; }
            if (s2 > target - tolerance) { t1 = t2; break; }
            float s1 = fcn.eval(iA, iB, t1);
            if (s1 < target - tolerance) { state = kToiFailed; *tOut = t1; done = true; break; }
            if (s1 <= target + tolerance) { state = kToiTouching; *tOut = t1; done = true; break; }
            float a1 = t1, a2 = t2;
            for (int rootIters = 0;;) {
                const float t = (rootIters & 1) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : 0.5f * (a1 + a2);
                ++rootIters;
                const float sv = fcn.eval(iA, iB, t);
                if (absf(sv - target) < tolerance) { t2 = t; break; }
                if (sv > target) { a1 = t; s1 = sv; }
                else { a2 = t; s2 = sv; }
                if (rootIters == 50) break;
            }
            if (++pushBack == 8) break;
        }
        ++iter;
        if (done) break;
        if (iter == 20) { state = kToiFailed; *tOut = t1; break; }
    }
    return state;
}

struct Body {
    int type;
    int flags;
    Xf xf;
    V2 localCenter, c0, c;
    float a0, a;
    float alpha0;  // sweep time of c0 / a0 inside a TOI pass
    V2 v;
    float w;
    V2 force;
    float torque;
    float mass, invMass, I, invI;
    float linearDamping, angularDamping, gravityScale;
    float sleepTime;
    int islandIndex;
    std::vector<int> fixtures;  // creation order; iterate reversed (newest first)
    int edgeHead;               // contact edge list head: contact*2 + side
    int jointHead;              // joint edge list head: joint*2 + side
    bool alive;
};

struct Fixture {
    int body;
    b2g_shape shape;
    float density, friction, restitution;
    bool sensor;
    uint16_t category, mask;
    int16_t group;
    AABB aabb;   // swept tight AABB of the last synchronize
    AABB fat;    // fat AABB in the broad-phase tree
    bool proxy;  // has a broad-phase proxy
    int proxyId; // its node id in the broad-phase tree (-1 without a proxy)
    int owner;   // first child fixture of the same b2Fixture (chain edges share one)
};

struct Edge { int prev, next; };  // contact-edge links (edge ids)

struct Contact {
    int fA, fB;
    int flags;
    float friction, restitution, tangentSpeed;
    b2g_manifold manifold;
    float toi;       // cached time of impact (valid with cToi)
    int toiCount;    // TOI events of this contact in the step
    int prev, next;  // world list
    Edge node[2];    // node[0] on body A (other = B), node[1] on body B
    bool alive;
};

// b2ContactVelocityConstraint / b2ContactPositionConstraint,
// REF/Dynamics/Contacts/b2ContactSolver.h:33-84 and b2ContactSolver.cpp:35-50
struct VelPoint { V2 rA, rB; float normalImpulse, tangentImpulse, normalMass, tangentMass, velocityBias; };
struct VelConstraint {
    VelPoint points[2];
    V2 normal;
    float K[4];           // ex.x ex.y ey.x ey.y
    float normalMass[4];
    int indexA, indexB;
    float invMassA, invMassB, invIA, invIB;
    float friction, restitution, tangentSpeed;
    int pointCount, contactIndex;
};
struct PosConstraint {
    V2 localPoints[2];
    V2 localNormal, localPoint;
    int indexA, indexB;
    float invMassA, invMassB;
    V2 localCenterA, localCenterB;
    float invIA, invIB;
    int type;
    float radiusA, radiusB;
    int pointCount;
};

// b2RevoluteJoint, REF/Dynamics/Joints/b2RevoluteJoint.h:92-200 and
// b2Joint.h (edges, island flag, collideConnected).
struct V3 { float x, y, z; };
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
struct Mat33 { V3 ex, ey, ez; };
// b2Mat33::Solve33 / Solve22, REF/Common/b2Math.cpp:28-56
inline V3 solve33(const Mat33& m, V3 b) {
    float det = dot3(m.ex, cross3(m.ey, m.ez));
    if (det != 0.0f) det = 1.0f / det;
    V3 x;
    x.x = det * dot3(b, cross3(m.ey, m.ez));
    x.y = det * dot3(m.ex, cross3(b, m.ez));
    x.z = det * dot3(m.ex, cross3(m.ey, b));
    return x;
}
// b2Mat22 with GetInverse and b2Mul(b2Mat22, b2Vec2), REF/Common/b2Math.h:226-238, 436-439
struct Mat22 { V2 ex, ey; };
inline Mat22 inverse(const Mat22& K) {
    const float a = K.ex.x, b = K.ey.x, c = K.ex.y, d = K.ey.y;
    float det = a * d - b * c;
    if (det != 0.0f) det = 1.0f / det;
    Mat22 B;
    B.ex.x = det * d; B.ey.x = -det * b;
    B.ex.y = -det * c; B.ey.y = det * a;
    return B;
}
inline V2 mul(const Mat22& A, V2 v) { return V2{A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y}; }

// b2Mat33::GetInverse22 and GetSymInverse33, REF/Common/b2Math.cpp:58-97
inline Mat33 inverse22(const Mat33& K) {
    const float a = K.ex.x, b = K.ey.x, c = K.ex.y, d = K.ey.y;
    float det = a * d - b * c;
    if (det != 0.0f) det = 1.0f / det;
    Mat33 M;
    M.ex = V3{det * d, -det * c, 0.0f};
    M.ey = V3{-det * b, det * a, 0.0f};
    M.ez = V3{0.0f, 0.0f, 0.0f};
    return M;
}
inline Mat33 symInverse33(const Mat33& K) {
    float det = dot3(K.ex, cross3(K.ey, K.ez));
    if (det != 0.0f) det = 1.0f / det;
    const float a11 = K.ex.x, a12 = K.ey.x, a13 = K.ez.x;
    const float a22 = K.ey.y, a23 = K.ez.y;
    const float a33 = K.ez.z;
    Mat33 M;
    M.ex.x = det * (a22 * a33 - a23 * a23);
    M.ex.y = det * (a13 * a23 - a12 * a33);
    M.ex.z = det * (a12 * a23 - a13 * a22);
    M.ey.x = M.ex.y;
    M.ey.y = det * (a11 * a33 - a13 * a13);
    M.ey.z = det * (a13 * a12 - a11 * a23);
    M.ez.x = M.ex.z;
    M.ez.y = M.ey.z;
    M.ez.z = det * (a11 * a22 - a12 * a12);
    return M;
}
// b2Mul(b2Mat33, b2Vec3) and b2Mul22, REF/Common/b2Math.h:531-540
inline V3 mul33(const Mat33& A, V3 v) {
    return V3{v.x * A.ex.x + v.y * A.ey.x + v.z * A.ez.x, v.x * A.ex.y + v.y * A.ey.y + v.z * A.ez.y,
              v.x * A.ex.z + v.y * A.ey.z + v.z * A.ez.z};
}
inline V2 mul22(const Mat33& A, V2 v) { return V2{A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y}; }

inline V2 solve22(const Mat33& m, V2 b) {
    float a11 = m.ex.x, a12 = m.ey.x, a21 = m.ex.y, a22 = m.ey.y;
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return mk(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}
enum { kInactiveLimit = 0, kAtLowerLimit = 1, kAtUpperLimit = 2, kEqualLimits = 3 };
struct Joint {
    int bodyA, bodyB;
    bool collideConnected;
    bool islandFlag;
    int prev, next;        // world list (newest first)
    Edge node[2];          // node[0] on body A (other = B), node[1] on body B
    V2 localAnchorA, localAnchorB;
    float referenceAngle;
    V3 impulse;
    float motorImpulse;
    bool enableMotor, enableLimit;
    float maxMotorTorque, motorSpeed, lowerAngle, upperAngle;
    int limitState;
    bool removed;
    // b2DistanceJoint (type 1), REF/Dynamics/Joints/b2DistanceJoint.h:88-113;
    // b2RopeJoint (type 2: length is maxLength, distMass the rope's mass,
    // limitState its state), REF/Dynamics/Joints/b2RopeJoint.h:85-110;
    // b2WeldJoint (type 3: referenceAngle, frequencyHz, dampingRatio, mass,
    // gamma, bias), REF/Dynamics/Joints/b2WeldJoint.h:99-124
    int type;
    float length, frequencyHz, dampingRatio;
    V2 u;
    float distMass, gamma, bias;
    float ropeLength;
    // b2FrictionJoint (type 4) and b2MotorJoint (type 5: linearOffset in
    // localAnchorA, angularOffset in referenceAngle),
    // REF/Dynamics/Joints/b2FrictionJoint.h:88-106, b2MotorJoint.h:98-119
    float maxForce, maxTorque, correctionFactor;
    V2 linearImpulse;
    float angularImpulse;
    Mat22 linearMass;
    float angularMass;
    V2 linearError;
    float angularError;
    // b2PulleyJoint (type 6: length = lengthA), REF/Dynamics/Joints/b2PulleyJoint.h:125-149;
    // b2MouseJoint (type 7: target in localAnchorA, maxForce, frequencyHz,
    // dampingRatio, gamma), REF/Dynamics/Joints/b2MouseJoint.h:110-130
    V2 groundAnchorA, groundAnchorB, uA, uB;
    float lengthB, ratio, constant, pulleyMass, beta;
    V2 mouseC;
    // b2PrismaticJoint (type 8: lowerAngle / upperAngle hold the
    // translations, maxMotorTorque the force), REF/Dynamics/Joints/b2PrismaticJoint.h:150-178;
    // b2WheelJoint (type 9), REF/Dynamics/Joints/b2WheelJoint.h:150-190
    V2 localXAxisA, localYAxisA, axis, perp;
    float s1, s2, a1, a2;
    V2 ax, ay;
    float sAx, sBx, sAy, sBy, wheelMass, springMass, springImpulse;
    // b2GearJoint (type 10; bodyA / bodyB, ratio and constant above),
    // REF/Dynamics/Joints/b2GearJoint.h:80-120
    int bodyC, bodyD, typeA, typeB;
    V2 localAnchorC, localAnchorD, localAxisC, localAxisD, JvAC, JvBD;
    float referenceAngleA, referenceAngleB, JwA, JwB, JwC, JwD, gearMass;
    // solver temporaries
    int indexA, indexB;
    V2 rA, rB, localCenterA, localCenterB;
    float invMassA, invMassB, invIA, invIB;
    Mat33 mass;
    float motorMass;
};

struct Events { long begin = 0, end = 0; };

struct World {
    V2 gravity;
    std::vector<Body> bodies;
    std::vector<Fixture> fixtures;
    std::vector<Contact> contacts;
    std::vector<int> freeContacts;
    int contactHead = -1;
    int contactCount = 0;
    std::vector<int> moveBuffer;
    // Node ids of the broad-phase tree's pool, b2DynamicTree::AllocateNode /
    // FreeNode / CreateProxy / DestroyProxy / InsertLeaf / RemoveLeaf
    // (REF/Collision/b2DynamicTree.cpp:25-53, 56-131, 179-265, 319-376): a
    // free list that starts as 0..15 and doubles when it runs out (the new
    // ids in ascending order), taking freed nodes at its head. A proxy takes
    // a leaf node, then a parent node unless the tree was empty; destroying
    // one frees its parent (unless it was the root) and then its leaf.
    // Moving a proxy frees its parent and takes the same node straight back,
    // and nothing else in a step allocates, so only creation and destruction
    // change the free list. Freed parents are kept as -1: which internal node
    // a leaf hangs from depends on the tree's shape, but such a node is only
    // ever taken again as a parent (a freed leaf always lies above it).
    std::vector<int> treeFree;  // free node ids, the head at the back
    int treeCapacity = 0, treeLeaves = 0;
    std::vector<int> proxyFixture;  // fixture of each live proxy id
    bool newFixture = false;
    bool allowSleep = true, warmStarting = true, continuous = false, subStepping = false;
    bool autoClearForces = true;
    bool stepComplete = true;
    float inv_dt0 = 0.0f;
    int unsupported = 0;
    Events events;

    std::vector<Joint> joints;
    int jointHead = -1;

    // island scratch
    std::vector<int> islandBodies, islandContacts, islandJoints, stack;
    std::vector<V2> posC, velV;
    std::vector<float> posA, velW;
    std::vector<VelConstraint> vcs;
    std::vector<PosConstraint> pcs;

    // ---- body state (REF/Dynamics/b2Body.h:635-654)
    void setAwake(Body& b, bool flag) {
        if (flag) {
            if ((b.flags & kAwake) == 0) { b.flags |= kAwake; b.sleepTime = 0.0f; }
        } else {
            b.flags &= ~kAwake;
            b.sleepTime = 0.0f;
            b.v = mk(0.0f, 0.0f);
            b.w = 0.0f;
            b.force = mk(0.0f, 0.0f);
            b.torque = 0.0f;
        }
    }

    // b2Body::ResetMassData, REF/Dynamics/b2Body.cpp:285-352
    void resetMassData(Body& b) {
        b.mass = 0.0f; b.invMass = 0.0f; b.I = 0.0f; b.invI = 0.0f;
        b.localCenter = mk(0.0f, 0.0f);
        if (b.type == kStatic || b.type == kKinematic) {
            b.c0 = b.xf.p; b.c = b.xf.p; b.a0 = b.a;
            return;
        }
        V2 localCenter = mk(0.0f, 0.0f);
        for (int k = (int)b.fixtures.size() - 1; k >= 0; --k) {
            const Fixture& f = fixtures[b.fixtures[k]];
            if (f.density == 0.0f) continue;
            MassData md = shapeMass(&f.shape, f.density);
            b.mass += md.mass;
            localCenter += md.mass * md.center;
            b.I += md.I;
        }
        if (b.mass > 0.0f) {
            b.invMass = 1.0f / b.mass;
            localCenter *= b.invMass;
        } else {
            b.mass = 1.0f;
            b.invMass = 1.0f;
        }
        if (b.I > 0.0f && (b.flags & kFixedRotation) == 0) {
            b.I -= b.mass * dot(localCenter, localCenter);
            b.invI = 1.0f / b.I;
        } else {
            b.I = 0.0f;
            b.invI = 0.0f;
        }
        V2 oldCenter = b.c;
        b.localCenter = localCenter;
        b.c0 = b.c = mul(b.xf, b.localCenter);
        b.v += cross(b.w, b.c - oldCenter);
    }

    int treeAllocate() {
        if (treeFree.empty()) {
            const int old = treeCapacity;
            treeCapacity = old == 0 ? 16 : 2 * old;
            for (int id = treeCapacity - 1; id >= old; --id) treeFree.push_back(id);
        }
        const int id = treeFree.back();
        treeFree.pop_back();
        return id;
    }
    int createProxyId(int fi) {
        const int leaf = treeAllocate();
        if (treeLeaves > 0) (void)treeAllocate();
        ++treeLeaves;
        if ((int)proxyFixture.size() <= leaf) proxyFixture.resize(leaf + 1, -1);
        proxyFixture[leaf] = fi;
        return leaf;
    }
    void destroyProxyId(int leaf) {
        if (treeLeaves > 1) treeFree.push_back(-1);
        treeFree.push_back(leaf);
        --treeLeaves;
        proxyFixture[leaf] = -1;
    }

    // b2World::CreateBody + b2Body::b2Body, REF/Dynamics/b2World.cpp:109-127,
    // REF/Dynamics/b2Body.cpp:25-90
    int createBody(const b2scene::BodyDef& d) {
        Body b;
        b.flags = 0;
        if (d.bullet) b.flags |= kBullet;
        if (d.fixedRotation) b.flags |= kFixedRotation;
        if (d.allowSleep) b.flags |= kAutoSleep;
        if (d.awake) b.flags |= kAwake;
        b.flags |= kActive;
        b.xf.p = mk(d.px, d.py);
        b.xf.q = rotOf(d.angle);
        b.localCenter = mk(0.0f, 0.0f);
        b.c0 = b.xf.p; b.c = b.xf.p;
        b.a0 = d.angle; b.a = d.angle;
        b.alpha0 = 0.0f;
        b.v = mk(d.vx, d.vy);
        b.w = d.w;
        b.linearDamping = d.linearDamping;
        b.angularDamping = d.angularDamping;
        b.gravityScale = d.gravityScale;
        b.force = mk(0.0f, 0.0f);
        b.torque = 0.0f;
        b.sleepTime = 0.0f;
        b.type = d.type;
        if (b.type == kDynamic) { b.mass = 1.0f; b.invMass = 1.0f; }
        else { b.mass = 0.0f; b.invMass = 0.0f; }
        b.I = 0.0f; b.invI = 0.0f;
        b.islandIndex = 0;
        b.edgeHead = -1;
        b.jointHead = -1;
        b.alive = true;
        bodies.push_back(b);
        return (int)bodies.size() - 1;
    }

    // b2Body::CreateFixture + b2Fixture::CreateProxies + b2BroadPhase::CreateProxy,
    // REF/Dynamics/b2Body.cpp:164-203, b2Fixture.cpp:124-139, b2DynamicTree.cpp:112-126
    // owner: the first child fixture of a chain's b2Fixture (-1: this one).
    int createFixture(int bi, const b2g_shape& shape, const b2scene::FixtureDef& d, int owner = -1) {
        Fixture f;
        f.body = bi;
        f.shape = shape;
        f.density = d.density;
        f.friction = d.friction;
        f.restitution = d.restitution;
        f.sensor = d.sensor != 0;
        f.category = (uint16_t)d.categoryBits;
        f.mask = (uint16_t)d.maskBits;
        f.group = (int16_t)d.groupIndex;
        Body& b = bodies[bi];
        f.proxy = (b.flags & kActive) != 0;
        if (f.proxy) {
            f.aabb = shapeAABB(&f.shape, b.xf);
            V2 r = mk(kAabbExtension, kAabbExtension);
            f.fat.lo = f.aabb.lo - r;
            f.fat.hi = f.aabb.hi + r;
        }
        const int fi = (int)fixtures.size();
        f.proxyId = f.proxy ? createProxyId(fi) : -1;
        f.owner = owner < 0 ? fi : owner;
        fixtures.push_back(f);
        if (f.proxy) moveBuffer.push_back(fi);
        b.fixtures.push_back(fi);
        if (f.density > 0.0f) resetMassData(b);
        newFixture = true;
        return fi;
    }

    // b2Body::ShouldCollide, REF/Dynamics/b2Body.cpp:396-417
    bool shouldCollide(int bi, int oi) const {
        const Body& b = bodies[bi];
        const Body& other = bodies[oi];
        if (b.type != kDynamic && other.type != kDynamic) return false;
        for (int e = b.jointHead; e != -1; e = joints[e >> 1].node[e & 1].next) {
            const Joint& j = joints[e >> 1];
            const int o = (e & 1) ? j.bodyA : j.bodyB;
            if (o == oi && !j.collideConnected) return false;
        }
        return true;
    }

    int jointBase = 0;  // index of the loading scene's first joint

    // b2GearJoint ctor, REF/Dynamics/Joints/b2GearJoint.cpp:45-128: the
    // coordinates of its two joints at creation fix the constant.
    void gearCreate(Joint& j, const b2scene::JointDef& d) {
        const Joint& j1 = joints[jointBase + d.joint1];
        const Joint& j2 = joints[jointBase + d.joint2];
        j.typeA = j1.type;
        j.typeB = j2.type;
        j.bodyC = j1.bodyA;
        j.bodyA = j1.bodyB;
        const Body& bA = bodies[j.bodyA];
        const Body& bC = bodies[j.bodyC];
        float coordinateA, coordinateB;
        j.localAnchorC = j1.localAnchorA;
        j.localAnchorA = j1.localAnchorB;
        j.referenceAngleA = j1.referenceAngle;
        if (j.typeA == 0) {
            j.localAxisC = mk(0.0f, 0.0f);
            coordinateA = bA.a - bC.a - j.referenceAngleA;
        } else {
            j.localAxisC = j1.localXAxisA;
            const V2 pC = j.localAnchorC;
            const V2 pA = mulT(bC.xf.q, mul(bA.xf.q, j.localAnchorA) + (bA.xf.p - bC.xf.p));
            coordinateA = dot(pA - pC, j.localAxisC);
        }
        j.bodyD = j2.bodyA;
        j.bodyB = j2.bodyB;
        const Body& bB = bodies[j.bodyB];
        const Body& bD = bodies[j.bodyD];
        j.localAnchorD = j2.localAnchorA;
        j.localAnchorB = j2.localAnchorB;
        j.referenceAngleB = j2.referenceAngle;
        if (j.typeB == 0) {
            j.localAxisD = mk(0.0f, 0.0f);
            coordinateB = bB.a - bD.a - j.referenceAngleB;
        } else {
            j.localAxisD = j2.localXAxisA;
            const V2 pD = j.localAnchorD;
            const V2 pB = mulT(bD.xf.q, mul(bB.xf.q, j.localAnchorB) + (bB.xf.p - bD.xf.p));
            coordinateB = dot(pB - pD, j.localAxisD);
        }
        j.ratio = d.ratio;
        j.constant = coordinateA + j.ratio * coordinateB;
        j.impulse.x = 0.0f;
        j.JvAC = j.JvBD = mk(0.0f, 0.0f);
        j.JwA = j.JwB = j.JwC = j.JwD = j.gearMass = 0.0f;
    }

    // b2GearJoint::InitVelocityConstraints, REF b2GearJoint.cpp:147-240
    void gearInitVelocity(Joint& j) {
        const Body &bA = bodies[j.bodyA], &bB = bodies[j.bodyB], &bC = bodies[j.bodyC], &bD = bodies[j.bodyD];
        const int iA = bA.islandIndex, iB = bB.islandIndex, iC = bC.islandIndex, iD = bD.islandIndex;
        const float mA = bA.invMass, mB = bB.invMass, mC = bC.invMass, mD = bD.invMass;
        const float IA = bA.invI, IB = bB.invI, IC = bC.invI, ID = bD.invI;
        const float aA = posA[iA], aB = posA[iB], aC = posA[iC], aD = posA[iD];
        V2 vA = velV[iA], vB = velV[iB], vC = velV[iC], vD = velV[iD];
        float wA = velW[iA], wB = velW[iB], wC = velW[iC], wD = velW[iD];
        const Rot qA = stepRot(aA), qB = stepRot(aB), qC = stepRot(aC), qD = stepRot(aD);
        j.gearMass = 0.0f;
        if (j.typeA == 0) {
            j.JvAC = mk(0.0f, 0.0f);
            j.JwA = 1.0f;
            j.JwC = 1.0f;
            j.gearMass += IA + IC;
        } else {
            const V2 u = mul(qC, j.localAxisC);
            const V2 rC = mul(qC, j.localAnchorC - bC.localCenter);
            const V2 rA = mul(qA, j.localAnchorA - bA.localCenter);
            j.JvAC = u;
            j.JwC = cross(rC, u);
            j.JwA = cross(rA, u);
            j.gearMass += mC + mA + IC * j.JwC * j.JwC + IA * j.JwA * j.JwA;
        }
        if (j.typeB == 0) {
            j.JvBD = mk(0.0f, 0.0f);
            j.JwB = j.ratio;
            j.JwD = j.ratio;
            j.gearMass += j.ratio * j.ratio * (IB + ID);
        } else {
            const V2 u = mul(qD, j.localAxisD);
            const V2 rD = mul(qD, j.localAnchorD - bD.localCenter);
            const V2 rB = mul(qB, j.localAnchorB - bB.localCenter);
            j.JvBD = j.ratio * u;
            j.JwD = j.ratio * cross(rD, u);
            j.JwB = j.ratio * cross(rB, u);
            j.gearMass += j.ratio * j.ratio * (mD + mB) + ID * j.JwD * j.JwD + IB * j.JwB * j.JwB;
        }
        j.gearMass = j.gearMass > 0.0f ? 1.0f / j.gearMass : 0.0f;
        if (warmStarting) {
            const float im = j.impulse.x;
            vA += (mA * im) * j.JvAC;
            wA += IA * im * j.JwA;
            vB += (mB * im) * j.JvBD;
            wB += IB * im * j.JwB;
            vC -= (mC * im) * j.JvAC;
            wC -= IC * im * j.JwC;
            vD -= (mD * im) * j.JvBD;
            wD -= ID * im * j.JwD;
        } else {
            j.impulse.x = 0.0f;
        }
        velV[iA] = vA; velW[iA] = wA;
        velV[iB] = vB; velW[iB] = wB;
        velV[iC] = vC; velW[iC] = wC;
        velV[iD] = vD; velW[iD] = wD;
    }

    // b2GearJoint::SolveVelocityConstraints, REF b2GearJoint.cpp:242-274
    void gearSolveVelocity(Joint& j) {
        const Body &bA = bodies[j.bodyA], &bB = bodies[j.bodyB], &bC = bodies[j.bodyC], &bD = bodies[j.bodyD];
        const int iA = bA.islandIndex, iB = bB.islandIndex, iC = bC.islandIndex, iD = bD.islandIndex;
        V2 vA = velV[iA], vB = velV[iB], vC = velV[iC], vD = velV[iD];
        float wA = velW[iA], wB = velW[iB], wC = velW[iC], wD = velW[iD];
        float Cdot = dot(j.JvAC, vA - vC) + dot(j.JvBD, vB - vD);
        Cdot += (j.JwA * wA - j.JwC * wC) + (j.JwB * wB - j.JwD * wD);
        const float impulse = -j.gearMass * Cdot;
        j.impulse.x += impulse;
        vA += (bA.invMass * impulse) * j.JvAC;
        wA += bA.invI * impulse * j.JwA;
        vB += (bB.invMass * impulse) * j.JvBD;
        wB += bB.invI * impulse * j.JwB;
        vC -= (bC.invMass * impulse) * j.JvAC;
        wC -= bC.invI * impulse * j.JwC;
        vD -= (bD.invMass * impulse) * j.JvBD;
        wD -= bD.invI * impulse * j.JwD;
        velV[iA] = vA; velW[iA] = wA;
        velV[iB] = vB; velW[iB] = wB;
        velV[iC] = vC; velW[iC] = wC;
        velV[iD] = vD; velW[iD] = wD;
    }

    // b2GearJoint::SolvePositionConstraints, REF b2GearJoint.cpp:276-393
    bool gearSolvePosition(Joint& j) {
        const Body &bA = bodies[j.bodyA], &bB = bodies[j.bodyB], &bC = bodies[j.bodyC], &bD = bodies[j.bodyD];
        const int iA = bA.islandIndex, iB = bB.islandIndex, iC = bC.islandIndex, iD = bD.islandIndex;
        V2 cA = posC[iA], cB = posC[iB], cC = posC[iC], cD = posC[iD];
        float aA = posA[iA], aB = posA[iB], aC = posA[iC], aD = posA[iD];
        const Rot qA = stepRot(aA), qB = stepRot(aB), qC = stepRot(aC), qD = stepRot(aD);
        const float mA = bA.invMass, mB = bB.invMass, mC = bC.invMass, mD = bD.invMass;
        const float IA = bA.invI, IB = bB.invI, IC = bC.invI, ID = bD.invI;
        float coordinateA, coordinateB;
        V2 JvAC, JvBD;
        float JwA, JwB, JwC, JwD;
        float mass = 0.0f;
        if (j.typeA == 0) {
            JvAC = mk(0.0f, 0.0f);
            JwA = 1.0f;
            JwC = 1.0f;
            mass += IA + IC;
            coordinateA = aA - aC - j.referenceAngleA;
        } else {
            const V2 u = mul(qC, j.localAxisC);
            const V2 rC = mul(qC, j.localAnchorC - bC.localCenter);
            const V2 rA = mul(qA, j.localAnchorA - bA.localCenter);
            JvAC = u;
            JwC = cross(rC, u);
            JwA = cross(rA, u);
            mass += mC + mA + IC * JwC * JwC + IA * JwA * JwA;
            const V2 pC = j.localAnchorC - bC.localCenter;
            const V2 pA = mulT(qC, rA + (cA - cC));
            coordinateA = dot(pA - pC, j.localAxisC);
        }
        if (j.typeB == 0) {
            JvBD = mk(0.0f, 0.0f);
            JwB = j.ratio;
            JwD = j.ratio;
            mass += j.ratio * j.ratio * (IB + ID);
            coordinateB = aB - aD - j.referenceAngleB;
        } else {
            const V2 u = mul(qD, j.localAxisD);
            const V2 rD = mul(qD, j.localAnchorD - bD.localCenter);
            const V2 rB = mul(qB, j.localAnchorB - bB.localCenter);
            JvBD = j.ratio * u;
            JwD = j.ratio * cross(rD, u);
            JwB = j.ratio * cross(rB, u);
            mass += j.ratio * j.ratio * (mD + mB) + ID * JwD * JwD + IB * JwB * JwB;
            const V2 pD = j.localAnchorD - bD.localCenter;
            const V2 pB = mulT(qD, rB + (cB - cD));
            coordinateB = dot(pB - pD, j.localAxisD);
        }
        const float C = (coordinateA + j.ratio * coordinateB) - j.constant;
        float impulse = 0.0f;
        if (mass > 0.0f) impulse = -C / mass;
        cA += mA * impulse * JvAC;
        aA += IA * impulse * JwA;
        cB += mB * impulse * JvBD;
        aB += IB * impulse * JwB;
        cC -= mC * impulse * JvAC;
        aC -= IC * impulse * JwC;
        cD -= mD * impulse * JvBD;
        aD -= ID * impulse * JwD;
        posC[iA] = cA; posA[iA] = aA;
        posC[iB] = cB; posA[iB] = aB;
        posC[iC] = cC; posA[iC] = aC;
        posC[iD] = cD; posA[iD] = aD;
        return true;
    }

    // b2World::CreateJoint + b2RevoluteJoint ctor,
    // REF/Dynamics/b2World.cpp:212-268, REF/Dynamics/Joints/b2RevoluteJoint.cpp:47-64
    int createJoint(const b2scene::JointDef& d) {
        Joint j{};
        j.bodyA = d.bodyA; j.bodyB = d.bodyB;
        j.collideConnected = d.collideConnected != 0;
        j.islandFlag = false;
        j.localAnchorA = mk(d.ax, d.ay); j.localAnchorB = mk(d.bx, d.by);
        j.referenceAngle = d.referenceAngle;
        j.impulse = V3{0.0f, 0.0f, 0.0f};
        j.motorImpulse = 0.0f;
        j.lowerAngle = d.lowerAngle; j.upperAngle = d.upperAngle;
        j.maxMotorTorque = d.maxMotorTorque; j.motorSpeed = d.motorSpeed;
        j.enableLimit = d.enableLimit != 0; j.enableMotor = d.enableMotor != 0;
        j.limitState = kInactiveLimit;
        j.type = d.type;
        j.length = d.length; j.frequencyHz = d.frequencyHz; j.dampingRatio = d.dampingRatio;
        j.u = mk(0.0f, 0.0f);
        j.distMass = j.gamma = j.bias = 0.0f;
        j.ropeLength = 0.0f;
        j.maxForce = d.maxForce; j.maxTorque = d.maxTorque; j.correctionFactor = d.correctionFactor;
        j.linearImpulse = mk(0.0f, 0.0f);
        j.angularImpulse = 0.0f;
        j.groundAnchorA = mk(d.gax, d.gay); j.groundAnchorB = mk(d.gbx, d.gby);
        j.lengthB = d.lengthB; j.ratio = d.ratio;
        j.constant = d.length + d.ratio * d.lengthB;   // b2PulleyJoint ctor, REF b2PulleyJoint.cpp:57-75
        j.uA = j.uB = mk(0.0f, 0.0f);
        j.pulleyMass = j.beta = 0.0f;
        j.mouseC = mk(0.0f, 0.0f);
        j.axis = j.perp = j.ax = j.ay = mk(0.0f, 0.0f);
        j.s1 = j.s2 = j.a1 = j.a2 = 0.0f;
        j.sAx = j.sBx = j.sAy = j.sBy = j.wheelMass = j.springMass = j.springImpulse = 0.0f;
        j.localXAxisA = mk(d.axisX, d.axisY);
        if (d.type == 8) normalize(j.localXAxisA);  // b2PrismaticJoint ctor, REF b2PrismaticJoint.cpp:107-109
        j.localYAxisA = cross(1.0f, j.localXAxisA);
        j.bodyC = j.bodyD = -1;
        j.typeA = j.typeB = 0;
        if (d.type == 10) gearCreate(j, d);
        if (d.type == 7) {
            // b2MouseJoint ctor (REF/Dynamics/Joints/b2MouseJoint.cpp:34-53) with
            // the target on the grab point, then SetTarget (:55-62)
            const Body& bB = bodies[d.bodyB];
            const V2 grab = mul(bB.xf, mk(d.bx, d.by));
            j.localAnchorB = mulT(bB.xf, grab);
            j.localAnchorA = mk(d.ax, d.ay);
        }
        const int ji = (int)joints.size();
        j.prev = -1;
        j.next = jointHead;
        joints.push_back(j);
        if (jointHead != -1) joints[jointHead].prev = ji;
        jointHead = ji;
        for (int side = 0; side < 2; ++side) {
            Body& b = bodies[side == 0 ? d.bodyA : d.bodyB];
            Edge& e = joints[ji].node[side];
            e.prev = -1;
            e.next = b.jointHead;
            if (b.jointHead != -1) joints[b.jointHead >> 1].node[b.jointHead & 1].prev = ji * 2 + side;
            b.jointHead = ji * 2 + side;
        }
        if (!j.collideConnected) {
            // flag contacts between the two bodies for filtering (bodyB's list)
            for (int e = bodies[d.bodyB].edgeHead; e != -1; e = edgeRef(e).next)
                if (edgeOther(e) == d.bodyA) contacts[e >> 1].flags |= cFilter;
        }
        return ji;
    }

    // b2World::DestroyJoint, REF/Dynamics/b2World.cpp:267-358: off the world
    // list and both bodies' joint lists, both bodies woken. The slot stays
    // (marked removed) so joint indices keep their creation order.
    void destroyJoint(int ji) {
        Joint& j = joints[ji];
        if (j.removed) return;
        if (j.prev != -1) joints[j.prev].next = j.next;
        if (j.next != -1) joints[j.next].prev = j.prev;
        if (jointHead == ji) jointHead = j.next;
        setAwake(bodies[j.bodyA], true);
        setAwake(bodies[j.bodyB], true);
        for (int side = 0; side < 2; ++side) {
            Edge& e = j.node[side];
            Body& b = bodies[side == 0 ? j.bodyA : j.bodyB];
            if (e.prev != -1) joints[e.prev >> 1].node[e.prev & 1].next = e.next;
            if (e.next != -1) joints[e.next >> 1].node[e.next & 1].prev = e.prev;
            if (b.jointHead == ji * 2 + side) b.jointHead = e.next;
            e.prev = e.next = -1;
        }
        j.removed = true;
        if (!j.collideConnected) {
            for (int e = bodies[j.bodyB].edgeHead; e != -1; e = edgeRef(e).next)
                if (edgeOther(e) == j.bodyA) contacts[e >> 1].flags |= cFilter;
        }
    }

    // b2World::DestroyBody, REF/Dynamics/b2World.cpp:133-215: the body's
    // joints (each DestroyJoint), its contacts (b2ContactManager::Destroy:
    // EndContact if touching, the bodies woken when the manifold had points),
    // then its fixtures newest first, each one's proxies in child order
    // (b2Fixture::DestroyProxies, REF/Dynamics/b2Fixture.cpp:141-152, and
    // b2BroadPhase::DestroyProxy, REF/Collision/b2BroadPhase.cpp:50-55, which
    // takes the proxy out of the move buffer). The slot stays, marked dead,
    // so body and fixture indices keep their creation order.
    void destroyBody(int bi) {
        Body& b = bodies[bi];
        if (!b.alive) return;
        while (b.jointHead != -1) destroyJoint(b.jointHead >> 1);
        while (b.edgeHead != -1) destroyContact(b.edgeHead >> 1);
        int k = (int)b.fixtures.size();
        while (k > 0) {
            const int owner = fixtures[b.fixtures[k - 1]].owner;
            int first = k - 1;
            while (first > 0 && fixtures[b.fixtures[first - 1]].owner == owner) --first;
            for (int i = first; i < k; ++i) {
                Fixture& f = fixtures[b.fixtures[i]];
                if (!f.proxy) continue;
                for (int& m : moveBuffer)
                    if (m == b.fixtures[i]) m = -1;
                destroyProxyId(f.proxyId);
                f.proxy = false;
                f.proxyId = -1;
            }
            k = first;
        }
        b.alive = false;
    }

    // b2Body::SetType, REF/Dynamics/b2Body.cpp:115-172: mass data reset; a
    // body made static stops and re-synchronises its fixtures from its equal
    // sweep ends; woken, forces cleared, its contacts destroyed, its proxies
    // touched (b2BroadPhase::TouchProxy, REF/Collision/b2BroadPhase.cpp:66-69).
    void setType(int bi, int type) {
        Body& b = bodies[bi];
        if (!b.alive || b.type == type) return;
        b.type = type;
        resetMassData(b);
        if (type == kStatic) {
            b.v = mk(0.0f, 0.0f);
            b.w = 0.0f;
            b.a0 = b.a;
            b.c0 = b.c;
            synchronizeFixtures(b);
        }
        setAwake(b, true);
        b.force = mk(0.0f, 0.0f);
        b.torque = 0.0f;
        while (b.edgeHead != -1) destroyContact(b.edgeHead >> 1);
        for (int k = (int)b.fixtures.size() - 1; k >= 0; --k)
            if (fixtures[b.fixtures[k]].proxy) moveBuffer.push_back(b.fixtures[k]);
    }

    // b2Body::SetActive, REF/Dynamics/b2Body.cpp:456-496: proxies created
    // (b2Fixture::CreateProxies at the body's transform, REF/Dynamics/
    // b2Fixture.cpp:124-139) or destroyed and the contacts with them, the
    // fixtures newest first and each one's children in order.
    void setActive(int bi, bool flag) {
        Body& b = bodies[bi];
        if (!b.alive || ((b.flags & kActive) != 0) == flag) return;
        if (flag) b.flags |= kActive;
        else b.flags &= ~kActive;
        int k = (int)b.fixtures.size();
        while (k > 0) {
            const int owner = fixtures[b.fixtures[k - 1]].owner;
            int first = k - 1;
            while (first > 0 && fixtures[b.fixtures[first - 1]].owner == owner) --first;
            for (int i = first; i < k; ++i) {
                const int fi = b.fixtures[i];
                Fixture& f = fixtures[fi];
                if (flag) {
                    f.proxy = true;
                    f.aabb = shapeAABB(&f.shape, b.xf);
                    const V2 r = mk(kAabbExtension, kAabbExtension);
                    f.fat.lo = f.aabb.lo - r;
                    f.fat.hi = f.aabb.hi + r;
                    f.proxyId = createProxyId(fi);
                    moveBuffer.push_back(fi);
                } else if (f.proxy) {
                    for (int& m : moveBuffer)
                        if (m == fi) m = -1;
                    destroyProxyId(f.proxyId);
                    f.proxy = false;
                    f.proxyId = -1;
                }
            }
            k = first;
        }
        if (!flag)
            while (b.edgeHead != -1) destroyContact(b.edgeHead >> 1);
    }

    // b2Body::DestroyFixture, REF/Dynamics/b2Body.cpp:216-281: the fixture
    // (every child of a chain) leaves the body's list, the body's contacts
    // through it are destroyed in the body's edge-list order, then its
    // proxies in child order, and the mass data is reset.
    void destroyFixture(int owner) {
        Fixture& f0 = fixtures[owner];
        Body& b = bodies[f0.body];
        if (!b.alive) return;
        std::vector<int> children;
        for (int fi : b.fixtures)
            if (fixtures[fi].owner == owner) children.push_back(fi);
        if (children.empty()) return;
        auto mine = [&](int fi) { return fixtures[fi].owner == owner; };
        for (int e = b.edgeHead; e != -1;) {
            const int next = edgeRef(e).next;
            const Contact& c = contacts[e >> 1];
            if ((fixtures[c.fA].body == f0.body && mine(c.fA)) || (fixtures[c.fB].body == f0.body && mine(c.fB)))
                destroyContact(e >> 1);
            e = next;
        }
        for (int fi : children) {
            Fixture& f = fixtures[fi];
            if (!f.proxy) continue;
            for (int& m : moveBuffer)
                if (m == fi) m = -1;
            destroyProxyId(f.proxyId);
            f.proxy = false;
            f.proxyId = -1;
        }
        b.fixtures.erase(std::remove_if(b.fixtures.begin(), b.fixtures.end(), mine), b.fixtures.end());
        resetMassData(b);
    }

    // b2RevoluteJoint::EnableLimit, SetLimits, EnableMotor, SetMotorSpeed,
    // SetMaxMotorTorque in that order, REF/Dynamics/Joints/b2RevoluteJoint.cpp:419-471.
    void setJoint(int ji, bool enableLimit, float lower, float upper, bool enableMotor, float speed, float torque) {
        Joint& j = joints[ji];
        if (j.removed) return;
        if (enableLimit != j.enableLimit) {
            setAwake(bodies[j.bodyA], true);
            setAwake(bodies[j.bodyB], true);
            j.enableLimit = enableLimit;
            j.impulse.z = 0.0f;
        }
        if (lower != j.lowerAngle || upper != j.upperAngle) {
            setAwake(bodies[j.bodyA], true);
            setAwake(bodies[j.bodyB], true);
            j.impulse.z = 0.0f;
            j.lowerAngle = lower;
            j.upperAngle = upper;
        }
        setAwake(bodies[j.bodyA], true);
        setAwake(bodies[j.bodyB], true);
        j.enableMotor = enableMotor;
        j.motorSpeed = speed;
        j.maxMotorTorque = torque;
    }
    // b2ContactFilter::ShouldCollide, REF/Dynamics/b2WorldCallbacks.cpp:26-38
    static bool filter(const Fixture& a, const Fixture& b) {
        if (a.group == b.group && a.group != 0) return a.group > 0;
        return (a.mask & b.category) != 0 && (a.category & b.mask) != 0;
    }

    // ---- contact-edge lists (edge id = contact*2 + side)
    Edge& edgeRef(int e) { return contacts[e >> 1].node[e & 1]; }
    int edgeOther(int e) {
        const Contact& c = contacts[e >> 1];
        return (e & 1) == 0 ? fixtures[c.fB].body : fixtures[c.fA].body;
    }
    void edgePrepend(int bi, int e) {
        Edge& n = edgeRef(e);
        n.prev = -1;
        n.next = bodies[bi].edgeHead;
        if (n.next != -1) edgeRef(n.next).prev = e;
        bodies[bi].edgeHead = e;
    }
    void edgeUnlink(int bi, int e) {
        Edge& n = edgeRef(e);
        if (n.prev != -1) edgeRef(n.prev).next = n.next;
        if (n.next != -1) edgeRef(n.next).prev = n.prev;
        if (bodies[bi].edgeHead == e) bodies[bi].edgeHead = n.next;
    }

    // b2ContactManager::AddPair + b2Contact::Create,
    // REF/Dynamics/b2ContactManager.cpp:179-296, b2Contact.cpp:64-95, b2Contact.cpp:136-168
    void addPair(int fa, int fb) {
        const int ba = fixtures[fa].body, bb = fixtures[fb].body;
        if (ba == bb) return;
        for (int e = bodies[bb].edgeHead; e != -1; e = edgeRef(e).next) {
            if (edgeOther(e) == ba) {
                const Contact& c = contacts[e >> 1];
                if (c.fA == fa && c.fB == fb) return;
                if (c.fA == fb && c.fB == fa) return;
            }
        }
        if (!shouldCollide(bb, ba)) return;
        if (!filter(fixtures[fa], fixtures[fb])) return;
        // Registry: circle-circle, polygon-circle, polygon-polygon, edge-circle,
        // edge-polygon are primary; their mirrors swap the fixtures.
        int ta = fixtures[fa].shape.type, tb = fixtures[fb].shape.type;
        int A = fa, B = fb;
        auto primary = [](int x, int y) {
            return (x == B2G_CIRCLE && y == B2G_CIRCLE) || (x == B2G_POLYGON && y == B2G_CIRCLE) ||
                   (x == B2G_POLYGON && y == B2G_POLYGON) || (x == B2G_EDGE && y == B2G_CIRCLE) ||
                   (x == B2G_EDGE && y == B2G_POLYGON);
        };
        if (!primary(ta, tb)) {
            if (!primary(tb, ta)) return;
            A = fb; B = fa;
        }
        int ci;
        if (!freeContacts.empty()) { ci = freeContacts.back(); freeContacts.pop_back(); }
        else { contacts.push_back(Contact()); ci = (int)contacts.size() - 1; }
        Contact& c = contacts[ci];
        c.fA = A; c.fB = B;
        c.flags = cEnabled;
        clearManifold(&c.manifold);
        c.friction = sqrtf(fixtures[A].friction * fixtures[B].friction);
        c.restitution = fixtures[A].restitution > fixtures[B].restitution ? fixtures[A].restitution
                                                                          : fixtures[B].restitution;
        c.tangentSpeed = 0.0f;
        c.toi = 1.0f;
        c.toiCount = 0;
        c.alive = true;
        c.prev = -1;
        c.next = contactHead;
        if (contactHead != -1) contacts[contactHead].prev = ci;
        contactHead = ci;
        const int bodyA = fixtures[A].body, bodyB = fixtures[B].body;
        edgePrepend(bodyA, ci * 2 + 0);
        edgePrepend(bodyB, ci * 2 + 1);
        if (!fixtures[A].sensor && !fixtures[B].sensor) {
            setAwake(bodies[bodyA], true);
            setAwake(bodies[bodyB], true);
        }
        ++contactCount;
    }

    // b2ContactManager::Destroy + b2Contact::Destroy,
    // REF/Dynamics/b2ContactManager.cpp:38-103, b2Contact.cpp:112-134
    void destroyContact(int ci) {
        Contact& c = contacts[ci];
        const int bodyA = fixtures[c.fA].body, bodyB = fixtures[c.fB].body;
        if (c.flags & cTouching) events.end++;
        if (c.prev != -1) contacts[c.prev].next = c.next;
        if (c.next != -1) contacts[c.next].prev = c.prev;
        if (contactHead == ci) contactHead = c.next;
        edgeUnlink(bodyA, ci * 2 + 0);
        edgeUnlink(bodyB, ci * 2 + 1);
        if (c.manifold.pointCount > 0 && !fixtures[c.fA].sensor && !fixtures[c.fB].sensor) {
            setAwake(bodies[bodyA], true);
            setAwake(bodies[bodyB], true);
        }
        c.alive = false;
        freeContacts.push_back(ci);
        --contactCount;
    }

    // b2TestOverlap, REF/Collision/b2Collision.cpp:237-255: GJK distance from
    // an empty cache, with the radii taken off as b2Distance's useRadii does
    // (REF/Collision/b2Distance.cpp:579-602), under 10 epsilon.
    static bool testOverlap(const b2g_shape* a, const Xf& xfA, const b2g_shape* b, const Xf& xfB) {
        const Proxy pA = proxyOf(a), pB = proxyOf(b);
        SimplexCache cache;
        cache.count = 0;
        cache.metric = 0.0f;
        V2 wA, wB;
        float d;
        gjk(&wA, &wB, &d, &cache, pA, xfA, pB, xfB);
        if (d > pA.radius + pB.radius && d > FLT_EPSILON) d -= pA.radius + pB.radius;
        else d = 0.0f;
        return d < 10.0f * FLT_EPSILON;
    }

    // b2Contact::Update, REF/Dynamics/Contacts/b2Contact.cpp:171-251
    void updateContact(int ci) {
        Contact& c = contacts[ci];
        b2g_manifold old = c.manifold;
        c.flags |= cEnabled;
        bool wasTouching = (c.flags & cTouching) != 0;
        Body& bA = bodies[fixtures[c.fA].body];
        Body& bB = bodies[fixtures[c.fB].body];
        if (fixtures[c.fA].sensor || fixtures[c.fB].sensor) {
            // the sensor branch (:189-197): overlap only, no manifold, no wake-up
            const bool touching = testOverlap(&fixtures[c.fA].shape, bA.xf, &fixtures[c.fB].shape, bB.xf);
            c.manifold.pointCount = 0;
            if (touching) c.flags |= cTouching;
            else c.flags &= ~cTouching;
            if (!wasTouching && touching) events.begin++;
            if (wasTouching && !touching) events.end++;
            return;
        }
        collide(&c.manifold, &fixtures[c.fA].shape, bA.xf, &fixtures[c.fB].shape, bB.xf);
        bool touching = c.manifold.pointCount > 0;
        for (int i = 0; i < c.manifold.pointCount; ++i) {
            b2g_manifold_point* mp2 = c.manifold.points + i;
            mp2->normalImpulse = 0.0f;
            mp2->tangentImpulse = 0.0f;
            for (int j = 0; j < old.pointCount; ++j) {
                if (old.points[j].id == mp2->id) {
                    mp2->normalImpulse = old.points[j].normalImpulse;
                    mp2->tangentImpulse = old.points[j].tangentImpulse;
                    break;
                }
            }
        }
        if (touching != wasTouching) {
            setAwake(bA, true);
            setAwake(bB, true);
        }
        if (touching) c.flags |= cTouching;
        else c.flags &= ~cTouching;
        if (!wasTouching && touching) events.begin++;
        if (wasTouching && !touching) events.end++;
    }

    // b2ContactManager::Collide, REF/Dynamics/b2ContactManager.cpp:107-172
    void collideAll() {
        int ci = contactHead;
        while (ci != -1) {
            Contact& c = contacts[ci];
            const Fixture& fA = fixtures[c.fA];
            const Fixture& fB = fixtures[c.fB];
            Body& bA = bodies[fA.body];
            Body& bB = bodies[fB.body];
            if (c.flags & cFilter) {
                if (!shouldCollide(fB.body, fA.body) || !filter(fA, fB)) {
                    int nx = c.next;
                    destroyContact(ci);
                    ci = nx;
                    continue;
                }
                c.flags &= ~cFilter;
            }
            bool activeA = (bA.flags & kAwake) && bA.type != kStatic;
            bool activeB = (bB.flags & kAwake) && bB.type != kStatic;
            if (!activeA && !activeB) { ci = c.next; continue; }
            if (!overlap(fA.fat, fB.fat)) {
                int nx = c.next;
                destroyContact(ci);
                ci = nx;
                continue;
            }
            updateContact(ci);
            ci = contacts[ci].next;
        }
    }

    // b2BroadPhase::UpdatePairs + QueryCallback, REF/Collision/b2BroadPhase.h:188-241,
    // b2BroadPhase.cpp:98-119. The pair set is every proxy whose fat AABB overlaps a
    // moved proxy's fat AABB; a pair is (smaller, larger) tree proxy id, the pairs are
    // sorted by those ids, and the fixture of the smaller id becomes AddPair's A.
    void findNewContacts() {
        std::vector<std::pair<int, int>> pairs;
        const int nf = (int)fixtures.size();
        // Candidate index: proxies sorted by fat lower x; proxies wider than
        // `wide` are checked against every query.
        std::vector<int> order, wideList;
        std::vector<float> widths;
        for (int i = 0; i < nf; ++i) {
            if (!fixtures[i].proxy) continue;
            widths.push_back(fixtures[i].fat.hi.x - fixtures[i].fat.lo.x);
        }
        float wide = 1e30f;
        if (!widths.empty()) {
            std::vector<float> w2 = widths;
            std::nth_element(w2.begin(), w2.begin() + w2.size() / 2, w2.end());
            wide = 8.0f * w2[w2.size() / 2] + 1.0f;
        }
        float maxNarrow = 0.0f;
        for (int i = 0; i < nf; ++i) {
            if (!fixtures[i].proxy) continue;
            float wd = fixtures[i].fat.hi.x - fixtures[i].fat.lo.x;
            if (wd > wide) wideList.push_back(i);
            else { order.push_back(i); maxNarrow = vmax(maxNarrow, wd); }
        }
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return fixtures[a].fat.lo.x < fixtures[b].fat.lo.x || (fixtures[a].fat.lo.x == fixtures[b].fat.lo.x && a < b);
        });
        std::vector<float> keys(order.size());
        for (size_t i = 0; i < order.size(); ++i) keys[i] = fixtures[order[i]].fat.lo.x;
        for (int q : moveBuffer) {
            if (q < 0) continue;
            const AABB& fq = fixtures[q].fat;
            auto consider = [&](int p) {
                if (p == q) return;
                if (!overlap(fq, fixtures[p].fat)) return;
                const int ip = fixtures[p].proxyId, iq = fixtures[q].proxyId;
                pairs.push_back(std::make_pair(std::min(ip, iq), std::max(ip, iq)));
            };
            const float qw = fq.hi.x - fq.lo.x;
            if (qw > wide) {
                for (int p = 0; p < nf; ++p) if (fixtures[p].proxy) consider(p);
                continue;
            }
            const float lo = fq.lo.x - maxNarrow;
            size_t k = std::lower_bound(keys.begin(), keys.end(), lo) - keys.begin();
            for (; k < order.size() && keys[k] <= fq.hi.x; ++k) consider(order[k]);
            for (int p : wideList) consider(p);
        }
        moveBuffer.clear();
        std::sort(pairs.begin(), pairs.end());
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        for (auto& pr : pairs) addPair(proxyFixture[pr.first], proxyFixture[pr.second]);
    }

    // b2DynamicTree::MoveProxy fat-AABB rule, REF/Collision/b2DynamicTree.cpp:133-181
    void moveProxy(int fi, const AABB& aabb, V2 displacement) {
        Fixture& f = fixtures[fi];
        if (contains(f.fat, aabb)) return;
        AABB b = aabb;
        V2 r = mk(kAabbExtension, kAabbExtension);
        b.lo = b.lo - r;
        b.hi = b.hi + r;
        V2 d = kAabbMultiplier * displacement;
        if (d.x < 0.0f) b.lo.x += d.x; else b.hi.x += d.x;
        if (d.y < 0.0f) b.lo.y += d.y; else b.hi.y += d.y;
        f.fat = b;
        moveBuffer.push_back(fi);
    }

    // b2Body::SynchronizeFixtures + b2Fixture::Synchronize,
    // REF/Dynamics/b2Body.cpp:443-454, REF/Dynamics/b2Fixture.cpp:154-176
    void synchronizeFixtures(Body& b) {
        Xf xf1;
        xf1.q = stepRot(b.a0);
        xf1.p = b.c0 - mul(xf1.q, b.localCenter);
        for (int k = (int)b.fixtures.size() - 1; k >= 0; --k) {
            const int fi = b.fixtures[k];
            Fixture& f = fixtures[fi];
            if (!f.proxy) continue;
            AABB a1 = shapeAABB(&f.shape, xf1);
            AABB a2 = shapeAABB(&f.shape, b.xf);
            f.aabb.lo = vmin(a1.lo, a2.lo);
            f.aabb.hi = vmax(a1.hi, a2.hi);
            V2 disp = b.xf.p - xf1.p;
            moveProxy(fi, f.aabb, disp);
        }
    }

    // b2Body::SynchronizeTransform, REF/Dynamics/b2Body.h:836-840
    void synchronizeTransform(Body& b) {
        b.xf.q = stepRot(b.a);
        b.xf.p = b.c - mul(b.xf.q, b.localCenter);
    }

    // b2ContactSolver ctor over the island's contacts, REF/Dynamics/Contacts/b2ContactSolver.cpp:52-140
    void contactSolverInit(float dtRatio, bool warm) {
        const int nc = (int)islandContacts.size();
        vcs.resize(nc); pcs.resize(nc);
        for (int i = 0; i < nc; ++i) {
            Contact& ct = contacts[islandContacts[i]];
            const Fixture& fA = fixtures[ct.fA];
            const Fixture& fB = fixtures[ct.fB];
            const Body& bA = bodies[fA.body];
            const Body& bB = bodies[fB.body];
            const b2g_manifold& m = ct.manifold;
            VelConstraint& vc = vcs[i];
            vc.friction = ct.friction;
            vc.restitution = ct.restitution;
            vc.tangentSpeed = ct.tangentSpeed;
            vc.indexA = bA.islandIndex;
            vc.indexB = bB.islandIndex;
            vc.invMassA = bA.invMass; vc.invMassB = bB.invMass;
            vc.invIA = bA.invI; vc.invIB = bB.invI;
            vc.contactIndex = i;
            vc.pointCount = m.pointCount;
            for (int k = 0; k < 4; ++k) { vc.K[k] = 0.0f; vc.normalMass[k] = 0.0f; }
            PosConstraint& pc = pcs[i];
            pc.indexA = bA.islandIndex; pc.indexB = bB.islandIndex;
            pc.invMassA = bA.invMass; pc.invMassB = bB.invMass;
            pc.localCenterA = bA.localCenter; pc.localCenterB = bB.localCenter;
            pc.invIA = bA.invI; pc.invIB = bB.invI;
            pc.localNormal = v(m.localNormal);
            pc.localPoint = v(m.localPoint);
            pc.pointCount = m.pointCount;
            pc.radiusA = fA.shape.radius;
            pc.radiusB = fB.shape.radius;
            pc.type = m.type;
            for (int j = 0; j < m.pointCount; ++j) {
                VelPoint& vp = vc.points[j];
                if (warm) {
                    vp.normalImpulse = dtRatio * m.points[j].normalImpulse;
                    vp.tangentImpulse = dtRatio * m.points[j].tangentImpulse;
                } else {
                    vp.normalImpulse = 0.0f;
                    vp.tangentImpulse = 0.0f;
                }
                vp.rA = mk(0.0f, 0.0f); vp.rB = mk(0.0f, 0.0f);
                vp.normalMass = 0.0f; vp.tangentMass = 0.0f; vp.velocityBias = 0.0f;
                pc.localPoints[j] = v(m.points[j].localPoint);
            }
        }
    }

    // b2ContactSolver::InitializeVelocityConstraints, REF/Dynamics/Contacts/b2ContactSolver.cpp:143-252
    void initVelocityConstraints() {
        const int nc = (int)islandContacts.size();
        for (int i = 0; i < nc; ++i) {
            VelConstraint& vc = vcs[i];
            PosConstraint& pc = pcs[i];
            const b2g_manifold& m = contacts[islandContacts[vc.contactIndex]].manifold;
            const int iA = vc.indexA, iB = vc.indexB;
            const float mA = vc.invMassA, mB = vc.invMassB, IA = vc.invIA, IB = vc.invIB;
            V2 cA = posC[iA]; float aA = posA[iA];
            V2 vA = velV[iA]; float wA = velW[iA];
            V2 cB = posC[iB]; float aB = posA[iB];
            V2 vB = velV[iB]; float wB = velW[iB];
            Xf xfA, xfB;
            xfA.q = stepRot(aA);
            xfB.q = stepRot(aB);
            xfA.p = cA - mul(xfA.q, pc.localCenterA);
            xfB.p = cB - mul(xfB.q, pc.localCenterB);
            WorldManifold wm;
            worldManifold(&wm, &m, xfA, pc.radiusA, xfB, pc.radiusB);
            vc.normal = wm.normal;
            for (int j = 0; j < vc.pointCount; ++j) {
                VelPoint& vp = vc.points[j];
                vp.rA = wm.points[j] - cA;
                vp.rB = wm.points[j] - cB;
                float rnA = cross(vp.rA, vc.normal);
                float rnB = cross(vp.rB, vc.normal);
                float kNormal = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
                vp.normalMass = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
                V2 tangent = cross(vc.normal, 1.0f);
                float rtA = cross(vp.rA, tangent);
                float rtB = cross(vp.rB, tangent);
                float kTangent = mA + mB + IA * rtA * rtA + IB * rtB * rtB;
                vp.tangentMass = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
                vp.velocityBias = 0.0f;
                float vRel = dot(vc.normal, vB + cross(wB, vp.rB) - vA - cross(wA, vp.rA));
                if (vRel < -kVelocityThreshold) vp.velocityBias = -vc.restitution * vRel;
            }
            if (vc.pointCount == 2) {
                VelPoint& p1 = vc.points[0];
                VelPoint& p2 = vc.points[1];
                float rn1A = cross(p1.rA, vc.normal), rn1B = cross(p1.rB, vc.normal);
                float rn2A = cross(p2.rA, vc.normal), rn2B = cross(p2.rB, vc.normal);
                float k11 = mA + mB + IA * rn1A * rn1A + IB * rn1B * rn1B;
                float k22 = mA + mB + IA * rn2A * rn2A + IB * rn2B * rn2B;
                float k12 = mA + mB + IA * rn1A * rn2A + IB * rn1B * rn2B;
                const float k_maxConditionNumber = 1000.0f;
                if (k11 * k11 < k_maxConditionNumber * (k11 * k22 - k12 * k12)) {
                    vc.K[0] = k11; vc.K[1] = k12; vc.K[2] = k12; vc.K[3] = k22;
                    // b2Mat22::GetInverse (REF b2Math.h:198-212)
                    float a = vc.K[0], b = vc.K[2], c = vc.K[1], d = vc.K[3];
                    float det = a * d - b * c;
                    if (det != 0.0f) det = 1.0f / det;
                    vc.normalMass[0] = det * d;
                    vc.normalMass[2] = -det * b;
                    vc.normalMass[1] = -det * c;
                    vc.normalMass[3] = det * a;
                } else {
                    vc.pointCount = 1;
                }
            }
        }
    }

    // b2Island::Solve and the b2ContactSolver it drives,
    // REF/Dynamics/b2Island.cpp:182-383, REF/Dynamics/Contacts/b2ContactSolver.cpp:52-744
    void solveIsland(float h, float inv_dt, float dtRatio, int velIters, int posIters) {
        (void)inv_dt;
        const int nb = (int)islandBodies.size();
        const int nc = (int)islandContacts.size();
        posC.resize(nb); posA.resize(nb); velV.resize(nb); velW.resize(nb);
        for (int i = 0; i < nb; ++i) {
            Body& b = bodies[islandBodies[i]];
            V2 c = b.c;
            float a = b.a;
            V2 vv = b.v;
            float w = b.w;
            b.c0 = b.c;
            b.a0 = b.a;
            if (b.type == kDynamic) {
                vv += h * (b.gravityScale * gravity + b.invMass * b.force);
                w += h * b.invI * b.torque;
                vv *= 1.0f / (1.0f + h * b.linearDamping);
                w *= 1.0f / (1.0f + h * b.angularDamping);
            }
            posC[i] = c; posA[i] = a; velV[i] = vv; velW[i] = w;
        }
        contactSolverInit(dtRatio, warmStarting);
        initVelocityConstraints();
        // WarmStart (REF :254-292)
        if (warmStarting) {
            for (int i = 0; i < nc; ++i) {
                VelConstraint& vc = vcs[i];
                const int iA = vc.indexA, iB = vc.indexB;
                const float mA = vc.invMassA, IA = vc.invIA, mB = vc.invMassB, IB = vc.invIB;
                V2 vA = velV[iA]; float wA = velW[iA];
                V2 vB = velV[iB]; float wB = velW[iB];
                V2 normal = vc.normal;
                V2 tangent = cross(normal, 1.0f);
                for (int j = 0; j < vc.pointCount; ++j) {
                    VelPoint& vp = vc.points[j];
                    V2 P = vp.normalImpulse * normal + vp.tangentImpulse * tangent;
                    wA -= IA * cross(vp.rA, P);
                    vA -= mA * P;
                    wB += IB * cross(vp.rB, P);
                    vB += mB * P;
                }
                velV[iA] = vA; velW[iA] = wA;
                velV[iB] = vB; velW[iB] = wB;
            }
        }
        // Joint velocity constraints (REF b2Island.cpp:252-255), then the
        // iterations: joints, then contacts (REF b2Island.cpp:261-270)
        for (int ji : islandJoints) jointInitVelocity(joints[ji], h, dtRatio);
        for (int it = 0; it < velIters; ++it) {
            for (int ji : islandJoints) jointSolveVelocity(joints[ji], h);
            for (int i = 0; i < nc; ++i) solveVelocity(vcs[i]);
        }
        // StoreImpulses (REF :599-612)
        for (int i = 0; i < nc; ++i) {
            VelConstraint& vc = vcs[i];
            b2g_manifold& m = contacts[islandContacts[vc.contactIndex]].manifold;
            for (int j = 0; j < vc.pointCount; ++j) {
                m.points[j].normalImpulse = vc.points[j].normalImpulse;
                m.points[j].tangentImpulse = vc.points[j].tangentImpulse;
            }
        }
        // Integrate positions (REF b2Island.cpp:276-306)
        for (int i = 0; i < nb; ++i) {
            V2 c = posC[i]; float a = posA[i];
            V2 vv = velV[i]; float w = velW[i];
            V2 translation = h * vv;
            if (dot(translation, translation) > kMaxTranslationSq) {
                float ratio = kMaxTranslation / length(translation);
                vv *= ratio;
            }
            float rotation = h * w;
            if (rotation * rotation > kMaxRotationSq) {
                float ratio = kMaxRotation / absf(rotation);
                w *= ratio;
            }
            c += h * vv;
            a += h * w;
            posC[i] = c; posA[i] = a; velV[i] = vv; velW[i] = w;
        }
        // Position iterations with early exit (REF b2Island.cpp:308-328)
        bool positionSolved = false;
        for (int it = 0; it < posIters; ++it) {
            const bool contactsOkay = solvePositions();
            bool jointsOkay = true;
            for (int ji : islandJoints) {
                const bool jointOkay = jointSolvePosition(joints[ji]);
                jointsOkay = jointsOkay && jointOkay;
            }
            if (contactsOkay && jointsOkay) { positionSolved = true; break; }
        }
        // Copy back (REF b2Island.cpp:330-339)
        for (int i = 0; i < nb; ++i) {
            Body& b = bodies[islandBodies[i]];
            b.c = posC[i]; b.a = posA[i];
            b.v = velV[i]; b.w = velW[i];
            synchronizeTransform(b);
        }
        // Sleep (REF b2Island.cpp:345-382)
        if (allowSleep) {
            float minSleepTime = FLT_MAX;
            const float linTolSqr = kLinearSleepTol * kLinearSleepTol;
            const float angTolSqr = kAngularSleepTol * kAngularSleepTol;
            for (int i = 0; i < nb; ++i) {
                Body& b = bodies[islandBodies[i]];
                if (b.type == kStatic) continue;
                if ((b.flags & kAutoSleep) == 0 || b.w * b.w > angTolSqr || dot(b.v, b.v) > linTolSqr) {
                    b.sleepTime = 0.0f;
                    minSleepTime = 0.0f;
                } else {
                    b.sleepTime += h;
                    minSleepTime = vmin(minSleepTime, b.sleepTime);
                }
            }
            if (minSleepTime >= kTimeToSleep && positionSolved) {
                for (int i = 0; i < nb; ++i) setAwake(bodies[islandBodies[i]], false);
            }
        }
    }

    // b2RevoluteJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2RevoluteJoint.cpp:66-198
    void jointInitVelocity(Joint& j, float h, float dtRatio) {
        (void)h;
        const Body& bA = bodies[j.bodyA];
        const Body& bB = bodies[j.bodyB];
        j.indexA = bA.islandIndex; j.indexB = bB.islandIndex;
        j.localCenterA = bA.localCenter; j.localCenterB = bB.localCenter;
        j.invMassA = bA.invMass; j.invMassB = bB.invMass;
        j.invIA = bA.invI; j.invIB = bB.invI;
        if (j.type == 1) {
            distanceInitVelocity(j, h, dtRatio);
            return;
        }
        if (j.type == 2) {
            ropeInitVelocity(j, dtRatio);
            return;
        }
        if (j.type == 3) {
            weldInitVelocity(j, h, dtRatio);
            return;
        }
        if (j.type == 4 || j.type == 5) {
            frictionMotorInitVelocity(j, dtRatio);
            return;
        }
        if (j.type == 6) {
            pulleyInitVelocity(j, dtRatio);
            return;
        }
        if (j.type == 7) {
            mouseInitVelocity(j, h, dtRatio);
            return;
        }
        if (j.type == 8) {
            prismaticInitVelocity(j, dtRatio);
            return;
        }
        if (j.type == 10) {
            gearInitVelocity(j);
            return;
        }
        if (j.type == 9) {
            wheelInitVelocity(j, h, dtRatio);
            return;
        }
        const float aA = posA[j.indexA];
        V2 vA = velV[j.indexA]; float wA = velW[j.indexA];
        const float aB = posA[j.indexB];
        V2 vB = velV[j.indexB]; float wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const bool fixedRotation = (iA + iB == 0.0f);
        j.mass.ex.x = mA + mB + j.rA.y * j.rA.y * iA + j.rB.y * j.rB.y * iB;
        j.mass.ey.x = -j.rA.y * j.rA.x * iA - j.rB.y * j.rB.x * iB;
        j.mass.ez.x = -j.rA.y * iA - j.rB.y * iB;
        j.mass.ex.y = j.mass.ey.x;
        j.mass.ey.y = mA + mB + j.rA.x * j.rA.x * iA + j.rB.x * j.rB.x * iB;
        j.mass.ez.y = j.rA.x * iA + j.rB.x * iB;
        j.mass.ex.z = j.mass.ez.x;
        j.mass.ey.z = j.mass.ez.y;
        j.mass.ez.z = iA + iB;
        j.motorMass = iA + iB;
        if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        if (!j.enableMotor || fixedRotation) j.motorImpulse = 0.0f;
        if (j.enableLimit && !fixedRotation) {
            const float jointAngle = aB - aA - j.referenceAngle;
            if (absf(j.upperAngle - j.lowerAngle) < 2.0f * kAngularSlop) {
                j.limitState = kEqualLimits;
            } else if (jointAngle <= j.lowerAngle) {
                if (j.limitState != kAtLowerLimit) j.impulse.z = 0.0f;
                j.limitState = kAtLowerLimit;
            } else if (jointAngle >= j.upperAngle) {
                if (j.limitState != kAtUpperLimit) j.impulse.z = 0.0f;
                j.limitState = kAtUpperLimit;
            } else {
                j.limitState = kInactiveLimit;
                j.impulse.z = 0.0f;
            }
        } else {
            j.limitState = kInactiveLimit;
        }
        if (warmStarting) {
            j.impulse.x *= dtRatio; j.impulse.y *= dtRatio; j.impulse.z *= dtRatio;
            j.motorImpulse *= dtRatio;
            const V2 P = mk(j.impulse.x, j.impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.motorImpulse + j.impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.motorImpulse + j.impulse.z);
        } else {
            j.impulse = V3{0.0f, 0.0f, 0.0f};
            j.motorImpulse = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2RevoluteJoint::SolveVelocityConstraints, REF b2RevoluteJoint.cpp:200-292
    void jointSolveVelocity(Joint& j, float h) {
        if (j.type == 1) {
            distanceSolveVelocity(j);
            return;
        }
        if (j.type == 2) {
            ropeSolveVelocity(j, 1.0f / h);
            return;
        }
        if (j.type == 3) {
            weldSolveVelocity(j);
            return;
        }
        if (j.type == 4 || j.type == 5) {
            frictionMotorSolveVelocity(j, h, 1.0f / h);
            return;
        }
        if (j.type == 6) {
            pulleySolveVelocity(j);
            return;
        }
        if (j.type == 7) {
            mouseSolveVelocity(j, h);
            return;
        }
        if (j.type == 8) {
            prismaticSolveVelocity(j, h);
            return;
        }
        if (j.type == 10) {
            gearSolveVelocity(j);
            return;
        }
        if (j.type == 9) {
            wheelSolveVelocity(j, h);
            return;
        }
        V2 vA = velV[j.indexA]; float wA = velW[j.indexA];
        V2 vB = velV[j.indexB]; float wB = velW[j.indexB];
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const bool fixedRotation = (iA + iB == 0.0f);
        if (j.enableMotor && j.limitState != kEqualLimits && !fixedRotation) {
            const float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = h * j.maxMotorTorque;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        if (j.enableLimit && j.limitState != kInactiveLimit && !fixedRotation) {
            const V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            const float Cdot2 = wB - wA;
            const V3 Cdot = V3{Cdot1.x, Cdot1.y, Cdot2};
            V3 impulse = solve33(j.mass, Cdot);
            impulse = V3{-impulse.x, -impulse.y, -impulse.z};
            if (j.limitState == kEqualLimits) {
                j.impulse.x += impulse.x; j.impulse.y += impulse.y; j.impulse.z += impulse.z;
            } else if (j.limitState == kAtLowerLimit || j.limitState == kAtUpperLimit) {
                const float newImpulse = j.impulse.z + impulse.z;
                const bool reduce = j.limitState == kAtLowerLimit ? newImpulse < 0.0f : newImpulse > 0.0f;
                if (reduce) {
                    const V2 rhs = -Cdot1 + j.impulse.z * mk(j.mass.ez.x, j.mass.ez.y);
                    const V2 reduced = solve22(j.mass, rhs);
                    impulse.x = reduced.x;
                    impulse.y = reduced.y;
                    impulse.z = -j.impulse.z;
                    j.impulse.x += reduced.x;
                    j.impulse.y += reduced.y;
                    j.impulse.z = 0.0f;
                } else {
                    j.impulse.x += impulse.x; j.impulse.y += impulse.y; j.impulse.z += impulse.z;
                }
            }
            const V2 P = mk(impulse.x, impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + impulse.z);
        } else {
            const V2 Cdot = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            const V2 impulse = solve22(j.mass, -Cdot);
            j.impulse.x += impulse.x;
            j.impulse.y += impulse.y;
            vA -= mA * impulse;
            wA -= iA * cross(j.rA, impulse);
            vB += mB * impulse;
            wB += iB * cross(j.rB, impulse);
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2RevoluteJoint::SolvePositionConstraints, REF b2RevoluteJoint.cpp:294-378
    // b2DistanceJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2DistanceJoint.cpp:63-154
    void distanceInitVelocity(Joint& j, float h, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.u = cB + j.rB - cA - j.rA;
        const float len = length(j.u);
        if (len > kLinearSlop) j.u *= 1.0f / len;
        else j.u = mk(0.0f, 0.0f);
        const float crAu = cross(j.rA, j.u), crBu = cross(j.rB, j.u);
        float invMass = j.invMassA + j.invIA * crAu * crAu + j.invMassB + j.invIB * crBu * crBu;
        j.distMass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (j.frequencyHz > 0.0f) {
            const float C = len - j.length;
            const float omega = 2.0f * kPi * j.frequencyHz;
            const float d = 2.0f * j.distMass * j.dampingRatio * omega;
            const float k = j.distMass * omega * omega;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invMass += j.gamma;
            j.distMass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        } else {
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        if (warmStarting) {
            j.impulse.x *= dtRatio;
            const V2 P = j.impulse.x * j.u;
            vA -= j.invMassA * P;
            wA -= j.invIA * cross(j.rA, P);
            vB += j.invMassB * P;
            wB += j.invIB * cross(j.rB, P);
        } else {
            j.impulse.x = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2DistanceJoint::SolveVelocityConstraints, REF b2DistanceJoint.cpp:156-181
    void distanceSolveVelocity(Joint& j) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const V2 vpA = vA + cross(wA, j.rA);
        const V2 vpB = vB + cross(wB, j.rB);
        const float Cdot = dot(j.u, vpB - vpA);
        const float impulse = -j.distMass * (Cdot + j.bias + j.gamma * j.impulse.x);
        j.impulse.x += impulse;
        const V2 P = impulse * j.u;
        vA -= j.invMassA * P;
        wA -= j.invIA * cross(j.rA, P);
        vB += j.invMassB * P;
        wB += j.invIB * cross(j.rB, P);
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2DistanceJoint::SolvePositionConstraints, REF b2DistanceJoint.cpp:183-221
    bool distanceSolvePosition(Joint& j) {
        if (j.frequencyHz > 0.0f) return true;
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 u = cB + rB - cA - rA;
        const float len = normalize(u);
        float C = len - j.length;
        C = clampf(C, -kMaxLinearCorrection, kMaxLinearCorrection);
        const float impulse = -j.distMass * C;
        const V2 P = impulse * u;
        cA -= j.invMassA * P;
        aA -= j.invIA * cross(rA, P);
        cB += j.invMassB * P;
        aB += j.invIB * cross(rB, P);
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return absf(C) < kLinearSlop;
    }

    // b2RopeJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2RopeJoint.cpp:48-126
    void ropeInitVelocity(Joint& j, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.u = cB + j.rB - cA - j.rA;
        j.ropeLength = length(j.u);
        const float C = j.ropeLength - j.length;
        j.limitState = C > 0.0f ? kAtUpperLimit : kInactiveLimit;
        if (j.ropeLength > kLinearSlop) {
            j.u *= 1.0f / j.ropeLength;
        } else {
            j.u = mk(0.0f, 0.0f);
            j.distMass = 0.0f;
            j.impulse.x = 0.0f;
            return;
        }
        const float crA = cross(j.rA, j.u), crB = cross(j.rB, j.u);
        const float invMass = j.invMassA + j.invIA * crA * crA + j.invMassB + j.invIB * crB * crB;
        j.distMass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
        if (warmStarting) {
            j.impulse.x *= dtRatio;
            const V2 P = j.impulse.x * j.u;
            vA -= j.invMassA * P;
            wA -= j.invIA * cross(j.rA, P);
            vB += j.invMassB * P;
            wB += j.invIB * cross(j.rB, P);
        } else {
            j.impulse.x = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2RopeJoint::SolveVelocityConstraints, REF b2RopeJoint.cpp:128-162
    void ropeSolveVelocity(Joint& j, float invDt) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const V2 vpA = vA + cross(wA, j.rA);
        const V2 vpB = vB + cross(wB, j.rB);
        const float C = j.ropeLength - j.length;
        float Cdot = dot(j.u, vpB - vpA);
        if (C < 0.0f) Cdot += invDt * C;
        float impulse = -j.distMass * Cdot;
        const float oldImpulse = j.impulse.x;
        j.impulse.x = vmin(0.0f, j.impulse.x + impulse);
        impulse = j.impulse.x - oldImpulse;
        const V2 P = impulse * j.u;
        vA -= j.invMassA * P;
        wA -= j.invIA * cross(j.rA, P);
        vB += j.invMassB * P;
        wB += j.invIB * cross(j.rB, P);
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2RopeJoint::SolvePositionConstraints, REF b2RopeJoint.cpp:164-196
    bool ropeSolvePosition(Joint& j) {
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 u = cB + rB - cA - rA;
        const float len = normalize(u);
        float C = len - j.length;
        C = clampf(C, 0.0f, kMaxLinearCorrection);
        const float impulse = -j.distMass * C;
        const V2 P = impulse * u;
        cA -= j.invMassA * P;
        aA -= j.invIA * cross(rA, P);
        cB += j.invMassB * P;
        aB += j.invIB * cross(rB, P);
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return len - j.length < kLinearSlop;
    }

    // b2FrictionJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2FrictionJoint.cpp:59-132,
    // and b2MotorJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2MotorJoint.cpp:63-141
    void frictionMotorInitVelocity(Joint& j, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        if (j.type == 5) {
            j.rA = mul(qA, -j.localCenterA);
            j.rB = mul(qB, -j.localCenterB);
        } else {
            j.rA = mul(qA, j.localAnchorA - j.localCenterA);
            j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        }
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        Mat22 K;
        K.ex.x = mA + mB + iA * j.rA.y * j.rA.y + iB * j.rB.y * j.rB.y;
        K.ex.y = -iA * j.rA.x * j.rA.y - iB * j.rB.x * j.rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = mA + mB + iA * j.rA.x * j.rA.x + iB * j.rB.x * j.rB.x;
        j.linearMass = inverse(K);
        j.angularMass = iA + iB;
        if (j.angularMass > 0.0f) j.angularMass = 1.0f / j.angularMass;
        if (j.type == 5) {
            j.linearError = cB + j.rB - cA - j.rA - mul(qA, j.localAnchorA);
            j.angularError = aB - aA - j.referenceAngle;
        }
        if (warmStarting) {
            j.linearImpulse *= dtRatio;
            j.angularImpulse *= dtRatio;
            const V2 P = j.linearImpulse;
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.angularImpulse);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.angularImpulse);
        } else {
            j.linearImpulse = mk(0.0f, 0.0f);
            j.angularImpulse = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2FrictionJoint::SolveVelocityConstraints, REF b2FrictionJoint.cpp:134-189,
    // and b2MotorJoint::SolveVelocityConstraints, REF b2MotorJoint.cpp:143-199
    void frictionMotorSolveVelocity(Joint& j, float h, float invH) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const bool motor = j.type == 5;
        {
            const float Cdot = motor ? wB - wA + invH * j.correctionFactor * j.angularError : wB - wA;
            float impulse = -j.angularMass * Cdot;
            const float oldImpulse = j.angularImpulse;
            const float maxImpulse = h * j.maxTorque;
            j.angularImpulse = clampf(j.angularImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.angularImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            V2 Cdot = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            if (motor) Cdot = Cdot + invH * j.correctionFactor * j.linearError;
            V2 impulse = -mul(j.linearMass, Cdot);
            const V2 oldImpulse = j.linearImpulse;
            j.linearImpulse += impulse;
            const float maxImpulse = h * j.maxForce;
            if (lengthSq(j.linearImpulse) > maxImpulse * maxImpulse) {
                normalize(j.linearImpulse);
                j.linearImpulse *= maxImpulse;
            }
            impulse = j.linearImpulse - oldImpulse;
            vA -= mA * impulse;
            wA -= iA * cross(j.rA, impulse);
            vB += mB * impulse;
            wB += iB * cross(j.rB, impulse);
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PulleyJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2PulleyJoint.cpp:76-164
    void pulleyInitVelocity(Joint& j, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        j.uA = cA + j.rA - j.groundAnchorA;
        j.uB = cB + j.rB - j.groundAnchorB;
        const float lengthA = length(j.uA), lengthB = length(j.uB);
        if (lengthA > 10.0f * kLinearSlop) j.uA *= 1.0f / lengthA;
        else j.uA = mk(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) j.uB *= 1.0f / lengthB;
        else j.uB = mk(0.0f, 0.0f);
        const float ruA = cross(j.rA, j.uA), ruB = cross(j.rB, j.uB);
        const float mA = j.invMassA + j.invIA * ruA * ruA;
        const float mB = j.invMassB + j.invIB * ruB * ruB;
        j.pulleyMass = mA + j.ratio * j.ratio * mB;
        if (j.pulleyMass > 0.0f) j.pulleyMass = 1.0f / j.pulleyMass;
        if (warmStarting) {
            j.impulse.x *= dtRatio;
            const V2 PA = -(j.impulse.x) * j.uA;
            const V2 PB = (-j.ratio * j.impulse.x) * j.uB;
            vA += j.invMassA * PA;
            wA += j.invIA * cross(j.rA, PA);
            vB += j.invMassB * PB;
            wB += j.invIB * cross(j.rB, PB);
        } else {
            j.impulse.x = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PulleyJoint::SolveVelocityConstraints, REF b2PulleyJoint.cpp:166-191
    void pulleySolveVelocity(Joint& j) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const V2 vpA = vA + cross(wA, j.rA);
        const V2 vpB = vB + cross(wB, j.rB);
        const float Cdot = -dot(j.uA, vpA) - j.ratio * dot(j.uB, vpB);
        const float impulse = -j.pulleyMass * Cdot;
        j.impulse.x += impulse;
        const V2 PA = -impulse * j.uA;
        const V2 PB = -j.ratio * impulse * j.uB;
        vA += j.invMassA * PA;
        wA += j.invIA * cross(j.rA, PA);
        vB += j.invMassB * PB;
        wB += j.invIB * cross(j.rB, PB);
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PulleyJoint::SolvePositionConstraints, REF b2PulleyJoint.cpp:193-263
    bool pulleySolvePosition(Joint& j) {
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        V2 uA = cA + rA - j.groundAnchorA;
        V2 uB = cB + rB - j.groundAnchorB;
        const float lengthA = length(uA), lengthB = length(uB);
        if (lengthA > 10.0f * kLinearSlop) uA *= 1.0f / lengthA;
        else uA = mk(0.0f, 0.0f);
        if (lengthB > 10.0f * kLinearSlop) uB *= 1.0f / lengthB;
        else uB = mk(0.0f, 0.0f);
        const float ruA = cross(rA, uA), ruB = cross(rB, uB);
        const float mA = j.invMassA + j.invIA * ruA * ruA;
        const float mB = j.invMassB + j.invIB * ruB * ruB;
        float mass = mA + j.ratio * j.ratio * mB;
        if (mass > 0.0f) mass = 1.0f / mass;
        const float C = j.constant - lengthA - j.ratio * lengthB;
        const float linearError = absf(C);
        const float impulse = -mass * C;
        const V2 PA = -impulse * uA;
        const V2 PB = -j.ratio * impulse * uB;
        cA += j.invMassA * PA;
        aA += j.invIA * cross(rA, PA);
        cB += j.invMassB * PB;
        aB += j.invIB * cross(rB, PB);
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return linearError < kLinearSlop;
    }

    // b2MouseJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2MouseJoint.cpp:99-169
    void mouseInitVelocity(Joint& j, float h, float dtRatio) {
        const V2 cB = posC[j.indexB];
        const float aB = posA[j.indexB];
        V2 vB = velV[j.indexB];
        float wB = velW[j.indexB];
        const Rot qB = stepRot(aB);
        const float mass = bodies[j.bodyB].mass;
        const float omega = 2.0f * kPi * j.frequencyHz;
        const float d = 2.0f * mass * j.dampingRatio * omega;
        const float k = mass * (omega * omega);
        j.gamma = h * (d + h * k);
        if (j.gamma != 0.0f) j.gamma = 1.0f / j.gamma;
        j.beta = h * k * j.gamma;
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        Mat22 K;
        K.ex.x = j.invMassB + j.invIB * j.rB.y * j.rB.y + j.gamma;
        K.ex.y = -j.invIB * j.rB.x * j.rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = j.invMassB + j.invIB * j.rB.x * j.rB.x + j.gamma;
        j.linearMass = inverse(K);
        j.mouseC = cB + j.rB - j.localAnchorA;
        j.mouseC *= j.beta;
        wB *= 0.98f;
        if (warmStarting) {
            j.linearImpulse *= dtRatio;
            vB += j.invMassB * j.linearImpulse;
            wB += j.invIB * cross(j.rB, j.linearImpulse);
        } else {
            j.linearImpulse = mk(0.0f, 0.0f);
        }
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2MouseJoint::SolveVelocityConstraints, REF b2MouseJoint.cpp:171-194
    void mouseSolveVelocity(Joint& j, float h) {
        V2 vB = velV[j.indexB];
        float wB = velW[j.indexB];
        const V2 Cdot = vB + cross(wB, j.rB);
        V2 impulse = mul(j.linearMass, -(Cdot + j.mouseC + j.gamma * j.linearImpulse));
        const V2 oldImpulse = j.linearImpulse;
        j.linearImpulse += impulse;
        const float maxImpulse = h * j.maxForce;
        if (lengthSq(j.linearImpulse) > maxImpulse * maxImpulse) j.linearImpulse *= maxImpulse / length(j.linearImpulse);
        impulse = j.linearImpulse - oldImpulse;
        vB += j.invMassB * impulse;
        wB += j.invIB * cross(j.rB, impulse);
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PrismaticJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2PrismaticJoint.cpp:128-266
    void prismaticInitVelocity(Joint& j, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        const V2 d = (cB - cA) + rB - rA;
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        j.axis = mul(qA, j.localXAxisA);
        j.a1 = cross(d + rA, j.axis);
        j.a2 = cross(rB, j.axis);
        j.motorMass = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
        if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        j.perp = mul(qA, j.localYAxisA);
        j.s1 = cross(d + rA, j.perp);
        j.s2 = cross(rB, j.perp);
        const float k11 = mA + mB + iA * j.s1 * j.s1 + iB * j.s2 * j.s2;
        const float k12 = iA * j.s1 + iB * j.s2;
        const float k13 = iA * j.s1 * j.a1 + iB * j.s2 * j.a2;
        float k22 = iA + iB;
        if (k22 == 0.0f) k22 = 1.0f;
        const float k23 = iA * j.a1 + iB * j.a2;
        const float k33 = mA + mB + iA * j.a1 * j.a1 + iB * j.a2 * j.a2;
        j.mass.ex = V3{k11, k12, k13};
        j.mass.ey = V3{k12, k22, k23};
        j.mass.ez = V3{k13, k23, k33};
        if (j.enableLimit) {
            const float jointTranslation = dot(j.axis, d);
            if (absf(j.upperAngle - j.lowerAngle) < 2.0f * kLinearSlop) {
                j.limitState = kEqualLimits;
            } else if (jointTranslation <= j.lowerAngle) {
                if (j.limitState != kAtLowerLimit) {
                    j.limitState = kAtLowerLimit;
                    j.impulse.z = 0.0f;
                }
            } else if (jointTranslation >= j.upperAngle) {
                if (j.limitState != kAtUpperLimit) {
                    j.limitState = kAtUpperLimit;
                    j.impulse.z = 0.0f;
                }
            } else {
                j.limitState = kInactiveLimit;
                j.impulse.z = 0.0f;
            }
        } else {
            j.limitState = kInactiveLimit;
            j.impulse.z = 0.0f;
        }
        if (!j.enableMotor) j.motorImpulse = 0.0f;
        if (warmStarting) {
            j.impulse.x *= dtRatio; j.impulse.y *= dtRatio; j.impulse.z *= dtRatio;
            j.motorImpulse *= dtRatio;
            const V2 P = j.impulse.x * j.perp + (j.motorImpulse + j.impulse.z) * j.axis;
            const float LA = j.impulse.x * j.s1 + j.impulse.y + (j.motorImpulse + j.impulse.z) * j.a1;
            const float LB = j.impulse.x * j.s2 + j.impulse.y + (j.motorImpulse + j.impulse.z) * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        } else {
            j.impulse = V3{0.0f, 0.0f, 0.0f};
            j.motorImpulse = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PrismaticJoint::SolveVelocityConstraints, REF b2PrismaticJoint.cpp:268-363
    void prismaticSolveVelocity(Joint& j, float h) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        if (j.enableMotor && j.limitState != kEqualLimits) {
            const float Cdot = dot(j.axis, vB - vA) + j.a2 * wB - j.a1 * wA;
            float impulse = j.motorMass * (j.motorSpeed - Cdot);
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = h * j.maxMotorTorque;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            const V2 P = impulse * j.axis;
            const float LA = impulse * j.a1, LB = impulse * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        V2 Cdot1;
        Cdot1.x = dot(j.perp, vB - vA) + j.s2 * wB - j.s1 * wA;
        Cdot1.y = wB - wA;
        if (j.enableLimit && j.limitState != kInactiveLimit) {
            const float Cdot2 = dot(j.axis, vB - vA) + j.a2 * wB - j.a1 * wA;
            const V3 f1 = j.impulse;
            const V3 df0 = solve33(j.mass, V3{-Cdot1.x, -Cdot1.y, -Cdot2});
            j.impulse = V3{j.impulse.x + df0.x, j.impulse.y + df0.y, j.impulse.z + df0.z};
            if (j.limitState == kAtLowerLimit) j.impulse.z = vmax(j.impulse.z, 0.0f);
            else if (j.limitState == kAtUpperLimit) j.impulse.z = vmin(j.impulse.z, 0.0f);
            const V2 b = -Cdot1 - (j.impulse.z - f1.z) * mk(j.mass.ez.x, j.mass.ez.y);
            const V2 f2r = solve22(j.mass, b) + mk(f1.x, f1.y);
            j.impulse.x = f2r.x;
            j.impulse.y = f2r.y;
            const V3 df = V3{j.impulse.x - f1.x, j.impulse.y - f1.y, j.impulse.z - f1.z};
            const V2 P = df.x * j.perp + df.z * j.axis;
            const float LA = df.x * j.s1 + df.y + df.z * j.a1;
            const float LB = df.x * j.s2 + df.y + df.z * j.a2;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        } else {
            const V2 df = solve22(j.mass, -Cdot1);
            j.impulse.x += df.x;
            j.impulse.y += df.y;
            const V2 P = df.x * j.perp;
            const float LA = df.x * j.s1 + df.y;
            const float LB = df.x * j.s2 + df.y;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2PrismaticJoint::SolvePositionConstraints, REF b2PrismaticJoint.cpp:365-488
    bool prismaticSolvePosition(Joint& j) {
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        const V2 d = cB + rB - cA - rA;
        const V2 axis = mul(qA, j.localXAxisA);
        const float a1 = cross(d + rA, axis), a2 = cross(rB, axis);
        const V2 perp = mul(qA, j.localYAxisA);
        const float s1 = cross(d + rA, perp), s2 = cross(rB, perp);
        V3 impulse;
        V2 C1;
        C1.x = dot(perp, d);
        C1.y = aB - aA - j.referenceAngle;
        float linearError = absf(C1.x);
        const float angularError = absf(C1.y);
        bool active = false;
        float C2 = 0.0f;
        if (j.enableLimit) {
            const float translation = dot(axis, d);
            if (absf(j.upperAngle - j.lowerAngle) < 2.0f * kLinearSlop) {
                C2 = clampf(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
                linearError = vmax(linearError, absf(translation));
                active = true;
            } else if (translation <= j.lowerAngle) {
                C2 = clampf(translation - j.lowerAngle + kLinearSlop, -kMaxLinearCorrection, 0.0f);
                linearError = vmax(linearError, j.lowerAngle - translation);
                active = true;
            } else if (translation >= j.upperAngle) {
                C2 = clampf(translation - j.upperAngle - kLinearSlop, 0.0f, kMaxLinearCorrection);
                linearError = vmax(linearError, translation - j.upperAngle);
                active = true;
            }
        }
        if (active) {
            const float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            const float k12 = iA * s1 + iB * s2;
            const float k13 = iA * s1 * a1 + iB * s2 * a2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            const float k23 = iA * a1 + iB * a2;
            const float k33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
            Mat33 K;
            K.ex = V3{k11, k12, k13};
            K.ey = V3{k12, k22, k23};
            K.ez = V3{k13, k23, k33};
            impulse = solve33(K, V3{-C1.x, -C1.y, -C2});
        } else {
            const float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
            const float k12 = iA * s1 + iB * s2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            // b2Mat22::Solve, REF/Common/b2Math.h:242-254
            float det = k11 * k22 - k12 * k12;
            if (det != 0.0f) det = 1.0f / det;
            const V2 b = -C1;
            impulse = V3{det * (k22 * b.x - k12 * b.y), det * (k11 * b.y - k12 * b.x), 0.0f};
        }
        const V2 P = impulse.x * perp + impulse.z * axis;
        const float LA = impulse.x * s1 + impulse.y + impulse.z * a1;
        const float LB = impulse.x * s2 + impulse.y + impulse.z * a2;
        cA -= mA * P;
        aA -= iA * LA;
        cB += mB * P;
        aB += iB * LB;
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return linearError <= kLinearSlop && angularError <= kAngularSlop;
    }

    // b2WheelJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2WheelJoint.cpp:79-216
    void wheelInitVelocity(Joint& j, float h, float dtRatio) {
        const V2 cA = posC[j.indexA], cB = posC[j.indexB];
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        const V2 d = cB + rB - cA - rA;
        j.ay = mul(qA, j.localYAxisA);
        j.sAy = cross(d + rA, j.ay);
        j.sBy = cross(rB, j.ay);
        j.wheelMass = mA + mB + iA * j.sAy * j.sAy + iB * j.sBy * j.sBy;
        if (j.wheelMass > 0.0f) j.wheelMass = 1.0f / j.wheelMass;
        j.springMass = 0.0f;
        j.bias = 0.0f;
        j.gamma = 0.0f;
        if (j.frequencyHz > 0.0f) {
            j.ax = mul(qA, j.localXAxisA);
            j.sAx = cross(d + rA, j.ax);
            j.sBx = cross(rB, j.ax);
            const float invMass = mA + mB + iA * j.sAx * j.sAx + iB * j.sBx * j.sBx;
            if (invMass > 0.0f) {
                j.springMass = 1.0f / invMass;
                const float C = dot(d, j.ax);
                const float omega = 2.0f * kPi * j.frequencyHz;
                const float dd = 2.0f * j.springMass * j.dampingRatio * omega;
                const float k = j.springMass * omega * omega;
                j.gamma = h * (dd + h * k);
                if (j.gamma > 0.0f) j.gamma = 1.0f / j.gamma;
                j.bias = C * h * k * j.gamma;
                j.springMass = invMass + j.gamma;
                if (j.springMass > 0.0f) j.springMass = 1.0f / j.springMass;
            }
        } else {
            j.springImpulse = 0.0f;
        }
        if (j.enableMotor) {
            j.motorMass = iA + iB;
            if (j.motorMass > 0.0f) j.motorMass = 1.0f / j.motorMass;
        } else {
            j.motorMass = 0.0f;
            j.motorImpulse = 0.0f;
        }
        if (warmStarting) {
            j.impulse.x *= dtRatio;
            j.springImpulse *= dtRatio;
            j.motorImpulse *= dtRatio;
            const V2 P = j.impulse.x * j.ay + j.springImpulse * j.ax;
            const float LA = j.impulse.x * j.sAy + j.springImpulse * j.sAx + j.motorImpulse;
            const float LB = j.impulse.x * j.sBy + j.springImpulse * j.sBx + j.motorImpulse;
            vA -= j.invMassA * P;
            wA -= j.invIA * LA;
            vB += j.invMassB * P;
            wB += j.invIB * LB;
        } else {
            j.impulse.x = 0.0f;
            j.springImpulse = 0.0f;
            j.motorImpulse = 0.0f;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2WheelJoint::SolveVelocityConstraints, REF b2WheelJoint.cpp:218-280
    void wheelSolveVelocity(Joint& j, float h) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        {
            const float Cdot = dot(j.ax, vB - vA) + j.sBx * wB - j.sAx * wA;
            const float impulse = -j.springMass * (Cdot + j.bias + j.gamma * j.springImpulse);
            j.springImpulse += impulse;
            const V2 P = impulse * j.ax;
            const float LA = impulse * j.sAx, LB = impulse * j.sBx;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        {
            const float Cdot = wB - wA - j.motorSpeed;
            float impulse = -j.motorMass * Cdot;
            const float oldImpulse = j.motorImpulse;
            const float maxImpulse = h * j.maxMotorTorque;
            j.motorImpulse = clampf(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            wA -= iA * impulse;
            wB += iB * impulse;
        }
        {
            const float Cdot = dot(j.ay, vB - vA) + j.sBy * wB - j.sAy * wA;
            const float impulse = -j.wheelMass * Cdot;
            j.impulse.x += impulse;
            const V2 P = impulse * j.ay;
            const float LA = impulse * j.sAy, LB = impulse * j.sBy;
            vA -= mA * P;
            wA -= iA * LA;
            vB += mB * P;
            wB += iB * LB;
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2WheelJoint::SolvePositionConstraints, REF b2WheelJoint.cpp:282-327
    // (its effective mass uses the sAy, sBy of the velocity pass)
    bool wheelSolvePosition(Joint& j) {
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        const V2 d = (cB - cA) + rB - rA;
        const V2 ay = mul(qA, j.localYAxisA);
        const float sAy = cross(d + rA, ay), sBy = cross(rB, ay);
        const float C = dot(d, ay);
        const float k = j.invMassA + j.invMassB + j.invIA * j.sAy * j.sAy + j.invIB * j.sBy * j.sBy;
        const float impulse = k != 0.0f ? -C / k : 0.0f;
        const V2 P = impulse * ay;
        const float LA = impulse * sAy, LB = impulse * sBy;
        cA -= j.invMassA * P;
        aA -= j.invIA * LA;
        cB += j.invMassB * P;
        aB += j.invIB * LB;
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return absf(C) <= kLinearSlop;
    }

    // The 3x3 K of b2WeldJoint (REF/Dynamics/Joints/b2WeldJoint.cpp:95-104)
    static Mat33 weldK(float mA, float iA, float mB, float iB, V2 rA, V2 rB) {
        Mat33 K;
        K.ex.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
        K.ey.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
        K.ez.x = -rA.y * iA - rB.y * iB;
        K.ex.y = K.ey.x;
        K.ey.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
        K.ez.y = rA.x * iA + rB.x * iB;
        K.ex.z = K.ez.x;
        K.ey.z = K.ez.y;
        K.ez.z = iA + iB;
        return K;
    }

    // b2WeldJoint::InitVelocityConstraints, REF b2WeldJoint.cpp:60-169
    void weldInitVelocity(Joint& j, float h, float dtRatio) {
        const float aA = posA[j.indexA], aB = posA[j.indexB];
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        j.rA = mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = mul(qB, j.localAnchorB - j.localCenterB);
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const Mat33 K = weldK(mA, iA, mB, iB, j.rA, j.rB);
        if (j.frequencyHz > 0.0f) {
            j.mass = inverse22(K);
            float invM = iA + iB;
            const float m = invM > 0.0f ? 1.0f / invM : 0.0f;
            const float C = aB - aA - j.referenceAngle;
            const float omega = 2.0f * kPi * j.frequencyHz;
            const float d = 2.0f * m * j.dampingRatio * omega;
            const float k = m * omega * omega;
            j.gamma = h * (d + h * k);
            j.gamma = j.gamma != 0.0f ? 1.0f / j.gamma : 0.0f;
            j.bias = C * h * k * j.gamma;
            invM += j.gamma;
            j.mass.ez.z = invM != 0.0f ? 1.0f / invM : 0.0f;
        } else if (K.ez.z == 0.0f) {
            j.mass = inverse22(K);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        } else {
            j.mass = symInverse33(K);
            j.gamma = 0.0f;
            j.bias = 0.0f;
        }
        if (warmStarting) {
            j.impulse.x *= dtRatio; j.impulse.y *= dtRatio; j.impulse.z *= dtRatio;
            const V2 P = mk(j.impulse.x, j.impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + j.impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + j.impulse.z);
        } else {
            j.impulse = V3{0.0f, 0.0f, 0.0f};
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2WeldJoint::SolveVelocityConstraints, REF b2WeldJoint.cpp:171-227
    void weldSolveVelocity(Joint& j) {
        V2 vA = velV[j.indexA], vB = velV[j.indexB];
        float wA = velW[j.indexA], wB = velW[j.indexB];
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        if (j.frequencyHz > 0.0f) {
            const float Cdot2 = wB - wA;
            const float impulse2 = -j.mass.ez.z * (Cdot2 + j.bias + j.gamma * j.impulse.z);
            j.impulse.z += impulse2;
            wA -= iA * impulse2;
            wB += iB * impulse2;
            const V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            const V2 impulse1 = -mul22(j.mass, Cdot1);
            j.impulse.x += impulse1.x;
            j.impulse.y += impulse1.y;
            const V2 P = impulse1;
            vA -= mA * P;
            wA -= iA * cross(j.rA, P);
            vB += mB * P;
            wB += iB * cross(j.rB, P);
        } else {
            const V2 Cdot1 = vB + cross(wB, j.rB) - vA - cross(wA, j.rA);
            const float Cdot2 = wB - wA;
            V3 impulse = mul33(j.mass, V3{Cdot1.x, Cdot1.y, Cdot2});
            impulse = V3{-impulse.x, -impulse.y, -impulse.z};
            j.impulse.x += impulse.x; j.impulse.y += impulse.y; j.impulse.z += impulse.z;
            const V2 P = mk(impulse.x, impulse.y);
            vA -= mA * P;
            wA -= iA * (cross(j.rA, P) + impulse.z);
            vB += mB * P;
            wB += iB * (cross(j.rB, P) + impulse.z);
        }
        velV[j.indexA] = vA; velW[j.indexA] = wA;
        velV[j.indexB] = vB; velW[j.indexB] = wB;
    }

    // b2WeldJoint::SolvePositionConstraints, REF b2WeldJoint.cpp:229-308
    bool weldSolvePosition(Joint& j) {
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        const Rot qA = stepRot(aA), qB = stepRot(aB);
        const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
        const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
        const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
        const Mat33 K = weldK(mA, iA, mB, iB, rA, rB);
        float positionError, angularError;
        if (j.frequencyHz > 0.0f) {
            const V2 C1 = cB + rB - cA - rA;
            positionError = length(C1);
            angularError = 0.0f;
            const V2 P = -solve22(K, C1);
            cA -= mA * P;
            aA -= iA * cross(rA, P);
            cB += mB * P;
            aB += iB * cross(rB, P);
        } else {
            const V2 C1 = cB + rB - cA - rA;
            const float C2 = aB - aA - j.referenceAngle;
            positionError = length(C1);
            angularError = absf(C2);
            V3 impulse;
            if (K.ez.z > 0.0f) {
                impulse = solve33(K, V3{C1.x, C1.y, C2});
                impulse = V3{-impulse.x, -impulse.y, -impulse.z};
            } else {
                const V2 impulse2 = -solve22(K, C1);
                impulse = V3{impulse2.x, impulse2.y, 0.0f};
            }
            const V2 P = mk(impulse.x, impulse.y);
            cA -= mA * P;
            aA -= iA * (cross(rA, P) + impulse.z);
            cB += mB * P;
            aB += iB * (cross(rB, P) + impulse.z);
        }
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return positionError <= kLinearSlop && angularError <= kAngularSlop;
    }

    bool jointSolvePosition(Joint& j) {
        if (j.type == 1) return distanceSolvePosition(j);
        if (j.type == 2) return ropeSolvePosition(j);
        if (j.type == 3) return weldSolvePosition(j);
        if (j.type == 4 || j.type == 5 || j.type == 7) return true;  // friction, motor, mouse: no position pass
        if (j.type == 6) return pulleySolvePosition(j);
        if (j.type == 8) return prismaticSolvePosition(j);
        if (j.type == 10) return gearSolvePosition(j);
        if (j.type == 9) return wheelSolvePosition(j);
        V2 cA = posC[j.indexA]; float aA = posA[j.indexA];
        V2 cB = posC[j.indexB]; float aB = posA[j.indexB];
        float angularError = 0.0f, positionError = 0.0f;
        const bool fixedRotation = (j.invIA + j.invIB == 0.0f);
        if (j.enableLimit && j.limitState != kInactiveLimit && !fixedRotation) {
            const float angle = aB - aA - j.referenceAngle;
            float limitImpulse = 0.0f;
            if (j.limitState == kEqualLimits) {
                const float C = clampf(angle - j.lowerAngle, -kMaxAngularCorrection, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
                angularError = absf(C);
            } else if (j.limitState == kAtLowerLimit) {
                float C = angle - j.lowerAngle;
                angularError = -C;
                C = clampf(C + kAngularSlop, -kMaxAngularCorrection, 0.0f);
                limitImpulse = -j.motorMass * C;
            } else if (j.limitState == kAtUpperLimit) {
                float C = angle - j.upperAngle;
                angularError = C;
                C = clampf(C - kAngularSlop, 0.0f, kMaxAngularCorrection);
                limitImpulse = -j.motorMass * C;
            }
            aA -= j.invIA * limitImpulse;
            aB += j.invIB * limitImpulse;
        }
        {
            const Rot qA = stepRot(aA), qB = stepRot(aB);
            const V2 rA = mul(qA, j.localAnchorA - j.localCenterA);
            const V2 rB = mul(qB, j.localAnchorB - j.localCenterB);
            const V2 C = cB + rB - cA - rA;
            positionError = length(C);
            const float mA = j.invMassA, mB = j.invMassB, iA = j.invIA, iB = j.invIB;
            float k11 = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
            float k21 = -iA * rA.x * rA.y - iB * rB.x * rB.y;
            float k12 = k21;
            float k22 = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
            // b2Mat22::Solve (REF b2Math.h:242-254): ex = (k11, k21), ey = (k12, k22)
            float det = k11 * k22 - k12 * k21;
            if (det != 0.0f) det = 1.0f / det;
            V2 impulse = mk(det * (k22 * C.x - k12 * C.y), det * (k11 * C.y - k21 * C.x));
            impulse = -impulse;
            cA -= mA * impulse;
            aA -= iA * cross(rA, impulse);
            cB += mB * impulse;
            aB += iB * cross(rB, impulse);
        }
        posC[j.indexA] = cA; posA[j.indexA] = aA;
        posC[j.indexB] = cB; posA[j.indexB] = aB;
        return positionError <= kLinearSlop && angularError <= kAngularSlop;
    }

    void solveVelocity(VelConstraint& vc) {
        const int iA = vc.indexA, iB = vc.indexB;
        const float mA = vc.invMassA, IA = vc.invIA, mB = vc.invMassB, IB = vc.invIB;
        const int pointCount = vc.pointCount;
        V2 vA = velV[iA]; float wA = velW[iA];
        V2 vB = velV[iB]; float wB = velW[iB];
        V2 normal = vc.normal;
        V2 tangent = cross(normal, 1.0f);
        const float friction = vc.friction;
        for (int j = 0; j < pointCount; ++j) {
            VelPoint& vp = vc.points[j];
            V2 dv = vB + cross(wB, vp.rB) - vA - cross(wA, vp.rA);
            float vt = dot(dv, tangent) - vc.tangentSpeed;
            float lambda = vp.tangentMass * (-vt);
            float maxFriction = friction * vp.normalImpulse;
            float newImpulse = clampf(vp.tangentImpulse + lambda, -maxFriction, maxFriction);
            lambda = newImpulse - vp.tangentImpulse;
            vp.tangentImpulse = newImpulse;
            V2 P = lambda * tangent;
            vA -= mA * P;
            wA -= IA * cross(vp.rA, P);
            vB += mB * P;
            wB += IB * cross(vp.rB, P);
        }
        if (pointCount == 1) {
            VelPoint& vp = vc.points[0];
            V2 dv = vB + cross(wB, vp.rB) - vA - cross(wA, vp.rA);
            float vn = dot(dv, normal);
            float lambda = -vp.normalMass * (vn - vp.velocityBias);
            float newImpulse = vmax(vp.normalImpulse + lambda, 0.0f);
            lambda = newImpulse - vp.normalImpulse;
            vp.normalImpulse = newImpulse;
            V2 P = lambda * normal;
            vA -= mA * P;
            wA -= IA * cross(vp.rA, P);
            vB += mB * P;
            wB += IB * cross(vp.rB, P);
        } else {
            // Block solver by total enumeration (REF :386-589)
            VelPoint& cp1 = vc.points[0];
            VelPoint& cp2 = vc.points[1];
            V2 a = mk(cp1.normalImpulse, cp2.normalImpulse);
            V2 dv1 = vB + cross(wB, cp1.rB) - vA - cross(wA, cp1.rA);
            V2 dv2 = vB + cross(wB, cp2.rB) - vA - cross(wA, cp2.rA);
            float vn1 = dot(dv1, normal);
            float vn2 = dot(dv2, normal);
            V2 b;
            b.x = vn1 - cp1.velocityBias;
            b.y = vn2 - cp2.velocityBias;
            // b -= K * a  (b2Mul(b2Mat22, b2Vec2), REF b2Math.h:389-392)
            b -= mk(vc.K[0] * a.x + vc.K[2] * a.y, vc.K[1] * a.x + vc.K[3] * a.y);
            auto apply = [&](V2 x) {
                V2 d = x - a;
                V2 P1 = d.x * normal;
                V2 P2 = d.y * normal;
                vA -= mA * (P1 + P2);
                wA -= IA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
                vB += mB * (P1 + P2);
                wB += IB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
                cp1.normalImpulse = x.x;
                cp2.normalImpulse = x.y;
            };
            for (;;) {
                // x = -M * b
                V2 Mb = mk(vc.normalMass[0] * b.x + vc.normalMass[2] * b.y,
                           vc.normalMass[1] * b.x + vc.normalMass[3] * b.y);
                V2 x = -Mb;
                if (x.x >= 0.0f && x.y >= 0.0f) { apply(x); break; }
                x.x = -cp1.normalMass * b.x;
                x.y = 0.0f;
                vn1 = 0.0f;
                vn2 = vc.K[1] * x.x + b.y;
                if (x.x >= 0.0f && vn2 >= 0.0f) { apply(x); break; }
                x.x = 0.0f;
                x.y = -cp2.normalMass * b.y;
                vn1 = vc.K[2] * x.y + b.x;
                vn2 = 0.0f;
                if (x.y >= 0.0f && vn1 >= 0.0f) { apply(x); break; }
                x.x = 0.0f;
                x.y = 0.0f;
                vn1 = b.x;
                vn2 = b.y;
                if (vn1 >= 0.0f && vn2 >= 0.0f) { apply(x); break; }
                break;
            }
        }
        velV[iA] = vA; velW[iA] = wA;
        velV[iB] = vB; velW[iB] = wB;
    }

    // b2PositionSolverManifold::Initialize, REF/Dynamics/Contacts/b2ContactSolver.cpp:622-671
    static void positionManifold(const PosConstraint& pc, const Xf& xfA, const Xf& xfB, int j, V2* normalOut,
                                 V2* pointOut, float* separationOut) {
        V2 normal, point;
        float separation;
        if (pc.type == B2G_MANIFOLD_CIRCLES) {
            V2 pointA = mul(xfA, pc.localPoint);
            V2 pointB = mul(xfB, pc.localPoints[0]);
            normal = pointB - pointA;
            normalize(normal);
            point = 0.5f * (pointA + pointB);
            separation = dot(pointB - pointA, normal) - pc.radiusA - pc.radiusB;
        } else if (pc.type == B2G_MANIFOLD_FACE_A) {
            normal = mul(xfA.q, pc.localNormal);
            V2 planePoint = mul(xfA, pc.localPoint);
            V2 clipPoint = mul(xfB, pc.localPoints[j]);
            separation = dot(clipPoint - planePoint, normal) - pc.radiusA - pc.radiusB;
            point = clipPoint;
        } else {
            normal = mul(xfB.q, pc.localNormal);
            V2 planePoint = mul(xfB, pc.localPoint);
            V2 clipPoint = mul(xfA, pc.localPoints[j]);
            separation = dot(clipPoint - planePoint, normal) - pc.radiusA - pc.radiusB;
            point = clipPoint;
            normal = -normal;
        }
        *normalOut = normal;
        *pointOut = point;
        *separationOut = separation;
    }

    // SolvePositionConstraints + b2PositionSolverManifold (REF :614-744)
    bool solvePositions() {
        float minSeparation = 0.0f;
        for (size_t i = 0; i < pcs.size(); ++i) {
            PosConstraint& pc = pcs[i];
            const int iA = pc.indexA, iB = pc.indexB;
            V2 lcA = pc.localCenterA, lcB = pc.localCenterB;
            const float mA = pc.invMassA, IA = pc.invIA, mB = pc.invMassB, IB = pc.invIB;
            V2 cA = posC[iA]; float aA = posA[iA];
            V2 cB = posC[iB]; float aB = posA[iB];
            for (int j = 0; j < pc.pointCount; ++j) {
                Xf xfA, xfB;
                xfA.q = stepRot(aA);
                xfB.q = stepRot(aB);
                xfA.p = cA - mul(xfA.q, lcA);
                xfB.p = cB - mul(xfB.q, lcB);
                V2 normal, point;
                float separation;
                positionManifold(pc, xfA, xfB, j, &normal, &point, &separation);
                V2 rA = point - cA;
                V2 rB = point - cB;
                minSeparation = vmin(minSeparation, separation);
                float C = clampf(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
                float rnA = cross(rA, normal);
                float rnB = cross(rB, normal);
                float K = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
                float impulse = K > 0.0f ? -C / K : 0.0f;
                V2 P = impulse * normal;
                cA -= mA * P;
                aA -= IA * cross(rA, P);
                cB += mB * P;
                aB += IB * cross(rB, P);
            }
            posC[iA] = cA; posA[iA] = aA;
            posC[iB] = cB; posA[iB] = aB;
        }
        return minSeparation >= -3.0f * kLinearSlop;
    }

    // b2World::Solve, REF/Dynamics/b2World.cpp:388-577
    void solve(float dt, float inv_dt, float dtRatio, int velIters, int posIters) {
        const int nb = (int)bodies.size();
        for (auto& b : bodies) b.flags &= ~kIsland;
        for (int ci = contactHead; ci != -1; ci = contacts[ci].next) contacts[ci].flags &= ~cIsland;
        for (auto& j : joints) j.islandFlag = false;
        stack.resize(nb > 0 ? nb : 1);
        for (int seed = nb - 1; seed >= 0; --seed) {
            Body& sb = bodies[seed];
            if (!sb.alive) continue;
            if (sb.flags & kIsland) continue;
            if ((sb.flags & kAwake) == 0 || (sb.flags & kActive) == 0) continue;
            if (sb.type == kStatic) continue;
            islandBodies.clear();
            islandContacts.clear();
            islandJoints.clear();
            int sc = 0;
            stack[sc++] = seed;
            sb.flags |= kIsland;
            while (sc > 0) {
                const int bi = stack[--sc];
                Body& b = bodies[bi];
                b.islandIndex = (int)islandBodies.size();
                islandBodies.push_back(bi);
                setAwake(b, true);
                if (b.type == kStatic) continue;
                for (int e = b.edgeHead; e != -1; e = edgeRef(e).next) {
                    Contact& c = contacts[e >> 1];
                    if (c.flags & cIsland) continue;
                    if ((c.flags & cEnabled) == 0 || (c.flags & cTouching) == 0) continue;
                    if (fixtures[c.fA].sensor || fixtures[c.fB].sensor) continue;
                    islandContacts.push_back(e >> 1);
                    c.flags |= cIsland;
                    const int other = edgeOther(e);
                    if (bodies[other].flags & kIsland) continue;
                    stack[sc++] = other;
                    bodies[other].flags |= kIsland;
                }
                // joints of this body (REF b2World.cpp:502-529)
                for (int e = b.jointHead; e != -1; e = joints[e >> 1].node[e & 1].next) {
                    Joint& j = joints[e >> 1];
                    if (j.islandFlag) continue;
                    const int other = (e & 1) ? j.bodyA : j.bodyB;
                    if ((bodies[other].flags & kActive) == 0) continue;
                    islandJoints.push_back(e >> 1);
                    j.islandFlag = true;
                    if (bodies[other].flags & kIsland) continue;
                    stack[sc++] = other;
                    bodies[other].flags |= kIsland;
                }
            }
            solveIsland(dt, inv_dt, dtRatio, velIters, posIters);
            for (int bi : islandBodies) {
                if (bodies[bi].type == kStatic) bodies[bi].flags &= ~kIsland;
            }
        }
        for (int bi = nb - 1; bi >= 0; --bi) {
            Body& b = bodies[bi];
            if (!b.alive) continue;
            if ((b.flags & kIsland) == 0) continue;
            if (b.type == kStatic) continue;
            synchronizeFixtures(b);
        }
        findNewContacts();
    }

    Sweep sweepOf(const Body& b) const {
        Sweep sw;
        sw.localCenter = b.localCenter; sw.c0 = b.c0; sw.c = b.c; sw.a0 = b.a0; sw.a = b.a; sw.alpha0 = b.alpha0;
        return sw;
    }
    // b2Sweep::Advance on a body's sweep, REF/Common/b2Math.h:705-712
    void advanceSweep(Body& b, float alpha) {
        Sweep sw = sweepOf(b);
        sweepAdvance(sw, alpha);
        b.c0 = sw.c0; b.a0 = sw.a0; b.alpha0 = sw.alpha0;
    }
    // b2Body::Advance, REF/Dynamics/b2Body.h:842-850
    void advanceBody(Body& b, float alpha) {
        advanceSweep(b, alpha);
        b.c = b.c0;
        b.a = b.a0;
        synchronizeTransform(b);
    }

    // b2ContactSolver::SolveTOIPositionConstraints, REF/Dynamics/Contacts/b2ContactSolver.cpp:753-841:
    // only the two TOI bodies move, with the TOI Baumgarte factor.
    bool solveTOIPositions(int toiA, int toiB) {
        float minSeparation = 0.0f;
        for (size_t i = 0; i < pcs.size(); ++i) {
            PosConstraint& pc = pcs[i];
            const int iA = pc.indexA, iB = pc.indexB;
            const V2 lcA = pc.localCenterA, lcB = pc.localCenterB;
            float mA = 0.0f, IA = 0.0f, mB = 0.0f, IB = 0.0f;
            if (iA == toiA || iA == toiB) { mA = pc.invMassA; IA = pc.invIA; }
            if (iB == toiA || iB == toiB) { mB = pc.invMassB; IB = pc.invIB; }
            V2 cA = posC[iA]; float aA = posA[iA];
            V2 cB = posC[iB]; float aB = posA[iB];
            for (int j = 0; j < pc.pointCount; ++j) {
                Xf xfA, xfB;
                xfA.q = stepRot(aA);
                xfB.q = stepRot(aB);
                xfA.p = cA - mul(xfA.q, lcA);
                xfB.p = cB - mul(xfB.q, lcB);
                V2 normal, point;
                float separation;
                positionManifold(pc, xfA, xfB, j, &normal, &point, &separation);
                const V2 rA = point - cA;
                const V2 rB = point - cB;
                minSeparation = vmin(minSeparation, separation);
                const float C = clampf(kToiBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
                const float rnA = cross(rA, normal);
                const float rnB = cross(rB, normal);
                const float K = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
                const float impulse = K > 0.0f ? -C / K : 0.0f;
                const V2 P = impulse * normal;
                cA -= mA * P;
                aA -= IA * cross(rA, P);
                cB += mB * P;
                aB += IB * cross(rB, P);
            }
            posC[iA] = cA; posA[iA] = aA;
            posC[iB] = cB; posA[iB] = aB;
        }
        return minSeparation >= -1.5f * kLinearSlop;
    }

    // b2Island::SolveTOI, REF/Dynamics/b2Island.cpp:385-515 (no warm start,
    // impulses not stored; the TOI pair is islandBodies[0] and [1]).
    void solveTOIIsland(float h, int velIters) {
        const int nb = (int)islandBodies.size();
        posC.resize(nb); posA.resize(nb); velV.resize(nb); velW.resize(nb);
        for (int i = 0; i < nb; ++i) {
            const Body& b = bodies[islandBodies[i]];
            posC[i] = b.c; posA[i] = b.a; velV[i] = b.v; velW[i] = b.w;
        }
        contactSolverInit(1.0f, false);
        for (int i = 0; i < 20; ++i) {
            if (solveTOIPositions(0, 1)) break;
        }
        for (int k = 0; k < 2; ++k) {
            Body& b = bodies[islandBodies[k]];
            b.c0 = posC[k];
            b.a0 = posA[k];
        }
        initVelocityConstraints();
        for (int it = 0; it < velIters; ++it)
            for (size_t i = 0; i < vcs.size(); ++i) solveVelocity(vcs[i]);
        for (int i = 0; i < nb; ++i) {
            V2 c = posC[i]; float a = posA[i];
            V2 vv = velV[i]; float w = velW[i];
            const V2 translation = h * vv;
            if (dot(translation, translation) > kMaxTranslationSq) {
                const float ratio = kMaxTranslation / length(translation);
                vv *= ratio;
            }
            const float rotation = h * w;
            if (rotation * rotation > kMaxRotationSq) {
                const float ratio = kMaxRotation / absf(rotation);
                w *= ratio;
            }
            c += h * vv;
            a += h * w;
            Body& b = bodies[islandBodies[i]];
            b.c = c; b.a = a; b.v = vv; b.w = w;
            synchronizeTransform(b);
        }
    }

    // TOI of one contact over the rest of the step, the two sweeps first
    // brought to a common start (REF/Dynamics/b2World.cpp:645-689).
    float contactTOI(Contact& c) {
        Body& bA = bodies[fixtures[c.fA].body];
        Body& bB = bodies[fixtures[c.fB].body];
        float alpha0 = bA.alpha0;
        if (bA.alpha0 < bB.alpha0) {
            alpha0 = bB.alpha0;
            advanceSweep(bA, alpha0);
        } else if (bB.alpha0 < bA.alpha0) {
            alpha0 = bA.alpha0;
            advanceSweep(bB, alpha0);
        }
        const Proxy pA = proxyOf(&fixtures[c.fA].shape), pB = proxyOf(&fixtures[c.fB].shape);
        float beta;
        const int state = timeOfImpact(&beta, pA, pB, sweepOf(bA), sweepOf(bB), 1.0f);
        return state == kToiTouching ? vmin(alpha0 + (1.0f - alpha0) * beta, 1.0f) : 1.0f;
    }

    // b2World::SolveTOI, REF/Dynamics/b2World.cpp:579-897: repeatedly take
    // the earliest TOI over the contact list, move its two bodies there,
    // solve a small island of them and the static, kinematic and bullet
    // bodies they touch for the rest of the step, and find new pairs.
    void solveTOI(float dt, int velIters) {
        if (stepComplete) {
            for (auto& b : bodies) { b.flags &= ~kIsland; b.alpha0 = 0.0f; }
            for (int ci = contactHead; ci != -1; ci = contacts[ci].next) {
                Contact& c = contacts[ci];
                c.flags &= ~(cToi | cIsland);
                c.toiCount = 0;
                c.toi = 1.0f;
            }
        }
        for (;;) {
            int minContact = -1;
            float minAlpha = 1.0f;
            for (int ci = contactHead; ci != -1; ci = contacts[ci].next) {
                Contact& c = contacts[ci];
                if ((c.flags & cEnabled) == 0) continue;
                if (c.toiCount > kMaxSubSteps) continue;
                float alpha = 1.0f;
                if (c.flags & cToi) {
                    alpha = c.toi;
                } else {
                    const Fixture& fA = fixtures[c.fA];
                    const Fixture& fB = fixtures[c.fB];
                    if (fA.sensor || fB.sensor) continue;
                    const Body& bA = bodies[fA.body];
                    const Body& bB = bodies[fB.body];
                    const bool activeA = (bA.flags & kAwake) && bA.type != kStatic;
                    const bool activeB = (bB.flags & kAwake) && bB.type != kStatic;
                    if (!activeA && !activeB) continue;
                    const bool collideA = (bA.flags & kBullet) || bA.type != kDynamic;
                    const bool collideB = (bB.flags & kBullet) || bB.type != kDynamic;
                    if (!collideA && !collideB) continue;
                    alpha = contactTOI(c);
                    c.toi = alpha;
                    c.flags |= cToi;
                }
                if (alpha < minAlpha) { minContact = ci; minAlpha = alpha; }
            }
            if (minContact == -1 || 1.0f - 10.0f * FLT_EPSILON < minAlpha) {
                stepComplete = true;
                break;
            }
            Contact& mc = contacts[minContact];
            const int ia = fixtures[mc.fA].body, ib = fixtures[mc.fB].body;
            Body& bA = bodies[ia];
            Body& bB = bodies[ib];
            const Sweep backupA = sweepOf(bA), backupB = sweepOf(bB);
            advanceBody(bA, minAlpha);
            advanceBody(bB, minAlpha);
            updateContact(minContact);
            mc.flags &= ~cToi;
            ++mc.toiCount;
            if ((mc.flags & cEnabled) == 0 || (mc.flags & cTouching) == 0) {
                mc.flags &= ~cEnabled;
                restoreSweep(bA, backupA);
                restoreSweep(bB, backupB);
                continue;
            }
            setAwake(bA, true);
            setAwake(bB, true);
            islandBodies.clear();
            islandContacts.clear();
            bA.islandIndex = 0; islandBodies.push_back(ia);
            bB.islandIndex = 1; islandBodies.push_back(ib);
            islandContacts.push_back(minContact);
            bA.flags |= kIsland;
            bB.flags |= kIsland;
            mc.flags |= cIsland;
            const int pair[2] = {ia, ib};
            for (int k = 0; k < 2; ++k) {
                Body& body = bodies[pair[k]];
                if (body.type != kDynamic) continue;
                for (int e = body.edgeHead; e != -1; e = edgeRef(e).next) {
                    if ((int)islandBodies.size() == 2 * kMaxTOIContacts) break;
                    if ((int)islandContacts.size() == kMaxTOIContacts) break;
                    const int ci = e >> 1;
                    Contact& c = contacts[ci];
                    if (c.flags & cIsland) continue;
                    const int oi = edgeOther(e);
                    Body& other = bodies[oi];
                    if (other.type == kDynamic && (body.flags & kBullet) == 0 && (other.flags & kBullet) == 0) continue;
                    if (fixtures[c.fA].sensor || fixtures[c.fB].sensor) continue;
                    const Sweep backup = sweepOf(other);
                    if ((other.flags & kIsland) == 0) advanceBody(other, minAlpha);
                    updateContact(ci);
                    if ((c.flags & cEnabled) == 0 || (c.flags & cTouching) == 0) {
                        restoreSweep(other, backup);
                        continue;
                    }
                    c.flags |= cIsland;
                    islandContacts.push_back(ci);
                    if (other.flags & kIsland) continue;
                    other.flags |= kIsland;
                    if (other.type != kStatic) setAwake(other, true);
                    other.islandIndex = (int)islandBodies.size();
                    islandBodies.push_back(oi);
                }
            }
            solveTOIIsland((1.0f - minAlpha) * dt, velIters);
            for (int bi : islandBodies) {
                Body& b = bodies[bi];
                b.flags &= ~kIsland;
                if (b.type != kDynamic) continue;
                synchronizeFixtures(b);
                for (int e = b.edgeHead; e != -1; e = edgeRef(e).next) contacts[e >> 1].flags &= ~(cToi | cIsland);
            }
            findNewContacts();
            if (subStepping) {
                stepComplete = false;
                break;
            }
        }
    }
    // Restores a body's sweep from a copy and re-derives its transform.
    void restoreSweep(Body& b, const Sweep& saved) {
        b.c0 = saved.c0; b.c = saved.c; b.a0 = saved.a0; b.a = saved.a; b.alpha0 = saved.alpha0;
        synchronizeTransform(b);
    }

    // b2World::Step, REF/Dynamics/b2World.cpp:899-964
    void step(float dt, int velIters, int posIters) {
        if (newFixture) {
            findNewContacts();
            newFixture = false;
        }
        float inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
        float dtRatio = inv_dt0 * dt;
        collideAll();
        if (stepComplete && dt > 0.0f) solve(dt, inv_dt, dtRatio, velIters, posIters);
        if (continuous && dt > 0.0f) solveTOI(dt, velIters);
        if (dt > 0.0f) inv_dt0 = inv_dt;
        if (autoClearForces) {
            for (auto& b : bodies) { b.force = mk(0.0f, 0.0f); b.torque = 0.0f; }
        }
    }

    int load(const b2scene::BodyDef* bd, int nb, const b2scene::FixtureDef* fd, int nf,
             const b2scene::JointDef* jd = nullptr, int nj = 0) {
        const int base = (int)bodies.size();
        for (int i = 0; i < nb; ++i) {
            const int bi = createBody(bd[i]);
            for (int k = 0; k < bd[i].fixtureCount; ++k) {
                const int fi = bd[i].firstFixture + k;
                if (fi < 0 || fi >= nf) return B2G_ERR_ARG;
                const b2scene::FixtureDef& f = fd[fi];
                b2g_shape s;
                if (f.shape == b2scene::kPolygon) {
                    if (f.polyMode == b2scene::kPolyBox) setAsBox(&s, f.pts[0], f.pts[1]);
                    else if (f.polyMode == b2scene::kPolyBoxRotated)
                        setAsBoxRotated(&s, f.pts[0], f.pts[1], mk(f.pts[2], f.pts[3]), f.pts[4]);
                    else {
                        V2 p[8];
                        for (int j = 0; j < f.n && j < 8; ++j) p[j] = mk(f.pts[2 * j], f.pts[2 * j + 1]);
                        polygonSet(&s, p, f.n);
                    }
                } else if (f.shape == b2scene::kCircle) {
                    circleShape(&s, mk(f.cx, f.cy), f.radius);
                } else if (f.shape == b2scene::kChain) {
                    // b2ChainShape: one proxy per child edge, in child order
                    // (REF/Collision/Shapes/b2ChainShape.cpp:39-143)
                    std::vector<V2> vs;
                    for (int j = 0; j < f.n && j < 8; ++j) vs.push_back(mk(f.pts[2 * j], f.pts[2 * j + 1]));
                    V2 prev = mk(f.e[0], f.e[1]), next = mk(f.e[6], f.e[7]);
                    bool hasPrev = f.has0 != 0, hasNext = f.has3 != 0;
                    if (f.polyMode == 1) {
                        vs.push_back(vs[0]);
                        prev = vs[vs.size() - 2];
                        next = vs[1];
                        hasPrev = hasNext = true;
                    }
                    const int m = (int)vs.size();
                    int owner = -1;
                    for (int c = 0; c + 1 < m; ++c) {
                        edgeShape(&s, c > 0 ? vs[c - 1] : prev, vs[c], vs[c + 1], c < m - 2 ? vs[c + 2] : next,
                                  c > 0 || hasPrev, c < m - 2 || hasNext);
                        s.flags |= 4;
                        const int fi = createFixture(bi, s, f, owner);
                        if (owner < 0) owner = fi;
                    }
                    continue;
                } else {
                    edgeShape(&s, mk(f.e[0], f.e[1]), mk(f.e[2], f.e[3]), mk(f.e[4], f.e[5]), mk(f.e[6], f.e[7]),
                              f.has0 != 0, f.has3 != 0);
                }
                createFixture(bi, s, f);
            }
        }
        jointBase = (int)joints.size();
        for (int k = 0; k < nj; ++k) {
            b2scene::JointDef d = jd[k];
            if (d.bodyA < 0 || d.bodyB < 0 || d.bodyA >= nb || d.bodyB >= nb || d.bodyA == d.bodyB)
                return B2G_ERR_ARG;
            d.bodyA += base;
            d.bodyB += base;
            createJoint(d);
        }
        return B2G_OK;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

void* orc_world_create(float gx, float gy) {
    World* w = new World;
    w->gravity = mk(gx, gy);
    return w;
}
void orc_world_destroy(void* h) { delete (World*)h; }

// b2World::SetAllowSleeping (REF b2World.cpp:370-386) wakes every body when
// sleeping is switched off.
void orc_world_set_flags(void* h, int allowSleep, int warmStarting, int continuous, int subStepping) {
    World* w = (World*)h;
    bool s = allowSleep != 0;
    if (s != w->allowSleep) {
        w->allowSleep = s;
        if (!s) for (auto& b : w->bodies) w->setAwake(b, true);
    }
    w->warmStarting = warmStarting != 0;
    w->continuous = continuous != 0;
    w->subStepping = subStepping != 0;
}

int orc_world_load(void* h, const b2scene::BodyDef* bd, int nb, const b2scene::FixtureDef* fd, int nf) {
    return ((World*)h)->load(bd, nb, fd, nf);
}

int orc_world_load_scene(void* h, const char* kind, int n, unsigned seed) {
    b2scene::Scene s = b2scene::byName(kind, n, seed);
    return ((World*)h)->load(s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size(),
                             s.joints.data(), (int)s.joints.size());
}

int orc_world_joint_count(void* h) { return (int)((World*)h)->joints.size(); }

int orc_world_set_joint(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor, float speed,
                        float maxTorque) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed) return -1;
    w->setJoint(joint, enableLimit != 0, lower, upper, enableMotor != 0, speed, maxTorque);
    return 0;
}

int orc_world_destroy_body(void* h, int body) {
    World* w = (World*)h;
    if (body < 0 || body >= (int)w->bodies.size() || !w->bodies[body].alive) return -1;
    w->destroyBody(body);
    return 0;
}

int orc_world_set_active(void* h, int body, int flag) {
    World* w = (World*)h;
    if (body < 0 || body >= (int)w->bodies.size() || !w->bodies[body].alive) return -1;
    w->setActive(body, flag != 0);
    return 0;
}

// type: 0 static, 1 kinematic, 2 dynamic.
int orc_world_set_type(void* h, int body, int type) {
    World* w = (World*)h;
    if (body < 0 || body >= (int)w->bodies.size() || !w->bodies[body].alive || type < 0 || type > 2) return -1;
    w->setType(body, type);
    return 0;
}

// fixture: the first device-order slot of the b2Fixture (its first child).
int orc_world_destroy_fixture(void* h, int fixture) {
    World* w = (World*)h;
    if (fixture < 0 || fixture >= (int)w->fixtures.size() || w->fixtures[fixture].owner != fixture) return -1;
    const orc::Fixture& f = w->fixtures[fixture];
    if (!w->bodies[f.body].alive) return -1;
    bool attached = false;
    for (int fi : w->bodies[f.body].fixtures) attached = attached || fi == fixture;
    if (!attached) return -1;
    w->destroyFixture(fixture);
    return 0;
}

// Per fixture (chain edges one each) in creation order: its broad-phase
// tree proxy id, -1 once destroyed.
int orc_world_proxy_ids(void* h, int32_t* out, int cap) {
    World* w = (World*)h;
    const int n = (int)w->fixtures.size();
    for (int i = 0; i < n && i < cap; ++i) out[i] = w->fixtures[i].proxyId;
    return n;
}

int orc_world_destroy_joint(void* h, int joint) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed) return -1;
    w->destroyJoint(joint);
    return 0;
}

// Per joint in creation order: impulse x, y, z and motor impulse.
// b2DistanceJoint::SetLength / SetFrequency / SetDampingRatio.
// b2PrismaticJoint::EnableLimit, SetLimits (dropping the limit impulse when
// they change something), EnableMotor, SetMotorSpeed, SetMaxMotorForce, all
// waking both bodies, REF/Dynamics/Joints/b2PrismaticJoint.cpp:546-609.
int orc_world_set_prismatic(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor,
                            float speed, float maxForce) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 8)
        return B2G_ERR_ARG;
    Joint& j = w->joints[joint];
    Body& bA = w->bodies[j.bodyA];
    Body& bB = w->bodies[j.bodyB];
    if ((enableLimit != 0) != j.enableLimit) {
        w->setAwake(bA, true);
        w->setAwake(bB, true);
        j.enableLimit = enableLimit != 0;
        j.impulse.z = 0.0f;
    }
    if (lower != j.lowerAngle || upper != j.upperAngle) {
        w->setAwake(bA, true);
        w->setAwake(bB, true);
        j.lowerAngle = lower;
        j.upperAngle = upper;
        j.impulse.z = 0.0f;
    }
    w->setAwake(bA, true);
    w->setAwake(bB, true);
    j.enableMotor = enableMotor != 0;
    j.motorSpeed = speed;
    j.maxMotorTorque = maxForce;
    return B2G_OK;
}

// b2WheelJoint::EnableMotor, SetMotorSpeed, SetMaxMotorTorque (waking both
// bodies), SetSpringFrequencyHz, SetSpringDampingRatio,
// REF/Dynamics/Joints/b2WheelJoint.cpp:377-401.
int orc_world_set_wheel(void* h, int joint, int enableMotor, float speed, float maxTorque, float hz, float ratio) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 9)
        return B2G_ERR_ARG;
    Joint& j = w->joints[joint];
    w->setAwake(w->bodies[j.bodyA], true);
    w->setAwake(w->bodies[j.bodyB], true);
    j.enableMotor = enableMotor != 0;
    j.motorSpeed = speed;
    j.maxMotorTorque = maxTorque;
    j.frequencyHz = hz;
    j.dampingRatio = ratio;
    return B2G_OK;
}

// b2MouseJoint::SetTarget (waking body B), SetMaxForce, SetFrequency,
// SetDampingRatio, REF/Dynamics/Joints/b2MouseJoint.cpp:55-97.
int orc_world_set_mouse(void* h, int joint, float tx, float ty, float maxForce, float hz, float ratio) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 7)
        return B2G_ERR_ARG;
    Joint& j = w->joints[joint];
    Body& bB = w->bodies[j.bodyB];
    if (!(bB.flags & kAwake)) w->setAwake(bB, true);
    j.localAnchorA = mk(tx, ty);
    j.maxForce = maxForce;
    j.frequencyHz = hz;
    j.dampingRatio = ratio;
    return B2G_OK;
}

// b2FrictionJoint::SetMaxForce / SetMaxTorque.
int orc_world_set_friction(void* h, int joint, float maxForce, float maxTorque) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 4)
        return B2G_ERR_ARG;
    w->joints[joint].maxForce = maxForce;
    w->joints[joint].maxTorque = maxTorque;
    return B2G_OK;
}

// b2MotorJoint::SetLinearOffset / SetAngularOffset (waking both bodies
// when the offset changes, REF/Dynamics/Joints/b2MotorJoint.cpp:261-284),
// SetMaxForce, SetMaxTorque, SetCorrectionFactor.
int orc_world_set_motor(void* h, int joint, float ox, float oy, float angle, float maxForce, float maxTorque,
                        float factor) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 5)
        return B2G_ERR_ARG;
    Joint& j = w->joints[joint];
    if (ox != j.localAnchorA.x || oy != j.localAnchorA.y) {
        w->setAwake(w->bodies[j.bodyA], true);
        w->setAwake(w->bodies[j.bodyB], true);
        j.localAnchorA = mk(ox, oy);
    }
    if (angle != j.referenceAngle) {
        w->setAwake(w->bodies[j.bodyA], true);
        w->setAwake(w->bodies[j.bodyB], true);
        j.referenceAngle = angle;
    }
    j.maxForce = maxForce;
    j.maxTorque = maxTorque;
    j.correctionFactor = factor;
    return B2G_OK;
}

// b2RopeJoint::SetMaxLength.
int orc_world_set_rope(void* h, int joint, float maxLength) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 2)
        return B2G_ERR_ARG;
    w->joints[joint].length = maxLength;
    return B2G_OK;
}

// b2WeldJoint::SetFrequency / SetDampingRatio.
int orc_world_set_weld(void* h, int joint, float hz, float ratio) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 3)
        return B2G_ERR_ARG;
    w->joints[joint].frequencyHz = hz;
    w->joints[joint].dampingRatio = ratio;
    return B2G_OK;
}

int orc_world_set_distance(void* h, int joint, float length, float hz, float ratio) {
    World* w = (World*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint].removed || w->joints[joint].type != 1)
        return B2G_ERR_ARG;
    Joint& j = w->joints[joint];
    j.length = length;
    j.frequencyHz = hz;
    j.dampingRatio = ratio;
    return B2G_OK;
}

void orc_world_get_joints(void* h, float* out) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->joints.size(); ++i) {
        const Joint& j = w->joints[i];
        if (j.removed) {  // destroyed
            for (int k = 0; k < 4; ++k) out[4 * i + k] = 0.0f;
            continue;
        }
        if (j.type == 8) {  // b2PrismaticJoint::GetReactionForce / Torque at inv_dt = 1
            const V2 P = j.impulse.x * j.perp + (j.motorImpulse + j.impulse.z) * j.axis;
            out[4 * i] = P.x;
            out[4 * i + 1] = P.y;
            out[4 * i + 2] = j.impulse.y;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        if (j.type == 9) {  // b2WheelJoint::GetReactionForce / Torque at inv_dt = 1
            const V2 P = j.impulse.x * j.ay + j.springImpulse * j.ax;
            out[4 * i] = P.x;
            out[4 * i + 1] = P.y;
            out[4 * i + 2] = j.motorImpulse;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        if (j.type == 10) {  // b2GearJoint::GetReactionForce / Torque at inv_dt = 1
            const V2 P = j.impulse.x * j.JvAC;
            out[4 * i] = P.x;
            out[4 * i + 1] = P.y;
            out[4 * i + 2] = j.impulse.x * j.JwA;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        if (j.type == 6) {  // b2PulleyJoint::GetReactionForce at inv_dt = 1: impulse along uB
            const V2 P = j.impulse.x * j.uB;
            out[4 * i] = P.x;
            out[4 * i + 1] = P.y;
            out[4 * i + 2] = 0.0f;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        if (j.type == 4 || j.type == 5 || j.type == 7) {  // reaction force and torque at inv_dt = 1
            out[4 * i] = j.linearImpulse.x;
            out[4 * i + 1] = j.linearImpulse.y;
            out[4 * i + 2] = j.angularImpulse;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        if (j.type == 1 || j.type == 2) {  // reaction force at inv_dt = 1: impulse along u
            out[4 * i] = j.impulse.x * j.u.x;
            out[4 * i + 1] = j.impulse.x * j.u.y;
            out[4 * i + 2] = 0.0f;
            out[4 * i + 3] = 0.0f;
            continue;
        }
        out[4 * i] = j.impulse.x;
        out[4 * i + 1] = j.impulse.y;
        out[4 * i + 2] = j.impulse.z;
        out[4 * i + 3] = j.motorImpulse;
    }
}

int orc_world_step(void* h, float dt, int vel, int pos, int nsteps) {
    World* w = (World*)h;
    for (int i = 0; i < nsteps; ++i) w->step(dt, vel, pos);
    return w->unsupported ? B2G_ERR_UNSUPPORTED : B2G_OK;
}

int orc_world_body_count(void* h) { return (int)((World*)h)->bodies.size(); }
int orc_world_fixture_count(void* h) { return (int)((World*)h)->fixtures.size(); }

void orc_world_get_bodies(void* h, float* out) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        const Body& b = w->bodies[i];
        float* o = out + 12 * i;
        o[0] = b.xf.p.x; o[1] = b.xf.p.y; o[2] = b.xf.q.s; o[3] = b.xf.q.c;
        o[4] = b.c.x; o[5] = b.c.y; o[6] = b.a;
        o[7] = b.v.x; o[8] = b.v.y; o[9] = b.w;
        o[10] = (b.flags & kAwake) ? 1.0f : 0.0f;
        o[11] = b.sleepTime;
    }
}

int orc_world_contact_count(void* h) { return ((World*)h)->contactCount; }

int orc_world_get_contacts(void* h, int32_t* meta, b2g_manifold* manifolds, int cap) {
    World* w = (World*)h;
    int n = 0;
    for (int ci = w->contactHead; ci != -1; ci = w->contacts[ci].next) {
        const Contact& c = w->contacts[ci];
        if (n < cap) {
            meta[4 * n + 0] = c.fA;
            meta[4 * n + 1] = c.fB;
            meta[4 * n + 2] = ((c.flags & cTouching) ? 1 : 0) | ((c.flags & cEnabled) ? 2 : 0);
            meta[4 * n + 3] = 0;
            b2g_manifold m = c.manifold;
            for (int i = m.pointCount; i < 2; ++i) memset(&m.points[i], 0, sizeof(m.points[i]));
            if (m.pointCount == 0) { m.type = 0; m.localNormal = g(mk(0, 0)); m.localPoint = g(mk(0, 0)); }
            manifolds[n] = m;
        }
        ++n;
    }
    return n;
}

void orc_world_get_fat_aabbs(void* h, b2g_aabb* out) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->fixtures.size(); ++i) {
        const AABB& a = w->fixtures[i].fat;
        out[i] = {a.lo.x, a.lo.y, a.hi.x, a.hi.y};
    }
}

void orc_world_get_shapes(void* h, b2g_shape* out) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->fixtures.size(); ++i) out[i] = w->fixtures[i].shape;
}

// Per fixture proxy: owning body index; radius, friction, restitution.
void orc_world_fixture_info(void* h, int32_t* body, float* material) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->fixtures.size(); ++i) {
        const Fixture& f = w->fixtures[i];
        body[i] = f.body;
        material[3 * i + 0] = f.shape.radius;
        material[3 * i + 1] = f.friction;
        material[3 * i + 2] = f.restitution;
    }
}

// Per body: mass and rotational inertia about the centre of mass.
void orc_world_body_mass(void* h, float* out) {
    World* w = (World*)h;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        out[2 * i + 0] = w->bodies[i].mass;
        out[2 * i + 1] = w->bodies[i].I;
    }
}

void orc_world_penetration(void* h, double* out) {
    World* w = (World*)h;
    double sum = 0.0, mx = 0.0;
    long np = 0;
    for (int ci = w->contactHead; ci != -1; ci = w->contacts[ci].next) {
        const Contact& c = w->contacts[ci];
        if (!(c.flags & cTouching)) continue;
        const Fixture& fA = w->fixtures[c.fA];
        const Fixture& fB = w->fixtures[c.fB];
        WorldManifold wm;
        worldManifold(&wm, &c.manifold, w->bodies[fA.body].xf, fA.shape.radius, w->bodies[fB.body].xf,
                      fB.shape.radius);
        for (int i = 0; i < c.manifold.pointCount; ++i) {
            double p = wm.separations[i] < 0.0f ? -(double)wm.separations[i] : 0.0;
            sum += p;
            mx = p > mx ? p : mx;
            ++np;
        }
    }
    out[0] = np ? sum / np : 0.0;
    out[1] = mx;
    out[2] = (double)np;
}

void orc_world_events(void* h, long* out) {
    World* w = (World*)h;
    out[0] = w->events.begin;
    out[1] = w->events.end;
}

int orc_collide(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b, const b2g_transform* xfB,
                b2g_manifold* out) {
    clearManifold(out);
    return collide(out, a, xfOf(xfA), b, xfOf(xfB)) ? 0 : -1;
}

float orc_find_max_separation(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                              const b2g_transform* xfB, int* edge) {
    return findMaxSeparation(edge, a, xfOf(xfA), b, xfOf(xfB));
}

void orc_find_incident_edge(const b2g_shape* a, const b2g_transform* xfA, int edge, const b2g_shape* b,
                            const b2g_transform* xfB, float* out, uint32_t* ids) {
    ClipVertex c[2];
    findIncidentEdge(c, a, xfOf(xfA), edge, b, xfOf(xfB));
    out[0] = c[0].v.x; out[1] = c[0].v.y; out[2] = c[1].v.x; out[3] = c[1].v.y;
    ids[0] = c[0].id;
    ids[1] = c[1].id;
}

void orc_transform(float x, float y, float angle, b2g_transform* out) {
    Rot q = rotOf(angle);
    out->px = x; out->py = y; out->s = q.s; out->c = q.c;
}

void orc_world_manifold(const b2g_manifold* m, const b2g_transform* xfA, float radiusA, const b2g_transform* xfB,
                        float radiusB, float* out) {
    for (int i = 0; i < 8; ++i) out[i] = 0.0f;
    if (m->pointCount == 0) return;
    WorldManifold wm;
    worldManifold(&wm, m, xfOf(xfA), radiusA, xfOf(xfB), radiusB);
    out[0] = wm.normal.x; out[1] = wm.normal.y;
    for (int i = 0; i < m->pointCount; ++i) {
        out[2 + 2 * i] = wm.points[i].x;
        out[3 + 2 * i] = wm.points[i].y;
        out[6 + i] = wm.separations[i];
    }
}

void orc_polygon_set(const float* pts, int n, b2g_shape* out) {
    V2 p[8];
    for (int i = 0; i < n && i < 8; ++i) p[i] = mk(pts[2 * i], pts[2 * i + 1]);
    polygonSet(out, p, n);
}

void orc_polygon_box(float hx, float hy, float cx, float cy, float angle, int rotated, b2g_shape* out) {
    if (rotated) setAsBoxRotated(out, hx, hy, mk(cx, cy), angle);
    else setAsBox(out, hx, hy);
}

void orc_shape_mass(const b2g_shape* s, float density, float* out) {
    MassData md = shapeMass(s, density);
    out[0] = md.mass; out[1] = md.center.x; out[2] = md.center.y; out[3] = md.I;
}

void orc_shape_aabb(const b2g_shape* s, const b2g_transform* xf, b2g_aabb* out) {
    AABB a = shapeAABB(s, xfOf(xf));
    *out = {a.lo.x, a.lo.y, a.hi.x, a.hi.y};
}

// Selects the trig used inside Step (0 glibc, 1 rounded double); process-wide.
void orc_set_trig_mode(int mode) { g_trigMode = mode; }

// glibc sinf/cosf, the functions b2Rot::Set calls (REF b2Math.h:302-311).
// GJK distance without radii from an empty cache; out = pointA(2),
// pointB(2), distance, (unused), simplex count.
void orc_distance(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b, const b2g_transform* xfB,
                  float* out) {
    const Proxy pa = proxyOf(a), pb = proxyOf(b);
    Xf ta, tb;
    ta.p = mk(xfA->px, xfA->py); ta.q.s = xfA->s; ta.q.c = xfA->c;
    tb.p = mk(xfB->px, xfB->py); tb.q.s = xfB->s; tb.q.c = xfB->c;
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    V2 wa, wb;
    float d;
    gjk(&wa, &wb, &d, &cache, pa, ta, pb, tb);
    out[0] = wa.x; out[1] = wa.y; out[2] = wb.x; out[3] = wb.y;
    out[4] = d; out[5] = 0.0f; out[6] = (float)cache.count;
}

// Time of impact of two swept shapes; returns the state (unknown, failed,
// overlapped, touching, separated), *t the time. Trig as set by
// orc_set_trig_mode.
int orc_time_of_impact(const b2g_shape* a, const b2g_sweep* swA, const b2g_shape* b, const b2g_sweep* swB,
                       float tMax, float* t) {
    Sweep s[2];
    const b2g_sweep* in[2] = {swA, swB};
    for (int k = 0; k < 2; ++k) {
        s[k].localCenter = mk(in[k]->localCenter[0], in[k]->localCenter[1]);
        s[k].c0 = mk(in[k]->c0[0], in[k]->c0[1]);
        s[k].c = mk(in[k]->c[0], in[k]->c[1]);
        s[k].a0 = in[k]->a0;
        s[k].a = in[k]->a;
        s[k].alpha0 = in[k]->alpha0;
    }
    return timeOfImpact(t, proxyOf(a), proxyOf(b), s[0], s[1], tMax);
}

void orc_sincos(const float* a, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) { s[i] = sinf(a[i]); c[i] = cosf(a[i]); }
}

// Broad-phase checker: the pair set UpdatePairs reports for the given fat
// AABBs and moved flags, by brute force over all proxies.
int orc_pairs_bruteforce(const b2g_aabb* fat, const uint8_t* moved, int n, int32_t* out, int cap) {
    std::vector<std::pair<int, int>> pairs;
    for (int q = 0; q < n; ++q) {
        if (!moved[q]) continue;
        AABB a = {mk(fat[q].lx, fat[q].ly), mk(fat[q].ux, fat[q].uy)};
        for (int p = 0; p < n; ++p) {
            if (p == q) continue;
            AABB b = {mk(fat[p].lx, fat[p].ly), mk(fat[p].ux, fat[p].uy)};
            if (overlap(a, b)) pairs.push_back(std::make_pair(std::min(p, q), std::max(p, q)));
        }
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    const int np = (int)pairs.size();
    for (int i = 0; i < np && i < cap; ++i) { out[2 * i] = pairs[i].first; out[2 * i + 1] = pairs[i].second; }
    return np;
}

// The fat-AABB rule alone: returns 1 and writes the new fat AABB when the
// proxy must be reinserted (moved), 0 when the old one still contains it.
int orc_fat_rule(const b2g_aabb* fat, const b2g_aabb* tight, float dx, float dy, b2g_aabb* out) {
    World w;
    Fixture f;
    f.fat = {mk(fat->lx, fat->ly), mk(fat->ux, fat->uy)};
    w.fixtures.push_back(f);
    AABB t = {mk(tight->lx, tight->ly), mk(tight->ux, tight->uy)};
    w.moveProxy(0, t, mk(dx, dy));
    const AABB& r = w.fixtures[0].fat;
    *out = {r.lo.x, r.lo.y, r.hi.x, r.hi.y};
    return (int)w.moveBuffer.size();
}

}  // extern "C"
