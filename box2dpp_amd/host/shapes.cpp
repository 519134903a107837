// Host-side geometry of the box2d:: API: polygon hull construction, boxes,
// AABBs, mass properties and world manifolds. These run once per shape at
// creation (or on demand for queries); their FP32 arithmetic follows
// upstream Box2D 2.3.2 so shapes built here are bit-identical to the
// reference's.
#include "box2d/b2Collision.h"

namespace box2d {

// b2PolygonShape::SetAsBox, REF/Collision/Shapes/b2PolygonShape.cpp:32-61
void b2PolygonShape::SetAsBox(float hx, float hy) {
    m_count = 4;
    m_vertices[0] = b2Vec2(-hx, -hy);
    m_vertices[1] = b2Vec2(hx, -hy);
    m_vertices[2] = b2Vec2(hx, hy);
    m_vertices[3] = b2Vec2(-hx, hy);
    m_normals[0] = b2Vec2(0.0f, -1.0f);
    m_normals[1] = b2Vec2(1.0f, 0.0f);
    m_normals[2] = b2Vec2(0.0f, 1.0f);
    m_normals[3] = b2Vec2(-1.0f, 0.0f);
    m_centroid.SetZero();
}

void b2PolygonShape::SetAsBox(float hx, float hy, const b2Vec2& center, float angle) {
    SetAsBox(hx, hy);
    m_centroid = center;
    b2Transform xf;
    xf.p = center;
    xf.q.Set(angle);
    for (int32_t i = 0; i < m_count; ++i) {
        m_vertices[i] = b2Mul(xf, m_vertices[i]);
        m_normals[i] = b2Mul(xf.q, m_normals[i]);
    }
}

namespace {
// Area-weighted centroid about the origin (b2PolygonShape.cpp:96-142).
b2Vec2 centroidOf(const b2Vec2* vs, int32_t count) {
    b2Vec2 c(0.0f, 0.0f);
    float area = 0.0f;
    const b2Vec2 ref(0.0f, 0.0f);
    const float third = 1.0f / 3.0f;
    for (int32_t i = 0; i < count; ++i) {
        const b2Vec2 p2 = vs[i];
        const b2Vec2 p3 = i + 1 < count ? vs[i + 1] : vs[0];
        const b2Vec2 e1 = p2 - ref, e2 = p3 - ref;
        const float half = 0.5f * b2Cross(e1, e2);
        area += half;
        c += half * third * (ref + p2 + p3);
    }
    c *= 1.0f / area;
    return c;
}
}  // namespace

// b2PolygonShape::Set, REF/Collision/Shapes/b2PolygonShape.cpp:144-245
void b2PolygonShape::Set(const b2Vec2* points, int32_t count) {
    if (count < 3) {
        SetAsBox(1.0f, 1.0f);
        return;
    }
    const int32_t n0 = b2Min(count, MAX_POLYGON_VERTICES);
    b2Vec2 ps[MAX_POLYGON_VERTICES];
    int32_t n = 0;
    for (int32_t i = 0; i < n0; ++i) {
        bool unique = true;
        for (int32_t j = 0; j < n && unique; ++j)
            if (b2DistanceSquared(points[i], ps[j]) < 0.5f * LINEAR_SLOP) unique = false;
        if (unique) ps[n++] = points[i];
    }
    if (n < 3) {
        SetAsBox(1.0f, 1.0f);
        return;
    }
    // Rightmost point (lowest y on ties) starts the gift wrap.
    int32_t start = 0;
    float x0 = ps[0][0];
    for (int32_t i = 1; i < n; ++i) {
        const float x = ps[i][0];
        if (x > x0 || (x == x0 && ps[i][1] < ps[start][1])) {
            start = i;
            x0 = x;
        }
    }
    int32_t hull[MAX_POLYGON_VERTICES];
    int32_t m = 0;
    int32_t cur = start;
    for (;;) {
        hull[m] = cur;
        int32_t next = 0;
        for (int32_t j = 1; j < n; ++j) {
            if (next == cur) {
                next = j;
                continue;
            }
            const b2Vec2 r = ps[next] - ps[hull[m]];
            const b2Vec2 v = ps[j] - ps[hull[m]];
            const float c = b2Cross(r, v);
            if (c < 0.0f) next = j;
            if (c == 0.0f && v.LengthSquared() > r.LengthSquared()) next = j;
        }
        ++m;
        cur = next;
        if (next == start) break;
    }
    if (m < 3) {
        SetAsBox(1.0f, 1.0f);
        return;
    }
    m_count = m;
    for (int32_t i = 0; i < m; ++i) m_vertices[i] = ps[hull[i]];
    for (int32_t i = 0; i < m; ++i) {
        const b2Vec2 edge = m_vertices[i + 1 < m ? i + 1 : 0] - m_vertices[i];
        m_normals[i] = b2Cross(edge, 1.0f);
        b2Normalize(m_normals[i]);
    }
    m_centroid = centroidOf(m_vertices, m);
}

void b2PolygonShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const {
    B2_NOT_USED(childIndex);
    b2Vec2 lower = b2Mul(xf, m_vertices[0]);
    b2Vec2 upper = lower;
    for (int32_t i = 1; i < m_count; ++i) {
        const b2Vec2 v = b2Mul(xf, m_vertices[i]);
        lower = b2Min(lower, v);
        upper = b2Max(upper, v);
    }
    const b2Vec2 r(m_radius, m_radius);
    aabb->lowerBound = lower - r;
    aabb->upperBound = upper + r;
}

// b2PolygonShape::ComputeMass, REF/Collision/Shapes/b2PolygonShape.cpp:354-437:
// triangle fan about the vertex mean.
void b2PolygonShape::ComputeMass(b2MassData* md, float density) const {
    b2Vec2 center(0.0f, 0.0f);
    float area = 0.0f;
    float I = 0.0f;
    b2Vec2 s(0.0f, 0.0f);
    for (int32_t i = 0; i < m_count; ++i) s += m_vertices[i];
    s *= 1.0f / m_count;
    const float third = 1.0f / 3.0f;
    for (int32_t i = 0; i < m_count; ++i) {
        const b2Vec2 e1 = m_vertices[i] - s;
        const b2Vec2 e2 = i + 1 < m_count ? m_vertices[i + 1] - s : m_vertices[0] - s;
        const float D = b2Cross(e1, e2);
        const float tri = 0.5f * D;
        area += tri;
        center += tri * third * (e1 + e2);
        const float ex1 = e1[0], ey1 = e1[1], ex2 = e2[0], ey2 = e2[1];
        const float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
        const float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
        I += (0.25f * third * D) * (intx2 + inty2);
    }
    md->mass = density * area;
    center *= 1.0f / area;
    md->center = center + s;
    md->I = density * I;
    md->I += md->mass * (b2Dot(md->center, md->center) - b2Dot(center, center));
}

void b2CircleShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const {
    B2_NOT_USED(childIndex);
    const b2Vec2 p = xf.p + b2Mul(xf.q, m_p);
    aabb->lowerBound = b2Vec2(p[0] - m_radius, p[1] - m_radius);
    aabb->upperBound = b2Vec2(p[0] + m_radius, p[1] + m_radius);
}

void b2CircleShape::ComputeMass(b2MassData* md, float density) const {
    md->mass = density * B2_PI * m_radius * m_radius;
    md->center = m_p;
    md->I = md->mass * (0.5f * m_radius * m_radius + b2Dot(m_p, m_p));
}

void b2EdgeShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const {
    B2_NOT_USED(childIndex);
    const b2Vec2 v1 = b2Mul(xf, m_vertex1);
    const b2Vec2 v2 = b2Mul(xf, m_vertex2);
    const b2Vec2 r(m_radius, m_radius);
    aabb->lowerBound = b2Min(v1, v2) - r;
    aabb->upperBound = b2Max(v1, v2) + r;
}

void b2EdgeShape::ComputeMass(b2MassData* md, float density) const {
    B2_NOT_USED(density);
    md->mass = 0.0f;
    md->center = 0.5f * (m_vertex1 + m_vertex2);
    md->I = 0.0f;
}

// b2ChainShape, REF/Collision/Shapes/b2ChainShape.cpp:39-196
void b2ChainShape::CreateLoop(const b2Vec2* vertices, int32_t count) {
    m_vertices.assign(vertices, vertices + count);
    m_vertices.push_back(vertices[0]);
    m_count = count + 1;
    m_prevVertex = m_vertices[m_count - 2];
    m_nextVertex = m_vertices[1];
    m_hasPrevVertex = true;
    m_hasNextVertex = true;
}

void b2ChainShape::CreateChain(const b2Vec2* vertices, int32_t count) {
    m_vertices.assign(vertices, vertices + count);
    m_count = count;
    m_hasPrevVertex = false;
    m_hasNextVertex = false;
    m_prevVertex = b2Vec2(0.0f, 0.0f);
    m_nextVertex = b2Vec2(0.0f, 0.0f);
}

void b2ChainShape::GetChildEdge(b2EdgeShape* edge, int32_t index) const {
    edge->m_radius = m_radius;
    edge->m_vertex1 = m_vertices[index];
    edge->m_vertex2 = m_vertices[index + 1];
    if (index > 0) {
        edge->m_vertex0 = m_vertices[index - 1];
        edge->m_hasVertex0 = true;
    } else {
        edge->m_vertex0 = m_prevVertex;
        edge->m_hasVertex0 = m_hasPrevVertex;
    }
    if (index < m_count - 2) {
        edge->m_vertex3 = m_vertices[index + 2];
        edge->m_hasVertex3 = true;
    } else {
        edge->m_vertex3 = m_nextVertex;
        edge->m_hasVertex3 = m_hasNextVertex;
    }
}

void b2ChainShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const {
    const int32_t i1 = childIndex;
    const int32_t i2 = childIndex + 1 == m_count ? 0 : childIndex + 1;
    const b2Vec2 v1 = b2Mul(xf, m_vertices[i1]);
    const b2Vec2 v2 = b2Mul(xf, m_vertices[i2]);
    aabb->lowerBound = b2Min(v1, v2);
    aabb->upperBound = b2Max(v1, v2);
}

void b2ChainShape::ComputeMass(b2MassData* md, float density) const {
    B2_NOT_USED(density);
    md->mass = 0.0f;
    md->center = b2Vec2(0.0f, 0.0f);
    md->I = 0.0f;
}

// b2WorldManifold::Initialize, REF/Collision/b2Collision.cpp:22-86
void b2WorldManifold::Initialize(const b2Manifold* m, const b2Transform& xfA, float radiusA, const b2Transform& xfB,
                                 float radiusB) {
    if (m->pointCount == 0) return;
    if (m->type == b2Manifold::e_circles) {
        normal = b2Vec2(1.0f, 0.0f);
        const b2Vec2 pointA = b2Mul(xfA, m->localPoint);
        const b2Vec2 pointB = b2Mul(xfB, m->points[0].localPoint);
        if (b2DistanceSquared(pointA, pointB) > FLT_EPSILON * FLT_EPSILON) {
            normal = pointB - pointA;
            b2Normalize(normal);
        }
        const b2Vec2 cA = pointA + radiusA * normal;
        const b2Vec2 cB = pointB - radiusB * normal;
        points[0] = 0.5f * (cA + cB);
        separations[0] = b2Dot(cB - cA, normal);
    } else if (m->type == b2Manifold::e_faceA) {
        normal = b2Mul(xfA.q, m->localNormal);
        const b2Vec2 planePoint = b2Mul(xfA, m->localPoint);
        for (int32_t i = 0; i < m->pointCount; ++i) {
            const b2Vec2 clipPoint = b2Mul(xfB, m->points[i].localPoint);
            const b2Vec2 cA = clipPoint + (radiusA - b2Dot(clipPoint - planePoint, normal)) * normal;
            const b2Vec2 cB = clipPoint - radiusB * normal;
            points[i] = 0.5f * (cA + cB);
            separations[i] = b2Dot(cB - cA, normal);
        }
    } else {
        normal = b2Mul(xfB.q, m->localNormal);
        const b2Vec2 planePoint = b2Mul(xfB, m->localPoint);
        for (int32_t i = 0; i < m->pointCount; ++i) {
            const b2Vec2 clipPoint = b2Mul(xfA, m->points[i].localPoint);
            const b2Vec2 cB = clipPoint + (radiusB - b2Dot(clipPoint - planePoint, normal)) * normal;
            const b2Vec2 cA = clipPoint - radiusA * normal;
            points[i] = 0.5f * (cA + cB);
            separations[i] = b2Dot(cA - cB, normal);
        }
        normal = -normal;
    }
}

// ------------------------------------------------------------ ray casts
// The ray p1 + t (p2 - p1), t in [0, maxFraction], against each shape; the
// reported normal points against the ray.

// Circle (Box2D/Collision/Shapes/b2CircleShape.cpp:56-89): the smaller root
// of |s + t r|^2 = radius^2, s = p1 - centre, r = p2 - p1.
bool b2CircleShape::RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                            int32_t childIndex) const {
    (void)childIndex;
    const b2Vec2 centre = xf.p + b2Mul(xf.q, m_p);
    const b2Vec2 sv = input.p1 - centre;
    const float b = b2Dot(sv, sv) - m_radius * m_radius;
    const b2Vec2 r = input.p2 - input.p1;
    const float c = b2Dot(sv, r);
    const float rr = b2Dot(r, r);
    const float sigma = c * c - rr * b;
    if (sigma < 0.0f || rr < FLT_EPSILON) return false;
    float a = -(c + b2Sqrt(sigma));
    if (0.0f <= a && a <= input.maxFraction * rr) {
        a /= rr;
        output->fraction = a;
        output->normal = sv + a * r;
        b2Normalize(output->normal);
        return true;
    }
    return false;
}

// Edge, both sides (Box2D/Collision/Shapes/b2EdgeShape.cpp:88-137): the ray
// meets the edge's line at t, the hit must lie between the vertices.
bool b2EdgeShape::RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                          int32_t childIndex) const {
    (void)childIndex;
    const b2Vec2 p1 = b2MulT(xf.q, input.p1 - xf.p);
    const b2Vec2 p2 = b2MulT(xf.q, input.p2 - xf.p);
    const b2Vec2 d = p2 - p1;
    const b2Vec2 e = m_vertex2 - m_vertex1;
    b2Vec2 normal(e[1], -e[0]);
    b2Normalize(normal);
    const float numerator = b2Dot(normal, m_vertex1 - p1);
    const float denominator = b2Dot(normal, d);
    if (denominator == 0.0f) return false;
    const float t = numerator / denominator;
    if (t < 0.0f || input.maxFraction < t) return false;
    const b2Vec2 q = p1 + t * d;
    const float ee = b2Dot(e, e);
    if (ee == 0.0f) return false;
    const float sPos = b2Dot(q - m_vertex1, e) / ee;
    if (sPos < 0.0f || 1.0f < sPos) return false;
    output->fraction = t;
    output->normal = numerator > 0.0f ? -b2Mul(xf.q, normal) : b2Mul(xf.q, normal);
    return true;
}

// A chain child is its edge (Box2D/Collision/Shapes/b2ChainShape.cpp:145-163).
bool b2ChainShape::RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                           int32_t childIndex) const {
    b2EdgeShape edge;
    GetChildEdge(&edge, childIndex);
    return edge.RayCast(output, input, xf, 0);
}

// Convex polygon (Box2D/Collision/Shapes/b2PolygonShape.cpp:279-333): the
// ray clipped against each edge's half-plane in the polygon frame; the last
// edge that raised the entry bound is the one hit.
bool b2PolygonShape::RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                             int32_t childIndex) const {
    (void)childIndex;
    const b2Vec2 p1 = b2MulT(xf.q, input.p1 - xf.p);
    const b2Vec2 p2 = b2MulT(xf.q, input.p2 - xf.p);
    const b2Vec2 d = p2 - p1;
    float lower = 0.0f, upper = input.maxFraction;
    int32_t index = -1;
    for (int32_t i = 0; i < m_count; ++i) {
        const float numerator = b2Dot(m_normals[i], m_vertices[i] - p1);
        const float denominator = b2Dot(m_normals[i], d);
        if (denominator == 0.0f) {
            if (numerator < 0.0f) return false;
        } else if (denominator < 0.0f && numerator < lower * denominator) {
            lower = numerator / denominator;
            index = i;
        } else if (denominator > 0.0f && numerator < upper * denominator) {
            upper = numerator / denominator;
        }
        if (upper < lower) return false;
    }
    if (index < 0) return false;
    output->fraction = lower;
    output->normal = b2Mul(xf.q, m_normals[index]);
    return true;
}

}  // namespace box2d
