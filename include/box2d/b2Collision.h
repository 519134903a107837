// box2d:: collision types of the host API (the fork's
// Box2D/Collision/b2Collision.h and Shapes/*.h surface): AABBs, contact ids,
// manifolds, world manifolds and the circle, edge and polygon shapes.
#pragma once
#include <vector>
#include "b2Math.h"

namespace box2d {

struct b2AABB {
    b2Vec2 lowerBound, upperBound;
    b2Vec2 GetCenter() const { return 0.5f * (lowerBound + upperBound); }
    b2Vec2 GetExtents() const { return 0.5f * (upperBound - lowerBound); }
    float GetPerimeter() const {
        return 2.0f * ((upperBound[0] - lowerBound[0]) + (upperBound[1] - lowerBound[1]));
    }
    void Combine(const b2AABB& a, const b2AABB& b) {
        lowerBound = b2Min(a.lowerBound, b.lowerBound);
        upperBound = b2Max(a.upperBound, b.upperBound);
    }
    bool Contains(const b2AABB& a) const {
        return lowerBound[0] <= a.lowerBound[0] && lowerBound[1] <= a.lowerBound[1] &&
               a.upperBound[0] <= upperBound[0] && a.upperBound[1] <= upperBound[1];
    }
};

inline bool b2TestOverlap(const b2AABB& a, const b2AABB& b) {
    const b2Vec2 d1 = b.lowerBound - a.upperBound;
    const b2Vec2 d2 = a.lowerBound - b.upperBound;
    if (d1[0] > 0.0f || d1[1] > 0.0f) return false;
    if (d2[0] > 0.0f || d2[1] > 0.0f) return false;
    return true;
}

struct b2ContactFeature {
    enum Type { e_vertex = 0, e_face = 1 };
    uint8_t indexA, indexB, typeA, typeB;
};

union b2ContactID {
    b2ContactFeature cf;
    uint32_t key;
};

struct b2ManifoldPoint {
    b2Vec2 localPoint;
    float normalImpulse = 0.0f;
    float tangentImpulse = 0.0f;
    b2ContactID id;
};

struct b2Manifold {
    enum Type { e_circles, e_faceA, e_faceB };
    b2ManifoldPoint points[MAX_MANIFOLD_POINTS];
    b2Vec2 localNormal;
    b2Vec2 localPoint;
    Type type = e_circles;
    int32_t pointCount = 0;
};

struct b2WorldManifold {
    void Initialize(const b2Manifold* manifold, const b2Transform& xfA, float radiusA, const b2Transform& xfB,
                    float radiusB);
    b2Vec2 normal;
    b2Vec2 points[MAX_MANIFOLD_POINTS];
    float separations[MAX_MANIFOLD_POINTS] = {0.0f, 0.0f};
};

struct b2MassData {
    float mass = 0.0f;
    b2Vec2 center;
    float I = 0.0f;
};

// Box2D/Collision/b2Collision.h:158-171: a ray from p1 through p2, cut at
// p1 + maxFraction * (p2 - p1); a hit's normal and fraction along it.
struct b2RayCastInput {
    b2Vec2 p1, p2;
    float maxFraction = 1.0f;
};
struct b2RayCastOutput {
    b2Vec2 normal;
    float fraction = 0.0f;
};

class b2Shape {
public:
    enum Type { e_circle = 0, e_edge = 1, e_polygon = 2, e_chain = 3, e_typeCount = 4 };
    virtual ~b2Shape() {}
    virtual b2Shape* Clone() const = 0;
    virtual int32_t GetChildCount() const { return 1; }
    virtual void ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const = 0;
    virtual void ComputeMass(b2MassData* massData, float density) const = 0;
    // Box2D/Collision/Shapes/b2Shape.h:83-89: the ray against the shape's
    // child at transform xf (the shapes' own RayCast in b2CircleShape.cpp,
    // b2EdgeShape.cpp, b2PolygonShape.cpp, b2ChainShape.cpp).
    virtual bool RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                         int32_t childIndex) const = 0;
    Type GetType() const { return m_type; }
    Type m_type = e_circle;
    float m_radius = 0.0f;
};

class b2CircleShape : public b2Shape {
public:
    b2CircleShape() { m_type = e_circle; m_radius = 0.0f; }
    b2Shape* Clone() const override { return new b2CircleShape(*this); }
    void ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const override;
    void ComputeMass(b2MassData* massData, float density) const override;
    bool RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                 int32_t childIndex) const override;
    b2Vec2 m_p;
};

class b2EdgeShape : public b2Shape {
public:
    b2EdgeShape() { m_type = e_edge; m_radius = POLYGON_RADIUS; }
    void Set(const b2Vec2& v1, const b2Vec2& v2) {
        m_vertex1 = v1;
        m_vertex2 = v2;
        m_hasVertex0 = false;
        m_hasVertex3 = false;
    }
    b2Shape* Clone() const override { return new b2EdgeShape(*this); }
    void ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const override;
    void ComputeMass(b2MassData* massData, float density) const override;
    bool RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                 int32_t childIndex) const override;
    b2Vec2 m_vertex1, m_vertex2;
    b2Vec2 m_vertex0, m_vertex3;
    bool m_hasVertex0 = false, m_hasVertex3 = false;
};

// Box2D/Collision/Shapes/b2ChainShape.h:32-99: a sequence of segments,
// collided two-sided as edges with their neighbours as ghost vertices. Each
// segment is a child with its own broad-phase proxy.
class b2ChainShape : public b2Shape {
public:
    b2ChainShape() { m_type = e_chain; m_radius = POLYGON_RADIUS; }
    b2Shape* Clone() const override { return new b2ChainShape(*this); }
    void Clear() {
        m_vertices.clear();
        m_count = 0;
        m_hasPrevVertex = m_hasNextVertex = false;
    }
    // A closed loop: the first vertex repeated at the end, and the
    // neighbours across the seam as ghosts.
    void CreateLoop(const b2Vec2* vertices, int32_t count);
    // An open chain with free ends (ghosts set with SetPrevVertex / SetNextVertex).
    void CreateChain(const b2Vec2* vertices, int32_t count);
    void SetPrevVertex(const b2Vec2& prevVertex) {
        m_prevVertex = prevVertex;
        m_hasPrevVertex = true;
    }
    void SetNextVertex(const b2Vec2& nextVertex) {
        m_nextVertex = nextVertex;
        m_hasNextVertex = true;
    }
    int32_t GetChildCount() const override { return m_count - 1; }
    void GetChildEdge(b2EdgeShape* edge, int32_t index) const;
    // Child AABB: the segment's box without the skin (b2ChainShape::ComputeAABB).
    void ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const override;
    void ComputeMass(b2MassData* massData, float density) const override;
    bool RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                 int32_t childIndex) const override;
    std::vector<b2Vec2> m_vertices;
    int32_t m_count = 0;
    b2Vec2 m_prevVertex, m_nextVertex;
    bool m_hasPrevVertex = false, m_hasNextVertex = false;
};

class b2PolygonShape : public b2Shape {
public:
    b2PolygonShape() { m_type = e_polygon; m_radius = POLYGON_RADIUS; }
    b2Shape* Clone() const override { return new b2PolygonShape(*this); }
    // Convex hull of up to MAX_POLYGON_VERTICES points (gift wrapping after
    // welding points closer than sqrt(LINEAR_SLOP / 2)).
    void Set(const b2Vec2* points, int32_t count);
    void SetAsBox(float hx, float hy);
    void SetAsBox(float hx, float hy, const b2Vec2& center, float angle);
    void ComputeAABB(b2AABB* aabb, const b2Transform& xf, int32_t childIndex) const override;
    void ComputeMass(b2MassData* massData, float density) const override;
    bool RayCast(b2RayCastOutput* output, const b2RayCastInput& input, const b2Transform& xf,
                 int32_t childIndex) const override;
    int32_t GetVertexCount() const { return m_count; }
    const b2Vec2& GetVertex(int32_t i) const { return m_vertices[i]; }
    b2Vec2 m_centroid;
    b2Vec2 m_vertices[MAX_POLYGON_VERTICES];
    b2Vec2 m_normals[MAX_POLYGON_VERTICES];
    int32_t m_count = 0;
};

}  // namespace box2d
