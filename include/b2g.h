/*
 * b2g.h - C ABI of the MI355X (gfx950) Box2D step library, libb2g.so.
 *
 * This is the thin boundary between the host C++ `box2d::` API
 * (include/box2d/*.h) and the HIP kernels. Only plain C types cross it:
 * pointers, sizes, POD structs. Every call returns an int status
 * (B2G_OK = 0, negative on error); b2g_last_error() gives the message.
 * Device memory is owned by the context. One context per host thread;
 * calls on one context are not re-entrant.
 *
 * Each entry point names the reference interface it stands behind
 * (paths relative to the reference tree; REF = Regression/box2d-ref/Box2DRef,
 * the upstream Box2D 2.3.2 whose arithmetic the fork's hot path keeps).
 */
#ifndef B2G_H
#define B2G_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define B2G_OK 0
#define B2G_ERR_ARG (-1)        /* bad argument (null pointer, bad index) */
#define B2G_ERR_CAPACITY (-2)   /* more bodies/fixtures/contacts than reserved */
#define B2G_ERR_HIP (-3)        /* a HIP runtime call failed */
#define B2G_ERR_UNSUPPORTED (-4)/* a feature the GPU path does not implement */
#define B2G_ERR_STATE (-5)      /* call not valid in the current state */

/* ----------------------------------------------------------- data layout */
/* Shape kinds follow b2Shape::Type (Box2D/Collision/Shapes/b2Shape.h:47-54). */
#define B2G_CIRCLE 0
#define B2G_EDGE 1
#define B2G_POLYGON 2

/* Manifold kinds follow b2Manifold::Type (Box2D/Collision/b2Collision.h:100-105). */
#define B2G_MANIFOLD_CIRCLES 0
#define B2G_MANIFOLD_FACE_A 1
#define B2G_MANIFOLD_FACE_B 2

typedef struct b2g_vec2 { float x, y; } b2g_vec2;

/* b2Transform {p, q(s, c)} (REF/Common/b2Math.h, b2Transform / b2Rot). */
typedef struct b2g_transform { float px, py, s, c; } b2g_transform;

/* One convex shape in local body coordinates, 176 bytes.
 * polygon: count vertices/normals (<= 8), centroid, radius = 2*linearSlop
 *          (b2PolygonShape, REF/Collision/Shapes/b2PolygonShape.h:80-96)
 * circle : centroid = m_p, radius (b2CircleShape.h:55-62)
 * edge   : verts[0] = m_vertex1, verts[1] = m_vertex2, v0 = m_vertex0,
 *          v3 = m_vertex3, flags bit0 = m_hasVertex0, bit1 = m_hasVertex3
 *          (b2EdgeShape.h:50-62) */
typedef struct b2g_shape {
    int32_t type;
    int32_t count;
    float radius;
    int32_t flags;
    b2g_vec2 centroid;
    b2g_vec2 v0;
    b2g_vec2 v3;
    b2g_vec2 pad;
    b2g_vec2 verts[8];
    b2g_vec2 normals[8];
} b2g_shape;

/* b2ManifoldPoint (REF/Collision/b2Collision.h:72-78); id = b2ContactID.key,
 * bytes (indexA, indexB, typeA, typeB) little-endian. */
typedef struct b2g_manifold_point {
    b2g_vec2 localPoint;
    float normalImpulse;
    float tangentImpulse;
    uint32_t id;
} b2g_manifold_point;

/* b2Manifold (REF/Collision/b2Collision.h:94-112), 64 bytes. */
typedef struct b2g_manifold {
    b2g_manifold_point points[2];
    b2g_vec2 localNormal;
    b2g_vec2 localPoint;
    int32_t type;
    int32_t pointCount;
} b2g_manifold;

/* Fixture-level AABB {lower, upper}. */
typedef struct b2g_aabb { float lx, ly, ux, uy; } b2g_aabb;

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
