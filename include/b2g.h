/*
 * b2g.h - C ABI of the MI355X (gfx950) Box2D step library, libb2g.so.
 *
 * This is the thin boundary between the host C++ `box2d::` API
 * (headers under include/box2d) and the HIP kernels. Only plain C types cross it:
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

/* b2Sweep (REF/Common/b2Math.h:383-405): the motion of a body over a step,
 * centre of mass from c0 / a0 at time alpha0 to c / a at time 1. 36 bytes. */
typedef struct b2g_sweep {
    float localCenter[2];
    float c0[2];
    float c[2];
    float a0;
    float a;
    float alpha0;
} b2g_sweep;

/* One convex shape in local body coordinates, 176 bytes.
 * polygon: count vertices/normals (<= 8), centroid, radius = 2*linearSlop
 *          (b2PolygonShape, REF/Collision/Shapes/b2PolygonShape.h:80-96)
 * circle : centroid = m_p, radius (b2CircleShape.h:55-62)
 * edge   : verts[0] = m_vertex1, verts[1] = m_vertex2, v0 = m_vertex0,
 *          v3 = m_vertex3, flags bit0 = m_hasVertex0, bit1 = m_hasVertex3
 *          (b2EdgeShape.h:50-62)
 * chain  : one edge record per child (b2ChainShape::GetChildEdge,
 *          REF/Collision/Shapes/b2ChainShape.cpp:112-143) with flags bit2
 *          set: its AABB has no skin (b2ChainShape::ComputeAABB, :171-187).
 *          A chain fixture is added as its children, in child order. */
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

/* Full simulation state of one body: b2Body's transform, sweep, velocity,
 * accumulated force and mass data (REF/Dynamics/b2Body.h:430-470). The host
 * computes mass data (ResetMassData) at creation; the device owns the rest
 * between steps. flags use b2Body::m_flags bits: 0x2 awake, 0x4 auto-sleep,
 * 0x8 bullet, 0x10 fixed rotation, 0x20 active. type: 0 static,
 * 1 kinematic, 2 dynamic (b2BodyType order). */
typedef struct b2g_body_state {
    int32_t type;
    int32_t flags;
    float xf[4];      /* p.x p.y q.s q.c */
    float lc[2];      /* sweep localCenter */
    float c0[2];
    float a0;
    float c[2];
    float a;
    float v[2];
    float w;
    float force[2];
    float torque;
    float invMass;
    float invI;
    float linearDamping;
    float angularDamping;
    float gravityScale;
    float sleepTime;
} b2g_body_state;

/* One fixture: body index, material, filter and shape
 * (b2FixtureDef, Box2D/Dynamics/b2Fixture.h:59-94). Fixtures get broad-phase
 * proxy ids in the order they are added. */
typedef struct b2g_fixture_def {
    int32_t body;
    int32_t sensor;
    float friction;
    float restitution;
    float density;
    uint32_t categoryBits;
    uint32_t maskBits;
    int32_t groupIndex;
    b2g_shape shape;
} b2g_fixture_def;

/* World-level switches (b2World::SetAllowSleeping / SetWarmStarting /
 * SetContinuousPhysics / SetSubStepping / SetAutoClearForces,
 * Box2D/Dynamics/b2World.h:146-217). continuous runs the time-of-impact
 * phase after the discrete solve (b2World::SolveTOI,
 * REF/Dynamics/b2World.cpp:579-897); subStepping stops it after one event
 * per step. */
typedef struct b2g_world_def {
    float gravity[2];
    int32_t allowSleep;
    int32_t warmStarting;
    int32_t continuous;
    int32_t subStepping;
    int32_t autoClearForces;
} b2g_world_def;

/* Solver modes. COLOURED: graph-coloured sequential impulses over all
 * contacts at once (the throughput path). REFERENCE: one wavefront per
 * island visiting constraints in Box2D's island order. */
#define B2G_MODE_COLOURED 0
#define B2G_MODE_REFERENCE 1

/* b2World::Step arguments (Box2D/Dynamics/b2World.h:96) plus the solver
 * mode and the learned warm-start switches. */
typedef struct b2g_step_params {
    float dt;
    int32_t velocityIterations;
    int32_t positionIterations;
    int32_t mode;
    int32_t learnedWarmStart;   /* 0 off, 1 on: fresh contact points start from the
                                   warm-start network's impulses (needs a network,
                                   b2g_mlp_init / b2g_mlp_set_params) */
    float residualThreshold;    /* > 0: if the learned start leaves an RMS
                                   complementarity residual above this (m/s) after
                                   the velocity iterations, the step is solved again
                                   from the plain start; <= 0: no check */
} b2g_step_params;

/* Counters of the last step. */
typedef struct b2g_step_stats {
    int32_t bodies;
    int32_t fixtures;
    int32_t contacts;          /* contacts alive after the step (b2World::GetContactCount) */
    int32_t constraints;       /* touching contacts solved this step */
    int32_t points;            /* manifold points of those contacts */
    int32_t colours;
    int32_t colouringRounds;
    int32_t newPairs;          /* contacts created by the broad-phase this step */
    int32_t positionIterations;/* position iterations actually run */
    int32_t beginEvents;       /* BeginContact count this step */
    int32_t endEvents;         /* EndContact count this step */
    int32_t overflow;          /* nonzero if a device buffer overflowed */
    float minSeparation;       /* of the last position iteration run */
    float cellSize;            /* broad-phase grid cell */
    int32_t freshPoints;       /* points without a carried impulse (when the learned
                                  start, the training log or a trace is on) */
    int32_t fallback;          /* 1 if the learned start was replaced by the plain one */
    float residual;            /* RMS residual checked for the fallback, -1 if none */
    int32_t toiEvents;         /* TOI events solved this step (continuous collision) */
    int32_t queries;           /* moved proxies whose broad-phase queries ran */
    int32_t scanned;           /* grid entries and wide proxies those queries examined */
    int32_t wideConstraints;   /* constraints in the colours that get a solver launch
                                  each (not the fused trailing colours, not the
                                  overflow colour) */
} b2g_step_stats;

typedef struct b2g_context b2g_context;

/* ---------------------------------------------------------- world calls */

/* Creates a context on HIP device `device`. Replaces the construction of
 * b2World's device-side state (Box2D/Dynamics/b2World.h:48). */
int b2g_create(b2g_context** out, int device);
int b2g_destroy(b2g_context* ctx);
/* Message of the last failing call on ctx (static storage owned by ctx). */
const char* b2g_last_error(const b2g_context* ctx);

/* Pre-sizes device arrays (grows automatically otherwise). */
int b2g_reserve(b2g_context* ctx, int maxBodies, int maxFixtures, int maxContacts);

/* World switches and gravity; see b2g_world_def. */
int b2g_set_world(b2g_context* ctx, const b2g_world_def* def);

/* Appends bodies (b2World::CreateBody, Box2D/Dynamics/b2World.h:74);
 * *first receives the index of the first new body. */
int b2g_add_bodies(b2g_context* ctx, const b2g_body_state* bodies, int count, int* first);
/* Overwrites bodies [first, first+count) (SetTransform, SetLinearVelocity,
 * ApplyForce, SetAwake, mass changes: Box2D/Dynamics/b2Body.h:163-330). */
int b2g_set_bodies(b2g_context* ctx, int first, int count, const b2g_body_state* bodies);
/* Reads bodies [first, first+count) back to the host. */
int b2g_get_bodies(b2g_context* ctx, int first, int count, b2g_body_state* out);

/* Appends fixtures with broad-phase proxies (b2Body::CreateFixture,
 * Box2D/Dynamics/b2Body.h:138); new proxies join the move buffer so the next
 * step finds their pairs (b2World::Step's e_newFixture path,
 * REF/Dynamics/b2World.cpp:903-907). */
int b2g_add_fixtures(b2g_context* ctx, const b2g_fixture_def* fixtures, int count, int* first);

/* Removes fixtures' proxies from the broad-phase (b2Body::DestroyFixture /
 * b2World::DestroyBody, Box2D/Dynamics/b2Body.h:156, b2World.h:80): the
 * bodies of their contacts with manifold points wake now, as the reference's
 * immediate contact destruction wakes them; the contacts themselves end
 * first thing in the next step's collide (EndContact then). */
int b2g_remove_fixtures(b2g_context* ctx, const int32_t* ids, int count);
/* The contact reset of b2Body::SetType (Box2D/Dynamics/b2Body.h:305) for
 * bodies whose type the host changed (their new state uploaded first): each
 * of their contacts is destroyed, its bodies woken now when its manifold has
 * points and the contact dropped first thing in the next step's collide
 * (EndContact then), and their proxies join the move buffer, so the step's
 * broad-phase recreates the pairs that still overlap. */
int b2g_drop_body_contacts(b2g_context* ctx, const int32_t* bodies, int count);

/* The broad-phase order of the proxies: ids[f] is the reference tree's
 * proxy id of fixture f (-1 for a removed one) for every fixture (count =
 * fixtures added so far; call again after adding more). New pairs are then
 * (smaller id, larger id), created in id order with the smaller id's fixture
 * as A, as b2BroadPhase::UpdatePairs sorts its pair buffer
 * (Box2D/Collision/b2BroadPhase.h:193-247, b2BroadPhase.cpp:98-121). Unset,
 * the fixture index is the id: the reference's ids grow with creation order
 * until a proxy is destroyed (b2DynamicTree::AllocateNode / FreeNode,
 * Box2D/Collision/b2DynamicTree.cpp:59-110). One world per context. */
int b2g_set_proxy_ids(b2g_context* ctx, const int32_t* ids, int count);

/* Gives removed fixtures their proxies back (b2Body::SetActive(true),
 * Box2D/Dynamics/b2Body.h:353): each fat box from the shape at its body's
 * device transform, the proxy in the move buffer, so the next step's
 * closing broad-phase finds its pairs. */
int b2g_restore_fixtures(b2g_context* ctx, const int32_t* ids, int count);

/* Updates material and filter of fixtures (b2Fixture::SetFriction /
 * SetRestitution / SetFilterData, Box2D/Dynamics/b2Fixture.h:131-170); a
 * filter change flags the fixture's contacts for filtering and re-queries
 * its proxy, as b2Fixture::Refilter does. */
int b2g_update_fixtures(b2g_context* ctx, const int32_t* ids, const b2g_fixture_def* defs, int count);

/* One time step: b2World::Step(dt, velocityIterations, positionIterations)
 * (Box2D/Dynamics/b2World.cpp:888-954; REF/Dynamics/b2World.cpp:899-964),
 * with the continuous (TOI) phase when the world def asks for it.
 * Asynchronous on the context's stream without TOI: returns once the work
 * is queued; b2g_get_* and b2g_sync wait for it. The TOI phase reads one
 * small record back per event. */
int b2g_step(b2g_context* ctx, const b2g_step_params* params);
/* Blocks until queued work on ctx has finished. */
int b2g_sync(b2g_context* ctx);
int b2g_get_stats(b2g_context* ctx, b2g_step_stats* out);

/* Contacts in table order: meta[4*i] = fixtureA, [4*i+1] = fixtureB (after
 * the registry's shape-order swap), [4*i+2] = flags (bit0 touching,
 * bit1 enabled), [4*i+3] = solver colour (-1: none; coloured mode).
 * *count receives the number of contacts (may exceed cap). */
int b2g_get_contacts(b2g_context* ctx, int32_t* meta, b2g_manifold* manifolds, int cap, int* count);
/* Fat AABB of every proxy, fixture order. */
int b2g_get_fat_aabbs(b2g_context* ctx, b2g_aabb* out, int cap);
/* Mean and max penetration over touching manifold points (from
 * b2WorldManifold separations), and the point count: out[0..2]. */
int b2g_penetration(b2g_context* ctx, double* out);
/* Joint drift: mean and max position error over the joints, and the count
 * of joints measured: |cB + rB - cA - rA| for revolute and weld joints (the
 * error b2RevoluteJoint::SolvePositionConstraints measures,
 * Box2D/Dynamics/Joints/b2RevoluteJoint.cpp:336-343), the error in length
 * for rigid distance joints, the stretch beyond the maximum length for rope
 * joints, |lengthA + ratio lengthB - constant| for pulley joints, the
 * distance of anchor B from the axis line for prismatic and wheel joints,
 * |coordinateA + ratio coordinateB - constant| for gear joints; soft
 * distance, friction, motor and mouse joints are not measured. out[3], FP64 sums on the
 * host; no counterpart in the reference. */
int b2g_joint_drift(b2g_context* ctx, double* out);
/* Kinetic (0.5 m v.v + 0.5 I w^2) and gravitational potential (-m g.c)
 * energy summed over the bodies in FP64; out[2]. No counterpart in the
 * reference. */
int b2g_energy(b2g_context* ctx, double* out);

/* Per-phase GPU time of the steps run while profiling is on, measured with
 * HIP events on the context's stream around every launch of the phase. */
typedef struct b2g_phase_profile {
    char name[24];
    int32_t launches;
    int32_t pad;
    double ms;        /* summed launch durations */
} b2g_phase_profile;
/* on != 0 starts profiling and clears the accumulated times. */
int b2g_set_profiling(b2g_context* ctx, int on);
int b2g_get_profile(b2g_context* ctx, b2g_phase_profile* out, int cap, int* count);

/* ------------------------------------------------------------- joints
 * b2RevoluteJointDef (Box2D/Dynamics/Joints/b2RevoluteJoint.h:37-90) with
 * body indices; joints connect bodies of the same sub-world. */
typedef struct b2g_revolute_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float referenceAngle;
    float lowerAngle, upperAngle;
    float maxMotorTorque, motorSpeed;
    int32_t enableLimit, enableMotor, collideConnected;
} b2g_revolute_def;

/* Appends revolute joints (b2World::CreateJoint, Box2D/Dynamics/b2World.h:85);
 * *first receives the index of the first new joint. Contacts between bodies
 * that a joint keeps from colliding end at the next step. */
int b2g_add_joints(b2g_context* ctx, const b2g_revolute_def* defs, int count, int* first);
/* Joint state for joints [first, first + count): impulse x, y, z and motor
 * impulse, 4 floats each (b2RevoluteJoint::GetReactionForce / Torque /
 * GetMotorTorque read these, Box2D/Dynamics/Joints/b2RevoluteJoint.h:158-166). */
int b2g_get_joints(b2g_context* ctx, int first, int count, float* out);
/* Motor and limit switches of one joint (b2RevoluteJoint::EnableMotor,
 * SetMotorSpeed, SetMaxMotorTorque, EnableLimit, SetLimits,
 * Box2D/Dynamics/Joints/b2RevoluteJoint.cpp:419-471); bodies, anchors and
 * the reference angle keep their creation values. A changed limit switch or
 * limit range drops the limit impulse, as those setters do. */
int b2g_set_joint(b2g_context* ctx, int joint, const b2g_revolute_def* def);

/* b2DistanceJointDef (Box2D/Dynamics/Joints/b2DistanceJoint.h:32-63) with
 * body indices: a rod (frequencyHz 0) or a soft spring-damper. 40 bytes. */
typedef struct b2g_distance_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float length, frequencyHz, dampingRatio;
    int32_t collideConnected;
} b2g_distance_def;
/* Appends distance joints to the same joint table as the revolute ones;
 * b2g_get_joints then reports the impulse along the rod and the rod's unit
 * direction of the last step (impulse, u.x, u.y, 0).
 * They run with the coloured solver (b2DistanceJoint,
 * REF/Dynamics/Joints/b2DistanceJoint.cpp:63-221). */
int b2g_add_distance_joints(b2g_context* ctx, const b2g_distance_def* defs, int count, int* first);
/* b2DistanceJoint::SetLength / SetFrequency / SetDampingRatio. */
int b2g_set_distance_joint(b2g_context* ctx, int joint, float length, float frequencyHz, float dampingRatio);

/* b2RopeJointDef (Box2D/Dynamics/Joints/b2RopeJoint.h:30-50) with body
 * indices: the anchors may come no further apart than maxLength. 32 bytes. */
typedef struct b2g_rope_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float maxLength;
    int32_t collideConnected;
} b2g_rope_def;
/* Appends rope joints (b2RopeJoint, REF/Dynamics/Joints/b2RopeJoint.cpp:48-196)
 * to the joint table; b2g_get_joints reports the impulse along the rope, the
 * rope's unit direction of the last step and its limit state (impulse, u.x,
 * u.y, state: 0 slack, 2 taut as b2LimitState counts). */
int b2g_add_rope_joints(b2g_context* ctx, const b2g_rope_def* defs, int count, int* first);
/* b2RopeJoint::SetMaxLength. */
int b2g_set_rope_joint(b2g_context* ctx, int joint, float maxLength);

/* b2WeldJointDef (Box2D/Dynamics/Joints/b2WeldJoint.h:29-60) with body
 * indices: rigid (frequencyHz 0) or with a soft angular spring. 40 bytes. */
typedef struct b2g_weld_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float referenceAngle, frequencyHz, dampingRatio;
    int32_t collideConnected;
} b2g_weld_def;
/* Appends weld joints (b2WeldJoint, REF/Dynamics/Joints/b2WeldJoint.cpp:60-308);
 * b2g_get_joints reports the linear and angular impulse (x, y, z, 0). */
int b2g_add_weld_joints(b2g_context* ctx, const b2g_weld_def* defs, int count, int* first);
/* b2WeldJoint::SetFrequency / SetDampingRatio. */
int b2g_set_weld_joint(b2g_context* ctx, int joint, float frequencyHz, float dampingRatio);

/* b2FrictionJointDef (Box2D/Dynamics/Joints/b2FrictionJoint.h:27-53) with
 * body indices: top-down friction, a force and a torque up to the limits
 * against relative motion. 36 bytes. */
typedef struct b2g_friction_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float maxForce, maxTorque;
    int32_t collideConnected;
} b2g_friction_def;
/* Appends friction joints (b2FrictionJoint,
 * REF/Dynamics/Joints/b2FrictionJoint.cpp:59-195); b2g_get_joints reports
 * the linear and angular impulse (x, y, angular, 0). */
int b2g_add_friction_joints(b2g_context* ctx, const b2g_friction_def* defs, int count, int* first);
/* b2FrictionJoint::SetMaxForce / SetMaxTorque. */
int b2g_set_friction_joint(b2g_context* ctx, int joint, float maxForce, float maxTorque);

/* b2MotorJointDef (Box2D/Dynamics/Joints/b2MotorJoint.h:27-56) with body
 * indices: drives body B to an offset from body A (in A's frame) with a
 * force and torque up to the limits. 36 bytes. */
typedef struct b2g_motor_def {
    int32_t bodyA, bodyB;
    float linearOffset[2];
    float angularOffset;
    float maxForce, maxTorque, correctionFactor;
    int32_t collideConnected;
} b2g_motor_def;
/* Appends motor joints (b2MotorJoint, REF/Dynamics/Joints/b2MotorJoint.cpp:63-205);
 * b2g_get_joints reports the linear and angular impulse (x, y, angular, 0). */
int b2g_add_motor_joints(b2g_context* ctx, const b2g_motor_def* defs, int count, int* first);
/* b2MotorJoint::SetLinearOffset, SetAngularOffset, SetMaxForce, SetMaxTorque
 * and SetCorrectionFactor in one call (bodies and collideConnected of `def`
 * are ignored). */
int b2g_set_motor_joint(b2g_context* ctx, int joint, const b2g_motor_def* def);

/* b2PulleyJointDef (Box2D/Dynamics/Joints/b2PulleyJoint.h:30-70) with body
 * indices: lengthA + ratio * lengthB stays at its value at creation. 56
 * bytes. */
typedef struct b2g_pulley_def {
    int32_t bodyA, bodyB;
    float groundAnchorA[2];
    float groundAnchorB[2];
    float localAnchorA[2];
    float localAnchorB[2];
    float lengthA, lengthB, ratio;
    int32_t collideConnected;
} b2g_pulley_def;
/* Appends pulley joints (b2PulleyJoint, REF/Dynamics/Joints/b2PulleyJoint.cpp:57-263);
 * b2g_get_joints reports the impulse, body B's rope direction of the last
 * step and the pulley's constant lengthA + ratio * lengthB (impulse, uB.x,
 * uB.y, constant); ratio must be positive. */
int b2g_add_pulley_joints(b2g_context* ctx, const b2g_pulley_def* defs, int count, int* first);

/* b2MouseJointDef (Box2D/Dynamics/Joints/b2MouseJoint.h:28-53) with body
 * indices, plus what b2MouseJoint takes from body B at creation and step
 * time: the local anchor (b2MulT of body B's transform and the target) and
 * body B's mass (b2Body::GetMass). 44 bytes. */
typedef struct b2g_mouse_def {
    int32_t bodyA, bodyB;
    float target[2];
    float localAnchorB[2];
    float maxForce, frequencyHz, dampingRatio;
    float massB;
    int32_t collideConnected;
} b2g_mouse_def;
/* Appends mouse joints (b2MouseJoint, REF/Dynamics/Joints/b2MouseJoint.cpp:34-200);
 * b2g_get_joints reports the impulse (x, y, 0, 0). */
int b2g_add_mouse_joints(b2g_context* ctx, const b2g_mouse_def* defs, int count, int* first);
/* b2MouseJoint::SetTarget, SetMaxForce, SetFrequency, SetDampingRatio, and
 * body B's mass when it changed (bodies, localAnchorB and collideConnected
 * of `def` are ignored). */
int b2g_set_mouse_joint(b2g_context* ctx, int joint, const b2g_mouse_def* def);

/* b2PrismaticJointDef (Box2D/Dynamics/Joints/b2PrismaticJoint.h:30-92) with
 * body indices: body B slides along localAxisA of body A (normalized on
 * creation), with an optional translation limit and motor. 64 bytes. */
typedef struct b2g_prismatic_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float localAxisA[2];
    float referenceAngle;
    float lowerTranslation, upperTranslation;
    float maxMotorForce, motorSpeed;
    int32_t enableLimit, enableMotor, collideConnected;
} b2g_prismatic_def;
/* Appends prismatic joints (b2PrismaticJoint,
 * REF/Dynamics/Joints/b2PrismaticJoint.cpp:102-488); b2g_get_joints reports
 * the impulses (x, y, limit, motor), b2g_get_joint_reactions the force. */
int b2g_add_prismatic_joints(b2g_context* ctx, const b2g_prismatic_def* defs, int count, int* first);
/* b2PrismaticJoint::EnableLimit, SetLimits, EnableMotor, SetMotorSpeed,
 * SetMaxMotorForce (bodies, anchors, axis and reference angle keep their
 * creation values); a changed limit drops the limit impulse. */
int b2g_set_prismatic_joint(b2g_context* ctx, int joint, const b2g_prismatic_def* def);

/* b2WheelJointDef (Box2D/Dynamics/Joints/b2WheelJoint.h:30-72) with body
 * indices: body B (the wheel) rides along localAxisA of body A on a soft
 * spring, with an optional motor. 56 bytes. */
typedef struct b2g_wheel_def {
    int32_t bodyA, bodyB;
    float localAnchorA[2];
    float localAnchorB[2];
    float localAxisA[2];
    float maxMotorTorque, motorSpeed;
    float frequencyHz, dampingRatio;
    int32_t enableMotor, collideConnected;
} b2g_wheel_def;
/* Appends wheel joints (b2WheelJoint, REF/Dynamics/Joints/b2WheelJoint.cpp:50-327);
 * b2g_get_joints reports (impulse, spring impulse, motor impulse, 0),
 * b2g_get_joint_reactions the force and torque. */
int b2g_add_wheel_joints(b2g_context* ctx, const b2g_wheel_def* defs, int count, int* first);
/* b2WheelJoint::EnableMotor, SetMotorSpeed, SetMaxMotorTorque,
 * SetSpringFrequencyHz, SetSpringDampingRatio. */
int b2g_set_wheel_joint(b2g_context* ctx, int joint, const b2g_wheel_def* def);

/* b2GearJointDef (Box2D/Dynamics/Joints/b2GearJoint.h:24-42) resolved the
 * way the b2GearJoint constructor resolves it
 * (REF/Dynamics/Joints/b2GearJoint.cpp:45-128): joint1 joins bodies C and A
 * (its bodyA and bodyB), joint2 bodies D and B; each is revolute (prismatic
 * flag 0, coordinate = angle difference) or prismatic (coordinate along
 * localAxis); constant = coordinateA + ratio * coordinateB at creation.
 * 92 bytes. */
typedef struct b2g_gear_def {
    int32_t bodyA, bodyB, bodyC, bodyD;
    int32_t prismaticA, prismaticB;
    float localAnchorA[2];
    float localAnchorB[2];
    float localAnchorC[2];
    float localAnchorD[2];
    float localAxisC[2];
    float localAxisD[2];
    float referenceAngleA, referenceAngleB;
    float ratio, constant;
    int32_t collideConnected;
} b2g_gear_def;
/* Appends gear joints (b2GearJoint, REF/Dynamics/Joints/b2GearJoint.cpp:147-393);
 * a gear is coloured with all four of its bodies; b2g_get_joints reports
 * (impulse, JvAC.x, JvAC.y, JwA), b2g_get_joint_reactions the force and
 * torque. */
int b2g_add_gear_joints(b2g_context* ctx, const b2g_gear_def* defs, int count, int* first);
/* b2GearJoint::SetRatio (the constant keeps its creation value). */
int b2g_set_gear_joint(b2g_context* ctx, int joint, float ratio);

/* b2Joint::GetReactionForce and GetReactionTorque at inv_dt = 1 for joints
 * [first, first + count), every type: force x, y, torque, 0 (4 floats each);
 * zeros for removed joints. Multiply by inv_dt for other steps. */
int b2g_get_joint_reactions(b2g_context* ctx, int first, int count, float* out);
/* Removes joints (b2World::DestroyJoint, Box2D/Dynamics/b2World.cpp:267):
 * they stop constraining and connecting islands at the next step; indices of
 * the other joints do not change. */
int b2g_remove_joints(b2g_context* ctx, const int* ids, int count);

/* ------------------------------------------------ learned warm start
 * No counterpart in the reference: its solver starts every new contact
 * point from zero impulse (b2ContactSolver ctor, REF/Dynamics/Contacts/
 * b2ContactSolver.cpp:100-109, with b2Contact::Update's id matching,
 * REF/Dynamics/Contacts/b2Contact.cpp:190-205, carrying the rest).
 *
 * A fresh point is a solved manifold point that starts the step without a
 * carried impulse. Its 16 features: relative normal and tangent velocity,
 * normal (x, y), separation, inverse mass A, B, inverse inertia A, B, point
 * count, normal mass, tangent mass, friction, cross(rA, n), cross(rB, n),
 * restitution bias. The network (16-64-64-64-2, GELU, BF16 weights on the
 * matrix cores, FP32 accumulation) maps them to the normal and tangent
 * impulse in velocity-change units (impulse over the point's normal or
 * tangent mass); the point starts from those times its masses, clamped to
 * the friction cone. Log targets are in the same units. */
#define B2G_MLP_FEATURES 16
#define B2G_MLP_HIDDEN 64
typedef struct b2g_mlp_params {
    float W1[64 * 16];  /* row-major [out][in] */
    float b1[64];
    float W2[64 * 64];
    float b2[64];
    float W3[64 * 64];
    float b3[64];
    float W4[2 * 64];
    float b4[2];
    float xmean[16];    /* feature normalisation; xstd 0 ignores a feature */
    float xstd[16];
    float ymean[2];     /* target normalisation (normal, tangent impulse) */
    float ystd[2];
} b2g_mlp_params;

/* Logs the candidate points of the following steps with the impulses the
 * solver leaves on them, up to `capacity` points (0 stops and frees the
 * log). */
int b2g_log_fresh(b2g_context* ctx, int capacity);
/* Which points the network serves (logging, the learned start and the
 * learned trace curve):
 *   scope 0 (default): fresh points only, the features above;
 *   scope 1: every solved point, the network replacing the carried impulse;
 *            features 9 and 15 are then the carried normal and tangent
 *            impulse in velocity-change units (0 for a fresh point), and the
 *            log takes every logStride-th step (whole steps).
 * A network is trained for one scope. */
int b2g_set_learned_scope(b2g_context* ctx, int scope, int logStride);
int b2g_log_size(b2g_context* ctx, int* count);
int b2g_get_log(b2g_context* ctx, int first, int count, float* features, float* targets);
/* Fresh network with seeded random weights; the normalisation is taken from
 * the log (identity if the log is empty). Resets the optimiser. */
int b2g_mlp_init(b2g_context* ctx, uint32_t seed);
int b2g_mlp_set_params(b2g_context* ctx, const b2g_mlp_params* p);
int b2g_mlp_get_params(b2g_context* ctx, b2g_mlp_params* p);
/* Trains on the log with Adam (mean squared error on normalised targets),
 * minibatches of `batch` points in a seeded order per epoch; loss[e] gets
 * epoch e's mean loss (may be null). */
int b2g_mlp_train(b2g_context* ctx, int epochs, int batch, float lr, uint32_t seed, double* loss);
/* Network outputs for n feature rows (host arrays; y: 2 per row). */
int b2g_mlp_forward(b2g_context* ctx, const float* x, int n, float* y);
/* Gradient of L = 1/(2n) sum |y - t|^2 (normalised targets) with respect to
 * the weights and biases, in b2g_mlp_params layout (normalisation fields
 * zero); *loss receives L. */
int b2g_mlp_gradient(b2g_context* ctx, const float* x, const float* t, int n, b2g_mlp_params* grad, double* loss);

/* Residual trace: while maxIters > 0, each step also solves its velocity
 * constraints from three starts (cold: all impulses zero; warm: carried
 * impulses, fresh points zero; learned: the candidate points of the scope
 * from the network, the others carried) for maxIters iterations each,
 * recording the RMS
 * complementarity residual (FP64 on the host) after the warm start and after
 * every iteration; the step's own result is unchanged. */
int b2g_set_trace(b2g_context* ctx, int maxIters);
/* curves: 3 x (iters + 1) values (cold, warm, learned; NaN where a start
 * was not available); cap in doubles. */
int b2g_get_trace(b2g_context* ctx, double* curves, int cap, int* iters);

/* ---------------------------------------------- per-function batch calls
 * Run one hot-path stage over host arrays (copied in and out); used for
 * per-function parity. */

/* GJK distance between the core shapes (no radii) of n pairs from an empty
 * simplex cache (b2Distance with useRadii = false,
 * Box2D/Collision/b2Distance.h:99; REF/Collision/b2Distance.cpp:447-606): out[2i] = distance, out[2i+1] = simplex vertex count. */
int b2g_distance_batch(b2g_context* ctx, const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                       const b2g_transform* xfB, int n, float* out);
/* Time of impact of n swept pairs over [0, tMax] (b2TimeOfImpact,
 * Box2D/Collision/b2TimeOfImpact.h:70; REF/Collision/b2TimeOfImpact.cpp:
 * 257-489): state[i] in b2TOIOutput::State order (0 unknown, 1 failed,
 * 2 overlapped, 3 touching, 4 separated), t[i] the time. */
int b2g_toi_batch(b2g_context* ctx, const b2g_shape* a, const b2g_sweep* sweepA, const b2g_shape* b,
                  const b2g_sweep* sweepB, int n, float tMax, int32_t* state, float* t);

/* Narrowphase for n shape pairs already in registry order
 * (b2CollidePolygons / b2CollideCircles / b2CollidePolygonAndCircle /
 * b2CollideEdgeAndCircle / b2CollideEdgeAndPolygon,
 * Box2D/Collision/b2Collision.h:237-260). Unsupported pairs give
 * pointCount 0. */
int b2g_collide_batch(b2g_context* ctx, const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                      const b2g_transform* xfB, int n, b2g_manifold* out);
/* b2WorldManifold::Initialize (REF/Collision/b2Collision.cpp:22) for n
 * manifolds; out: 8 floats each (normal, 2 points, 2 separations). */
int b2g_world_manifold_batch(b2g_context* ctx, const b2g_manifold* m, const b2g_transform* xfA,
                             const float* radiusA, const b2g_transform* xfB, const float* radiusB, int n,
                             float* out);
/* b2BroadPhase::UpdatePairs (Box2D/Collision/b2BroadPhase.h:193-247): the
 * pair set of proxies whose fat AABBs overlap a moved proxy's, as sorted
 * unique (min, max) proxy index pairs. *count may exceed cap. */
int b2g_pairs_batch(b2g_context* ctx, const b2g_aabb* fat, const uint8_t* moved, int n, int32_t* pairs, int cap,
                    int* count);
/* b2g_pairs_batch without allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same pair set,
 * computed in a fixed scratch array that each GPU holds for the life of the
 * process (about 72 MiB), with at most 65536 proxies and 1048576 pairs per
 * call (B2G_ERR_CAPACITY beyond). Inputs and outputs are the caller's host
 * buffers, as for b2g_pairs_batch. Calls that overlap on one GPU share its
 * scratch and race: serialise them. */
int b2g_pairs_batch_noalloc(b2g_context* ctx, const b2g_aabb* fat, const uint8_t* moved, int n, int32_t* pairs,
                            int cap, int* count);
/* b2DynamicTree::MoveProxy's fat-AABB rule (Box2D/Collision/b2DynamicTree.cpp:137-181):
 * per proxy, fatOut = fat unchanged and moved = 0 when fat contains tight,
 * else the re-fattened, displacement-extended box and moved = 1. */
int b2g_fat_rule_batch(b2g_context* ctx, const b2g_aabb* fat, const b2g_aabb* tight, const float* disp, int n,
                       b2g_aabb* fatOut, uint8_t* moved);
/* Rotation used inside the step for n angles (sine, cosine). */
int b2g_sincos_batch(b2g_context* ctx, const float* angles, int n, float* s, float* c);

/* Batched worlds: sets the sub-world of bodies [first, first + count) (ids
 * 0 .. 2^24 - 1; every body starts in sub-world 0). Fixtures of bodies in
 * different sub-worlds never form contacts, and the broad-phase keys its
 * cells by sub-world, so one context advances many independent worlds in one
 * step. Set before the bodies' first step. No reference counterpart (the
 * reference steps one b2World at a time, Box2D/Dynamics/b2World.h:96). */
int b2g_set_body_worlds(b2g_context* ctx, int first, int count, const int32_t* worlds);

/* ------------------------------------------------------------- schedule
 * How the coloured step runs, not what it computes: every choice gives the
 * same results bit for bit. No counterpart in the reference (its step is
 * single-threaded). */
typedef struct b2g_schedule {
    int32_t islands;       /* island labelling: 0 label-propagation rounds, 1 one round
                              then union-find, 2 union-find */
    int32_t collideOrder;  /* 0 contact-table order, 1 grouped by shape-pair class,
                              2 table order with polygon pairs by eight-lane tiles */
    int32_t solverBlock;   /* workgroup size of the per-colour solver kernels */
    int32_t collideBlock;  /* workgroup size of the narrowphase kernel */
    int32_t queryBlock;    /* workgroup size of the broad-phase query */
    int32_t tailColours;   /* trailing colours of at most this many constraints are
                              solved by one 1024-lane workgroup, colour after colour
                              (0: every colour its own launch) */
    int32_t clusterBodies; /* clustered solve: awake bodies are grouped into clusters of
                              this many (a power of two, 64-2048) by the Morton order of
                              their positions; constraints inside a cluster are solved by
                              one workgroup with the cluster's bodies in LDS, colour after
                              colour, then the constraints between clusters by colour.
                              0: off. Unlike the other fields this changes the order in
                              which a pass visits constraints (each still once per pass),
                              so results differ from the unclustered solve. */
    int32_t persistentSolve; /* one-launch solve: the warm start, velocity and position
                              passes of the step in one cooperative launch of this many
                              workgroups per compute unit (at most what fits), colours
                              separated by a grid-wide barrier instead of a launch each;
                              results are those of the per-colour launches bit for bit.
                              0: a launch per colour. Not used by the clustered solve or
                              a residual trace. On MI355X a grid barrier with the
                              cross-XCD cache write-back and invalidation it needs costs
                              more than a kernel boundary (the 100k pile steps in 4.9 ms
                              against 1.7 ms), so it is opt-in and not autotuned. */
} b2g_schedule;
/* Workgroup sizes are multiples of 64 from 64 to 1024; tailColours 0 to 8192;
   persistentSolve 0 to 16. */
int b2g_set_schedule(b2g_context* ctx, const b2g_schedule* s);
int b2g_get_schedule(b2g_context* ctx, b2g_schedule* s);
/* Autotuner: picks each schedule field in turn (islands, collideOrder,
 * solverBlock, collideBlock, queryBlock, tailColours; the clustered and the
 * one-launch solve, both measured slower on the piles, are not tried) by
 * timing the candidates on real steps of this world (the event time of the
 * phases the field affects, candidates interleaved and their order rotated
 * each of `rounds` rounds), keeps the fastest, and returns the result in
 * *out. The steps advance the world as ordinary steps do; since no choice
 * changes results, the trajectory is the one the world would have had
 * anyway. (3 + 3 + 4 + 3 + 3 + 4) * rounds steps. */
int b2g_autotune(b2g_context* ctx, const b2g_step_params* p, int rounds, b2g_schedule* out);

/* ------------------------------------------------------------- split step
 * b2g_step in two halves, so that the host can run b2ContactListener's
 * BeginContact, EndContact and PreSolve inside the step as
 * b2ContactManager::Collide does (Box2D/Dynamics/b2ContactManager.cpp:110-175,
 * Contacts/b2Contact.cpp:170-256). Either solver (the reference-order
 * narrowphase commits in list order before the split).
 *   b2g_step_collide: the narrowphase over the contact table. Until
 *     b2g_step_solve, b2g_get_contacts returns the updated table including
 *     the contacts Collide destroys, with meta[2] bits 1 touching,
 *     2 enabled, 4 destroyed this step, 8 updated this step, 16 touching
 *     before the update, and
 *     b2g_get_old_manifolds the manifolds from before the update (in the
 *     same order);
 *   b2g_override_contacts: PreSolve's edits per contact index: enabled for
 *     this step (b2Contact::SetEnabled), friction and restitution (kept);
 *   b2g_step_solve: the rest of the step. */
int b2g_step_collide(b2g_context* ctx, const b2g_step_params* p);
/* 1 for each contact of [0, cap) that the last step solved (an island
 * contact of an awake island: the contacts b2Island::Report passes to
 * PostSolve, Box2D/Dynamics/b2Island.cpp:517-540), else 0. */
int b2g_get_contact_solved(b2g_context* ctx, int32_t* out, int cap);
/* Friction and restitution of contacts [0, cap), 2 floats each, in the
 * order of b2g_get_contacts (b2Contact::GetFriction / GetRestitution). */
int b2g_get_contact_materials(b2g_context* ctx, float* out, int cap);
int b2g_get_old_manifolds(b2g_context* ctx, b2g_manifold* out, int cap, int* count);
/* Tangent speeds of contacts [0, cap) in the order of b2g_get_contacts
 * (b2Contact::GetTangentSpeed, Box2D/Dynamics/Contacts/b2Contact.h:142,
 * 351-354); 0 for a new contact. */
int b2g_get_contact_tangent_speeds(b2g_context* ctx, float* out, int cap);
/* b2Contact::SetTangentSpeed (b2Contact.h:139, 346-349) per contact index:
 * the speed (m/s along the contact tangent, a conveyor belt) holds until
 * changed or the contact ends. Between steps, or inside a split step before
 * b2g_step_solve (from PreSolve). Both solvers read it. */
int b2g_set_contact_tangent_speeds(b2g_context* ctx, const int32_t* ids, const float* speeds, int n);
int b2g_override_contacts(b2g_context* ctx, const int32_t* ids, const int32_t* enabled, const float* friction,
                          const float* restitution, int n);
int b2g_step_solve(b2g_context* ctx);

/* ------------------------------------------------------------- snapshot
 * The world state a step reads (bodies, fixtures with fat boxes and move
 * flags, the contact table with manifolds, warm-start impulses and colours,
 * joints with their impulses and order, world settings and the warm-start
 * ratio) as one byte buffer. A context restored from it steps exactly as
 * the original does (the reference has only a text dump, b2World::Dump,
 * Box2D/Dynamics/b2World.cpp:1271-1328). The network and logs are not
 * included. b2g_snapshot with out == NULL returns the size in *size;
 * b2g_restore needs a new context. */
int b2g_snapshot(b2g_context* ctx, void* out, size_t cap, size_t* size);
int b2g_restore(b2g_context* ctx, const void* in, size_t size);

/* Captured solves (default off): the plain coloured solve of a step (warm
 * start, velocity passes, store, position integration, position passes,
 * transform sync) runs as one hipGraph launch once its layout has repeated
 * (views, iterations, time step, non-empty colours and their size classes);
 * steps with the learned start, a trace, a log or profiling launch
 * directly. Results are the same either way. On a settling 100k-body pile
 * the layout changes often enough that capturing costs more than replaying
 * saves, so direct launches are the default. No reference counterpart. */
int b2g_set_graphs(b2g_context* ctx, int on);
/* Graph launches and captures since the context was created. */
int b2g_get_graph_counts(b2g_context* ctx, int64_t* launches, int64_t* captures);

/* Sizes in bytes of the ABI structs, in this order: b2g_shape,
 * b2g_manifold, b2g_body_state, b2g_fixture_def, b2g_world_def,
 * b2g_step_params, b2g_step_stats, b2g_phase_profile, b2g_mlp_params,
 * b2g_aabb, b2g_transform, b2g_revolute_def, b2g_schedule, b2g_sweep,
 * b2g_distance_def, b2g_rope_def, b2g_weld_def, b2g_friction_def, b2g_motor_def,
 * b2g_pulley_def, b2g_mouse_def, b2g_prismatic_def, b2g_wheel_def, b2g_gear_def.
 * Returns the count (out may be null). Needs no device; lets bindings check
 * their struct layouts. */
int b2g_abi_sizes(int32_t* out, int cap);

/* Checks the colouring of the last coloured step: *conflicts counts dynamic
 * bodies met twice within one colour (must be 0: a colour is solved in
 * parallel), *maskErrors bodies whose colour mask differs from the colours
 * of their constraints. No reference counterpart (the reference solves
 * sequentially). */
int b2g_check_colouring(b2g_context* ctx, int* conflicts, int* maskErrors);

/* Colour demotion rounds per coloured step (0-64, default 1). A constraint
 * keeps its colour for its life; each round moves constraints down to the
 * lowest colour that has since become free at both of their dynamic bodies
 * (each body admitting one mover per round, so a colour still never holds
 * two constraints on one dynamic body). Fewer, larger colours mean fewer
 * launches per pass. This changes the order in which a pass visits
 * constraints, so results differ with the number of rounds. */
int b2g_set_colour_demotion(b2g_context* ctx, int rounds);

/* Constraints per colour slot of the last coloured step (slots 0-63, then
 * the overflow slot 64): out[c] for c < min(cap, 65); *count = 65. No
 * reference counterpart (the reference solves sequentially); it tells how
 * many launches a pass takes. */
int b2g_get_colour_counts(b2g_context* ctx, int32_t* out, int cap, int* count);

/* The device scan and sorts the step is built on, run once on host buffers
 * (device buffers allocated per call). No reference counterpart: the
 * reference sorts its pair buffer on the host with std::sort
 * (Box2D/Collision/b2BroadPhase.h:220) and needs no scan; these let tests
 * check the primitives themselves against a host sort.
 *  - b2g_prim_exclusive_sum: out[i] = in[0] + ... + in[i-1].
 *  - b2g_prim_sort_pairs: stable sort of (key, value) pairs; method 0 by the
 *    low `bits` bits of key (1-32, higher bits ignored), method 1 by the low
 *    7 bits (one pass) with totals[k] += count of key k for k < totalsKeys
 *    (0-128; totals may be null when 0).
 *  - b2g_prim_sort_keys_u64: ascending sort of 64-bit keys whose bits at and
 *    above endBit (1-64) are all equal. */
int b2g_prim_exclusive_sum(b2g_context* ctx, const int32_t* in, int n, int32_t* out);
int b2g_prim_sort_pairs(b2g_context* ctx, const uint32_t* keys, const int32_t* vals, int n, int bits, int method,
                        uint32_t* outKeys, int32_t* outVals, int32_t* totals, int totalsKeys);
int b2g_prim_sort_keys_u64(b2g_context* ctx, const uint64_t* keys, int n, int endBit, uint64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
