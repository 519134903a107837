// libb2g: the MI355X (gfx950) Box2D step behind the C ABI of include/b2g.h.
//
// One context holds one world in HBM as structure-of-arrays. A step is a
// sequence of kernels on the context's stream:
//
//   [new fixtures] hash + broad-phase pairs -> new contacts
//   collide       per contact: fat-AABB overlap test (destroy), narrowphase,
//                 feature-id warm-start matching, touching flag, wakes
//   mark          constraint set of this step; release colours of contacts
//                 that stopped being constraints
//   compact       stable removal of destroyed contacts; hash of live pairs
//   colour        greedy edge colouring of new constraints (persistent
//                 colours per contact, one winner per body per round)
//   order         constraint list sorted by colour
//   integrate v   gravity, forces, damping
//   init          velocity/position constraints (world manifold, masses,
//                 restitution bias, 2-point block matrix), warm-start scale
//   warm start, velocity iterations, store impulses   (per colour)
//   integrate x   translation/rotation clamps
//   position iterations (per colour, early exit on min separation)
//   sync          transforms from sweeps
//   fat AABBs     swept AABB and the fat-box reinsert rule -> move flags
//   broad-phase   uniform hash grid over fat AABBs, moved-proxy queries,
//                 existing-contact filter, new contacts in pair-key order
//
// Contacts are created only from moved proxies and destroyed only when their
// fat AABBs stop overlapping, exactly as b2ContactManager does, so the
// contact set equals the reference's at every step.
#include "b2g_device.h"
#include "collide.h"
#include "prims.h"
#include "mlp.h"
#include "toi.h"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <math.h>
#include <string.h>
#include <stdio.h>
#include <string>
#include <vector>
#include <limits>
#include <random>
#include <algorithm>
#include <chrono>
#include <mutex>

using namespace b2g;

// ====================================================================
// Device counters shared by the step kernels.
// ====================================================================
namespace {

constexpr int kMaxColours = 64;       // colours with a body-mask bit
constexpr int kOverflowColour = 64;   // solved serially when masks are full
constexpr int kColourSlots = 65;
constexpr int kMaxPosIters = 32;
constexpr int kMaxCellsPerProxy = 16;
constexpr int kColourRoundsUnchecked = 2;  // fewest colouring rounds launched before the first readback
constexpr int kColourRoundSlots = 16;      // rounds whose leftovers are recorded (later rounds share the last)
constexpr uint32_t kNotConstraintKey = 0x7Fu;  // order key of a contact outside the constraint set

struct Counters {
    int nContacts;
    int nNewPairs;
    int nConstraints;
    int remaining;
    int overflow;
    int nBig;
    int nQueries;          // moved proxies binned in the grid (their queries)
    int nTiles;            // polygon-polygon contacts deferred to k_collide_tiles
    int nScanned;          // grid entries and wide proxies the queries examined
    int roundLeft[kColourRoundSlots];  // constraints left uncoloured by colouring round r of the step
    int beginEvents;
    int endEvents;
    int points;
    int nEdges;
    int nIslands;
    int islandChanged;     // set by a propagation round that lowered a label
    int nFresh;            // fresh contact points of the step (learned warm start, log, trace)
    int residualPoints;
    float residualSum;     // sum of squared complementarity residuals (k_residual)
    int colourCount[kColourSlots + 3];
    int minSep[kMaxPosIters];
    int nClustered;        // bodies placed in clusters (clustered solve)
    int boundaryStart;     // first constraint between clusters (clustered solve)
    int maxColour;         // highest constraint colour of the step
    unsigned barrierCount; // one-launch solve: workgroups arrived at the grid barrier
    unsigned barrierGen;   // one-launch solve: barrier generation
    int barrierAbort;      // one-launch solve: a barrier wait passed its bound
};

// Constraint arrays (b2ContactVelocityConstraint / b2ContactPositionConstraint,
// REF/Dynamics/Contacts/b2ContactSolver.h:33-84, b2ContactSolver.cpp:35-50),
// one slot per constraint in colour order.
struct ConView {
    int* contact;     // contact table slot
    int2* bodies;     // body a, body b
    float4* mass;     // invMassA invIA invMassB invIB
    float4* normal;   // normal.x normal.y friction tangentSpeed
    float4* r0;       // point 0: rA.x rA.y rB.x rB.y
    float4* r1;       // point 1
    float4* pmass;    // normalMass0 tangentMass0 normalMass1 tangentMass1
    float2* bias;     // velocityBias0 velocityBias1
    float4* imp;      // normalImpulse0 tangentImpulse0 normalImpulse1 tangentImpulse1
    float4* K;        // block matrix ex.x ex.y ey.x ey.y
    float4* NM;       // its inverse
    int* info;        // bits 0-1 velocity point count, 2-3 position point count,
                      // 4 body A dynamic, 5 body B dynamic, 8-9 manifold type
    float4* lp;       // localPoints[0], localPoints[1]
    float4* lnp;      // localNormal, localPoint
    float4* lc;       // localCenterA, localCenterB
    float2* rad;      // radiusA radiusB
    int* island;      // island root of the constraint's non-static bodies
    int2* slot;       // clustered solve: LDS slot of body A, B in the cluster (-1: not staged)
};

// Revolute joints (b2RevoluteJoint, REF/Dynamics/Joints/b2RevoluteJoint.h),
// indexed by creation order; `order` lists them by colour.
enum : int {
    JF_LIMIT = 1, JF_MOTOR = 2, JF_COLLIDE = 4, JF_REMOVED = 8,
    JL_INACTIVE = 0, JL_LOWER = 1, JL_UPPER = 2, JL_EQUAL = 3,  // limit state in bits 8-9
    JT_REVOLUTE = 0, JT_DISTANCE = 1, JT_ROPE = 2, JT_WELD = 3,  // joint type in bits 12-15
    JT_FRICTION = 4, JT_MOTOR = 5, JT_PULLEY = 6, JT_MOUSE = 7, JT_PRISMATIC = 8, JT_WHEEL = 9, JT_GEAR = 10,
    JF_GEAR_PRISM_A = 1 << 16, JF_GEAR_PRISM_B = 1 << 17,  // a gear's joints: prismatic when set
};
__host__ __device__ inline int jointType(int flags) { return (flags >> 12) & 15; }
struct JointView {
    int n;
    const int* order;  // joint per slot, colour by colour
    int2* bodies;
    float4* anchors;   // localAnchorA, localAnchorB
    float4* limits;    // referenceAngle, lowerAngle, upperAngle, maxMotorTorque
    float* motorSpeed;
    int* flags;        // JF_* | limit state << 8
    float4* impulse;   // impulse x y z, motor impulse
    float4* r;         // rA, rB of the step
    float4* masses;    // invMassA invIA invMassB invIB
    float4* K0;        // mass matrix m11 m12 m13 m22
    float2* K1;        // m23 m33
    float* motorMass;
    float4* axis;      // localXAxisA (prismatic, wheel)
    float4* extra;     // two rows per joint: step directions kept across steps
    float h;           // time step of the solve (soft joints)
};
// Per-type use of the arrays. Revolute: limits = referenceAngle, lower,
// upper, maxMotorTorque; impulse = (x, y, limit, motor). Distance: limits =
// length, frequencyHz, dampingRatio; impulse.x; K0 = (u.x, u.y, mass, bias),
// K1.x = gamma. Rope: limits.x = maxLength; impulse.x; K0 = (u.x, u.y, mass,
// length of the step); impulse = (impulse, u.x, u.y, limit state), the state
// also in flags bits 8-9. Weld: limits =
// referenceAngle, frequencyHz, dampingRatio; impulse x y z; K0, K1 = the
// effective mass matrix; motorMass = gamma, motorSpeed = bias. Friction:
// limits = maxForce, maxTorque; impulse = linear x y, angular; K0 = linear
// mass m11 m12 m22, angular mass. Motor: anchors = linearOffset x y,
// angularOffset, correctionFactor; limits, impulse and K0 as friction; K1 =
// linear error, motorMass = angular error. Pulley: limits = ground anchors A,
// B; motorSpeed = ratio; impulse = (impulse, uB.x, uB.y, the constant
// lengthA + ratio lengthB); K0 = (uA, uB), K1.x = mass. Mouse:
// anchors = target, localAnchorB; limits = maxForce, frequencyHz,
// dampingRatio, mass of body B; impulse x y; K0 = mass m11 m12 m22, gamma;
// K1 = beta C. Prismatic: limits = referenceAngle, lowerTranslation,
// upperTranslation, maxMotorForce; axis = localXAxisA; impulse x y z, motor;
// extra[2j] = (axis, perp); r = (s1, s2, a1, a2); K0, K1 = K; limit state in
// flags bits 8-9. Wheel: limits = frequencyHz, dampingRatio, maxMotorTorque;
// axis = localXAxisA; impulse = (impulse, spring, motor); extra[2j] = (ax,
// sAx, sBx), extra[2j+1] = (ay, sAy, sBy); K0 = (mass, springMass, bias,
// gamma).

__device__ inline int encodeOrdered(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__host__ __device__ inline float decodeOrdered(int i) {
    const int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
    union { int i; float f; } u;
    u.i = b;
    return u.f;
}

__host__ __device__ inline void zeroManifold(b2g_manifold& m) {
    for (int i = 0; i < 2; ++i) {
        m.points[i].localPoint.x = m.points[i].localPoint.y = 0.0f;
        m.points[i].normalImpulse = m.points[i].tangentImpulse = 0.0f;
        m.points[i].id = 0;
    }
    m.localNormal.x = m.localNormal.y = 0.0f;
    m.localPoint.x = m.localPoint.y = 0.0f;
    m.type = 0;
    m.pointCount = 0;
}

__device__ inline unsigned long long pairKey(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((unsigned long long)(unsigned)lo << 32) | (unsigned)hi;
}

// A proxy's broad-phase order key and back (WorldView::ford / ordf).
__device__ inline int orderOf(const WorldView& w, int f) { return w.ford ? w.ford[f] : f; }
__device__ inline int fixtureOf(const WorldView& w, int k) { return w.ordf ? w.ordf[k] : k; }

__device__ inline uint32_t hashU64(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return (uint32_t)k;
}

// Open-addressing hash set of 64-bit keys (~0 marks an empty slot).
__device__ inline bool hashContains(const unsigned long long* table, uint32_t mask, unsigned long long key) {
    uint32_t h = hashU64(key) & mask;
    for (;;) {
        const unsigned long long v = table[h];
        if (v == key) return true;
        if (v == ~0ull) return false;
        h = (h + 1) & mask;
    }
}

// b2Body::ShouldCollide's joint test (REF/Dynamics/b2Body.cpp:404-414): a
// joint with collideConnected off keeps its two bodies from colliding.
__device__ inline bool jointKeepsApart(const WorldView& w, int bA, int bB) {
    return w.jointPairs && hashContains(w.jointPairs, w.jointPairMask, pairKey(bA, bB));
}

// Sets the awake bit; a body that was asleep restarts its sleep timer
// (b2Body::SetAwake(true), REF/Dynamics/b2Body.h:635-644).
__device__ inline void wakeBody(const WorldView& w, int b) {
    const int old = atomicOr(&w.bflags[b], BF_AWAKE);
    if ((old & BF_AWAKE) == 0) w.gsleep[b].y = 0.0f;
}

__device__ inline bool sensorOf(const WorldView& w, int f) { return w.fmat[f].w != 0.0f; }

// A removed proxy's fat box is inverted (k_remove_fixtures); a live one
// never is.
__device__ inline bool removedFixture(const WorldView& w, int f) { return w.fat[f].x > w.fat[f].z; }

// b2ContactFilter::ShouldCollide, REF/Dynamics/b2WorldCallbacks.cpp:26-38.
__device__ inline bool filterPass(const WorldView& w, int fa, int fb) {
    const int ga = w.fgroup[fa], gb = w.fgroup[fb];
    if (ga == gb && ga != 0) return ga > 0;
    const uint32_t a = w.ffilter[fa], b = w.ffilter[fb];
    const uint32_t catA = a & 0xFFFF, maskA = a >> 16, catB = b & 0xFFFF, maskB = b >> 16;
    return (maskA & catB) != 0 && (catA & maskB) != 0;
}

// ====================================================================
// Collide (b2ContactManager::Collide + b2Contact::Update,
// REF/Dynamics/b2ContactManager.cpp:107-172, REF/Dynamics/Contacts/b2Contact.cpp:171-251)
// ====================================================================
// Shape-pair class of a contact (registry order puts fixture A's type
// first: circle-circle, polygon-circle, polygon-polygon, edge-circle,
// edge-polygon); below 16.
__device__ inline int collideClass(const WorldView& w, int c) {
    return (w.shape[w.cfA[c]].type << 2) | w.shape[w.cfB[c]].type;
}

// Collide order: contacts grouped by shape-pair class (a stable radix sort
// of these keys), so that a wavefront runs one narrowphase routine. The
// order does not change results: contacts are collided independently.
__global__ void k_collide_keys(WorldView w, uint32_t* keys, int* vals) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    keys[c] = (uint32_t)(collideClass(w, c) & 15);
    vals[c] = c;
}

// PreSolve's edits (b2Contact::SetEnabled, SetFriction, SetRestitution,
// REF/Dynamics/Contacts/b2Contact.h:98-131): the enabled flag holds for this
// step (Update re-enables every step), friction and restitution stay.
__global__ void k_override_contacts(WorldView w, const int* ids, const int* enabled, const float2* mat, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = ids[i];
    if (c < 0 || c >= w.nContacts) return;
    w.cflags[c] = enabled[i] ? (w.cflags[c] | CF_ENABLED) : (w.cflags[c] & ~CF_ENABLED);
    w.cmat[c] = mat[i];
}

// b2Contact::SetTangentSpeed (REF/Dynamics/Contacts/b2Contact.h:346-349): the
// speed stays with the contact until changed or the contact ends.
__global__ void k_set_tangent_speeds(WorldView w, const int* ids, const float* speeds, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = ids[i];
    if (c >= 0 && c < w.nContacts) w.ctan[c] = speeds[i];
}

// Slot of this lane in a list that every lane with `want` appends to: one
// atomic per wavefront, lanes in lane order.
__device__ inline int waveAppend(int* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0ull) return -1;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    return want ? base + __popcll(mask & ((1ull << lane) - 1ull)) : -1;
}

// b2Contact::Update after the manifold is known: impulses carried over by
// feature id, touching flag, wakes and begin/end counts.
__device__ inline void collideFinish(const WorldView& w, Counters* cnt, int c, int flags, b2g_manifold& m,
                                     const b2g_manifold& old, int bA, int bB) {
    const bool wasTouching = (flags & CF_TOUCHING) != 0;
    const bool touching = m.pointCount > 0;
    for (int i = 0; i < 2; ++i) {
        m.points[i].normalImpulse = 0.0f;
        m.points[i].tangentImpulse = 0.0f;
        if (i >= m.pointCount) continue;
        for (int j = 0; j < old.pointCount; ++j) {
            if (old.points[j].id == m.points[i].id) {
                m.points[i].normalImpulse = old.points[j].normalImpulse;
                m.points[i].tangentImpulse = old.points[j].tangentImpulse;
                break;
            }
        }
    }
    if (touching != wasTouching) {
        wakeBody(w, bA);
        wakeBody(w, bB);
    }
    flags |= CF_ENABLED | CF_UPDATED;
    if (touching) flags |= CF_TOUCHING;
    else flags &= ~CF_TOUCHING;
    if (wasTouching) flags |= CF_WAS_TOUCHING;
    else flags &= ~CF_WAS_TOUCHING;
    if (!wasTouching && touching) atomicAdd(&cnt->beginEvents, 1);
    if (wasTouching && !touching) atomicAdd(&cnt->endEvents, 1);
    w.cman[c] = m;
    w.cflags[c] = flags;
}

// b2TestOverlap (REF/Collision/b2Collision.cpp:237-255): GJK distance from an
// empty cache, the radii taken off as b2Distance's useRadii does
// (REF/Collision/b2Distance.cpp:579-602), under 10 epsilon.
__device__ inline bool testOverlap(const b2g_shape& a, const Xf& xfA, const b2g_shape& b, const Xf& xfB) {
    const Proxy pA = proxyOf(a), pB = proxyOf(b);
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    float d = gjkDistance(cache, pA, xfA, pB, xfB);
    if (d > pA.radius + pB.radius && d > FLT_EPSILON) d -= pA.radius + pB.radius;
    else d = 0.0f;
    return d < 10.0f * FLT_EPSILON;
}

// b2Contact::Update's sensor branch (REF/Dynamics/Contacts/b2Contact.cpp:189-197):
// touching is the shapes' overlap, the manifold keeps no points, and no
// body is woken.
__device__ inline void sensorFinish(const WorldView& w, Counters* cnt, int c, int flags, bool touching) {
    const bool wasTouching = (flags & CF_TOUCHING) != 0;
    flags |= CF_ENABLED | CF_UPDATED;
    if (touching) flags |= CF_TOUCHING;
    else flags &= ~CF_TOUCHING;
    if (wasTouching) flags |= CF_WAS_TOUCHING;
    else flags &= ~CF_WAS_TOUCHING;
    if (!wasTouching && touching) atomicAdd(&cnt->beginEvents, 1);
    if (wasTouching && !touching) atomicAdd(&cnt->endEvents, 1);
    w.cman[c].pointCount = 0;
    w.cflags[c] = flags;
}

// One contact of b2ContactManager::Collide: filtering, the activity test,
// the fat-box test, then the narrowphase. With `deferPolygons` a
// polygon-polygon contact stops after the tests (flags written) and returns
// true: k_collide_tiles takes its narrowphase.
__device__ inline bool collideOne(const WorldView& w, Counters* cnt, int c, bool deferPolygons) {
    int flags = w.cflags[c] & ~CF_UPDATED;
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    if (removedFixture(w, fA) || removedFixture(w, fB) || (flags & CF_DESTROY)) {
        // destroyed with its body or fixture, or by a body type change (its
        // wake-ups done when that happened)
        if (flags & CF_TOUCHING) atomicAdd(&cnt->endEvents, 1);
        w.cflags[c] = flags | CF_DEAD;
        return false;
    }
    if (flags & CF_FILTER) {
        const bool dynAny = w.btype[bA] == BT_DYNAMIC || w.btype[bB] == BT_DYNAMIC;
        if (!dynAny || jointKeepsApart(w, bA, bB) || !filterPass(w, fA, fB)) {
            if (flags & CF_TOUCHING) atomicAdd(&cnt->endEvents, 1);
            if (w.cman[c].pointCount > 0 && !sensorOf(w, fA) && !sensorOf(w, fB)) {
                wakeBody(w, bA);
                wakeBody(w, bB);
            }
            w.cflags[c] = flags | CF_DEAD;
            return false;
        }
        flags &= ~CF_FILTER;
    }
    const bool activeA = (w.bflags[bA] & BF_AWAKE) && w.btype[bA] != BT_STATIC;
    const bool activeB = (w.bflags[bB] & BF_AWAKE) && w.btype[bB] != BT_STATIC;
    if (!activeA && !activeB) {
        w.cflags[c] = flags;
        return false;
    }
    if (!overlaps(w.fat[fA], w.fat[fB])) {
        if (flags & CF_TOUCHING) atomicAdd(&cnt->endEvents, 1);
        if (w.cman[c].pointCount > 0 && !sensorOf(w, fA) && !sensorOf(w, fB)) {
            wakeBody(w, bA);
            wakeBody(w, bB);
        }
        w.cflags[c] = flags | CF_DEAD;
        return false;
    }
    if (sensorOf(w, fA) || sensorOf(w, fB)) {
        sensorFinish(w, cnt, c, flags, testOverlap(w.shape[fA], xfOf(w.xf[bA]), w.shape[fB], xfOf(w.xf[bB])));
        return false;
    }
    if (deferPolygons && w.shape[fA].type == B2G_POLYGON && w.shape[fB].type == B2G_POLYGON) {
        w.cflags[c] = flags;
        return true;
    }
    const b2g_manifold old = w.cman[c];
    b2g_manifold m;
    zeroManifold(m);
    const Xf xfA = xfOf(w.xf[bA]);
    const Xf xfB = xfOf(w.xf[bB]);
    collideShapes(m, w.shape[fA], xfA, w.shape[fB], xfB);
    collideFinish(w, cnt, c, flags, m, old, bA, bB);
    return false;
}

// tiles: null computes every contact here; else polygon-polygon contacts go
// to that list (count in cnt->nTiles) for k_collide_tiles.
__global__ void k_collide(WorldView w, Counters* cnt, const int* order, int* tiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    bool deferred = false;
    if (i < w.nContacts) {
        c = order ? order[i] : i;
        deferred = collideOne(w, cnt, c, tiles != nullptr);
    }
    if (tiles) {
        const int slot = waveAppend(&cnt->nTiles, deferred);
        if (deferred) tiles[slot] = c;
    }
}

// The shapes of one tile in LDS: kTileLanes lanes stage the two b2g_shape
// records of their contact with 16-byte loads.
constexpr int kShapeWords = (int)(sizeof(b2g_shape) / sizeof(float4));
static_assert(sizeof(b2g_shape) % sizeof(float4) == 0, "b2g_shape is a whole number of float4");

__device__ inline void stageShapes(b2g_shape* dst, const b2g_shape& a, const b2g_shape& b, int lane) {
    float4* d = (float4*)dst;
    const float4* sa = (const float4*)&a;
    const float4* sb = (const float4*)&b;
    for (int k = lane; k < 2 * kShapeWords; k += kTileLanes) d[k] = k < kShapeWords ? sa[k] : sb[k - kShapeWords];
}

// Narrowphase of the deferred polygon-polygon contacts: a group of
// kTileLanes lanes per contact (collidePolygonsTile), the two shapes staged
// in LDS; the workgroups walk the list in steps of their tile count.
constexpr int kTileBlock = 64;
__global__ void __launch_bounds__(kTileBlock) k_collide_tiles(WorldView w, Counters* cnt, const int* tiles) {
    __shared__ b2g_shape sh[kTileBlock / kTileLanes][2];
    const int grp = threadIdx.x / kTileLanes, lane = threadIdx.x % kTileLanes;
    const int perBlock = kTileBlock / kTileLanes;
    const int n = cnt->nTiles;
    for (int base = blockIdx.x * perBlock; base < n; base += gridDim.x * perBlock) {
        const int t = base + grp;
        int c = -1, fA = 0, fB = 0;
        if (t < n) {
            c = tiles[t];
            fA = w.cfA[c];
            fB = w.cfB[c];
            stageShapes(sh[grp], w.shape[fA], w.shape[fB], lane);
        }
        __syncthreads();
        if (c >= 0) {
            const int bA = w.fbody[fA], bB = w.fbody[fB];
            const Xf xfA = xfOf(w.xf[bA]);
            const Xf xfB = xfOf(w.xf[bB]);
            b2g_manifold m;
            zeroManifold(m);
            collidePolygonsTile(m, sh[grp][0], xfA, sh[grp][1], xfB, lane);
            if (lane == 0) {
                const b2g_manifold old = w.cman[c];
                collideFinish(w, cnt, c, w.cflags[c], m, old, bA, bB);
            }
        }
        __syncthreads();
    }
}
// Constraint set of the step: touching, enabled, non-sensor contacts with an
// awake non-static body (the contacts b2World::Solve's island search adds,
// REF/Dynamics/b2World.cpp:451-475). Contacts leaving the set give their
// colour back to both bodies' masks.
__device__ inline void releaseColour(const WorldView& w, int c, int bA, int bB) {
    const int col = w.ccolour[c];
    if (col < 0) return;
    if (col < kMaxColours) {
        const unsigned long long bit = 1ull << col;
        if (w.btype[bA] == BT_DYNAMIC) atomicAnd(&w.colourMask[bA], ~bit);
        if (w.btype[bB] == BT_DYNAMIC) atomicAnd(&w.colourMask[bB], ~bit);
    }
    w.ccolour[c] = -1;
}

// Destroyed contacts give their colour back and leave the table.
__global__ void k_mark_dead(WorldView w, int* keep) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    const int flags = w.cflags[c];
    if (flags & CF_DEAD) releaseColour(w, c, w.fbody[w.cfA[c]], w.fbody[w.cfB[c]]);
    keep[c] = (flags & CF_DEAD) ? 0 : 1;
}

// A contact joins an island when it is touching, enabled and not a sensor
// (the contact filter of b2World::Solve's search, REF/Dynamics/b2World.cpp:451-469).
__device__ inline bool islandContact(const WorldView& w, int c) {
    const int f = w.cflags[c];
    return (f & CF_TOUCHING) && (f & CF_ENABLED) && !(f & CF_DEAD) && !sensorOf(w, w.cfA[c]) &&
           !sensorOf(w, w.cfB[c]);
}

// ---------------------------------------------------------------- islands
// b2World::Solve grows islands by depth-first search from every awake,
// active, non-static body through island contacts, without crossing static
// bodies, and wakes every body it reaches (REF/Dynamics/b2World.cpp:408-519).
// Here the same components come from one round of minimum-label
// propagation (each island contact lowers the larger label to the smaller
// with atomicMin, then pointer jumping), which merges most of each island
// cheaply, and a lock-free union-find pass over the island contacts and
// joints that finishes every island, then one more pointer-jumping pass:
//  * every body starts as its own root, and a parent index is never larger
//    than the body's own, so parent chains end;
//  * a union finds both roots and hooks the larger root under the smaller
//    with a compare-and-swap that succeeds only while it is still a root;
//    a failed swap returns the root's new parent, from which the search
//    goes on. Each retry lowers the sum of the two indices, so it ends;
//  * plain loads of parents may be stale across XCDs, but every value a
//    parent ever held is an ancestor, so a stale read only shortens the
//    walk, and the swap itself reads the current value. Path halving stores
//    such an ancestor, so it keeps every invariant above.
// Roots only ever move under smaller indices, so each component ends up
// labelled with its smallest body index, for any thread schedule.
__global__ void k_island_init(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    w.islandOf[b] = b;
    w.islandAwake[b] = 0;
    w.islandSleep[b] = 0;
}

__device__ inline int islandFind(int* L, int x) {
    for (;;) {
        const int p = L[x];
        if (p == x) return x;
        const int g = L[p];
        if (g != p) L[x] = g;  // path halving: g is an ancestor of x
        x = p;
    }
}

__device__ inline void islandUnion(int* L, int a, int b) {
    for (;;) {
        a = islandFind(L, a);
        b = islandFind(L, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicCAS(&L[a], a, b);
        if (old == a) return;
        a = old;
    }
}

// A propagation round: each island contact lowers the larger of its
// bodies' labels to the smaller (atomicMin only).
__global__ void k_island_hook(WorldView w, int* changed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts || !islandContact(w, c)) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    if (w.btype[bA] == BT_STATIC || w.btype[bB] == BT_STATIC) return;
    if (!(w.bflags[bA] & BF_ACTIVE) || !(w.bflags[bB] & BF_ACTIVE)) return;
    const int la = w.islandOf[bA], lb = w.islandOf[bB];
    if (la == lb) return;
    atomicMin(&w.islandOf[la > lb ? la : lb], la < lb ? la : lb);
    if (changed) *changed = 1;
}

__global__ void k_island_union(WorldView w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts || !islandContact(w, c)) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    if (w.btype[bA] == BT_STATIC || w.btype[bB] == BT_STATIC) return;
    if (!(w.bflags[bA] & BF_ACTIVE) || !(w.bflags[bB] & BF_ACTIVE)) return;
    islandUnion(w.islandOf, bA, bB);
}

// Chains only lead to smaller indices, so a chain ends within nBodies hops.
__global__ void k_island_jump(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    int l = w.islandOf[b];
    for (;;) {
        const int n = w.islandOf[l];
        if (n == l) break;
        l = n;
    }
    if (l != w.islandOf[b]) w.islandOf[b] = l;
}

__global__ void k_island_roots(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    if (w.btype[b] != BT_STATIC && (w.bflags[b] & BF_ACTIVE) && (w.bflags[b] & BF_AWAKE))
        w.islandAwake[w.islandOf[b]] = 1;
}

// Bodies of awake islands are woken and take part in this step's solve.
__global__ void k_island_mark(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    int flags = w.bflags[b] & ~BF_ISLAND;
    if (w.btype[b] != BT_STATIC && (flags & BF_ACTIVE) && w.islandAwake[w.islandOf[b]]) {
        if ((flags & BF_AWAKE) == 0) w.gsleep[b].y = 0.0f;
        flags |= BF_AWAKE | BF_ISLAND;
    }
    w.bflags[b] = flags;
}

// Constraint set of the step: island contacts of awake islands. Contacts
// leaving the set give their colour back.
__global__ void k_mark_constraints(WorldView w, Counters* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    int flags = w.cflags[c] & ~CF_CONSTRAINT;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool isCons = islandContact(w, c) && ((w.bflags[bA] & BF_ISLAND) || (w.bflags[bB] & BF_ISLAND));
    if (!isCons) releaseColour(w, c, bA, bB);
    if (isCons) {
        flags |= CF_CONSTRAINT;
        if (w.ccolour[c] < 0) atomicAdd(&cnt->remaining, 1);
    }
    w.cflags[c] = flags;
}

struct ContactArrays {
    int* cfA;
    int* cfB;
    int* cflags;
    float2* cmat;
    float* ctan;
    b2g_manifold* cman;
    int* ccolour;
    float* ctoi;
    int* ctoiCount;
};

__global__ void k_compact(ContactArrays src, ContactArrays dst, const int* keep, const int* idx, int n,
                          Counters* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (c == n - 1) cnt->nContacts = idx[c] + keep[c];
    if (!keep[c]) return;
    const int d = idx[c];
    dst.cfA[d] = src.cfA[c];
    dst.cfB[d] = src.cfB[c];
    dst.cflags[d] = src.cflags[c];
    dst.cmat[d] = src.cmat[c];
    dst.ctan[d] = src.ctan[c];
    dst.cman[d] = src.cman[c];
    dst.ccolour[d] = src.ccolour[c];
    dst.ctoi[d] = src.ctoi[c];
    dst.ctoiCount[d] = src.ctoiCount[c];
}

template <typename T>
__global__ void k_fill(T* p, int n, T v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Open-addressing set of live contact pair keys (the existence test of
// b2ContactManager::AddPair, REF/Dynamics/b2ContactManager.cpp:196-221).
// live: the contact count when the host holds only a bound (w.nContacts).
__global__ void k_hash_insert(WorldView w, unsigned long long* table, uint32_t mask, const Counters* live) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts || (live && c >= live->nContacts)) return;
    const unsigned long long key = pairKey(w.cfA[c], w.cfB[c]);
    uint32_t h = hashU64(key) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[h], ~0ull, key);
        if (prev == ~0ull || prev == key) return;
        h = (h + 1) & mask;
    }
}

// ====================================================================
// Colouring: a new constraint takes the lowest colour free at both of its
// dynamic bodies. Per round, each body admits at most one new constraint
// (the highest priority one claiming it), so no two constraints of a colour
// share a dynamic body. Priorities are a hash of the contact's pair key, so
// the result does not depend on thread scheduling.
// ====================================================================
__device__ inline unsigned long long claimValue(uint32_t tag, const WorldView& w, int c) {
    const unsigned long long key = pairKey(w.cfA[c], w.cfB[c]);
    const unsigned long long prio = ((unsigned long long)(hashU64(key) & 0xFFFFFFu) << 24) | (unsigned)(c & 0xFFFFFF);
    return ((unsigned long long)tag << 48) | (0xFFFFFFFFFFFFull - prio);
}

// Also clears the count of losers that k_colour_commit then takes.
__global__ void k_colour_claim(WorldView w, unsigned long long* claim, uint32_t tag, Counters* cnt, int round) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) cnt->roundLeft[round] = 0;
    if (c >= w.nContacts) return;
    if (!(w.cflags[c] & CF_CONSTRAINT) || w.ccolour[c] >= 0) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const unsigned long long v = claimValue(tag, w, c);
    if (w.btype[bA] == BT_DYNAMIC) atomicMax(&claim[bA], v);
    if (w.btype[bB] == BT_DYNAMIC) atomicMax(&claim[bB], v);
}

__global__ void k_colour_commit(WorldView w, const unsigned long long* claim, uint32_t tag, Counters* cnt, int round) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (!(w.cflags[c] & CF_CONSTRAINT) || w.ccolour[c] >= 0) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool dA = w.btype[bA] == BT_DYNAMIC, dB = w.btype[bB] == BT_DYNAMIC;
    const unsigned long long v = claimValue(tag, w, c);
    const bool won = (!dA || claim[bA] == v) && (!dB || claim[bB] == v);
    if (!won) {
        atomicAdd(&cnt->roundLeft[round], 1);
        return;
    }
    const unsigned long long used = (dA ? w.colourMask[bA] : 0ull) | (dB ? w.colourMask[bB] : 0ull);
    const unsigned long long freeBits = ~used;
    if (freeBits == 0ull) {
        w.ccolour[c] = kOverflowColour;
        return;
    }
    const int col = __ffsll((long long)freeBits) - 1;
    const unsigned long long bit = 1ull << col;
    if (dA) atomicOr(&w.colourMask[bA], bit);
    if (dB) atomicOr(&w.colourMask[bB], bit);
    w.ccolour[c] = col;
}

// Demotion: a constraint keeps its colour for its life, so when the
// contacts around it end, lower colours free up at both of its bodies while
// it stays high. Per round, a constraint above the lowest colour free at
// both of its dynamic bodies (any colour, from the overflow slot) claims
// them, each body admits its highest-priority claimant as in the colouring,
// and the winners move down. A winner is the only constraint changing colour
// at its bodies that round, so no colour ever holds two constraints on one
// dynamic body. The pass visits fewer, larger colours.
__device__ inline unsigned long long demotionBits(const WorldView& w, int c, int bA, int bB, bool dA, bool dB) {
    const int col = w.ccolour[c];
    const unsigned long long used = (dA ? w.colourMask[bA] : 0ull) | (dB ? w.colourMask[bB] : 0ull);
    const unsigned long long below = col >= kMaxColours ? ~0ull : ((1ull << col) - 1ull);
    return ~used & below;
}

__global__ void k_demote_claim(WorldView w, unsigned long long* claim, uint32_t tag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (!(w.cflags[c] & CF_CONSTRAINT) || w.ccolour[c] <= 0) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool dA = w.btype[bA] == BT_DYNAMIC, dB = w.btype[bB] == BT_DYNAMIC;
    if (demotionBits(w, c, bA, bB, dA, dB) == 0ull) return;
    const unsigned long long v = claimValue(tag, w, c);
    if (dA) atomicMax(&claim[bA], v);
    if (dB) atomicMax(&claim[bB], v);
}

__global__ void k_demote_commit(WorldView w, const unsigned long long* claim, uint32_t tag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (!(w.cflags[c] & CF_CONSTRAINT) || w.ccolour[c] <= 0) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool dA = w.btype[bA] == BT_DYNAMIC, dB = w.btype[bB] == BT_DYNAMIC;
    const unsigned long long v = claimValue(tag, w, c);
    if ((dA && claim[bA] != v) || (dB && claim[bB] != v)) return;
    const unsigned long long freeBits = demotionBits(w, c, bA, bB, dA, dB);
    if (freeBits == 0ull) return;
    const int col = w.ccolour[c];
    const int to = __ffsll((long long)freeBits) - 1;
    const unsigned long long keep = col >= kMaxColours ? ~0ull : ~(1ull << col), bit = 1ull << to;
    if (dA) {
        atomicAnd(&w.colourMask[bA], keep);
        atomicOr(&w.colourMask[bA], bit);
    }
    if (dB) {
        atomicAnd(&w.colourMask[bB], keep);
        atomicOr(&w.colourMask[bB], bit);
    }
    w.ccolour[c] = to;
}

// Order key of every contact: a constraint's colour; contacts outside the
// constraint set sort after every colour.
// Also clears the colour counts that the order's counting sort then takes.
__global__ void k_cons_keys(WorldView w, uint32_t* keys, int* vals, Counters* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < kColourSlots + 3) cnt->colourCount[c] = 0;
    if (c >= w.nContacts) return;
    keys[c] = (w.cflags[c] & CF_CONSTRAINT) ? (uint32_t)w.ccolour[c] : kNotConstraintKey;
    vals[c] = c;
}


// ====================================================================
// Island solve stages (b2Island::Solve, REF/Dynamics/b2Island.cpp:182-383)
// ====================================================================
__device__ inline bool islandBody(const WorldView& w, int b) { return (w.bflags[b] & BF_ISLAND) != 0; }

__device__ inline void integrateVelocityOne(const WorldView& w, int b, V2 gravity, float h) {
    const float4 p = w.pos[b];
    w.pos0[b] = make_float4(p.x, p.y, p.z, w.pos0[b].w);
    if (w.btype[b] != BT_DYNAMIC) return;
    float4 vv = w.vel[b];
    V2 v = mk(vv.x, vv.y);
    float wv = vv.z;
    const float4 ms = w.mass[b];
    const float4 f = w.force[b];
    const float gs = w.gsleep[b].x;
    v += h * (gs * gravity + ms.x * mk(f.x, f.y));
    wv += h * ms.y * f.z;
    v *= 1.0f / (1.0f + h * ms.z);
    wv *= 1.0f / (1.0f + h * ms.w);
    w.vel[b] = make_float4(v.x, v.y, wv, 0.0f);
}

__global__ void k_integrate_velocities(WorldView w, V2 gravity, float h) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || !islandBody(w, b)) return;
    integrateVelocityOne(w, b, gravity, h);
}

// Fresh contact points: points that start the step without a carried
// impulse (new contacts, new feature ids, or warm starting off). Their
// features feed the warm-start network and the training log.
constexpr int kFeatures = 16;
struct FreshView {
    int* mask;    // per constraint: bit j set if point j is a network candidate
    int* num;     // popcount of mask
    float* feat;  // [2 * constraints][kFeatures], written for candidate points only
    int scope;    // 0: fresh points; 1: every solved point, carried impulse as input
};

// b2ContactSolver ctor + InitializeVelocityConstraints,
// REF/Dynamics/Contacts/b2ContactSolver.cpp:52-252.
__device__ inline void initConstraintOne(const WorldView& w, const ConView& cv, int k, int c, float dtRatio, int warm,
                                         Counters* cnt, const FreshView& fv) {
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    const b2g_manifold& m = w.cman[c];
    const float4 msA = w.mass[bA], msB = w.mass[bB];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float radiusA = w.shape[fA].radius, radiusB = w.shape[fB].radius;
    const float2 lcA2 = w.lcenter[bA], lcB2 = w.lcenter[bB];
    const V2 lcA = mk(lcA2.x, lcA2.y), lcB = mk(lcB2.x, lcB2.y);
    const float4 pA = w.pos[bA], pB = w.pos[bB];
    const float4 vA4 = w.vel[bA], vB4 = w.vel[bB];
    const V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    const V2 vA = mk(vA4.x, vA4.y), vB = mk(vB4.x, vB4.y);
    const float wA = vA4.z, wB = vB4.z;
    const float2 mat = w.cmat[c];
    const int pc = m.pointCount;
    atomicAdd(&cnt->points, pc);

    Xf xfA, xfB;
    xfA.q = rotOf(pA.z);
    xfB.q = rotOf(pB.z);
    xfA.p = cA - mul(xfA.q, lcA);
    xfB.p = cB - mul(xfB.q, lcB);
    WorldManifold wm;
    worldManifold(wm, m, xfA, radiusA, xfB, radiusB);
    const V2 normal = wm.normal;
    const V2 tangent = cross(normal, 1.0f);

    float nm[2] = {0.0f, 0.0f}, tm[2] = {0.0f, 0.0f}, vb[2] = {0.0f, 0.0f};
    V2 rA[2] = {mk(0.0f, 0.0f), mk(0.0f, 0.0f)}, rB[2] = {mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
    float ni[2] = {0.0f, 0.0f}, ti[2] = {0.0f, 0.0f};
    for (int j = 0; j < pc; ++j) {
        if (warm) {
            ni[j] = dtRatio * m.points[j].normalImpulse;
            ti[j] = dtRatio * m.points[j].tangentImpulse;
        }
        rA[j] = wm.points[j] - cA;
        rB[j] = wm.points[j] - cB;
        const float rnA = cross(rA[j], normal);
        const float rnB = cross(rB[j], normal);
        const float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
        nm[j] = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
        const float rtA = cross(rA[j], tangent);
        const float rtB = cross(rB[j], tangent);
        const float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
        tm[j] = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
        const float vRel = dot(normal, vB + cross(wB, rB[j]) - vA - cross(wA, rA[j]));
        if (vRel < -kVelocityThreshold) vb[j] = -mat.y * vRel;
    }
    int vpc = pc;
    float4 K = make_float4(0.0f, 0.0f, 0.0f, 0.0f), NM = K;
    if (pc == 2) {
        const float rn1A = cross(rA[0], normal), rn1B = cross(rB[0], normal);
        const float rn2A = cross(rA[1], normal), rn2B = cross(rB[1], normal);
        const float k11 = mA + mB + iA * rn1A * rn1A + iB * rn1B * rn1B;
        const float k22 = mA + mB + iA * rn2A * rn2A + iB * rn2B * rn2B;
        const float k12 = mA + mB + iA * rn1A * rn2A + iB * rn1B * rn2B;
        if (k11 * k11 < 1000.0f * (k11 * k22 - k12 * k12)) {
            K = make_float4(k11, k12, k12, k22);
            // b2Mat22::GetInverse: a = ex.x, b = ey.x, c = ex.y, d = ey.y
            float det = k11 * k22 - k12 * k12;
            if (det != 0.0f) det = 1.0f / det;
            NM = make_float4(det * k22, -det * k12, -det * k12, det * k11);
        } else {
            vpc = 1;
        }
    }
    const bool dynA = w.btype[bA] == BT_DYNAMIC, dynB = w.btype[bB] == BT_DYNAMIC;
    cv.contact[k] = c;
    cv.bodies[k] = make_int2(bA, bB);
    cv.mass[k] = make_float4(mA, iA, mB, iB);
    cv.normal[k] = make_float4(normal.x, normal.y, mat.x, w.ctan[c]);
    cv.r0[k] = make_float4(rA[0].x, rA[0].y, rB[0].x, rB[0].y);
    cv.r1[k] = make_float4(rA[1].x, rA[1].y, rB[1].x, rB[1].y);
    cv.pmass[k] = make_float4(nm[0], tm[0], nm[1], tm[1]);
    cv.bias[k] = make_float2(vb[0], vb[1]);
    cv.imp[k] = make_float4(ni[0], ti[0], ni[1], ti[1]);
    cv.K[k] = K;
    cv.NM[k] = NM;
    cv.info[k] = vpc | (pc << 2) | (dynA ? 16 : 0) | (dynB ? 32 : 0) | (m.type << 8);
    cv.lp[k] = make_float4(m.points[0].localPoint.x, m.points[0].localPoint.y, m.points[1].localPoint.x,
                           m.points[1].localPoint.y);
    cv.lnp[k] = make_float4(m.localNormal.x, m.localNormal.y, m.localPoint.x, m.localPoint.y);
    cv.lc[k] = make_float4(lcA.x, lcA.y, lcB.x, lcB.y);
    cv.rad[k] = make_float2(radiusA, radiusB);
    cv.island[k] = w.islandOf[w.btype[bA] != BT_STATIC ? bA : bB];
    if (fv.mask) {
        // Features of the candidate points: relative normal and tangent
        // velocity, normal, separation, inverse masses and inertias, point
        // count, normal and tangent mass, friction, r x n of both bodies,
        // restitution bias. In scope 1 (every point) features 9 and 15 are
        // the carried normal and tangent impulse instead, in velocity-change
        // units (impulse over the point's normal or tangent mass).
        int fm = 0;
        for (int j = 0; j < vpc; ++j) {
            if (fv.scope == 0 && (ni[j] != 0.0f || ti[j] != 0.0f)) continue;
            fm |= 1 << j;
            const V2 dv = vB + cross(wB, rB[j]) - vA - cross(wA, rA[j]);
            const float f9 = fv.scope == 0 ? (float)vpc : (nm[j] > 0.0f ? ni[j] / nm[j] : 0.0f);
            const float f15 = fv.scope == 0 ? vb[j] : (tm[j] > 0.0f ? ti[j] / tm[j] : 0.0f);
            float4* f = (float4*)(fv.feat + (size_t)(2 * k + j) * kFeatures);
            f[0] = make_float4(dot(dv, normal), dot(dv, tangent), normal.x, normal.y);
            f[1] = make_float4(wm.separations[j], mA, mB, iA);
            f[2] = make_float4(iB, f9, nm[j], tm[j]);
            f[3] = make_float4(mat.x, cross(rA[j], normal), cross(rB[j], normal), f15);
        }
        fv.mask[k] = fm;
        fv.num[k] = __popc(fm);
    }
}

__global__ void k_init_constraints(WorldView w, ConView cv, const int* order, int n, float dtRatio, int warm,
                                   Counters* cnt, FreshView fv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    initConstraintOne(w, cv, k, order[k], dtRatio, warm, cnt, fv);
}

// Compacts the fresh points in constraint order: ref = (k << 1) | j.
__global__ void k_fresh_gather(int n, FreshView fv, const int* off, float* feat, int* ref, Counters* cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int m = fv.mask[k];
    if (k == n - 1) cnt->nFresh = off[k] + __popc(m);
    int o = off[k];
    for (int j = 0; j < 2; ++j) {
        if (!((m >> j) & 1)) continue;
        const float4* src = (const float4*)(fv.feat + (size_t)(2 * k + j) * kFeatures);
        float4* dst = (float4*)(feat + (size_t)o * kFeatures);
        for (int q = 0; q < kFeatures / 4; ++q) dst[q] = src[q];
        ref[o] = 2 * k + j;
        ++o;
    }
}

// Appends the candidate points of the step to the training log: features
// and the impulses the solver gave them, in velocity-change units (the
// normal impulse over the point's normal mass, the tangent impulse over its
// tangent mass), so that an impact's target is about its approach speed. In
// scope 1 the target is the change from the carried impulse (features 9 and
// 15).
__global__ void k_log_append(int nFresh, const int* ref, const float* feat, const float4* imp, float* logFeat,
                             float* logTarget, int base, int cap, int scope) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nFresh || base + e >= cap) return;
    const size_t d = (size_t)(base + e);
    const float* f = feat + (size_t)e * kFeatures;
    const float4* src = (const float4*)f;
    float4* dst = (float4*)(logFeat + d * kFeatures);
    for (int q = 0; q < kFeatures / 4; ++q) dst[q] = src[q];
    const int r = ref[e];
    const float4 v = imp[r >> 1];
    const float nm = f[10], tm = f[11];
    const float ln = (r & 1) ? v.z : v.x, lt = (r & 1) ? v.w : v.y;
    logTarget[2 * d] = (nm > 0.0f ? ln / nm : 0.0f) - (scope == 1 ? f[9] : 0.0f);
    logTarget[2 * d + 1] = (tm > 0.0f ? lt / tm : 0.0f) - (scope == 1 ? f[15] : 0.0f);
}

// Complementarity residual of the velocity constraints, summed over the
// solved points: r_n = vn - bias where the normal impulse is positive, else
// min(vn - bias, 0); r_t = vt where the tangent impulse is inside the
// friction cone, else 0.
__global__ void k_residual(WorldView w, ConView cv, int n, Counters* cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float sum = 0.0f;
    int pts = 0;
    if (k < n) {
        const int2 bi = cv.bodies[k];
        const int vpc = cv.info[k] & 3;
        const float4 va = w.vel[bi.x], vbv = w.vel[bi.y];
        const V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
        const float4 nf = cv.normal[k];
        const V2 normal = mk(nf.x, nf.y), tangent = cross(normal, 1.0f);
        const float4 imp = cv.imp[k];
        const float2 bias = cv.bias[k];
        for (int j = 0; j < vpc; ++j) {
            const float4 r = j == 0 ? cv.r0[k] : cv.r1[k];
            const V2 dv = vB + cross(vbv.z, mk(r.z, r.w)) - vA - cross(va.z, mk(r.x, r.y));
            const float ln = j == 0 ? imp.x : imp.z, lt = j == 0 ? imp.y : imp.w;
            const float vn = dot(dv, normal) - (j == 0 ? bias.x : bias.y);
            const float rn = ln > 0.0f ? vn : fminB(vn, 0.0f);
            const float rt = absB(lt) < nf.z * ln ? dot(dv, tangent) : 0.0f;
            sum += rn * rn + rt * rt;
            ++pts;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_down(sum, off);
        pts += __shfl_down(pts, off);
    }
    if ((threadIdx.x & 63) == 0 && pts > 0) {
        atomicAdd(&cnt->residualSum, sum);
        atomicAdd(&cnt->residualPoints, pts);
    }
}

// b2ContactSolver::WarmStart, REF/Dynamics/Contacts/b2ContactSolver.cpp:254-292.
// The body velocities are passed in and returned (w = 0), so the same
// arithmetic serves global and LDS-staged bodies.
__device__ inline void warmStartCore(const ConView& cv, int k, int info, float4& va, float4& vbv) {
    const float4 ms = cv.mass[k];
    const int pc = info & 3;
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 nf = cv.normal[k];
    const V2 normal = mk(nf.x, nf.y);
    const V2 tangent = cross(normal, 1.0f);
    const float4 imp = cv.imp[k];
    const float4 r0 = cv.r0[k], r1 = cv.r1[k];
    for (int j = 0; j < pc; ++j) {
        const float ni = j == 0 ? imp.x : imp.z;
        const float ti = j == 0 ? imp.y : imp.w;
        const float4 r = j == 0 ? r0 : r1;
        const V2 P = ni * normal + ti * tangent;
        wA -= ms.y * cross(mk(r.x, r.y), P);
        vA -= ms.x * P;
        wB += ms.w * cross(mk(r.z, r.w), P);
        vB += ms.z * P;
    }
    va = make_float4(vA.x, vA.y, wA, 0.0f);
    vbv = make_float4(vB.x, vB.y, wB, 0.0f);
}

__device__ inline void warmStartOne(const WorldView& w, const ConView& cv, int k) {
    const int2 bi = cv.bodies[k];
    const int info = cv.info[k];
    float4 va = w.vel[bi.x], vb = w.vel[bi.y];
    warmStartCore(cv, k, info, va, vb);
    if (info & 16) w.vel[bi.x] = va;
    if (info & 32) w.vel[bi.y] = vb;
}

__global__ void k_warm_start(WorldView w, ConView cv, int begin, int end) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    warmStartOne(w, cv, k);
}

// The overflow colour holds constraints that share bodies: one lane solves
// them in order (launched as a single one-lane workgroup).
__global__ void k_warm_start_serial(WorldView w, ConView cv, int begin, int end) {
    for (int k = begin; k < end; ++k) warmStartOne(w, cv, k);
}

// b2ContactSolver::SolveVelocityConstraints, REF/Dynamics/Contacts/b2ContactSolver.cpp:294-597:
// friction per point clamped to the friction cone, then the normal impulse,
// by a 2x2 linear complementarity solve (total enumeration) for two points.
__device__ inline void solveVelocityCore(const ConView& cv, int k, int info, float4& va, float4& vbv) {
    const float4 ms = cv.mass[k];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const int pc = info & 3;
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 nf = cv.normal[k];
    const V2 normal = mk(nf.x, nf.y);
    const V2 tangent = cross(normal, 1.0f);
    const float friction = nf.z;
    const float tangentSpeed = nf.w;
    float4 imp = cv.imp[k];
    const float4 pm = cv.pmass[k];
    const float4 r0 = cv.r0[k], r1 = cv.r1[k];
    const V2 rA0 = mk(r0.x, r0.y), rB0 = mk(r0.z, r0.w);
    const V2 rA1 = mk(r1.x, r1.y), rB1 = mk(r1.z, r1.w);

    // Tangent constraints.
    {
        const V2 dv = vB + cross(wB, rB0) - vA - cross(wA, rA0);
        const float vt = dot(dv, tangent) - tangentSpeed;
        float lambda = pm.y * (-vt);
        const float maxF = friction * imp.x;
        const float ni = clampB(imp.y + lambda, -maxF, maxF);
        lambda = ni - imp.y;
        imp.y = ni;
        const V2 P = lambda * tangent;
        vA -= mA * P;
        wA -= iA * cross(rA0, P);
        vB += mB * P;
        wB += iB * cross(rB0, P);
    }
    if (pc == 2) {
        const V2 dv = vB + cross(wB, rB1) - vA - cross(wA, rA1);
        const float vt = dot(dv, tangent) - tangentSpeed;
        float lambda = pm.w * (-vt);
        const float maxF = friction * imp.z;
        const float ni = clampB(imp.w + lambda, -maxF, maxF);
        lambda = ni - imp.w;
        imp.w = ni;
        const V2 P = lambda * tangent;
        vA -= mA * P;
        wA -= iA * cross(rA1, P);
        vB += mB * P;
        wB += iB * cross(rB1, P);
    }
    // Normal constraints.
    if (pc == 1) {
        const V2 dv = vB + cross(wB, rB0) - vA - cross(wA, rA0);
        const float vn = dot(dv, normal);
        float lambda = -pm.x * (vn - cv.bias[k].x);
        const float ni = fmaxB(imp.x + lambda, 0.0f);
        lambda = ni - imp.x;
        imp.x = ni;
        const V2 P = lambda * normal;
        vA -= mA * P;
        wA -= iA * cross(rA0, P);
        vB += mB * P;
        wB += iB * cross(rB0, P);
    } else {
        const float2 vbias = cv.bias[k];
        const float4 K = cv.K[k], NM = cv.NM[k];
        const V2 a = mk(imp.x, imp.z);
        const V2 dv1 = vB + cross(wB, rB0) - vA - cross(wA, rA0);
        const V2 dv2 = vB + cross(wB, rB1) - vA - cross(wA, rA1);
        const float vn1 = dot(dv1, normal);
        const float vn2 = dot(dv2, normal);
        V2 b;
        b.x = vn1 - vbias.x;
        b.y = vn2 - vbias.y;
        b -= mk(K.x * a.x + K.z * a.y, K.y * a.x + K.w * a.y);
        V2 x;
        bool found = false;
        // Case 1: both points active.
        {
            const V2 Mb = mk(NM.x * b.x + NM.z * b.y, NM.y * b.x + NM.w * b.y);
            x = -Mb;
            if (x.x >= 0.0f && x.y >= 0.0f) found = true;
        }
        // Case 2: point 1 active, point 2 separating.
        if (!found) {
            x.x = -pm.x * b.x;
            x.y = 0.0f;
            const float vn2b = K.y * x.x + b.y;
            if (x.x >= 0.0f && vn2b >= 0.0f) found = true;
        }
        // Case 3: point 2 active, point 1 separating.
        if (!found) {
            x.x = 0.0f;
            x.y = -pm.z * b.y;
            const float vn1b = K.z * x.y + b.x;
            if (x.y >= 0.0f && vn1b >= 0.0f) found = true;
        }
        // Case 4: both separating.
        if (!found) {
            x.x = 0.0f;
            x.y = 0.0f;
            if (b.x >= 0.0f && b.y >= 0.0f) found = true;
        }
        if (found) {
            const V2 d = x - a;
            const V2 P1 = d.x * normal;
            const V2 P2 = d.y * normal;
            vA -= mA * (P1 + P2);
            wA -= iA * (cross(rA0, P1) + cross(rA1, P2));
            vB += mB * (P1 + P2);
            wB += iB * (cross(rB0, P1) + cross(rB1, P2));
            imp.x = x.x;
            imp.z = x.y;
        }
    }
    cv.imp[k] = imp;
    va = make_float4(vA.x, vA.y, wA, 0.0f);
    vbv = make_float4(vB.x, vB.y, wB, 0.0f);
}

__device__ inline void solveVelocityOne(const WorldView& w, const ConView& cv, int k) {
    const int2 bi = cv.bodies[k];
    const int info = cv.info[k];
    float4 va = w.vel[bi.x], vb = w.vel[bi.y];
    solveVelocityCore(cv, k, info, va, vb);
    if (info & 16) w.vel[bi.x] = va;
    if (info & 32) w.vel[bi.y] = vb;
}

__global__ void k_solve_velocity(WorldView w, ConView cv, int begin, int end) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    solveVelocityOne(w, cv, k);
}

__global__ void k_solve_velocity_serial(WorldView w, ConView cv, int begin, int end) {
    for (int k = begin; k < end; ++k) solveVelocityOne(w, cv, k);
}

// b2ContactSolver::StoreImpulses, REF/Dynamics/Contacts/b2ContactSolver.cpp:599-612.
__device__ inline void storeImpulsesOne(const WorldView& w, const ConView& cv, int k) {
    const int c = cv.contact[k];
    const int vpc = cv.info[k] & 3;
    const float4 imp = cv.imp[k];
    w.cman[c].points[0].normalImpulse = imp.x;
    w.cman[c].points[0].tangentImpulse = imp.y;
    if (vpc == 2) {
        w.cman[c].points[1].normalImpulse = imp.z;
        w.cman[c].points[1].tangentImpulse = imp.w;
    }
}

__global__ void k_store_impulses(WorldView w, ConView cv, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    storeImpulsesOne(w, cv, k);
}

// Kernels of a captured solve: colour c's range [ranges[c], ranges[c + 1])
// is read from device memory at run time (ranges[kColourSlots] is the
// constraint count), and any grid covers it by striding, so one captured
// launch sequence serves every step with the same colour layout.
__global__ void k_warm_start_r(WorldView w, ConView cv, const int* ranges, int c) {
    const int b1 = ranges[c + 1];
    for (int k = ranges[c] + blockIdx.x * blockDim.x + threadIdx.x; k < b1; k += gridDim.x * blockDim.x)
        warmStartOne(w, cv, k);
}

__global__ void k_warm_start_serial_r(WorldView w, ConView cv, const int* ranges, int c) {
    for (int k = ranges[c]; k < ranges[c + 1]; ++k) warmStartOne(w, cv, k);
}

__global__ void k_solve_velocity_r(WorldView w, ConView cv, const int* ranges, int c) {
    const int b1 = ranges[c + 1];
    for (int k = ranges[c] + blockIdx.x * blockDim.x + threadIdx.x; k < b1; k += gridDim.x * blockDim.x)
        solveVelocityOne(w, cv, k);
}

__global__ void k_solve_velocity_serial_r(WorldView w, ConView cv, const int* ranges, int c) {
    for (int k = ranges[c]; k < ranges[c + 1]; ++k) solveVelocityOne(w, cv, k);
}

__global__ void k_store_impulses_r(WorldView w, ConView cv, const int* ranges) {
    const int n = ranges[kColourSlots];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) storeImpulsesOne(w, cv, k);
}

// Position integration with translation and rotation clamps,
// REF/Dynamics/b2Island.cpp:276-306.
__device__ inline void integratePositionOne(const WorldView& w, int b, float h) {
    float4 p = w.pos[b];
    const float4 vv = w.vel[b];
    V2 c = mk(p.x, p.y);
    float a = p.z;
    V2 v = mk(vv.x, vv.y);
    float wv = vv.z;
    const V2 translation = h * v;
    if (dot(translation, translation) > kMaxTranslationSq) {
        const float ratio = kMaxTranslation / lengthOf(translation);
        v *= ratio;
    }
    const float rotation = h * wv;
    if (rotation * rotation > kMaxRotationSq) {
        const float ratio = kMaxRotation / absB(rotation);
        wv *= ratio;
    }
    c += h * v;
    a += h * wv;
    w.pos[b] = make_float4(c.x, c.y, a, 0.0f);
    w.vel[b] = make_float4(v.x, v.y, wv, 0.0f);
}

__global__ void k_integrate_positions(WorldView w, float h) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || !islandBody(w, b)) return;
    integratePositionOne(w, b, h);
}

// b2ContactSolver::SolvePositionConstraints with b2PositionSolverManifold,
// REF/Dynamics/Contacts/b2ContactSolver.cpp:614-744. Skips itself when the
// previous iteration already met the early-exit test of b2Island::Solve.
__device__ inline float solvePositionCore(const ConView& cv, int k, int info, float4& pA, float4& pB) {
    const float4 ms = cv.mass[k];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const int pc = (info >> 2) & 3;
    const int type = (info >> 8) & 3;
    const float4 lc = cv.lc[k];
    const V2 lcA = mk(lc.x, lc.y), lcB = mk(lc.z, lc.w);
    const float4 lp = cv.lp[k];
    const float4 lnp = cv.lnp[k];
    const float2 rad = cv.rad[k];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    float minSep = 0.0f;
    for (int j = 0; j < pc; ++j) {
        Xf xfA, xfB;
        xfA.q = rotOf(aA);
        xfB.q = rotOf(aB);
        xfA.p = cA - mul(xfA.q, lcA);
        xfB.p = cB - mul(xfB.q, lcB);
        V2 normal, point;
        float separation;
        if (type == B2G_MANIFOLD_CIRCLES) {
            const V2 pointA = mul(xfA, mk(lnp.z, lnp.w));
            const V2 pointB = mul(xfB, mk(lp.x, lp.y));
            normal = pointB - pointA;
            normalizeV(normal);
            point = 0.5f * (pointA + pointB);
            separation = dot(pointB - pointA, normal) - rad.x - rad.y;
        } else if (type == B2G_MANIFOLD_FACE_A) {
            normal = mul(xfA.q, mk(lnp.x, lnp.y));
            const V2 planePoint = mul(xfA, mk(lnp.z, lnp.w));
            const V2 clipPoint = mul(xfB, j == 0 ? mk(lp.x, lp.y) : mk(lp.z, lp.w));
            separation = dot(clipPoint - planePoint, normal) - rad.x - rad.y;
            point = clipPoint;
        } else {
            normal = mul(xfB.q, mk(lnp.x, lnp.y));
            const V2 planePoint = mul(xfB, mk(lnp.z, lnp.w));
            const V2 clipPoint = mul(xfA, j == 0 ? mk(lp.x, lp.y) : mk(lp.z, lp.w));
            separation = dot(clipPoint - planePoint, normal) - rad.x - rad.y;
            point = clipPoint;
            normal = -normal;
        }
        const V2 rA = point - cA;
        const V2 rB = point - cB;
        minSep = fminB(minSep, separation);
        const float C = clampB(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
        const float rnA = cross(rA, normal);
        const float rnB = cross(rB, normal);
        const float Kn = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
        const float impulse = Kn > 0.0f ? -C / Kn : 0.0f;
        const V2 P = impulse * normal;
        cA -= mA * P;
        aA -= iA * cross(rA, P);
        cB += mB * P;
        aB += iB * cross(rB, P);
    }
    pA = make_float4(cA.x, cA.y, aA, 0.0f);
    pB = make_float4(cB.x, cB.y, aB, 0.0f);
    return minSep;
}

__device__ inline float solvePositionOne(const WorldView& w, const ConView& cv, int k) {
    const int2 bi = cv.bodies[k];
    const int info = cv.info[k];
    float4 pA = w.pos[bi.x], pB = w.pos[bi.y];
    const float minSep = solvePositionCore(cv, k, info, pA, pB);
    if (info & 16) w.pos[bi.x] = pA;
    if (info & 32) w.pos[bi.y] = pB;
    return minSep;
}

// Early exit per island (b2Island::Solve, REF/Dynamics/b2Island.cpp:310-328):
// a pass leaves an island solved when none of its points is deeper than
// 3 * linearSlop, so per island and pass one "failed" flag suffices. Lanes
// set it only while it is still clear, which keeps the writes to a large
// island's flag few.
__global__ void k_solve_position(WorldView w, ConView cv, int begin, int end, Counters* cnt, int it) {
    __shared__ int blockMin;
    if (threadIdx.x == 0) blockMin = 0;
    __syncthreads();
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k < end) {
        const int root = cv.island[k];
        const bool done = it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0;
        if (!done) {
            const float minSep = solvePositionOne(w, cv, k);
            if (minSep < -3.0f * kLinearSlop) {
                int* f = &w.islandMinSep[it * w.nBodies + root];
                if (*f == 0) *f = 1;
            }
            if (minSep < 0.0f) atomicMin(&blockMin, encodeOrdered(minSep));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && blockMin < 0) atomicMin(&cnt->minSep[it], blockMin);
}

__global__ void k_solve_position_r(WorldView w, ConView cv, const int* ranges, int c, Counters* cnt, int it) {
    __shared__ int blockMin;
    if (threadIdx.x == 0) blockMin = 0;
    __syncthreads();
    const int b1 = ranges[c + 1];
    int lowest = 0;
    for (int k = ranges[c] + blockIdx.x * blockDim.x + threadIdx.x; k < b1; k += gridDim.x * blockDim.x) {
        const int root = cv.island[k];
        if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) continue;
        const float minSep = solvePositionOne(w, cv, k);
        if (minSep < -3.0f * kLinearSlop) {
            int* f = &w.islandMinSep[it * w.nBodies + root];
            if (*f == 0) *f = 1;
        }
        if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
    }
    if (lowest < 0) atomicMin(&blockMin, lowest);
    __syncthreads();
    if (threadIdx.x == 0 && blockMin < 0) atomicMin(&cnt->minSep[it], blockMin);
}

__global__ void k_solve_position_serial_r(WorldView w, ConView cv, const int* ranges, int c, Counters* cnt, int it) {
    int lowest = 0;
    for (int k = ranges[c]; k < ranges[c + 1]; ++k) {
        const int root = cv.island[k];
        if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) continue;
        const float minSep = solvePositionOne(w, cv, k);
        if (minSep < -3.0f * kLinearSlop) w.islandMinSep[it * w.nBodies + root] = 1;
        if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
    }
    if (lowest < 0) atomicMin(&cnt->minSep[it], lowest);
}

// Trailing colours in one workgroup: colours [c0, c1) in order, the
// workgroup's lanes sharing each colour and a barrier between colours (the
// barrier orders the workgroup's global writes, which the next colour
// reads); the overflow colour is solved by lane 0 alone, in order.
struct ColourRanges {
    int s[kColourSlots + 1];
};

__global__ void k_warm_start_tail(WorldView w, ConView cv, ColourRanges r, int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
        if (c == kOverflowColour) {
            if (threadIdx.x == 0)
                for (int k = r.s[c]; k < r.s[c + 1]; ++k) warmStartOne(w, cv, k);
        } else {
            for (int k = r.s[c] + threadIdx.x; k < r.s[c + 1]; k += blockDim.x) warmStartOne(w, cv, k);
        }
        __syncthreads();
    }
}

__global__ void k_solve_velocity_tail(WorldView w, ConView cv, ColourRanges r, int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
        if (c == kOverflowColour) {
            if (threadIdx.x == 0)
                for (int k = r.s[c]; k < r.s[c + 1]; ++k) solveVelocityOne(w, cv, k);
        } else {
            for (int k = r.s[c] + threadIdx.x; k < r.s[c + 1]; k += blockDim.x) solveVelocityOne(w, cv, k);
        }
        __syncthreads();
    }
}

__global__ void k_solve_position_tail(WorldView w, ConView cv, ColourRanges r, int c0, int c1, Counters* cnt,
                                      int it) {
    __shared__ int blockMin;
    if (threadIdx.x == 0) blockMin = 0;
    __syncthreads();
    int lowest = 0;
    for (int c = c0; c < c1; ++c) {
        const bool serial = c == kOverflowColour;
        if (!serial || threadIdx.x == 0) {
            for (int k = r.s[c] + (serial ? 0 : threadIdx.x); k < r.s[c + 1]; k += serial ? 1 : blockDim.x) {
                const int root = cv.island[k];
                if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) continue;
                const float minSep = solvePositionOne(w, cv, k);
                if (minSep < -3.0f * kLinearSlop) {
                    int* f = &w.islandMinSep[it * w.nBodies + root];
                    if (*f == 0) *f = 1;
                }
                if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
            }
        }
        __syncthreads();
    }
    if (lowest < 0) atomicMin(&blockMin, lowest);
    __syncthreads();
    if (threadIdx.x == 0 && blockMin < 0) atomicMin(&cnt->minSep[it], blockMin);
}

// ====================================================================
// Clustered solve. Awake island bodies are sorted by (sub-world, Morton code
// of their 1 m cell) and cut into clusters of a fixed size. A constraint
// whose dynamic bodies lie in one cluster is interior: one workgroup per
// cluster solves its interior constraints with the cluster's body states
// staged in LDS, colour after colour with a workgroup barrier between
// colours. The constraints between clusters (the boundary set) are solved
// after them by colour, as in the unclustered solve. Each constraint is
// still visited once per pass and no two constraints solved at once share a
// dynamic body, so a pass remains a Gauss-Seidel sweep in another order.
// ====================================================================
constexpr int kBoundaryCluster = 0xFFFFFF;  // cluster field of a boundary constraint's order key
constexpr int kMaxClusterBodies = 1 << 21;  // body index bits of the cluster sort key

// Sort key per body: (sub-world, Morton code of its cell, body index); bodies
// outside the islands sort last.
__device__ inline uint32_t spreadBits15(uint32_t v) {
    v &= 0x7FFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_body_cluster_keys(WorldView w, unsigned long long* keys, int2* cluster, Counters* cnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    cluster[b] = make_int2(-1, -1);
    if (w.btype[b] != BT_DYNAMIC || !(w.bflags[b] & BF_ISLAND)) {
        keys[b] = ~0ull;
        return;
    }
    atomicAdd(&cnt->nClustered, 1);
    const float4 p = w.pos[b];
    const int cx = min(max((int)floorf(p.x) + 16384, 0), 32767);
    const int cy = min(max((int)floorf(p.y) + 16384, 0), 32767);
    const uint32_t morton = spreadBits15((uint32_t)cx) | (spreadBits15((uint32_t)cy) << 1);
    const unsigned long long world = w.bworld ? (unsigned long long)(w.bworld[b] & 0xFFF) : 0ull;
    keys[b] = (world << 51) | ((unsigned long long)morton << 21) | (unsigned long long)b;
}

// Cluster and LDS slot of each clustered body, and the bodies by rank.
__global__ void k_body_clusters(const unsigned long long* sorted, int n, int size, const Counters* cnt,
                                int2* cluster, int* bodies) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || r >= cnt->nClustered) return;
    const int b = (int)(sorted[r] & (unsigned long long)(kMaxClusterBodies - 1));
    cluster[b] = make_int2(r / size, r % size);
    bodies[r] = b;
}

// Order key of the clustered solve: (cluster or the boundary set, colour)
// for a constraint; contacts outside the constraint set sort last.
__global__ void k_cons_keys_clustered(WorldView w, const int2* cluster, uint32_t* keys, int* vals, Counters* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < kColourSlots + 3) cnt->colourCount[c] = 0;
    if (c >= w.nContacts) return;
    vals[c] = c;
    if (!(w.cflags[c] & CF_CONSTRAINT)) {
        keys[c] = 0x7FFFFFFFu;
        return;
    }
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool dA = w.btype[bA] == BT_DYNAMIC, dB = w.btype[bB] == BT_DYNAMIC;
    const int ca = dA ? cluster[bA].x : -1, cb = dB ? cluster[bB].x : -1;
    const bool boundary = (dA && ca < 0) || (dB && cb < 0) || (dA && dB && ca != cb);
    const uint32_t q = boundary ? (uint32_t)kBoundaryCluster : (uint32_t)max(ca, cb);
    const int colour = w.ccolour[c];
    atomicMax(&cnt->maxColour, colour);
    keys[c] = (q << 7) | (uint32_t)colour;
}

// start[q * 66 + c]: first constraint of cluster q with colour >= c (q =
// nClusters: the boundary set); start[q * 66 + 65] ends the cluster. The
// boundary set's colour sizes and start go to the counters for the host.
__global__ void k_cluster_ranges(const uint32_t* keys, int n, int nClusters, int* start, Counters* cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = t / (kColourSlots + 1), c = t % (kColourSlots + 1);
    if (q > nClusters) return;
    const uint32_t cl = q == nClusters ? (uint32_t)kBoundaryCluster : (uint32_t)q;
    auto lower = [&](uint32_t key) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int s0 = lower((cl << 7) | (uint32_t)c);
    start[t] = s0;
    if (q == nClusters) {
        if (c == 0) cnt->boundaryStart = s0;
        if (c < kColourSlots) cnt->colourCount[c] = lower((cl << 7) | (uint32_t)(c + 1)) - s0;
    }
}

struct ClusterView {
    const int* start;   // per cluster: kColourSlots + 1 constraint offsets
    const int* bodies;  // clustered bodies by rank
    const int* count;   // number of clustered bodies (device counter)
    int size;           // bodies per cluster
};

// Shared set-up of the cluster kernels: the cluster's body rows (velocities
// or positions) into LDS and its colour offsets; returns the number of bodies
// staged, or -1 for a workgroup past the last cluster.
__device__ inline int stageCluster(const ClusterView& cl, const float4* rows, float4* st, int* rs) {
    const int base = blockIdx.x * cl.size;
    const int total = *cl.count;
    if (base >= total) return -1;
    const int nb = min(cl.size, total - base);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) st[i] = rows[cl.bodies[base + i]];
    for (int i = threadIdx.x; i <= kColourSlots; i += blockDim.x)
        rs[i] = cl.start[blockIdx.x * (kColourSlots + 1) + i];
    __syncthreads();
    return nb;
}

__device__ inline void unstageCluster(const ClusterView& cl, float4* rows, const float4* st, int nb) {
    __syncthreads();
    const int base = blockIdx.x * cl.size;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) rows[cl.bodies[base + i]] = st[i];
}

// Runs fn(k) over the cluster's interior constraints colour by colour; the
// overflow colour by lane 0 in order.
template <typename F>
__device__ inline void clusterColours(const int* rs, F fn) {
    for (int c = 0; c < kColourSlots; ++c) {
        const int b0 = rs[c], b1 = rs[c + 1];
        if (b1 <= b0) continue;
        if (c == kOverflowColour) {
            if (threadIdx.x == 0)
                for (int k = b0; k < b1; ++k) fn(k);
        } else {
            for (int k = b0 + threadIdx.x; k < b1; k += blockDim.x) fn(k);
        }
        __syncthreads();
    }
}

__device__ inline float4 stagedRow(const float4* st, const float4* rows, int slot, int body) {
    return slot >= 0 ? st[slot] : rows[body];
}

// LDS slots of each constraint's dynamic bodies (used by interior constraints).
__global__ void k_cons_slots(ConView cv, int n, const int2* cluster) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int2 bi = cv.bodies[k];
    const int info = cv.info[k];
    cv.slot[k] = make_int2((info & 16) ? cluster[bi.x].y : -1, (info & 32) ? cluster[bi.y].y : -1);
}

__global__ void k_warm_start_cluster(WorldView w, ConView cv, ClusterView cl) {
    extern __shared__ float4 st[];
    __shared__ int rs[kColourSlots + 1];
    const int nb = stageCluster(cl, w.vel, st, rs);
    if (nb < 0) return;
    clusterColours(rs, [&](int k) {
        const int2 bi = cv.bodies[k];
        const int2 sl = cv.slot[k];
        const int info = cv.info[k];
        float4 va = stagedRow(st, w.vel, sl.x, bi.x), vb = stagedRow(st, w.vel, sl.y, bi.y);
        warmStartCore(cv, k, info, va, vb);
        if ((info & 16) && sl.x >= 0) st[sl.x] = va;
        if ((info & 32) && sl.y >= 0) st[sl.y] = vb;
    });
    unstageCluster(cl, w.vel, st, nb);
}

__global__ void k_solve_velocity_cluster(WorldView w, ConView cv, ClusterView cl) {
    extern __shared__ float4 st[];
    __shared__ int rs[kColourSlots + 1];
    const int nb = stageCluster(cl, w.vel, st, rs);
    if (nb < 0) return;
    clusterColours(rs, [&](int k) {
        const int2 bi = cv.bodies[k];
        const int2 sl = cv.slot[k];
        const int info = cv.info[k];
        float4 va = stagedRow(st, w.vel, sl.x, bi.x), vb = stagedRow(st, w.vel, sl.y, bi.y);
        solveVelocityCore(cv, k, info, va, vb);
        if ((info & 16) && sl.x >= 0) st[sl.x] = va;
        if ((info & 32) && sl.y >= 0) st[sl.y] = vb;
    });
    unstageCluster(cl, w.vel, st, nb);
}

// Position pass with the per-island early exit of k_solve_position.
__global__ void k_solve_position_cluster(WorldView w, ConView cv, ClusterView cl, Counters* cnt, int it) {
    extern __shared__ float4 st[];
    __shared__ int rs[kColourSlots + 1];
    __shared__ int blockMin;
    if (threadIdx.x == 0) blockMin = 0;
    const int nb = stageCluster(cl, w.pos, st, rs);
    if (nb < 0) return;
    int lowest = 0;
    clusterColours(rs, [&](int k) {
        const int root = cv.island[k];
        if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) return;
        const int2 bi = cv.bodies[k];
        const int2 sl = cv.slot[k];
        const int info = cv.info[k];
        float4 pA = stagedRow(st, w.pos, sl.x, bi.x), pB = stagedRow(st, w.pos, sl.y, bi.y);
        const float minSep = solvePositionCore(cv, k, info, pA, pB);
        if ((info & 16) && sl.x >= 0) st[sl.x] = pA;
        if ((info & 32) && sl.y >= 0) st[sl.y] = pB;
        if (minSep < -3.0f * kLinearSlop) {
            int* f = &w.islandMinSep[it * w.nBodies + root];
            if (*f == 0) *f = 1;
        }
        if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
    });
    if (lowest < 0) atomicMin(&blockMin, lowest);
    unstageCluster(cl, w.pos, st, nb);
    if (threadIdx.x == 0 && blockMin < 0) atomicMin(&cnt->minSep[it], blockMin);
}

__global__ void k_solve_position_serial(WorldView w, ConView cv, int begin, int end, Counters* cnt, int it) {
    int lowest = 0;
    for (int k = begin; k < end; ++k) {
        const int root = cv.island[k];
        if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) continue;
        const float minSep = solvePositionOne(w, cv, k);
        if (minSep < -3.0f * kLinearSlop) w.islandMinSep[it * w.nBodies + root] = 1;
        if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
    }
    if (lowest < 0) atomicMin(&cnt->minSep[it], lowest);
}

// ---------------------------------------------------------------- joints
// A joint is solved when one of its bodies is in an awake island and both
// are active (the joints b2World::Solve's island search adds,
// REF/Dynamics/b2World.cpp:502-529).
__device__ inline bool jointActive(const WorldView& w, int2 b) {
    return ((w.bflags[b.x] & BF_ISLAND) || (w.bflags[b.y] & BF_ISLAND)) && (w.bflags[b.x] & BF_ACTIVE) &&
           (w.bflags[b.y] & BF_ACTIVE);
}

// b2Mat33::Solve33 / Solve22 (REF/Common/b2Math.cpp:28-56) for the symmetric
// mass matrix ex = (m11, m12, m13), ey = (m12, m22, m23), ez = (m13, m23, m33).
struct Sym33 { float m11, m12, m13, m22, m23, m33; };
__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}
__device__ inline void solve33(const Sym33& m, float bx, float by, float bz, float& x, float& y, float& z) {
    // ey x ez, b x ez, ey x b (b2Cross of b2Vec3)
    const float c1x = m.m22 * m.m33 - m.m23 * m.m23, c1y = m.m23 * m.m13 - m.m12 * m.m33,
                c1z = m.m12 * m.m23 - m.m22 * m.m13;
    float det = dot3(m.m11, m.m12, m.m13, c1x, c1y, c1z);
    if (det != 0.0f) det = 1.0f / det;
    const float c2x = by * m.m33 - bz * m.m23, c2y = bz * m.m13 - bx * m.m33, c2z = bx * m.m23 - by * m.m13;
    const float c3x = m.m22 * bz - m.m23 * by, c3y = m.m23 * bx - m.m12 * bz, c3z = m.m12 * by - m.m22 * bx;
    x = det * dot3(bx, by, bz, c1x, c1y, c1z);
    y = det * dot3(m.m11, m.m12, m.m13, c2x, c2y, c2z);
    z = det * dot3(m.m11, m.m12, m.m13, c3x, c3y, c3z);
}
__device__ inline V2 solve22(const Sym33& m, V2 b) {
    const float a11 = m.m11, a12 = m.m12, a21 = m.m12, a22 = m.m22;
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return mk(det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x));
}

// b2DistanceJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2DistanceJoint.cpp:63-154
__device__ inline void distanceInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio,
                                       int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 u = mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA;
    const float length = lengthOf(u);
    if (length > kLinearSlop) u *= 1.0f / length;
    else u = mk(0.0f, 0.0f);
    const float crAu = cross(rA, u), crBu = cross(rB, u);
    float invMass = mA + iA * crAu * crAu + mB + iB * crBu * crBu;
    float mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
    const float4 par = jv.limits[j];  // length, frequencyHz, dampingRatio
    float gamma = 0.0f, bias = 0.0f;
    if (par.y > 0.0f) {
        const float C = length - par.x;
        const float omega = 2.0f * kPi * par.y;
        const float d = 2.0f * mass * par.z * omega;
        const float k = mass * omega * omega;
        const float h = jv.h;
        gamma = h * (d + h * k);
        gamma = gamma != 0.0f ? 1.0f / gamma : 0.0f;
        bias = C * h * k * gamma;
        invMass += gamma;
        mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
    }
    float4 imp = jv.impulse[j];
    if (warm) {
        imp.x *= dtRatio;
        const V2 P = imp.x * u;
        vA -= mA * P;
        wA -= iA * cross(rA, P);
        vB += mB * P;
        wB += iB * cross(rB, P);
    } else {
        imp.x = 0.0f;
    }
    imp.y = u.x;  // the step's direction, for b2DistanceJoint::GetReactionForce
    imp.z = u.y;
    jv.impulse[j] = imp;
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(u.x, u.y, mass, bias);
    jv.K1[j] = make_float2(gamma, 0.0f);
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2DistanceJoint::SolveVelocityConstraints, REF b2DistanceJoint.cpp:156-181
__device__ inline void distanceSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const V2 u = mk(k0.x, k0.y);
    const V2 vpA = vA + cross(wA, rA);
    const V2 vpB = vB + cross(wB, rB);
    const float Cdot = dot(u, vpB - vpA);
    float4 imp = jv.impulse[j];
    const float impulse = -k0.z * (Cdot + k0.w + jv.K1[j].x * imp.x);
    imp.x += impulse;
    const V2 P = impulse * u;
    vA -= mA * P;
    wA -= iA * cross(rA, P);
    vB += mB * P;
    wB += iB * cross(rB, P);
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2DistanceJoint::SolvePositionConstraints, REF b2DistanceJoint.cpp:183-221:
// none for springs; rods report an error of |C| against the linear slop.
__device__ inline bool distanceSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 par = jv.limits[j];
    if (par.y > 0.0f) return true;
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 u = cB + rB - cA - rA;
    const float length = normalizeV(u);
    float C = length - par.x;
    C = clampB(C, -kMaxLinearCorrection, kMaxLinearCorrection);
    const float impulse = -jv.K0[j].z * C;
    const V2 P = impulse * u;
    cA -= mA * P;
    aA -= iA * cross(rA, P);
    cB += mB * P;
    aB += iB * cross(rB, P);
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return absB(C) < kLinearSlop;
}

// b2RopeJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2RopeJoint.cpp:48-126
__device__ inline void ropeInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 u = mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA;
    const float length = lengthOf(u);
    const int state = length - jv.limits[j].x > 0.0f ? JL_UPPER : JL_INACTIVE;
    jv.flags[j] = (jv.flags[j] & ~0x300) | (state << 8);
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    float4 imp = jv.impulse[j];
    if (!(length > kLinearSlop)) {
        // anchors together: no direction, no mass, no impulse this step
        jv.impulse[j] = make_float4(0.0f, 0.0f, 0.0f, (float)state);
        jv.K0[j] = make_float4(0.0f, 0.0f, 0.0f, length);
        return;
    }
    u *= 1.0f / length;
    const float crA = cross(rA, u), crB = cross(rB, u);
    const float invMass = mA + iA * crA * crA + mB + iB * crB * crB;
    const float mass = invMass != 0.0f ? 1.0f / invMass : 0.0f;
    if (warm) {
        imp.x *= dtRatio;
        const V2 P = imp.x * u;
        vA -= mA * P;
        wA -= iA * cross(rA, P);
        vB += mB * P;
        wB += iB * cross(rB, P);
    } else {
        imp.x = 0.0f;
    }
    // u for b2RopeJoint::GetReactionForce, the state for GetLimitState
    jv.impulse[j] = make_float4(imp.x, u.x, u.y, (float)state);
    jv.K0[j] = make_float4(u.x, u.y, mass, length);
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2RopeJoint::SolveVelocityConstraints, REF b2RopeJoint.cpp:128-162: a
// one-sided impulse, predictive while the rope is slack.
__device__ inline void ropeSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b, float h) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const V2 u = mk(k0.x, k0.y);
    const V2 vpA = vA + cross(wA, rA);
    const V2 vpB = vB + cross(wB, rB);
    const float C = k0.w - jv.limits[j].x;
    float Cdot = dot(u, vpB - vpA);
    if (C < 0.0f) Cdot += (1.0f / h) * C;
    float impulse = -k0.z * Cdot;
    float4 imp = jv.impulse[j];
    const float oldImpulse = imp.x;
    const float sum = imp.x + impulse;
    imp.x = 0.0f < sum ? 0.0f : sum;  // b2Min(0, sum)
    impulse = imp.x - oldImpulse;
    const V2 P = impulse * u;
    vA -= mA * P;
    wA -= iA * cross(rA, P);
    vB += mB * P;
    wB += iB * cross(rB, P);
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2RopeJoint::SolvePositionConstraints, REF b2RopeJoint.cpp:164-196
__device__ inline bool ropeSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 u = cB + rB - cA - rA;
    const float length = normalizeV(u);
    const float maxLength = jv.limits[j].x;
    const float C = clampB(length - maxLength, 0.0f, kMaxLinearCorrection);
    const float impulse = -jv.K0[j].z * C;
    const V2 P = impulse * u;
    cA -= mA * P;
    aA -= iA * cross(rA, P);
    cB += mB * P;
    aB += iB * cross(rB, P);
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return length - maxLength < kLinearSlop;
}

// The weld's 3x3 K (REF/Dynamics/Joints/b2WeldJoint.cpp:95-104; the same
// matrix b2RevoluteJoint builds).
__device__ inline Sym33 weldK(float mA, float iA, float mB, float iB, V2 rA, V2 rB) {
    Sym33 K;
    K.m11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
    K.m12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    K.m13 = -rA.y * iA - rB.y * iB;
    K.m22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
    K.m23 = rA.x * iA + rB.x * iB;
    K.m33 = iA + iB;
    return K;
}

// b2Mat33::GetInverse22 (REF/Common/b2Math.cpp:58-71): the inverse of the
// upper-left 2x2 block, zeros in the third row and column.
__device__ inline Sym33 inverse22(const Sym33& K) {
    const float a = K.m11, bb = K.m12, c = K.m12, d = K.m22;
    float det = a * d - bb * c;
    if (det != 0.0f) det = 1.0f / det;
    return Sym33{det * d, -det * c, 0.0f, det * a, 0.0f, 0.0f};
}

// b2Mat33::GetSymInverse33 (REF/Common/b2Math.cpp:73-97)
__device__ inline Sym33 symInverse33(const Sym33& K) {
    const float c1x = K.m22 * K.m33 - K.m23 * K.m23, c1y = K.m23 * K.m13 - K.m12 * K.m33,
                c1z = K.m12 * K.m23 - K.m22 * K.m13;
    float det = dot3(K.m11, K.m12, K.m13, c1x, c1y, c1z);
    if (det != 0.0f) det = 1.0f / det;
    const float a11 = K.m11, a12 = K.m12, a13 = K.m13, a22 = K.m22, a23 = K.m23, a33 = K.m33;
    Sym33 M;
    M.m11 = det * (a22 * a33 - a23 * a23);
    M.m12 = det * (a13 * a23 - a12 * a33);
    M.m13 = det * (a12 * a23 - a13 * a22);
    M.m22 = det * (a11 * a33 - a13 * a13);
    M.m23 = det * (a13 * a12 - a11 * a23);
    M.m33 = det * (a11 * a22 - a12 * a12);
    return M;
}

// b2WeldJoint::InitVelocityConstraints, REF b2WeldJoint.cpp:60-169
__device__ inline void weldInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float aA = w.pos[b.x].z, aB = w.pos[b.y].z;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const Sym33 K = weldK(mA, iA, mB, iB, rA, rB);
    const float4 par = jv.limits[j];  // referenceAngle, frequencyHz, dampingRatio
    Sym33 M;
    float gamma = 0.0f, bias = 0.0f;
    if (par.y > 0.0f) {
        M = inverse22(K);
        float invM = iA + iB;
        const float m = invM > 0.0f ? 1.0f / invM : 0.0f;
        const float C = aB - aA - par.x;
        const float omega = 2.0f * kPi * par.y;
        const float d = 2.0f * m * par.z * omega;
        const float k = m * omega * omega;
        const float h = jv.h;
        gamma = h * (d + h * k);
        gamma = gamma != 0.0f ? 1.0f / gamma : 0.0f;
        bias = C * h * k * gamma;
        invM += gamma;
        M.m33 = invM != 0.0f ? 1.0f / invM : 0.0f;
    } else if (K.m33 == 0.0f) {
        M = inverse22(K);
    } else {
        M = symInverse33(K);
    }
    float4 imp = jv.impulse[j];
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        imp.z *= dtRatio;
        const V2 P = mk(imp.x, imp.y);
        vA -= mA * P;
        wA -= iA * (cross(rA, P) + imp.z);
        vB += mB * P;
        wB += iB * (cross(rB, P) + imp.z);
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.impulse[j] = imp;
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(M.m11, M.m12, M.m13, M.m22);
    jv.K1[j] = make_float2(M.m23, M.m33);
    jv.motorMass[j] = gamma;
    jv.motorSpeed[j] = bias;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2WeldJoint::SolveVelocityConstraints, REF b2WeldJoint.cpp:171-227
__device__ inline void weldSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const float2 k1 = jv.K1[j];
    float4 imp = jv.impulse[j];
    if (jv.limits[j].y > 0.0f) {
        const float Cdot2 = wB - wA;
        const float impulse2 = -k1.y * (Cdot2 + jv.motorSpeed[j] + jv.motorMass[j] * imp.z);
        imp.z += impulse2;
        wA -= iA * impulse2;
        wB += iB * impulse2;
        const V2 Cdot1 = vB + cross(wB, rB) - vA - cross(wA, rA);
        // -b2Mul22(mass, Cdot1)
        const V2 impulse1 = -mk(k0.x * Cdot1.x + k0.y * Cdot1.y, k0.y * Cdot1.x + k0.w * Cdot1.y);
        imp.x += impulse1.x;
        imp.y += impulse1.y;
        const V2 P = impulse1;
        vA -= mA * P;
        wA -= iA * cross(rA, P);
        vB += mB * P;
        wB += iB * cross(rB, P);
    } else {
        const V2 Cdot1 = vB + cross(wB, rB) - vA - cross(wA, rA);
        const float Cdot2 = wB - wA;
        // -b2Mul(mass, Cdot): Cdot.x * ex + Cdot.y * ey + Cdot.z * ez
        const float ix = -(Cdot1.x * k0.x + Cdot1.y * k0.y + Cdot2 * k0.z);
        const float iy = -(Cdot1.x * k0.y + Cdot1.y * k0.w + Cdot2 * k1.x);
        const float iz = -(Cdot1.x * k0.z + Cdot1.y * k1.x + Cdot2 * k1.y);
        imp.x += ix;
        imp.y += iy;
        imp.z += iz;
        const V2 P = mk(ix, iy);
        vA -= mA * P;
        wA -= iA * (cross(rA, P) + iz);
        vB += mB * P;
        wB += iB * (cross(rB, P) + iz);
    }
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2WeldJoint::SolvePositionConstraints, REF b2WeldJoint.cpp:229-308: the
// spring leaves the angle to the velocity pass.
__device__ inline bool weldSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const Sym33 K = weldK(mA, iA, mB, iB, rA, rB);
    const float4 par = jv.limits[j];
    float positionError, angularError;
    const V2 C1 = cB + rB - cA - rA;
    if (par.y > 0.0f) {
        positionError = lengthOf(C1);
        angularError = 0.0f;
        const V2 P = -solve22(K, C1);
        cA -= mA * P;
        aA -= iA * cross(rA, P);
        cB += mB * P;
        aB += iB * cross(rB, P);
    } else {
        const float C2 = aB - aA - par.x;
        positionError = lengthOf(C1);
        angularError = absB(C2);
        float ix, iy, iz;
        if (K.m33 > 0.0f) {
            solve33(K, C1.x, C1.y, C2, ix, iy, iz);
            ix = -ix;
            iy = -iy;
            iz = -iz;
        } else {
            const V2 i2 = -solve22(K, C1);
            ix = i2.x;
            iy = i2.y;
            iz = 0.0f;
        }
        const V2 P = mk(ix, iy);
        cA -= mA * P;
        aA -= iA * (cross(rA, P) + iz);
        cB += mB * P;
        aB += iB * (cross(rB, P) + iz);
    }
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return positionError <= kLinearSlop && angularError <= kAngularSlop;
}

// The 2x2 point-to-point K of the friction and motor joints, inverted
// (b2Mat22::GetInverse, REF/Common/b2Math.h:226-238): m11, m12, m22.
__device__ inline float3 linearMass22(float mA, float iA, float mB, float iB, V2 rA, V2 rB) {
    const float a = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
    const float c = -iA * rA.x * rA.y - iB * rB.x * rB.y;
    const float bb = c;
    const float d = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
    float det = a * d - bb * c;
    if (det != 0.0f) det = 1.0f / det;
    return make_float3(det * d, -det * bb, det * a);
}

// b2FrictionJoint::InitVelocityConstraints and b2MotorJoint's
// (REF/Dynamics/Joints/b2FrictionJoint.cpp:59-132,
// REF/Dynamics/Joints/b2MotorJoint.cpp:63-141): the same rows, the motor
// anchored at the centres with an error against its offsets.
__device__ inline void frictionMotorInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio,
                                            int warm, bool motor) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    V2 rA, rB;
    if (motor) {
        rA = mul(qA, -mk(lcA.x, lcA.y));
        rB = mul(qB, -mk(lcB.x, lcB.y));
    } else {
        rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
        rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    }
    const float3 lm = linearMass22(mA, iA, mB, iB, rA, rB);
    float angularMass = iA + iB;
    if (angularMass > 0.0f) angularMass = 1.0f / angularMass;
    if (motor) {
        const V2 linearError = mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA - mul(qA, mk(an.x, an.y));
        jv.K1[j] = make_float2(linearError.x, linearError.y);
        jv.motorMass[j] = pB.z - pA.z - an.z;
    }
    float4 imp = jv.impulse[j];
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        imp.z *= dtRatio;
        const V2 P = mk(imp.x, imp.y);
        vA -= mA * P;
        wA -= iA * (cross(rA, P) + imp.z);
        vB += mB * P;
        wB += iB * (cross(rB, P) + imp.z);
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.impulse[j] = imp;
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(lm.x, lm.y, lm.z, angularMass);
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2FrictionJoint::SolveVelocityConstraints and b2MotorJoint's
// (REF b2FrictionJoint.cpp:134-189, REF b2MotorJoint.cpp:143-199): an
// angular impulse within h maxTorque, then a linear one within h maxForce.
__device__ inline void frictionMotorSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b, float h,
                                                     bool motor) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const float4 lim = jv.limits[j];  // maxForce, maxTorque
    float4 imp = jv.impulse[j];
    const float invH = 1.0f / h;
    const float cf = jv.anchors[j].w;
    {
        float Cdot = wB - wA;
        if (motor) Cdot = Cdot + invH * cf * jv.motorMass[j];
        float impulse = -k0.w * Cdot;
        const float oldImpulse = imp.z;
        const float maxImpulse = h * lim.y;
        imp.z = clampB(imp.z + impulse, -maxImpulse, maxImpulse);
        impulse = imp.z - oldImpulse;
        wA -= iA * impulse;
        wB += iB * impulse;
    }
    {
        V2 Cdot = vB + cross(wB, rB) - vA - cross(wA, rA);
        if (motor) {
            const float2 e = jv.K1[j];
            Cdot = Cdot + (invH * cf) * mk(e.x, e.y);
        }
        // -b2Mul(linearMass, Cdot)
        V2 impulse = -mk(k0.x * Cdot.x + k0.y * Cdot.y, k0.y * Cdot.x + k0.z * Cdot.y);
        const V2 oldImpulse = mk(imp.x, imp.y);
        V2 acc = oldImpulse + impulse;
        const float maxImpulse = h * lim.x;
        if (lengthSq(acc) > maxImpulse * maxImpulse) {
            normalizeV(acc);
            acc *= maxImpulse;
        }
        impulse = acc - oldImpulse;
        imp.x = acc.x;
        imp.y = acc.y;
        vA -= mA * impulse;
        wA -= iA * cross(rA, impulse);
        vB += mB * impulse;
        wB += iB * cross(rB, impulse);
    }
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// Unit vector from a ground anchor to a body anchor, zero when closer than
// 10 linearSlop, and the distance (b2PulleyJoint,
// REF/Dynamics/Joints/b2PulleyJoint.cpp:103-126).
__device__ inline float pulleyDir(V2& u) {
    const float len = lengthOf(u);
    if (len > 10.0f * kLinearSlop) u *= 1.0f / len;
    else u = mk(0.0f, 0.0f);
    return len;
}

// b2PulleyJoint::InitVelocityConstraints, REF b2PulleyJoint.cpp:76-164
__device__ inline void pulleyInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float imA = msA.x, iiA = msA.y, imB = msB.x, iiB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    const float4 g = jv.limits[j];  // ground anchors A, B
    const float ratio = jv.motorSpeed[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 uA = mk(pA.x, pA.y) + rA - mk(g.x, g.y);
    V2 uB = mk(pB.x, pB.y) + rB - mk(g.z, g.w);
    pulleyDir(uA);
    pulleyDir(uB);
    const float ruA = cross(rA, uA), ruB = cross(rB, uB);
    const float mA = imA + iiA * ruA * ruA;
    const float mB = imB + iiB * ruB * ruB;
    float mass = mA + ratio * ratio * mB;
    if (mass > 0.0f) mass = 1.0f / mass;
    float4 imp = jv.impulse[j];
    if (warm) {
        imp.x *= dtRatio;
        const V2 PA = -(imp.x) * uA;
        const V2 PB = (-ratio * imp.x) * uB;
        vA += imA * PA;
        wA += iiA * cross(rA, PA);
        vB += imB * PB;
        wB += iiB * cross(rB, PB);
    } else {
        imp.x = 0.0f;
    }
    jv.impulse[j] = make_float4(imp.x, uB.x, uB.y, imp.w);  // uB for b2PulleyJoint::GetReactionForce
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(imA, iiA, imB, iiB);
    jv.K0[j] = make_float4(uA.x, uA.y, uB.x, uB.y);
    jv.K1[j] = make_float2(mass, 0.0f);
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2PulleyJoint::SolveVelocityConstraints, REF b2PulleyJoint.cpp:166-191
__device__ inline void pulleySolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 ms = jv.masses[j];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const V2 uA = mk(k0.x, k0.y), uB = mk(k0.z, k0.w);
    const float ratio = jv.motorSpeed[j];
    const V2 vpA = vA + cross(wA, rA);
    const V2 vpB = vB + cross(wB, rB);
    const float Cdot = -dot(uA, vpA) - ratio * dot(uB, vpB);
    const float impulse = -jv.K1[j].x * Cdot;
    float4 imp = jv.impulse[j];
    imp.x += impulse;
    const V2 PA = -impulse * uA;
    const V2 PB = -ratio * impulse * uB;
    vA += ms.x * PA;
    wA += ms.y * cross(rA, PA);
    vB += ms.z * PB;
    wB += ms.w * cross(rB, PB);
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2PulleyJoint::SolvePositionConstraints, REF b2PulleyJoint.cpp:193-263:
// the constant is lengthA + ratio lengthB at creation.
__device__ inline bool pulleySolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const float4 g = jv.limits[j];
    const float ratio = jv.motorSpeed[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    V2 uA = cA + rA - mk(g.x, g.y);
    V2 uB = cB + rB - mk(g.z, g.w);
    const float lengthA = pulleyDir(uA);
    const float lengthB = pulleyDir(uB);
    const float ruA = cross(rA, uA), ruB = cross(rB, uB);
    const float mA = ms.x + ms.y * ruA * ruA;
    const float mB = ms.z + ms.w * ruB * ruB;
    float mass = mA + ratio * ratio * mB;
    if (mass > 0.0f) mass = 1.0f / mass;
    const float C = jv.impulse[j].w - lengthA - ratio * lengthB;
    const float linearError = absB(C);
    const float impulse = -mass * C;
    const V2 PA = -impulse * uA;
    const V2 PB = -ratio * impulse * uB;
    cA += ms.x * PA;
    aA += ms.y * cross(rA, PA);
    cB += ms.z * PB;
    aB += ms.w * cross(rB, PB);
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return linearError < kLinearSlop;
}

// b2MouseJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2MouseJoint.cpp:99-169:
// a soft point constraint pulling body B's anchor to the target; body A is
// not touched.
__device__ inline void mouseInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const float4 msB = w.mass[b.y];
    const float imB = msB.x, iiB = msB.y;
    const float2 lcB = w.lcenter[b.y];
    const float4 pB = w.pos[b.y];
    const float4 vbv = w.vel[b.y];
    V2 vB = mk(vbv.x, vbv.y);
    float wB = vbv.z;
    const Rot qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];   // target, localAnchorB
    const float4 par = jv.limits[j];   // maxForce, frequencyHz, dampingRatio, mass of body B
    const float mass = par.w;
    const float omega = 2.0f * kPi * par.y;
    const float d = 2.0f * mass * par.z * omega;
    const float k = mass * (omega * omega);
    const float h = jv.h;
    float gamma = h * (d + h * k);
    if (gamma != 0.0f) gamma = 1.0f / gamma;
    const float beta = h * k * gamma;
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const float a = imB + iiB * rB.y * rB.y + gamma;
    const float c = -iiB * rB.x * rB.y;
    const float bb = c;
    const float dd = imB + iiB * rB.x * rB.x + gamma;
    float det = a * dd - bb * c;
    if (det != 0.0f) det = 1.0f / det;
    V2 C = mk(pB.x, pB.y) + rB - mk(an.x, an.y);
    C *= beta;
    wB *= 0.98f;
    float4 imp = jv.impulse[j];
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        vB += imB * mk(imp.x, imp.y);
        wB += iiB * cross(rB, mk(imp.x, imp.y));
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.impulse[j] = imp;
    jv.r[j] = make_float4(0.0f, 0.0f, rB.x, rB.y);
    jv.masses[j] = make_float4(0.0f, 0.0f, imB, iiB);
    jv.K0[j] = make_float4(det * dd, -det * bb, det * a, gamma);
    jv.K1[j] = make_float2(C.x, C.y);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2MouseJoint::SolveVelocityConstraints, REF b2MouseJoint.cpp:171-194
__device__ inline void mouseSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b, float h) {
    const float4 ms = jv.masses[j];
    const float4 vbv = w.vel[b.y];
    V2 vB = mk(vbv.x, vbv.y);
    float wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rB = mk(r.z, r.w);
    const float4 k0 = jv.K0[j];
    const float2 C = jv.K1[j];
    float4 imp = jv.impulse[j];
    const V2 Cdot = vB + cross(wB, rB);
    const V2 rhs = -(Cdot + mk(C.x, C.y) + k0.w * mk(imp.x, imp.y));
    V2 impulse = mk(k0.x * rhs.x + k0.y * rhs.y, k0.y * rhs.x + k0.z * rhs.y);  // b2Mul(mass, rhs)
    const V2 oldImpulse = mk(imp.x, imp.y);
    V2 acc = oldImpulse + impulse;
    const float maxImpulse = h * jv.limits[j].x;
    if (lengthSq(acc) > maxImpulse * maxImpulse) acc *= maxImpulse / lengthOf(acc);
    impulse = acc - oldImpulse;
    imp.x = acc.x;
    imp.y = acc.y;
    vB += ms.z * impulse;
    wB += ms.w * cross(rB, impulse);
    jv.impulse[j] = imp;
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2Vec2::Normalize and b2Cross(1, v) for a joint's local axes, as the
// b2PrismaticJoint constructor sets them (REF/Dynamics/Joints/b2PrismaticJoint.cpp:107-109).
__device__ inline V2 perpOf(V2 x) { return mk(-1.0f * x.y, 1.0f * x.x); }

// b2PrismaticJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2PrismaticJoint.cpp:128-266
__device__ inline void prismaticInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio,
                                        int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    const float4 ax4 = jv.axis[j];
    const V2 lx = mk(ax4.x, ax4.y);
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const V2 d = (mk(pB.x, pB.y) - mk(pA.x, pA.y)) + rB - rA;
    const V2 axis = mul(qA, lx);
    const float a1 = cross(d + rA, axis), a2 = cross(rB, axis);
    float motorMass = mA + mB + iA * a1 * a1 + iB * a2 * a2;
    if (motorMass > 0.0f) motorMass = 1.0f / motorMass;
    const V2 perp = mul(qA, perpOf(lx));
    const float s1 = cross(d + rA, perp), s2 = cross(rB, perp);
    Sym33 K;
    K.m11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
    K.m12 = iA * s1 + iB * s2;
    K.m13 = iA * s1 * a1 + iB * s2 * a2;
    K.m22 = iA + iB;
    if (K.m22 == 0.0f) K.m22 = 1.0f;
    K.m23 = iA * a1 + iB * a2;
    K.m33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
    int flags = jv.flags[j];
    float4 imp = jv.impulse[j];
    int state = (flags >> 8) & 3;
    const float4 lim = jv.limits[j];  // referenceAngle, lower, upper, maxMotorForce
    if (flags & JF_LIMIT) {
        const float jointTranslation = dot(axis, d);
        if (absB(lim.z - lim.y) < 2.0f * kLinearSlop) {
            state = JL_EQUAL;
        } else if (jointTranslation <= lim.y) {
            if (state != JL_LOWER) {
                state = JL_LOWER;
                imp.z = 0.0f;
            }
        } else if (jointTranslation >= lim.z) {
            if (state != JL_UPPER) {
                state = JL_UPPER;
                imp.z = 0.0f;
            }
        } else {
            state = JL_INACTIVE;
            imp.z = 0.0f;
        }
    } else {
        state = JL_INACTIVE;
        imp.z = 0.0f;
    }
    if (!(flags & JF_MOTOR)) imp.w = 0.0f;
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        imp.z *= dtRatio;
        imp.w *= dtRatio;
        const V2 P = imp.x * perp + (imp.w + imp.z) * axis;
        const float LA = imp.x * s1 + imp.y + (imp.w + imp.z) * a1;
        const float LB = imp.x * s2 + imp.y + (imp.w + imp.z) * a2;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.flags[j] = (flags & ~0x300) | (state << 8);
    jv.impulse[j] = imp;
    jv.extra[2 * j] = make_float4(axis.x, axis.y, perp.x, perp.y);
    jv.r[j] = make_float4(s1, s2, a1, a2);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(K.m11, K.m12, K.m13, K.m22);
    jv.K1[j] = make_float2(K.m23, K.m33);
    jv.motorMass[j] = motorMass;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2PrismaticJoint::SolveVelocityConstraints, REF b2PrismaticJoint.cpp:268-363
__device__ inline void prismaticSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b, float h) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 e = jv.extra[2 * j];
    const V2 axis = mk(e.x, e.y), perp = mk(e.z, e.w);
    const float4 sa = jv.r[j];
    const float s1 = sa.x, s2 = sa.y, a1 = sa.z, a2 = sa.w;
    const int flags = jv.flags[j];
    const int state = (flags >> 8) & 3;
    float4 imp = jv.impulse[j];
    if ((flags & JF_MOTOR) && state != JL_EQUAL) {
        const float Cdot = dot(axis, vB - vA) + a2 * wB - a1 * wA;
        float impulse = jv.motorMass[j] * (jv.motorSpeed[j] - Cdot);
        const float oldImpulse = imp.w;
        const float maxImpulse = h * jv.limits[j].w;
        imp.w = clampB(imp.w + impulse, -maxImpulse, maxImpulse);
        impulse = imp.w - oldImpulse;
        const V2 P = impulse * axis;
        const float LA = impulse * a1, LB = impulse * a2;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    }
    const float c1x = dot(perp, vB - vA) + s2 * wB - s1 * wA;
    const float c1y = wB - wA;
    const float4 k0 = jv.K0[j];
    const float2 k1 = jv.K1[j];
    const Sym33 K{k0.x, k0.y, k0.z, k0.w, k1.x, k1.y};
    if ((flags & JF_LIMIT) && state != JL_INACTIVE) {
        const float Cdot2 = dot(axis, vB - vA) + a2 * wB - a1 * wA;
        const float4 f1 = imp;
        float dx, dy, dz;
        solve33(K, -c1x, -c1y, -Cdot2, dx, dy, dz);
        imp.x += dx;
        imp.y += dy;
        imp.z += dz;
        if (state == JL_LOWER) imp.z = imp.z > 0.0f ? imp.z : 0.0f;       // b2Max
        else if (state == JL_UPPER) imp.z = imp.z < 0.0f ? imp.z : 0.0f;  // b2Min
        const float sz = imp.z - f1.z;
        const V2 rhs = mk(-c1x - sz * K.m13, -c1y - sz * K.m23);
        const V2 f2r = solve22(K, rhs) + mk(f1.x, f1.y);
        imp.x = f2r.x;
        imp.y = f2r.y;
        const float ddx = imp.x - f1.x, ddy = imp.y - f1.y, ddz = imp.z - f1.z;
        const V2 P = ddx * perp + ddz * axis;
        const float LA = ddx * s1 + ddy + ddz * a1;
        const float LB = ddx * s2 + ddy + ddz * a2;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    } else {
        const V2 df = solve22(K, mk(-c1x, -c1y));
        imp.x += df.x;
        imp.y += df.y;
        const V2 P = df.x * perp;
        const float LA = df.x * s1 + df.y;
        const float LB = df.x * s2 + df.y;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    }
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2PrismaticJoint::SolvePositionConstraints, REF b2PrismaticJoint.cpp:365-488
__device__ inline bool prismaticSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const float4 ax4 = jv.axis[j];
    const V2 lx = mk(ax4.x, ax4.y);
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const V2 d = cB + rB - cA - rA;
    const V2 axis = mul(qA, lx);
    const float a1 = cross(d + rA, axis), a2 = cross(rB, axis);
    const V2 perp = mul(qA, perpOf(lx));
    const float s1 = cross(d + rA, perp), s2 = cross(rB, perp);
    const float4 lim = jv.limits[j];
    const float C1x = dot(perp, d), C1y = aB - aA - lim.x;
    float linearError = absB(C1x);
    const float angularError = absB(C1y);
    bool active = false;
    float C2 = 0.0f;
    if (jv.flags[j] & JF_LIMIT) {
        const float translation = dot(axis, d);
        if (absB(lim.z - lim.y) < 2.0f * kLinearSlop) {
            C2 = clampB(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
            const float e = absB(translation);
            linearError = linearError > e ? linearError : e;
            active = true;
        } else if (translation <= lim.y) {
            C2 = clampB(translation - lim.y + kLinearSlop, -kMaxLinearCorrection, 0.0f);
            const float e = lim.y - translation;
            linearError = linearError > e ? linearError : e;
            active = true;
        } else if (translation >= lim.z) {
            C2 = clampB(translation - lim.z - kLinearSlop, 0.0f, kMaxLinearCorrection);
            const float e = translation - lim.z;
            linearError = linearError > e ? linearError : e;
            active = true;
        }
    }
    Sym33 K;
    K.m11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
    K.m12 = iA * s1 + iB * s2;
    K.m22 = iA + iB;
    if (K.m22 == 0.0f) K.m22 = 1.0f;
    float ix, iy, iz;
    if (active) {
        K.m13 = iA * s1 * a1 + iB * s2 * a2;
        K.m23 = iA * a1 + iB * a2;
        K.m33 = mA + mB + iA * a1 * a1 + iB * a2 * a2;
        solve33(K, -C1x, -C1y, -C2, ix, iy, iz);
    } else {
        K.m13 = K.m23 = K.m33 = 0.0f;
        const V2 i1 = solve22(K, mk(-C1x, -C1y));  // b2Mat22::Solve
        ix = i1.x;
        iy = i1.y;
        iz = 0.0f;
    }
    const V2 P = ix * perp + iz * axis;
    const float LA = ix * s1 + iy + iz * a1;
    const float LB = ix * s2 + iy + iz * a2;
    cA -= mA * P;
    aA -= iA * LA;
    cB += mB * P;
    aB += iB * LB;
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return linearError <= kLinearSlop && angularError <= kAngularSlop;
}

// b2WheelJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2WheelJoint.cpp:79-216:
// a point-on-line row along the perpendicular axis, a soft spring along the
// axis and a motor on the relative rotation.
__device__ inline void wheelInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(pA.z), qB = rotOf(pB.z);
    const float4 an = jv.anchors[j];
    const float4 ax4 = jv.axis[j];
    const V2 lx = mk(ax4.x, ax4.y);
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const V2 d = mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA;
    const V2 ay = mul(qA, perpOf(lx));
    const float sAy = cross(d + rA, ay), sBy = cross(rB, ay);
    float mass = mA + mB + iA * sAy * sAy + iB * sBy * sBy;
    if (mass > 0.0f) mass = 1.0f / mass;
    const float4 par = jv.limits[j];  // frequencyHz, dampingRatio, maxMotorTorque
    float4 ex = jv.extra[2 * j];     // ax, sAx, sBx: kept while the spring is off
    float springMass = 0.0f, bias = 0.0f, gamma = 0.0f;
    float4 imp = jv.impulse[j];      // impulse, springImpulse, motorImpulse
    if (par.x > 0.0f) {
        const V2 ax = mul(qA, lx);
        const float sAx = cross(d + rA, ax), sBx = cross(rB, ax);
        ex = make_float4(ax.x, ax.y, sAx, sBx);
        const float invMass = mA + mB + iA * sAx * sAx + iB * sBx * sBx;
        if (invMass > 0.0f) {
            springMass = 1.0f / invMass;
            const float C = dot(d, ax);
            const float omega = 2.0f * kPi * par.x;
            const float dd = 2.0f * springMass * par.y * omega;
            const float k = springMass * omega * omega;
            const float h = jv.h;
            gamma = h * (dd + h * k);
            if (gamma > 0.0f) gamma = 1.0f / gamma;
            bias = C * h * k * gamma;
            springMass = invMass + gamma;
            if (springMass > 0.0f) springMass = 1.0f / springMass;
        }
    } else {
        imp.y = 0.0f;
    }
    float motorMass = 0.0f;
    if (jv.flags[j] & JF_MOTOR) {
        motorMass = iA + iB;
        if (motorMass > 0.0f) motorMass = 1.0f / motorMass;
    } else {
        imp.z = 0.0f;
    }
    const V2 ax = mk(ex.x, ex.y);
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        imp.z *= dtRatio;
        const V2 P = imp.x * ay + imp.y * ax;
        const float LA = imp.x * sAy + imp.y * ex.z + imp.z;
        const float LB = imp.x * sBy + imp.y * ex.w + imp.z;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.impulse[j] = imp;
    jv.extra[2 * j] = ex;
    jv.extra[2 * j + 1] = make_float4(ay.x, ay.y, sAy, sBy);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(mass, springMass, bias, gamma);
    jv.motorMass[j] = motorMass;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2WheelJoint::SolveVelocityConstraints, REF b2WheelJoint.cpp:218-280
__device__ inline void wheelSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b, float h) {
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 ex = jv.extra[2 * j], ey = jv.extra[2 * j + 1];
    const V2 ax = mk(ex.x, ex.y), ay = mk(ey.x, ey.y);
    const float4 k0 = jv.K0[j];  // mass, springMass, bias, gamma
    float4 imp = jv.impulse[j];
    {
        const float Cdot = dot(ax, vB - vA) + ex.w * wB - ex.z * wA;
        const float impulse = -k0.y * (Cdot + k0.z + k0.w * imp.y);
        imp.y += impulse;
        const V2 P = impulse * ax;
        const float LA = impulse * ex.z, LB = impulse * ex.w;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    }
    {
        const float Cdot = wB - wA - jv.motorSpeed[j];
        float impulse = -jv.motorMass[j] * Cdot;
        const float oldImpulse = imp.z;
        const float maxImpulse = h * jv.limits[j].z;
        imp.z = clampB(imp.z + impulse, -maxImpulse, maxImpulse);
        impulse = imp.z - oldImpulse;
        wA -= iA * impulse;
        wB += iB * impulse;
    }
    {
        const float Cdot = dot(ay, vB - vA) + ey.w * wB - ey.z * wA;
        const float impulse = -k0.x * Cdot;
        imp.x += impulse;
        const V2 P = impulse * ay;
        const float LA = impulse * ey.z, LB = impulse * ey.w;
        vA -= mA * P;
        wA -= iA * LA;
        vB += mB * P;
        wB += iB * LB;
    }
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

// b2WheelJoint::SolvePositionConstraints, REF b2WheelJoint.cpp:282-327; its
// effective mass uses the sAy, sBy of the velocity pass, as the reference's.
__device__ inline bool wheelSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const float4 ax4 = jv.axis[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const V2 d = (cB - cA) + rB - rA;
    const V2 ay = mul(qA, perpOf(mk(ax4.x, ax4.y)));
    const float sAy = cross(d + rA, ay), sBy = cross(rB, ay);
    const float C = dot(d, ay);
    const float4 ey = jv.extra[2 * j + 1];
    const float k = ms.x + ms.z + ms.y * ey.z * ey.z + ms.w * ey.w * ey.w;
    const float impulse = k != 0.0f ? -C / k : 0.0f;
    const V2 P = impulse * ay;
    const float LA = impulse * sAy, LB = impulse * sBy;
    cA -= ms.x * P;
    aA -= ms.y * LA;
    cB += ms.z * P;
    aB += ms.w * LB;
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return absB(C) <= kLinearSlop;
}

// b2Joint::GetReactionForce and GetReactionTorque at inv_dt = 1 for every
// joint type, from the last step's impulses and directions
// (REF/Dynamics/Joints/b2*Joint.cpp, GetReactionForce / GetReactionTorque).
__global__ void k_joint_reactions(JointView jv, int first, int count, float4* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int j = first + k;
    const int flags = jv.flags[j];
    const float4 imp = jv.impulse[j];
    float fx = 0.0f, fy = 0.0f, t = 0.0f;
    if (!(flags & JF_REMOVED)) {
        switch (jointType(flags)) {
            case JT_REVOLUTE: fx = imp.x; fy = imp.y; t = imp.z; break;
            case JT_DISTANCE: case JT_ROPE: case JT_PULLEY: {
                const V2 P = imp.x * mk(imp.y, imp.z);
                fx = P.x; fy = P.y;
                break;
            }
            case JT_WELD: case JT_FRICTION: case JT_MOTOR: fx = imp.x; fy = imp.y; t = imp.z; break;
            case JT_MOUSE: fx = imp.x; fy = imp.y; t = 1.0f * 0.0f; break;
            case JT_PRISMATIC: {
                const float4 e = jv.extra[2 * j];
                const V2 P = imp.x * mk(e.z, e.w) + (imp.w + imp.z) * mk(e.x, e.y);
                fx = P.x; fy = P.y; t = imp.y;
                break;
            }
            case JT_GEAR: {
                const V2 P = imp.x * mk(imp.y, imp.z);
                fx = P.x; fy = P.y; t = imp.x * imp.w;
                break;
            }
            case JT_WHEEL: {
                const float4 ex = jv.extra[2 * j], ey = jv.extra[2 * j + 1];
                const V2 P = imp.x * mk(ey.x, ey.y) + imp.y * mk(ex.x, ex.y);
                fx = P.x; fy = P.y; t = imp.z;
                break;
            }
            default: break;
        }
    }
    out[k] = make_float4(fx, fy, t, 0.0f);
}

// Gear joint (type 10): bodies A, B in the joint row, C, D (the other bodies
// of its two joints) as integer bits in axis.x / axis.y, the two joints'
// kinds in flags (JF_GEAR_PRISM_A / _B set for prismatic, else revolute);
// anchors = localAnchorA, B; extra[2j] = localAnchorC, D; extra[2j+1] =
// localAxisC, D; limits = referenceAngleA, referenceAngleB, ratio, constant.
// Per step: r = (JvAC, JvBD), K0 = (JwA, JwB, JwC, JwD), motorMass = mass;
// impulse = (impulse, JvAC.x, JvAC.y, JwA) for the reaction.
struct GearRows {
    V2 JvAC, JvBD;
    float JwA, JwB, JwC, JwD, mass, coordinateA, coordinateB;
};

// The Jacobian rows, effective mass and coordinates of a gear joint at the
// given poses (b2GearJoint::InitVelocityConstraints and
// SolvePositionConstraints, REF/Dynamics/Joints/b2GearJoint.cpp:147-393).
__device__ inline GearRows gearRows(const WorldView& w, const JointView& jv, int j, int4 q, const float4* p,
                                    bool coords) {
    const float4 an = jv.anchors[j], lcd = jv.extra[2 * j], axes = jv.extra[2 * j + 1], par = jv.limits[j];
    const int flags = jv.flags[j];
    const float2 lcA = w.lcenter[q.x], lcB = w.lcenter[q.y], lcC = w.lcenter[q.z], lcDv = w.lcenter[q.w];
    const float4 mA = w.mass[q.x], mB = w.mass[q.y], mC = w.mass[q.z], mD = w.mass[q.w];
    const Rot qA = rotOf(p[0].z), qB = rotOf(p[1].z), qC = rotOf(p[2].z), qD = rotOf(p[3].z);
    const float ratio = par.z;
    GearRows g;
    g.mass = 0.0f;
    g.coordinateA = g.coordinateB = 0.0f;
    if (!(flags & JF_GEAR_PRISM_A)) {
        g.JvAC = mk(0.0f, 0.0f);
        g.JwA = 1.0f;
        g.JwC = 1.0f;
        g.mass += mA.y + mC.y;
        if (coords) g.coordinateA = p[0].z - p[2].z - par.x;
    } else {
        const V2 axisC = mk(axes.x, axes.y);
        const V2 u = mul(qC, axisC);
        const V2 rC = mul(qC, mk(lcd.x, lcd.y) - mk(lcC.x, lcC.y));
        const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
        g.JvAC = u;
        g.JwC = cross(rC, u);
        g.JwA = cross(rA, u);
        g.mass += mC.x + mA.x + mC.y * g.JwC * g.JwC + mA.y * g.JwA * g.JwA;
        if (coords) {
            const V2 pC = mk(lcd.x, lcd.y) - mk(lcC.x, lcC.y);
            const V2 pA = mulT(qC, rA + (mk(p[0].x, p[0].y) - mk(p[2].x, p[2].y)));
            g.coordinateA = dot(pA - pC, axisC);
        }
    }
    if (!(flags & JF_GEAR_PRISM_B)) {
        g.JvBD = mk(0.0f, 0.0f);
        g.JwB = ratio;
        g.JwD = ratio;
        g.mass += ratio * ratio * (mB.y + mD.y);
        if (coords) g.coordinateB = p[1].z - p[3].z - par.y;
    } else {
        const V2 axisD = mk(axes.z, axes.w);
        const V2 u = mul(qD, axisD);
        const V2 rD = mul(qD, mk(lcd.z, lcd.w) - mk(lcDv.x, lcDv.y));
        const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
        g.JvBD = ratio * u;
        g.JwD = ratio * cross(rD, u);
        g.JwB = ratio * cross(rB, u);
        g.mass += ratio * ratio * (mD.x + mB.x) + mD.y * g.JwD * g.JwD + mB.y * g.JwB * g.JwB;
        if (coords) {
            const V2 pD = mk(lcd.z, lcd.w) - mk(lcDv.x, lcDv.y);
            const V2 pB = mulT(qD, rB + (mk(p[1].x, p[1].y) - mk(p[3].x, p[3].y)));
            g.coordinateB = dot(pB - pD, axisD);
        }
    }
    return g;
}

__device__ inline int4 gearBodies(const JointView& jv, int j, int2 b) {
    const float4 ax = jv.axis[j];
    return make_int4(b.x, b.y, __float_as_int(ax.x), __float_as_int(ax.y));
}

// Writes the four bodies' velocities in the reference's order A, B, C, D
// (a body met twice keeps the last write, as b2GearJoint's does).
__device__ inline void gearStoreVel(const WorldView& w, int4 q, const float4* v) {
    const int ids[4] = {q.x, q.y, q.z, q.w};
    for (int k = 0; k < 4; ++k)
        if (w.btype[ids[k]] == BT_DYNAMIC) w.vel[ids[k]] = v[k];
}

// b2GearJoint::InitVelocityConstraints, REF b2GearJoint.cpp:147-240
__device__ inline void gearInitOne(const WorldView& w, const JointView& jv, int j, int2 b, float dtRatio, int warm) {
    const int4 q = gearBodies(jv, j, b);
    const float4 p[4] = {w.pos[q.x], w.pos[q.y], w.pos[q.z], w.pos[q.w]};
    const GearRows g = gearRows(w, jv, j, q, p, false);
    const float mass = g.mass > 0.0f ? 1.0f / g.mass : 0.0f;
    float4 v[4] = {w.vel[q.x], w.vel[q.y], w.vel[q.z], w.vel[q.w]};
    float4 imp = jv.impulse[j];
    if (warm) {
        const float4 mA = w.mass[q.x], mB = w.mass[q.y], mC = w.mass[q.z], mD = w.mass[q.w];
        const float im = imp.x;
        V2 vA = mk(v[0].x, v[0].y), vB = mk(v[1].x, v[1].y), vC = mk(v[2].x, v[2].y), vD = mk(v[3].x, v[3].y);
        vA += (mA.x * im) * g.JvAC;
        v[0].z += mA.y * im * g.JwA;
        vB += (mB.x * im) * g.JvBD;
        v[1].z += mB.y * im * g.JwB;
        vC -= (mC.x * im) * g.JvAC;
        v[2].z -= mC.y * im * g.JwC;
        vD -= (mD.x * im) * g.JvBD;
        v[3].z -= mD.y * im * g.JwD;
        v[0].x = vA.x; v[0].y = vA.y; v[1].x = vB.x; v[1].y = vB.y;
        v[2].x = vC.x; v[2].y = vC.y; v[3].x = vD.x; v[3].y = vD.y;
    } else {
        imp.x = 0.0f;
    }
    jv.impulse[j] = make_float4(imp.x, g.JvAC.x, g.JvAC.y, g.JwA);
    jv.r[j] = make_float4(g.JvAC.x, g.JvAC.y, g.JvBD.x, g.JvBD.y);
    jv.K0[j] = make_float4(g.JwA, g.JwB, g.JwC, g.JwD);
    jv.motorMass[j] = mass;
    for (int k = 0; k < 4; ++k) v[k].w = 0.0f;
    gearStoreVel(w, q, v);
}

// b2GearJoint::SolveVelocityConstraints, REF b2GearJoint.cpp:242-274
__device__ inline void gearSolveVelocityOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const int4 q = gearBodies(jv, j, b);
    float4 v[4] = {w.vel[q.x], w.vel[q.y], w.vel[q.z], w.vel[q.w]};
    const float4 r = jv.r[j], Jw = jv.K0[j];
    const V2 JvAC = mk(r.x, r.y), JvBD = mk(r.z, r.w);
    V2 vA = mk(v[0].x, v[0].y), vB = mk(v[1].x, v[1].y), vC = mk(v[2].x, v[2].y), vD = mk(v[3].x, v[3].y);
    float Cdot = dot(JvAC, vA - vC) + dot(JvBD, vB - vD);
    Cdot += (Jw.x * v[0].z - Jw.z * v[2].z) + (Jw.y * v[1].z - Jw.w * v[3].z);
    const float impulse = -jv.motorMass[j] * Cdot;
    float4 imp = jv.impulse[j];
    imp.x += impulse;
    const float4 mA = w.mass[q.x], mB = w.mass[q.y], mC = w.mass[q.z], mD = w.mass[q.w];
    vA += (mA.x * impulse) * JvAC;
    v[0].z += mA.y * impulse * Jw.x;
    vB += (mB.x * impulse) * JvBD;
    v[1].z += mB.y * impulse * Jw.y;
    vC -= (mC.x * impulse) * JvAC;
    v[2].z -= mC.y * impulse * Jw.z;
    vD -= (mD.x * impulse) * JvBD;
    v[3].z -= mD.y * impulse * Jw.w;
    v[0].x = vA.x; v[0].y = vA.y; v[1].x = vB.x; v[1].y = vB.y;
    v[2].x = vC.x; v[2].y = vC.y; v[3].x = vD.x; v[3].y = vD.y;
    jv.impulse[j] = imp;
    gearStoreVel(w, q, v);
}

// b2GearJoint::SolvePositionConstraints, REF b2GearJoint.cpp:276-393 (its
// linear error stays 0, so it never holds back the island's early exit).
__device__ inline bool gearSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const int4 q = gearBodies(jv, j, b);
    float4 p[4] = {w.pos[q.x], w.pos[q.y], w.pos[q.z], w.pos[q.w]};
    const GearRows g = gearRows(w, jv, j, q, p, true);
    const float4 par = jv.limits[j];
    const float C = (g.coordinateA + par.z * g.coordinateB) - par.w;
    float impulse = 0.0f;
    if (g.mass > 0.0f) impulse = -C / g.mass;
    const float4 mA = w.mass[q.x], mB = w.mass[q.y], mC = w.mass[q.z], mD = w.mass[q.w];
    V2 cA = mk(p[0].x, p[0].y), cB = mk(p[1].x, p[1].y), cC = mk(p[2].x, p[2].y), cD = mk(p[3].x, p[3].y);
    cA += mA.x * impulse * g.JvAC;
    p[0].z += mA.y * impulse * g.JwA;
    cB += mB.x * impulse * g.JvBD;
    p[1].z += mB.y * impulse * g.JwB;
    cC -= mC.x * impulse * g.JvAC;
    p[2].z -= mC.y * impulse * g.JwC;
    cD -= mD.x * impulse * g.JvBD;
    p[3].z -= mD.y * impulse * g.JwD;
    p[0].x = cA.x; p[0].y = cA.y; p[1].x = cB.x; p[1].y = cB.y;
    p[2].x = cC.x; p[2].y = cC.y; p[3].x = cD.x; p[3].y = cD.y;
    const int ids[4] = {q.x, q.y, q.z, q.w};
    for (int k = 0; k < 4; ++k)
        if (w.btype[ids[k]] == BT_DYNAMIC) w.pos[ids[k]] = p[k];
    return true;
}

// b2RevoluteJoint::InitVelocityConstraints, REF/Dynamics/Joints/b2RevoluteJoint.cpp:66-198
__device__ inline void jointInitOne(const WorldView& w, const JointView& jv, int j, float dtRatio, int warm) {
    const int2 b = jv.bodies[j];
    if (!jointActive(w, b)) return;
    const int type = jointType(jv.flags[j]);
    if (type == JT_DISTANCE) {
        distanceInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_ROPE) {
        ropeInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_WELD) {
        weldInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_FRICTION || type == JT_MOTOR) {
        frictionMotorInitOne(w, jv, j, b, dtRatio, warm, type == JT_MOTOR);
        return;
    }
    if (type == JT_PULLEY) {
        pulleyInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_MOUSE) {
        mouseInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_PRISMATIC) {
        prismaticInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_GEAR) {
        gearInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    if (type == JT_WHEEL) {
        wheelInitOne(w, jv, j, b, dtRatio, warm);
        return;
    }
    const float4 msA = w.mass[b.x], msB = w.mass[b.y];
    const float mA = msA.x, iA = msA.y, mB = msB.x, iB = msB.y;
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float aA = w.pos[b.x].z, aB = w.pos[b.y].z;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const Rot qA = rotOf(aA), qB = rotOf(aB);
    const float4 an = jv.anchors[j];
    const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const bool fixedRotation = (iA + iB == 0.0f);
    Sym33 K;
    K.m11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
    K.m12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    K.m13 = -rA.y * iA - rB.y * iB;
    K.m22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
    K.m23 = rA.x * iA + rB.x * iB;
    K.m33 = iA + iB;
    float motorMass = iA + iB;
    if (motorMass > 0.0f) motorMass = 1.0f / motorMass;
    int flags = jv.flags[j];
    float4 imp = jv.impulse[j];
    if (!(flags & JF_MOTOR) || fixedRotation) imp.w = 0.0f;
    int state = (flags >> 8) & 3;
    const float4 lim = jv.limits[j];
    if ((flags & JF_LIMIT) && !fixedRotation) {
        const float jointAngle = aB - aA - lim.x;
        if (absB(lim.z - lim.y) < 2.0f * kAngularSlop) {
            state = JL_EQUAL;
        } else if (jointAngle <= lim.y) {
            if (state != JL_LOWER) imp.z = 0.0f;
            state = JL_LOWER;
        } else if (jointAngle >= lim.z) {
            if (state != JL_UPPER) imp.z = 0.0f;
            state = JL_UPPER;
        } else {
            state = JL_INACTIVE;
            imp.z = 0.0f;
        }
    } else {
        state = JL_INACTIVE;
    }
    if (warm) {
        imp.x *= dtRatio;
        imp.y *= dtRatio;
        imp.z *= dtRatio;
        imp.w *= dtRatio;
        const V2 P = mk(imp.x, imp.y);
        vA -= mA * P;
        wA -= iA * (cross(rA, P) + imp.w + imp.z);
        vB += mB * P;
        wB += iB * (cross(rB, P) + imp.w + imp.z);
    } else {
        imp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    jv.flags[j] = (flags & ~0x300) | (state << 8);
    jv.impulse[j] = imp;
    jv.r[j] = make_float4(rA.x, rA.y, rB.x, rB.y);
    jv.masses[j] = make_float4(mA, iA, mB, iB);
    jv.K0[j] = make_float4(K.m11, K.m12, K.m13, K.m22);
    jv.K1[j] = make_float2(K.m23, K.m33);
    jv.motorMass[j] = motorMass;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

__global__ void k_joint_init(WorldView w, JointView jv, int begin, int end, float dtRatio, int warm) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    jointInitOne(w, jv, jv.order[k], dtRatio, warm);
}

__global__ void k_joint_init_serial(WorldView w, JointView jv, int begin, int end, float dtRatio, int warm) {
    for (int k = begin; k < end; ++k) jointInitOne(w, jv, jv.order[k], dtRatio, warm);
}

// b2RevoluteJoint::SolveVelocityConstraints, REF b2RevoluteJoint.cpp:200-292
__device__ inline void jointSolveVelocityOne(const WorldView& w, const JointView& jv, int j, float h) {
    const int2 b = jv.bodies[j];
    if (!jointActive(w, b)) return;
    const int type = jointType(jv.flags[j]);
    if (type == JT_DISTANCE) {
        distanceSolveVelocityOne(w, jv, j, b);
        return;
    }
    if (type == JT_ROPE) {
        ropeSolveVelocityOne(w, jv, j, b, h);
        return;
    }
    if (type == JT_WELD) {
        weldSolveVelocityOne(w, jv, j, b);
        return;
    }
    if (type == JT_FRICTION || type == JT_MOTOR) {
        frictionMotorSolveVelocityOne(w, jv, j, b, h, type == JT_MOTOR);
        return;
    }
    if (type == JT_PULLEY) {
        pulleySolveVelocityOne(w, jv, j, b);
        return;
    }
    if (type == JT_MOUSE) {
        mouseSolveVelocityOne(w, jv, j, b, h);
        return;
    }
    if (type == JT_PRISMATIC) {
        prismaticSolveVelocityOne(w, jv, j, b, h);
        return;
    }
    if (type == JT_GEAR) {
        gearSolveVelocityOne(w, jv, j, b);
        return;
    }
    if (type == JT_WHEEL) {
        wheelSolveVelocityOne(w, jv, j, b, h);
        return;
    }
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float4 va = w.vel[b.x], vbv = w.vel[b.y];
    V2 vA = mk(va.x, va.y), vB = mk(vbv.x, vbv.y);
    float wA = va.z, wB = vbv.z;
    const float4 r = jv.r[j];
    const V2 rA = mk(r.x, r.y), rB = mk(r.z, r.w);
    const int flags = jv.flags[j];
    const int state = (flags >> 8) & 3;
    float4 imp = jv.impulse[j];
    const bool fixedRotation = (iA + iB == 0.0f);
    if ((flags & JF_MOTOR) && state != JL_EQUAL && !fixedRotation) {
        const float Cdot = wB - wA - jv.motorSpeed[j];
        float impulse = -jv.motorMass[j] * Cdot;
        const float oldImpulse = imp.w;
        const float maxImpulse = h * jv.limits[j].w;
        imp.w = clampB(imp.w + impulse, -maxImpulse, maxImpulse);
        impulse = imp.w - oldImpulse;
        wA -= iA * impulse;
        wB += iB * impulse;
    }
    const float4 k0 = jv.K0[j];
    const float2 k1 = jv.K1[j];
    const Sym33 K{k0.x, k0.y, k0.z, k0.w, k1.x, k1.y};
    if ((flags & JF_LIMIT) && state != JL_INACTIVE && !fixedRotation) {
        const V2 Cdot1 = vB + cross(wB, rB) - vA - cross(wA, rA);
        const float Cdot2 = wB - wA;
        float ix, iy, iz;
        solve33(K, Cdot1.x, Cdot1.y, Cdot2, ix, iy, iz);
        ix = -ix;
        iy = -iy;
        iz = -iz;
        if (state == JL_EQUAL) {
            imp.x += ix;
            imp.y += iy;
            imp.z += iz;
        } else {
            const float newImpulse = imp.z + iz;
            const bool reduce = state == JL_LOWER ? newImpulse < 0.0f : newImpulse > 0.0f;
            if (reduce) {
                const V2 rhs = -Cdot1 + imp.z * mk(K.m13, K.m23);
                const V2 reduced = solve22(K, rhs);
                ix = reduced.x;
                iy = reduced.y;
                iz = -imp.z;
                imp.x += reduced.x;
                imp.y += reduced.y;
                imp.z = 0.0f;
            } else {
                imp.x += ix;
                imp.y += iy;
                imp.z += iz;
            }
        }
        const V2 P = mk(ix, iy);
        vA -= mA * P;
        wA -= iA * (cross(rA, P) + iz);
        vB += mB * P;
        wB += iB * (cross(rB, P) + iz);
    } else {
        const V2 Cdot = vB + cross(wB, rB) - vA - cross(wA, rA);
        const V2 impulse = solve22(K, -Cdot);
        imp.x += impulse.x;
        imp.y += impulse.y;
        vA -= mA * impulse;
        wA -= iA * cross(rA, impulse);
        vB += mB * impulse;
        wB += iB * cross(rB, impulse);
    }
    jv.impulse[j] = imp;
    if (w.btype[b.x] == BT_DYNAMIC) w.vel[b.x] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (w.btype[b.y] == BT_DYNAMIC) w.vel[b.y] = make_float4(vB.x, vB.y, wB, 0.0f);
}

__global__ void k_joint_solve_velocity(WorldView w, JointView jv, int begin, int end, float h) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    jointSolveVelocityOne(w, jv, jv.order[k], h);
}

__global__ void k_joint_solve_velocity_serial(WorldView w, JointView jv, int begin, int end, float h) {
    for (int k = begin; k < end; ++k) jointSolveVelocityOne(w, jv, jv.order[k], h);
}

// b2RevoluteJoint::SolvePositionConstraints, REF b2RevoluteJoint.cpp:294-378:
// true when the errors are within the slops.
__device__ inline bool revoluteSolvePositionOne(const WorldView& w, const JointView& jv, int j, int2 b) {
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    const float4 ms = jv.masses[j];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const int flags = jv.flags[j];
    const int state = (flags >> 8) & 3;
    const float4 lim = jv.limits[j];
    float angularError = 0.0f, positionError = 0.0f;
    const bool fixedRotation = (iA + iB == 0.0f);
    if ((flags & JF_LIMIT) && state != JL_INACTIVE && !fixedRotation) {
        const float angle = aB - aA - lim.x;
        float limitImpulse = 0.0f;
        const float motorMass = jv.motorMass[j];
        if (state == JL_EQUAL) {
            const float C = clampB(angle - lim.y, -kMaxAngularCorrection, kMaxAngularCorrection);
            limitImpulse = -motorMass * C;
            angularError = absB(C);
        } else if (state == JL_LOWER) {
            float C = angle - lim.y;
            angularError = -C;
            C = clampB(C + kAngularSlop, -kMaxAngularCorrection, 0.0f);
            limitImpulse = -motorMass * C;
        } else if (state == JL_UPPER) {
            float C = angle - lim.z;
            angularError = C;
            C = clampB(C - kAngularSlop, 0.0f, kMaxAngularCorrection);
            limitImpulse = -motorMass * C;
        }
        aA -= iA * limitImpulse;
        aB += iB * limitImpulse;
    }
    {
        const Rot qA = rotOf(aA), qB = rotOf(aB);
        const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
        const float4 an = jv.anchors[j];
        const V2 rA = mul(qA, mk(an.x, an.y) - mk(lcA.x, lcA.y));
        const V2 rB = mul(qB, mk(an.z, an.w) - mk(lcB.x, lcB.y));
        const V2 C = cB + rB - cA - rA;
        positionError = lengthOf(C);
        const float k11 = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
        const float k21 = -iA * rA.x * rA.y - iB * rB.x * rB.y;
        const float k12 = k21;
        const float k22 = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
        // b2Mat22::Solve (REF/Common/b2Math.h:242-254)
        float det = k11 * k22 - k12 * k21;
        if (det != 0.0f) det = 1.0f / det;
        const V2 impulse = -mk(det * (k22 * C.x - k12 * C.y), det * (k11 * C.y - k21 * C.x));
        cA -= mA * impulse;
        aA -= iA * cross(rA, impulse);
        cB += mB * impulse;
        aB += iB * cross(rB, impulse);
    }
    if (w.btype[b.x] == BT_DYNAMIC) w.pos[b.x] = make_float4(cA.x, cA.y, aA, pA.w);
    if (w.btype[b.y] == BT_DYNAMIC) w.pos[b.y] = make_float4(cB.x, cB.y, aB, pB.w);
    return positionError <= kLinearSlop && angularError <= kAngularSlop;
}

// One joint's SolvePositionConstraints by type; friction, motor and mouse
// joints have no position pass and report true.
__device__ inline bool jointPositionCore(const WorldView& w, const JointView& jv, int j, int2 b) {
    const int type = jointType(jv.flags[j]);
    switch (type) {
    case JT_FRICTION:
    case JT_MOTOR:
    case JT_MOUSE: return true;
    case JT_DISTANCE: return distanceSolvePositionOne(w, jv, j, b);
    case JT_ROPE: return ropeSolvePositionOne(w, jv, j, b);
    case JT_PULLEY: return pulleySolvePositionOne(w, jv, j, b);
    case JT_PRISMATIC: return prismaticSolvePositionOne(w, jv, j, b);
    case JT_WHEEL: return wheelSolvePositionOne(w, jv, j, b);
    case JT_GEAR: return gearSolvePositionOne(w, jv, j, b);
    case JT_WELD: return weldSolvePositionOne(w, jv, j, b);
    default: return revoluteSolvePositionOne(w, jv, j, b);
    }
}

// A joint left with an error above the slops marks its island's pass as
// failed (the jointsOkay of b2Island::Solve, REF/Dynamics/b2Island.cpp:310-328).
__device__ inline void jointSolvePositionOne(const WorldView& w, const JointView& jv, int j, int it) {
    const int2 b = jv.bodies[j];
    if (!jointActive(w, b)) return;
    const int root = w.islandOf[w.btype[b.x] != BT_STATIC ? b.x : b.y];
    if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) return;  // island done
    if (!jointPositionCore(w, jv, j, b)) {
        int* f = &w.islandMinSep[it * w.nBodies + root];
        if (*f == 0) *f = 1;
    }
}

__global__ void k_joint_solve_position(WorldView w, JointView jv, int begin, int end, int it) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    jointSolvePositionOne(w, jv, jv.order[k], it);
}

__global__ void k_joint_solve_position_serial(WorldView w, JointView jv, int begin, int end, int it) {
    for (int k = begin; k < end; ++k) jointSolvePositionOne(w, jv, jv.order[k], it);
}

// Island union through joints whose bodies are both non-static and active.
__global__ void k_island_union_joints(WorldView w, JointView jv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= jv.n) return;
    const int2 b = jv.bodies[j];
    if (jv.flags[j] & JF_REMOVED) return;
    if (w.btype[b.x] == BT_STATIC || w.btype[b.y] == BT_STATIC) return;
    if (!(w.bflags[b.x] & BF_ACTIVE) || !(w.bflags[b.y] & BF_ACTIVE)) return;
    islandUnion(w.islandOf, b.x, b.y);
}

// Position error of each joint, |cB + rB - cA - rA| as
// b2RevoluteJoint::SolvePositionConstraints measures it
// (REF/Dynamics/Joints/b2RevoluteJoint.cpp:336-343); -1 for removed joints.
__global__ void k_joint_drift(WorldView w, JointView jv, float* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= jv.n) return;
    if (jv.flags[j] & JF_REMOVED) {
        out[j] = -1.0f;
        return;
    }
    const int2 b = jv.bodies[j];
    const float4 pA = w.pos[b.x], pB = w.pos[b.y];
    const float2 lcA = w.lcenter[b.x], lcB = w.lcenter[b.y];
    const float4 an = jv.anchors[j];
    const V2 rA = mul(rotOf(pA.z), mk(an.x, an.y) - mk(lcA.x, lcA.y));
    const V2 rB = mul(rotOf(pB.z), mk(an.z, an.w) - mk(lcB.x, lcB.y));
    const float gap = lengthOf(mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA);
    const int type = jointType(jv.flags[j]);
    if (type == JT_DISTANCE) {
        // rods: the error in length; springs are meant to stretch
        out[j] = jv.limits[j].y > 0.0f ? -1.0f : absB(gap - jv.limits[j].x);
        return;
    }
    if (type == JT_ROPE) {  // stretch beyond the maximum length
        out[j] = fmaxf(gap - jv.limits[j].x, 0.0f);
        return;
    }
    if (type == JT_FRICTION || type == JT_MOTOR || type == JT_MOUSE) {  // no positional constraint
        out[j] = -1.0f;
        return;
    }
    if (type == JT_GEAR) {  // |coordinateA + ratio coordinateB - constant|
        const int4 q = gearBodies(jv, j, b);
        const float4 p[4] = {w.pos[q.x], w.pos[q.y], w.pos[q.z], w.pos[q.w]};
        const GearRows g = gearRows(w, jv, j, q, p, true);
        out[j] = absB((g.coordinateA + jv.limits[j].z * g.coordinateB) - jv.limits[j].w);
        return;
    }
    if (type == JT_PRISMATIC || type == JT_WHEEL) {  // distance of anchor B from the line through anchor A
        const V2 axis = mul(rotOf(pA.z), mk(jv.axis[j].x, jv.axis[j].y));
        const V2 d = mk(pB.x, pB.y) + rB - mk(pA.x, pA.y) - rA;
        out[j] = absB(cross(axis, d)) / lengthOf(axis);
        return;
    }
    if (type == JT_PULLEY) {  // |lengthA + ratio lengthB - constant|
        const float4 g = jv.limits[j];
        const float la = lengthOf(mk(pA.x, pA.y) + rA - mk(g.x, g.y));
        const float lb = lengthOf(mk(pB.x, pB.y) + rB - mk(g.z, g.w));
        out[j] = absB(la + jv.motorSpeed[j] * lb - jv.impulse[j].w);
        return;
    }
    out[j] = gap;  // revolute and weld: the anchors' separation
}

// Kinetic (0.5 m v.v + 0.5 I w^2) and potential (-m g.c) energy per body.
__global__ void k_body_energy(WorldView w, V2 g, double2* out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    const float4 ms = w.mass[b];
    const double m = ms.x > 0.0f ? 1.0 / (double)ms.x : 0.0;
    const double I = ms.y > 0.0f ? 1.0 / (double)ms.y : 0.0;
    const float4 v = w.vel[b], p = w.pos[b];
    const double ke = 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y) + 0.5 * I * (double)v.z * v.z;
    const double pe = -m * ((double)g.x * p.x + (double)g.y * p.y);
    out[b] = make_double2(ke, pe);
}

// Copy-back transform of every solved body (b2Body::SynchronizeTransform,
// REF/Dynamics/b2Body.h:836-840). Static bodies keep their creation-time
// transform, which the reference's recomputation reproduces exactly.
__device__ inline void syncTransformOne(const WorldView& w, int b) {
    const float4 p = w.pos[b];
    const float2 lc = w.lcenter[b];
    Xf xf;
    xf.q = rotOf(p.z);
    xf.p = mk(p.x, p.y) - mul(xf.q, mk(lc.x, lc.y));
    w.xf[b] = f4(xf);
}

__global__ void k_sync_transforms(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || !islandBody(w, b)) return;
    syncTransformOne(w, b);
}

// ====================================================================
// One-launch solve: the passes of b2Island::Solve (warm start, velocity
// iterations, impulse store and position integration, position
// iterations, transform sync) in one cooperative launch whose workgroups
// are all resident, the colours separated by a grid-wide barrier instead
// of kernel boundaries. Every constraint is solved by one lane with the
// per-colour kernels' functions, colour after colour, so the results are
// theirs bit for bit; what goes is the dispatch gap between two colours.
// ====================================================================
enum : int { SP_WARM = 1, SP_VELOCITY = 2, SP_POSITION = 4 };

struct SolvePlan {
    int cs[kColourSlots + 1];  // contact colour c holds constraints [cs[c], cs[c + 1])
    int js[kColourSlots + 1];  // joint colour c holds order slots [js[c], js[c + 1])
    int nColours;              // contact colour slots in use
    int tailBegin;             // colours [tailBegin, nColours) by workgroup 0, a workgroup barrier apart
    int nJointColours;         // joint colour slots to visit (0 without joints)
    int stages;                // SP_* passes of this launch
    int contactWarm;           // warm start the contacts (SP_WARM)
    int jointWarm;             // the joints' warm starting flag (SP_WARM)
    float dtRatio;
    float h;                   // time step
    int velIters, posIters;
    int nCons;
    int pad;
};

// Spin bound of the grid barrier (wall clock at 100 MHz: 2 s). A workgroup
// that waits longer sets the abort word, after which every barrier of the
// launch returns at once, so the launch always ends; the host reports it.
constexpr long long kBarrierSpinTicks = 200000000ll;

// Grid-wide barrier. The workgroup barrier first waits for every lane's
// stores; lane 0 then arrives with an agent-scope release (L2 written back
// for the other XCDs) and leaves with an agent-scope acquire (stale lines
// invalidated), which the workgroup's lanes share through the closing
// workgroup barrier. The last workgroup to arrive resets the count and
// opens the next generation.
__device__ inline void gridBarrier(Counters* cnt) {
    __syncthreads();
    if (threadIdx.x == 0 && __hip_atomic_load(&cnt->barrierAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        const unsigned g = __hip_atomic_load(&cnt->barrierGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived =
            __hip_atomic_fetch_add(&cnt->barrierCount, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&cnt->barrierCount, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&cnt->barrierGen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const long long t0 = wall_clock64();
            while (__hip_atomic_load(&cnt->barrierGen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > kBarrierSpinTicks) {
                    __hip_atomic_store(&cnt->barrierAbort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                if (__hip_atomic_load(&cnt->barrierAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            }
        }
    }
    __syncthreads();
}

__device__ inline int waveMin(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

// kJoints: the joint passes are compiled in (worlds with joints), with at
// most kPersistJointBlock lanes per workgroup so the joint code keeps its
// registers; without them the kernel keeps the contact passes' budget.
constexpr int kPersistJointBlock = 256;
template <bool kJoints>
__global__ void __launch_bounds__(kJoints ? kPersistJointBlock : 1024) k_solve_persistent(WorldView w, ConView cv, JointView jv, SolvePlan p,
                                                           Counters* cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    // Contact colours: the wide ones shared by the grid (the overflow colour
    // by one lane in order), then the trailing ones by workgroup 0.
    auto contactColours = [&](auto one) {
        for (int c = 0; c < p.tailBegin; ++c) {
            const int b0 = p.cs[c], b1 = p.cs[c + 1];
            if (b1 <= b0) continue;
            if (c == kOverflowColour) {
                if (t == 0)
                    for (int k = b0; k < b1; ++k) one(k);
            } else {
                for (int k = b0 + t; k < b1; k += stride) one(k);
            }
            gridBarrier(cnt);
        }
        if (p.tailBegin < p.nColours) {
            if (blockIdx.x == 0) {
                for (int c = p.tailBegin; c < p.nColours; ++c) {
                    if (c == kOverflowColour) {
                        if (threadIdx.x == 0)
                            for (int k = p.cs[c]; k < p.cs[c + 1]; ++k) one(k);
                    } else {
                        for (int k = p.cs[c] + (int)threadIdx.x; k < p.cs[c + 1]; k += blockDim.x) one(k);
                    }
                    __syncthreads();
                }
            }
            gridBarrier(cnt);
        }
    };
    auto jointColours = [&](auto one) {
        if (!kJoints) return;
        for (int c = 0; c < p.nJointColours; ++c) {
            const int b0 = p.js[c], b1 = p.js[c + 1];
            if (b1 <= b0) continue;
            if (c == kOverflowColour) {
                if (t == 0)
                    for (int k = b0; k < b1; ++k) one(jv.order[k]);
            } else {
                for (int k = b0 + t; k < b1; k += stride) one(jv.order[k]);
            }
            gridBarrier(cnt);
        }
    };
    // b2Island::Solve, REF/Dynamics/b2Island.cpp:246-270: contact warm
    // start, the joints' init, then per iteration the joints and the contacts.
    if (p.stages & SP_WARM) {
        if (p.contactWarm) contactColours([&](int k) { warmStartOne(w, cv, k); });
        jointColours([&](int j) { jointInitOne(w, jv, j, p.dtRatio, p.jointWarm); });
    }
    if (p.stages & SP_VELOCITY) {
        for (int it = 0; it < p.velIters; ++it) {
            jointColours([&](int j) { jointSolveVelocityOne(w, jv, j, p.h); });
            contactColours([&](int k) { solveVelocityOne(w, cv, k); });
        }
    }
    if (!(p.stages & SP_POSITION)) return;
    // StoreImpulses and the position integration (b2Island.cpp:272-306).
    for (int k = t; k < p.nCons; k += stride) storeImpulsesOne(w, cv, k);
    for (int b = t; b < w.nBodies; b += stride)
        if (islandBody(w, b)) integratePositionOne(w, b, p.h);
    gridBarrier(cnt);
    // Position iterations: contacts, then joints (b2Island.cpp:310-328), each
    // skipped for islands the previous pass left solved.
    for (int it = 0; it < p.posIters; ++it) {
        int lowest = 0;
        contactColours([&](int k) {
            const int root = cv.island[k];
            if (it > 0 && w.islandMinSep[(it - 1) * w.nBodies + root] == 0) return;
            const float minSep = solvePositionOne(w, cv, k);
            if (minSep < -3.0f * kLinearSlop) {
                int* f = &w.islandMinSep[it * w.nBodies + root];
                if (*f == 0) *f = 1;
            }
            if (minSep < 0.0f) lowest = min(lowest, encodeOrdered(minSep));
        });
        lowest = waveMin(lowest);
        if ((threadIdx.x & 63) == 0 && lowest < 0) atomicMin(&cnt->minSep[it], lowest);
        jointColours([&](int j) { jointSolvePositionOne(w, jv, j, it); });
    }
    for (int b = t; b < w.nBodies; b += stride)
        if (islandBody(w, b)) syncTransformOne(w, b);
}

// Sleep (b2Island::Solve, REF/Dynamics/b2Island.cpp:345-382): per island,
// the minimum over its bodies of the time spent below the sleep tolerances;
// an island that has rested for b2_timeToSleep and whose positions were
// solved goes to sleep as a whole.
__global__ void k_sleep_accum(WorldView w, float h) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || !islandBody(w, b)) return;
    const float4 v = w.vel[b];
    const float linTolSqr = kLinearSleepTol * kLinearSleepTol;
    const float angTolSqr = kAngularSleepTol * kAngularSleepTol;
    float t;
    if ((w.bflags[b] & BF_AUTOSLEEP) == 0 || v.z * v.z > angTolSqr || dot(mk(v.x, v.y), mk(v.x, v.y)) > linTolSqr) {
        t = 0.0f;
    } else {
        t = w.gsleep[b].y + h;
    }
    w.gsleep[b].y = t;
    if (t < kTimeToSleep) {
        int* f = &w.islandSleep[w.islandOf[b]];
        if (*f == 0) *f = 1;  // an island body has not rested long enough
    }
}

// Static bodies reached by the island search (REF/Dynamics/b2World.cpp:
// 451-529): the search wakes every body it adds, static ones included
// (without searching on from them), and an island that falls asleep puts all
// its bodies to sleep (REF/Dynamics/b2Island.cpp:366-382). Islands are solved
// in the order of their seeds, the body list from its newest body, so by
// decreasing largest body index; a static body reached by several islands
// keeps the state that the island of the smallest largest index leaves.
// islandMax holds each island's largest body index (one atomic per
// wavefront when the wavefront's bodies share an island); staticKey the
// minimum over the solved islands reaching a static body of
// (largest index << 1) | (island stays awake).
__global__ void k_island_max(WorldView w, int* islandMax) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = b < w.nBodies && islandBody(w, b);
    const int root = valid ? w.islandOf[b] : -1;
    const int r0 = __shfl(root, 0);
    const unsigned long long same = __ballot(root == r0);
    if (r0 >= 0 && same == ~0ull) {
        // the wavefront's bodies all belong to island r0: its last lane has the largest index
        if (__lane_id() == 63) atomicMax(&islandMax[r0], b);
        return;
    }
    if (valid) atomicMax(&islandMax[root], b);
}

__device__ inline int staticReachKey(const WorldView& w, const int* islandMax, int root, int allowSleep,
                                     int posIters) {
    bool solved = false;
    for (int it = 0; it < posIters; ++it)
        if (w.islandMinSep[it * w.nBodies + root] == 0) solved = true;
    const bool slept = allowSleep && w.islandSleep[root] == 0 && solved;
    return (islandMax[root] << 1) | (slept ? 0 : 1);
}

__global__ void k_static_reach(WorldView w, const int* islandMax, int* staticKey, int allowSleep, int posIters) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts || !(w.cflags[c] & CF_CONSTRAINT)) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const bool sA = w.btype[bA] == BT_STATIC, sB = w.btype[bB] == BT_STATIC;
    if (sA == sB) return;
    const int st = sA ? bA : bB, dyn = sA ? bB : bA;
    atomicMin(&staticKey[st], staticReachKey(w, islandMax, w.islandOf[dyn], allowSleep, posIters));
}

__global__ void k_static_reach_joints(WorldView w, JointView jv, const int* islandMax, int* staticKey,
                                      int allowSleep, int posIters) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= jv.n || (jv.flags[j] & JF_REMOVED)) return;
    const int2 b = jv.bodies[j];
    if (!jointActive(w, b)) return;
    const bool sA = w.btype[b.x] == BT_STATIC, sB = w.btype[b.y] == BT_STATIC;
    if (sA == sB) return;
    const int st = sA ? b.x : b.y, dyn = sA ? b.y : b.x;
    if (!islandBody(w, dyn)) return;
    atomicMin(&staticKey[st], staticReachKey(w, islandMax, w.islandOf[dyn], allowSleep, posIters));
}

// b2Body::SetAwake (REF/Dynamics/b2Body.h:635-654) on the reached static bodies.
__global__ void k_static_apply(WorldView w, int* staticKey) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    const int key = staticKey[b];
    if (key == 0x7FFFFFFF) return;
    staticKey[b] = 0x7FFFFFFF;
    if (key & 1) {
        if (!(w.bflags[b] & BF_AWAKE)) {
            w.bflags[b] |= BF_AWAKE;
            w.gsleep[b].y = 0.0f;
        }
    } else {
        w.bflags[b] &= ~BF_AWAKE;
        w.gsleep[b].y = 0.0f;
        w.vel[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        w.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_sleep_apply(WorldView w, int posIters) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || !islandBody(w, b)) return;
    const int root = w.islandOf[b];
    bool solved = false;
    for (int it = 0; it < posIters; ++it)
        if (w.islandMinSep[it * w.nBodies + root] == 0) solved = true;
    if (!(w.islandSleep[root] == 0 && solved)) return;
    // b2Body::SetAwake(false)
    w.bflags[b] &= ~BF_AWAKE;
    w.gsleep[b].y = 0.0f;
    w.vel[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ inline float4 shapeAABB(const b2g_shape& s, const Xf& xf) {
    if (s.type == B2G_CIRCLE) {
        const V2 p = xf.p + mul(xf.q, ld(s.centroid));
        return make_float4(p.x - s.radius, p.y - s.radius, p.x + s.radius, p.y + s.radius);
    }
    const int n = s.type == B2G_EDGE ? 2 : s.count;
    V2 lo = mul(xf, ld(s.verts[0]));
    V2 hi = lo;
    for (int i = 1; i < n; ++i) {
        const V2 v = mul(xf, ld(s.verts[i]));
        lo = vmin(lo, v);
        hi = vmax(hi, v);
    }
    if (s.type == B2G_EDGE && (s.flags & 4)) return make_float4(lo.x, lo.y, hi.x, hi.y);  // chain child
    const V2 r = mk(s.radius, s.radius);
    const V2 l = lo - r, u = hi + r;
    return make_float4(l.x, l.y, u.x, u.y);
}

// b2DynamicTree::MoveProxy's fat-box rule (REF/Collision/b2DynamicTree.cpp:133-181):
// keep the fat box while it contains the tight one, else re-fatten by the
// AABB extension and stretch by twice the displacement toward the motion.
__device__ inline bool fatRule(float4& fat, float4 tight, V2 disp) {
    if (containsBox(fat, tight)) return false;
    float4 b = make_float4(tight.x - kAabbExtension, tight.y - kAabbExtension, tight.z + kAabbExtension,
                           tight.w + kAabbExtension);
    const V2 d = kAabbMultiplier * disp;
    if (d.x < 0.0f) b.x += d.x; else b.z += d.x;
    if (d.y < 0.0f) b.y += d.y; else b.w += d.y;
    fat = b;
    return true;
}

// b2Body::SynchronizeFixtures + b2Fixture::Synchronize,
// REF/Dynamics/b2Body.cpp:443-454, REF/Dynamics/b2Fixture.cpp:154-176: the
// swept box from the sweep start to the current transform, run through the
// fat-box rule. Removed proxies (inverted fat boxes) stay removed.
__device__ inline bool syncFixtureOne(const WorldView& w, int f, int b) {
    float4 fat = w.fat[f];
    if (!(fat.x <= fat.z && fat.y <= fat.w)) return false;
    const float4 p0 = w.pos0[b];
    const float2 lc = w.lcenter[b];
    Xf xf1;
    xf1.q = rotOf(p0.z);
    xf1.p = mk(p0.x, p0.y) - mul(xf1.q, mk(lc.x, lc.y));
    const Xf xf2 = xfOf(w.xf[b]);
    const b2g_shape& s = w.shape[f];
    const float4 a1 = shapeAABB(s, xf1);
    const float4 a2 = shapeAABB(s, xf2);
    const float4 tight = make_float4(fminB(a1.x, a2.x), fminB(a1.y, a2.y), fmaxB(a1.z, a2.z), fmaxB(a1.w, a2.w));
    if (!fatRule(fat, tight, xf2.p - xf1.p)) return false;
    w.fat[f] = fat;
    w.moved[f] = 1;
    return true;
}

__global__ void k_sync_fixtures(WorldView w) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= w.nFixtures) return;
    const int b = w.fbody[f];
    if (!islandBody(w, b)) return;
    syncFixtureOne(w, f, b);
}

// ====================================================================
// Broad-phase: uniform hash grid over fat AABBs. Every proxy is binned into
// the cells its fat box covers; proxies covering more than
// kMaxCellsPerProxy cells go to a short "wide" list tested by brute force.
// Each moved proxy reports the pairs b2BroadPhase::UpdatePairs would
// (REF/Collision/b2BroadPhase.h:188-241): every other proxy whose fat box
// overlaps its own. A pair of two moved proxies is reported by the lower id
// only, and a pair seen in several cells only in the cell holding the lower
// corner of the two boxes' intersection, so each pair appears once.
// ====================================================================
struct GridView {
    float invCell;
    uint32_t bucketMask;
    int* counts;     // nBuckets + 1
    int* starts;     // exclusive scan of counts
    int* cursor;     // fill cursor per bucket
    int* entries;    // proxy ids by bucket; bit 31 set: the proxy moved
    float4* efat;    // each entry's fat box, next to its id
    int* queries;    // moved proxies binned in the grid, in no fixed order
    unsigned long long* ecell;  // the cell each entry was binned for (buckets mix cells)
    unsigned long long* big;  // wide proxies: (sub-world << 32) | proxy id, sorted when bigSorted
    int bigSorted;
    int4* range;     // per proxy cell range (x0, y0, x1, y1); x0 > x1 marks a wide proxy
    const int* fbody;   // proxy -> body, with bworld: the proxy's sub-world
    const int* bworld;  // null: every proxy in sub-world 0
};

// Sub-world of a proxy: proxies of different sub-worlds share no cell.
__device__ inline int proxyWorld(const GridView& g, int f) { return g.bworld ? g.bworld[g.fbody[f]] : 0; }

// Entries [b0, b1) of the wide-proxy list that may belong to `world`: the
// whole list when unsorted, else the sub-world's run found by binary search.
__device__ inline void bigRange(const GridView& g, int nBig, int world, int& b0, int& b1) {
    if (!g.bigSorted) {
        b0 = 0;
        b1 = nBig;
        return;
    }
    auto lower = [&](unsigned long long key) {
        int lo = 0, hi = nBig;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (g.big[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    b0 = lower((unsigned long long)world << 32);
    b1 = lower((unsigned long long)(world + 1) << 32);
}

__device__ inline int cellCoord(float x, float inv) {
    float c = floorf(x * inv);
    c = fminB(fmaxB(c, -1.0e9f), 1.0e9f);
    return (int)c;
}
// Cell identity within a sub-world: 24 bits of sub-world, 20 bits of each
// coordinate (coordinates that alias only cost extra overlap tests).
__device__ inline unsigned long long cellKey(int ix, int iy, int world) {
    return ((unsigned long long)(uint32_t)(world & 0xFFFFFF) << 40) |
           ((unsigned long long)((uint32_t)ix & 0xFFFFFu) << 20) | ((uint32_t)iy & 0xFFFFFu);
}
__device__ inline uint32_t cellHash(int ix, int iy, int world) {
    return ((uint32_t)ix * 0x8DA6B343u) ^ ((uint32_t)iy * 0xD8163841u) ^ ((uint32_t)world * 0xCB1AB31Fu);
}


__global__ void k_grid_count(const float4* fat, const uint8_t* moved, int n, const uint8_t* hasProxy, GridView g,
                             Counters* cnt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    bool query = false;
    if (f < n) {
        if (hasProxy && !hasProxy[f]) {
            g.range[f] = make_int4(1, 0, 0, 0);
        } else {
            const float4 a = fat[f];
            if (!(a.x <= a.z && a.y <= a.w)) {
                g.range[f] = make_int4(1, 0, 0, -1);  // removed proxy: not binned, never queried
            } else {
                const int x0 = cellCoord(a.x, g.invCell), y0 = cellCoord(a.y, g.invCell);
                const int x1 = cellCoord(a.z, g.invCell), y1 = cellCoord(a.w, g.invCell);
                const long long cells = (long long)(x1 - x0 + 1) * (long long)(y1 - y0 + 1);
                if (cells > kMaxCellsPerProxy) {
                    g.range[f] = make_int4(1, 0, 0, 0);
                    const int slot = atomicAdd(&cnt->nBig, 1);
                    g.big[slot] = ((unsigned long long)proxyWorld(g, f) << 32) | (uint32_t)f;
                } else {
                    g.range[f] = make_int4(x0, y0, x1, y1);
                    query = moved && moved[f];
                    const int wf = proxyWorld(g, f);
                    for (int iy = y0; iy <= y1; ++iy)
                        for (int ix = x0; ix <= x1; ++ix) atomicAdd(&g.counts[cellHash(ix, iy, wf) & g.bucketMask], 1);
                }
            }
        }
    }
    const int slot = waveAppend(&cnt->nQueries, query);
    if (query) g.queries[slot] = f;
}

__global__ void k_grid_scatter(int n, GridView g, const float4* fat, const uint8_t* moved) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int4 r = g.range[f];
    if (r.x > r.z) return;
    const int wf = proxyWorld(g, f);
    const float4 box = fat[f];
    const int id = f | (moved[f] ? (int)0x80000000 : 0);
    for (int iy = r.y; iy <= r.w; ++iy)
        for (int ix = r.x; ix <= r.z; ++ix) {
            const uint32_t bkt = cellHash(ix, iy, wf) & g.bucketMask;
            const int e = g.starts[bkt] + atomicAdd(&g.cursor[bkt], 1);
            g.entries[e] = id;
            g.efat[e] = box;
            g.ecell[e] = cellKey(ix, iy, wf);
        }
}

// Pair acceptance of b2ContactManager::AddPair (REF/Dynamics/b2ContactManager.cpp:179-230):
// distinct bodies, no existing contact, at least one dynamic body, filter
// pass, and a contact type for the two shapes.
__device__ inline bool acceptPair(const WorldView& w, int a, int b, const unsigned long long* table, uint32_t hmask) {
    // the existing-contact test first: it needs only the proxy ids and
    // rejects most overlapping pairs of a settled scene
    if (table && hashContains(table, hmask, pairKey(a, b))) return false;
    const int ba = w.fbody[a], bb = w.fbody[b];
    if (ba == bb) return false;
    if (w.bworld && w.bworld[ba] != w.bworld[bb]) return false;
    if (jointKeepsApart(w, ba, bb)) return false;
    if (w.btype[ba] != BT_DYNAMIC && w.btype[bb] != BT_DYNAMIC) return false;
    if (!filterPass(w, a, b)) return false;
    const int ta = w.shape[a].type, tb = w.shape[b].type;
    return primaryPair(ta, tb) || primaryPair(tb, ta);
}

// The broad-phase pass's zeroing in one launch: bucket counts and cursors,
// and the pass's counters (wide proxies, queries, new pairs, overflow).
__global__ void k_grid_reset(int* counts, int* cursor, int n, Counters* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        counts[i] = 0;
        cursor[i] = 0;
    }
    if (i == 0) {
        cnt->nBig = 0;
        cnt->nQueries = 0;
        cnt->nNewPairs = 0;
        cnt->overflow = 0;
    }
}

// Where queries put the pairs they find. With packBase = n (the proxy
// count, or the order-key span with `ord`) a pair is stored as lo * n + hi,
// lo < hi its two fixtures' order keys (`ord`, or the fixture indices when
// null), which orders pairs as pairKey does in fewer bits (fewer radix
// passes); with packBase 0 as pairKey of the fixture indices.
struct PairSink {
    unsigned long long* keys;
    int cap;
    unsigned long long packBase;
    const int* ord;
};

__device__ inline void emitPair(PairSink sink, Counters* cnt, int a, int b) {
    unsigned long long key;
    if (sink.packBase) {
        if (sink.ord) {
            const int ka = sink.ord[a], kb = sink.ord[b];
            if (ka < 0 || kb < 0) return;  // no proxy (only inverted boxes, which overlap nothing)
            a = ka < kb ? ka : kb;
            b = ka < kb ? kb : ka;
        }
        key = (unsigned long long)a * sink.packBase + (unsigned long long)b;
    } else {
        key = pairKey(a, b);
    }
    const int slot = atomicAdd(&cnt->nNewPairs, 1);
    if (slot < sink.cap) sink.keys[slot] = key;
    else atomicOr(&cnt->overflow, 1);
}

// Queries of the moved proxies binned in the grid (the list k_grid_count
// builds, in no fixed order): kQueryLanes lanes per query, visiting the
// query's cells one after another with the lanes striding over each cell's
// entries (a cell holds about as many entries as a group has lanes), and a
// share of the wide proxies each.
// A pair is reported from the one cell holding the corner of the two boxes'
// intersection, so lanes never report a pair twice. Each candidate's id,
// move flag and fat box come from its grid entry. `world` is false for the
// bare pair-set computation (no contact filtering).
constexpr int kQueryLanes = 16;  // a divisor of 64

__device__ inline void queryOne(const float4* fat, const uint8_t* moved, const GridView& g, const WorldView& w,
                                bool world, const unsigned long long* table, uint32_t hmask, const PairSink& sink,
                                Counters* cnt, int i, int lane) {
    const int q = g.queries[i];
    const float4 fq = fat[q];
    const int4 r = g.range[q];
    const int nBig = cnt->nBig;
    auto consider = [&](int p, bool movedP, const float4& fp, bool checkCell, int ix, int iy) {
        if (p == q) return;
        if (movedP && p < q) return;  // reported by p's own query
        if (!overlaps(fq, fp)) return;
        if (checkCell) {
            const int ox = cellCoord(fmaxB(fq.x, fp.x), g.invCell);
            const int oy = cellCoord(fmaxB(fq.y, fp.y), g.invCell);
            if (ox != ix || oy != iy) return;
        }
        const int a = q < p ? q : p, b = q < p ? p : q;
        if (world && !acceptPair(w, a, b, table, hmask)) return;
        emitPair(sink, cnt, a, b);
    };
    const int wq = proxyWorld(g, q);
    const int width = r.z - r.x + 1, ncell = width * (r.w - r.y + 1);
    // cell after cell, the group's lanes striding over the cell's entries
    for (int k = 0; k < ncell; ++k) {
        const int ix = r.x + k % width, iy = r.y + k / width;
        const uint32_t bkt = cellHash(ix, iy, wq) & g.bucketMask;
        const int s0 = g.starts[bkt], s1 = g.starts[bkt + 1];
        const unsigned long long ck = cellKey(ix, iy, wq);
        for (int e = s0 + lane; e < s1; e += kQueryLanes) {
            if (g.ecell[e] != ck) continue;
            const int id = g.entries[e];
            consider(id & 0x7FFFFFFF, id < 0, g.efat[e], true, ix, iy);
        }
    }
    int b0, b1;
    bigRange(g, nBig, wq, b0, b1);
    for (int k = b0 + lane; k < b1; k += kQueryLanes) {
        const unsigned long long e = g.big[k];
        if ((int)(e >> 32) != wq) continue;
        const int p = (int)(e & 0xFFFFFFFFu);
        consider(p, moved[p] != 0, fat[p], false, 0, 0);
    }
}

// The grid may be smaller than the query list (a launch sized for every
// proxy leaves most lanes idle when few moved): groups stride over it.
__global__ void k_pairs_query(const float4* fat, const uint8_t* moved, int n, GridView g, WorldView w, bool world,
                              const unsigned long long* table, uint32_t hmask, PairSink sink, Counters* cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = t % kQueryLanes;
    const int nq = cnt->nQueries < n ? cnt->nQueries : n;
    const int groups = (int)(gridDim.x * blockDim.x / kQueryLanes);
    for (int i = t / kQueryLanes; i < nq; i += groups) queryOne(fat, moved, g, w, world, table, hmask, sink, cnt, i, lane);
}

// Grid entries and wide proxies that k_pairs_query examines (the buckets of
// each query's cells and its sub-world's wide list), one thread per query:
// the unit count of the query's byte model, for the timing pass only.
__global__ void k_count_scanned(int n, GridView g, Counters* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int scanned = 0;
    if (i < n && i < cnt->nQueries) {
        const int q = g.queries[i];
        const int4 r = g.range[q];
        const int wq = proxyWorld(g, q);
        const int width = r.z - r.x + 1, ncell = width * (r.w - r.y + 1);
        for (int k = 0; k < ncell; ++k) {
            const uint32_t bkt = cellHash(r.x + k % width, r.y + k / width, wq) & g.bucketMask;
            scanned += g.starts[bkt + 1] - g.starts[bkt];
        }
        int b0, b1;
        bigRange(g, cnt->nBig, wq, b0, b1);
        scanned += b1 - b0;
    }
    for (int off = 32; off > 0; off >>= 1) scanned += __shfl_down(scanned, off);
    if ((threadIdx.x & 63) == 0 && scanned) atomicAdd(&cnt->nScanned, scanned);
}

// Queries of moved wide proxies (those not binned in the grid), one thread
// per candidate proxy: the same pairs as the wide proxy testing every proxy
// itself, under the same rule for two moved proxies (the lower id reports).
// Removed proxies fail every overlap test.
__global__ void k_pairs_wide(const float4* fat, const uint8_t* moved, int n, GridView g, WorldView w, bool world,
                             const unsigned long long* table, uint32_t hmask, PairSink sink, Counters* cnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int nBig = cnt->nBig;
    if (nBig == 0) return;
    const float4 fp = fat[p];
    const bool movedP = moved[p] != 0;
    const int wp = proxyWorld(g, p);
    int b0, b1;
    bigRange(g, nBig, wp, b0, b1);
    for (int i = b0; i < b1; ++i) {
        const unsigned long long e = g.big[i];
        const int q = (int)(e & 0xFFFFFFFFu);
        if (q == p || !moved[q] || (int)(e >> 32) != wp) continue;
        if (movedP && p < q) continue;  // reported by p's own query
        if (!overlaps(fat[q], fp)) continue;
        const int a = q < p ? q : p, b = q < p ? p : q;
        if (world && !acceptPair(w, a, b, table, hmask)) continue;
        emitPair(sink, cnt, a, b);
    }
}

// New contacts from the sorted pair keys (b2Contact::Create with the
// registry's fixture swap, b2Contact ctor, AddPair's wake-up;
// REF/Dynamics/Contacts/b2Contact.cpp:64-95, 136-168, REF/Dynamics/b2ContactManager.cpp:284-293).
// b2ContactManager::AddPair's contact creation (REF/Dynamics/b2ContactManager.cpp:
// 179-296): registry order, mixed friction and restitution, no manifold, and
// both bodies woken unless a fixture is a sensor.
// a: the fixture of the pair's smaller order key (AddPair's fixture A).
__device__ inline void appendContactOne(const WorldView& w, int a, int b, int c) {
    if (!primaryPair(w.shape[a].type, w.shape[b].type)) {
        const int t = a;
        a = b;
        b = t;
    }
    w.cfA[c] = a;
    w.cfB[c] = b;
    w.cflags[c] = CF_ENABLED;
    const float4 ma = w.fmat[a], mb = w.fmat[b];
    w.cmat[c] = make_float2(sqrtf(ma.x * mb.x), ma.y > mb.y ? ma.y : mb.y);
    w.ctan[c] = 0.0f;
    b2g_manifold m;
    zeroManifold(m);
    w.cman[c] = m;
    w.ccolour[c] = -1;
    w.ctoi[c] = 1.0f;
    w.ctoiCount[c] = 0;
    if (ma.w == 0.0f && mb.w == 0.0f) {
        wakeBody(w, w.fbody[a]);
        wakeBody(w, w.fbody[b]);
    }
}

// keys packed as lo * packBase + hi of the two order keys (PairSink).
__global__ void k_append_contacts(WorldView w, const unsigned long long* keys, int nNew, int base,
                                  unsigned long long packBase) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nNew) return;
    const unsigned long long key = keys[k];
    appendContactOne(w, fixtureOf(w, (int)(key / packBase)), fixtureOf(w, (int)(key % packBase)), base + k);
}

// Contacts between bodies that a new joint keeps apart are flagged for
// filtering (b2World::CreateJoint, REF/Dynamics/b2World.cpp:250-266).
__global__ void k_flag_joint_contacts(WorldView w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (jointKeepsApart(w, w.fbody[w.cfA[c]], w.fbody[w.cfB[c]])) w.cflags[c] |= CF_FILTER;
}

// The pair keys of contacts first..first+n-1 into the live-contact set.
__global__ void k_hash_insert_range(WorldView w, int first, int n, unsigned long long* table, uint32_t mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = pairKey(w.cfA[first + i], w.cfB[first + i]);
    uint32_t h = hashU64(key) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[h], ~0ull, key);
        if (prev == ~0ull || prev == key) return;
        h = (h + 1) & mask;
    }
}

__global__ void k_hash_insert_keys(const unsigned long long* keys, int n, unsigned long long* table, uint32_t mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = keys[i];
    uint32_t h = hashU64(key) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[h], ~0ull, key);
        if (prev == ~0ull || prev == key) return;
        h = (h + 1) & mask;
    }
}

// Proxy removal: an inverted fat box overlaps nothing.
__global__ void k_remove_fixtures(WorldView w, const int* ids, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int f = ids[i];
    w.fat[f] = make_float4(FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX);
    w.moved[f] = 0;
}

// b2World::DestroyBody destroys the body's contacts at once, and destroying
// a contact with manifold points wakes both bodies (b2ContactManager::Destroy,
// b2Contact::Destroy; REF/Dynamics/b2World.cpp:159-167,
// b2ContactManager.cpp:39-101, Contacts/b2Contact.cpp:104-118). The wake-ups
// happen here, as the proxies go; the next step's collide drops the
// contacts before the activity test reads any body.
// b2Body::SetType destroys the body's contacts (the same wake-ups as above)
// and touches its proxies, so the step's broad-phase recreates the pairs that
// still overlap (REF/Dynamics/b2Body.cpp:115-172, b2BroadPhase::TouchProxy
// REF/Collision/b2BroadPhase.cpp:66-69). mark: per body, 1 for the bodies
// whose type changed.
__global__ void k_drop_body_contacts(WorldView w, const uint8_t* mark) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    if (!mark[bA] && !mark[bB]) return;
    const int flags = w.cflags[c];
    if (flags & (CF_DEAD | CF_DESTROY)) return;
    if (w.cman[c].pointCount > 0 && !sensorOf(w, fA) && !sensorOf(w, fB)) {
        wakeBody(w, bA);
        wakeBody(w, bB);
    }
    w.cflags[c] = flags | CF_DESTROY;
}

__global__ void k_touch_body_proxies(WorldView w, const uint8_t* mark) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= w.nFixtures || !mark[w.fbody[f]] || removedFixture(w, f)) return;
    w.moved[f] = 1;
}

__global__ void k_wake_removed_contacts(WorldView w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    const int fA = w.cfA[c], fB = w.cfB[c];
    if (!removedFixture(w, fA) && !removedFixture(w, fB)) return;
    if ((w.cflags[c] & CF_DEAD) || w.cman[c].pointCount == 0 || sensorOf(w, fA) || sensorOf(w, fB)) return;
    wakeBody(w, w.fbody[fA]);
    wakeBody(w, w.fbody[fB]);
}

// b2Fixture::Refilter (REF/Dynamics/b2Fixture.cpp:183-213): flag contacts of
// the given fixtures for filtering.
__global__ void k_flag_filter(WorldView w, const uint8_t* mark) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (mark[w.cfA[c]] || mark[w.cfB[c]]) w.cflags[c] |= CF_FILTER;
}

// ====================================================================
// Reference-order mode: Box2D's own visiting order, for parity
// (b2ContactManager::Collide over the newest-first contact list, islands by
// depth-first search from the newest-first body list, constraints solved
// sequentially in island order, REF/Dynamics/b2World.cpp:388-577,
// REF/Dynamics/b2Island.cpp:182-383). The contact table holds contacts in
// creation order (new pairs are appended in pair-key order and compaction
// keeps order), so "newest first" is descending table index.
// ====================================================================
struct RefView {
    b2g_manifold* tmpMan;   // narrowphase result per contact
    uint8_t* overlap;       // fat AABBs overlap
    int* edgeCount;         // per body (+1)
    int* edgeStart;         // exclusive scan of edgeCount
    int* edgeCursor;
    unsigned long long* edgeKeys;
    unsigned long long* edgeSorted;
    int* edgeContact;       // contact c, or ~j for joint j
    int* bvisited;
    int* cvisited;
    int* stack;
    int* islandBodies;
    int* islandContacts;
    int4* islandRange;      // bodyStart bodyCount contactStart contactCount
    int* islandSlept;       // per island: went to sleep this step
    int* lastIsland;        // per static body: last island (in search order) holding it
    int* jvisited;          // per joint: in an island this step
    int* islandJoints;
    int2* islandJointRange; // per island: jointStart jointCount
};

__global__ void k_collide_compute(WorldView w, RefView r) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    const bool ov = overlaps(w.fat[fA], w.fat[fB]);
    r.overlap[c] = ov ? 1 : 0;
    b2g_manifold m;
    zeroManifold(m);
    if (ov && (sensorOf(w, fA) || sensorOf(w, fB))) {
        if (testOverlap(w.shape[fA], xfOf(w.xf[bA]), w.shape[fB], xfOf(w.xf[bB]))) r.overlap[c] = 3;
    } else if (ov) {
        collideShapes(m, w.shape[fA], xfOf(w.xf[bA]), w.shape[fB], xfOf(w.xf[bB]));
    }
    r.tmpMan[c] = m;
}

// Sequential commit in the reference's list order: activity is read as
// earlier contacts' wake-ups leave it.
__global__ void k_collide_serial(WorldView w, RefView r, Counters* cnt) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int c = w.nContacts - 1; c >= 0; --c) {
        int flags = w.cflags[c] & ~CF_UPDATED;
        const int fA = w.cfA[c], fB = w.cfB[c];
        const int bA = w.fbody[fA], bB = w.fbody[fB];
        const bool destroyWake = w.cman[c].pointCount > 0 && !sensorOf(w, fA) && !sensorOf(w, fB);
        if (removedFixture(w, fA) || removedFixture(w, fB) || (flags & CF_DESTROY)) {
            if (flags & CF_TOUCHING) cnt->endEvents++;
            w.cflags[c] = flags | CF_DEAD;
            continue;
        }
        if (flags & CF_FILTER) {
            const bool dynAny = w.btype[bA] == BT_DYNAMIC || w.btype[bB] == BT_DYNAMIC;
            if (!dynAny || jointKeepsApart(w, bA, bB) || !filterPass(w, fA, fB)) {
                if (flags & CF_TOUCHING) cnt->endEvents++;
                if (destroyWake) { wakeBody(w, bA); wakeBody(w, bB); }
                w.cflags[c] = flags | CF_DEAD;
                continue;
            }
            flags &= ~CF_FILTER;
        }
        const bool activeA = (w.bflags[bA] & BF_AWAKE) && w.btype[bA] != BT_STATIC;
        const bool activeB = (w.bflags[bB] & BF_AWAKE) && w.btype[bB] != BT_STATIC;
        if (!activeA && !activeB) {
            w.cflags[c] = flags;
            continue;
        }
        if (!r.overlap[c]) {
            if (flags & CF_TOUCHING) cnt->endEvents++;
            if (destroyWake) { wakeBody(w, bA); wakeBody(w, bB); }
            w.cflags[c] = flags | CF_DEAD;
            continue;
        }
        if (sensorOf(w, fA) || sensorOf(w, fB)) {
            sensorFinish(w, cnt, c, flags, r.overlap[c] == 3);
            continue;
        }
        const b2g_manifold old = w.cman[c];
        b2g_manifold m = r.tmpMan[c];
        const bool wasTouching = (flags & CF_TOUCHING) != 0;
        const bool touching = m.pointCount > 0;
        for (int i = 0; i < m.pointCount; ++i) {
            for (int j = 0; j < old.pointCount; ++j) {
                if (old.points[j].id == m.points[i].id) {
                    m.points[i].normalImpulse = old.points[j].normalImpulse;
                    m.points[i].tangentImpulse = old.points[j].tangentImpulse;
                    break;
                }
            }
        }
        if (touching != wasTouching) { wakeBody(w, bA); wakeBody(w, bB); }
        flags |= CF_ENABLED | CF_UPDATED;
        if (touching) flags |= CF_TOUCHING;
        else flags &= ~CF_TOUCHING;
        if (wasTouching) flags |= CF_WAS_TOUCHING;
        else flags &= ~CF_WAS_TOUCHING;
        if (!wasTouching && touching) cnt->beginEvents++;
        if (wasTouching && !touching) cnt->endEvents++;
        w.cman[c] = m;
        w.cflags[c] = flags;
    }
}

__global__ void k_clear_island_flags(WorldView w, RefView r) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    w.bflags[b] &= ~BF_ISLAND;
    r.bvisited[b] = 0;
    r.edgeCount[b] = 0;
    r.edgeCursor[b] = 0;
    if (b == 0) r.edgeCount[w.nBodies] = 0;
}

__global__ void k_edge_count(WorldView w, RefView r, Counters* cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    r.cvisited[c] = 0;
    w.cflags[c] &= ~CF_CONSTRAINT;  // set again for the island contacts by k_dfs
    if (!islandContact(w, c)) return;
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    if (w.btype[bA] != BT_STATIC) { atomicAdd(&r.edgeCount[bA], 1); atomicAdd(&cnt->nEdges, 1); }
    if (w.btype[bB] != BT_STATIC) { atomicAdd(&r.edgeCount[bB], 1); atomicAdd(&cnt->nEdges, 1); }
}

__global__ void k_edge_fill(WorldView w, RefView r) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts || !islandContact(w, c)) return;
    const int bs[2] = {w.fbody[w.cfA[c]], w.fbody[w.cfB[c]]};
    for (int k = 0; k < 2; ++k) {
        const int b = bs[k];
        if (w.btype[b] == BT_STATIC) continue;
        const int slot = r.edgeStart[b] + atomicAdd(&r.edgeCursor[b], 1);
        r.edgeKeys[slot] = ((unsigned long long)(unsigned)b << 32) | (0x7FFFFFFFu - (unsigned)c);
    }
}

// Joint edges of the non-static bodies (b2Body::m_jointList): a body's
// edges sort after its contact edges, newest joint first, as the island
// search of b2World::Solve walks them (REF/Dynamics/b2World.cpp:470-531).
__device__ inline bool islandJoint(const WorldView& w, const JointView& jv, int j) {
    return (jv.flags[j] & JF_REMOVED) == 0;
}

__global__ void k_joint_edge_count(WorldView w, JointView jv, RefView r, Counters* cnt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= jv.n) return;
    r.jvisited[j] = 0;
    if (!islandJoint(w, jv, j)) return;
    const int2 b = jv.bodies[j];
    if (w.btype[b.x] != BT_STATIC) { atomicAdd(&r.edgeCount[b.x], 1); atomicAdd(&cnt->nEdges, 1); }
    if (w.btype[b.y] != BT_STATIC) { atomicAdd(&r.edgeCount[b.y], 1); atomicAdd(&cnt->nEdges, 1); }
}

__global__ void k_joint_edge_fill(WorldView w, JointView jv, RefView r) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= jv.n || !islandJoint(w, jv, j)) return;
    const int2 bb = jv.bodies[j];
    const int bs[2] = {bb.x, bb.y};
    for (int k = 0; k < 2; ++k) {
        const int b = bs[k];
        if (w.btype[b] == BT_STATIC) continue;
        const int slot = r.edgeStart[b] + atomicAdd(&r.edgeCursor[b], 1);
        r.edgeKeys[slot] = ((unsigned long long)(unsigned)b << 32) | (0xFFFFFFFFu - (unsigned)j);
    }
}

__global__ void k_edge_extract(RefView r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned low = (unsigned)(r.edgeSorted[i] & 0xFFFFFFFFull);
    r.edgeContact[i] = low < 0x80000000u ? (int)(0x7FFFFFFFu - low) : ~(int)(0xFFFFFFFFu - low);
}

// Island search of b2World::Solve (REF/Dynamics/b2World.cpp:429-539), one
// thread: contact edges, then joint edges to active bodies.
__global__ void k_dfs(WorldView w, JointView jv, RefView r, Counters* cnt) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int nI = 0, nb = 0, nc = 0, nj = 0;
    for (int seed = w.nBodies - 1; seed >= 0; --seed) {
        if (r.bvisited[seed]) continue;
        const int sf = w.bflags[seed];
        if (!(sf & BF_AWAKE) || !(sf & BF_ACTIVE)) continue;
        if (w.btype[seed] == BT_STATIC) continue;
        const int bStart = nb, cStart = nc, jStart = nj;
        int sp = 0;
        r.stack[sp++] = seed;
        r.bvisited[seed] = 1;
        while (sp > 0) {
            const int b = r.stack[--sp];
            r.islandBodies[nb++] = b;
            if ((w.bflags[b] & BF_AWAKE) == 0) {
                w.bflags[b] |= BF_AWAKE;
                w.gsleep[b].y = 0.0f;
            }
            if (w.btype[b] == BT_STATIC) continue;
            w.bflags[b] |= BF_ISLAND;
            for (int e = r.edgeStart[b]; e < r.edgeStart[b + 1]; ++e) {
                const int c = r.edgeContact[e];
                if (c < 0) {
                    const int j = ~c;
                    if (r.jvisited[j]) continue;
                    const int2 jb = jv.bodies[j];
                    const int other = jb.x == b ? jb.y : jb.x;
                    if ((w.bflags[other] & BF_ACTIVE) == 0) continue;
                    r.islandJoints[nj++] = j;
                    r.jvisited[j] = 1;
                    if (r.bvisited[other]) continue;
                    r.stack[sp++] = other;
                    r.bvisited[other] = 1;
                    continue;
                }
                if (r.cvisited[c]) continue;
                r.islandContacts[nc++] = c;
                r.cvisited[c] = 1;
                w.cflags[c] |= CF_CONSTRAINT;
                const int bA = w.fbody[w.cfA[c]];
                const int other = bA == b ? w.fbody[w.cfB[c]] : bA;
                if (r.bvisited[other]) continue;
                r.stack[sp++] = other;
                r.bvisited[other] = 1;
            }
        }
        r.islandJointRange[nI] = make_int2(jStart, nj - jStart);
        r.islandRange[nI++] = make_int4(bStart, nb - bStart, cStart, nc - cStart);
        for (int i = bStart; i < nb; ++i)
            if (w.btype[r.islandBodies[i]] == BT_STATIC) r.bvisited[r.islandBodies[i]] = 0;
    }
    cnt->nIslands = nI;
    cnt->nConstraints = nc;
}

// One wavefront per island: parallel per-body and per-constraint set-up,
// the sequential passes of b2ContactSolver on lane 0.
// Joints as b2Island::Solve orders them (REF/Dynamics/b2Island.cpp:246-328):
// their InitVelocityConstraints after the contact warm start, each velocity
// iteration joints then contacts, each position iteration contacts then
// joints, and the early exit when both are within tolerance.
__global__ void k_island_solve(WorldView w, ConView cv, JointView jv, RefView r, V2 gravity, float h, float dtRatio,
                               int warm, int velIters, int posIters, int allowSleep, Counters* cnt) {
    const int4 rg = r.islandRange[blockIdx.x];
    const int2 jr = r.islandJointRange[blockIdx.x];
    const int* joints = r.islandJoints + jr.x;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int* bodies = r.islandBodies + rg.x;
    const int k0 = rg.z, k1 = rg.z + rg.w;
    for (int i = tid; i < rg.y; i += nt) integrateVelocityOne(w, bodies[i], gravity, h);
    __syncthreads();
    const FreshView noFresh{nullptr, nullptr, nullptr};
    for (int k = k0 + tid; k < k1; k += nt)
        initConstraintOne(w, cv, k, r.islandContacts[k], dtRatio, warm, cnt, noFresh);
    __syncthreads();
    if (tid == 0) {
        if (warm)
            for (int k = k0; k < k1; ++k) warmStartOne(w, cv, k);
        for (int i = 0; i < jr.y; ++i) jointInitOne(w, jv, joints[i], dtRatio, warm);
        for (int it = 0; it < velIters; ++it) {
            for (int i = 0; i < jr.y; ++i) jointSolveVelocityOne(w, jv, joints[i], h);
            for (int k = k0; k < k1; ++k) solveVelocityOne(w, cv, k);
        }
    }
    __syncthreads();
    for (int k = k0 + tid; k < k1; k += nt) storeImpulsesOne(w, cv, k);
    for (int i = tid; i < rg.y; i += nt)
        if (w.btype[bodies[i]] != BT_STATIC) integratePositionOne(w, bodies[i], h);
    __syncthreads();
    __shared__ int solved;
    __shared__ int minSleep;
    if (tid == 0) {
        solved = 0;
        minSleep = encodeOrdered(FLT_MAX);
        for (int it = 0; it < posIters; ++it) {
            float minSep = 0.0f;
            for (int k = k0; k < k1; ++k) minSep = fminB(minSep, solvePositionOne(w, cv, k));
            bool jointsOkay = true;
            for (int i = 0; i < jr.y; ++i) {
                const bool ok = jointPositionCore(w, jv, joints[i], jv.bodies[joints[i]]);
                jointsOkay = jointsOkay && ok;
            }
            if (minSep >= -3.0f * kLinearSlop && jointsOkay) {
                solved = 1;
                break;
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < rg.y; i += nt)
        if (w.btype[bodies[i]] != BT_STATIC) syncTransformOne(w, bodies[i]);
    if (tid == 0) r.islandSlept[blockIdx.x] = 0;
    if (!allowSleep) return;
    const float linTolSqr = kLinearSleepTol * kLinearSleepTol;
    const float angTolSqr = kAngularSleepTol * kAngularSleepTol;
    for (int i = tid; i < rg.y; i += nt) {
        const int b = bodies[i];
        if (w.btype[b] == BT_STATIC) continue;
        const float4 v = w.vel[b];
        float t;
        if ((w.bflags[b] & BF_AUTOSLEEP) == 0 || v.z * v.z > angTolSqr ||
            dot(mk(v.x, v.y), mk(v.x, v.y)) > linTolSqr) {
            t = 0.0f;
        } else {
            t = w.gsleep[b].y + h;
        }
        w.gsleep[b].y = t;
        atomicMin(&minSleep, encodeOrdered(t));
    }
    __syncthreads();
    const bool sleeps = decodeOrdered(minSleep) >= kTimeToSleep && solved;
    if (tid == 0) r.islandSlept[blockIdx.x] = sleeps ? 1 : 0;
    if (sleeps) {
        for (int i = tid; i < rg.y; i += nt) {
            const int b = bodies[i];
            if (w.btype[b] == BT_STATIC) continue;  // shared static bodies: see k_static_awake
            w.bflags[b] &= ~BF_AWAKE;
            w.gsleep[b].y = 0.0f;
            w.vel[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            w.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// A static body joins every island that touches it; the reference wakes it
// when each island is built and puts it to sleep with its island, one island
// after another, so its final awake flag is that of the last island in
// search order that holds it.
__global__ void k_static_last_island(WorldView w, RefView r, int nIslands) {
    const int isl = blockIdx.x;
    if (isl >= nIslands) return;
    const int4 rg = r.islandRange[isl];
    for (int i = threadIdx.x; i < rg.y; i += blockDim.x) {
        const int b = r.islandBodies[rg.x + i];
        if (w.btype[b] == BT_STATIC) atomicMax(&r.lastIsland[b], isl);
    }
}

__global__ void k_static_awake(WorldView w, RefView r) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies || w.btype[b] != BT_STATIC) return;
    const int isl = r.lastIsland[b];
    if (isl < 0) return;
    if (r.islandSlept[isl]) w.bflags[b] &= ~BF_AWAKE;
    r.lastIsland[b] = -1;
}

// ====================================================================
// Batch kernels for per-function parity
// ====================================================================
// Per-pair distance and time of impact for the C ABI's batch entry points.
__global__ void k_distance_batch(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                                 const b2g_transform* xfB, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Proxy pA = proxyOf(a[i]), pB = proxyOf(b[i]);
    Xf ta, tb;
    ta.p = mk(xfA[i].px, xfA[i].py); ta.q.s = xfA[i].s; ta.q.c = xfA[i].c;
    tb.p = mk(xfB[i].px, xfB[i].py); tb.q.s = xfB[i].s; tb.q.c = xfB[i].c;
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    out[2 * i] = gjkDistance(cache, pA, ta, pB, tb);
    out[2 * i + 1] = (float)cache.count;
}

__global__ void k_toi_batch(const b2g_shape* a, const b2g_sweep* sa, const b2g_shape* b, const b2g_sweep* sb, int n,
                            float tMax, int32_t* state, float* t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sweep s[2];
    const b2g_sweep* in[2] = {sa + i, sb + i};
    for (int k = 0; k < 2; ++k) {
        s[k].localCenter = mk(in[k]->localCenter[0], in[k]->localCenter[1]);
        s[k].c0 = mk(in[k]->c0[0], in[k]->c0[1]);
        s[k].c = mk(in[k]->c[0], in[k]->c[1]);
        s[k].a0 = in[k]->a0;
        s[k].a = in[k]->a;
        s[k].alpha0 = in[k]->alpha0;
    }
    float ti;
    state[i] = timeOfImpact(&ti, proxyOf(a[i]), proxyOf(b[i]), s[0], s[1], tMax);
    t[i] = ti;
}

__global__ void k_collide_batch(const b2g_shape* A, const b2g_transform* xa, const b2g_shape* B,
                                const b2g_transform* xb, int n, b2g_manifold* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b2g_manifold m;
    zeroManifold(m);
    Xf xfA, xfB;
    xfA.p = mk(xa[i].px, xa[i].py); xfA.q.s = xa[i].s; xfA.q.c = xa[i].c;
    xfB.p = mk(xb[i].px, xb[i].py); xfB.q.s = xb[i].s; xfB.q.c = xb[i].c;
    if (A[i].type == B2G_POLYGON && B[i].type == B2G_POLYGON) return;  // k_collide_batch_tiles
    collideShapes(m, A[i], xfA, B[i], xfB);
    if (m.pointCount == 0) zeroManifold(m);
    out[i] = m;
}

// The polygon pairs of a batch, eight lanes per pair (collidePolygonsTile).
__global__ void __launch_bounds__(kTileBlock) k_collide_batch_tiles(const b2g_shape* A, const b2g_transform* xa,
                                                                    const b2g_shape* B, const b2g_transform* xb, int n,
                                                                    b2g_manifold* out) {
    __shared__ b2g_shape sh[kTileBlock / kTileLanes][2];
    const int grp = threadIdx.x / kTileLanes, lane = threadIdx.x % kTileLanes;
    const int i = blockIdx.x * (kTileBlock / kTileLanes) + grp;
    const bool mine = i < n && A[i].type == B2G_POLYGON && B[i].type == B2G_POLYGON;
    if (mine) stageShapes(sh[grp], A[i], B[i], lane);
    __syncthreads();
    if (!mine) return;
    Xf xfA, xfB;
    xfA.p = mk(xa[i].px, xa[i].py); xfA.q.s = xa[i].s; xfA.q.c = xa[i].c;
    xfB.p = mk(xb[i].px, xb[i].py); xfB.q.s = xb[i].s; xfB.q.c = xb[i].c;
    b2g_manifold m;
    zeroManifold(m);
    collidePolygonsTile(m, sh[grp][0], xfA, sh[grp][1], xfB, lane);
    if (m.pointCount == 0) zeroManifold(m);
    if (lane == 0) out[i] = m;
}

__global__ void k_world_manifold_batch(const b2g_manifold* m, const b2g_transform* xa, const float* ra,
                                       const b2g_transform* xb, const float* rb, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WorldManifold wm;
    wm.normal = mk(0.0f, 0.0f);
    wm.points[0] = wm.points[1] = mk(0.0f, 0.0f);
    wm.separations[0] = wm.separations[1] = 0.0f;
    Xf xfA, xfB;
    xfA.p = mk(xa[i].px, xa[i].py); xfA.q.s = xa[i].s; xfA.q.c = xa[i].c;
    xfB.p = mk(xb[i].px, xb[i].py); xfB.q.s = xb[i].s; xfB.q.c = xb[i].c;
    worldManifold(wm, m[i], xfA, ra[i], xfB, rb[i]);
    float* o = out + 8 * i;
    o[0] = wm.normal.x; o[1] = wm.normal.y;
    o[2] = wm.points[0].x; o[3] = wm.points[0].y; o[4] = wm.points[1].x; o[5] = wm.points[1].y;
    o[6] = wm.separations[0]; o[7] = wm.separations[1];
}

__global__ void k_fat_rule_batch(const float4* fat, const float4* tight, const float2* disp, int n, float4* out,
                                 uint8_t* moved) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 f = fat[i];
    moved[i] = fatRule(f, tight[i], mk(disp[i].x, disp[i].y)) ? 1 : 0;
    out[i] = f;
}

__global__ void k_sincos_batch(const float* a, int n, float* s, float* c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rot q = rotOf(a[i]);
    s[i] = q.s;
    c[i] = q.c;
}

__global__ void k_penetration(WorldView w, double* acc, int* npts, unsigned long long* maxBits) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    if (!(w.cflags[c] & CF_TOUCHING)) return;
    const int fA = w.cfA[c], fB = w.cfB[c];
    const b2g_manifold& m = w.cman[c];
    WorldManifold wm;
    worldManifold(wm, m, xfOf(w.xf[w.fbody[fA]]), w.shape[fA].radius, xfOf(w.xf[w.fbody[fB]]), w.shape[fB].radius);
    for (int i = 0; i < m.pointCount; ++i) {
        const double p = wm.separations[i] < 0.0f ? -(double)wm.separations[i] : 0.0;
        atomicAdd(acc, p);
        atomicMax(maxBits, (unsigned long long)__double_as_longlong(p));
        atomicAdd(npts, 1);
    }
}

// ====================================================================
// Continuous collision: b2World::SolveTOI, REF/Dynamics/b2World.cpp:579-897.
// One event per pass of the host loop: the earliest time of impact over the
// contact list (list order = reverse table order breaks ties), the two
// bodies moved to it, a mini-island of them and the static, kinematic and
// bullet bodies they touch (b2Island::SolveTOI, REF/Dynamics/b2Island.cpp:
// 385-515) solved for the rest of the step, fixtures synchronised and new
// pairs found. Times of impact are computed in parallel when every pending
// contact's bodies share a sweep start (always on the first scan of a
// step); otherwise in list order by one lane, since aligning two sweeps
// moves the later-starting one for the contacts that follow.
// ====================================================================
constexpr int kToiGatherCap = 2048;  // edges of the TOI pair sorted in LDS
constexpr int kToiMovedCap = 256;    // proxies one event may move before the grid broad-phase takes over
constexpr int kToiNewCap = 1024;     // new pairs one event may create likewise

struct ToiState {
    // read back after every event
    unsigned long long minKey;  // (time bits << 32) | list position of the earliest TOI
    int mismatch;               // a pending contact's sweeps start at different times
    int active;                 // the event built an island
    int done;                   // no event left below 1 - 10 eps
    int events;                 // events solved in the pass
    int nMoved;                 // proxies moved by the event (list below, or overflow)
    int nNew;                   // new pairs found for them
    int fallback;               // lists overflowed: the host runs the grid broad-phase
    int live;                   // contacts in the table (new pairs append on the device)
    int added;                  // contacts appended by the pass
    int halt;                   // later events of a batch do nothing (done, fallback, sub-step)
    int subStop;                // sub-stepping: the pass ends after one island event
    // device only
    int bodyA, bodyB, contact;
    float alpha;
    int nGather;
    int nBodies, nContacts;     // island sizes
    int bodies[2 * kMaxTOIContacts];
    int contacts[kMaxTOIContacts];
    float4 backup[4];           // pos0 / pos of body A and B before the advance
    int movedList[kToiMovedCap];
    unsigned long long newKeys[kToiNewCap];
    Counters scratch;           // counters the constraint set-up writes
};
constexpr size_t kToiReadback = offsetof(ToiState, bodyA);

__device__ inline bool toiPending(const WorldView& w, int c, int flags) {
    if (!(flags & CF_ENABLED)) return false;
    if (w.ctoiCount[c] > kMaxSubSteps) return false;
    if (flags & CF_TOI) return false;
    const int fA = w.cfA[c], fB = w.cfB[c];
    if (sensorOf(w, fA) || sensorOf(w, fB)) return false;
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    const int tA = w.btype[bA], tB = w.btype[bB];
    const bool activeA = (w.bflags[bA] & BF_AWAKE) && tA != BT_STATIC;
    const bool activeB = (w.bflags[bB] & BF_AWAKE) && tB != BT_STATIC;
    if (!activeA && !activeB) return false;
    const bool collideA = (w.bflags[bA] & BF_BULLET) || tA != BT_DYNAMIC;
    const bool collideB = (w.bflags[bB] & BF_BULLET) || tB != BT_DYNAMIC;
    return collideA || collideB;
}

__device__ inline Sweep sweepOf(const WorldView& w, int b) {
    const float4 p0 = w.pos0[b], p = w.pos[b];
    const float2 lc = w.lcenter[b];
    Sweep sw;
    sw.localCenter = mk(lc.x, lc.y);
    sw.c0 = mk(p0.x, p0.y);
    sw.a0 = p0.z;
    sw.alpha0 = p0.w;
    sw.c = mk(p.x, p.y);
    sw.a = p.z;
    return sw;
}

// b2Sweep::Advance on a body's sweep start.
__device__ inline void advanceSweepStart(const WorldView& w, int b, float alpha) {
    Sweep sw = sweepOf(w, b);
    sweepAdvance(sw, alpha);
    w.pos0[b] = make_float4(sw.c0.x, sw.c0.y, sw.a0, sw.alpha0);
}

// b2Body::Advance, REF/Dynamics/b2Body.h:842-850: the body to its sweep at alpha.
__device__ inline void advanceBodyTo(const WorldView& w, int b, float alpha) {
    advanceSweepStart(w, b, alpha);
    const float4 p0 = w.pos0[b];
    w.pos[b] = make_float4(p0.x, p0.y, p0.z, 0.0f);
    syncTransformOne(w, b);
}

// The two sweeps of a contact brought to a common start: the one that
// starts earlier is advanced (and stays advanced), REF/Dynamics/b2World.cpp:
// 645-660.
__device__ inline void alignSweeps(const WorldView& w, int c) {
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    const float alphaA = w.pos0[bA].w, alphaB = w.pos0[bB].w;
    if (alphaA < alphaB) advanceSweepStart(w, bA, alphaB);
    else if (alphaB < alphaA) advanceSweepStart(w, bB, alphaA);
}

// TOI of a contact over the rest of the step from aligned sweeps,
// REF/Dynamics/b2World.cpp:662-689.
__device__ inline float toiFromSweeps(const WorldView& w, int c, const Sweep& sA, const Sweep& sB) {
    const Proxy pA = proxyOf(w.shape[w.cfA[c]]), pB = proxyOf(w.shape[w.cfB[c]]);
    float beta;
    const int state = timeOfImpact(&beta, pA, pB, sA, sB, 1.0f);
    const float alpha0 = sA.alpha0;
    return state == TOI_TOUCHING ? fminB(alpha0 + (1.0f - alpha0) * beta, 1.0f) : 1.0f;
}

// b2Contact::Update without a listener (REF/Dynamics/Contacts/b2Contact.cpp:
// 171-251) for one contact on one lane.
__device__ inline void toiUpdateContact(const WorldView& w, int c) {
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int bA = w.fbody[fA], bB = w.fbody[fB];
    int flags = w.cflags[c] | CF_ENABLED;
    const b2g_manifold old = w.cman[c];
    b2g_manifold m;
    zeroManifold(m);
    collideShapes(m, w.shape[fA], xfOf(w.xf[bA]), w.shape[fB], xfOf(w.xf[bB]));
    for (int i = 0; i < m.pointCount; ++i) {
        m.points[i].normalImpulse = 0.0f;
        m.points[i].tangentImpulse = 0.0f;
        for (int j = 0; j < old.pointCount; ++j) {
            if (old.points[j].id == m.points[i].id) {
                m.points[i].normalImpulse = old.points[j].normalImpulse;
                m.points[i].tangentImpulse = old.points[j].tangentImpulse;
                break;
            }
        }
    }
    const bool wasTouching = (flags & CF_TOUCHING) != 0;
    const bool touching = m.pointCount > 0;
    if (touching != wasTouching) {
        wakeBody(w, bA);
        wakeBody(w, bB);
    }
    if (touching) flags |= CF_TOUCHING;
    else flags &= ~CF_TOUCHING;
    w.cman[c] = m;
    w.cflags[c] = flags;
}

// b2ContactSolver::SolveTOIPositionConstraints, REF/Dynamics/Contacts/
// b2ContactSolver.cpp:753-841, one constraint: only the TOI pair moves.
__device__ inline float solveTOIPositionOne(const WorldView& w, const ConView& cv, int k, int toiA, int toiB) {
    const int2 bi = cv.bodies[k];
    const float4 ms = cv.mass[k];
    const bool movesA = bi.x == toiA || bi.x == toiB, movesB = bi.y == toiA || bi.y == toiB;
    const float mA = movesA ? ms.x : 0.0f, iA = movesA ? ms.y : 0.0f;
    const float mB = movesB ? ms.z : 0.0f, iB = movesB ? ms.w : 0.0f;
    const int info = cv.info[k];
    const int pc = (info >> 2) & 3;
    const int type = (info >> 8) & 3;
    const float4 lc = cv.lc[k];
    const V2 lcA = mk(lc.x, lc.y), lcB = mk(lc.z, lc.w);
    const float4 lp = cv.lp[k];
    const float4 lnp = cv.lnp[k];
    const float2 rad = cv.rad[k];
    const float4 pA = w.pos[bi.x], pB = w.pos[bi.y];
    V2 cA = mk(pA.x, pA.y), cB = mk(pB.x, pB.y);
    float aA = pA.z, aB = pB.z;
    float minSep = 0.0f;
    for (int j = 0; j < pc; ++j) {
        Xf xfA, xfB;
        xfA.q = rotOf(aA);
        xfB.q = rotOf(aB);
        xfA.p = cA - mul(xfA.q, lcA);
        xfB.p = cB - mul(xfB.q, lcB);
        V2 normal, point;
        float separation;
        if (type == B2G_MANIFOLD_CIRCLES) {
            const V2 pointA = mul(xfA, mk(lnp.z, lnp.w));
            const V2 pointB = mul(xfB, mk(lp.x, lp.y));
            normal = pointB - pointA;
            normalizeV(normal);
            point = 0.5f * (pointA + pointB);
            separation = dot(pointB - pointA, normal) - rad.x - rad.y;
        } else if (type == B2G_MANIFOLD_FACE_A) {
            normal = mul(xfA.q, mk(lnp.x, lnp.y));
            const V2 planePoint = mul(xfA, mk(lnp.z, lnp.w));
            const V2 clipPoint = mul(xfB, j == 0 ? mk(lp.x, lp.y) : mk(lp.z, lp.w));
            separation = dot(clipPoint - planePoint, normal) - rad.x - rad.y;
            point = clipPoint;
        } else {
            normal = mul(xfB.q, mk(lnp.x, lnp.y));
            const V2 planePoint = mul(xfB, mk(lnp.z, lnp.w));
            const V2 clipPoint = mul(xfA, j == 0 ? mk(lp.x, lp.y) : mk(lp.z, lp.w));
            separation = dot(clipPoint - planePoint, normal) - rad.x - rad.y;
            point = clipPoint;
            normal = -normal;
        }
        const V2 rA = point - cA;
        const V2 rB = point - cB;
        minSep = fminB(minSep, separation);
        const float C = clampB(kToiBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
        const float rnA = cross(rA, normal);
        const float rnB = cross(rB, normal);
        const float Kn = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
        const float impulse = Kn > 0.0f ? -C / Kn : 0.0f;
        const V2 P = impulse * normal;
        cA -= mA * P;
        aA -= iA * cross(rA, P);
        cB += mB * P;
        aB += iB * cross(rB, P);
    }
    if (movesA) w.pos[bi.x] = make_float4(cA.x, cA.y, aA, 0.0f);
    if (movesB) w.pos[bi.y] = make_float4(cB.x, cB.y, aB, 0.0f);
    return minSep;
}

// Start of a TOI pass after a completed step: island flags off, sweeps
// start at 0, cached TOIs dropped.
__global__ void k_toi_reset_bodies(WorldView w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= w.nBodies) return;
    w.bflags[b] &= ~BF_ISLAND;
    w.pos0[b].w = 0.0f;
}
__global__ void k_toi_reset_contacts(WorldView w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w.nContacts) return;
    w.cflags[c] &= ~(CF_TOI | CF_ISLAND);
    w.ctoiCount[c] = 0;
    w.ctoi[c] = 1.0f;
}

// Per-event fields cleared for the next event (by the last kernel of an
// event, and when a pass starts or resumes).
__device__ inline void toiResetEvent(ToiState* st) {
    st->minKey = ~0ull;
    st->mismatch = 0;
    st->active = 0;
    st->nGather = 0;
    st->nBodies = 0;
    st->nContacts = 0;
    st->nMoved = 0;
    st->nNew = 0;
    st->fallback = 0;
}

// Contacts whose TOI the scan computes, and whether their sweeps need aligning.
__global__ void k_toi_classify(WorldView w, int* pending, ToiState* st) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st->live || st->halt) return;
    const bool p = toiPending(w, c, w.cflags[c]);
    pending[c] = p ? 1 : 0;
    if (p && w.pos0[w.fbody[w.cfA[c]]].w != w.pos0[w.fbody[w.cfB[c]]].w) atomicOr(&st->mismatch, 1);
}


// Otherwise the alignments run in list order on one lane, since an advance
// moves a body for the contacts after it, and each contact keeps the two
// sweeps it saw; the TOIs then run in parallel from those.
__global__ void k_toi_align_serial(WorldView w, const int* pending, const ToiState* st, Sweep* seen) {
    if (blockIdx.x != 0 || threadIdx.x != 0 || st->halt || !st->mismatch) return;
    for (int c = st->live - 1; c >= 0; --c) {
        if (!pending[c]) continue;
        alignSweeps(w, c);
        seen[2 * c] = sweepOf(w, w.fbody[w.cfA[c]]);
        seen[2 * c + 1] = sweepOf(w, w.fbody[w.cfB[c]]);
    }
}

// One lane per pending contact: from the current sweeps when every pending
// pair's sweeps already start together (each TOI then depends on its own two
// bodies only), else from the sweeps the list-order alignment recorded.
__global__ void k_toi_compute(WorldView w, const int* pending, const ToiState* st, const Sweep* seen) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st->live || st->halt || !pending[c]) return;
    if (st->mismatch) {
        w.ctoi[c] = toiFromSweeps(w, c, seen[2 * c], seen[2 * c + 1]);
    } else {
        const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
        w.ctoi[c] = toiFromSweeps(w, c, sweepOf(w, bA), sweepOf(w, bB));
    }
    w.cflags[c] |= CF_TOI;
}

// Earliest cached TOI below 1, first in list order among equal times.
__global__ void k_toi_min(WorldView w, ToiState* st) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (st->halt) return;
    unsigned long long key = ~0ull;
    if (c < st->live) {
        const int flags = w.cflags[c];
        if ((flags & CF_ENABLED) && (flags & CF_TOI) && w.ctoiCount[c] <= kMaxSubSteps) {
            const float alpha = w.ctoi[c];
            if (alpha < 1.0f)
                key = ((unsigned long long)__float_as_uint(alpha) << 32) | (unsigned)(st->live - 1 - c);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(key, off);
        key = o < key ? o : key;
    }
    if ((threadIdx.x & 63) == 0 && key != ~0ull) atomicMin(&st->minKey, key);
}

// The event's TOI pair: advance both bodies, update the contact; an island
// starts when it touches.
__global__ void k_toi_event(WorldView w, ToiState* st) {
    if (blockIdx.x != 0 || threadIdx.x != 0 || st->halt) return;
    const unsigned long long key = st->minKey;
    const float alpha = key == ~0ull ? 1.0f : __uint_as_float((unsigned)(key >> 32));
    if (key == ~0ull || 1.0f - 10.0f * FLT_EPSILON < alpha) {
        st->done = 1;
        st->halt = 1;
        return;
    }
    const int c = st->live - 1 - (int)(key & 0xFFFFFFFFu);
    const int bA = w.fbody[w.cfA[c]], bB = w.fbody[w.cfB[c]];
    st->backup[0] = w.pos0[bA];
    st->backup[1] = w.pos[bA];
    st->backup[2] = w.pos0[bB];
    st->backup[3] = w.pos[bB];
    advanceBodyTo(w, bA, alpha);
    advanceBodyTo(w, bB, alpha);
    toiUpdateContact(w, c);
    int flags = w.cflags[c] & ~CF_TOI;
    w.ctoiCount[c] += 1;
    if (!(flags & CF_ENABLED) || !(flags & CF_TOUCHING)) {
        w.cflags[c] = flags & ~CF_ENABLED;
        w.pos0[bA] = st->backup[0];
        w.pos[bA] = st->backup[1];
        w.pos0[bB] = st->backup[2];
        w.pos[bB] = st->backup[3];
        syncTransformOne(w, bA);
        syncTransformOne(w, bB);
        return;
    }
    w.cflags[c] = flags | CF_ISLAND;
    wakeBody(w, bA);
    wakeBody(w, bB);
    w.bflags[bA] |= BF_ISLAND;
    w.bflags[bB] |= BF_ISLAND;
    st->bodyA = bA;
    st->bodyB = bB;
    st->contact = c;
    st->alpha = alpha;
    st->bodies[0] = bA;
    st->bodies[1] = bB;
    st->nBodies = 2;
    st->contacts[0] = c;
    st->nContacts = 1;
    st->active = 1;
    st->events += 1;
}

// Contact edges of the TOI pair's dynamic bodies, keyed for list order:
// (0 for A / 1 for B) << 62 | list position << 11 | gather slot.
constexpr int kGatherSlotBits = 11;
static_assert((1 << kGatherSlotBits) >= kToiGatherCap, "gather slot field");
__global__ void k_toi_gather(WorldView w, ToiState* st, unsigned long long* gather) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st->live || st->halt || !st->active) return;
    const int x = w.fbody[w.cfA[c]], y = w.fbody[w.cfB[c]];
    const unsigned pos = (unsigned)(st->live - 1 - c);
    for (int k = 0; k < 2; ++k) {
        const int body = k == 0 ? st->bodyA : st->bodyB;
        if (w.btype[body] != BT_DYNAMIC || (x != body && y != body)) continue;
        const int slot = atomicAdd(&st->nGather, 1);
        if (slot < kToiGatherCap)
            gather[slot] = ((unsigned long long)k << 62) | ((unsigned long long)pos << kGatherSlotBits) | (unsigned)slot;
    }
}

// One contact edge of the island search (REF/Dynamics/b2World.cpp:760-843).
__device__ inline void toiVisitEdge(const WorldView& w, ToiState* st, int body, int c) {
    if (w.cflags[c] & CF_ISLAND) return;
    const int x = w.fbody[w.cfA[c]], y = w.fbody[w.cfB[c]];
    const int other = x == body ? y : x;
    if (w.btype[other] == BT_DYNAMIC && !(w.bflags[body] & BF_BULLET) && !(w.bflags[other] & BF_BULLET)) return;
    if (sensorOf(w, w.cfA[c]) || sensorOf(w, w.cfB[c])) return;
    const float4 save0 = w.pos0[other], save = w.pos[other];
    if (!(w.bflags[other] & BF_ISLAND)) advanceBodyTo(w, other, st->alpha);
    toiUpdateContact(w, c);
    const int flags = w.cflags[c];
    if (!(flags & CF_ENABLED) || !(flags & CF_TOUCHING)) {
        w.pos0[other] = save0;
        w.pos[other] = save;
        syncTransformOne(w, other);
        return;
    }
    w.cflags[c] = flags | CF_ISLAND;
    st->contacts[st->nContacts++] = c;
    if (w.bflags[other] & BF_ISLAND) return;
    w.bflags[other] |= BF_ISLAND;
    if (w.btype[other] != BT_STATIC) wakeBody(w, other);
    st->bodies[st->nBodies++] = other;
}

// The narrowphase of every gathered edge ahead of the island search, one
// lane each. What b2Contact::Update computes for an edge depends only on the
// two bodies' transforms at the visit, and those are known in advance: the
// TOI pair as advanced, any other body as advanced to the event's time
// (kept if its contact touches, else restored, so a later edge of the same
// body sees the same advance again); a contact met twice (between the two
// TOI bodies) is not touching at either visit or is skipped at the second.
__global__ void k_toi_prep(WorldView w, ToiState* st, const unsigned long long* gather, b2g_manifold* man,
                           int* touching) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (st->halt || !st->active || i >= st->nGather || i >= kToiGatherCap) return;
    const unsigned long long key = gather[i];
    const int k = (int)(key >> 62);
    const int pos = (int)((key >> kGatherSlotBits) & ((1ull << 51) - 1));
    const int c = st->live - 1 - pos;
    const int body = k == 0 ? st->bodyA : st->bodyB;
    const int fA = w.cfA[c], fB = w.cfB[c];
    const int x = w.fbody[fA], y = w.fbody[fB];
    const int other = x == body ? y : x;
    Xf xfOther;
    if (w.bflags[other] & BF_ISLAND) {
        xfOther = xfOf(w.xf[other]);
    } else {
        Sweep sw = sweepOf(w, other);
        sweepAdvance(sw, st->alpha);
        xfOther.q = rotOf(sw.a0);
        xfOther.p = sw.c0 - mul(xfOther.q, sw.localCenter);
    }
    const Xf xfBody = xfOf(w.xf[body]);
    const Xf xfX = x == body ? xfBody : xfOther;
    const Xf xfY = y == body ? xfBody : xfOther;
    const b2g_manifold old = w.cman[c];
    b2g_manifold m;
    zeroManifold(m);
    collideShapes(m, w.shape[fA], xfX, w.shape[fB], xfY);
    for (int p = 0; p < m.pointCount; ++p) {
        m.points[p].normalImpulse = 0.0f;
        m.points[p].tangentImpulse = 0.0f;
        for (int j = 0; j < old.pointCount; ++j) {
            if (old.points[j].id == m.points[p].id) {
                m.points[p].normalImpulse = old.points[j].normalImpulse;
                m.points[p].tangentImpulse = old.points[j].tangentImpulse;
                break;
            }
        }
    }
    const int slot = (int)(key & ((1u << kGatherSlotBits) - 1));
    man[slot] = m;
    touching[slot] = m.pointCount > 0 ? 1 : 0;
}

// One edge of the island search with its narrowphase done ahead: commits
// what toiVisitEdge computes.
__device__ inline void toiCommitEdge(const WorldView& w, ToiState* st, int body, int c, const b2g_manifold& m,
                                     bool touching) {
    if (w.cflags[c] & CF_ISLAND) return;
    const int x = w.fbody[w.cfA[c]], y = w.fbody[w.cfB[c]];
    const int other = x == body ? y : x;
    if (w.btype[other] == BT_DYNAMIC && !(w.bflags[body] & BF_BULLET) && !(w.bflags[other] & BF_BULLET)) return;
    if (sensorOf(w, w.cfA[c]) || sensorOf(w, w.cfB[c])) return;
    if (touching && !(w.bflags[other] & BF_ISLAND)) advanceBodyTo(w, other, st->alpha);
    const int flags = w.cflags[c] | CF_ENABLED;
    const bool wasTouching = (flags & CF_TOUCHING) != 0;
    if (touching != wasTouching) {
        wakeBody(w, x);
        wakeBody(w, y);
    }
    w.cman[c] = m;
    if (!touching) {
        w.cflags[c] = flags & ~CF_TOUCHING;
        syncTransformOne(w, other);  // the restored sweep's transform
        return;
    }
    w.cflags[c] = flags | CF_TOUCHING | CF_ISLAND;
    st->contacts[st->nContacts++] = c;
    if (w.bflags[other] & BF_ISLAND) return;
    w.bflags[other] |= BF_ISLAND;
    if (w.btype[other] != BT_STATIC) wakeBody(w, other);
    st->bodies[st->nBodies++] = other;
}

// Island search over the TOI pair's edges in list order, then
// b2Island::SolveTOI on lane 0 (Gauss-Seidel order is the island's).
__global__ __launch_bounds__(256) void k_toi_island(WorldView w, ConView cv, ToiState* st,
                                                    const unsigned long long* gather, const b2g_manifold* man,
                                                    const int* touching, float dt, int velIters) {
    __shared__ unsigned long long keys[kToiGatherCap];
    if (st->halt || !st->active) return;
    const int n = st->nGather;
    const bool sorted = n <= kToiGatherCap;
    if (sorted) {
        int size = 1;
        while (size < n) size <<= 1;
        for (int i = threadIdx.x; i < size; i += blockDim.x) keys[i] = i < n ? gather[i] : ~0ull;
        __syncthreads();
        for (int k = 2; k <= size; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < size; i += blockDim.x) {
                    const int l = i ^ j;
                    if (l > i) {
                        const unsigned long long a = keys[i], b = keys[l];
                        const bool up = (i & k) == 0;
                        if ((a > b) == up) {
                            keys[i] = b;
                            keys[l] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    if (threadIdx.x != 0) return;
    const int bA = st->bodyA, bB = st->bodyB;
    const int nc = st->live;
    if (sorted) {
        for (int i = 0; i < n; ++i) {
            if (st->nBodies == 2 * kMaxTOIContacts || st->nContacts == kMaxTOIContacts) break;
            const unsigned long long key = keys[i];
            const int k = (int)(key >> 62);
            const int c = nc - 1 - (int)((key >> kGatherSlotBits) & ((1ull << 51) - 1));
            const int slot = (int)(key & ((1u << kGatherSlotBits) - 1));
            toiCommitEdge(w, st, k == 0 ? bA : bB, c, man[slot], touching[slot] != 0);
        }
    } else {
        // more edges than the sort holds: walk the list once per body
        for (int k = 0; k < 2; ++k) {
            const int body = k == 0 ? bA : bB;
            if (w.btype[body] != BT_DYNAMIC) continue;
            for (int c = nc - 1; c >= 0; --c) {
                if (st->nBodies == 2 * kMaxTOIContacts || st->nContacts == kMaxTOIContacts) break;
                if (w.fbody[w.cfA[c]] != body && w.fbody[w.cfB[c]] != body) continue;
                toiVisitEdge(w, st, body, c);
            }
        }
    }
    // b2Island::SolveTOI
    const FreshView noFresh{nullptr, nullptr, nullptr, 0};
    const int ncons = st->nContacts;
    for (int k = 0; k < ncons; ++k) initConstraintOne(w, cv, k, st->contacts[k], 1.0f, 0, &st->scratch, noFresh);
    for (int it = 0; it < 20; ++it) {
        float minSep = 0.0f;
        for (int k = 0; k < ncons; ++k) minSep = fminB(minSep, solveTOIPositionOne(w, cv, k, bA, bB));
        if (minSep >= -1.5f * kLinearSlop) break;
    }
    const float4 pa = w.pos[bA], pb = w.pos[bB];
    w.pos0[bA] = make_float4(pa.x, pa.y, pa.z, w.pos0[bA].w);
    w.pos0[bB] = make_float4(pb.x, pb.y, pb.z, w.pos0[bB].w);
    for (int k = 0; k < ncons; ++k) initConstraintOne(w, cv, k, st->contacts[k], 1.0f, 0, &st->scratch, noFresh);
    for (int it = 0; it < velIters; ++it)
        for (int k = 0; k < ncons; ++k) solveVelocityOne(w, cv, k);
    const float h = (1.0f - st->alpha) * dt;
    for (int i = 0; i < st->nBodies; ++i) {
        const int b = st->bodies[i];
        integratePositionOne(w, b, h);
        syncTransformOne(w, b);
    }
}

// Fixtures of the island's dynamic bodies to their swept boxes, and the
// cached TOIs and island marks of every contact on such a body dropped.
__global__ void k_toi_displaced(WorldView w, ToiState* st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (st->halt || !st->active) return;
    if (i < w.nFixtures) {
        const int b = w.fbody[i];
        if ((w.bflags[b] & BF_ISLAND) && w.btype[b] == BT_DYNAMIC && syncFixtureOne(w, i, b)) {
            const int slot = atomicAdd(&st->nMoved, 1);
            if (slot < kToiMovedCap) st->movedList[slot] = i;
        }
    }
    if (i < st->live) {
        const int x = w.fbody[w.cfA[i]], y = w.fbody[w.cfB[i]];
        const bool hit = ((w.bflags[x] & BF_ISLAND) && w.btype[x] == BT_DYNAMIC) ||
                         ((w.bflags[y] & BF_ISLAND) && w.btype[y] == BT_DYNAMIC);
        if (hit) w.cflags[i] &= ~(CF_TOI | CF_ISLAND);
    }
}

// b2BroadPhase::UpdatePairs for the event's moved proxies only: every proxy
// whose fat box overlaps a moved one's, a pair of two moved proxies found by
// the lower id's thread, filtered as AddPair filters.
__global__ void k_toi_pairs(WorldView w, ToiState* st, const unsigned long long* table, uint32_t hmask) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= w.nFixtures || st->halt || !st->active) return;
    const int nm = st->nMoved;
    if (nm == 0 || nm > kToiMovedCap) return;
    const float4 ff = w.fat[f];
    const bool fMoved = w.moved[f] != 0;
    for (int i = 0; i < nm; ++i) {
        const int q = st->movedList[i];
        if (q == f || (fMoved && f > q)) continue;
        if (!overlaps(ff, w.fat[q])) continue;
        const int a = f < q ? f : q, b = f < q ? q : f;
        if (!acceptPair(w, a, b, table, hmask)) continue;
        if (orderOf(w, a) < 0 || orderOf(w, b) < 0) continue;  // no proxy
        const int slot = atomicAdd(&st->nNew, 1);
        if (slot < kToiNewCap) st->newKeys[slot] = pairKey(orderOf(w, a), orderOf(w, b));
    }
}

// New contacts of the event in pair order (sorted keys, as the reference
// sorts its pair buffer), appended to the table and its hash; the moved
// flags cleared. Lists that overflowed leave everything to the host's grid
// broad-phase instead.
__global__ __launch_bounds__(256) void k_toi_append(WorldView w, ToiState* st, unsigned long long* table,
                                                   uint32_t hmask, int capContacts, int subStepping) {
    __shared__ unsigned long long keys[kToiNewCap];
    if (st->halt) return;
    if (!st->active) {
        __syncthreads();
        if (threadIdx.x == 0) toiResetEvent(st);
        return;
    }
    for (int i = threadIdx.x; i < st->nBodies; i += blockDim.x) w.bflags[st->bodies[i]] &= ~BF_ISLAND;
    const int nm = st->nMoved, nn = st->nNew, base = st->live;
    if (nm > kToiMovedCap || nn > kToiNewCap || base + nn > capContacts) {
        __syncthreads();
        if (threadIdx.x == 0) {
            st->fallback = 1;
            st->halt = 1;
        }
        return;
    }
    if (nn > 0) {
        int size = 1;
        while (size < nn) size <<= 1;
        for (int i = threadIdx.x; i < size; i += blockDim.x) keys[i] = i < nn ? st->newKeys[i] : ~0ull;
        __syncthreads();
        for (int k = 2; k <= size; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < size; i += blockDim.x) {
                    const int l = i ^ j;
                    if (l > i) {
                        const unsigned long long a = keys[i], b = keys[l];
                        if ((a > b) == ((i & k) == 0)) {
                            keys[i] = b;
                            keys[l] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < nn; i += blockDim.x) {
            const int a = fixtureOf(w, (int)(keys[i] >> 32)), b = fixtureOf(w, (int)(keys[i] & 0xFFFFFFFFu));
            appendContactOne(w, a, b, base + i);
            const unsigned long long key = pairKey(a, b);
            uint32_t h = hashU64(key) & hmask;
            for (;;) {
                const unsigned long long prev = atomicCAS(&table[h], ~0ull, key);
                if (prev == ~0ull || prev == key) break;
                h = (h + 1) & hmask;
            }
        }
    }
    for (int i = threadIdx.x; i < nm; i += blockDim.x) w.moved[st->movedList[i]] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        st->live = base + nn;
        st->added += nn;
        toiResetEvent(st);
        if (subStepping) {
            st->subStop = 1;
            st->halt = 1;
        }
    }
}

inline int blocks(int n, int bs = 256) { return (n + bs - 1) / bs; }

}  // namespace

// ====================================================================
// Host side
// ====================================================================
struct b2g_context {
    int device = 0;
    b2g_schedule sched{0, 0, 256, 256, 256, 1024, 0, 0};
    hipStream_t stream = nullptr;
    std::string err;
    int cuCount = 0;            // compute units of the device (one-launch solve)
    int persistOcc[2][17] = {};  // resident one-launch solve workgroups per CU, without / with joints, by
                                 // workgroup size / 64

    // capacities
    int capBodies = 0, capFixtures = 0, capContacts = 0;
    int nBodies = 0, nFixtures = 0, nContacts = 0;

    // body arrays
    int* btype = nullptr;
    int* bflags = nullptr;
    float4 *xf = nullptr, *pos = nullptr, *pos0 = nullptr, *vel = nullptr, *force = nullptr, *mass = nullptr;
    float2 *lcenter = nullptr, *gsleep = nullptr;
    unsigned long long* colourMask = nullptr;
    unsigned long long* claim = nullptr;
    int* islandOf = nullptr;
    int* islandAwake = nullptr;
    int* islandSleep = nullptr;
    int* islandMax = nullptr;   // largest body index per island root
    int* staticKey = nullptr;   // per static body: see k_static_reach
    int staticKeyCap = 0;
    bool staticAsleep = false;  // a static body was uploaded asleep
    int* islandMinSep = nullptr;
    int islandIterCap = 0;
    int* bworld = nullptr;          // sub-world of each body (batched worlds)
    bool subWorlds = false;         // some body is outside sub-world 0
    // clustered solve: bodies sorted by (sub-world, Morton code of the cell)
    unsigned long long *bodyKey = nullptr, *bodyKeySorted = nullptr;
    int2* bodyCluster = nullptr;    // per body: cluster, slot (-1, -1: not clustered)
    int* clusterBodies = nullptr;   // bodies by rank
    int* clusterStart = nullptr;    // per cluster (and the boundary set): 66 constraint offsets
    int clusterStartCap = 0;

    // fixture arrays
    int* fbody = nullptr;
    b2g_shape* shape = nullptr;
    float4* fmat = nullptr;
    uint32_t* ffilter = nullptr;
    int* fgroup = nullptr;
    float4* fat = nullptr;
    uint8_t* moved = nullptr;
    int4* range = nullptr;

    // contacts (two sets for compaction)
    ContactArrays cur{}, alt{};
    int* keep = nullptr;
    int* idx = nullptr;
    unsigned long long* hashTable = nullptr;
    uint32_t hashMask = 0;
    int hashCap = 0;
    unsigned long long* newKeys = nullptr;
    unsigned long long* newKeysSorted = nullptr;
    // broad-phase order keys (b2g_set_proxy_ids): key per fixture for the
    // first orderCount fixtures, fixture per key below orderSpan; orderSpan
    // 0 while unset (the fixture index is the key)
    int* fixOrder = nullptr;
    int* orderFix = nullptr;
    int orderCount = 0, orderSpan = 0;

    // constraints
    ConView cv{};
    uint32_t* consKeys = nullptr;
    uint32_t* consKeysSorted = nullptr;
    int* consVals = nullptr;
    int* consOrder = nullptr;
    int* tileList = nullptr;   // polygon-polygon contacts of the step's narrowphase

    // grid
    int nBuckets = 0, capEntries = 0;
    int *gCounts = nullptr, *gStarts = nullptr, *gCursor = nullptr, *gEntries = nullptr, *gQueries = nullptr;
    float4* gEfat = nullptr;
    unsigned long long *gBig = nullptr, *gBigSorted = nullptr;
    unsigned long long* gCell = nullptr;
    float cellSize = 2.0f;
    bool cellSizeSet = false;

    // scratch
    void* scratch = nullptr;
    size_t scratchBytes = 0;
    Counters* cnt = nullptr;       // device
    Counters* hcnt = nullptr;      // pinned host mirror
    double* penAcc = nullptr;

    // world switches
    V2 gravity = mk(0.0f, -10.0f);
    bool allowSleep = true, warmStarting = true, continuous = false, subStepping = false, autoClearForces = true;
    bool newFixture = false;
    float inv_dt0 = 0.0f;
    uint32_t claimTag = 0;
    b2g_step_stats stats{};

    // reference-order mode buffers
    RefView ref{};
    int refCapJoints = 0;
    int refCapBodies = 0, refCapContacts = 0;

    // revolute joints
    int nJoints = 0, capJoints = 0;
    int2* jBodies = nullptr;
    float4 *jAnchors = nullptr, *jLimits = nullptr, *jImpulse = nullptr, *jR = nullptr, *jMasses = nullptr,
           *jK0 = nullptr, *jImpulseSnap = nullptr;
    float2* jK1 = nullptr;
    float *jMotorSpeed = nullptr, *jMotorMass = nullptr;
    float4* jAxis = nullptr;   // persistent: local axis of prismatic and wheel joints
    float4* jExtra = nullptr;  // persistent: two rows of step directions per joint
    int *jFlags = nullptr, *jFlagsSnap = nullptr, *jOrder = nullptr;
    std::vector<int2> hJointBodies;
    std::vector<int2> hJointBodies2;   // a gear joint's bodies C, D (-1 otherwise)
    std::vector<int> hJointKeepApart;
    std::vector<int> hJointRemoved;
    std::vector<int> jointColourStart;  // colour c: slots [start[c], start[c + 1]) of jOrder
    bool jointsDirty = false;
    unsigned long long* jointPairs = nullptr;
    uint32_t jointPairMask = 0;
    int jointPairCap = 0;
    bool anyJointPair = false;

    // learned warm start: network, fresh points of the step, training log,
    // snapshots for the fallback and the residual trace
    b2mlp::Net mlp;
    bool mlpAllocated = false;
    int capFresh = 0;               // constraint capacity of the fresh-point arrays
    int* freshMask = nullptr;
    int* freshNum = nullptr;
    int* freshOff = nullptr;
    float* freshAll = nullptr;      // [2 * capFresh][16]
    float* freshFeat = nullptr;     // compacted [2 * capFresh][16]
    int* freshRef = nullptr;        // compacted (k << 1) | j
    float4* impSnap = nullptr;      // [capFresh]
    float4* impAlt = nullptr;       // [capFresh]
    float4* velSnap = nullptr;
    int capVelSnap = 0;
    float* logFeat = nullptr;
    float* logTarget = nullptr;
    int logCap = 0, logCount = 0;
    int learnedScope = 0;   // 0 fresh points, 1 every solved point
    // Captured solves: launch sequences from the warm start to the
    // transform sync, keyed by everything their launches depend on.
    struct SolveGraph {
        std::vector<unsigned long long> key;
        hipGraphExec_t exec = nullptr;
        long long used = 0;
    };
    bool graphs = false;
    std::vector<SolveGraph> solveGraphs;
    std::vector<std::vector<unsigned long long>> recentKeys;  // keys of the last steps not captured
    long long graphClock = 0;
    int* dRanges = nullptr;  // colour ranges of the step (kColourSlots + 1 ints)
    int* hRanges = nullptr;  // pinned host staging of dRanges
    // continuous collision (b2World::SolveTOI)
    bool stepComplete = true;       // the last TOI pass ran to the end (sub-stepping may stop it early)
    ToiState* toiState = nullptr;   // device
    ToiState* hToi = nullptr;       // pinned host mirror
    unsigned long long* toiGather = nullptr;
    b2g_manifold* toiMan = nullptr; // narrowphase of the gathered edges
    int* toiTouch = nullptr;
    Sweep* toiSeen = nullptr;       // per contact: the aligned sweeps of a list-order scan
    int toiSeenCap = 0;
    ConView toiCv{};                // constraints of one TOI island
    // split step (b2g_step_collide / b2g_step_solve) for the host's PreSolve
    bool keepOldManifolds = false;
    b2g_manifold* oldMan = nullptr;
    int oldManCap = 0;
    bool stepOpen = false;
    b2g_step_params openParams{};
    long long graphLaunches = 0, graphCaptures = 0;
    int logStride = 1;      // scope 1 logs every logStride-th step
    long long logSteps = 0;
    void* trainScratch = nullptr;
    size_t trainScratchBytes = 0;
    int traceIters = 0;
    std::vector<double> trace;      // [3][traceIters + 1]
    int traceStepIters = 0;         // iterations of the last traced step
    int lastColourStart[kColourSlots + 1] = {};  // colour ranges of the last coloured step
    int demoteRounds = 1;           // colour demotion rounds per coloured step
    int compactedFrom = -1;  // table length before the last coloured step's compaction (-1: none)
    int colourRoundsNeeded = 0;  // colouring rounds the last coloured step needed
    int lastConstraints = 0;

    // profiling (HIP events around launches)
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    int evUsed = 0;
    std::vector<std::pair<int, int>> evMarks;  // (phase, start event index)
    double profMs[32] = {};
    int profLaunches[32] = {};

    JointView jview() const {
        JointView j;
        j.n = nJoints;
        j.order = jOrder;
        j.bodies = jBodies; j.anchors = jAnchors; j.limits = jLimits; j.motorSpeed = jMotorSpeed;
        j.flags = jFlags; j.impulse = jImpulse; j.r = jR; j.masses = jMasses; j.K0 = jK0; j.K1 = jK1;
        j.motorMass = jMotorMass;
        j.axis = jAxis;
        j.extra = jExtra;
        j.h = 0.0f;
        return j;
    }

    WorldView view() const {
        WorldView w;
        w.nBodies = nBodies;
        w.btype = btype; w.bflags = bflags; w.xf = xf; w.pos = pos; w.pos0 = pos0; w.lcenter = lcenter;
        w.vel = vel; w.force = force; w.mass = mass; w.gsleep = gsleep; w.colourMask = colourMask;
        w.islandOf = islandOf;
        w.islandAwake = islandAwake;
        w.islandSleep = islandSleep;
        w.islandMinSep = islandMinSep;
        w.bworld = subWorlds ? bworld : nullptr;
        w.jointPairs = anyJointPair ? jointPairs : nullptr;
        w.jointPairMask = jointPairMask;
        w.nFixtures = nFixtures;
        w.fbody = fbody; w.shape = shape; w.fmat = fmat; w.ffilter = ffilter; w.fgroup = fgroup; w.fat = fat;
        w.moved = moved;
        w.ford = orderSpan ? fixOrder : nullptr;
        w.ordf = orderSpan ? orderFix : nullptr;
        w.nContacts = nContacts;
        w.cfA = cur.cfA; w.cfB = cur.cfB; w.cflags = cur.cflags; w.cmat = cur.cmat; w.ctan = cur.ctan; w.cman = cur.cman;
        w.ccolour = cur.ccolour;
        w.ctoi = cur.ctoi;
        w.ctoiCount = cur.ctoiCount;
        return w;
    }
};

namespace {

int fail(b2g_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, B2G_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define LAUNCHCHK(ctx)                                                                             \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, B2G_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
hipError_t growArray(T*& p, int oldCount, int newCap, hipStream_t s) {
    T* q = nullptr;
    hipError_t e = hipMalloc((void**)&q, sizeof(T) * (size_t)(newCap > 0 ? newCap : 1));
    if (e != hipSuccess) return e;
    if (p && oldCount > 0) {
        e = hipMemcpyAsync(q, p, sizeof(T) * (size_t)oldCount, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    if (p) (void)hipFree(p);
    p = q;
    return hipSuccess;
}

template <typename T>
void freeArray(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int sizeScratch(b2g_context* ctx);

int growBodies(b2g_context* ctx, int need) {
    if (need <= ctx->capBodies) return B2G_OK;
    const int cap = std::max(need, ctx->capBodies * 2 + 64);
    const int n = ctx->nBodies;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, growArray(ctx->btype, n, cap, s));
    HIPCHK(ctx, growArray(ctx->bflags, n, cap, s));
    HIPCHK(ctx, growArray(ctx->xf, n, cap, s));
    HIPCHK(ctx, growArray(ctx->pos, n, cap, s));
    HIPCHK(ctx, growArray(ctx->pos0, n, cap, s));
    HIPCHK(ctx, growArray(ctx->vel, n, cap, s));
    HIPCHK(ctx, growArray(ctx->force, n, cap, s));
    HIPCHK(ctx, growArray(ctx->mass, n, cap, s));
    HIPCHK(ctx, growArray(ctx->lcenter, n, cap, s));
    HIPCHK(ctx, growArray(ctx->gsleep, n, cap, s));
    HIPCHK(ctx, growArray(ctx->colourMask, n, cap, s));
    HIPCHK(ctx, growArray(ctx->claim, n, cap, s));
    HIPCHK(ctx, growArray(ctx->bworld, n, cap, s));
    HIPCHK(ctx, growArray(ctx->islandOf, n, cap, s));
    HIPCHK(ctx, growArray(ctx->islandAwake, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->islandSleep, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->islandMax, 0, cap, s));
    {
        // staticKey starts "not reached" and k_static_apply restores that after use
        const bool fresh = ctx->staticKeyCap < cap;
        HIPCHK(ctx, growArray(ctx->staticKey, 0, cap, s));
        if (fresh) {
            hipLaunchKernelGGL(k_fill<int>, dim3(blocks(cap)), dim3(256), 0, s, ctx->staticKey, cap, 0x7FFFFFFF);
            ctx->staticKeyCap = cap;
        }
    }
    HIPCHK(ctx, growArray(ctx->islandMinSep, 0, cap * std::max(ctx->islandIterCap, 4), s));
    HIPCHK(ctx, growArray(ctx->bodyKey, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->bodyKeySorted, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->bodyCluster, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->clusterBodies, 0, cap, s));
    ctx->clusterStartCap = (cap / 64 + 2) * (kColourSlots + 1);  // clusters of at least 64 bodies, plus the boundary
    HIPCHK(ctx, growArray(ctx->clusterStart, 0, ctx->clusterStartCap, s));
    ctx->islandIterCap = std::max(ctx->islandIterCap, 4);
    // claims start at zero (tag 0 never wins)
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(cap - n)), dim3(256), 0, s, ctx->claim + n, cap - n,
                       0ull);
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(cap - n)), dim3(256), 0, s, ctx->colourMask + n,
                       cap - n, 0ull);
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks(cap - n)), dim3(256), 0, s, ctx->bworld + n, cap - n, 0);
    LAUNCHCHK(ctx);
    ctx->capBodies = cap;
    return ctx->scratch ? sizeScratch(ctx) : B2G_OK;  // the clustered solve sorts the bodies
}

int growContacts(b2g_context* ctx, int need);

// Scan/sort scratch for the largest array the step scans or sorts.
int sizeScratch(b2g_context* ctx) {
    const int sn = std::max(std::max(std::max(ctx->capContacts, ctx->nBuckets + 1), ctx->capBodies), 1024);
    const size_t need = prims_scratch_bytes(sn);
    if (need > ctx->scratchBytes) {
        freeArray(ctx->scratch);
        HIPCHK(ctx, hipMalloc(&ctx->scratch, need));
        ctx->scratchBytes = need;
    }
    return B2G_OK;
}

int growFixtures(b2g_context* ctx, int need) {
    if (need <= ctx->capFixtures) return B2G_OK;
    const int cap = std::max(need, ctx->capFixtures * 2 + 64);
    const int n = ctx->nFixtures;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, growArray(ctx->fbody, n, cap, s));
    HIPCHK(ctx, growArray(ctx->shape, n, cap, s));
    HIPCHK(ctx, growArray(ctx->fmat, n, cap, s));
    HIPCHK(ctx, growArray(ctx->ffilter, n, cap, s));
    HIPCHK(ctx, growArray(ctx->fgroup, n, cap, s));
    HIPCHK(ctx, growArray(ctx->fat, n, cap, s));
    HIPCHK(ctx, growArray(ctx->moved, n, cap, s));
    HIPCHK(ctx, growArray(ctx->range, 0, cap, s));
    ctx->capFixtures = cap;
    // grid sized for the fixture capacity
    int nb = 4096;
    while (nb < 4 * cap) nb <<= 1;
    freeArray(ctx->gCounts); freeArray(ctx->gStarts); freeArray(ctx->gCursor); freeArray(ctx->gEntries);
    freeArray(ctx->gBig); freeArray(ctx->gBigSorted); freeArray(ctx->gCell); freeArray(ctx->gEfat);
    freeArray(ctx->gQueries);
    ctx->nBuckets = nb;
    ctx->capEntries = kMaxCellsPerProxy * cap;
    HIPCHK(ctx, hipMalloc((void**)&ctx->gCounts, sizeof(int) * (nb + 1)));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gStarts, sizeof(int) * (nb + 1)));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gCursor, sizeof(int) * (nb + 1)));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gEntries, sizeof(int) * ctx->capEntries));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gCell, sizeof(unsigned long long) * ctx->capEntries));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gEfat, sizeof(float4) * ctx->capEntries));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gQueries, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gBig, sizeof(unsigned long long) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->gBigSorted, sizeof(unsigned long long) * cap));
    const int rc = growContacts(ctx, 8 * cap + 4096);
    if (rc != B2G_OK) return rc;
    return sizeScratch(ctx);
}

int growContacts(b2g_context* ctx, int need) {
    if (need <= ctx->capContacts) return B2G_OK;
    const int cap = std::max(need, ctx->capContacts * 2);
    const int n = ctx->nContacts;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, growArray(ctx->cur.cfA, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.cfB, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.cflags, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.cmat, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.ctan, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.cman, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.ccolour, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.ctoi, n, cap, s));
    HIPCHK(ctx, growArray(ctx->cur.ctoiCount, n, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.cfA, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.cfB, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.cflags, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.cmat, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.ctan, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.cman, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.ccolour, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.ctoi, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->alt.ctoiCount, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->keep, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->idx, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->newKeys, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->newKeysSorted, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->consKeys, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->consKeysSorted, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->consVals, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->consOrder, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->tileList, 0, cap, s));
    ConView& c = ctx->cv;
    HIPCHK(ctx, growArray(c.contact, 0, cap, s));
    HIPCHK(ctx, growArray(c.bodies, 0, cap, s));
    HIPCHK(ctx, growArray(c.mass, 0, cap, s));
    HIPCHK(ctx, growArray(c.normal, 0, cap, s));
    HIPCHK(ctx, growArray(c.r0, 0, cap, s));
    HIPCHK(ctx, growArray(c.r1, 0, cap, s));
    HIPCHK(ctx, growArray(c.pmass, 0, cap, s));
    HIPCHK(ctx, growArray(c.bias, 0, cap, s));
    HIPCHK(ctx, growArray(c.imp, 0, cap, s));
    HIPCHK(ctx, growArray(c.K, 0, cap, s));
    HIPCHK(ctx, growArray(c.NM, 0, cap, s));
    HIPCHK(ctx, growArray(c.info, 0, cap, s));
    HIPCHK(ctx, growArray(c.lp, 0, cap, s));
    HIPCHK(ctx, growArray(c.lnp, 0, cap, s));
    HIPCHK(ctx, growArray(c.lc, 0, cap, s));
    HIPCHK(ctx, growArray(c.rad, 0, cap, s));
    HIPCHK(ctx, growArray(c.island, 0, cap, s));
    HIPCHK(ctx, growArray(c.slot, 0, cap, s));
    int hc = 1024;
    while (hc < 2 * cap) hc <<= 1;
    freeArray(ctx->hashTable);
    HIPCHK(ctx, hipMalloc((void**)&ctx->hashTable, sizeof(unsigned long long) * hc));
    ctx->hashCap = hc;
    ctx->hashMask = (uint32_t)(hc - 1);
    ctx->capContacts = cap;
    return sizeScratch(ctx);
}

// Cell size of the broad-phase grid: the largest proxy extent among
// non-static fixtures plus the fat margins, so a typical fat box spans at
// most 2x2 cells.
void chooseCellSize(b2g_context* ctx, const b2g_fixture_def* defs, int n, const std::vector<int>& types) {
    float ext = ctx->cellSizeSet ? ctx->cellSize - 2.0f * kAabbExtension : 0.0f;
    for (int i = 0; i < n; ++i) {
        const b2g_shape& s = defs[i].shape;
        const int b = defs[i].body;
        if (b < 0 || b >= (int)types.size() || types[b] == BT_STATIC) continue;
        float r = 0.0f;
        if (s.type == B2G_CIRCLE) {
            r = sqrtf(s.centroid.x * s.centroid.x + s.centroid.y * s.centroid.y) + s.radius;
        } else {
            const int nv = s.type == B2G_EDGE ? 2 : s.count;
            for (int k = 0; k < nv; ++k)
                r = std::max(r, sqrtf(s.verts[k].x * s.verts[k].x + s.verts[k].y * s.verts[k].y));
            r += s.radius;
        }
        ext = std::max(ext, 2.0f * r);
    }
    if (ext > 0.0f) {
        ctx->cellSize = ext + 2.0f * kAabbExtension;
        ctx->cellSizeSet = true;
    }
}

int rebuildHash(b2g_context* ctx, bool liveOnDevice = false);

// Keeps a third of the contact capacity free ahead of the step.
int reserveContacts(b2g_context* ctx) {
    if (3 * (long long)ctx->nContacts > 2 * (long long)ctx->capContacts)
        return growContacts(ctx, 2 * ctx->capContacts);
    return B2G_OK;
}

// Runs the broad-phase over the current fat boxes and appends the new
// contacts. Requires the hash of live contacts to be current.
// Phases are timed with events recorded on the step stream around their
// launches (so they include dispatch gaps); entries named after a kernel
// ("k_...") are that kernel's own execution time, from events that the
// dispatch packet itself records (hipExtLaunchKernelGGL).
enum Phase {
    PH_COLLIDE, PH_COMPACT, PH_COLOUR, PH_ORDER, PH_INTEGRATE_V, PH_INIT, PH_WARM, PH_SOLVE_V, PH_STORE,
    PH_INTEGRATE_X, PH_SOLVE_X, PH_SYNC, PH_BROADPHASE, PH_ISLANDS, PH_LEARNED, PH_K_SOLVE_V, PH_K_SOLVE_X,
    PH_K_MLP_INFER, PH_K_MLP_GRAD, PH_TOI, PH_K_SOLVE_V_CLUSTER, PH_K_SOLVE_X_CLUSTER, PH_K_COLLIDE,
    PH_K_PAIRS_QUERY, PH_K_SOLVE_PERSIST, PH_COUNT
};
const char* kPhaseNames[PH_COUNT] = {"collide", "compact", "colour", "order", "integrate_velocities",
                                     "init_constraints", "warm_start", "solve_velocity", "store_impulses",
                                     "integrate_positions", "solve_position", "sync", "broadphase", "islands",
                                     "fresh_points", "k_solve_velocity", "k_solve_position", "k_mlp_infer",
                                     "k_mlp_grad", "toi", "k_solve_velocity_cluster",
                                     "k_solve_position_cluster", "k_collide", "k_pairs_query",
                                     "k_solve_persistent"};

int readCounters(b2g_context* ctx);
int profKernelEvents(b2g_context* ctx, hipEvent_t* start, hipEvent_t* stop);
void profKernel(b2g_context* ctx, int phase, int i);

// liveOnDevice: the contact table was compacted without a readback
// (compactContacts); the query's readback brings the live count.
int runBroadphase(b2g_context* ctx, bool liveOnDevice = false) {
    hipStream_t s = ctx->stream;
    const int nf = ctx->nFixtures;
    if (ctx->orderSpan > 0 && ctx->orderCount != nf)
        return fail(ctx, B2G_ERR_STATE,
                    "proxy ids cover " + std::to_string(ctx->orderCount) + " of " + std::to_string(nf) + " fixtures");
    if (nf == 0) {
        if (liveOnDevice) {
            const int rc = readCounters(ctx);
            if (rc != B2G_OK) return rc;
            ctx->nContacts = ctx->hcnt->nContacts;
        }
        return B2G_OK;
    }
    GridView g;
    g.invCell = 1.0f / ctx->cellSize;
    g.bucketMask = (uint32_t)(ctx->nBuckets - 1);
    g.counts = ctx->gCounts; g.starts = ctx->gStarts; g.cursor = ctx->gCursor; g.entries = ctx->gEntries;
    g.ecell = ctx->gCell;
    g.efat = ctx->gEfat;
    g.queries = ctx->gQueries;
    g.big = ctx->gBig; g.range = ctx->range;
    g.bigSorted = 0;
    g.fbody = ctx->fbody;
    g.bworld = ctx->subWorlds ? ctx->bworld : nullptr;
    hipLaunchKernelGGL(k_grid_reset, dim3(blocks(ctx->nBuckets + 1)), dim3(256), 0, s, ctx->gCounts, ctx->gCursor,
                       ctx->nBuckets + 1, ctx->cnt);
    hipLaunchKernelGGL(k_grid_count, dim3(blocks(nf)), dim3(256), 0, s, ctx->fat, ctx->moved, nf,
                       (const uint8_t*)nullptr, g, ctx->cnt);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, prims_exclusive_sum(ctx->scratch, ctx->scratchBytes, ctx->gCounts, ctx->gStarts, ctx->nBuckets + 1, s));
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks(nf)), dim3(256), 0, s, nf, g, ctx->fat, ctx->moved);
    if (ctx->subWorlds) {
        // Wide proxies sorted by sub-world, so a query scans its own world's only.
        int nBig = 0;
        HIPCHK(ctx, hipMemcpyAsync(&ctx->hcnt->nBig, &ctx->cnt->nBig, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        nBig = ctx->hcnt->nBig;
        if (nBig > 0) {
            HIPCHK(ctx, prims_sort_keys_u64(ctx->scratch, ctx->scratchBytes, ctx->gBig, ctx->gBigSorted, nBig, s));
            g.big = ctx->gBigSorted;
            g.bigSorted = 1;
        }
    }
    // The query only reads the grid, the fat boxes and the move flags, so it
    // is re-run after growing the contact arrays when its pairs do not fit.
    // Pairs are packed by the proxies' order keys when the host set them.
    const bool ordered = ctx->orderSpan > 0;
    const int packBase = ordered ? ctx->orderSpan : nf;
    int nNew = 0;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) {  // k_grid_reset cleared them for the first
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->nNewPairs, 1, 0);
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->overflow, 1, 0);
        }
        PairSink sink{ctx->newKeys, ctx->capContacts, (unsigned long long)packBase, ordered ? ctx->fixOrder : nullptr};
        const int qb = ctx->sched.queryBlock;
        hipEvent_t e0, e1;
        const int pk = profKernelEvents(ctx, &e0, &e1);
        // at most as many query groups as the chip holds at once (256 CUs x 32
        // wavefronts); the groups stride over the moved proxies' queries
        const long long queryThreads = std::min((long long)nf * kQueryLanes, 32768LL * kQueryLanes);
        hipExtLaunchKernelGGL(k_pairs_query, dim3((unsigned)((queryThreads + qb - 1) / qb)), dim3(qb), 0,
                              s, e0, e1, 0, ctx->fat, ctx->moved, nf, g, ctx->view(), true, ctx->hashTable,
                              ctx->hashMask, sink, ctx->cnt);
        if (ctx->profiling)
            hipLaunchKernelGGL(k_count_scanned, dim3(blocks(nf)), dim3(256), 0, s, nf, g, ctx->cnt);
        profKernel(ctx, PH_K_PAIRS_QUERY, pk);
        hipLaunchKernelGGL(k_pairs_wide, dim3(blocks(nf)), dim3(256), 0, s, ctx->fat, ctx->moved, nf, g, ctx->view(),
                           true, ctx->hashTable, ctx->hashMask, sink, ctx->cnt);
        LAUNCHCHK(ctx);
        HIPCHK(ctx, hipMemcpyAsync(ctx->hcnt, ctx->cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        if (liveOnDevice && attempt == 0) ctx->nContacts = ctx->hcnt->nContacts;
        nNew = ctx->hcnt->nNewPairs;
        if (!ctx->hcnt->overflow && ctx->nContacts + nNew <= ctx->capContacts) break;
        if (attempt >= 2) return fail(ctx, B2G_ERR_CAPACITY, "contact capacity exceeded in the broad-phase");
        int rc = growContacts(ctx, std::max(2 * ctx->capContacts, ctx->nContacts + nNew + 4096));
        if (rc != B2G_OK) return rc;
        if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
    }
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3(blocks(nf)), dim3(256), 0, s, ctx->moved, nf, (uint8_t)0);
    WorldView w = ctx->view();
    if (nNew > 0) {
        // packed keys lo * packBase + hi < packBase^2
        int endBit = 1;
        while (endBit < 64 && (1ull << endBit) < (unsigned long long)packBase * (unsigned long long)packBase) ++endBit;
        HIPCHK(ctx, prims_sort_keys_u64(ctx->scratch, ctx->scratchBytes, ctx->newKeys, ctx->newKeysSorted, nNew, s,
                                        endBit));
        hipLaunchKernelGGL(k_append_contacts, dim3(blocks(nNew)), dim3(256), 0, s, w, ctx->newKeysSorted, nNew,
                           ctx->nContacts, (unsigned long long)packBase);
        LAUNCHCHK(ctx);
        ctx->nContacts += nNew;
    }
    ctx->stats.newPairs += nNew;
    ctx->stats.queries = ctx->hcnt->nQueries;
    ctx->stats.scanned = ctx->hcnt->nScanned;
    return B2G_OK;
}

// liveOnDevice: ctx->nContacts only bounds the table; the live count is the
// device's (set by k_compact).
int rebuildHash(b2g_context* ctx, bool liveOnDevice) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(ctx->hashCap)), dim3(256), 0, s, ctx->hashTable,
                       ctx->hashCap, ~0ull);
    if (ctx->nContacts > 0)
        hipLaunchKernelGGL(k_hash_insert, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), ctx->hashTable,
                           ctx->hashMask, liveOnDevice ? (const Counters*)ctx->cnt : (const Counters*)nullptr);
    LAUNCHCHK(ctx);
    return B2G_OK;
}



int readCounters(b2g_context* ctx) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->hcnt, ctx->cnt, sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

// Workgroups of a one-launch solve of `block` lanes each: sched.persistentSolve
// per compute unit, at most as many as can be resident together (a
// cooperative launch needs every workgroup resident, and is refused
// otherwise).
int persistentBlocks(b2g_context* ctx, int block, int* out) {
    if (ctx->cuCount == 0)
        HIPCHK(ctx, hipDeviceGetAttribute(&ctx->cuCount, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const bool joints = ctx->nJoints > 0;
    int& occ = ctx->persistOcc[joints ? 1 : 0][block / 64];
    if (occ == 0)
        HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ, joints ? k_solve_persistent<true> : k_solve_persistent<false>, block, 0));
    if (occ <= 0 || ctx->cuCount <= 0) return fail(ctx, B2G_ERR_STATE, "one-launch solve: no resident workgroup");
    *out = ctx->cuCount * std::min(occ, ctx->sched.persistentSolve);
    return B2G_OK;
}

// ---------------------------------------------------------------- profiling
int profStart(b2g_context* ctx) {
    if (!ctx->profiling) return -1;
    while ((int)ctx->evPool.size() < ctx->evUsed + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->evPool.push_back(e);
    }
    const int i = ctx->evUsed;
    ctx->evUsed += 2;
    (void)hipEventRecord(ctx->evPool[i], ctx->stream);
    return i;
}

void profEnd(b2g_context* ctx, int phase, int i) {
    if (i < 0) return;
    (void)hipEventRecord(ctx->evPool[i + 1], ctx->stream);
    ctx->evMarks.push_back(std::make_pair(phase, i));
}

// Event pair for a kernel-level entry: the caller passes both events to
// hipExtLaunchKernelGGL, then calls profKernel.
int profKernelEvents(b2g_context* ctx, hipEvent_t* start, hipEvent_t* stop) {
    *start = nullptr;
    *stop = nullptr;
    if (!ctx->profiling) return -1;
    while ((int)ctx->evPool.size() < ctx->evUsed + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->evPool.push_back(e);
    }
    const int i = ctx->evUsed;
    ctx->evUsed += 2;
    *start = ctx->evPool[i];
    *stop = ctx->evPool[i + 1];
    return i;
}

void profKernel(b2g_context* ctx, int phase, int i) {
    if (i >= 0) ctx->evMarks.push_back(std::make_pair(phase, i));
}

void profCollect(b2g_context* ctx) {
    if (!ctx->profiling || ctx->evMarks.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& m : ctx->evMarks) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->evPool[m.second], ctx->evPool[m.second + 1]) == hipSuccess) {
            ctx->profMs[m.first] += ms;
            ctx->profLaunches[m.first] += 1;
        }
    }
    ctx->evMarks.clear();
    ctx->evUsed = 0;
}

// Removes the contacts flagged dead, keeping the others in table order; the
// live count is left in the device counters (no readback).
int compactContacts(b2g_context* ctx) {
    hipStream_t s = ctx->stream;
    if (ctx->nContacts == 0) {
        ctx->compactedFrom = 0;
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->nContacts, 1, 0);
        LAUNCHCHK(ctx);
        return B2G_OK;
    }
    const int pe = profStart(ctx);
    ctx->compactedFrom = ctx->nContacts;  // idx[] maps the step's contact indices to the compacted table
    hipLaunchKernelGGL(k_mark_dead, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), ctx->keep);
    HIPCHK(ctx, prims_exclusive_sum(ctx->scratch, ctx->scratchBytes, ctx->keep, ctx->idx, ctx->nContacts, s));
    hipLaunchKernelGGL(k_compact, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->cur, ctx->alt, ctx->keep, ctx->idx,
                       ctx->nContacts, ctx->cnt);
    profEnd(ctx, PH_COMPACT, pe);
    LAUNCHCHK(ctx);
    std::swap(ctx->cur, ctx->alt);
    return B2G_OK;
}

// ---------------------------------------------------------------- joints
int growJoints(b2g_context* ctx, int need) {
    if (need <= ctx->capJoints) return B2G_OK;
    const int cap = std::max(need, ctx->capJoints * 2 + 64);
    const int n = ctx->nJoints;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, growArray(ctx->jBodies, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jAnchors, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jLimits, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jMotorSpeed, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jFlags, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jImpulse, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jAxis, n, cap, s));
    HIPCHK(ctx, growArray(ctx->jExtra, 2 * n, 2 * cap, s));
    HIPCHK(ctx, growArray(ctx->jR, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jMasses, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jK0, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jK1, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jMotorMass, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jOrder, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jImpulseSnap, 0, cap, s));
    HIPCHK(ctx, growArray(ctx->jFlagsSnap, 0, cap, s));
    ctx->capJoints = cap;
    ctx->jointsDirty = true;
    return B2G_OK;
}

// Joint colours: joints in creation order take the lowest colour free at
// their dynamic bodies (joint topology changes only when joints are added
// or removed, so this runs on the host then). Colour 64 collects joints
// whose bodies have no free colour left; it is solved one joint at a time.
// Removed joints take no slot.
// The bits of a float as an int, on the host (a gear joint keeps its bodies
// C and D as such bits in its axis row).
static int floatBits(float f) {
    int i;
    memcpy(&i, &f, sizeof(i));
    return i;
}

int recolourJoints(b2g_context* ctx) {
    const int nj = ctx->nJoints, nb = ctx->nBodies;
    std::vector<int> btype(nb);
    if (nb > 0) {
        HIPCHK(ctx, hipMemcpyAsync(btype.data(), ctx->btype, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    std::vector<unsigned long long> used(nb, 0ull);
    std::vector<int> colour(nj);
    std::vector<int> count(kColourSlots, 0);
    for (int j = 0; j < nj; ++j) {
        colour[j] = -1;
        if (ctx->hJointRemoved[j]) continue;
        const int2 b = ctx->hJointBodies[j];
        const int2 b2 = j < (int)ctx->hJointBodies2.size() ? ctx->hJointBodies2[j] : make_int2(-1, -1);
        // a gear joint also writes bodies C and D
        const int ids[4] = {b.x, b.y, b2.x, b2.y};
        unsigned long long u = 0ull;
        for (int k = 0; k < 4; ++k)
            if (ids[k] >= 0 && btype[ids[k]] == BT_DYNAMIC) u |= used[ids[k]];
        int c = kOverflowColour;
        if (~u != 0ull) {
            c = __builtin_ctzll(~u);
            for (int k = 0; k < 4; ++k)
                if (ids[k] >= 0 && btype[ids[k]] == BT_DYNAMIC) used[ids[k]] |= 1ull << c;
        }
        colour[j] = c;
        ++count[c];
    }
    ctx->jointColourStart.assign(kColourSlots + 1, 0);
    for (int c = 0; c < kColourSlots; ++c) ctx->jointColourStart[c + 1] = ctx->jointColourStart[c] + count[c];
    std::vector<int> order(nj), cursor(ctx->jointColourStart.begin(), ctx->jointColourStart.end() - 1);
    for (int j = 0; j < nj; ++j)
        if (colour[j] >= 0) order[cursor[colour[j]]++] = j;
    if (nj > 0)
        HIPCHK(ctx, hipMemcpyAsync(ctx->jOrder, order.data(), sizeof(int) * nj, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->jointsDirty = false;
    return B2G_OK;
}

// Hash set of body pairs that joints keep from colliding.
int rebuildJointPairs(b2g_context* ctx) {
    std::vector<unsigned long long> keys;
    for (int j = 0; j < ctx->nJoints; ++j) {
        if (!ctx->hJointKeepApart[j] || ctx->hJointRemoved[j]) continue;
        const int2 b = ctx->hJointBodies[j];
        const unsigned lo = (unsigned)std::min(b.x, b.y), hi = (unsigned)std::max(b.x, b.y);
        keys.push_back(((unsigned long long)lo << 32) | hi);
    }
    ctx->anyJointPair = !keys.empty();
    if (keys.empty()) return B2G_OK;
    int cap = 1024;
    while (cap < 2 * (int)keys.size()) cap <<= 1;
    hipStream_t s = ctx->stream;
    if (cap > ctx->jointPairCap) {
        freeArray(ctx->jointPairs);
        HIPCHK(ctx, hipMalloc((void**)&ctx->jointPairs, sizeof(unsigned long long) * cap));
        ctx->jointPairCap = cap;
    }
    ctx->jointPairMask = (uint32_t)(ctx->jointPairCap - 1);
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(ctx->jointPairCap)), dim3(256), 0, s, ctx->jointPairs,
                       ctx->jointPairCap, ~0ull);
    unsigned long long* dk = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&dk, sizeof(unsigned long long) * keys.size()));
    HIPCHK(ctx, hipMemcpyAsync(dk, keys.data(), sizeof(unsigned long long) * keys.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_hash_insert_keys, dim3(blocks((int)keys.size())), dim3(256), 0, s, dk, (int)keys.size(),
                       ctx->jointPairs, ctx->jointPairMask);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipStreamSynchronize(s));
    (void)hipFree(dk);
    if (ctx->nContacts > 0)
        hipLaunchKernelGGL(k_flag_joint_contacts, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view());
    LAUNCHCHK(ctx);
    return B2G_OK;
}

// ---------------------------------------------------------------- learned warm start
int ensureOldManifolds(b2g_context* ctx) {
    if (ctx->oldManCap >= ctx->capContacts) return B2G_OK;
    freeArray(ctx->oldMan);
    HIPCHK(ctx, hipMalloc((void**)&ctx->oldMan, sizeof(b2g_manifold) * (size_t)ctx->capContacts));
    ctx->oldManCap = ctx->capContacts;
    return B2G_OK;
}

int ensureFresh(b2g_context* ctx) {
    const int cap = ctx->capContacts;
    if (cap <= ctx->capFresh) return B2G_OK;
    freeArray(ctx->freshMask); freeArray(ctx->freshNum); freeArray(ctx->freshOff); freeArray(ctx->freshAll);
    freeArray(ctx->freshFeat); freeArray(ctx->freshRef); freeArray(ctx->impSnap); freeArray(ctx->impAlt);
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshMask, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshNum, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshOff, sizeof(int) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshAll, sizeof(float) * kFeatures * 2 * (size_t)cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshFeat, sizeof(float) * kFeatures * 2 * (size_t)cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->freshRef, sizeof(int) * 2 * (size_t)cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->impSnap, sizeof(float4) * cap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->impAlt, sizeof(float4) * cap));
    ctx->capFresh = cap;
    return B2G_OK;
}

int ensureVelSnap(b2g_context* ctx) {
    if (ctx->capVelSnap >= ctx->capBodies) return B2G_OK;
    freeArray(ctx->velSnap);
    HIPCHK(ctx, hipMalloc((void**)&ctx->velSnap, sizeof(float4) * std::max(ctx->capBodies, 1)));
    ctx->capVelSnap = ctx->capBodies;
    return B2G_OK;
}

int ensureMlp(b2g_context* ctx) {
    if (ctx->mlpAllocated) return B2G_OK;
    HIPCHK(ctx, b2mlp::mlpAlloc(ctx->mlp));
    ctx->mlpAllocated = true;
    return B2G_OK;
}

// RMS complementarity residual of the current velocities and impulses
// (k_residual, FP32 on the device).
int deviceResidual(b2g_context* ctx, const WorldView& w, int nCons, double* rms) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->residualPoints, 1, 0);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, (int*)&ctx->cnt->residualSum, 1, 0);
    hipLaunchKernelGGL(k_residual, dim3(blocks(nCons)), dim3(256), 0, s, w, ctx->cv, nCons, ctx->cnt);
    LAUNCHCHK(ctx);
    int rc = readCounters(ctx);
    if (rc != B2G_OK) return rc;
    const int n = ctx->hcnt->residualPoints;
    *rms = n > 0 ? sqrt((double)ctx->hcnt->residualSum / n) : 0.0;
    return B2G_OK;
}

// Host copy of the velocity-constraint geometry for the FP64 residual.
struct HostConstraints {
    std::vector<int2> bodies;
    std::vector<float4> normal, r0, r1, imp, vel;
    std::vector<float2> bias;
    std::vector<int> info;
};

// RMS over the solved points of the complementarity residual, in FP64:
// r_n = vn - bias where the normal impulse is positive, else min(vn - bias, 0);
// r_t = vt where |tangent impulse| < friction * normal impulse, else 0.
double hostResidual(const HostConstraints& h, int nCons) {
    double sum = 0.0;
    long long pts = 0;
    for (int k = 0; k < nCons; ++k) {
        const int vpc = h.info[k] & 3;
        const float4 va = h.vel[h.bodies[k].x], vb = h.vel[h.bodies[k].y];
        const double nx = h.normal[k].x, ny = h.normal[k].y, mu = h.normal[k].z;
        const double tx = ny, ty = -nx;
        for (int j = 0; j < vpc; ++j) {
            const float4 r = j == 0 ? h.r0[k] : h.r1[k];
            const double dvx = (double)vb.x - (double)vb.z * r.w - (double)va.x + (double)va.z * r.y;
            const double dvy = (double)vb.y + (double)vb.z * r.z - (double)va.y - (double)va.z * r.x;
            const double bias = j == 0 ? h.bias[k].x : h.bias[k].y;
            const double vn = dvx * nx + dvy * ny - bias;
            const double vt = dvx * tx + dvy * ty;
            const double ln = j == 0 ? h.imp[k].x : h.imp[k].z;
            const double lt = j == 0 ? h.imp[k].y : h.imp[k].w;
            const double rn = ln > 0.0 ? vn : std::min(vn, 0.0);
            const double rt = fabs(lt) < mu * ln ? vt : 0.0;
            sum += rn * rn + rt * rt;
            ++pts;
        }
    }
    return pts > 0 ? sqrt(sum / (double)pts) : 0.0;
}

// Residual curves of the three starts (cold, warm, learned) for the
// current step; velocities and impulses are restored afterwards.
// warmStart(bool) starts the velocity phase (contact warm start when true,
// then the joints' initialisation).
template <typename WarmFn, typename SolveFn, typename PredictFn>
int traceResiduals(b2g_context* ctx, const WorldView& w, int nCons, int nFresh, WarmFn& warmStart,
                   SolveFn& solveVelocity, PredictFn& predict) {
    hipStream_t s = ctx->stream;
    const int T = ctx->traceIters, nb = ctx->nBodies;
    int rc;
    if ((rc = ensureFresh(ctx)) != B2G_OK) return rc;
    if ((rc = ensureVelSnap(ctx)) != B2G_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->velSnap, ctx->vel, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->impSnap, ctx->cv.imp, sizeof(float4) * nCons, hipMemcpyDeviceToDevice, s));
    const bool haveLearned = ctx->mlp.ready;
    if (haveLearned) {
        HIPCHK(ctx, hipMemcpyAsync(ctx->impAlt, ctx->impSnap, sizeof(float4) * nCons, hipMemcpyDeviceToDevice, s));
        if (nFresh > 0 && (rc = predict(ctx->impAlt)) != B2G_OK) return rc;
    }
    HostConstraints h;
    h.bodies.resize(nCons); h.normal.resize(nCons); h.r0.resize(nCons); h.r1.resize(nCons); h.imp.resize(nCons);
    h.bias.resize(nCons); h.info.resize(nCons); h.vel.resize(nb);
    const ConView& cv = ctx->cv;
    HIPCHK(ctx, hipMemcpyAsync(h.bodies.data(), cv.bodies, sizeof(int2) * nCons, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(h.normal.data(), cv.normal, sizeof(float4) * nCons, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(h.r0.data(), cv.r0, sizeof(float4) * nCons, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(h.r1.data(), cv.r1, sizeof(float4) * nCons, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(h.bias.data(), cv.bias, sizeof(float2) * nCons, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(h.info.data(), cv.info, sizeof(int) * nCons, hipMemcpyDeviceToHost, s));
    auto sample = [&](double* out) -> int {
        HIPCHK(ctx, hipMemcpyAsync(h.vel.data(), ctx->vel, sizeof(float4) * nb, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(h.imp.data(), cv.imp, sizeof(float4) * nCons, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        *out = hostResidual(h, nCons);
        return B2G_OK;
    };
    ctx->trace.assign(3 * (size_t)(T + 1), std::numeric_limits<double>::quiet_NaN());
    for (int v = 0; v < 3; ++v) {
        if (v == 2 && !haveLearned) continue;
        HIPCHK(ctx, hipMemcpyAsync(ctx->vel, ctx->velSnap, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s));
        if (v == 0) {
            hipLaunchKernelGGL(k_fill<float4>, dim3(blocks(nCons)), dim3(256), 0, s, cv.imp, nCons,
                               make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            warmStart(false);
        } else {
            HIPCHK(ctx, hipMemcpyAsync(cv.imp, v == 1 ? ctx->impSnap : ctx->impAlt, sizeof(float4) * nCons,
                                       hipMemcpyDeviceToDevice, s));
            warmStart(true);
        }
        LAUNCHCHK(ctx);
        double* curve = &ctx->trace[(size_t)v * (T + 1)];
        if ((rc = sample(&curve[0])) != B2G_OK) return rc;
        for (int it = 1; it <= T; ++it) {
            solveVelocity(1);
            LAUNCHCHK(ctx);
            if ((rc = sample(&curve[it])) != B2G_OK) return rc;
        }
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->vel, ctx->velSnap, sizeof(float4) * nb, hipMemcpyDeviceToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(cv.imp, ctx->impSnap, sizeof(float4) * nCons, hipMemcpyDeviceToDevice, s));
    ctx->traceStepIters = T;
    return B2G_OK;
}

// The plain solve of a step -- contact warm start and joint init, the
// velocity passes (joint colours, then contact colours), store impulses,
// position integration, the position passes (contact colours, then joint
// colours) and the transform sync -- as one captured graph. Contact colour
// ranges come from ctx->dRanges, written before each launch, and each
// colour's grid is its size class (the next power of two of the workgroups
// it needs), so a graph stays valid while the layout key repeats: views,
// iteration counts, time step, the non-empty colours and their size
// classes, the joint colours. A key is captured when it has been seen in
// one of the last three steps; until then the step launches directly. The
// kernels are the direct path's, reading their ranges from memory, so both
// give the same results.
int solveByGraph(b2g_context* ctx, const WorldView& w, const JointView& jv, const int* colourStart, int nColours,
                 int nCons, float dt, float dtRatio, int velIters, int posIters, bool* ran) {
    *ran = false;
    hipStream_t s = ctx->stream;
    const int sb = ctx->sched.solverBlock;
    auto sizeClass = [sb](int n) {
        const int g = (n + sb - 1) / sb;
        int c = 1;
        while (c < g) c <<= 1;
        return c;
    };
    const bool warm = nCons > 0 && ctx->warmStarting;
    const bool positions = nCons > 0 || ctx->nJoints > 0;
    std::vector<unsigned long long> key;
    auto put = [&](unsigned long long v) { key.push_back(v); };
    auto ptr = [&](const void* q) { put((unsigned long long)(uintptr_t)q); };
    // every pointer and count the captured kernels read
    put((unsigned)w.nBodies);
    ptr(w.btype); ptr(w.bflags); ptr(w.xf); ptr(w.pos); ptr(w.pos0); ptr(w.lcenter); ptr(w.vel); ptr(w.force);
    ptr(w.mass); ptr(w.gsleep); ptr(w.colourMask); ptr(w.islandOf); ptr(w.islandAwake); ptr(w.islandSleep);
    ptr(w.islandMinSep); ptr(w.bworld); ptr(w.jointPairs); put(w.jointPairMask);
    ptr(w.fbody); ptr(w.shape); ptr(w.fmat); ptr(w.ffilter); ptr(w.fgroup); ptr(w.fat); ptr(w.moved);
    ptr(w.cfA); ptr(w.cfB); ptr(w.cflags); ptr(w.cmat); ptr(w.ctan); ptr(w.cman); ptr(w.ccolour);
    const ConView& cv = ctx->cv;
    ptr(cv.contact); ptr(cv.bodies); ptr(cv.mass); ptr(cv.normal); ptr(cv.r0); ptr(cv.r1); ptr(cv.pmass);
    ptr(cv.bias); ptr(cv.imp); ptr(cv.K); ptr(cv.NM); ptr(cv.info); ptr(cv.lp); ptr(cv.lnp); ptr(cv.lc);
    ptr(cv.rad); ptr(cv.island);
    put((unsigned)jv.n); ptr(jv.order); ptr(jv.bodies); ptr(jv.anchors); ptr(jv.limits); ptr(jv.motorSpeed);
    ptr(jv.flags); ptr(jv.impulse); ptr(jv.r); ptr(jv.masses); ptr(jv.K0); ptr(jv.K1); ptr(jv.motorMass);
    ptr(ctx->cnt); ptr(ctx->dRanges);
    unsigned dtBits, ratioBits;
    memcpy(&dtBits, &dt, 4);
    memcpy(&ratioBits, &dtRatio, 4);
    put(dtBits); put(ratioBits); put((unsigned)velIters); put((unsigned)posIters); put(warm ? 1 : 0);
    put((unsigned)sb); put((unsigned)nColours); put(sizeClass(nCons));
    for (int c = 0; c < nColours; ++c) {
        const int n = colourStart[c + 1] - colourStart[c];
        put(n > 0 ? (unsigned long long)sizeClass(n) : 0ull);
    }
    if (ctx->nJoints > 0)
        for (int c = 0; c <= kColourSlots; ++c) put((unsigned)ctx->jointColourStart[c]);

    using SolveGraph = b2g_context::SolveGraph;
    SolveGraph* hit = nullptr;
    for (auto& g : ctx->solveGraphs)
        if (g.key == key) hit = &g;
    if (!hit) {
        bool recent = false;
        for (auto& k : ctx->recentKeys) recent = recent || k == key;
        if (!recent) {
            ctx->recentKeys.push_back(key);
            if (ctx->recentKeys.size() > 3) ctx->recentKeys.erase(ctx->recentKeys.begin());
            return B2G_OK;  // launched directly this time
        }
    }
    if (!ctx->dRanges) HIPCHK(ctx, hipMalloc((void**)&ctx->dRanges, sizeof(int) * (kColourSlots + 1)));
    if (!ctx->hRanges)
        HIPCHK(ctx, hipHostMalloc((void**)&ctx->hRanges, sizeof(int) * (kColourSlots + 1), hipHostMallocDefault));
    // the previous step's copy has completed: the step synchronises with the
    // stream several times before reaching here
    memcpy(ctx->hRanges, colourStart, sizeof(int) * (kColourSlots + 1));
    HIPCHK(ctx, hipMemcpyAsync(ctx->dRanges, ctx->hRanges, sizeof(int) * (kColourSlots + 1), hipMemcpyHostToDevice,
                               s));
    if (!hit) {
        // capture the launch sequence
        auto jointLaunch = [&](auto parallel, auto serial) {
            for (int c = 0; ctx->nJoints > 0 && c < kColourSlots; ++c) {
                const int b0 = ctx->jointColourStart[c], b1 = ctx->jointColourStart[c + 1];
                if (b1 <= b0) continue;
                if (c == kOverflowColour) serial(b0, b1);
                else parallel(b0, b1);
            }
        };
        auto colourLaunch = [&](auto parallel, auto serial) {
            for (int c = 0; c < nColours; ++c) {
                const int n = colourStart[c + 1] - colourStart[c];
                if (n <= 0) continue;
                if (c == kOverflowColour) serial(c, dim3(1));
                else parallel(c, dim3((unsigned)sizeClass(n)));
            }
        };
        const int* R = ctx->dRanges;
        HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        if (warm)
            colourLaunch([&](int c, dim3 g) { hipLaunchKernelGGL(k_warm_start_r, g, dim3(sb), 0, s, w, cv, R, c); },
                         [&](int c, dim3) { hipLaunchKernelGGL(k_warm_start_serial_r, dim3(1), dim3(1), 0, s, w, cv, R, c); });
        const int wj = ctx->warmStarting ? 1 : 0;
        jointLaunch(
            [&](int b0, int b1) {
                hipLaunchKernelGGL(k_joint_init, dim3((b1 - b0 + sb - 1) / sb), dim3(sb), 0, s, w, jv, b0, b1, dtRatio, wj);
            },
            [&](int b0, int b1) { hipLaunchKernelGGL(k_joint_init_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, dtRatio, wj); });
        for (int it = 0; it < velIters; ++it) {
            jointLaunch(
                [&](int b0, int b1) {
                    hipLaunchKernelGGL(k_joint_solve_velocity, dim3((b1 - b0 + sb - 1) / sb), dim3(sb), 0, s, w, jv, b0,
                                       b1, dt);
                },
                [&](int b0, int b1) {
                    hipLaunchKernelGGL(k_joint_solve_velocity_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, dt);
                });
            colourLaunch([&](int c, dim3 g) { hipLaunchKernelGGL(k_solve_velocity_r, g, dim3(sb), 0, s, w, cv, R, c); },
                         [&](int c, dim3) { hipLaunchKernelGGL(k_solve_velocity_serial_r, dim3(1), dim3(1), 0, s, w, cv, R, c); });
        }
        if (nCons > 0)
            hipLaunchKernelGGL(k_store_impulses_r, dim3((unsigned)sizeClass(nCons)), dim3(sb), 0, s, w, cv, R);
        hipLaunchKernelGGL(k_integrate_positions, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, dt);
        for (int it = 0; it < posIters && positions; ++it) {
            colourLaunch(
                [&](int c, dim3 g) { hipLaunchKernelGGL(k_solve_position_r, g, dim3(sb), 0, s, w, cv, R, c, ctx->cnt, it); },
                [&](int c, dim3) {
                    hipLaunchKernelGGL(k_solve_position_serial_r, dim3(1), dim3(1), 0, s, w, cv, R, c, ctx->cnt, it);
                });
            jointLaunch(
                [&](int b0, int b1) {
                    hipLaunchKernelGGL(k_joint_solve_position, dim3((b1 - b0 + sb - 1) / sb), dim3(sb), 0, s, w, jv, b0,
                                       b1, it);
                },
                [&](int b0, int b1) {
                    hipLaunchKernelGGL(k_joint_solve_position_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, it);
                });
        }
        hipLaunchKernelGGL(k_sync_transforms, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        const hipError_t le = hipGetLastError();
        if (ce != hipSuccess || le != hipSuccess || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(ctx, B2G_ERR_HIP, std::string("solve capture: ") + hipGetErrorString(ce != hipSuccess ? ce : le));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ctx, ie);
        if (ctx->solveGraphs.size() >= 6) {
            auto lru = std::min_element(ctx->solveGraphs.begin(), ctx->solveGraphs.end(),
                                        [](const SolveGraph& a, const SolveGraph& b) { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            ctx->solveGraphs.erase(lru);
        }
        ctx->solveGraphs.push_back(SolveGraph{key, exec, 0});
        hit = &ctx->solveGraphs.back();
        ++ctx->graphCaptures;
    }
    hit->used = ++ctx->graphClock;
    HIPCHK(ctx, hipGraphLaunch(hit->exec, s));
    ++ctx->graphLaunches;
    *ran = true;
    return B2G_OK;
}

// Island labels (each body labelled with the smallest body index of its
// island) by the schedule's method; all three give the same labels.
//  0: propagation rounds (two per host check of the change flag) until
//     none lowers a label; best for a few large islands;
//  1: one propagation round, then the union-find pass;
//  2: the union-find pass alone; best for many small islands.
// Joints join islands by the union-find pass in every method.
int labelIslands(b2g_context* ctx, const WorldView& w) {
    hipStream_t s = ctx->stream;
    const int nb = ctx->nBodies, nc = ctx->nContacts;
    const int method = ctx->sched.islands;
    if (nc > 0 && method == 0) {
        for (int pass = 0;; ++pass) {
            // rounds before the checked one: four on the first pass (most
            // scenes settle within them), one after; a checked round that
            // lowers no label means the labels are final
            const int unchecked = pass == 0 ? 3 : 1;
            for (int r = 0; r < unchecked; ++r) {
                hipLaunchKernelGGL(k_island_hook, dim3(blocks(nc)), dim3(256), 0, s, w, (int*)nullptr);
                hipLaunchKernelGGL(k_island_jump, dim3(blocks(nb)), dim3(256), 0, s, w);
            }
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->islandChanged, 1, 0);
            hipLaunchKernelGGL(k_island_hook, dim3(blocks(nc)), dim3(256), 0, s, w, &ctx->cnt->islandChanged);
            hipLaunchKernelGGL(k_island_jump, dim3(blocks(nb)), dim3(256), 0, s, w);
            LAUNCHCHK(ctx);
            HIPCHK(ctx, hipMemcpyAsync(&ctx->hcnt->islandChanged, &ctx->cnt->islandChanged, sizeof(int),
                                       hipMemcpyDeviceToHost, s));
            HIPCHK(ctx, hipStreamSynchronize(s));
            if (!ctx->hcnt->islandChanged) break;
            // every round that changes something lowers a label, so this is a safeguard only
            if (pass >= 1000000) return fail(ctx, B2G_ERR_STATE, "island labelling did not converge");
        }
    } else if (nc > 0) {
        if (method == 1) {
            hipLaunchKernelGGL(k_island_hook, dim3(blocks(nc)), dim3(256), 0, s, w, (int*)nullptr);
            hipLaunchKernelGGL(k_island_jump, dim3(blocks(nb)), dim3(256), 0, s, w);
        }
        hipLaunchKernelGGL(k_island_union, dim3(blocks(nc)), dim3(256), 0, s, w);
    }
    if (ctx->nJoints > 0)
        hipLaunchKernelGGL(k_island_union_joints, dim3(blocks(ctx->nJoints)), dim3(256), 0, s, w, ctx->jview());
    if (method != 0 || ctx->nJoints > 0) hipLaunchKernelGGL(k_island_jump, dim3(blocks(nb)), dim3(256), 0, s, w);
    LAUNCHCHK(ctx);
    return B2G_OK;
}

// First half of a coloured step: counters, pending broad-phase, and the
// narrowphase over the contact table (b2ContactManager::Collide's Update of
// every contact, REF/Dynamics/b2ContactManager.cpp:107-175). The table still
// holds the contacts that Collide destroys (flagged dead) until the second
// half compacts it. With ctx->keepOldManifolds the manifolds from before the
// update are kept for PreSolve.
int solveTOI(b2g_context* ctx, const b2g_step_params* p);

int stepColouredCollide(b2g_context* ctx, const b2g_step_params* p) {
    (void)p;
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = reserveContacts(ctx)) != B2G_OK) return rc;
    memset(&ctx->stats, 0, sizeof(ctx->stats));
    ctx->stats.residual = -1.0f;
    // reset per-step counters
    Counters zero;
    memset(&zero, 0, sizeof(zero));
    for (int i = 0; i < kMaxPosIters; ++i) {
        float z = 0.0f;
        int b;
        memcpy(&b, &z, 4);
        zero.minSep[i] = b;  // encodeOrdered(0.0f) == 0
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->cnt, &zero, sizeof(Counters), hipMemcpyHostToDevice, s));

    if (ctx->newFixture) {
        if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
        if ((rc = runBroadphase(ctx)) != B2G_OK) return rc;
        ctx->newFixture = false;
    }
    // Collide: narrowphase, destruction of non-overlapping contacts.
    if (ctx->nContacts > 0) {
        if (ctx->keepOldManifolds) {
            if ((rc = ensureOldManifolds(ctx)) != B2G_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(ctx->oldMan, ctx->cur.cman, sizeof(b2g_manifold) * ctx->nContacts,
                                       hipMemcpyDeviceToDevice, s));
        }
        int pe = profStart(ctx);
        const int* order = nullptr;
        if (ctx->sched.collideOrder == 1) {
            // the constraint-order buffers are free until the order phase
            hipLaunchKernelGGL(k_collide_keys, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(),
                               ctx->consKeys, ctx->consVals);
            HIPCHK(ctx, prims_count_sort_pairs(ctx->scratch, ctx->scratchBytes, ctx->consKeys, ctx->consKeysSorted,
                                               ctx->consVals, ctx->consOrder, ctx->nContacts, nullptr, 0, s));
            order = ctx->consOrder;
        }
        const int cb = ctx->sched.collideBlock;
        const bool tiles = ctx->sched.collideOrder == 2;
        hipEvent_t e0, e1;
        const int pk = profKernelEvents(ctx, &e0, &e1);
        hipExtLaunchKernelGGL(k_collide, dim3((ctx->nContacts + cb - 1) / cb), dim3(cb), 0, s, e0, e1, 0, ctx->view(),
                              ctx->cnt, order, tiles ? ctx->tileList : (int*)nullptr);
        profKernel(ctx, PH_K_COLLIDE, pk);
        if (tiles) {
            // the deferred polygon pairs, eight lanes each: one pass of the grid
            // covers the last step's count (a longer list takes more passes)
            const int perBlock = kTileBlock / kTileLanes;
            const int expect = std::min(ctx->nContacts, ctx->hcnt->nTiles + ctx->hcnt->nTiles / 4 + 1024);
            const int tileBlocks = std::max((expect + perBlock - 1) / perBlock, 1);
            hipLaunchKernelGGL(k_collide_tiles, dim3(tileBlocks), dim3(kTileBlock), 0, s, ctx->view(), ctx->cnt,
                               ctx->tileList);
        }
        profEnd(ctx, PH_COLLIDE, pe);
        LAUNCHCHK(ctx);
    }
    return B2G_OK;
}

// Second half: compaction, islands, colouring, the solve, broad-phase.
int stepColouredSolve(b2g_context* ctx, const b2g_step_params* p) {
    hipStream_t s = ctx->stream;
    const float dt = p->dt;
    const float inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
    const float dtRatio = ctx->inv_dt0 * dt;
    int rc;
    // The contacts Collide destroyed stay in the table, flagged dead (every
    // stage below skips them), until the compaction before the broad-phase;
    // their colours go back to the bodies now. ctx->nContacts stays the
    // table's length until the broad-phase reads back the live count.
    if (ctx->nContacts > 0) {
        const int pe = profStart(ctx);
        hipLaunchKernelGGL(k_mark_dead, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), ctx->keep);
        profEnd(ctx, PH_COMPACT, pe);
        LAUNCHCHK(ctx);
    }
    if (dt > 0.0f && ctx->stepComplete) {
        const int posIters = std::min(p->positionIterations, kMaxPosIters);
        if (posIters > ctx->islandIterCap) {
            freeArray(ctx->islandMinSep);
            HIPCHK(ctx, hipMalloc((void**)&ctx->islandMinSep, sizeof(int) * (size_t)ctx->capBodies * posIters));
            ctx->islandIterCap = posIters;
        }
        WorldView w = ctx->view();
        // Islands of awake bodies and the constraint set.
        {
            const int pe = profStart(ctx);
            const int nb = ctx->nBodies;
            hipLaunchKernelGGL(k_fill<int>, dim3(blocks(nb * std::max(posIters, 1))), dim3(256), 0, s,
                               ctx->islandMinSep, nb * std::max(posIters, 1), 0);
            hipLaunchKernelGGL(k_island_init, dim3(blocks(nb)), dim3(256), 0, s, w);
            if ((rc = labelIslands(ctx, w)) != B2G_OK) return rc;
            hipLaunchKernelGGL(k_island_roots, dim3(blocks(nb)), dim3(256), 0, s, w);
            hipLaunchKernelGGL(k_island_mark, dim3(blocks(nb)), dim3(256), 0, s, w);
            hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->remaining, 1, 0);
            if (ctx->nContacts > 0)
                hipLaunchKernelGGL(k_mark_constraints, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->cnt);
            profEnd(ctx, PH_ISLANDS, pe);
            LAUNCHCHK(ctx);
        }
        // Constraints move down to lower colours freed since they were
        // coloured (rounds launched without a readback).
        for (int r = 0; r < ctx->demoteRounds && ctx->nContacts > 0; ++r) {
            if (ctx->claimTag >= 0xFFFF) {
                hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(ctx->nBodies)), dim3(256), 0, s,
                                   ctx->claim, ctx->nBodies, 0ull);
                ctx->claimTag = 0;
            }
            const uint32_t tag = ++ctx->claimTag;
            const int pe = profStart(ctx);
            hipLaunchKernelGGL(k_demote_claim, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->claim, tag);
            hipLaunchKernelGGL(k_demote_commit, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->claim, tag);
            profEnd(ctx, PH_COLOUR, pe);
            LAUNCHCHK(ctx);
        }
        // Colour new constraints: rounds are launched without a readback
        // (a round with nothing left to colour changes nothing), and the count
        // still uncoloured comes back with the order's counts below.
        int rounds = 0;
        auto colourRound = [&]() -> int {
            if (ctx->claimTag >= 0xFFFF) {
                hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(ctx->nBodies)), dim3(256), 0, s,
                                   ctx->claim, ctx->nBodies, 0ull);
                ctx->claimTag = 0;
            }
            const uint32_t tag = ++ctx->claimTag;
            const int slot = std::min(rounds, kColourRoundSlots - 1);
            const int pe = profStart(ctx);
            hipLaunchKernelGGL(k_colour_claim, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->claim, tag,
                               ctx->cnt, slot);
            hipLaunchKernelGGL(k_colour_commit, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->claim, tag,
                               ctx->cnt, slot);
            profEnd(ctx, PH_COLOUR, pe);
            LAUNCHCHK(ctx);
            if (++rounds > 100000) return fail(ctx, B2G_ERR_STATE, "colouring did not converge");
            return B2G_OK;
        };
        // as many unchecked rounds as the last step needed, at least two
        const int unchecked =
            std::min(std::max(ctx->colourRoundsNeeded, kColourRoundsUnchecked), kColourRoundSlots - 1);
        if (ctx->nContacts > 0)
            for (int r = 0; r < unchecked; ++r)
                if ((rc = colourRound()) != B2G_OK) return rc;
        auto leftAfterLast = [&]() { return ctx->hcnt->roundLeft[std::min(rounds - 1, kColourRoundSlots - 1)]; };
        // Constraint order by colour; in the clustered solve by (cluster or
        // the boundary set, colour), and colourStart then holds the boundary
        // set's colour ranges.
        const int clusterSize = ctx->sched.clusterBodies;
        const bool clustered = clusterSize > 0 && ctx->nContacts > 0 && ctx->nBodies < kMaxClusterBodies;
        const int nClusters = clustered ? (ctx->nBodies + clusterSize - 1) / clusterSize : 0;
        int nCons = 0;
        int colourStart[kColourSlots + 1];
        memset(colourStart, 0, sizeof(colourStart));
        // Every contact gets an order key (constraints by colour, or by
        // cluster and colour; the others after them) and the table is
        // sorted; one readback brings the counts and what the colouring left.
        auto order = [&]() -> int {
            const int pe = profStart(ctx);
            if (clustered) {
                if ((nClusters + 1) * (kColourSlots + 1) > ctx->clusterStartCap)
                    return fail(ctx, B2G_ERR_STATE, "cluster table too small");
                hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, &ctx->cnt->nClustered, 1, 0);
                hipLaunchKernelGGL(k_body_cluster_keys, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, ctx->bodyKey,
                                   ctx->bodyCluster, ctx->cnt);
                HIPCHK(ctx, prims_sort_keys_u64(ctx->scratch, ctx->scratchBytes, ctx->bodyKey, ctx->bodyKeySorted,
                                                ctx->nBodies, s));
                hipLaunchKernelGGL(k_body_clusters, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, ctx->bodyKeySorted,
                                   ctx->nBodies, clusterSize, ctx->cnt, ctx->bodyCluster, ctx->clusterBodies);
                hipLaunchKernelGGL(k_cons_keys_clustered, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w,
                                   ctx->bodyCluster, ctx->consKeys, ctx->consVals, ctx->cnt);
            } else {
                hipLaunchKernelGGL(k_cons_keys, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->consKeys,
                                   ctx->consVals, ctx->cnt);
            }
            if (clustered) {
                HIPCHK(ctx, prims_sort_pairs_u32(ctx->scratch, ctx->scratchBytes, ctx->consKeys, ctx->consKeysSorted,
                                                 ctx->consVals, ctx->consOrder, ctx->nContacts, 31, s));
                hipLaunchKernelGGL(k_cluster_ranges, dim3(blocks((nClusters + 1) * (kColourSlots + 1))), dim3(256), 0,
                                   s, ctx->consKeysSorted, ctx->nContacts, nClusters, ctx->clusterStart, ctx->cnt);
            } else {
                // colours and the key of the contacts outside the set fit 7 bits
                HIPCHK(ctx, prims_count_sort_pairs(ctx->scratch, ctx->scratchBytes, ctx->consKeys, ctx->consKeysSorted,
                                                   ctx->consVals, ctx->consOrder, ctx->nContacts,
                                                   ctx->cnt->colourCount, kColourSlots, s));
            }
            profEnd(ctx, PH_ORDER, pe);
            LAUNCHCHK(ctx);
            return readCounters(ctx);
        };
        if (ctx->nContacts > 0) {
            if ((rc = order()) != B2G_OK) return rc;
            if (leftAfterLast() > 0) {
                // the colouring needed more rounds: finish it, then order again
                while (leftAfterLast() > 0) {
                    if ((rc = colourRound()) != B2G_OK) return rc;
                    if ((rc = readCounters(ctx)) != B2G_OK) return rc;
                }
                if ((rc = order()) != B2G_OK) return rc;
            }
            int needed = rounds;
            for (int r = 0; r < std::min(rounds, kColourRoundSlots); ++r)
                if (ctx->hcnt->roundLeft[r] == 0) {
                    needed = r + 1;
                    break;
                }
            ctx->colourRoundsNeeded = needed;
            if (clustered) colourStart[0] = ctx->hcnt->boundaryStart;
            for (int c = 0; c < kColourSlots; ++c) colourStart[c + 1] = colourStart[c] + ctx->hcnt->colourCount[c];
            nCons = colourStart[kColourSlots];
        }
        ctx->stats.colouringRounds = rounds;
        int nColours = 0;
        for (int c = 0; c < kColourSlots; ++c)
            if (colourStart[c + 1] > colourStart[c]) nColours = c + 1;
        memcpy(ctx->lastColourStart, colourStart, sizeof(colourStart));
        ctx->lastConstraints = nCons;
        ctx->stats.constraints = nCons;
        ctx->stats.colours = clustered ? (nCons > 0 ? ctx->hcnt->maxColour + 1 : 0) : nColours;
        const ClusterView cview{ctx->clusterStart, ctx->clusterBodies, &ctx->cnt->nClustered, clusterSize};
        const size_t clusterLds = sizeof(float4) * (size_t)std::max(clusterSize, 1);
        const bool clusterLaunch = clustered && nCons > 0 && colourStart[0] > 0;  // some interior constraint

        // Island solve.
        if (ctx->jointsDirty && (rc = recolourJoints(ctx)) != B2G_OK) return rc;
        const bool learned = p->learnedWarmStart != 0;
        if (learned && !ctx->mlp.ready)
            return fail(ctx, B2G_ERR_STATE, "learned warm start requested without a warm-start network");
        const bool wantFresh = nCons > 0 && (learned || ctx->logCap > 0 || ctx->traceIters > 0);
        if (wantFresh && (rc = ensureFresh(ctx)) != B2G_OK) return rc;
        const FreshView fv{wantFresh ? ctx->freshMask : nullptr, ctx->freshNum, ctx->freshAll, ctx->learnedScope};
        int pe = profStart(ctx);
        hipLaunchKernelGGL(k_integrate_velocities, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, ctx->gravity, dt);
        profEnd(ctx, PH_INTEGRATE_V, pe);
        if (nCons > 0) {
            pe = profStart(ctx);
            hipLaunchKernelGGL(k_init_constraints, dim3(blocks(nCons)), dim3(256), 0, s, w, ctx->cv, ctx->consOrder,
                               nCons, dtRatio, ctx->warmStarting ? 1 : 0, ctx->cnt, fv);
            if (clusterLaunch)
                hipLaunchKernelGGL(k_cons_slots, dim3(blocks(nCons)), dim3(256), 0, s, ctx->cv, nCons, ctx->bodyCluster);
            profEnd(ctx, PH_INIT, pe);
        }
        LAUNCHCHK(ctx);
        int nFresh = 0;
        if (wantFresh) {
            pe = profStart(ctx);
            HIPCHK(ctx, prims_exclusive_sum(ctx->scratch, ctx->scratchBytes, ctx->freshNum, ctx->freshOff, nCons, s));
            hipLaunchKernelGGL(k_fresh_gather, dim3(blocks(nCons)), dim3(256), 0, s, nCons, fv, ctx->freshOff,
                               ctx->freshFeat, ctx->freshRef, ctx->cnt);
            profEnd(ctx, PH_LEARNED, pe);
            LAUNCHCHK(ctx);
            if ((rc = readCounters(ctx)) != B2G_OK) return rc;
            nFresh = ctx->hcnt->nFresh;
        }
        ctx->stats.freshPoints = nFresh;
        ctx->stats.residual = -1.0f;
        const int sb = ctx->sched.solverBlock;
        auto sgrid = [sb](int n) { return dim3((unsigned)((n + sb - 1) / sb)); };
        // Trailing colours with at most sched.tailColours constraints each go to
        // one workgroup (tailLaunch); the colours before them get a launch each.
        int tailBegin = nColours;
        if (ctx->sched.tailColours > 0) {
            while (tailBegin > 0 && colourStart[tailBegin] - colourStart[tailBegin - 1] <= ctx->sched.tailColours)
                --tailBegin;
            if (nColours - tailBegin < 2) tailBegin = nColours;  // one colour: no gain
        }
        ctx->stats.wideConstraints = 0;
        for (int c = 0; c < tailBegin; ++c)
            if (c != kOverflowColour) ctx->stats.wideConstraints += colourStart[c + 1] - colourStart[c];
        ColourRanges ranges;
        memcpy(ranges.s, colourStart, sizeof(ranges.s));
        auto perColour = [&](int phase, auto kernelLaunch, auto tailLaunch) {
            for (int c = 0; c < tailBegin; ++c) {
                const int b0 = colourStart[c], b1 = colourStart[c + 1];
                if (b1 <= b0) continue;
                const int pc = profStart(ctx);
                kernelLaunch(b0, b1, c == kOverflowColour);
                profEnd(ctx, phase, pc);
            }
            if (tailBegin < nColours) {
                const int pc = profStart(ctx);
                tailLaunch(tailBegin, nColours);
                profEnd(ctx, phase, pc);
            }
        };
        // Joints: colour by colour; the overflow colour by one lane in order.
        JointView jv = ctx->jview();
        jv.h = dt;
        auto perJointColour = [&](auto kernelLaunch) {
            const int nc = ctx->nJoints > 0 ? kColourSlots : 0;
            for (int c = 0; c < nc; ++c) {
                const int b0 = ctx->jointColourStart[c], b1 = ctx->jointColourStart[c + 1];
                if (b1 > b0) kernelLaunch(b0, b1, c == kOverflowColour);
            }
        };
        // b2Island::Solve's start of the velocity phase (REF/Dynamics/b2Island.cpp:246-255):
        // contact warm start, then the joints' InitVelocityConstraints.
        auto startSolve = [&](bool contactWarm) {
            if (contactWarm && clusterLaunch) {
                const int pc = profStart(ctx);
                hipLaunchKernelGGL(k_warm_start_cluster, dim3(nClusters), dim3(sb), clusterLds, s, w, ctx->cv, cview);
                profEnd(ctx, PH_WARM, pc);
            }
            if (contactWarm) {
                perColour(
                    PH_WARM,
                    [&](int b0, int b1, bool serial) {
                        if (serial)
                            hipLaunchKernelGGL(k_warm_start_serial, dim3(1), dim3(1), 0, s, w, ctx->cv, b0, b1);
                        else
                            hipLaunchKernelGGL(k_warm_start, sgrid(b1 - b0), dim3(sb), 0, s, w, ctx->cv, b0, b1);
                    },
                    [&](int c0, int c1) {
                        hipLaunchKernelGGL(k_warm_start_tail, dim3(1), dim3(1024), 0, s, w, ctx->cv, ranges, c0, c1);
                    });
            }
            perJointColour([&](int b0, int b1, bool serial) {
                if (serial)
                    hipLaunchKernelGGL(k_joint_init_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, dtRatio,
                                       ctx->warmStarting ? 1 : 0);
                else
                    hipLaunchKernelGGL(k_joint_init, sgrid(b1 - b0), dim3(sb), 0, s, w, jv, b0, b1, dtRatio,
                                       ctx->warmStarting ? 1 : 0);
            });
        };
        // Each iteration: joints, then contacts (REF/Dynamics/b2Island.cpp:261-270).
        auto solveVelocity = [&](int iters) {
            for (int it = 0; it < iters; ++it) {
                perJointColour([&](int b0, int b1, bool serial) {
                    if (serial)
                        hipLaunchKernelGGL(k_joint_solve_velocity_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, dt);
                    else
                        hipLaunchKernelGGL(k_joint_solve_velocity, sgrid(b1 - b0), dim3(sb), 0, s, w, jv, b0, b1, dt);
                });
                if (clusterLaunch) {
                    const int pc = profStart(ctx);
                    hipEvent_t e0, e1;
                    const int pk = profKernelEvents(ctx, &e0, &e1);
                    hipExtLaunchKernelGGL(k_solve_velocity_cluster, dim3(nClusters), dim3(sb), clusterLds, s, e0, e1, 0,
                                          w, ctx->cv, cview);
                    profKernel(ctx, PH_K_SOLVE_V_CLUSTER, pk);
                    profEnd(ctx, PH_SOLVE_V, pc);
                }
                perColour(
                    PH_SOLVE_V,
                    [&](int b0, int b1, bool serial) {
                        if (serial) {
                            hipLaunchKernelGGL(k_solve_velocity_serial, dim3(1), dim3(1), 0, s, w, ctx->cv, b0, b1);
                            return;
                        }
                        hipEvent_t e0, e1;
                        const int pk = profKernelEvents(ctx, &e0, &e1);
                        hipExtLaunchKernelGGL(k_solve_velocity, sgrid(b1 - b0), dim3(sb), 0, s, e0, e1, 0, w,
                                              ctx->cv, b0, b1);
                        profKernel(ctx, PH_K_SOLVE_V, pk);
                    },
                    [&](int c0, int c1) {
                        hipLaunchKernelGGL(k_solve_velocity_tail, dim3(1), dim3(1024), 0, s, w, ctx->cv, ranges, c0,
                                           c1);
                    });
            }
        };
        auto predict = [&](float4* imp) -> int {
            hipEvent_t e0, e1;
            const int pk = profKernelEvents(ctx, &e0, &e1);
            HIPCHK(ctx, b2mlp::mlpPredictImpulses(ctx->mlp, ctx->freshFeat, ctx->freshRef, &ctx->cnt->nFresh, nFresh,
                                                  ctx->cv.normal, imp, ctx->learnedScope == 1, s, e0, e1));
            profKernel(ctx, PH_K_MLP_INFER, pk);
            return B2G_OK;
        };
        auto saveJoints = [&]() -> int {
            if (ctx->nJoints == 0) return B2G_OK;
            HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulseSnap, ctx->jImpulse, sizeof(float4) * ctx->nJoints,
                                       hipMemcpyDeviceToDevice, s));
            HIPCHK(ctx, hipMemcpyAsync(ctx->jFlagsSnap, ctx->jFlags, sizeof(int) * ctx->nJoints,
                                       hipMemcpyDeviceToDevice, s));
            return B2G_OK;
        };
        auto restoreJoints = [&]() -> int {
            if (ctx->nJoints == 0) return B2G_OK;
            HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse, ctx->jImpulseSnap, sizeof(float4) * ctx->nJoints,
                                       hipMemcpyDeviceToDevice, s));
            HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags, ctx->jFlagsSnap, sizeof(int) * ctx->nJoints,
                                       hipMemcpyDeviceToDevice, s));
            return B2G_OK;
        };
        // One-launch solve (sched.persistentSolve): the passes below in one
        // cooperative launch, or two where the host acts between the velocity
        // and the position passes (the residual fallback, the impulse log).
        const bool persist = ctx->sched.persistentSolve > 0 && !clusterLaunch && ctx->traceIters == 0 &&
                             (nCons > 0 || ctx->nJoints > 0);
        auto persistentLaunch = [&](int stages, bool contactWarm) -> int {
            SolvePlan plan;
            memset(&plan, 0, sizeof(plan));
            memcpy(plan.cs, colourStart, sizeof(plan.cs));
            if (ctx->nJoints > 0) {
                for (int c = 0; c <= kColourSlots; ++c) plan.js[c] = ctx->jointColourStart[c];
                plan.nJointColours = kColourSlots;
            }
            plan.nColours = nColours;
            plan.tailBegin = tailBegin;
            plan.stages = stages;
            plan.contactWarm = contactWarm ? 1 : 0;
            plan.jointWarm = ctx->warmStarting ? 1 : 0;
            plan.dtRatio = dtRatio;
            plan.h = dt;
            plan.velIters = p->velocityIterations;
            plan.posIters = posIters;
            plan.nCons = nCons;
            int nb = 0;
            const int pb = ctx->nJoints > 0 ? std::min(sb, kPersistJointBlock) : sb;
            const int brc = persistentBlocks(ctx, pb, &nb);
            if (brc != B2G_OK) return brc;
            WorldView wa = w;
            ConView ca = ctx->cv;
            JointView ja = jv;
            Counters* cp = ctx->cnt;
            void* args[] = {(void*)&wa, (void*)&ca, (void*)&ja, (void*)&plan, (void*)&cp};
            const int pc = profStart(ctx);
            hipEvent_t e0, e1;
            const int pk = profKernelEvents(ctx, &e0, &e1);
            if (e0) (void)hipEventRecord(e0, s);
            const void* kern = ctx->nJoints > 0 ? (const void*)k_solve_persistent<true>
                                                : (const void*)k_solve_persistent<false>;
            HIPCHK(ctx, hipLaunchCooperativeKernel(kern, dim3((unsigned)nb), dim3(pb), args, 0, s));
            if (e1) (void)hipEventRecord(e1, s);
            profKernel(ctx, PH_K_SOLVE_PERSIST, pk);
            profEnd(ctx, (stages & SP_VELOCITY) ? PH_SOLVE_V : PH_SOLVE_X, pc);
            return B2G_OK;
        };
        // The plain solve (no learned start, trace, log or profiling) runs as
        // a captured graph once its layout has repeated.
        bool graphRan = false;
        if (ctx->graphs && !persist && !ctx->profiling && !learned && ctx->traceIters == 0 && ctx->logCap == 0 &&
            !clustered) {
            rc = solveByGraph(ctx, w, jv, colourStart, nColours, nCons, dt, dtRatio, p->velocityIterations, posIters,
                              &graphRan);
            if (rc != B2G_OK) return rc;
        }
        if (!graphRan) {
        if (nCons > 0 && ctx->traceIters > 0) {
            if ((rc = saveJoints()) != B2G_OK) return rc;
            auto traceStart = [&](bool contactWarm) {
                (void)restoreJoints();
                startSolve(contactWarm);
            };
            if ((rc = traceResiduals(ctx, w, nCons, nFresh, traceStart, solveVelocity, predict)) != B2G_OK) return rc;
            if ((rc = restoreJoints()) != B2G_OK) return rc;
        }
        const bool usePrediction = learned && nFresh > 0;
        const bool checkResidual = usePrediction && p->residualThreshold > 0.0f;
        if (checkResidual) {
            if ((rc = ensureVelSnap(ctx)) != B2G_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(ctx->velSnap, ctx->vel, sizeof(float4) * ctx->nBodies, hipMemcpyDeviceToDevice, s));
            HIPCHK(ctx, hipMemcpyAsync(ctx->impSnap, ctx->cv.imp, sizeof(float4) * nCons, hipMemcpyDeviceToDevice, s));
            if ((rc = saveJoints()) != B2G_OK) return rc;
        }
        if (usePrediction && (rc = predict(ctx->cv.imp)) != B2G_OK) return rc;
        // With nothing for the host between the velocity and the position
        // passes, the one-launch solve runs them all in one launch below.
        const bool velocityDeferred = persist && !checkResidual && ctx->logCap == 0;
        const bool contactWarm = nCons > 0 && (ctx->warmStarting || usePrediction);
        if (!persist) {
            startSolve(contactWarm);
            solveVelocity(p->velocityIterations);
        } else if (!velocityDeferred && (rc = persistentLaunch(SP_WARM | SP_VELOCITY, contactWarm)) != B2G_OK) {
            return rc;
        }
        if (checkResidual) {
            double rms = 0.0;
            if ((rc = deviceResidual(ctx, w, nCons, &rms)) != B2G_OK) return rc;
            ctx->stats.residual = (float)rms;
            if (!(rms <= (double)p->residualThreshold)) {
                // The plain start judges the learned one: solve again from it.
                HIPCHK(ctx, hipMemcpyAsync(ctx->vel, ctx->velSnap, sizeof(float4) * ctx->nBodies,
                                           hipMemcpyDeviceToDevice, s));
                HIPCHK(ctx, hipMemcpyAsync(ctx->cv.imp, ctx->impSnap, sizeof(float4) * nCons, hipMemcpyDeviceToDevice,
                                           s));
                if ((rc = restoreJoints()) != B2G_OK) return rc;
                if (!persist) {
                    startSolve(ctx->warmStarting);
                    solveVelocity(p->velocityIterations);
                } else if ((rc = persistentLaunch(SP_WARM | SP_VELOCITY, ctx->warmStarting)) != B2G_OK) {
                    return rc;
                }
                ctx->stats.fallback = 1;
            }
        }
        LAUNCHCHK(ctx);
        const bool logStep = ctx->learnedScope == 0 || ctx->logSteps % ctx->logStride == 0;
        if (ctx->logCap > 0) ++ctx->logSteps;
        if (ctx->logCap > 0 && logStep && nFresh > 0 && ctx->logCount < ctx->logCap) {
            hipLaunchKernelGGL(k_log_append, dim3(blocks(nFresh)), dim3(256), 0, s, nFresh, ctx->freshRef,
                               ctx->freshFeat, ctx->cv.imp, ctx->logFeat, ctx->logTarget, ctx->logCount, ctx->logCap,
                               ctx->learnedScope);
            LAUNCHCHK(ctx);
            ctx->logCount = std::min(ctx->logCap, ctx->logCount + nFresh);
        }
        if (persist) {
            // store, integrate, position passes and transforms (and the
            // velocity passes when deferred) in one launch
            const int stages = SP_POSITION | (velocityDeferred ? SP_WARM | SP_VELOCITY : 0);
            if ((rc = persistentLaunch(stages, contactWarm)) != B2G_OK) return rc;
        } else {
        if (nCons > 0) {
            pe = profStart(ctx);
            hipLaunchKernelGGL(k_store_impulses, dim3(blocks(nCons)), dim3(256), 0, s, w, ctx->cv, nCons);
            profEnd(ctx, PH_STORE, pe);
        }
        pe = profStart(ctx);
        hipLaunchKernelGGL(k_integrate_positions, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, dt);
        profEnd(ctx, PH_INTEGRATE_X, pe);
        LAUNCHCHK(ctx);
        // Each position iteration: contacts, then joints (REF/Dynamics/b2Island.cpp:310-328).
        for (int it = 0; it < posIters && (nCons > 0 || ctx->nJoints > 0); ++it) {
            if (clusterLaunch) {
                const int pc = profStart(ctx);
                hipEvent_t e0, e1;
                const int pk = profKernelEvents(ctx, &e0, &e1);
                hipExtLaunchKernelGGL(k_solve_position_cluster, dim3(nClusters), dim3(sb), clusterLds, s, e0, e1, 0, w,
                                      ctx->cv, cview, ctx->cnt, it);
                profKernel(ctx, PH_K_SOLVE_X_CLUSTER, pk);
                profEnd(ctx, PH_SOLVE_X, pc);
            }
            perColour(
                PH_SOLVE_X,
                [&](int b0, int b1, bool serial) {
                    if (serial) {
                        hipLaunchKernelGGL(k_solve_position_serial, dim3(1), dim3(1), 0, s, w, ctx->cv, b0, b1,
                                           ctx->cnt, it);
                        return;
                    }
                    hipEvent_t e0, e1;
                    const int pk = profKernelEvents(ctx, &e0, &e1);
                    hipExtLaunchKernelGGL(k_solve_position, sgrid(b1 - b0), dim3(sb), 0, s, e0, e1, 0, w, ctx->cv,
                                          b0, b1, ctx->cnt, it);
                    profKernel(ctx, PH_K_SOLVE_X, pk);
                },
                [&](int c0, int c1) {
                    hipLaunchKernelGGL(k_solve_position_tail, dim3(1), dim3(1024), 0, s, w, ctx->cv, ranges, c0, c1,
                                       ctx->cnt, it);
                });
            perJointColour([&](int b0, int b1, bool serial) {
                if (serial)
                    hipLaunchKernelGGL(k_joint_solve_position_serial, dim3(1), dim3(1), 0, s, w, jv, b0, b1, it);
                else
                    hipLaunchKernelGGL(k_joint_solve_position, sgrid(b1 - b0), dim3(sb), 0, s, w, jv, b0, b1, it);
            });
        }
        }  // !persist
        }  // !graphRan
        pe = profStart(ctx);
        if (!graphRan && !persist) hipLaunchKernelGGL(k_sync_transforms, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w);
        if (ctx->allowSleep) {
            hipLaunchKernelGGL(k_sleep_accum, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, dt);
        }
        // Static bodies reached by this step's islands (before the sleep below
        // changes the island flags the key reads). Without sleeping, only a
        // static body that is asleep can change (the search wakes it).
        if ((ctx->nContacts > 0 || ctx->nJoints > 0) && (ctx->allowSleep || ctx->staticAsleep)) {
            hipLaunchKernelGGL(k_fill<int>, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, ctx->islandMax, ctx->nBodies,
                               -1);
            hipLaunchKernelGGL(k_island_max, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, ctx->islandMax);
            if (ctx->nContacts > 0)
                hipLaunchKernelGGL(k_static_reach, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->islandMax,
                                   ctx->staticKey, ctx->allowSleep ? 1 : 0, posIters);
            if (ctx->nJoints > 0)
                hipLaunchKernelGGL(k_static_reach_joints, dim3(blocks(ctx->nJoints)), dim3(256), 0, s, w, jv,
                                   ctx->islandMax, ctx->staticKey, ctx->allowSleep ? 1 : 0, posIters);
            hipLaunchKernelGGL(k_static_apply, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, ctx->staticKey);
        }
        if (ctx->allowSleep) {
            hipLaunchKernelGGL(k_sleep_apply, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w, posIters);
        }
        hipLaunchKernelGGL(k_sync_fixtures, dim3(blocks(ctx->nFixtures)), dim3(256), 0, s, w);
        profEnd(ctx, PH_SYNC, pe);
        LAUNCHCHK(ctx);
        // Broad-phase against the compacted contact set (its readback brings
        // the live count).
        if ((rc = compactContacts(ctx)) != B2G_OK) return rc;
        pe = profStart(ctx);
        if ((rc = rebuildHash(ctx, true)) != B2G_OK) return rc;
        if ((rc = runBroadphase(ctx, true)) != B2G_OK) return rc;
        profEnd(ctx, PH_BROADPHASE, pe);
        // Stats of the position iterations.
        int ran = 0;
        float lastMin = 0.0f;
        for (int it = 0; it < posIters && nCons > 0; ++it) {
            ++ran;
            lastMin = decodeOrdered(ctx->hcnt->minSep[it]);
            if (lastMin >= -3.0f * kLinearSlop) break;
        }
        ctx->stats.positionIterations = ran;
        ctx->stats.points = ctx->hcnt->points;
        ctx->stats.minSeparation = lastMin;
        if (ctx->hcnt->barrierAbort)
            return fail(ctx, B2G_ERR_STATE, "one-launch solve: a grid barrier wait passed its bound");
    } else {
        if ((rc = compactContacts(ctx)) != B2G_OK) return rc;
        if ((rc = readCounters(ctx)) != B2G_OK) return rc;
        ctx->nContacts = ctx->hcnt->nContacts;
    }
    ctx->stats.beginEvents = ctx->hcnt->beginEvents;
    ctx->stats.endEvents = ctx->hcnt->endEvents;
    if (ctx->continuous && dt > 0.0f && (rc = solveTOI(ctx, p)) != B2G_OK) return rc;
    if (dt > 0.0f) ctx->inv_dt0 = inv_dt;
    if (ctx->autoClearForces && ctx->nBodies > 0)
        hipLaunchKernelGGL(k_fill<float4>, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, ctx->force, ctx->nBodies,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    LAUNCHCHK(ctx);
    profCollect(ctx);
    ctx->stats.bodies = ctx->nBodies;
    ctx->stats.fixtures = ctx->nFixtures;
    ctx->stats.contacts = ctx->nContacts;
    ctx->stats.overflow = ctx->hcnt->overflow;
    ctx->stats.cellSize = ctx->cellSize;
    return B2G_OK;
}

void freeRef(b2g_context* ctx) {
    RefView& r = ctx->ref;
    freeArray(r.tmpMan); freeArray(r.overlap); freeArray(r.edgeCount); freeArray(r.edgeStart);
    freeArray(r.edgeCursor); freeArray(r.edgeKeys); freeArray(r.edgeSorted); freeArray(r.edgeContact);
    freeArray(r.bvisited); freeArray(r.cvisited); freeArray(r.stack); freeArray(r.islandBodies);
    freeArray(r.islandContacts); freeArray(r.islandRange); freeArray(r.islandSlept); freeArray(r.lastIsland);
    freeArray(r.jvisited); freeArray(r.islandJoints); freeArray(r.islandJointRange);
    ctx->refCapBodies = ctx->refCapContacts = ctx->refCapJoints = 0;
}

int ensureRefBuffers(b2g_context* ctx) {
    if (ctx->refCapBodies >= ctx->capBodies && ctx->refCapContacts >= ctx->capContacts &&
        ctx->refCapJoints >= ctx->capJoints)
        return B2G_OK;
    freeRef(ctx);
    RefView& r = ctx->ref;
    const size_t nb = (size_t)ctx->capBodies, nc = (size_t)ctx->capContacts;
    const size_t nj = (size_t)std::max(ctx->capJoints, 1);
    HIPCHK(ctx, hipMalloc((void**)&r.tmpMan, sizeof(b2g_manifold) * nc));
    HIPCHK(ctx, hipMalloc((void**)&r.overlap, nc));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeCount, sizeof(int) * (nb + 1)));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeStart, sizeof(int) * (nb + 1)));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeCursor, sizeof(int) * nb));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeKeys, sizeof(unsigned long long) * 2 * (nc + nj)));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeSorted, sizeof(unsigned long long) * 2 * (nc + nj)));
    HIPCHK(ctx, hipMalloc((void**)&r.edgeContact, sizeof(int) * 2 * (nc + nj)));
    HIPCHK(ctx, hipMalloc((void**)&r.bvisited, sizeof(int) * nb));
    HIPCHK(ctx, hipMalloc((void**)&r.cvisited, sizeof(int) * nc));
    HIPCHK(ctx, hipMalloc((void**)&r.stack, sizeof(int) * nb));
    // a static body joins each island that reaches it through a contact or a joint
    HIPCHK(ctx, hipMalloc((void**)&r.islandBodies, sizeof(int) * (nb + nc + nj)));
    HIPCHK(ctx, hipMalloc((void**)&r.islandContacts, sizeof(int) * nc));
    HIPCHK(ctx, hipMalloc((void**)&r.islandRange, sizeof(int4) * nb));
    HIPCHK(ctx, hipMalloc((void**)&r.islandSlept, sizeof(int) * nb));
    HIPCHK(ctx, hipMalloc((void**)&r.lastIsland, sizeof(int) * nb));
    HIPCHK(ctx, hipMalloc((void**)&r.jvisited, sizeof(int) * nj));
    HIPCHK(ctx, hipMalloc((void**)&r.islandJoints, sizeof(int) * nj));
    HIPCHK(ctx, hipMalloc((void**)&r.islandJointRange, sizeof(int2) * nb));
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks((int)nb)), dim3(256), 0, ctx->stream, r.lastIsland, (int)nb, -1);
    ctx->refCapBodies = (int)nb;
    ctx->refCapContacts = (int)nc;
    ctx->refCapJoints = (int)nj;
    // sorting 2 edges per contact and joint
    const size_t need = prims_scratch_bytes((int)std::max<size_t>(2 * (nc + nj), 1024));
    if (need > ctx->scratchBytes) {
        freeArray(ctx->scratch);
        HIPCHK(ctx, hipMalloc(&ctx->scratch, need));
        ctx->scratchBytes = need;
    }
    return B2G_OK;
}

int stepColoured(b2g_context* ctx, const b2g_step_params* p) {
    const int rc = stepColouredCollide(ctx, p);
    return rc != B2G_OK ? rc : stepColouredSolve(ctx, p);
}

int ensureToi(b2g_context* ctx) {
    if (ctx->toiState) return B2G_OK;
    HIPCHK(ctx, hipMalloc((void**)&ctx->toiState, sizeof(ToiState)));
    HIPCHK(ctx, hipHostMalloc((void**)&ctx->hToi, sizeof(ToiState), hipHostMallocDefault));
    memset(ctx->hToi, 0, sizeof(ToiState));
    HIPCHK(ctx, hipMalloc((void**)&ctx->toiGather, sizeof(unsigned long long) * kToiGatherCap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->toiMan, sizeof(b2g_manifold) * kToiGatherCap));
    HIPCHK(ctx, hipMalloc((void**)&ctx->toiTouch, sizeof(int) * kToiGatherCap));
    ConView& c = ctx->toiCv;
    const size_t n = kMaxTOIContacts;
    HIPCHK(ctx, hipMalloc((void**)&c.contact, sizeof(int) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.bodies, sizeof(int2) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.mass, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.normal, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.r0, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.r1, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.pmass, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.bias, sizeof(float2) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.imp, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.K, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.NM, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.info, sizeof(int) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.lp, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.lnp, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.lc, sizeof(float4) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.rad, sizeof(float2) * n));
    HIPCHK(ctx, hipMalloc((void**)&c.island, sizeof(int) * n));
    return B2G_OK;
}

void freeToi(b2g_context* ctx) {
    freeArray(ctx->toiState);
    if (ctx->hToi) (void)hipHostFree(ctx->hToi);
    ctx->hToi = nullptr;
    freeArray(ctx->toiGather);
    freeArray(ctx->toiMan);
    freeArray(ctx->toiTouch);
    freeArray(ctx->toiSeen);
    ctx->toiSeenCap = 0;
    ConView& c = ctx->toiCv;
    freeArray(c.contact); freeArray(c.bodies); freeArray(c.mass); freeArray(c.normal); freeArray(c.r0);
    freeArray(c.r1); freeArray(c.pmass); freeArray(c.bias); freeArray(c.imp); freeArray(c.K); freeArray(c.NM);
    freeArray(c.info); freeArray(c.lp); freeArray(c.lnp); freeArray(c.lc); freeArray(c.rad); freeArray(c.island);
}

__global__ void k_toi_start(ToiState* st, int live) {
    if (threadIdx.x != 0) return;
    st->done = 0;
    st->events = 0;
    st->live = live;
    st->added = 0;
    st->halt = 0;
    st->subStop = 0;
    toiResetEvent(st);
}

// After the host ran the grid broad-phase for an event: resume the batch.
__global__ void k_toi_resume(ToiState* st, int live) {
    if (threadIdx.x != 0) return;
    st->live = live;
    st->halt = 0;
    toiResetEvent(st);
}

// b2World::SolveTOI, REF/Dynamics/b2World.cpp:579-897. The events run on the
// device, one launch sequence each, in batches that grow geometrically;
// every kernel of an event does nothing once the pass has halted (no event
// left, a list overflowed, or the sub-step is over), and the host reads one
// small record per batch. The contact count lives on the device during the
// pass, so contact kernels are launched over the table's capacity.
int solveTOI(b2g_context* ctx, const b2g_step_params* p) {
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = ensureToi(ctx)) != B2G_OK) return rc;
    if ((rc = reserveContacts(ctx)) != B2G_OK) return rc;
    // the step's broad-phase appended contacts the pair hash does not hold yet
    if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
    // room for the pairs events add; a grown table gets its hash rebuilt
    auto reserve = [&]() -> int {
        const int before = ctx->capContacts;
        int r = reserveContacts(ctx);
        if (r == B2G_OK && ctx->capContacts != before) r = rebuildHash(ctx);
        return r;
    };
    ctx->hToi->events = 0;
    ToiState* st = ctx->toiState;
    if (ctx->stepComplete) {
        WorldView w = ctx->view();
        if (ctx->nBodies > 0)
            hipLaunchKernelGGL(k_toi_reset_bodies, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, w);
        if (ctx->nContacts > 0)
            hipLaunchKernelGGL(k_toi_reset_contacts, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w);
    }
    hipLaunchKernelGGL(k_toi_start, dim3(1), dim3(64), 0, s, st, ctx->nContacts);
    LAUNCHCHK(ctx);
    const int pe = profStart(ctx);
    int batch = 1;
    for (;;) {
        // every event raises some contact's TOI count or disables it, so a
        // pass ends within kMaxSubSteps + 1 events per contact
        if ((long long)ctx->hToi->events > (long long)(kMaxSubSteps + 2) * ((long long)ctx->nContacts + 1))
            return fail(ctx, B2G_ERR_STATE, "TOI events did not converge");
        if (ctx->toiSeenCap < ctx->capContacts) {
            freeArray(ctx->toiSeen);
            HIPCHK(ctx, hipMalloc((void**)&ctx->toiSeen, sizeof(Sweep) * 2 * (size_t)ctx->capContacts));
            ctx->toiSeenCap = ctx->capContacts;
        }
        const WorldView w = ctx->view();
        const int cap = ctx->capContacts;
        for (int e = 0; e < batch; ++e) {
            hipLaunchKernelGGL(k_toi_classify, dim3(blocks(cap)), dim3(256), 0, s, w, ctx->keep, st);
            hipLaunchKernelGGL(k_toi_align_serial, dim3(1), dim3(64), 0, s, w, ctx->keep, st, ctx->toiSeen);
            hipLaunchKernelGGL(k_toi_compute, dim3(blocks(cap, 64)), dim3(64), 0, s, w, ctx->keep, st, ctx->toiSeen);
            hipLaunchKernelGGL(k_toi_min, dim3(blocks(cap)), dim3(256), 0, s, w, st);
            hipLaunchKernelGGL(k_toi_event, dim3(1), dim3(64), 0, s, w, st);
            hipLaunchKernelGGL(k_toi_gather, dim3(blocks(cap)), dim3(256), 0, s, w, st, ctx->toiGather);
            hipLaunchKernelGGL(k_toi_prep, dim3(blocks(kToiGatherCap, 64)), dim3(64), 0, s, w, st, ctx->toiGather,
                               ctx->toiMan, ctx->toiTouch);
            hipLaunchKernelGGL(k_toi_island, dim3(1), dim3(256), 0, s, w, ctx->toiCv, st, ctx->toiGather, ctx->toiMan,
                               ctx->toiTouch, p->dt, p->velocityIterations);
            hipLaunchKernelGGL(k_toi_displaced, dim3(blocks(std::max(cap, ctx->nFixtures))), dim3(256), 0, s, w, st);
            if (ctx->nFixtures > 0)
                hipLaunchKernelGGL(k_toi_pairs, dim3(blocks(ctx->nFixtures)), dim3(256), 0, s, w, st, ctx->hashTable,
                                   ctx->hashMask);
            hipLaunchKernelGGL(k_toi_append, dim3(1), dim3(256), 0, s, w, st, ctx->hashTable, ctx->hashMask, cap,
                               ctx->subStepping ? 1 : 0);
        }
        LAUNCHCHK(ctx);
        HIPCHK(ctx, hipMemcpyAsync(ctx->hToi, st, kToiReadback, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        const ToiState& h = *ctx->hToi;
        ctx->nContacts = h.live;
        if (h.done) {
            ctx->stepComplete = true;
            break;
        }
        if (h.fallback) {
            // too many moved proxies or new pairs for the event's lists, or
            // no room: the grid broad-phase over the moved flags, as at the
            // end of a step
            const int before = ctx->nContacts;
            if ((rc = runBroadphase(ctx)) != B2G_OK) return rc;
            const int added = ctx->nContacts - before;
            if (added > 0) {
                hipLaunchKernelGGL(k_hash_insert_range, dim3(blocks(added)), dim3(256), 0, s, ctx->view(), before,
                                   added, ctx->hashTable, ctx->hashMask);
            }
            if ((rc = reserve()) != B2G_OK) return rc;
            hipLaunchKernelGGL(k_toi_resume, dim3(1), dim3(64), 0, s, st, ctx->nContacts);
            LAUNCHCHK(ctx);
            if (ctx->subStepping) {
                ctx->stepComplete = false;
                break;
            }
            batch = 1;
            continue;
        }
        if (h.subStop) {
            ctx->stepComplete = false;
            break;
        }
        if ((rc = reserve()) != B2G_OK) return rc;
        batch = std::min(2 * batch, 16);
    }
    profEnd(ctx, PH_TOI, pe);
    ctx->stats.toiEvents = ctx->hToi->events;
    ctx->stats.newPairs += ctx->hToi->added;
    return B2G_OK;
}

// b2World::Step in the reference's visiting order (see the reference-order
// kernels above). Same collide/broad-phase as the coloured path; islands
// and constraint order follow Box2D exactly.
// First half of a reference-order step: counters, pending broad-phase and
// b2ContactManager::Collide in list order (the contacts it destroys stay in
// the table, flagged dead, until the second half compacts it). With
// ctx->keepOldManifolds the manifolds from before the update are kept for
// PreSolve.
int stepReferenceCollide(b2g_context* ctx, const b2g_step_params* p) {
    (void)p;
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = reserveContacts(ctx)) != B2G_OK) return rc;
    if ((rc = ensureRefBuffers(ctx)) != B2G_OK) return rc;
    memset(&ctx->stats, 0, sizeof(ctx->stats));
    Counters zero;
    memset(&zero, 0, sizeof(zero));
    HIPCHK(ctx, hipMemcpyAsync(ctx->cnt, &zero, sizeof(Counters), hipMemcpyHostToDevice, s));
    if (ctx->newFixture) {
        if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
        if ((rc = runBroadphase(ctx)) != B2G_OK) return rc;
        ctx->newFixture = false;
    }
    RefView r = ctx->ref;
    if (ctx->nContacts > 0) {
        WorldView w = ctx->view();
        if (ctx->keepOldManifolds) {
            if ((rc = ensureOldManifolds(ctx)) != B2G_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(ctx->oldMan, ctx->cur.cman, sizeof(b2g_manifold) * ctx->nContacts,
                                       hipMemcpyDeviceToDevice, s));
        }
        hipLaunchKernelGGL(k_collide_compute, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, r);
        hipLaunchKernelGGL(k_collide_serial, dim3(1), dim3(64), 0, s, w, r, ctx->cnt);
        LAUNCHCHK(ctx);
    }
    return B2G_OK;
}

// Second half: compaction, Box2D's island search and island solves, the
// broad-phase and continuous collision.
int stepReferenceSolve(b2g_context* ctx, const b2g_step_params* p) {
    hipStream_t s = ctx->stream;
    const float dt = p->dt;
    const float inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
    const float dtRatio = ctx->inv_dt0 * dt;
    int rc;
    RefView r = ctx->ref;
    if (ctx->nContacts > 0) {
        WorldView w = ctx->view();
        hipLaunchKernelGGL(k_mark_dead, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, w, ctx->keep);
        LAUNCHCHK(ctx);
        HIPCHK(ctx, prims_exclusive_sum(ctx->scratch, ctx->scratchBytes, ctx->keep, ctx->idx, ctx->nContacts, s));
        hipLaunchKernelGGL(k_compact, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->cur, ctx->alt, ctx->keep,
                           ctx->idx, ctx->nContacts, ctx->cnt);
        LAUNCHCHK(ctx);
        std::swap(ctx->cur, ctx->alt);
        if ((rc = readCounters(ctx)) != B2G_OK) return rc;
        ctx->nContacts = ctx->hcnt->nContacts;
    } else {
        if ((rc = readCounters(ctx)) != B2G_OK) return rc;
    }
    ctx->stats.beginEvents = ctx->hcnt->beginEvents;
    ctx->stats.endEvents = ctx->hcnt->endEvents;
    if (dt > 0.0f && ctx->stepComplete) {
        WorldView w = ctx->view();
        const int nb = ctx->nBodies, nc = ctx->nContacts;
        JointView jv = ctx->jview();
        jv.h = dt;
        const int nj = ctx->nJoints;
        hipLaunchKernelGGL(k_clear_island_flags, dim3(blocks(nb + 1)), dim3(256), 0, s, w, r);
        if (nc > 0) hipLaunchKernelGGL(k_edge_count, dim3(blocks(nc)), dim3(256), 0, s, w, r, ctx->cnt);
        if (nj > 0) hipLaunchKernelGGL(k_joint_edge_count, dim3(blocks(nj)), dim3(256), 0, s, w, jv, r, ctx->cnt);
        LAUNCHCHK(ctx);
        HIPCHK(ctx, prims_exclusive_sum(ctx->scratch, ctx->scratchBytes, r.edgeCount, r.edgeStart, nb + 1, s));
        if (nc > 0) hipLaunchKernelGGL(k_edge_fill, dim3(blocks(nc)), dim3(256), 0, s, w, r);
        if (nj > 0) hipLaunchKernelGGL(k_joint_edge_fill, dim3(blocks(nj)), dim3(256), 0, s, w, jv, r);
        LAUNCHCHK(ctx);
        if ((rc = readCounters(ctx)) != B2G_OK) return rc;
        const int nEdges = ctx->hcnt->nEdges;
        if (nEdges > 0) {
            HIPCHK(ctx, prims_sort_keys_u64(ctx->scratch, ctx->scratchBytes, r.edgeKeys, r.edgeSorted, nEdges, s));
            hipLaunchKernelGGL(k_edge_extract, dim3(blocks(nEdges)), dim3(256), 0, s, r, nEdges);
        }
        hipLaunchKernelGGL(k_dfs, dim3(1), dim3(64), 0, s, w, jv, r, ctx->cnt);
        LAUNCHCHK(ctx);
        if ((rc = readCounters(ctx)) != B2G_OK) return rc;
        const int nIslands = ctx->hcnt->nIslands;
        const int posIters = std::min(p->positionIterations, kMaxPosIters);
        if (nIslands > 0)
            hipLaunchKernelGGL(k_island_solve, dim3(nIslands), dim3(64), 0, s, w, ctx->cv, jv, r, ctx->gravity, dt,
                               dtRatio, ctx->warmStarting ? 1 : 0, p->velocityIterations, posIters,
                               ctx->allowSleep ? 1 : 0, ctx->cnt);
        if (nIslands > 0 && ctx->allowSleep) {
            hipLaunchKernelGGL(k_static_last_island, dim3(nIslands), dim3(64), 0, s, w, r, nIslands);
            hipLaunchKernelGGL(k_static_awake, dim3(blocks(nb)), dim3(256), 0, s, w, r);
        }
        hipLaunchKernelGGL(k_sync_fixtures, dim3(blocks(ctx->nFixtures)), dim3(256), 0, s, w);
        LAUNCHCHK(ctx);
        if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
        if ((rc = runBroadphase(ctx)) != B2G_OK) return rc;
        ctx->stats.constraints = ctx->hcnt->nConstraints;
        ctx->stats.points = ctx->hcnt->points;
    }
    if (ctx->continuous && dt > 0.0f && (rc = solveTOI(ctx, p)) != B2G_OK) return rc;
    if (dt > 0.0f) ctx->inv_dt0 = inv_dt;
    if (ctx->autoClearForces && ctx->nBodies > 0)
        hipLaunchKernelGGL(k_fill<float4>, dim3(blocks(ctx->nBodies)), dim3(256), 0, s, ctx->force, ctx->nBodies,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    LAUNCHCHK(ctx);
    ctx->stats.bodies = ctx->nBodies;
    ctx->stats.fixtures = ctx->nFixtures;
    ctx->stats.contacts = ctx->nContacts;
    ctx->stats.overflow = ctx->hcnt->overflow;
    ctx->stats.cellSize = ctx->cellSize;
    return B2G_OK;
}

int stepReference(b2g_context* ctx, const b2g_step_params* p) {
    const int rc = stepReferenceCollide(ctx, p);
    return rc != B2G_OK ? rc : stepReferenceSolve(ctx, p);
}

void bodyToDevice(const b2g_body_state& b, int& type, int& flags, float4& xf, float4& pos, float4& pos0,
                  float2& lc, float4& vel, float4& force, float4& mass, float2& gs) {
    type = b.type;
    flags = b.flags;
    xf = make_float4(b.xf[0], b.xf[1], b.xf[2], b.xf[3]);
    pos = make_float4(b.c[0], b.c[1], b.a, 0.0f);
    pos0 = make_float4(b.c0[0], b.c0[1], b.a0, 0.0f);
    lc = make_float2(b.lc[0], b.lc[1]);
    vel = make_float4(b.v[0], b.v[1], b.w, 0.0f);
    force = make_float4(b.force[0], b.force[1], b.torque, 0.0f);
    mass = make_float4(b.invMass, b.invI, b.linearDamping, b.angularDamping);
    gs = make_float2(b.gravityScale, b.sleepTime);
}

int uploadBodies(b2g_context* ctx, int first, int count, const b2g_body_state* bodies) {
    std::vector<int> type(count), flags(count);
    std::vector<float4> xf(count), pos(count), pos0(count), vel(count), force(count), mass(count);
    std::vector<float2> lc(count), gs(count);
    for (int i = 0; i < count; ++i) {
        bodyToDevice(bodies[i], type[i], flags[i], xf[i], pos[i], pos0[i], lc[i], vel[i], force[i], mass[i], gs[i]);
        if (type[i] == BT_STATIC && !(flags[i] & BF_AWAKE)) ctx->staticAsleep = true;
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->btype + first, type.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->bflags + first, flags.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->xf + first, xf.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->pos + first, pos.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->pos0 + first, pos0.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->vel + first, vel.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->force + first, force.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->mass + first, mass.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->lcenter + first, lc.data(), sizeof(float2) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->gsleep + first, gs.data(), sizeof(float2) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

}  // namespace

// ====================================================================
// C ABI
// ====================================================================
namespace {
void noallocInit(int device);
}  // namespace

extern "C" {

int b2g_create(b2g_context** out, int device) {
    if (!out) return B2G_ERR_ARG;
    *out = nullptr;
    b2g_context* ctx = new b2g_context;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return B2G_ERR_HIP;
    }
    noallocInit(device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return B2G_ERR_HIP;
    }
    if (hipMalloc((void**)&ctx->cnt, sizeof(Counters)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->hcnt, sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&ctx->penAcc, 64) != hipSuccess) {
        delete ctx;
        return B2G_ERR_HIP;
    }
    memset(ctx->hcnt, 0, sizeof(Counters));
    int rc = growBodies(ctx, 64);
    if (rc == B2G_OK) rc = growFixtures(ctx, 64);
    if (rc != B2G_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return B2G_OK;
}

int b2g_destroy(b2g_context* ctx) {
    if (!ctx) return B2G_ERR_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    freeArray(ctx->btype); freeArray(ctx->bflags); freeArray(ctx->xf); freeArray(ctx->pos); freeArray(ctx->pos0);
    freeArray(ctx->vel); freeArray(ctx->force); freeArray(ctx->mass); freeArray(ctx->lcenter);
    freeArray(ctx->gsleep); freeArray(ctx->colourMask); freeArray(ctx->claim); freeArray(ctx->islandOf);
    freeArray(ctx->bworld);
    freeArray(ctx->islandAwake); freeArray(ctx->islandSleep); freeArray(ctx->islandMinSep);
    freeArray(ctx->islandMax); freeArray(ctx->staticKey);
    freeArray(ctx->bodyKey); freeArray(ctx->bodyKeySorted); freeArray(ctx->bodyCluster);
    freeArray(ctx->clusterBodies); freeArray(ctx->clusterStart);
    for (auto& g : ctx->solveGraphs) (void)hipGraphExecDestroy(g.exec);
    ctx->solveGraphs.clear();
    freeArray(ctx->dRanges);
    freeArray(ctx->oldMan);
    if (ctx->hRanges) (void)hipHostFree(ctx->hRanges);
    ctx->hRanges = nullptr;
    freeArray(ctx->fbody); freeArray(ctx->shape); freeArray(ctx->fmat); freeArray(ctx->ffilter);
    freeArray(ctx->fgroup); freeArray(ctx->fat); freeArray(ctx->moved); freeArray(ctx->range);
    for (ContactArrays* a : {&ctx->cur, &ctx->alt}) {
        freeArray(a->cfA); freeArray(a->cfB); freeArray(a->cflags); freeArray(a->cmat); freeArray(a->ctan); freeArray(a->cman);
        freeArray(a->ccolour); freeArray(a->ctoi); freeArray(a->ctoiCount);
    }
    freeToi(ctx);
    freeArray(ctx->keep); freeArray(ctx->idx); freeArray(ctx->hashTable); freeArray(ctx->newKeys);
    freeArray(ctx->newKeysSorted); freeArray(ctx->consKeys); freeArray(ctx->consKeysSorted);
    freeArray(ctx->fixOrder); freeArray(ctx->orderFix);
    freeArray(ctx->consVals); freeArray(ctx->consOrder); freeArray(ctx->tileList);
    ConView& c = ctx->cv;
    freeArray(c.contact); freeArray(c.bodies); freeArray(c.mass); freeArray(c.normal); freeArray(c.r0);
    freeArray(c.r1); freeArray(c.pmass); freeArray(c.bias); freeArray(c.imp); freeArray(c.K); freeArray(c.NM);
    freeArray(c.info); freeArray(c.lp); freeArray(c.lnp); freeArray(c.lc); freeArray(c.rad); freeArray(c.island);
    freeArray(c.slot);
    freeArray(ctx->gCounts); freeArray(ctx->gStarts); freeArray(ctx->gCursor); freeArray(ctx->gEntries);
    freeArray(ctx->gBig); freeArray(ctx->gBigSorted); freeArray(ctx->gCell); freeArray(ctx->gEfat);
    freeArray(ctx->gQueries);
    freeRef(ctx);
    freeArray(ctx->freshMask); freeArray(ctx->freshNum); freeArray(ctx->freshOff); freeArray(ctx->freshAll);
    freeArray(ctx->freshFeat); freeArray(ctx->freshRef); freeArray(ctx->impSnap); freeArray(ctx->impAlt);
    freeArray(ctx->velSnap); freeArray(ctx->logFeat); freeArray(ctx->logTarget);
    freeArray(ctx->jBodies); freeArray(ctx->jAnchors); freeArray(ctx->jLimits); freeArray(ctx->jImpulse);
    freeArray(ctx->jR); freeArray(ctx->jMasses); freeArray(ctx->jK0); freeArray(ctx->jK1);
    freeArray(ctx->jMotorSpeed); freeArray(ctx->jMotorMass); freeArray(ctx->jAxis); freeArray(ctx->jExtra); freeArray(ctx->jFlags); freeArray(ctx->jOrder);
    freeArray(ctx->jImpulseSnap); freeArray(ctx->jFlagsSnap); freeArray(ctx->jointPairs);
    if (ctx->trainScratch) (void)hipFree(ctx->trainScratch);
    if (ctx->mlpAllocated) b2mlp::mlpFree(ctx->mlp);
    for (hipEvent_t e : ctx->evPool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->cnt) (void)hipFree(ctx->cnt);
    if (ctx->penAcc) (void)hipFree(ctx->penAcc);
    if (ctx->hcnt) (void)hipHostFree(ctx->hcnt);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return B2G_OK;
}

const char* b2g_last_error(const b2g_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int b2g_reserve(b2g_context* ctx, int maxBodies, int maxFixtures, int maxContacts) {
    if (!ctx || maxBodies < 0 || maxFixtures < 0 || maxContacts < 0) return B2G_ERR_ARG;
    int rc = growBodies(ctx, maxBodies);
    if (rc == B2G_OK) rc = growFixtures(ctx, maxFixtures);
    if (rc == B2G_OK) rc = growContacts(ctx, maxContacts);
    return rc;
}

int b2g_set_world(b2g_context* ctx, const b2g_world_def* def) {
    if (!ctx || !def) return B2G_ERR_ARG;
    ctx->gravity = mk(def->gravity[0], def->gravity[1]);
    ctx->allowSleep = def->allowSleep != 0;
    ctx->warmStarting = def->warmStarting != 0;
    ctx->continuous = def->continuous != 0;
    ctx->subStepping = def->subStepping != 0;
    ctx->autoClearForces = def->autoClearForces != 0;
    return B2G_OK;
}

int b2g_add_bodies(b2g_context* ctx, const b2g_body_state* bodies, int count, int* first) {
    if (!ctx || (count > 0 && !bodies) || count < 0) return B2G_ERR_ARG;
    int rc = growBodies(ctx, ctx->nBodies + count);
    if (rc != B2G_OK) return rc;
    const int f = ctx->nBodies;
    ctx->nBodies += count;
    if (count > 0) {
        rc = uploadBodies(ctx, f, count, bodies);
        if (rc != B2G_OK) return rc;
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(count)), dim3(256), 0, ctx->stream,
                           ctx->colourMask + f, count, 0ull);
        LAUNCHCHK(ctx);
    }
    if (first) *first = f;
    return B2G_OK;
}

int b2g_set_bodies(b2g_context* ctx, int first, int count, const b2g_body_state* bodies) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->nBodies || (count > 0 && !bodies)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    return uploadBodies(ctx, first, count, bodies);
}

int b2g_get_bodies(b2g_context* ctx, int first, int count, b2g_body_state* out) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->nBodies || (count > 0 && !out)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    std::vector<int> type(count), flags(count);
    std::vector<float4> xf(count), pos(count), pos0(count), vel(count), force(count), mass(count);
    std::vector<float2> lc(count), gs(count);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(type.data(), ctx->btype + first, sizeof(int) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(flags.data(), ctx->bflags + first, sizeof(int) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(xf.data(), ctx->xf + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(pos.data(), ctx->pos + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(pos0.data(), ctx->pos0 + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(vel.data(), ctx->vel + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(force.data(), ctx->force + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(mass.data(), ctx->mass + first, sizeof(float4) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(lc.data(), ctx->lcenter + first, sizeof(float2) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(gs.data(), ctx->gsleep + first, sizeof(float2) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    for (int i = 0; i < count; ++i) {
        b2g_body_state& b = out[i];
        b.type = type[i];
        b.flags = flags[i];
        b.xf[0] = xf[i].x; b.xf[1] = xf[i].y; b.xf[2] = xf[i].z; b.xf[3] = xf[i].w;
        b.lc[0] = lc[i].x; b.lc[1] = lc[i].y;
        b.c0[0] = pos0[i].x; b.c0[1] = pos0[i].y; b.a0 = pos0[i].z;
        b.c[0] = pos[i].x; b.c[1] = pos[i].y; b.a = pos[i].z;
        b.v[0] = vel[i].x; b.v[1] = vel[i].y; b.w = vel[i].z;
        b.force[0] = force[i].x; b.force[1] = force[i].y; b.torque = force[i].z;
        b.invMass = mass[i].x; b.invI = mass[i].y; b.linearDamping = mass[i].z; b.angularDamping = mass[i].w;
        b.gravityScale = gs[i].x; b.sleepTime = gs[i].y;
    }
    return B2G_OK;
}

}  // extern "C"

// b2Fixture::CreateProxies' box: the shape's AABB at the body transform,
// fattened by the AABB extension (REF/Collision/b2DynamicTree.cpp:112-126);
// chain children have no skin in theirs (b2ChainShape::ComputeAABB).
static float4 proxyFatBox(const b2g_shape& s, const float bxf[4]) {
    Xf xf;
    xf.p = mk(bxf[0], bxf[1]);
    xf.q.s = bxf[2];
    xf.q.c = bxf[3];
    float lx, ly, ux, uy;
    if (s.type == B2G_CIRCLE) {
        const V2 p = xf.p + mul(xf.q, ld(s.centroid));
        lx = p.x - s.radius; ly = p.y - s.radius; ux = p.x + s.radius; uy = p.y + s.radius;
    } else {
        const int nv = s.type == B2G_EDGE ? 2 : s.count;
        V2 lo = mul(xf, ld(s.verts[0])), hi = lo;
        for (int k = 1; k < nv; ++k) {
            const V2 v = mul(xf, ld(s.verts[k]));
            lo = vmin(lo, v);
            hi = vmax(hi, v);
        }
        const float rad = (s.type == B2G_EDGE && (s.flags & 4)) ? 0.0f : s.radius;
        const V2 r = mk(rad, rad);
        const V2 l = lo - r, u = hi + r;
        lx = l.x; ly = l.y; ux = u.x; uy = u.y;
    }
    return make_float4(lx - kAabbExtension, ly - kAabbExtension, ux + kAabbExtension, uy + kAabbExtension);
}

extern "C" {

int b2g_add_fixtures(b2g_context* ctx, const b2g_fixture_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        if (defs[i].body < 0 || defs[i].body >= ctx->nBodies) return fail(ctx, B2G_ERR_ARG, "fixture body index");
        const int t = defs[i].shape.type;
        if (t != B2G_CIRCLE && t != B2G_EDGE && t != B2G_POLYGON)
            return fail(ctx, B2G_ERR_UNSUPPORTED, "shape type");
        if (t == B2G_POLYGON && (defs[i].shape.count < 3 || defs[i].shape.count > 8))
            return fail(ctx, B2G_ERR_ARG, "polygon vertex count");
    }
    int rc = growFixtures(ctx, ctx->nFixtures + count);
    if (rc != B2G_OK) return rc;
    const int f = ctx->nFixtures;
    if (count > 0) {
        // body types and transforms for the initial AABBs
        std::vector<b2g_body_state> bodies(ctx->nBodies);
        rc = b2g_get_bodies(ctx, 0, ctx->nBodies, bodies.data());
        if (rc != B2G_OK) return rc;
        std::vector<int> types(ctx->nBodies);
        for (int i = 0; i < ctx->nBodies; ++i) types[i] = bodies[i].type;
        std::vector<int> fbody(count), fgroup(count);
        std::vector<uint32_t> ffilter(count);
        std::vector<float4> fmat(count), fat(count);
        std::vector<uint8_t> moved(count, 0);
        for (int i = 0; i < count; ++i) {
            const b2g_fixture_def& d = defs[i];
            fbody[i] = d.body;
            fgroup[i] = d.groupIndex;
            ffilter[i] = (d.categoryBits & 0xFFFF) | ((d.maskBits & 0xFFFF) << 16);
            fmat[i] = make_float4(d.friction, d.restitution, d.density, d.sensor ? 1.0f : 0.0f);
            const b2g_body_state& b = bodies[d.body];
            const bool active = (b.flags & BF_ACTIVE) != 0;
            fat[i] = proxyFatBox(d.shape, b.xf);
            // an inactive body's fixtures have no proxy (b2Body::CreateFixture,
            // REF/Dynamics/b2Body.cpp:182-186): the inverted box of a removed one
            if (!active) fat[i] = make_float4(FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX);
            moved[i] = active ? 1 : 0;
        }
        std::vector<b2g_shape> shapes(count);
        for (int i = 0; i < count; ++i) shapes[i] = defs[i].shape;
        hipStream_t s = ctx->stream;
        HIPCHK(ctx, hipMemcpyAsync(ctx->fbody + f, fbody.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->shape + f, shapes.data(), sizeof(b2g_shape) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->fmat + f, fmat.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->ffilter + f, ffilter.data(), sizeof(uint32_t) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->fgroup + f, fgroup.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->fat + f, fat.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->moved + f, moved.data(), count, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        chooseCellSize(ctx, defs, count, types);
        ctx->nFixtures += count;
        ctx->newFixture = true;
        rc = growContacts(ctx, 8 * ctx->nFixtures + 4096);
        if (rc != B2G_OK) return rc;
    }
    if (first) *first = f;
    return B2G_OK;
}

int b2g_restore_fixtures(b2g_context* ctx, const int32_t* ids, int count) {
    if (!ctx || count < 0 || (count > 0 && !ids)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    for (int i = 0; i < count; ++i)
        if (ids[i] < 0 || ids[i] >= ctx->nFixtures) return fail(ctx, B2G_ERR_ARG, "fixture index");
    hipStream_t s = ctx->stream;
    std::vector<int> body(count);
    std::vector<b2g_shape> shape(count);
    for (int i = 0; i < count; ++i) {
        HIPCHK(ctx, hipMemcpyAsync(&body[i], ctx->fbody + ids[i], sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(&shape[i], ctx->shape + ids[i], sizeof(b2g_shape), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    std::vector<float4> xf(count), fat(count);
    for (int i = 0; i < count; ++i)
        HIPCHK(ctx, hipMemcpyAsync(&xf[i], ctx->xf + body[i], sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    const uint8_t one = 1;
    for (int i = 0; i < count; ++i) {
        const float bxf[4] = {xf[i].x, xf[i].y, xf[i].z, xf[i].w};
        fat[i] = proxyFatBox(shape[i], bxf);
        HIPCHK(ctx, hipMemcpyAsync(ctx->fat + ids[i], &fat[i], sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->moved + ids[i], &one, 1, hipMemcpyHostToDevice, s));
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_remove_fixtures(b2g_context* ctx, const int32_t* ids, int count) {
    if (!ctx || count < 0 || (count > 0 && !ids)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    for (int i = 0; i < count; ++i)
        if (ids[i] < 0 || ids[i] >= ctx->nFixtures) return fail(ctx, B2G_ERR_ARG, "fixture index");
    int* d = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&d, sizeof(int) * count));
    hipError_t e = hipMemcpyAsync(d, ids, sizeof(int) * count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_remove_fixtures, dim3(blocks(count)), dim3(256), 0, ctx->stream, ctx->view(), d, count);
        if (ctx->nContacts > 0)
            hipLaunchKernelGGL(k_wake_removed_contacts, dim3(blocks(ctx->nContacts)), dim3(256), 0, ctx->stream,
                               ctx->view());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, B2G_ERR_HIP, hipGetErrorString(e));
    return B2G_OK;
}

int b2g_drop_body_contacts(b2g_context* ctx, const int32_t* bodies, int count) {
    if (!ctx || count < 0 || (count > 0 && !bodies)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    std::vector<uint8_t> mark(std::max(ctx->nBodies, 1), 0);
    for (int i = 0; i < count; ++i) {
        if (bodies[i] < 0 || bodies[i] >= ctx->nBodies) return fail(ctx, B2G_ERR_ARG, "body index");
        mark[bodies[i]] = 1;
    }
    uint8_t* d = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&d, mark.size()));
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpyAsync(d, mark.data(), mark.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && ctx->nContacts > 0)
        hipLaunchKernelGGL(k_drop_body_contacts, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), d);
    if (e == hipSuccess && ctx->nFixtures > 0)
        hipLaunchKernelGGL(k_touch_body_proxies, dim3(blocks(ctx->nFixtures)), dim3(256), 0, s, ctx->view(), d);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, B2G_ERR_HIP, hipGetErrorString(e));
    return B2G_OK;
}

int b2g_set_proxy_ids(b2g_context* ctx, const int32_t* ids, int count) {
    if (!ctx || count < 0 || (count > 0 && !ids)) return B2G_ERR_ARG;
    if (ctx->subWorlds) return fail(ctx, B2G_ERR_STATE, "proxy ids: one world per context only");
    if (count != ctx->nFixtures)
        return fail(ctx, B2G_ERR_ARG, "proxy ids: one per fixture (" + std::to_string(ctx->nFixtures) + ")");
    int span = 0;
    for (int i = 0; i < count; ++i) {
        if (ids[i] < -1 || ids[i] >= (1 << 30)) return fail(ctx, B2G_ERR_ARG, "proxy id out of range");
        span = std::max(span, ids[i] + 1);
    }
    std::vector<int> inverse(std::max(span, 1), -1);
    for (int i = 0; i < count; ++i) {
        if (ids[i] < 0) continue;
        if (inverse[ids[i]] >= 0) return fail(ctx, B2G_ERR_ARG, "proxy id " + std::to_string(ids[i]) + " repeated");
        inverse[ids[i]] = i;
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipStreamSynchronize(s));  // earlier steps may still read the old keys
    freeArray(ctx->fixOrder);
    freeArray(ctx->orderFix);
    ctx->orderCount = ctx->orderSpan = 0;
    HIPCHK(ctx, hipMalloc((void**)&ctx->fixOrder, sizeof(int) * std::max(count, 1)));
    HIPCHK(ctx, hipMalloc((void**)&ctx->orderFix, sizeof(int) * inverse.size()));
    if (count > 0) HIPCHK(ctx, hipMemcpyAsync(ctx->fixOrder, ids, sizeof(int) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->orderFix, inverse.data(), sizeof(int) * inverse.size(), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->orderCount = count;
    ctx->orderSpan = (int)inverse.size();
    return B2G_OK;
}

int b2g_update_fixtures(b2g_context* ctx, const int32_t* ids, const b2g_fixture_def* defs, int count) {
    if (!ctx || count < 0 || (count > 0 && (!ids || !defs))) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    std::vector<uint8_t> mark(ctx->nFixtures, 0);
    hipStream_t s = ctx->stream;
    for (int i = 0; i < count; ++i) {
        const int f = ids[i];
        if (f < 0 || f >= ctx->nFixtures) return fail(ctx, B2G_ERR_ARG, "fixture index");
        const b2g_fixture_def& d = defs[i];
        const float4 mat = make_float4(d.friction, d.restitution, d.density, d.sensor ? 1.0f : 0.0f);
        const uint32_t filt = (d.categoryBits & 0xFFFF) | ((d.maskBits & 0xFFFF) << 16);
        const int group = d.groupIndex;
        const uint8_t one = 1;
        HIPCHK(ctx, hipMemcpyAsync(ctx->fmat + f, &mat, sizeof(float4), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->ffilter + f, &filt, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->fgroup + f, &group, sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ctx->moved + f, &one, 1, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        mark[f] = 1;
    }
    if (ctx->nContacts > 0) {
        uint8_t* dm = nullptr;
        HIPCHK(ctx, hipMalloc((void**)&dm, mark.size()));
        hipError_t e = hipMemcpyAsync(dm, mark.data(), mark.size(), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_flag_filter, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), dm);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(dm);
        if (e != hipSuccess) return fail(ctx, B2G_ERR_HIP, hipGetErrorString(e));
    }
    ctx->newFixture = true;  // touched proxies are queried at the next step
    return B2G_OK;
}

int b2g_step(b2g_context* ctx, const b2g_step_params* p) {
    if (!ctx || !p) return B2G_ERR_ARG;
    if (p->velocityIterations < 0 || p->positionIterations < 0) return fail(ctx, B2G_ERR_ARG, "iterations");
    if (p->mode != B2G_MODE_COLOURED && p->mode != B2G_MODE_REFERENCE)
        return fail(ctx, B2G_ERR_ARG, "solver mode");
    if (p->learnedWarmStart && p->mode == B2G_MODE_REFERENCE)
        return fail(ctx, B2G_ERR_UNSUPPORTED, "the learned warm start runs with the coloured solver");
    if (ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "a split step is open: finish it with b2g_step_solve");
    (void)hipSetDevice(ctx->device);
    return p->mode == B2G_MODE_REFERENCE ? stepReference(ctx, p) : stepColoured(ctx, p);
}

int b2g_get_contact_solved(b2g_context* ctx, int32_t* out, int cap) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    const int m = std::min(ctx->nContacts, cap);
    if (m == 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->cur.cflags, sizeof(int) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < m; ++i) out[i] = (out[i] & CF_CONSTRAINT) ? 1 : 0;
    return B2G_OK;
}

int b2g_get_contact_materials(b2g_context* ctx, float* out, int cap) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    const int m = std::min(ctx->nContacts, cap);
    if (m == 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->cur.cmat, sizeof(float2) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_step_collide(b2g_context* ctx, const b2g_step_params* p) {
    if (!ctx || !p) return B2G_ERR_ARG;
    if (ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "a split step is already open");
    if (p->velocityIterations < 0 || p->positionIterations < 0) return fail(ctx, B2G_ERR_ARG, "iterations");
    if (p->mode != B2G_MODE_COLOURED && p->mode != B2G_MODE_REFERENCE) return fail(ctx, B2G_ERR_ARG, "solver mode");
    if (p->learnedWarmStart && p->mode == B2G_MODE_REFERENCE)
        return fail(ctx, B2G_ERR_UNSUPPORTED, "the learned warm start runs with the coloured solver");
    (void)hipSetDevice(ctx->device);
    ctx->keepOldManifolds = true;
    const int rc = p->mode == B2G_MODE_REFERENCE ? stepReferenceCollide(ctx, p) : stepColouredCollide(ctx, p);
    ctx->keepOldManifolds = false;
    if (rc != B2G_OK) return rc;
    ctx->openParams = *p;
    ctx->stepOpen = true;
    return B2G_OK;
}

int b2g_get_old_manifolds(b2g_context* ctx, b2g_manifold* out, int cap, int* count) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    if (!ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "no split step is open");
    const int n = ctx->nContacts;
    if (count) *count = n;
    const int m = std::min(n, cap);
    if (m == 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->oldMan, sizeof(b2g_manifold) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_override_contacts(b2g_context* ctx, const int32_t* ids, const int32_t* enabled, const float* friction,
                          const float* restitution, int n) {
    if (!ctx || n < 0 || (n > 0 && (!ids || !enabled || !friction || !restitution))) return B2G_ERR_ARG;
    if (!ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "no split step is open");
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= ctx->nContacts) return fail(ctx, B2G_ERR_ARG, "contact index out of range");
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    std::vector<float2> mat(n);
    for (int i = 0; i < n; ++i) mat[i] = make_float2(friction[i], restitution[i]);
    int* dIds = nullptr;
    int* dEn = nullptr;
    float2* dMat = nullptr;
    hipError_t e = hipMalloc((void**)&dIds, sizeof(int) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dEn, sizeof(int) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dMat, sizeof(float2) * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dIds, ids, sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dEn, enabled, sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dMat, mat.data(), sizeof(float2) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_override_contacts, dim3(blocks(n)), dim3(256), 0, s, ctx->view(), dIds, dEn, dMat, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dIds);
    (void)hipFree(dEn);
    (void)hipFree(dMat);
    HIPCHK(ctx, e);
    return B2G_OK;
}

int b2g_get_contact_tangent_speeds(b2g_context* ctx, float* out, int cap) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    const int m = std::min(ctx->nContacts, cap);
    if (m == 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->cur.ctan, sizeof(float) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_set_contact_tangent_speeds(b2g_context* ctx, const int32_t* ids, const float* speeds, int n) {
    if (!ctx || n < 0 || (n > 0 && (!ids || !speeds))) return B2G_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= ctx->nContacts) return fail(ctx, B2G_ERR_ARG, "contact index out of range");
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    int* dIds = nullptr;
    float* dSp = nullptr;
    hipError_t e = hipMalloc((void**)&dIds, sizeof(int) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dSp, sizeof(float) * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dIds, ids, sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dSp, speeds, sizeof(float) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_set_tangent_speeds, dim3(blocks(n)), dim3(256), 0, s, ctx->view(), dIds, dSp, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dIds);
    (void)hipFree(dSp);
    HIPCHK(ctx, e);
    return B2G_OK;
}

int b2g_step_solve(b2g_context* ctx) {
    if (!ctx) return B2G_ERR_ARG;
    if (!ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "no split step is open");
    ctx->stepOpen = false;
    (void)hipSetDevice(ctx->device);
    return ctx->openParams.mode == B2G_MODE_REFERENCE ? stepReferenceSolve(ctx, &ctx->openParams)
                                                       : stepColouredSolve(ctx, &ctx->openParams);
}

int b2g_sync(b2g_context* ctx) {
    if (!ctx) return B2G_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_set_profiling(b2g_context* ctx, int on) {
    if (!ctx) return B2G_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->profiling = on != 0;
    ctx->evMarks.clear();
    ctx->evUsed = 0;
    for (int i = 0; i < 32; ++i) {
        ctx->profMs[i] = 0.0;
        ctx->profLaunches[i] = 0;
    }
    return B2G_OK;
}

int b2g_get_profile(b2g_context* ctx, b2g_phase_profile* out, int cap, int* count) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    if (count) *count = PH_COUNT;
    for (int i = 0; i < PH_COUNT && i < cap; ++i) {
        memset(&out[i], 0, sizeof(out[i]));
        strncpy(out[i].name, kPhaseNames[i], sizeof(out[i].name) - 1);
        out[i].launches = ctx->profLaunches[i];
        out[i].ms = ctx->profMs[i];
    }
    return B2G_OK;
}

int b2g_get_stats(b2g_context* ctx, b2g_step_stats* out) {
    if (!ctx || !out) return B2G_ERR_ARG;
    *out = ctx->stats;
    return B2G_OK;
}

int b2g_get_contacts(b2g_context* ctx, int32_t* meta, b2g_manifold* manifolds, int cap, int* count) {
    if (!ctx || cap < 0) return B2G_ERR_ARG;
    const int n = ctx->nContacts;
    if (count) *count = n;
    const int m = std::min(n, cap);
    if (m == 0) return B2G_OK;
    std::vector<int> fa(m), fb(m), fl(m), col(m);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(fa.data(), ctx->cur.cfA, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(fb.data(), ctx->cur.cfB, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(fl.data(), ctx->cur.cflags, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(col.data(), ctx->cur.ccolour, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    if (manifolds)
        HIPCHK(ctx, hipMemcpyAsync(manifolds, ctx->cur.cman, sizeof(b2g_manifold) * m, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (meta) {
        for (int i = 0; i < m; ++i) {
            meta[4 * i + 0] = fa[i];
            meta[4 * i + 1] = fb[i];
            meta[4 * i + 2] = ((fl[i] & CF_TOUCHING) ? 1 : 0) | ((fl[i] & CF_ENABLED) ? 2 : 0);
            if (ctx->stepOpen)
                meta[4 * i + 2] |= ((fl[i] & CF_DEAD) ? 4 : 0) | ((fl[i] & CF_UPDATED) ? 8 : 0) |
                                   ((fl[i] & CF_WAS_TOUCHING) ? 16 : 0);
            meta[4 * i + 3] = col[i];
        }
    }
    return B2G_OK;
}

int b2g_get_fat_aabbs(b2g_context* ctx, b2g_aabb* out, int cap) {
    if (!ctx || !out) return B2G_ERR_ARG;
    const int m = std::min(cap, ctx->nFixtures);
    if (m <= 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->fat, sizeof(float4) * m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_penetration(b2g_context* ctx, double* out) {
    if (!ctx || !out) return B2G_ERR_ARG;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemsetAsync(ctx->penAcc, 0, 64, s));
    double* acc = ctx->penAcc;
    int* np = (int*)(ctx->penAcc + 1);
    unsigned long long* mx = (unsigned long long*)(ctx->penAcc + 2);
    if (ctx->nContacts > 0)
        hipLaunchKernelGGL(k_penetration, dim3(blocks(ctx->nContacts)), dim3(256), 0, s, ctx->view(), acc, np, mx);
    LAUNCHCHK(ctx);
    double h[3];
    HIPCHK(ctx, hipMemcpyAsync(h, ctx->penAcc, 24, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    int n;
    memcpy(&n, &h[1], 4);
    out[0] = n ? h[0] / n : 0.0;
    memcpy(&out[1], &h[2], 8);
    out[2] = (double)n;
    return B2G_OK;
}

int b2g_joint_drift(b2g_context* ctx, double* out) {
    if (!ctx || !out) return B2G_ERR_ARG;
    out[0] = out[1] = out[2] = 0.0;
    const int nj = ctx->nJoints;
    if (nj == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    float* d = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&d, sizeof(float) * nj));
    hipLaunchKernelGGL(k_joint_drift, dim3(blocks(nj)), dim3(256), 0, s, ctx->view(), ctx->jview(), d);
    std::vector<float> h(nj);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpyAsync(h.data(), d, sizeof(float) * nj, hipMemcpyDeviceToHost, s);
    const hipError_t e3 = hipStreamSynchronize(s);
    (void)hipFree(d);
    HIPCHK(ctx, e1);
    HIPCHK(ctx, e2);
    HIPCHK(ctx, e3);
    double sum = 0.0, mx = 0.0;
    int n = 0;
    for (float x : h) {
        if (x < 0.0f) continue;
        sum += x;
        mx = std::max(mx, (double)x);
        ++n;
    }
    out[0] = n ? sum / n : 0.0;
    out[1] = mx;
    out[2] = (double)n;
    return B2G_OK;
}

int b2g_energy(b2g_context* ctx, double* out) {
    if (!ctx || !out) return B2G_ERR_ARG;
    out[0] = out[1] = 0.0;
    const int nb = ctx->nBodies;
    if (nb == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    double2* d = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&d, sizeof(double2) * nb));
    hipLaunchKernelGGL(k_body_energy, dim3(blocks(nb)), dim3(256), 0, s, ctx->view(), ctx->gravity, d);
    std::vector<double2> h(nb);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpyAsync(h.data(), d, sizeof(double2) * nb, hipMemcpyDeviceToHost, s);
    const hipError_t e3 = hipStreamSynchronize(s);
    (void)hipFree(d);
    HIPCHK(ctx, e1);
    HIPCHK(ctx, e2);
    HIPCHK(ctx, e3);
    for (const double2& x : h) {
        out[0] += x.x;
        out[1] += x.y;
    }
    return B2G_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batches
namespace {
template <typename T>
struct DevBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" {

int b2g_collide_batch(b2g_context* ctx, const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                      const b2g_transform* xfB, int n, b2g_manifold* out) {
    if (!ctx || n < 0 || (n > 0 && (!a || !b || !xfA || !xfB || !out))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<b2g_shape> dA, dB;
    DevBuf<b2g_transform> dxa, dxb;
    DevBuf<b2g_manifold> dm;
    HIPCHK(ctx, dA.alloc(n)); HIPCHK(ctx, dB.alloc(n)); HIPCHK(ctx, dxa.alloc(n)); HIPCHK(ctx, dxb.alloc(n));
    HIPCHK(ctx, dm.alloc(n));
    HIPCHK(ctx, hipMemcpyAsync(dA.p, a, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dB.p, b, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxa.p, xfA, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxb.p, xfB, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_collide_batch, dim3(blocks(n, 64)), dim3(64), 0, s, dA.p, dxa.p, dB.p, dxb.p, n, dm.p);
    hipLaunchKernelGGL(k_collide_batch_tiles, dim3((n + kTileBlock / kTileLanes - 1) / (kTileBlock / kTileLanes)),
                       dim3(kTileBlock), 0, s, dA.p, dxa.p, dB.p, dxb.p, n, dm.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(out, dm.p, sizeof(b2g_manifold) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_distance_batch(b2g_context* ctx, const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                       const b2g_transform* xfB, int n, float* out) {
    if (!ctx || n < 0 || (n > 0 && (!a || !b || !xfA || !xfB || !out))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<b2g_shape> dA, dB;
    DevBuf<b2g_transform> dxa, dxb;
    DevBuf<float> dout;
    HIPCHK(ctx, dA.alloc(n)); HIPCHK(ctx, dB.alloc(n)); HIPCHK(ctx, dxa.alloc(n)); HIPCHK(ctx, dxb.alloc(n));
    HIPCHK(ctx, dout.alloc(2 * (size_t)n));
    HIPCHK(ctx, hipMemcpyAsync(dA.p, a, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dB.p, b, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxa.p, xfA, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxb.p, xfB, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_distance_batch, dim3(blocks(n, 64)), dim3(64), 0, s, dA.p, dxa.p, dB.p, dxb.p, n, dout.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(out, dout.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_toi_batch(b2g_context* ctx, const b2g_shape* a, const b2g_sweep* sweepA, const b2g_shape* b,
                  const b2g_sweep* sweepB, int n, float tMax, int32_t* state, float* t) {
    if (!ctx || n < 0 || (n > 0 && (!a || !b || !sweepA || !sweepB || !state || !t))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<b2g_shape> dA, dB;
    DevBuf<b2g_sweep> dsa, dsb;
    DevBuf<int32_t> dst;
    DevBuf<float> dt;
    HIPCHK(ctx, dA.alloc(n)); HIPCHK(ctx, dB.alloc(n)); HIPCHK(ctx, dsa.alloc(n)); HIPCHK(ctx, dsb.alloc(n));
    HIPCHK(ctx, dst.alloc(n)); HIPCHK(ctx, dt.alloc(n));
    HIPCHK(ctx, hipMemcpyAsync(dA.p, a, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dB.p, b, sizeof(b2g_shape) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dsa.p, sweepA, sizeof(b2g_sweep) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dsb.p, sweepB, sizeof(b2g_sweep) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_toi_batch, dim3(blocks(n, 64)), dim3(64), 0, s, dA.p, dsa.p, dB.p, dsb.p, n, tMax, dst.p, dt.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(state, dst.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(t, dt.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_world_manifold_batch(b2g_context* ctx, const b2g_manifold* m, const b2g_transform* xfA, const float* radiusA,
                             const b2g_transform* xfB, const float* radiusB, int n, float* out) {
    if (!ctx || n < 0 || (n > 0 && (!m || !xfA || !xfB || !radiusA || !radiusB || !out))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<b2g_manifold> dm;
    DevBuf<b2g_transform> dxa, dxb;
    DevBuf<float> dra, drb, dout;
    HIPCHK(ctx, dm.alloc(n)); HIPCHK(ctx, dxa.alloc(n)); HIPCHK(ctx, dxb.alloc(n)); HIPCHK(ctx, dra.alloc(n));
    HIPCHK(ctx, drb.alloc(n)); HIPCHK(ctx, dout.alloc(8 * (size_t)n));
    HIPCHK(ctx, hipMemcpyAsync(dm.p, m, sizeof(b2g_manifold) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxa.p, xfA, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dxb.p, xfB, sizeof(b2g_transform) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dra.p, radiusA, sizeof(float) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(drb.p, radiusB, sizeof(float) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_world_manifold_batch, dim3(blocks(n)), dim3(256), 0, s, dm.p, dxa.p, dra.p, dxb.p, drb.p, n,
                       dout.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(out, dout.p, sizeof(float) * 8 * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_pairs_batch(b2g_context* ctx, const b2g_aabb* fat, const uint8_t* moved, int n, int32_t* pairs, int cap,
                    int* count) {
    if (!ctx || n < 0 || cap < 0 || (n > 0 && (!fat || !moved))) return B2G_ERR_ARG;
    if (count) *count = 0;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    // Cell size from the median proxy extent.
    std::vector<float> ext(n);
    for (int i = 0; i < n; ++i) ext[i] = std::max(fat[i].ux - fat[i].lx, fat[i].uy - fat[i].ly);
    std::nth_element(ext.begin(), ext.begin() + n / 2, ext.end());
    const float cell = std::max(ext[n / 2], 1e-3f);
    int nb = 1024;
    while (nb < 2 * n) nb <<= 1;
    const int capPairs = std::max(cap, 1 << 16);
    DevBuf<float4> dfat;
    DevBuf<uint8_t> dmoved;
    DevBuf<int> counts, starts, cursor, entries;
    DevBuf<unsigned long long> big;
    DevBuf<unsigned long long> ecell;
    DevBuf<float4> efat;
    DevBuf<int> queries;
    DevBuf<int4> range;
    DevBuf<unsigned long long> keys, sorted;
    DevBuf<char> scratch;
    DevBuf<Counters> dc;
    const size_t sb = prims_scratch_bytes(std::max(std::max(nb + 1, capPairs), n));
    HIPCHK(ctx, dfat.alloc(n)); HIPCHK(ctx, dmoved.alloc(n)); HIPCHK(ctx, counts.alloc(nb + 1));
    HIPCHK(ctx, starts.alloc(nb + 1)); HIPCHK(ctx, cursor.alloc(nb + 1));
    HIPCHK(ctx, entries.alloc((size_t)kMaxCellsPerProxy * n)); HIPCHK(ctx, big.alloc(n)); HIPCHK(ctx, range.alloc(n));
    HIPCHK(ctx, ecell.alloc((size_t)kMaxCellsPerProxy * n));
    HIPCHK(ctx, efat.alloc((size_t)kMaxCellsPerProxy * n));
    HIPCHK(ctx, queries.alloc(n));
    HIPCHK(ctx, keys.alloc(capPairs)); HIPCHK(ctx, sorted.alloc(capPairs)); HIPCHK(ctx, scratch.alloc(sb));
    HIPCHK(ctx, dc.alloc(1));
    HIPCHK(ctx, hipMemcpyAsync(dfat.p, fat, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dmoved.p, moved, n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemsetAsync(dc.p, 0, sizeof(Counters), s));
    HIPCHK(ctx, hipMemsetAsync(counts.p, 0, sizeof(int) * (nb + 1), s));
    HIPCHK(ctx, hipMemsetAsync(cursor.p, 0, sizeof(int) * (nb + 1), s));
    GridView g;
    g.invCell = 1.0f / cell;
    g.bucketMask = (uint32_t)(nb - 1);
    g.counts = counts.p; g.starts = starts.p; g.cursor = cursor.p; g.entries = entries.p; g.big = big.p;
    g.bigSorted = 0;
    g.ecell = ecell.p;
    g.efat = efat.p;
    g.queries = queries.p;
    g.range = range.p;
    g.fbody = nullptr;
    g.bworld = nullptr;
    hipLaunchKernelGGL(k_grid_count, dim3(blocks(n)), dim3(256), 0, s, dfat.p, dmoved.p, n, (const uint8_t*)nullptr, g,
                       dc.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, prims_exclusive_sum(scratch.p, sb, counts.p, starts.p, nb + 1, s));
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks(n)), dim3(256), 0, s, n, g, dfat.p, dmoved.p);
    WorldView w;
    memset(&w, 0, sizeof(w));
    PairSink sink{keys.p, capPairs, 0ull, nullptr};
    hipLaunchKernelGGL(k_pairs_query, dim3(blocks(n * kQueryLanes)), dim3(256), 0, s, dfat.p, dmoved.p, n, g, w, false,
                       (const unsigned long long*)nullptr, 0u, sink, dc.p);
    hipLaunchKernelGGL(k_pairs_wide, dim3(blocks(n)), dim3(256), 0, s, dfat.p, dmoved.p, n, g, w, false,
                       (const unsigned long long*)nullptr, 0u, sink, dc.p);
    LAUNCHCHK(ctx);
    Counters hc;
    HIPCHK(ctx, hipMemcpyAsync(&hc, dc.p, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    const int np = hc.nNewPairs;
    if (count) *count = np;
    if (hc.overflow || np > capPairs) return fail(ctx, B2G_ERR_CAPACITY, "pair buffer");
    if (np > 0) {
        HIPCHK(ctx, prims_sort_keys_u64(scratch.p, sb, keys.p, sorted.p, np, s));
        std::vector<unsigned long long> hk(np);
        HIPCHK(ctx, hipMemcpyAsync(hk.data(), sorted.p, sizeof(unsigned long long) * np, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        for (int i = 0; i < np && i < cap; ++i) {
            pairs[2 * i] = (int32_t)(hk[i] >> 32);
            pairs[2 * i + 1] = (int32_t)(hk[i] & 0xFFFFFFFFu);
        }
    }
    return B2G_OK;
}

int b2g_fat_rule_batch(b2g_context* ctx, const b2g_aabb* fat, const b2g_aabb* tight, const float* disp, int n,
                       b2g_aabb* fatOut, uint8_t* moved) {
    if (!ctx || n < 0 || (n > 0 && (!fat || !tight || !disp || !fatOut || !moved))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<float4> df, dt, dout;
    DevBuf<float2> dd;
    DevBuf<uint8_t> dm;
    HIPCHK(ctx, df.alloc(n)); HIPCHK(ctx, dt.alloc(n)); HIPCHK(ctx, dout.alloc(n)); HIPCHK(ctx, dd.alloc(n));
    HIPCHK(ctx, dm.alloc(n));
    HIPCHK(ctx, hipMemcpyAsync(df.p, fat, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dt.p, tight, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dd.p, disp, sizeof(float2) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fat_rule_batch, dim3(blocks(n)), dim3(256), 0, s, df.p, dt.p, dd.p, n, dout.p, dm.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(fatOut, dout.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(moved, dm.p, n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_sincos_batch(b2g_context* ctx, const float* angles, int n, float* so, float* co) {
    if (!ctx || n < 0 || (n > 0 && (!angles || !so || !co))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<float> da, ds, dc;
    HIPCHK(ctx, da.alloc(n)); HIPCHK(ctx, ds.alloc(n)); HIPCHK(ctx, dc.alloc(n));
    HIPCHK(ctx, hipMemcpyAsync(da.p, angles, sizeof(float) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sincos_batch, dim3(blocks(n)), dim3(256), 0, s, da.p, n, ds.p, dc.p);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, hipMemcpyAsync(so, ds.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(co, dc.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

}  // extern "C"

// ------------------------------------------------------------ learned warm start (C ABI)
namespace {

// Per-block partial sums over the log: pass 0 sums of the 16 features and 2
// targets, pass 1 sums of squared deviations from `mean`.
__global__ void k_log_moments(const float* feat, const float* tgt, int n, const float* mean, float* partial) {
    constexpr int kQ = kFeatures + 2;
    float acc[kQ];
    for (int q = 0; q < kQ; ++q) acc[q] = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int q = 0; q < kQ; ++q) {
            float v = q < kFeatures ? feat[(size_t)i * kFeatures + q] : tgt[2 * (size_t)i + (q - kFeatures)];
            if (mean) {
                v -= mean[q];
                v *= v;
            }
            acc[q] += v;
        }
    }
    __shared__ float red[256];
    for (int q = 0; q < kQ; ++q) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x * kQ + q] = red[0];
        __syncthreads();
    }
}

// Epoch order of the log: sort keys from a hash of (seed, epoch, index).
__global__ void k_perm_keys(int n, uint32_t seed, uint32_t epoch, uint32_t* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ seed * 0x85EBCA77u ^ epoch * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    keys[i] = h;
    vals[i] = i;
}

void paramsToDevice(const b2g_mlp_params& p, std::vector<float>& dev, float* norm) {
    using namespace b2mlp;
    dev.assign(kParams, 0.0f);
    memcpy(&dev[kW1], p.W1, sizeof(p.W1));
    memcpy(&dev[kB1], p.b1, sizeof(p.b1));
    memcpy(&dev[kW2], p.W2, sizeof(p.W2));
    memcpy(&dev[kB2], p.b2, sizeof(p.b2));
    memcpy(&dev[kW3], p.W3, sizeof(p.W3));
    memcpy(&dev[kB3], p.b3, sizeof(p.b3));
    memcpy(&dev[kW4], p.W4, sizeof(p.W4));
    memcpy(&dev[kB4], p.b4, sizeof(p.b4));
    if (norm) {
        for (int f = 0; f < kIn; ++f) {
            norm[kNormXMean + f] = p.xmean[f];
            norm[kNormXInv + f] = p.xstd[f] > 0.0f ? 1.0f / p.xstd[f] : 0.0f;
        }
        for (int o = 0; o < kOut; ++o) {
            norm[kNormYMean + o] = p.ymean[o];
            norm[kNormYStd + o] = p.ystd[o];
        }
    }
}

void paramsFromDevice(const std::vector<float>& dev, const float* norm, b2g_mlp_params& p) {
    using namespace b2mlp;
    memset(&p, 0, sizeof(p));
    memcpy(p.W1, &dev[kW1], sizeof(p.W1));
    memcpy(p.b1, &dev[kB1], sizeof(p.b1));
    memcpy(p.W2, &dev[kW2], sizeof(p.W2));
    memcpy(p.b2, &dev[kB2], sizeof(p.b2));
    memcpy(p.W3, &dev[kW3], sizeof(p.W3));
    memcpy(p.b3, &dev[kB3], sizeof(p.b3));
    memcpy(p.W4, &dev[kW4], sizeof(p.W4));
    memcpy(p.b4, &dev[kB4], sizeof(p.b4));
    if (norm) {
        for (int f = 0; f < kIn; ++f) {
            p.xmean[f] = norm[kNormXMean + f];
            p.xstd[f] = norm[kNormXInv + f] > 0.0f ? 1.0f / norm[kNormXInv + f] : 0.0f;
        }
        for (int o = 0; o < kOut; ++o) {
            p.ymean[o] = norm[kNormYMean + o];
            p.ystd[o] = norm[kNormYStd + o];
        }
    }
}

int uploadNet(b2g_context* ctx, const std::vector<float>& dev, const float* norm) {
    using namespace b2mlp;
    int rc = ensureMlp(ctx);
    if (rc != B2G_OK) return rc;
    hipStream_t s = ctx->stream;
    Net& n = ctx->mlp;
    HIPCHK(ctx, hipMemcpyAsync(n.params, dev.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(n.norm, norm, sizeof(float) * kNorm, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemsetAsync(n.grad, 0, sizeof(float) * kParams, s));
    HIPCHK(ctx, hipMemsetAsync(n.m, 0, sizeof(float) * kParams, s));
    HIPCHK(ctx, hipMemsetAsync(n.v, 0, sizeof(float) * kParams, s));
    HIPCHK(ctx, mlpPrepare(n, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    n.adamStep = 0;
    n.ready = true;
    return B2G_OK;
}

}  // namespace

extern "C" {

int b2g_log_fresh(b2g_context* ctx, int capacity) {
    if (!ctx || capacity < 0) return B2G_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    freeArray(ctx->logFeat);
    freeArray(ctx->logTarget);
    ctx->logCap = 0;
    ctx->logCount = 0;
    if (capacity == 0) return B2G_OK;
    HIPCHK(ctx, hipMalloc((void**)&ctx->logFeat, sizeof(float) * kFeatures * (size_t)capacity));
    HIPCHK(ctx, hipMalloc((void**)&ctx->logTarget, sizeof(float) * 2 * (size_t)capacity));
    ctx->logCap = capacity;
    return B2G_OK;
}

int b2g_set_learned_scope(b2g_context* ctx, int scope, int logStride) {
    if (!ctx || scope < 0 || scope > 1 || logStride < 1) return B2G_ERR_ARG;
    ctx->learnedScope = scope;
    ctx->logStride = logStride;
    ctx->logSteps = 0;
    return B2G_OK;
}

int b2g_log_size(b2g_context* ctx, int* count) {
    if (!ctx || !count) return B2G_ERR_ARG;
    *count = ctx->logCount;
    return B2G_OK;
}

int b2g_get_log(b2g_context* ctx, int first, int count, float* features, float* targets) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->logCount) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    if (features)
        HIPCHK(ctx, hipMemcpyAsync(features, ctx->logFeat + (size_t)first * kFeatures,
                                   sizeof(float) * kFeatures * (size_t)count, hipMemcpyDeviceToHost, s));
    if (targets)
        HIPCHK(ctx, hipMemcpyAsync(targets, ctx->logTarget + 2 * (size_t)first, sizeof(float) * 2 * (size_t)count,
                                   hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_mlp_init(b2g_context* ctx, uint32_t seed) {
    using namespace b2mlp;
    if (!ctx) return B2G_ERR_ARG;
    b2g_mlp_params p;
    memset(&p, 0, sizeof(p));
    std::mt19937 rng(seed);
    auto fill = [&](float* w, int fanOut, int fanIn) {
        const float a = sqrtf(6.0f / (float)(fanIn + fanOut));
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < fanOut * fanIn; ++i) w[i] = u(rng);
    };
    fill(p.W1, kHidden, kIn);
    fill(p.W2, kHidden, kHidden);
    fill(p.W3, kHidden, kHidden);
    fill(p.W4, kOut, kHidden);
    for (int f = 0; f < kIn; ++f) p.xstd[f] = 1.0f;
    for (int o = 0; o < kOut; ++o) p.ystd[o] = 1.0f;
    const int n = ctx->logCount;
    if (n > 0) {
        // Normalisation from the log: means, then standard deviations (two
        // passes; FP32 partial sums per block, summed in FP64 here).
        constexpr int kQ = kFeatures + 2;
        const int grid = std::min(256, std::max(1, blocks(n)));
        hipStream_t s = ctx->stream;
        DevBuf<float> partial, mean;
        HIPCHK(ctx, partial.alloc((size_t)grid * kQ));
        HIPCHK(ctx, mean.alloc(kQ));
        std::vector<float> hp((size_t)grid * kQ);
        double mu[kQ], var[kQ];
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(k_log_moments, dim3(grid), dim3(256), 0, s, ctx->logFeat, ctx->logTarget, n,
                               pass ? (const float*)mean.p : (const float*)nullptr, partial.p);
            LAUNCHCHK(ctx);
            HIPCHK(ctx, hipMemcpyAsync(hp.data(), partial.p, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, s));
            HIPCHK(ctx, hipStreamSynchronize(s));
            for (int q = 0; q < kQ; ++q) {
                double acc = 0.0;
                for (int b = 0; b < grid; ++b) acc += hp[(size_t)b * kQ + q];
                if (pass == 0) mu[q] = acc / n; else var[q] = acc / n;
            }
            if (pass == 0) {
                float mf[kQ];
                for (int q = 0; q < kQ; ++q) mf[q] = (float)mu[q];
                HIPCHK(ctx, hipMemcpyAsync(mean.p, mf, sizeof(mf), hipMemcpyHostToDevice, s));
            }
        }
        for (int f = 0; f < kIn; ++f) {
            p.xmean[f] = (float)mu[f];
            const double sd = sqrt(var[f]);
            p.xstd[f] = sd > 1e-12 * std::max(1.0, fabs(mu[f])) ? (float)sd : 0.0f;
        }
        for (int o = 0; o < kOut; ++o) {
            p.ymean[o] = (float)mu[kIn + o];
            const double sd = sqrt(var[kIn + o]);
            p.ystd[o] = sd > 1e-12 ? (float)sd : 1.0f;
        }
    }
    return b2g_mlp_set_params(ctx, &p);
}

int b2g_mlp_set_params(b2g_context* ctx, const b2g_mlp_params* p) {
    if (!ctx || !p) return B2G_ERR_ARG;
    std::vector<float> dev;
    float norm[b2mlp::kNorm];
    paramsToDevice(*p, dev, norm);
    return uploadNet(ctx, dev, norm);
}

int b2g_mlp_get_params(b2g_context* ctx, b2g_mlp_params* p) {
    using namespace b2mlp;
    if (!ctx || !p) return B2G_ERR_ARG;
    if (!ctx->mlp.ready) return fail(ctx, B2G_ERR_STATE, "no warm-start network");
    std::vector<float> dev(kParams);
    float norm[kNorm];
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(dev.data(), ctx->mlp.params, sizeof(float) * kParams, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(norm, ctx->mlp.norm, sizeof(norm), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    paramsFromDevice(dev, norm, *p);
    return B2G_OK;
}

int b2g_mlp_train(b2g_context* ctx, int epochs, int batch, float lr, uint32_t seed, double* loss) {
    using namespace b2mlp;
    if (!ctx || epochs < 0 || batch <= 0 || !(lr > 0.0f)) return B2G_ERR_ARG;
    if (!ctx->mlp.ready) return fail(ctx, B2G_ERR_STATE, "no warm-start network (b2g_mlp_init first)");
    const int n = ctx->logCount;
    if (n == 0) return fail(ctx, B2G_ERR_STATE, "the training log is empty");
    hipStream_t s = ctx->stream;
    // order buffers: keys, sorted keys, values, order, then sort scratch
    const size_t sortBytes = prims_scratch_bytes(n);
    const size_t need = 4 * sizeof(int) * (size_t)n + sortBytes + 1024;
    if (need > ctx->trainScratchBytes) {
        if (ctx->trainScratch) (void)hipFree(ctx->trainScratch);
        ctx->trainScratch = nullptr;
        ctx->trainScratchBytes = 0;
        HIPCHK(ctx, hipMalloc(&ctx->trainScratch, need));
        ctx->trainScratchBytes = need;
    }
    char* base = (char*)ctx->trainScratch;
    uint32_t* keys = (uint32_t*)base;
    uint32_t* keysSorted = keys + n;
    int* vals = (int*)(keysSorted + n);
    int* order = vals + n;
    void* sortScratch = (void*)(((uintptr_t)(order + n) + 255) & ~(uintptr_t)255);
    Net& net = ctx->mlp;
    const int batches = (n + batch - 1) / batch;
    for (int e = 0; e < epochs; ++e) {
        hipLaunchKernelGGL(k_perm_keys, dim3(blocks(n)), dim3(256), 0, s, n, seed, (uint32_t)e, keys, vals);
        LAUNCHCHK(ctx);
        HIPCHK(ctx, prims_sort_pairs_u32(sortScratch, sortBytes, keys, keysSorted, vals, order, n, 32, s));
        HIPCHK(ctx, hipMemsetAsync(net.loss, 0, sizeof(float), s));
        for (int b = 0; b < batches; ++b) {
            const int first = b * batch;
            const int count = std::min(batch, n - first);
            hipEvent_t e0, e1;
            const int pk = profKernelEvents(ctx, &e0, &e1);
            HIPCHK(ctx, mlpGradient(net, ctx->logFeat, ctx->logTarget, order, first, count, s, e0, e1));
            profKernel(ctx, PH_K_MLP_GRAD, pk);
            HIPCHK(ctx, mlpAdam(net, lr, s));
        }
        float l = 0.0f;
        HIPCHK(ctx, hipMemcpyAsync(&l, net.loss, sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        profCollect(ctx);
        if (loss) loss[e] = (double)l / batches;
    }
    return B2G_OK;
}

int b2g_mlp_forward(b2g_context* ctx, const float* x, int n, float* y) {
    if (!ctx || n < 0 || (n > 0 && (!x || !y))) return B2G_ERR_ARG;
    if (!ctx->mlp.ready) return fail(ctx, B2G_ERR_STATE, "no warm-start network");
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<float> dx, dy;
    HIPCHK(ctx, dx.alloc((size_t)n * kFeatures));
    HIPCHK(ctx, dy.alloc((size_t)n * 2));
    HIPCHK(ctx, hipMemcpyAsync(dx.p, x, sizeof(float) * kFeatures * (size_t)n, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    const int pk = profKernelEvents(ctx, &e0, &e1);
    HIPCHK(ctx, b2mlp::mlpForward(ctx->mlp, dx.p, n, dy.p, s, e0, e1));
    profKernel(ctx, PH_K_MLP_INFER, pk);
    HIPCHK(ctx, hipMemcpyAsync(y, dy.p, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    profCollect(ctx);
    return B2G_OK;
}

int b2g_mlp_gradient(b2g_context* ctx, const float* x, const float* t, int n, b2g_mlp_params* grad, double* loss) {
    using namespace b2mlp;
    if (!ctx || n <= 0 || !x || !t || !grad) return B2G_ERR_ARG;
    if (!ctx->mlp.ready) return fail(ctx, B2G_ERR_STATE, "no warm-start network");
    hipStream_t s = ctx->stream;
    DevBuf<float> dx, dt;
    HIPCHK(ctx, dx.alloc((size_t)n * kFeatures));
    HIPCHK(ctx, dt.alloc((size_t)n * 2));
    HIPCHK(ctx, hipMemcpyAsync(dx.p, x, sizeof(float) * kFeatures * (size_t)n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dt.p, t, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemsetAsync(ctx->mlp.grad, 0, sizeof(float) * kParams, s));
    HIPCHK(ctx, hipMemsetAsync(ctx->mlp.loss, 0, sizeof(float), s));
    HIPCHK(ctx, mlpGradient(ctx->mlp, dx.p, dt.p, nullptr, 0, n, s));
    std::vector<float> g(kParams);
    float l = 0.0f;
    HIPCHK(ctx, hipMemcpyAsync(g.data(), ctx->mlp.grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&l, ctx->mlp.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemsetAsync(ctx->mlp.grad, 0, sizeof(float) * kParams, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    paramsFromDevice(g, nullptr, *grad);
    if (loss) *loss = l;
    return B2G_OK;
}

int b2g_set_trace(b2g_context* ctx, int maxIters) {
    if (!ctx || maxIters < 0 || maxIters > 4096) return B2G_ERR_ARG;
    ctx->traceIters = maxIters;  // the last recorded trace stays readable
    return B2G_OK;
}

int b2g_get_trace(b2g_context* ctx, double* curves, int cap, int* iters) {
    if (!ctx || cap < 0 || (cap > 0 && !curves)) return B2G_ERR_ARG;
    if (iters) *iters = ctx->trace.empty() ? 0 : ctx->traceStepIters;
    const int m = std::min(cap, (int)ctx->trace.size());
    if (m > 0) memcpy(curves, ctx->trace.data(), sizeof(double) * m);
    return B2G_OK;
}

}  // extern "C"

extern "C" {

int b2g_set_colour_demotion(b2g_context* ctx, int rounds) {
    if (!ctx || rounds < 0 || rounds > 64) return B2G_ERR_ARG;
    ctx->demoteRounds = rounds;
    return B2G_OK;
}

int b2g_get_colour_counts(b2g_context* ctx, int32_t* out, int cap, int* count) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return B2G_ERR_ARG;
    for (int c = 0; c < kColourSlots && c < cap; ++c) out[c] = ctx->lastColourStart[c + 1] - ctx->lastColourStart[c];
    if (count) *count = kColourSlots;
    return B2G_OK;
}

int b2g_prim_exclusive_sum(b2g_context* ctx, const int32_t* in, int n, int32_t* out) {
    if (!ctx || n < 0 || (n > 0 && (!in || !out))) return B2G_ERR_ARG;
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<int> din, dout;
    DevBuf<char> scratch;
    const size_t sb = prims_scratch_bytes(n);
    HIPCHK(ctx, din.alloc(n)); HIPCHK(ctx, dout.alloc(n)); HIPCHK(ctx, scratch.alloc(sb));
    HIPCHK(ctx, hipMemcpyAsync(din.p, in, sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, prims_exclusive_sum(scratch.p, sb, din.p, dout.p, n, s));
    HIPCHK(ctx, hipMemcpyAsync(out, dout.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_prim_sort_pairs(b2g_context* ctx, const uint32_t* keys, const int32_t* vals, int n, int bits, int method,
                        uint32_t* outKeys, int32_t* outVals, int32_t* totals, int totalsKeys) {
    if (!ctx || n < 0 || (n > 0 && (!keys || !vals || !outKeys || !outVals))) return B2G_ERR_ARG;
    if (method != 0 && method != 1) return fail(ctx, B2G_ERR_ARG, "sort method: 0 (radix) or 1 (7-bit)");
    if (method == 0 && (bits < 1 || bits > 32)) return fail(ctx, B2G_ERR_ARG, "sort bits: 1-32");
    if (method == 1 && (totalsKeys < 0 || totalsKeys > 128 || (totalsKeys > 0 && !totals)))
        return fail(ctx, B2G_ERR_ARG, "totals: 0-128 keys");
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> dk, dko;
    DevBuf<int> dv, dvo, dt;
    DevBuf<char> scratch;
    const size_t sb = prims_scratch_bytes(n);
    HIPCHK(ctx, dk.alloc(n)); HIPCHK(ctx, dko.alloc(n)); HIPCHK(ctx, dv.alloc(n)); HIPCHK(ctx, dvo.alloc(n));
    HIPCHK(ctx, dt.alloc(128)); HIPCHK(ctx, scratch.alloc(sb));
    HIPCHK(ctx, hipMemcpyAsync(dk.p, keys, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dv.p, vals, sizeof(int) * n, hipMemcpyHostToDevice, s));
    if (method == 0) {
        HIPCHK(ctx, prims_sort_pairs_u32(scratch.p, sb, dk.p, dko.p, dv.p, dvo.p, n, bits, s));
    } else {
        if (totalsKeys > 0)
            HIPCHK(ctx, hipMemcpyAsync(dt.p, totals, sizeof(int) * totalsKeys, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, prims_count_sort_pairs(scratch.p, sb, dk.p, dko.p, dv.p, dvo.p, n, totalsKeys > 0 ? dt.p : nullptr,
                                           totalsKeys, s));
        if (totalsKeys > 0)
            HIPCHK(ctx, hipMemcpyAsync(totals, dt.p, sizeof(int) * totalsKeys, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(ctx, hipMemcpyAsync(outKeys, dko.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(outVals, dvo.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_prim_sort_keys_u64(b2g_context* ctx, const uint64_t* keys, int n, int endBit, uint64_t* out) {
    if (!ctx || n < 0 || (n > 0 && (!keys || !out))) return B2G_ERR_ARG;
    if (endBit < 1 || endBit > 64) return fail(ctx, B2G_ERR_ARG, "sort endBit: 1-64");
    if (n == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    DevBuf<unsigned long long> dk, dko;
    DevBuf<char> scratch;
    const size_t sb = prims_scratch_bytes(n);
    HIPCHK(ctx, dk.alloc(n)); HIPCHK(ctx, dko.alloc(n)); HIPCHK(ctx, scratch.alloc(sb));
    HIPCHK(ctx, hipMemcpyAsync(dk.p, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, prims_sort_keys_u64(scratch.p, sb, dk.p, dko.p, n, s, endBit));
    HIPCHK(ctx, hipMemcpyAsync(out, dko.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_check_colouring(b2g_context* ctx, int* conflicts, int* maskErrors) {
    if (!ctx || !conflicts || !maskErrors) return B2G_ERR_ARG;
    *conflicts = 0;
    *maskErrors = 0;
    const int n = ctx->lastConstraints;
    hipStream_t s = ctx->stream;
    std::vector<int2> bodies(n);
    std::vector<int> contact(n), btype(ctx->nBodies), colour(ctx->nContacts);
    std::vector<unsigned long long> mask(ctx->nBodies);
    // the constraints index the table as it was during the solve; the
    // compaction after it moved contact c to idx[c]
    std::vector<int> moved(std::max(ctx->compactedFrom, 0));
    if (n > 0) {
        HIPCHK(ctx, hipMemcpyAsync(bodies.data(), ctx->cv.bodies, sizeof(int2) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(contact.data(), ctx->cv.contact, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        if (!moved.empty())
            HIPCHK(ctx, hipMemcpyAsync(moved.data(), ctx->idx, sizeof(int) * moved.size(), hipMemcpyDeviceToHost, s));
    }
    if (ctx->nBodies > 0) {
        HIPCHK(ctx, hipMemcpyAsync(btype.data(), ctx->btype, sizeof(int) * ctx->nBodies, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(mask.data(), ctx->colourMask, sizeof(unsigned long long) * ctx->nBodies,
                                   hipMemcpyDeviceToHost, s));
    }
    if (ctx->nContacts > 0)
        HIPCHK(ctx, hipMemcpyAsync(colour.data(), ctx->cur.ccolour, sizeof(int) * ctx->nContacts,
                                   hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    // no dynamic body twice in one colour (each constraint's colour is its
    // contact's persistent colour, whatever order the solve used)
    std::vector<unsigned long long> rebuilt(ctx->nBodies, 0ull);
    for (int k = 0; k < n; ++k) {
        const int ck = moved.empty() ? contact[k] : moved[contact[k]];
        if (ck < 0 || ck >= ctx->nContacts) continue;
        const int c = colour[ck];
        if (c < 0 || c >= kMaxColours) continue;
        for (int b : {bodies[k].x, bodies[k].y}) {
            if (b < 0 || b >= ctx->nBodies || btype[b] != BT_DYNAMIC) continue;
            if (rebuilt[b] & (1ull << c)) ++*conflicts;
            rebuilt[b] |= 1ull << c;
        }
    }
    for (int b = 0; b < ctx->nBodies; ++b)
        if (btype[b] == BT_DYNAMIC && rebuilt[b] != mask[b]) ++*maskErrors;
    return B2G_OK;
}

}  // extern "C"

extern "C" {

int b2g_set_body_worlds(b2g_context* ctx, int first, int count, const int32_t* worlds) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->nBodies || (count > 0 && !worlds))
        return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    bool any = false;
    for (int i = 0; i < count; ++i) {
        if (worlds[i] < 0 || worlds[i] > 0xFFFFFF) return fail(ctx, B2G_ERR_ARG, "sub-world id out of range");
        any = any || worlds[i] != 0;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->bworld + first, worlds, sizeof(int32_t) * count, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (any && !ctx->subWorlds) {
        ctx->subWorlds = true;
        ctx->newFixture = true;  // re-run the broad-phase with sub-world cells
    }
    return B2G_OK;
}

}  // extern "C"

extern "C" {

int b2g_abi_sizes(int32_t* out, int cap) {
    const int32_t sizes[] = {
        (int32_t)sizeof(b2g_shape),       (int32_t)sizeof(b2g_manifold),     (int32_t)sizeof(b2g_body_state),
        (int32_t)sizeof(b2g_fixture_def), (int32_t)sizeof(b2g_world_def),    (int32_t)sizeof(b2g_step_params),
        (int32_t)sizeof(b2g_step_stats),  (int32_t)sizeof(b2g_phase_profile), (int32_t)sizeof(b2g_mlp_params),
        (int32_t)sizeof(b2g_aabb),        (int32_t)sizeof(b2g_transform),    (int32_t)sizeof(b2g_revolute_def),
        (int32_t)sizeof(b2g_schedule),    (int32_t)sizeof(b2g_sweep),        (int32_t)sizeof(b2g_distance_def),
        (int32_t)sizeof(b2g_rope_def),    (int32_t)sizeof(b2g_weld_def),     (int32_t)sizeof(b2g_friction_def),
        (int32_t)sizeof(b2g_motor_def),   (int32_t)sizeof(b2g_pulley_def),   (int32_t)sizeof(b2g_mouse_def),
        (int32_t)sizeof(b2g_prismatic_def), (int32_t)sizeof(b2g_wheel_def), (int32_t)sizeof(b2g_gear_def),
    };
    const int n = (int)(sizeof(sizes) / sizeof(sizes[0]));
    if (!out) return n;
    for (int i = 0; i < n && i < cap; ++i) out[i] = sizes[i];
    return n;
}

}  // extern "C"

// ------------------------------------------------------------ snapshot
// A snapshot is the world state a step reads: header, then the body,
// fixture, contact and joint arrays in a fixed order. Everything else
// (islands, constraint rows, grids, hash sets, counters) is rebuilt by the
// step or by the restore, so a restored context steps exactly as the
// original would.
namespace {
constexpr uint32_t kSnapMagic = 0x53473242u;  // "B2GS"
constexpr uint32_t kSnapVersion = 5;
struct SnapHeader {
    uint32_t magic, version;
    int32_t nBodies, nFixtures, nContacts, nJoints;
    float gravity[2];
    float inv_dt0, cellSize;
    uint32_t claimTag;
    uint8_t allowSleep, warmStarting, continuous, subStepping, autoClearForces, newFixture, cellSizeSet, subWorlds;
    uint8_t staticAsleep, stepComplete, pad[2];
    int32_t jointColourStart[kColourSlots + 1];
    int32_t orderCount, orderSpan;  // broad-phase order keys (0, 0: unset)
};

// Walks the snapshot layout: `visit(devicePointer, bytes)` for each array.
template <typename F>
void snapArrays(b2g_context* ctx, int nb, int nf, int nc, int nj, bool subWorlds, int orderCount, int orderSpan,
                F visit) {
    visit(ctx->btype, sizeof(int) * nb);
    visit(ctx->bflags, sizeof(int) * nb);
    visit(ctx->xf, sizeof(float4) * nb);
    visit(ctx->pos, sizeof(float4) * nb);
    visit(ctx->pos0, sizeof(float4) * nb);
    visit(ctx->vel, sizeof(float4) * nb);
    visit(ctx->force, sizeof(float4) * nb);
    visit(ctx->mass, sizeof(float4) * nb);
    visit(ctx->lcenter, sizeof(float2) * nb);
    visit(ctx->gsleep, sizeof(float2) * nb);
    visit(ctx->colourMask, sizeof(unsigned long long) * nb);
    if (subWorlds) visit(ctx->bworld, sizeof(int) * nb);
    visit(ctx->fbody, sizeof(int) * nf);
    visit(ctx->shape, sizeof(b2g_shape) * nf);
    visit(ctx->fmat, sizeof(float4) * nf);
    visit(ctx->ffilter, sizeof(uint32_t) * nf);
    visit(ctx->fgroup, sizeof(int) * nf);
    visit(ctx->fat, sizeof(float4) * nf);
    visit(ctx->moved, sizeof(uint8_t) * nf);
    visit(ctx->cur.cfA, sizeof(int) * nc);
    visit(ctx->cur.cfB, sizeof(int) * nc);
    visit(ctx->cur.cflags, sizeof(int) * nc);
    visit(ctx->cur.cmat, sizeof(float2) * nc);
    visit(ctx->cur.ctan, sizeof(float) * nc);
    visit(ctx->cur.cman, sizeof(b2g_manifold) * nc);
    visit(ctx->cur.ccolour, sizeof(int) * nc);
    visit(ctx->cur.ctoi, sizeof(float) * nc);
    visit(ctx->cur.ctoiCount, sizeof(int) * nc);
    visit(ctx->jBodies, sizeof(int2) * nj);
    visit(ctx->jAnchors, sizeof(float4) * nj);
    visit(ctx->jLimits, sizeof(float4) * nj);
    visit(ctx->jMotorSpeed, sizeof(float) * nj);
    visit(ctx->jFlags, sizeof(int) * nj);
    visit(ctx->jImpulse, sizeof(float4) * nj);
    visit(ctx->jAxis, sizeof(float4) * nj);
    visit(ctx->jExtra, sizeof(float4) * 2 * nj);
    visit(ctx->jOrder, sizeof(int) * nj);
    if (orderSpan > 0) {
        visit(ctx->fixOrder, sizeof(int) * orderCount);
        visit(ctx->orderFix, sizeof(int) * orderSpan);
    }
}
}  // namespace

extern "C" {

int b2g_snapshot(b2g_context* ctx, void* out, size_t cap, size_t* size) {
    if (!ctx || !size) return B2G_ERR_ARG;
    if (ctx->stepOpen) return fail(ctx, B2G_ERR_STATE, "a split step is open");
    if (ctx->jointsDirty && ctx->nJoints > 0) {
        const int rc = recolourJoints(ctx);
        if (rc != B2G_OK) return rc;
    }
    const int nb = ctx->nBodies, nf = ctx->nFixtures, nc = ctx->nContacts, nj = ctx->nJoints;
    size_t total = sizeof(SnapHeader) + 3 * sizeof(int32_t) * (size_t)nj;  // host joint vectors
    snapArrays(ctx, nb, nf, nc, nj, ctx->subWorlds, ctx->orderCount, ctx->orderSpan,
               [&](const void*, size_t bytes) { total += bytes; });
    *size = total;
    if (!out) return B2G_OK;
    if (cap < total) return fail(ctx, B2G_ERR_CAPACITY, "snapshot buffer too small");
    SnapHeader h{};
    h.magic = kSnapMagic;
    h.version = kSnapVersion;
    h.nBodies = nb; h.nFixtures = nf; h.nContacts = nc; h.nJoints = nj;
    h.gravity[0] = ctx->gravity.x; h.gravity[1] = ctx->gravity.y;
    h.inv_dt0 = ctx->inv_dt0; h.cellSize = ctx->cellSize; h.claimTag = ctx->claimTag;
    h.allowSleep = ctx->allowSleep; h.warmStarting = ctx->warmStarting; h.continuous = ctx->continuous;
    h.subStepping = ctx->subStepping; h.autoClearForces = ctx->autoClearForces; h.newFixture = ctx->newFixture;
    h.cellSizeSet = ctx->cellSizeSet; h.subWorlds = ctx->subWorlds; h.staticAsleep = ctx->staticAsleep;
    h.stepComplete = ctx->stepComplete;
    h.orderCount = ctx->orderCount;
    h.orderSpan = ctx->orderSpan;
    for (int c = 0; c <= kColourSlots; ++c)
        h.jointColourStart[c] = nj > 0 && (int)ctx->jointColourStart.size() > c ? ctx->jointColourStart[c] : 0;
    uint8_t* o = (uint8_t*)out;
    memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    for (int j = 0; j < nj; ++j) {
        const int32_t v[3] = {ctx->hJointKeepApart[j], ctx->hJointRemoved[j], 0};
        memcpy(o, v, sizeof(v));
        o += sizeof(v);
    }
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    snapArrays(ctx, nb, nf, nc, nj, ctx->subWorlds, ctx->orderCount, ctx->orderSpan, [&](const void* p, size_t bytes) {
        if (e == hipSuccess && bytes > 0) e = hipMemcpyAsync(o, p, bytes, hipMemcpyDeviceToHost, s);
        o += bytes;
    });
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    HIPCHK(ctx, e);
    return B2G_OK;
}

int b2g_restore(b2g_context* ctx, const void* in, size_t size) {
    if (!ctx || !in) return B2G_ERR_ARG;
    if (ctx->nBodies != 0 || ctx->nFixtures != 0 || ctx->nJoints != 0)
        return fail(ctx, B2G_ERR_STATE, "restore needs a new context");
    SnapHeader h;
    if (size < sizeof(h)) return fail(ctx, B2G_ERR_ARG, "snapshot too short");
    memcpy(&h, in, sizeof(h));
    if (h.magic != kSnapMagic || h.version != kSnapVersion) return fail(ctx, B2G_ERR_ARG, "not a snapshot");
    if (h.nBodies < 0 || h.nFixtures < 0 || h.nContacts < 0 || h.nJoints < 0 || h.orderSpan < 0 ||
        (h.orderSpan > 0 && h.orderCount != h.nFixtures))
        return fail(ctx, B2G_ERR_ARG, "snapshot counts");
    const int nb = h.nBodies, nf = h.nFixtures, nc = h.nContacts, nj = h.nJoints;
    size_t total = sizeof(SnapHeader) + 3 * sizeof(int32_t) * (size_t)nj;
    snapArrays(ctx, nb, nf, nc, nj, h.subWorlds != 0, h.orderCount, h.orderSpan,
               [&](const void*, size_t bytes) { total += bytes; });
    if (size != total) return fail(ctx, B2G_ERR_ARG, "snapshot size does not match its counts");
    int rc;
    if ((rc = growBodies(ctx, std::max(nb, 1))) != B2G_OK) return rc;
    if ((rc = growFixtures(ctx, std::max(nf, 1))) != B2G_OK) return rc;
    if ((rc = growContacts(ctx, std::max(2 * nc, 1024))) != B2G_OK) return rc;
    if (nj > 0 && (rc = growJoints(ctx, nj)) != B2G_OK) return rc;
    if (h.subWorlds && !ctx->bworld) HIPCHK(ctx, hipMalloc((void**)&ctx->bworld, sizeof(int) * ctx->capBodies));
    if (h.orderSpan > 0) {
        freeArray(ctx->fixOrder);
        freeArray(ctx->orderFix);
        HIPCHK(ctx, hipMalloc((void**)&ctx->fixOrder, sizeof(int) * std::max(h.orderCount, 1)));
        HIPCHK(ctx, hipMalloc((void**)&ctx->orderFix, sizeof(int) * h.orderSpan));
    }
    const uint8_t* q = (const uint8_t*)in + sizeof(h);
    ctx->hJointBodies.assign(nj, make_int2(0, 0));
    ctx->hJointKeepApart.assign(nj, 0);
    ctx->hJointRemoved.assign(nj, 0);
    for (int j = 0; j < nj; ++j) {
        int32_t v[3];
        memcpy(v, q, sizeof(v));
        q += sizeof(v);
        ctx->hJointKeepApart[j] = v[0];
        ctx->hJointRemoved[j] = v[1];
    }
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    snapArrays(ctx, nb, nf, nc, nj, h.subWorlds != 0, h.orderCount, h.orderSpan, [&](const void* p, size_t bytes) {
        if (e == hipSuccess && bytes > 0) e = hipMemcpyAsync((void*)p, q, bytes, hipMemcpyHostToDevice, s);
        q += bytes;
    });
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    HIPCHK(ctx, e);
    ctx->hJointBodies2.assign(nj, make_int2(-1, -1));
    if (nj > 0) {
        HIPCHK(ctx, hipMemcpy(ctx->hJointBodies.data(), ctx->jBodies, sizeof(int2) * nj, hipMemcpyDeviceToHost));
        std::vector<int> fl(nj);
        std::vector<float4> ax(nj);
        HIPCHK(ctx, hipMemcpy(fl.data(), ctx->jFlags, sizeof(int) * nj, hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(ax.data(), ctx->jAxis, sizeof(float4) * nj, hipMemcpyDeviceToHost));
        for (int j = 0; j < nj; ++j)
            if (jointType(fl[j]) == JT_GEAR)
                ctx->hJointBodies2[j] = make_int2(floatBits(ax[j].x), floatBits(ax[j].y));
    }
    ctx->nBodies = nb;
    ctx->nFixtures = nf;
    ctx->nContacts = nc;
    ctx->nJoints = nj;
    ctx->gravity = mk(h.gravity[0], h.gravity[1]);
    ctx->inv_dt0 = h.inv_dt0;
    ctx->cellSize = h.cellSize;
    ctx->cellSizeSet = h.cellSizeSet != 0;
    ctx->claimTag = h.claimTag;
    ctx->allowSleep = h.allowSleep != 0;
    ctx->warmStarting = h.warmStarting != 0;
    ctx->continuous = h.continuous != 0;
    ctx->subStepping = h.subStepping != 0;
    ctx->autoClearForces = h.autoClearForces != 0;
    ctx->newFixture = h.newFixture != 0;
    ctx->subWorlds = h.subWorlds != 0;
    ctx->staticAsleep = h.staticAsleep != 0;
    ctx->stepComplete = h.stepComplete != 0;
    ctx->orderCount = h.orderSpan > 0 ? h.orderCount : 0;
    ctx->orderSpan = h.orderSpan;
    // colour claims only compare with the current tag: zero is older than any
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(blocks(ctx->capBodies)), dim3(256), 0, s, ctx->claim,
                       ctx->capBodies, 0ull);
    LAUNCHCHK(ctx);
    ctx->jointColourStart.assign(h.jointColourStart, h.jointColourStart + kColourSlots + 1);
    ctx->jointsDirty = false;
    if ((rc = rebuildHash(ctx)) != B2G_OK) return rc;
    if (nj > 0 && (rc = rebuildJointPairs(ctx)) != B2G_OK) return rc;
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_set_graphs(b2g_context* ctx, int on) {
    if (!ctx) return B2G_ERR_ARG;
    ctx->graphs = on != 0;
    return B2G_OK;
}

int b2g_get_graph_counts(b2g_context* ctx, int64_t* launches, int64_t* captures) {
    if (!ctx) return B2G_ERR_ARG;
    if (launches) *launches = ctx->graphLaunches;
    if (captures) *captures = ctx->graphCaptures;
    return B2G_OK;
}

int b2g_set_schedule(b2g_context* ctx, const b2g_schedule* sc) {
    if (!ctx || !sc) return B2G_ERR_ARG;
    auto blockOk = [](int b) { return b >= 64 && b <= 1024 && b % 64 == 0; };
    if (sc->islands < 0 || sc->islands > 2 || sc->collideOrder < 0 || sc->collideOrder > 2 ||
        !blockOk(sc->solverBlock) || !blockOk(sc->collideBlock) || !blockOk(sc->queryBlock) ||
        sc->tailColours < 0 || sc->tailColours > 8192 ||
        (sc->clusterBodies != 0 &&
         (sc->clusterBodies < 64 || sc->clusterBodies > 2048 || (sc->clusterBodies & (sc->clusterBodies - 1)) != 0)) ||
        sc->persistentSolve < 0 || sc->persistentSolve > 16)
        return fail(ctx, B2G_ERR_ARG, "schedule field out of range");
    ctx->sched = *sc;
    return B2G_OK;
}

int b2g_get_schedule(b2g_context* ctx, b2g_schedule* sc) {
    if (!ctx || !sc) return B2G_ERR_ARG;
    *sc = ctx->sched;
    return B2G_OK;
}

int b2g_autotune(b2g_context* ctx, const b2g_step_params* p, int rounds, b2g_schedule* out) {
    if (!ctx || !p || rounds <= 0) return B2G_ERR_ARG;
    if (p->mode != B2G_MODE_COLOURED)
        return fail(ctx, B2G_ERR_UNSUPPORTED, "the schedule applies to the coloured step");
    // Each field is judged by the time of the phases it affects, from the
    // events the profiler records on the step stream (dispatch gaps and
    // host round trips inside a phase included).
    struct Knob {
        int32_t b2g_schedule::*field;
        std::vector<int> values;
        std::vector<int> phases;
    };
    const Knob knobs[] = {
        {&b2g_schedule::islands, {0, 1, 2}, {PH_ISLANDS}},
        {&b2g_schedule::collideOrder, {0, 1, 2}, {PH_COLLIDE}},
        {&b2g_schedule::solverBlock, {64, 128, 256, 512}, {PH_WARM, PH_SOLVE_V, PH_SOLVE_X}},
        {&b2g_schedule::collideBlock, {64, 128, 256}, {PH_COLLIDE}},
        {&b2g_schedule::queryBlock, {64, 128, 256}, {PH_BROADPHASE}},
        {&b2g_schedule::tailColours, {0, 512, 1024, 2048}, {PH_WARM, PH_SOLVE_V, PH_SOLVE_X}},
    };
    (void)hipSetDevice(ctx->device);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    // the caller's profile survives the tuning
    const bool wasProfiling = ctx->profiling;
    double savedMs[32];
    int savedLaunches[32];
    for (int i = 0; i < 32; ++i) {
        savedMs[i] = ctx->profMs[i];
        savedLaunches[i] = ctx->profLaunches[i];
    }
    ctx->profiling = true;
    ctx->evMarks.clear();
    ctx->evUsed = 0;
    int rc = B2G_OK;
    for (const Knob& k : knobs) {
        const int nv = (int)k.values.size();
        std::vector<double> total(nv, 0.0);
        for (int r = 0; r < rounds && rc == B2G_OK; ++r) {
            for (int j = 0; j < nv && rc == B2G_OK; ++j) {
                const int i = (j + r) % nv;
                ctx->sched.*k.field = k.values[i];
                for (int ph : k.phases) ctx->profMs[ph] = 0.0;
                rc = b2g_step(ctx, p);
                profCollect(ctx);
                for (int ph : k.phases) total[i] += ctx->profMs[ph];
            }
        }
        if (rc != B2G_OK) break;
        ctx->sched.*k.field = k.values[std::min_element(total.begin(), total.end()) - total.begin()];
    }
    ctx->profiling = wasProfiling;
    ctx->evMarks.clear();
    ctx->evUsed = 0;
    for (int i = 0; i < 32; ++i) {
        ctx->profMs[i] = savedMs[i];
        ctx->profLaunches[i] = savedLaunches[i];
    }
    if (rc != B2G_OK) return rc;
    if (out) *out = ctx->sched;
    return B2G_OK;
}

int b2g_add_joints(b2g_context* ctx, const b2g_revolute_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        const b2g_revolute_def& d = defs[i];
        if (d.bodyA < 0 || d.bodyB < 0 || d.bodyA >= ctx->nBodies || d.bodyB >= ctx->nBodies || d.bodyA == d.bodyB)
            return fail(ctx, B2G_ERR_ARG, "joint bodies out of range or equal");
    }
    if (first) *first = ctx->nJoints;
    if (count == 0) return B2G_OK;
    int rc = growJoints(ctx, ctx->nJoints + count);
    if (rc != B2G_OK) return rc;
    std::vector<int2> b(count);
    std::vector<float4> anchors(count), limits(count), imp(count, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    std::vector<float> speed(count);
    std::vector<int> flags(count);
    bool keepApart = false;
    for (int i = 0; i < count; ++i) {
        const b2g_revolute_def& d = defs[i];
        b[i] = make_int2(d.bodyA, d.bodyB);
        anchors[i] = make_float4(d.localAnchorA[0], d.localAnchorA[1], d.localAnchorB[0], d.localAnchorB[1]);
        limits[i] = make_float4(d.referenceAngle, d.lowerAngle, d.upperAngle, d.maxMotorTorque);
        speed[i] = d.motorSpeed;
        flags[i] = (d.enableLimit ? JF_LIMIT : 0) | (d.enableMotor ? JF_MOTOR : 0) | (d.collideConnected ? JF_COLLIDE : 0);
        ctx->hJointBodies.push_back(b[i]);
        ctx->hJointKeepApart.push_back(d.collideConnected ? 0 : 1);
        ctx->hJointRemoved.push_back(0);
        keepApart = keepApart || !d.collideConnected;
    }
    hipStream_t s = ctx->stream;
    const int n0 = ctx->nJoints;
    HIPCHK(ctx, hipMemcpyAsync(ctx->jBodies + n0, b.data(), sizeof(int2) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jAnchors + n0, anchors.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jLimits + n0, limits.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jMotorSpeed + n0, speed.data(), sizeof(float) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags + n0, flags.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse + n0, imp.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->nJoints += count;
    ctx->jointsDirty = true;
    if (keepApart) return rebuildJointPairs(ctx);
    return B2G_OK;
}

int b2g_get_joints(b2g_context* ctx, int first, int count, float* out) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->nJoints || (count > 0 && !out)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->jImpulse + first, sizeof(float4) * count, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_set_joint(b2g_context* ctx, int joint, const b2g_revolute_def* d) {
    if (!ctx || !d || joint < 0 || joint >= ctx->nJoints) return B2G_ERR_ARG;
    if (ctx->hJointRemoved[joint]) return fail(ctx, B2G_ERR_ARG, "joint was removed");
    hipStream_t s = ctx->stream;
    int flags = 0;
    float4 lim0, imp;
    HIPCHK(ctx, hipMemcpyAsync(&flags, ctx->jFlags + joint, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&lim0, ctx->jLimits + joint, sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&imp, ctx->jImpulse + joint, sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (jointType(flags) != JT_REVOLUTE) return fail(ctx, B2G_ERR_ARG, "not a revolute joint");
    // EnableLimit and SetLimits drop the limit impulse when they change
    // something (REF/Dynamics/Joints/b2RevoluteJoint.cpp:440-471).
    const bool limitSwitched = ((flags & JF_LIMIT) != 0) != (d->enableLimit != 0);
    const bool limitsMoved = lim0.y != d->lowerAngle || lim0.z != d->upperAngle;
    if (limitSwitched || limitsMoved) imp.z = 0.0f;
    flags = (flags & ~(JF_LIMIT | JF_MOTOR)) | (d->enableLimit ? JF_LIMIT : 0) | (d->enableMotor ? JF_MOTOR : 0);
    const float4 lim = make_float4(lim0.x, d->lowerAngle, d->upperAngle, d->maxMotorTorque);
    HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags + joint, &flags, sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jLimits + joint, &lim, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jMotorSpeed + joint, &d->motorSpeed, sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse + joint, &imp, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

}  // extern "C"

// One new row of the joint table: bodies, anchors, per-type parameters
// (limits), motor speed and flags.
struct JointRow {
    int bodyA, bodyB;
    float4 anchors, params;
    int flags;
    float4 axis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float speed = 0.0f;
};

// Appends rows to the joint table (impulses start at zero).
static int appendJointRows(b2g_context* ctx, const std::vector<JointRow>& rows, int* first) {
    const int count = (int)rows.size();
    for (const JointRow& d : rows) {
        if (d.bodyA < 0 || d.bodyB < 0 || d.bodyA >= ctx->nBodies || d.bodyB >= ctx->nBodies || d.bodyA == d.bodyB)
            return fail(ctx, B2G_ERR_ARG, "joint bodies out of range or equal");
    }
    if (first) *first = ctx->nJoints;
    if (count == 0) return B2G_OK;
    int rc = growJoints(ctx, ctx->nJoints + count);
    if (rc != B2G_OK) return rc;
    std::vector<int2> b(count);
    std::vector<float4> anchors(count), params(count), imp(count, make_float4(0.0f, 0.0f, 0.0f, 0.0f)), axis(count);
    std::vector<float4> extra(2 * (size_t)count, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    std::vector<float> speed(count, 0.0f);
    std::vector<int> flags(count);
    bool keepApart = false;
    for (int i = 0; i < count; ++i) {
        const JointRow& d = rows[i];
        b[i] = make_int2(d.bodyA, d.bodyB);
        anchors[i] = d.anchors;
        params[i] = d.params;
        flags[i] = d.flags;
        axis[i] = d.axis;
        speed[i] = d.speed;
        const bool collide = (d.flags & JF_COLLIDE) != 0;
        ctx->hJointBodies.push_back(b[i]);
        ctx->hJointBodies2.push_back(jointType(d.flags) == JT_GEAR
                                         ? make_int2(floatBits(d.axis.x), floatBits(d.axis.y))
                                         : make_int2(-1, -1));
        ctx->hJointKeepApart.push_back(collide ? 0 : 1);
        ctx->hJointRemoved.push_back(0);
        keepApart = keepApart || !collide;
    }
    hipStream_t s = ctx->stream;
    const int n0 = ctx->nJoints;
    HIPCHK(ctx, hipMemcpyAsync(ctx->jBodies + n0, b.data(), sizeof(int2) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jAnchors + n0, anchors.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jLimits + n0, params.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jMotorSpeed + n0, speed.data(), sizeof(float) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags + n0, flags.data(), sizeof(int) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse + n0, imp.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jAxis + n0, axis.data(), sizeof(float4) * count, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jExtra + 2 * (size_t)n0, extra.data(), sizeof(float4) * 2 * count,
                               hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->nJoints += count;
    ctx->jointsDirty = true;
    if (keepApart) return rebuildJointPairs(ctx);
    return B2G_OK;
}

// b2Vec2::Normalize (REF/Common/b2Math.h:117-129), for the local axis of a
// prismatic joint.
static float4 normalizedAxis(const float* a) {
    float x = a[0], y = a[1];
    const float len = sqrtf(x * x + y * y);
    if (!(len < FLT_EPSILON)) {
        const float inv = 1.0f / len;
        x *= inv;
        y *= inv;
    }
    return make_float4(x, y, 0.0f, 0.0f);
}

// Replaces the parameters (limits row) of one joint of the given type.
static int setJointParams(b2g_context* ctx, int joint, int type, const char* what, float4 par) {
    if (!ctx || joint < 0 || joint >= ctx->nJoints) return B2G_ERR_ARG;
    if (ctx->hJointRemoved[joint]) return fail(ctx, B2G_ERR_ARG, "joint was removed");
    hipStream_t s = ctx->stream;
    int flags = 0;
    HIPCHK(ctx, hipMemcpyAsync(&flags, ctx->jFlags + joint, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (jointType(flags) != type) return fail(ctx, B2G_ERR_ARG, what);
    HIPCHK(ctx, hipMemcpyAsync(ctx->jLimits + joint, &par, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

template <class Def>
static JointRow jointRow(const Def& d, int type, float4 params) {
    return JointRow{d.bodyA, d.bodyB,
                    make_float4(d.localAnchorA[0], d.localAnchorA[1], d.localAnchorB[0], d.localAnchorB[1]), params,
                    (type << 12) | (d.collideConnected ? JF_COLLIDE : 0)};
}

extern "C" {

int b2g_add_distance_joints(b2g_context* ctx, const b2g_distance_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i)
        rows.push_back(jointRow(defs[i], JT_DISTANCE,
                                make_float4(defs[i].length, defs[i].frequencyHz, defs[i].dampingRatio, 0.0f)));
    return appendJointRows(ctx, rows, first);
}

int b2g_add_rope_joints(b2g_context* ctx, const b2g_rope_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i)
        rows.push_back(jointRow(defs[i], JT_ROPE, make_float4(defs[i].maxLength, 0.0f, 0.0f, 0.0f)));
    return appendJointRows(ctx, rows, first);
}

int b2g_set_rope_joint(b2g_context* ctx, int joint, float maxLength) {
    return setJointParams(ctx, joint, JT_ROPE, "not a rope joint", make_float4(maxLength, 0.0f, 0.0f, 0.0f));
}

int b2g_add_weld_joints(b2g_context* ctx, const b2g_weld_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i)
        rows.push_back(jointRow(defs[i], JT_WELD,
                                make_float4(defs[i].referenceAngle, defs[i].frequencyHz, defs[i].dampingRatio, 0.0f)));
    return appendJointRows(ctx, rows, first);
}

int b2g_add_friction_joints(b2g_context* ctx, const b2g_friction_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i)
        rows.push_back(jointRow(defs[i], JT_FRICTION, make_float4(defs[i].maxForce, defs[i].maxTorque, 0.0f, 0.0f)));
    return appendJointRows(ctx, rows, first);
}

int b2g_set_friction_joint(b2g_context* ctx, int joint, float maxForce, float maxTorque) {
    return setJointParams(ctx, joint, JT_FRICTION, "not a friction joint",
                          make_float4(maxForce, maxTorque, 0.0f, 0.0f));
}

int b2g_add_motor_joints(b2g_context* ctx, const b2g_motor_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_motor_def& d = defs[i];
        rows.push_back(JointRow{d.bodyA, d.bodyB,
                                make_float4(d.linearOffset[0], d.linearOffset[1], d.angularOffset, d.correctionFactor),
                                make_float4(d.maxForce, d.maxTorque, 0.0f, 0.0f),
                                (JT_MOTOR << 12) | (d.collideConnected ? JF_COLLIDE : 0)});
    }
    return appendJointRows(ctx, rows, first);
}

int b2g_set_motor_joint(b2g_context* ctx, int joint, const b2g_motor_def* d) {
    if (!ctx || !d) return B2G_ERR_ARG;
    int rc = setJointParams(ctx, joint, JT_MOTOR, "not a motor joint", make_float4(d->maxForce, d->maxTorque, 0.0f, 0.0f));
    if (rc != B2G_OK) return rc;
    const float4 an = make_float4(d->linearOffset[0], d->linearOffset[1], d->angularOffset, d->correctionFactor);
    HIPCHK(ctx, hipMemcpyAsync(ctx->jAnchors + joint, &an, sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_add_pulley_joints(b2g_context* ctx, const b2g_pulley_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_pulley_def& d = defs[i];
        if (!(d.ratio > 0.0f)) return fail(ctx, B2G_ERR_ARG, "pulley ratio must be positive");
        rows.push_back(jointRow(d, JT_PULLEY,
                                make_float4(d.groundAnchorA[0], d.groundAnchorA[1], d.groundAnchorB[0],
                                            d.groundAnchorB[1])));
    }
    const int n0 = ctx->nJoints;
    int rc = appendJointRows(ctx, rows, first);
    if (rc != B2G_OK || count == 0) return rc;
    // ratio and the constant lengthA + ratio lengthB (b2PulleyJoint ctor,
    // REF/Dynamics/Joints/b2PulleyJoint.cpp:57-75)
    std::vector<float> ratio(count);
    std::vector<float4> imp(count);
    for (int i = 0; i < count; ++i) {
        ratio[i] = defs[i].ratio;
        imp[i] = make_float4(0.0f, 0.0f, 0.0f, defs[i].lengthA + defs[i].ratio * defs[i].lengthB);
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->jMotorSpeed + n0, ratio.data(), sizeof(float) * count, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse + n0, imp.data(), sizeof(float4) * count, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_add_mouse_joints(b2g_context* ctx, const b2g_mouse_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_mouse_def& d = defs[i];
        rows.push_back(JointRow{d.bodyA, d.bodyB,
                                make_float4(d.target[0], d.target[1], d.localAnchorB[0], d.localAnchorB[1]),
                                make_float4(d.maxForce, d.frequencyHz, d.dampingRatio, d.massB),
                                (JT_MOUSE << 12) | (d.collideConnected ? JF_COLLIDE : 0)});
    }
    return appendJointRows(ctx, rows, first);
}

int b2g_set_mouse_joint(b2g_context* ctx, int joint, const b2g_mouse_def* d) {
    if (!ctx || !d) return B2G_ERR_ARG;
    int rc = setJointParams(ctx, joint, JT_MOUSE, "not a mouse joint",
                            make_float4(d->maxForce, d->frequencyHz, d->dampingRatio, d->massB));
    if (rc != B2G_OK) return rc;
    float4 an;
    HIPCHK(ctx, hipMemcpyAsync(&an, ctx->jAnchors + joint, sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    an.x = d->target[0];
    an.y = d->target[1];
    HIPCHK(ctx, hipMemcpyAsync(ctx->jAnchors + joint, &an, sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_add_prismatic_joints(b2g_context* ctx, const b2g_prismatic_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_prismatic_def& d = defs[i];
        JointRow r = jointRow(d, JT_PRISMATIC,
                              make_float4(d.referenceAngle, d.lowerTranslation, d.upperTranslation, d.maxMotorForce));
        r.flags |= (d.enableLimit ? JF_LIMIT : 0) | (d.enableMotor ? JF_MOTOR : 0);
        r.axis = normalizedAxis(d.localAxisA);
        r.speed = d.motorSpeed;
        rows.push_back(r);
    }
    return appendJointRows(ctx, rows, first);
}

// Shared by the prismatic and wheel setters: motor speed, switches, the
// limits row; a changed limit switch or range drops the limit impulse.
static int setAxisJoint(b2g_context* ctx, int joint, int type, const char* what, float4 par, int enableLimit,
                        int enableMotor, float speed) {
    if (!ctx || joint < 0 || joint >= ctx->nJoints) return B2G_ERR_ARG;
    if (ctx->hJointRemoved[joint]) return fail(ctx, B2G_ERR_ARG, "joint was removed");
    hipStream_t s = ctx->stream;
    int flags = 0;
    float4 lim0, imp;
    HIPCHK(ctx, hipMemcpyAsync(&flags, ctx->jFlags + joint, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&lim0, ctx->jLimits + joint, sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&imp, ctx->jImpulse + joint, sizeof(float4), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (jointType(flags) != type) return fail(ctx, B2G_ERR_ARG, what);
    if (type == JT_PRISMATIC) {
        // b2PrismaticJoint::EnableLimit / SetLimits, REF/Dynamics/Joints/b2PrismaticJoint.cpp:546-578
        const bool limitSwitched = ((flags & JF_LIMIT) != 0) != (enableLimit != 0);
        const bool limitsMoved = lim0.y != par.y || lim0.z != par.z;
        if (limitSwitched || limitsMoved) imp.z = 0.0f;
        flags = (flags & ~JF_LIMIT) | (enableLimit ? JF_LIMIT : 0);
    }
    flags = (flags & ~JF_MOTOR) | (enableMotor ? JF_MOTOR : 0);
    HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags + joint, &flags, sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jLimits + joint, &par, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jMotorSpeed + joint, &speed, sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->jImpulse + joint, &imp, sizeof(float4), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return B2G_OK;
}

int b2g_set_prismatic_joint(b2g_context* ctx, int joint, const b2g_prismatic_def* d) {
    if (!d) return B2G_ERR_ARG;
    return setAxisJoint(ctx, joint, JT_PRISMATIC, "not a prismatic joint",
                        make_float4(d->referenceAngle, d->lowerTranslation, d->upperTranslation, d->maxMotorForce),
                        d->enableLimit, d->enableMotor, d->motorSpeed);
}

int b2g_add_wheel_joints(b2g_context* ctx, const b2g_wheel_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_wheel_def& d = defs[i];
        JointRow r = jointRow(d, JT_WHEEL, make_float4(d.frequencyHz, d.dampingRatio, d.maxMotorTorque, 0.0f));
        r.flags |= d.enableMotor ? JF_MOTOR : 0;
        r.axis = make_float4(d.localAxisA[0], d.localAxisA[1], 0.0f, 0.0f);
        r.speed = d.motorSpeed;
        rows.push_back(r);
    }
    return appendJointRows(ctx, rows, first);
}

int b2g_set_wheel_joint(b2g_context* ctx, int joint, const b2g_wheel_def* d) {
    if (!d) return B2G_ERR_ARG;
    return setAxisJoint(ctx, joint, JT_WHEEL, "not a wheel joint",
                        make_float4(d->frequencyHz, d->dampingRatio, d->maxMotorTorque, 0.0f), 0, d->enableMotor,
                        d->motorSpeed);
}

int b2g_add_gear_joints(b2g_context* ctx, const b2g_gear_def* defs, int count, int* first) {
    if (!ctx || count < 0 || (count > 0 && !defs)) return B2G_ERR_ARG;
    std::vector<JointRow> rows;
    for (int i = 0; i < count; ++i) {
        const b2g_gear_def& d = defs[i];
        if (d.bodyC < 0 || d.bodyD < 0 || d.bodyC >= ctx->nBodies || d.bodyD >= ctx->nBodies)
            return fail(ctx, B2G_ERR_ARG, "gear bodies C, D out of range");
        JointRow r{d.bodyA, d.bodyB,
                   make_float4(d.localAnchorA[0], d.localAnchorA[1], d.localAnchorB[0], d.localAnchorB[1]),
                   make_float4(d.referenceAngleA, d.referenceAngleB, d.ratio, d.constant),
                   (JT_GEAR << 12) | (d.collideConnected ? JF_COLLIDE : 0) | (d.prismaticA ? JF_GEAR_PRISM_A : 0) |
                       (d.prismaticB ? JF_GEAR_PRISM_B : 0)};
        union { int i; float f; } c, dd;
        c.i = d.bodyC;
        dd.i = d.bodyD;
        r.axis = make_float4(c.f, dd.f, 0.0f, 0.0f);
        rows.push_back(r);
    }
    const int n0 = ctx->nJoints;
    int rc = appendJointRows(ctx, rows, first);
    if (rc != B2G_OK || count == 0) return rc;
    std::vector<float4> extra(2 * (size_t)count);
    for (int i = 0; i < count; ++i) {
        const b2g_gear_def& d = defs[i];
        extra[2 * i] = make_float4(d.localAnchorC[0], d.localAnchorC[1], d.localAnchorD[0], d.localAnchorD[1]);
        extra[2 * i + 1] = make_float4(d.localAxisC[0], d.localAxisC[1], d.localAxisD[0], d.localAxisD[1]);
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->jExtra + 2 * (size_t)n0, extra.data(), sizeof(float4) * 2 * count,
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return B2G_OK;
}

int b2g_set_gear_joint(b2g_context* ctx, int joint, float ratio) {
    if (!ctx || joint < 0 || joint >= ctx->nJoints) return B2G_ERR_ARG;
    float4 par;
    HIPCHK(ctx, hipMemcpyAsync(&par, ctx->jLimits + joint, sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    par.z = ratio;  // b2GearJoint::SetRatio leaves the constant
    return setJointParams(ctx, joint, JT_GEAR, "not a gear joint", par);
}

int b2g_get_joint_reactions(b2g_context* ctx, int first, int count, float* out) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->nJoints || (count > 0 && !out)) return B2G_ERR_ARG;
    if (count == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    float4* d = nullptr;
    HIPCHK(ctx, hipMalloc((void**)&d, sizeof(float4) * count));
    hipLaunchKernelGGL(k_joint_reactions, dim3(blocks(count)), dim3(256), 0, s, ctx->jview(), first, count, d);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpyAsync(out, d, sizeof(float4) * count, hipMemcpyDeviceToHost, s);
    const hipError_t e3 = hipStreamSynchronize(s);
    (void)hipFree(d);
    HIPCHK(ctx, e1);
    HIPCHK(ctx, e2);
    HIPCHK(ctx, e3);
    return B2G_OK;
}

int b2g_set_weld_joint(b2g_context* ctx, int joint, float frequencyHz, float dampingRatio) {
    if (!ctx || joint < 0 || joint >= ctx->nJoints) return B2G_ERR_ARG;
    if (ctx->hJointRemoved[joint]) return fail(ctx, B2G_ERR_ARG, "joint was removed");
    float4 par;
    HIPCHK(ctx, hipMemcpyAsync(&par, ctx->jLimits + joint, sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return setJointParams(ctx, joint, JT_WELD, "not a weld joint", make_float4(par.x, frequencyHz, dampingRatio, 0.0f));
}

int b2g_set_distance_joint(b2g_context* ctx, int joint, float length, float frequencyHz, float dampingRatio) {
    return setJointParams(ctx, joint, JT_DISTANCE, "not a distance joint",
                          make_float4(length, frequencyHz, dampingRatio, 0.0f));
}

int b2g_remove_joints(b2g_context* ctx, const int* ids, int count) {
    if (!ctx || count < 0 || (count > 0 && !ids)) return B2G_ERR_ARG;
    for (int i = 0; i < count; ++i)
        if (ids[i] < 0 || ids[i] >= ctx->nJoints) return fail(ctx, B2G_ERR_ARG, "joint index out of range");
    if (count == 0) return B2G_OK;
    hipStream_t s = ctx->stream;
    bool keptApart = false;
    for (int i = 0; i < count; ++i) {
        const int j = ids[i];
        if (ctx->hJointRemoved[j]) continue;
        keptApart = keptApart || ctx->hJointKeepApart[j];
        ctx->hJointRemoved[j] = 1;
        const int flags = JF_REMOVED;
        HIPCHK(ctx, hipMemcpyAsync(ctx->jFlags + j, &flags, sizeof(int), hipMemcpyHostToDevice, s));
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->jointsDirty = true;
    if (keptApart) return rebuildJointPairs(ctx);
    return B2G_OK;
}

}  // extern "C"

// ------------------------------------------------ pair set without allocation
// b2g_pairs_batch_noalloc: b2g_pairs_batch with every device buffer of the
// call (staged inputs, grid, pair keys, sort and scan temporaries, counters)
// carved out of one file-scope array per GPU, for callers that cannot
// allocate device memory. Limits: kNoAllocMaxProxies proxies and
// kNoAllocMaxPairs pairs per call.
namespace {

constexpr int kNoAllocMaxProxies = 65536;
constexpr int kNoAllocMaxPairs = 1 << 20;
constexpr int kNoAllocMaxBuckets = 2 * kNoAllocMaxProxies;  // power of two >= 2n, at least 1024
constexpr size_t kNoAllocSortBytes = size_t(24) << 20;
constexpr size_t noallocAlign(size_t b) { return (b + 255) & ~size_t(255); }
// byte offsets into the array
constexpr size_t kNaFat = 0;
constexpr size_t kNaMoved = kNaFat + noallocAlign(sizeof(float4) * kNoAllocMaxProxies);
constexpr size_t kNaCounts = kNaMoved + noallocAlign(kNoAllocMaxProxies);
constexpr size_t kNaStarts = kNaCounts + noallocAlign(sizeof(int) * (kNoAllocMaxBuckets + 1));
constexpr size_t kNaCursor = kNaStarts + noallocAlign(sizeof(int) * (kNoAllocMaxBuckets + 1));
constexpr size_t kNaEntries = kNaCursor + noallocAlign(sizeof(int) * (kNoAllocMaxBuckets + 1));
constexpr size_t kNaCell = kNaEntries + noallocAlign(sizeof(int) * kMaxCellsPerProxy * kNoAllocMaxProxies);
constexpr size_t kNaEfat = kNaCell + noallocAlign(sizeof(unsigned long long) * kMaxCellsPerProxy * kNoAllocMaxProxies);
constexpr size_t kNaQueries = kNaEfat + noallocAlign(sizeof(float4) * kMaxCellsPerProxy * kNoAllocMaxProxies);
constexpr size_t kNaBig = kNaQueries + noallocAlign(sizeof(int) * kNoAllocMaxProxies);
constexpr size_t kNaRange = kNaBig + noallocAlign(sizeof(unsigned long long) * kNoAllocMaxProxies);
constexpr size_t kNaKeys = kNaRange + noallocAlign(sizeof(int4) * kNoAllocMaxProxies);
constexpr size_t kNaSorted = kNaKeys + noallocAlign(sizeof(unsigned long long) * kNoAllocMaxPairs);
constexpr size_t kNaSort = kNaSorted + noallocAlign(sizeof(unsigned long long) * kNoAllocMaxPairs);
constexpr size_t kNaCounters = kNaSort + noallocAlign(kNoAllocSortBytes);
constexpr size_t kNoAllocBytes = kNaCounters + noallocAlign(sizeof(Counters));

}  // namespace

// One copy per GPU, held until the process ends (about 72 MiB). Host code
// never passes this name to a kernel: it uses the device address that
// hipGetSymbolAddress returned for the GPU (noallocBase).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noallocMutex;
void* g_noallocBase[kNoAllocMaxDevices] = {};
bool g_noallocTried[kNoAllocMaxDevices] = {};

// Resolves this GPU's copy of the scratch array; called by b2g_create with
// the device current, once per device.
void noallocInit(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_noallocMutex);
    if (g_noallocTried[device]) return;
    g_noallocTried[device] = true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_noallocBase[device] = p;
}

void* noallocBase(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_noallocMutex);
    return g_noallocBase[device];
}

// Sorted pair keys to (min, max) proxy index pairs.
__global__ void k_unpack_pairs(const unsigned long long* keys, int n, int32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    out[2 * i] = (int32_t)(k >> 32);
    out[2 * i + 1] = (int32_t)(k & 0xFFFFFFFFu);
}

}  // namespace

extern "C" {

int b2g_pairs_batch_noalloc(b2g_context* ctx, const b2g_aabb* fat, const uint8_t* moved, int n, int32_t* pairs,
                            int cap, int* count) {
    if (!ctx || n < 0 || cap < 0 || (n > 0 && (!fat || !moved)) || (cap > 0 && !pairs)) return B2G_ERR_ARG;
    if (count) *count = 0;
    if (n > kNoAllocMaxProxies)
        return fail(ctx, B2G_ERR_CAPACITY, "b2g_pairs_batch_noalloc takes at most 65536 proxies");
    if (n == 0) return B2G_OK;
    unsigned char* base = (unsigned char*)noallocBase(ctx->device);
    if (!base) return fail(ctx, B2G_ERR_HIP, "no scratch address for this device");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    // Cell size: median extent of up to 1024 evenly spaced proxies (any
    // cell size gives the same pair set; this one keeps cells near the
    // typical proxy).
    float ext[1024];
    const int ns = std::min(n, 1024);
    for (int i = 0; i < ns; ++i) {
        const b2g_aabb& a = fat[(long long)i * n / ns];
        ext[i] = std::max(a.ux - a.lx, a.uy - a.ly);
    }
    std::nth_element(ext, ext + ns / 2, ext + ns);
    const float cell = std::max(ext[ns / 2], 1e-3f);
    int nb = 1024;
    while (nb < 2 * n) nb <<= 1;
    const size_t sortBytes = prims_scratch_bytes(std::max(std::max(nb + 1, kNoAllocMaxPairs), n));
    if (sortBytes > kNoAllocSortBytes) return fail(ctx, B2G_ERR_CAPACITY, "sort temporaries exceed the scratch");
    float4* dfat = (float4*)(base + kNaFat);
    uint8_t* dmoved = base + kNaMoved;
    int* counts = (int*)(base + kNaCounts);
    int* starts = (int*)(base + kNaStarts);
    int* cursor = (int*)(base + kNaCursor);
    unsigned long long* keys = (unsigned long long*)(base + kNaKeys);
    unsigned long long* sorted = (unsigned long long*)(base + kNaSorted);
    void* sortTmp = base + kNaSort;
    Counters* dc = (Counters*)(base + kNaCounters);
    HIPCHK(ctx, hipMemcpyAsync(dfat, fat, sizeof(float4) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dmoved, moved, n, hipMemcpyHostToDevice, s));
    // what the previous call left: counters and bucket counts start at zero
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks(sizeof(Counters) / sizeof(int))), dim3(256), 0, s, (int*)dc,
                       (int)(sizeof(Counters) / sizeof(int)), 0);
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks(nb + 1)), dim3(256), 0, s, counts, nb + 1, 0);
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks(nb + 1)), dim3(256), 0, s, cursor, nb + 1, 0);
    GridView g;
    g.invCell = 1.0f / cell;
    g.bucketMask = (uint32_t)(nb - 1);
    g.counts = counts; g.starts = starts; g.cursor = cursor;
    g.entries = (int*)(base + kNaEntries);
    g.ecell = (unsigned long long*)(base + kNaCell);
    g.efat = (float4*)(base + kNaEfat);
    g.queries = (int*)(base + kNaQueries);
    g.big = (unsigned long long*)(base + kNaBig);
    g.bigSorted = 0;
    g.range = (int4*)(base + kNaRange);
    g.fbody = nullptr;
    g.bworld = nullptr;
    hipLaunchKernelGGL(k_grid_count, dim3(blocks(n)), dim3(256), 0, s, dfat, dmoved, n, (const uint8_t*)nullptr, g, dc);
    LAUNCHCHK(ctx);
    HIPCHK(ctx, prims_exclusive_sum(sortTmp, kNoAllocSortBytes, counts, starts, nb + 1, s));
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks(n)), dim3(256), 0, s, n, g, dfat, dmoved);
    WorldView w;
    memset(&w, 0, sizeof(w));
    PairSink sink{keys, kNoAllocMaxPairs, 0ull, nullptr};
    hipLaunchKernelGGL(k_pairs_query, dim3(blocks(n * kQueryLanes)), dim3(256), 0, s, dfat, dmoved, n, g, w, false,
                       (const unsigned long long*)nullptr, 0u, sink, dc);
    hipLaunchKernelGGL(k_pairs_wide, dim3(blocks(n)), dim3(256), 0, s, dfat, dmoved, n, g, w, false,
                       (const unsigned long long*)nullptr, 0u, sink, dc);
    LAUNCHCHK(ctx);
    Counters hc;
    HIPCHK(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    const int np = hc.nNewPairs;
    if (count) *count = np;
    if (hc.overflow || np > kNoAllocMaxPairs)
        return fail(ctx, B2G_ERR_CAPACITY, "b2g_pairs_batch_noalloc holds at most 1048576 pairs");
    const int out = std::min(np, cap);
    if (out > 0) {
        HIPCHK(ctx, prims_sort_keys_u64(sortTmp, kNoAllocSortBytes, keys, sorted, np, s));
        // the unsorted keys are dead: their slots take the unpacked pairs
        hipLaunchKernelGGL(k_unpack_pairs, dim3(blocks(out)), dim3(256), 0, s, sorted, out, (int32_t*)keys);
        LAUNCHCHK(ctx);
        HIPCHK(ctx, hipMemcpyAsync(pairs, keys, sizeof(int32_t) * 2 * (size_t)out, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
    }
    return B2G_OK;
}

}  // extern "C"
