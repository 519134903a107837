// Device-side building blocks of the gfx950 Box2D step.
//
// FP32 everywhere, evaluated in the order upstream Box2D 2.3.2 evaluates it
// (REF/Common/b2Math.h operators): the library is compiled with
// -ffp-contract=off so no multiply-add is fused, and packed FP32 is disabled.
// Division and sqrt are the correctly rounded IEEE forms (hipcc default), so
// every trig-free expression gives the same bits as the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include "b2g.h"

namespace b2g {

// ------------------------------------------------------------ constants
// Box2D/Common/b2Settings.h:36-122 (fork), REF/Common/b2Settings.h
constexpr float kPi = 3.14159265359f;
constexpr float kLinearSlop = 0.005f;
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

// Body flags (bit layout of b2Body::m_flags, REF/Dynamics/b2Body.h:395-404).
enum : int {
    BF_ISLAND = 0x0001, BF_AWAKE = 0x0002, BF_AUTOSLEEP = 0x0004, BF_BULLET = 0x0008,
    BF_FIXEDROT = 0x0010, BF_ACTIVE = 0x0020,
};
// Contact flags (b2Contact::m_flags, REF/Dynamics/Contacts/b2Contact.h:149-171)
// plus the device-only state bits above 0x100.
enum : int {
    CF_ISLAND = 0x0001, CF_TOUCHING = 0x0002, CF_ENABLED = 0x0004, CF_FILTER = 0x0008,
    CF_TOI = 0x0020,         // cached time of impact is valid (b2Contact::e_toiFlag)
    CF_DEAD = 0x0100,        // destroyed this step, removed by compaction
    CF_CONSTRAINT = 0x0200,  // solved this step (touching, enabled, not a sensor, awake)
    CF_NOOVERLAP = 0x0400,   // fat AABBs stopped overlapping
    CF_UPDATED = 0x0800,     // b2Contact::Update ran on it this step
    CF_WAS_TOUCHING = 0x1000,  // touching before this step's update
    CF_DESTROY = 0x2000,     // destroyed by a body type change; dropped first thing in the next collide
};
enum : int { BT_STATIC = 0, BT_KINEMATIC = 1, BT_DYNAMIC = 2 };

// ------------------------------------------------------------ FP32 vector
struct V2 { float x, y; };
__host__ __device__ inline V2 mk(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
__host__ __device__ inline V2 operator+(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline V2 operator-(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline V2 operator-(V2 a) { return mk(-a.x, -a.y); }
__host__ __device__ inline V2 operator*(float s, V2 a) { return mk(s * a.x, s * a.y); }
__host__ __device__ inline void operator+=(V2& a, V2 b) { a.x += b.x; a.y += b.y; }
__host__ __device__ inline void operator-=(V2& a, V2 b) { a.x -= b.x; a.y -= b.y; }
__host__ __device__ inline void operator*=(V2& a, float s) { a.x *= s; a.y *= s; }
__host__ __device__ inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
__host__ __device__ inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
__host__ __device__ inline V2 cross(V2 a, float s) { return mk(s * a.y, -s * a.x); }
__host__ __device__ inline V2 cross(float s, V2 a) { return mk(-s * a.y, s * a.x); }
__host__ __device__ inline float lengthSq(V2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline float distSq(V2 a, V2 b) { V2 c = a - b; return dot(c, c); }
__host__ __device__ inline float fminB(float a, float b) { return a < b ? a : b; }
__host__ __device__ inline float fmaxB(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline V2 vmin(V2 a, V2 b) { return mk(fminB(a.x, b.x), fminB(a.y, b.y)); }
__host__ __device__ inline V2 vmax(V2 a, V2 b) { return mk(fmaxB(a.x, b.x), fmaxB(a.y, b.y)); }
__host__ __device__ inline float clampB(float a, float lo, float hi) { return fmaxB(lo, fminB(a, hi)); }
__host__ __device__ inline float absB(float a) { return a > 0.0f ? a : -a; }
__device__ inline float lengthOf(V2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
// b2Vec2::Normalize: vectors shorter than FLT_EPSILON are left unchanged.
__device__ inline float normalizeV(V2& v) {
    float len = lengthOf(v);
    if (len < FLT_EPSILON) return 0.0f;
    float inv = 1.0f / len;
    v.x *= inv;
    v.y *= inv;
    return len;
}
__host__ __device__ inline V2 ld(const b2g_vec2& a) { return mk(a.x, a.y); }
__host__ __device__ inline b2g_vec2 st(V2 a) { b2g_vec2 r; r.x = a.x; r.y = a.y; return r; }

// ------------------------------------------------------------ rotations
struct Rot { float s, c; };
struct Xf { V2 p; Rot q; };

// Rotation of a body angle inside the step. FP32 rounding of the
// double-precision sine and cosine: correctly rounded except in rare
// double-rounding cases, within 1 ulp of glibc's sinf/cosf that b2Rot::Set
// calls on the host (REF/Common/b2Math.h:302-311).
__device__ inline Rot rotOf(float angle) {
    Rot q;
    double s, c;
    sincos((double)angle, &s, &c);
    q.s = (float)s;
    q.c = (float)c;
    return q;
}
__host__ __device__ inline V2 mul(Rot q, V2 v) { return mk(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
__host__ __device__ inline V2 mulT(Rot q, V2 v) { return mk(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }
__host__ __device__ inline V2 mul(const Xf& T, V2 v) {
    float x = (T.q.c * v.x - T.q.s * v.y) + T.p.x;
    float y = (T.q.s * v.x + T.q.c * v.y) + T.p.y;
    return mk(x, y);
}
__host__ __device__ inline V2 mulT(const Xf& T, V2 v) {
    float px = v.x - T.p.x;
    float py = v.y - T.p.y;
    float x = (T.q.c * px + T.q.s * py);
    float y = (-T.q.s * px + T.q.c * py);
    return mk(x, y);
}
__host__ __device__ inline Rot mulT(Rot q, Rot r) {
    Rot o;
    o.s = q.c * r.s - q.s * r.c;
    o.c = q.c * r.c + q.s * r.s;
    return o;
}
__host__ __device__ inline Xf mulT(const Xf& A, const Xf& B) {
    Xf C;
    C.q = mulT(A.q, B.q);
    C.p = mulT(A.q, B.p - A.p);
    return C;
}
__host__ __device__ inline Xf xfOf(float4 t) { Xf x; x.p = mk(t.x, t.y); x.q.s = t.z; x.q.c = t.w; return x; }
__host__ __device__ inline float4 f4(const Xf& x) { return make_float4(x.p.x, x.p.y, x.q.s, x.q.c); }

// ------------------------------------------------------------ AABBs
// b2TestOverlap(b2AABB, b2AABB): touching boxes overlap.
__host__ __device__ inline bool overlaps(float4 a, float4 b) {
    float d1x = b.x - a.z, d1y = b.y - a.w;
    float d2x = a.x - b.z, d2y = a.y - b.w;
    if (d1x > 0.0f || d1y > 0.0f) return false;
    if (d2x > 0.0f || d2y > 0.0f) return false;
    return true;
}
// b2AABB::Contains
__host__ __device__ inline bool containsBox(float4 a, float4 b) {
    return a.x <= b.x && a.y <= b.y && b.z <= a.z && b.w <= a.w;
}

// Contact feature id (b2ContactFeature packed as b2ContactID::key).
__host__ __device__ inline uint32_t makeId(int iA, int iB, int tA, int tB) {
    return (uint32_t)(iA & 0xFF) | ((uint32_t)(iB & 0xFF) << 8) | ((uint32_t)(tA & 0xFF) << 16) |
           ((uint32_t)(tB & 0xFF) << 24);
}
__host__ __device__ inline uint32_t swapId(uint32_t id) {
    return ((id >> 8) & 0xFF) | ((id & 0xFF) << 8) | (((id >> 24) & 0xFF) << 16) | (((id >> 16) & 0xFF) << 24);
}
constexpr int kFeatVertex = 0, kFeatFace = 1;

// ------------------------------------------------------------ world view
// Device pointers of one world's state, passed by value to kernels. All
// arrays are structure-of-arrays indexed by creation order (bodies,
// fixtures = broad-phase proxies) or by table slot (contacts, constraints).
struct WorldView {
    // bodies
    int nBodies;
    int* btype;
    int* bflags;
    float4* xf;        // p.x p.y q.s q.c
    float4* pos;       // sweep c.x c.y a, w: unused
    float4* pos0;      // sweep c0.x c0.y a0, w: alpha0
    float2* lcenter;   // sweep localCenter
    float4* vel;       // v.x v.y w, w: unused
    float4* force;     // f.x f.y torque
    float4* mass;      // invMass invI linearDamping angularDamping
    float2* gsleep;    // gravityScale sleepTime
    unsigned long long* colourMask;
    int* islandOf;     // island root (smallest body index of the island) per body
    int* islandAwake;  // per root: the island holds an awake body
    int* islandSleep;  // per root: 1 if a body has rested less than the time to sleep
    int* islandMinSep; // per position pass and root: 1 if a point is deeper than 3 * linearSlop
    int* bworld;       // sub-world per body (null: a single world)
    const unsigned long long* jointPairs;  // hash set of body pairs a joint keeps apart (null: none)
    uint32_t jointPairMask;
    // fixtures / proxies
    int nFixtures;
    int* fbody;
    b2g_shape* shape;
    float4* fmat;      // friction restitution density sensor(0/1)
    uint32_t* ffilter; // categoryBits | maskBits << 16
    int* fgroup;
    float4* fat;       // fat AABB lower.x lower.y upper.x upper.y
    uint8_t* moved;    // proxy is in the move buffer
    // Broad-phase order of the proxies: new pairs are (smaller, larger) key
    // and sorted by key, as the reference sorts its pair buffer by tree proxy
    // id; null: the fixture index is the key.
    const int* ford;   // key per fixture
    const int* ordf;   // fixture per key (-1: none)
    // contacts
    int nContacts;
    int* cfA;
    int* cfB;
    int* cflags;
    float2* cmat;      // friction restitution
    float* ctan;       // tangent speed (b2Contact::m_tangentSpeed)
    b2g_manifold* cman;
    int* ccolour;
    float* ctoi;       // cached time of impact (with CF_TOI)
    int* ctoiCount;    // TOI events of the contact in the step
};

}  // namespace b2g
