// Continuous collision for the gfx950 step: GJK distance between convex
// proxies and the time of impact of two swept shapes, evaluated per lane in
// FP32 in the reference's order of operations (contraction off), so the
// results equal the CPU's bit for bit wherever the trig does (rotations of
// swept angles use rotOf, the FP32 rounding of the double sine and cosine).
//
//   b2DistanceProxy / b2Distance   REF/Collision/b2Distance.cpp:31-606
//   b2SeparationFunction           REF/Collision/b2TimeOfImpact.cpp:35-253
//   b2TimeOfImpact                 REF/Collision/b2TimeOfImpact.cpp:257-489
//   b2Sweep                        REF/Common/b2Math.h:383-405, 695-721
#pragma once
#include "b2g_device.h"

namespace b2g {

constexpr float kToiBaumgarte = 0.75f;  // b2_toiBaugarte
constexpr int kMaxSubSteps = 8;         // b2_maxSubSteps
constexpr int kMaxTOIContacts = 32;     // b2_maxTOIContacts

// Convex proxy: a circle is its centre with its radius, a polygon its
// vertices with the polygon skin, an edge its two ends.
struct Proxy {
    V2 v[8];
    int count;
    float radius;
};

__device__ inline Proxy proxyOf(const b2g_shape& s) {
    Proxy p;
    p.radius = s.radius;
    if (s.type == B2G_CIRCLE) {
        p.count = 1;
        p.v[0] = ld(s.centroid);
    } else if (s.type == B2G_EDGE) {
        p.count = 2;
        p.v[0] = ld(s.verts[0]);
        p.v[1] = ld(s.verts[1]);
    } else {
        p.count = s.count;
        for (int i = 0; i < 8; ++i)
            if (i < s.count) p.v[i] = ld(s.verts[i]);
    }
    return p;
}

// First vertex of largest projection on d.
__device__ inline int support(const Proxy& p, V2 d) {
    int best = 0;
    float bestValue = dot(p.v[0], d);
    for (int i = 1; i < p.count; ++i) {
        const float value = dot(p.v[i], d);
        if (value > bestValue) {
            best = i;
            bestValue = value;
        }
    }
    return best;
}

struct Sweep {
    V2 localCenter, c0, c;
    float a0, a, alpha0;
};

// Transform at fraction beta of the sweep's remaining interval.
__device__ inline Xf sweepXf(const Sweep& sw, float beta) {
    Xf xf;
    xf.p = (1.0f - beta) * sw.c0 + beta * sw.c;
    const float angle = (1.0f - beta) * sw.a0 + beta * sw.a;
    xf.q = rotOf(angle);
    xf.p -= mul(xf.q, sw.localCenter);
    return xf;
}

// Moves the sweep's start to time alpha.
__device__ inline void sweepAdvance(Sweep& sw, float alpha) {
    const float beta = (alpha - sw.alpha0) / (1.0f - sw.alpha0);
    sw.c0 += beta * (sw.c - sw.c0);
    sw.a0 += beta * (sw.a - sw.a0);
    sw.alpha0 = alpha;
}

// Shifts both angles by whole turns so the start angle is in [0, 2 pi).
__device__ inline void sweepNormalize(Sweep& sw) {
    const float twoPi = 2.0f * kPi;
    const float d = twoPi * floorf(sw.a0 / twoPi);
    sw.a0 -= d;
    sw.a -= d;
}

struct SimplexCache {
    float metric;
    int count;
    int indexA[3], indexB[3];
};
struct SimplexVertex {
    V2 wA, wB, w;
    float a;
    int indexA, indexB;
};

__device__ inline float simplexMetric(const SimplexVertex* v, int count) {
    if (count == 2) return lengthOf(v[0].w - v[1].w);
    if (count == 3) return cross(v[1].w - v[0].w, v[2].w - v[0].w);
    return 0.0f;
}

// Closest points of two proxies (no radii), continuing from `cache`.
__device__ inline float gjkDistance(SimplexCache& cache, const Proxy& pA, const Xf& xfA, const Proxy& pB,
                                    const Xf& xfB) {
    SimplexVertex v[3];
    int count = cache.count;
    for (int i = 0; i < 3; ++i) {
        if (i >= count) break;
        v[i].indexA = cache.indexA[i];
        v[i].indexB = cache.indexB[i];
        v[i].wA = mul(xfA, pA.v[v[i].indexA]);
        v[i].wB = mul(xfB, pB.v[v[i].indexB]);
        v[i].w = v[i].wB - v[i].wA;
        v[i].a = 0.0f;
    }
    if (count > 1) {
        const float m1 = cache.metric, m2 = simplexMetric(v, count);
        if (m2 < 0.5f * m1 || 2.0f * m1 < m2 || m2 < FLT_EPSILON) count = 0;
    }
    if (count == 0) {
        v[0].indexA = 0;
        v[0].indexB = 0;
        v[0].wA = mul(xfA, pA.v[0]);
        v[0].wB = mul(xfB, pB.v[0]);
        v[0].w = v[0].wB - v[0].wA;
        v[0].a = 1.0f;
        count = 1;
    }
    int saveA[3], saveB[3];
    for (int iter = 0; iter < 20;) {
        const int saveCount = count;
        for (int i = 0; i < 3; ++i) {
            if (i >= saveCount) break;
            saveA[i] = v[i].indexA;
            saveB[i] = v[i].indexB;
        }
        if (count == 2) {
            const V2 w1 = v[0].w, w2 = v[1].w;
            const V2 e12 = w2 - w1;
            const float d12_2 = -dot(w1, e12);
            const float d12_1 = dot(w2, e12);
            if (d12_2 <= 0.0f) {
                v[0].a = 1.0f;
                count = 1;
            } else if (d12_1 <= 0.0f) {
                v[1].a = 1.0f;
                count = 1;
                v[0] = v[1];
            } else {
                const float inv = 1.0f / (d12_1 + d12_2);
                v[0].a = d12_1 * inv;
                v[1].a = d12_2 * inv;
            }
        } else if (count == 3) {
            const V2 w1 = v[0].w, w2 = v[1].w, w3 = v[2].w;
            const V2 e12 = w2 - w1, e13 = w3 - w1, e23 = w3 - w2;
            const float d12_1 = dot(w2, e12), d12_2 = -dot(w1, e12);
            const float d13_1 = dot(w3, e13), d13_2 = -dot(w1, e13);
            const float d23_1 = dot(w3, e23), d23_2 = -dot(w2, e23);
            const float n123 = cross(e12, e13);
            const float d123_1 = n123 * cross(w2, w3);
            const float d123_2 = n123 * cross(w3, w1);
            const float d123_3 = n123 * cross(w1, w2);
            if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
                v[0].a = 1.0f;
                count = 1;
            } else if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
                const float inv = 1.0f / (d12_1 + d12_2);
                v[0].a = d12_1 * inv;
                v[1].a = d12_2 * inv;
                count = 2;
            } else if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
                const float inv = 1.0f / (d13_1 + d13_2);
                v[0].a = d13_1 * inv;
                v[2].a = d13_2 * inv;
                count = 2;
                v[1] = v[2];
            } else if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
                v[1].a = 1.0f;
                count = 1;
                v[0] = v[1];
            } else if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
                v[2].a = 1.0f;
                count = 1;
                v[0] = v[2];
            } else if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
                const float inv = 1.0f / (d23_1 + d23_2);
                v[1].a = d23_1 * inv;
                v[2].a = d23_2 * inv;
                count = 2;
                v[0] = v[2];
            } else {
                const float inv = 1.0f / (d123_1 + d123_2 + d123_3);
                v[0].a = d123_1 * inv;
                v[1].a = d123_2 * inv;
                v[2].a = d123_3 * inv;
            }
        }
        if (count == 3) break;
        V2 d;
        if (count == 1) {
            d = -v[0].w;
        } else {
            const V2 e12 = v[1].w - v[0].w;
            d = cross(e12, -v[0].w) > 0.0f ? cross(1.0f, e12) : cross(e12, 1.0f);
        }
        if (dot(d, d) < FLT_EPSILON * FLT_EPSILON) break;
        SimplexVertex nv;
        nv.indexA = support(pA, mulT(xfA.q, -d));
        nv.wA = mul(xfA, pA.v[nv.indexA]);
        nv.indexB = support(pB, mulT(xfB.q, d));
        nv.wB = mul(xfB, pB.v[nv.indexB]);
        nv.w = nv.wB - nv.wA;
        nv.a = 0.0f;
        ++iter;
        bool duplicate = false;
        for (int i = 0; i < 3; ++i) {
            if (i >= saveCount) break;
            if (nv.indexA == saveA[i] && nv.indexB == saveB[i]) {
                duplicate = true;
                break;
            }
        }
        if (duplicate) break;
        v[count] = nv;
        ++count;
    }
    V2 pa, pb;
    if (count == 1) {
        pa = v[0].wA;
        pb = v[0].wB;
    } else if (count == 2) {
        pa = v[0].a * v[0].wA + v[1].a * v[1].wA;
        pb = v[0].a * v[0].wB + v[1].a * v[1].wB;
    } else {
        pa = v[0].a * v[0].wA + v[1].a * v[1].wA + v[2].a * v[2].wA;
        pb = pa;
    }
    cache.metric = simplexMetric(v, count);
    cache.count = count;
    for (int i = 0; i < 3; ++i) {
        if (i >= count) break;
        cache.indexA[i] = v[i].indexA;
        cache.indexB[i] = v[i].indexB;
    }
    return lengthOf(pa - pb);
}

// Separation along the axis the last distance query found, as a function
// of time: points (one vertex each), or a face of A or of B.
struct SepFn {
    Sweep sA, sB;
    int type;  // 0 points, 1 face A, 2 face B
    V2 localPoint, axis;
};

__device__ inline void sepInit(SepFn& f, const SimplexCache& cache, const Proxy& pA, const Proxy& pB, float t1) {
    const Xf xfA = sweepXf(f.sA, t1), xfB = sweepXf(f.sB, t1);
    if (cache.count == 1) {
        f.type = 0;
        f.axis = mul(xfB, pB.v[cache.indexB[0]]) - mul(xfA, pA.v[cache.indexA[0]]);
        normalizeV(f.axis);
        return;
    }
    const bool faceB = cache.indexA[0] == cache.indexA[1];
    const Proxy& pf = faceB ? pB : pA;
    const Xf& xff = faceB ? xfB : xfA;
    const int i1 = faceB ? cache.indexB[0] : cache.indexA[0];
    const int i2 = faceB ? cache.indexB[1] : cache.indexA[1];
    const V2 v1 = pf.v[i1], v2 = pf.v[i2];
    f.type = faceB ? 2 : 1;
    f.axis = cross(v2 - v1, 1.0f);
    normalizeV(f.axis);
    const V2 normal = mul(xff.q, f.axis);
    f.localPoint = 0.5f * (v1 + v2);
    const V2 onFace = mul(xff, f.localPoint);
    const V2 other = faceB ? mul(xfA, pA.v[cache.indexA[0]]) : mul(xfB, pB.v[cache.indexB[0]]);
    if (dot(other - onFace, normal) < 0.0f) f.axis = -f.axis;
}

__device__ inline float sepFindMin(const SepFn& f, const Proxy& pA, const Proxy& pB, int* iA, int* iB, float t) {
    const Xf xfA = sweepXf(f.sA, t), xfB = sweepXf(f.sB, t);
    if (f.type == 0) {
        *iA = support(pA, mulT(xfA.q, f.axis));
        *iB = support(pB, mulT(xfB.q, -f.axis));
        return dot(mul(xfB, pB.v[*iB]) - mul(xfA, pA.v[*iA]), f.axis);
    }
    if (f.type == 1) {
        const V2 normal = mul(xfA.q, f.axis);
        const V2 pointA = mul(xfA, f.localPoint);
        *iA = -1;
        *iB = support(pB, mulT(xfB.q, -normal));
        return dot(mul(xfB, pB.v[*iB]) - pointA, normal);
    }
    const V2 normal = mul(xfB.q, f.axis);
    const V2 pointB = mul(xfB, f.localPoint);
    *iB = -1;
    *iA = support(pA, mulT(xfA.q, -normal));
    return dot(mul(xfA, pA.v[*iA]) - pointB, normal);
}

__device__ inline float sepEval(const SepFn& f, const Proxy& pA, const Proxy& pB, int iA, int iB, float t) {
    const Xf xfA = sweepXf(f.sA, t), xfB = sweepXf(f.sB, t);
    if (f.type == 0) return dot(mul(xfB, pB.v[iB]) - mul(xfA, pA.v[iA]), f.axis);
    if (f.type == 1) {
        const V2 normal = mul(xfA.q, f.axis);
        return dot(mul(xfB, pB.v[iB]) - mul(xfA, f.localPoint), normal);
    }
    const V2 normal = mul(xfB.q, f.axis);
    return dot(mul(xfA, pA.v[iA]) - mul(xfB, f.localPoint), normal);
}

enum : int { TOI_UNKNOWN = 0, TOI_FAILED = 1, TOI_OVERLAPPED = 2, TOI_TOUCHING = 3, TOI_SEPARATED = 4 };

// Time of impact in [0, tMax] by conservative advancement on separating
// axes, each root found by alternating bisection and secant steps.
__device__ inline int timeOfImpact(float* tOut, const Proxy& pA, const Proxy& pB, Sweep sA, Sweep sB, float tMax) {
    int state = TOI_UNKNOWN;
    *tOut = tMax;
    sweepNormalize(sA);
    sweepNormalize(sB);
    const float totalRadius = pA.radius + pB.radius;
    const float target = fmaxB(kLinearSlop, totalRadius - 3.0f * kLinearSlop);
    const float tolerance = 0.25f * kLinearSlop;
    float t1 = 0.0f;
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    for (int iter = 0;;) {
        const Xf xfA = sweepXf(sA, t1), xfB = sweepXf(sB, t1);
        const float dist = gjkDistance(cache, pA, xfA, pB, xfB);
        if (dist <= 0.0f) {
            state = TOI_OVERLAPPED;
            *tOut = 0.0f;
            break;
        }
        if (dist < target + tolerance) {
            state = TOI_TOUCHING;
            *tOut = t1;
            break;
        }
        SepFn fcn;
        fcn.sA = sA;
        fcn.sB = sB;
        sepInit(fcn, cache, pA, pB, t1);
        bool done = false;
        float t2 = tMax;
        for (int pushBack = 0;;) {
            int iA, iB;
            float s2 = sepFindMin(fcn, pA, pB, &iA, &iB, t2);
            if (s2 > target + tolerance) {
                state = TOI_SEPARATED;
                *tOut = tMax;
                done = true;
                break;
            }
            if (s2 > target - tolerance) {
                t1 = t2;
                break;
            }
            float s1 = sepEval(fcn, pA, pB, iA, iB, t1);
            if (s1 < target - tolerance) {
                state = TOI_FAILED;
                *tOut = t1;
                done = true;
                break;
            }
            if (s1 <= target + tolerance) {
                state = TOI_TOUCHING;
                *tOut = t1;
                done = true;
                break;
            }
            float a1 = t1, a2 = t2;
            for (int rootIters = 0;;) {
                const float t = (rootIters & 1) ? a1 + (target - s1) * (a2 - a1) / (s2 - s1) : 0.5f * (a1 + a2);
                ++rootIters;
                const float sv = sepEval(fcn, pA, pB, iA, iB, t);
                if (absB(sv - target) < tolerance) {
                    t2 = t;
                    break;
                }
                if (sv > target) {
                    a1 = t;
                    s1 = sv;
                } else {
                    a2 = t;
                    s2 = sv;
                }
                if (rootIters == 50) break;
            }
            if (++pushBack == 8) break;
        }
        ++iter;
        if (done) break;
        if (iter == 20) {
            state = TOI_FAILED;
            *tOut = t1;
            break;
        }
    }
    return state;
}

}  // namespace b2g
