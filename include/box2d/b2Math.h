// box2d:: math and settings of the host API (the fork's Box2D/Common/b2Vec.h,
// b2Math.h and b2Settings.h surface: b2Vec<T,N>, b2Rot, b2Transform,
// b2Sweep, b2Mat22, b2Dot/b2Cross/b2Mul/b2MulT and the UPPER_CASE
// constants). FP32 operations evaluate in upstream Box2D's order; build
// with FP contraction off.
#pragma once
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <array>
#include <initializer_list>
#include <cstddef>

namespace box2d {

// Box2D/Common/b2Settings.h:36-122
constexpr float B2_PI = 3.14159265359f;
constexpr int32_t MAX_MANIFOLD_POINTS = 2;
constexpr int32_t MAX_POLYGON_VERTICES = 8;
constexpr float AABB_EXTENSION = 0.1f;
constexpr float AABB_MULTIPLIER = 2.0f;
constexpr float LINEAR_SLOP = 0.005f;
constexpr float ANGULAR_SLOP = 2.0f / 180.0f * B2_PI;
constexpr float POLYGON_RADIUS = 2.0f * LINEAR_SLOP;
constexpr float VELOCITY_THRESHOLD = 1.0f;
constexpr float MAX_LINEAR_CORRECTION = 0.2f;
constexpr float MAX_TRANSLATION = 2.0f;
constexpr float MAX_ROTATION = 0.5f * B2_PI;
constexpr float BAUMGARTE = 0.2f;
constexpr float TIME_TO_SLEEP = 0.5f;
constexpr float LINEAR_SLEEP_TOLERANCE = 0.01f;
constexpr float ANGULAR_SLEEP_TOLERANCE = 2.0f / 180.0f * B2_PI;

#ifndef B2_NOT_USED
#define B2_NOT_USED(x) ((void)(x))
#endif

constexpr int b2VecX = 0;
constexpr int b2VecY = 1;
constexpr int b2VecZ = 2;

inline bool b2IsValid(float x) { return isfinite(x) != 0; }
inline float b2Sqrt(float x) { return sqrtf(x); }
inline float b2Atan2(float y, float x) { return atan2f(y, x); }

// Fixed-size vector, indexed with b2VecX / b2VecY.
template <typename T, std::size_t N>
class b2Vec {
    std::array<T, N> d_;

public:
    b2Vec() : d_{} {}
    b2Vec(std::initializer_list<T> list) : d_{} {
        std::size_t i = 0;
        for (T v : list) {
            if (i < N) d_[i++] = v;
        }
    }
    template <typename... Args, typename = typename std::enable_if<sizeof...(Args) == N && (N > 1)>::type>
    b2Vec(Args... args) : d_{{static_cast<T>(args)...}} {}

    std::size_t size() const { return N; }
    T& operator[](std::size_t i) { return d_[i]; }
    const T& operator[](std::size_t i) const { return d_[i]; }
    auto begin() { return d_.begin(); }
    auto end() { return d_.end(); }

    void SetZero() { d_.fill(T(0)); }
    b2Vec operator-() const {
        b2Vec r;
        for (std::size_t i = 0; i < N; ++i) r[i] = -d_[i];
        return r;
    }
    b2Vec operator+(const b2Vec& o) const {
        b2Vec r;
        for (std::size_t i = 0; i < N; ++i) r[i] = d_[i] + o[i];
        return r;
    }
    b2Vec operator-(const b2Vec& o) const {
        b2Vec r;
        for (std::size_t i = 0; i < N; ++i) r[i] = d_[i] - o[i];
        return r;
    }
    b2Vec operator*(T s) const {
        b2Vec r;
        for (std::size_t i = 0; i < N; ++i) r[i] = d_[i] * s;
        return r;
    }
    b2Vec operator/(T s) const {
        b2Vec r;
        for (std::size_t i = 0; i < N; ++i) r[i] = d_[i] / s;
        return r;
    }
    b2Vec& operator+=(const b2Vec& o) {
        for (std::size_t i = 0; i < N; ++i) d_[i] += o[i];
        return *this;
    }
    b2Vec& operator-=(const b2Vec& o) {
        for (std::size_t i = 0; i < N; ++i) d_[i] -= o[i];
        return *this;
    }
    b2Vec& operator*=(T s) {
        for (auto& v : d_) v *= s;
        return *this;
    }
    bool operator==(const b2Vec& o) const { return d_ == o.d_; }
    bool operator!=(const b2Vec& o) const { return !(d_ == o.d_); }
    T LengthSquared() const {
        T r{};
        for (auto v : d_) r += v * v;
        return r;
    }
    T Length() const { return static_cast<T>(sqrt(static_cast<double>(LengthSquared()))); }
    bool IsValid() const {
        for (auto v : d_)
            if (!b2IsValid(static_cast<float>(v))) return false;
        return true;
    }
};

// float specialisation keeps FP32 arithmetic (sqrtf) and b2Vec2's Normalize.
template <>
inline float b2Vec<float, 2>::Length() const {
    return sqrtf(d_[0] * d_[0] + d_[1] * d_[1]);
}

using b2Vec2 = b2Vec<float, 2>;
using b2Vec3 = b2Vec<float, 3>;

inline float b2Normalize(b2Vec2& v) {
    const float len = v.Length();
    if (len < FLT_EPSILON) return 0.0f;
    const float inv = 1.0f / len;
    v[0] *= inv;
    v[1] *= inv;
    return len;
}

template <typename T, std::size_t N>
inline b2Vec<T, N> operator*(T s, const b2Vec<T, N>& v) {
    b2Vec<T, N> r;
    for (std::size_t i = 0; i < N; ++i) r[i] = s * v[i];
    return r;
}

inline float b2Dot(const b2Vec2& a, const b2Vec2& b) { return a[0] * b[0] + a[1] * b[1]; }
inline float b2Cross(const b2Vec2& a, const b2Vec2& b) { return a[0] * b[1] - a[1] * b[0]; }
inline b2Vec2 b2Cross(const b2Vec2& a, float s) { return b2Vec2(s * a[1], -s * a[0]); }
inline b2Vec2 b2Cross(float s, const b2Vec2& a) { return b2Vec2(-s * a[1], s * a[0]); }
inline float b2DistanceSquared(const b2Vec2& a, const b2Vec2& b) {
    b2Vec2 c = a - b;
    return b2Dot(c, c);
}
inline float b2Distance(const b2Vec2& a, const b2Vec2& b) { return (a - b).Length(); }
inline b2Vec2 Skew(const b2Vec2& v) { return b2Vec2(-v[1], v[0]); }

template <typename T>
inline T b2Min(T a, T b) { return a < b ? a : b; }
template <typename T>
inline T b2Max(T a, T b) { return a > b ? a : b; }
template <typename T>
inline T b2Abs(T a) { return a > T(0) ? a : -a; }
template <typename T>
inline T b2Clamp(T a, T lo, T hi) { return b2Max(lo, b2Min(a, hi)); }
inline b2Vec2 b2Min(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(b2Min(a[0], b[0]), b2Min(a[1], b[1])); }
inline b2Vec2 b2Max(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(b2Max(a[0], b[0]), b2Max(a[1], b[1])); }

// 2x2 matrix stored by columns ex, ey.
struct b2Mat22 {
    b2Mat22() {}
    b2Mat22(const b2Vec2& c1, const b2Vec2& c2) : ex(c1), ey(c2) {}
    void SetZero() { ex.SetZero(); ey.SetZero(); }
    b2Mat22 GetInverse() const {
        const float a = ex[0], b = ey[0], c = ex[1], d = ey[1];
        float det = a * d - b * c;
        if (det != 0.0f) det = 1.0f / det;
        return b2Mat22(b2Vec2(det * d, -det * c), b2Vec2(-det * b, det * a));
    }
    b2Vec2 ex, ey;
};
inline b2Vec2 b2Mul(const b2Mat22& A, const b2Vec2& v) {
    return b2Vec2(A.ex[0] * v[0] + A.ey[0] * v[1], A.ex[1] * v[0] + A.ey[1] * v[1]);
}

// Rotation (sine and cosine of the angle).
struct b2Rot {
    b2Rot() : s(0.0f), c(1.0f) {}
    explicit b2Rot(float angle) : s(sinf(angle)), c(cosf(angle)) {}
    void Set(float angle) {
        s = sinf(angle);
        c = cosf(angle);
    }
    void SetIdentity() { s = 0.0f; c = 1.0f; }
    float GetAngle() const { return b2Atan2(s, c); }
    b2Vec2 GetXAxis() const { return b2Vec2(c, s); }
    b2Vec2 GetYAxis() const { return b2Vec2(-s, c); }
    float s, c;
};

struct b2Transform {
    b2Transform() : p(0.0f, 0.0f), q() {}
    b2Transform(const b2Vec2& position, const b2Rot& rotation) : p(position), q(rotation) {}
    void SetIdentity() { p.SetZero(); q.SetIdentity(); }
    void Set(const b2Vec2& position, float angle) { p = position; q.Set(angle); }
    b2Vec2 p;
    b2Rot q;
};

// Body sweep (centre of mass path over the step).
struct b2Sweep {
    b2Vec2 localCenter;
    b2Vec2 c0, c;
    float a0 = 0.0f, a = 0.0f;
    float alpha0 = 0.0f;
};

inline b2Vec2 b2Mul(const b2Rot& q, const b2Vec2& v) { return b2Vec2(q.c * v[0] - q.s * v[1], q.s * v[0] + q.c * v[1]); }
inline b2Vec2 b2MulT(const b2Rot& q, const b2Vec2& v) { return b2Vec2(q.c * v[0] + q.s * v[1], -q.s * v[0] + q.c * v[1]); }
inline b2Vec2 b2Mul(const b2Transform& T, const b2Vec2& v) {
    const float x = (T.q.c * v[0] - T.q.s * v[1]) + T.p[0];
    const float y = (T.q.s * v[0] + T.q.c * v[1]) + T.p[1];
    return b2Vec2(x, y);
}
inline b2Vec2 b2MulT(const b2Transform& T, const b2Vec2& v) {
    const float px = v[0] - T.p[0];
    const float py = v[1] - T.p[1];
    return b2Vec2(T.q.c * px + T.q.s * py, -T.q.s * px + T.q.c * py);
}
inline b2Rot b2MulT(const b2Rot& q, const b2Rot& r) {
    b2Rot o;
    o.s = q.c * r.s - q.s * r.c;
    o.c = q.c * r.c + q.s * r.s;
    return o;
}
inline b2Transform b2MulT(const b2Transform& A, const b2Transform& B) {
    b2Transform C;
    C.q = b2MulT(A.q, B.q);
    C.p = b2MulT(A.q, B.p - A.p);
    return C;
}

}  // namespace box2d
