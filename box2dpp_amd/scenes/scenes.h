// Workload scenes shared by the product, its tests and benchmarks, and the
// CPU checkers. A scene is a plain list of body and fixture definitions in
// creation order; every backend replays it through its own API, so the same
// bodies, fixtures and proxy creation order reach every engine.
//
// Scenes (sizes and parameters are the workloads the engine is measured on):
//   hello   : HelloWorld (HelloWorld/HelloWorld.cpp:30-107): ground box
//             50x10 at (0,-10), one dynamic 1x1 box at (0,4), density 1,
//             friction 0.3.
//   pyramid : 20-row box pyramid, 210 dynamic boxes of half-extent 0.5,
//             density 1, friction 0.6, static ground edge (-40,0)-(40,0).
//   pile    : N bodies from seed on a 300x600 m grid, 50% circles
//             (r ~ U[0.25,0.75]), 50% convex polygons (3..8 points on a
//             circle of R ~ U[0.3,0.8] with random phase, hulled by
//             b2PolygonShape::Set), density 1, friction 0.6, restitution 0,
//             inside a static U-shaped container of three edges.
//   worlds  : many independent piles of log-uniform size in [16,1024],
//             each with its own container (one scene per world).
//   pendulums : N single-box pendulums (box 0.5x0.5, density 1) hinged by a
//             revolute joint to one static body at y = 20, 5 m apart, each
//             started with its own horizontal speed; they never touch.
//   chains  : N chains of 32 boxes 0.5x0.1 m (density 1, friction 0.2)
//             joined end to end by revolute joints, the first hinged to a
//             static anchor, laid out horizontally 4 m apart above a ground
//             edge so they fall, swing and hit the ground.
#pragma once
#include <stdint.h>
#include <math.h>
#include <vector>
#include <random>
#include <algorithm>
#include <string>

namespace b2scene {

enum ShapeKind { kCircle = 0, kEdge = 1, kPolygon = 2, kChain = 3 };
enum PolyMode { kPolySet = 0, kPolyBox = 1, kPolyBoxRotated = 2 };

struct BodyDef {
    int type;              // 0 static, 1 kinematic, 2 dynamic (b2BodyType order)
    float px, py, angle;
    float vx, vy, w;
    float linearDamping, angularDamping, gravityScale;
    int allowSleep, awake, fixedRotation, bullet;
    int firstFixture, fixtureCount;
};

struct FixtureDef {
    int shape;             // ShapeKind
    int polyMode;          // PolyMode, polygons only
    int n;                 // polygon: number of input points
    float pts[16];         // polygon input points (x0,y0,x1,y1,...) or box params
    float radius, cx, cy;  // circle: radius and local centre
    float e[8];            // edge: v0, v1, v2, v3 (x,y pairs); chain: prev vertex e[0..1], next e[6..7]
    int has0, has3;        // edge: ghost vertex flags; chain: prev / next vertex set
    float density, friction, restitution;
    int sensor;
    int categoryBits, maskBits, groupIndex;
};

// b2RevoluteJointDef (Box2D/Dynamics/Joints/b2RevoluteJoint.h:37-90) with
// body indices in place of body pointers.
struct JointDef {
    int bodyA, bodyB;
    float ax, ay, bx, by;  // local anchors on A and B
    float referenceAngle;
    float lowerAngle, upperAngle;
    float maxMotorTorque, motorSpeed;
    int enableLimit, enableMotor, collideConnected;
    // b2DistanceJointDef (Box2D/Dynamics/Joints/b2DistanceJoint.h:32-63) when
    // type is 1; b2RopeJointDef (b2RopeJoint.h:30-50, length = maxLength) when
    // 2; b2WeldJointDef (b2WeldJoint.h:29-60, referenceAngle, frequencyHz,
    // dampingRatio) when 3
    // b2FrictionJointDef (b2FrictionJoint.h:27-53, maxForce, maxTorque) when
    // 4; b2MotorJointDef (b2MotorJoint.h:27-56, linearOffset = (ax, ay),
    // angularOffset = referenceAngle, maxForce, maxTorque, correctionFactor)
    // when 5
    // b2PulleyJointDef (b2PulleyJoint.h:30-70, ground anchors, lengthA =
    // length, lengthB, ratio) when 6; b2MouseJointDef (b2MouseJoint.h:28-53,
    // target = (ax, ay), maxForce, frequencyHz, dampingRatio) when 7
    int type;              // 0 revolute, 1 distance, 2 rope, 3 weld, 4 friction, 5 motor, 6 pulley, 7 mouse
    float length, frequencyHz, dampingRatio;
    float maxForce, maxTorque, correctionFactor;
    float gax, gay, gbx, gby, lengthB, ratio;
    // b2PrismaticJointDef (b2PrismaticJoint.h:30-92: localAxisA, referenceAngle,
    // lower / upper translation in lowerAngle / upperAngle, maxMotorForce in
    // maxMotorTorque, motorSpeed) when 8; b2WheelJointDef (b2WheelJoint.h:30-72:
    // localAxisA, frequencyHz, dampingRatio, maxMotorTorque, motorSpeed) when 9
    float axisX, axisY;
    // b2GearJointDef (b2GearJoint.h:24-42: joint1, joint2 as indices of this
    // scene's joints, ratio) when 10
    int joint1, joint2;
};

struct Scene {
    float gx = 0.0f, gy = -10.0f;
    std::vector<BodyDef> bodies;
    std::vector<FixtureDef> fixtures;
    std::vector<JointDef> joints;
    int dynamicCount() const {
        int n = 0;
        for (auto& b : bodies) n += (b.type == 2);
        return n;
    }
};

inline BodyDef defaultBody(int type, float x, float y) {
    BodyDef b{};
    b.type = type; b.px = x; b.py = y; b.angle = 0.0f;
    b.vx = b.vy = b.w = 0.0f;
    b.linearDamping = 0.0f; b.angularDamping = 0.0f; b.gravityScale = 1.0f;
    b.allowSleep = 1; b.awake = 1; b.fixedRotation = 0; b.bullet = 0;
    b.firstFixture = 0; b.fixtureCount = 0;
    return b;
}

inline FixtureDef defaultFixture(int shape) {
    FixtureDef f{};
    f.shape = shape; f.polyMode = kPolySet; f.n = 0;
    f.radius = 0.0f; f.cx = f.cy = 0.0f;
    f.has0 = f.has3 = 0;
    f.density = 0.0f; f.friction = 0.2f; f.restitution = 0.0f;
    f.sensor = 0; f.categoryBits = 1; f.maskBits = 0xFFFF; f.groupIndex = 0;
    return f;
}

inline void addBody(Scene& s, const BodyDef& b) {
    s.bodies.push_back(b);
    s.bodies.back().firstFixture = (int)s.fixtures.size();
    s.bodies.back().fixtureCount = 0;
}

inline void addFixture(Scene& s, const FixtureDef& f) {
    s.fixtures.push_back(f);
    s.bodies.back().fixtureCount++;
}

inline FixtureDef boxFixture(float hx, float hy, float density, float friction) {
    FixtureDef f = defaultFixture(kPolygon);
    f.polyMode = kPolyBox; f.pts[0] = hx; f.pts[1] = hy;
    f.density = density; f.friction = friction;
    return f;
}

inline FixtureDef edgeFixture(float x1, float y1, float x2, float y2) {
    FixtureDef f = defaultFixture(kEdge);
    f.e[2] = x1; f.e[3] = y1; f.e[4] = x2; f.e[5] = y2;
    return f;
}

// HelloWorld/HelloWorld.cpp:30-107
inline Scene hello() {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, -10.0f));
    addFixture(s, boxFixture(50.0f, 10.0f, 0.0f, 0.2f));
    addBody(s, defaultBody(2, 0.0f, 4.0f));
    addFixture(s, boxFixture(1.0f, 1.0f, 1.0f, 0.3f));
    return s;
}

// 20-row pyramid on a ground edge; the layout accumulates the row offsets
// in FP32 exactly as a b2Vec2 loop would.
inline Scene pyramid(int rows = 20, float density = 1.0f, float friction = 0.6f) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-40.0f, 0.0f, 40.0f, 0.0f));
    float xx = -7.0f, xy = 0.75f;
    const float dxx = 0.5625f, dxy = 1.25f;
    const float dyx = 1.125f, dyy = 0.0f;
    for (int i = 0; i < rows; ++i) {
        float yx = xx, yy = xy;
        for (int j = i; j < rows; ++j) {
            addBody(s, defaultBody(2, yx, yy));
            addFixture(s, boxFixture(0.5f, 0.5f, density, friction));
            yx += dyx; yy += dyy;
        }
        xx += dxx; xy += dxy;
    }
    return s;
}

// Uniform float in [0,1) from 24 high bits of one MT19937 draw.
struct Uniform {
    std::mt19937 mt;
    explicit Uniform(uint32_t seed) : mt(seed) {}
    float operator()() { return (float)(mt() >> 8) * (1.0f / 16777216.0f); }
};

// Random pile: N dynamic bodies from `seed` on a width x height grid with
// its lower edge at y = 5, in a U container of half-width width/2 + 5 and
// wall height height + 50. `scale` shrinks the whole layout for the small
// per-world piles of the batched scene.
inline Scene pile(int n, uint32_t seed, float width = 300.0f, float height = 600.0f, float base = 5.0f) {
    Scene s;
    const float half = 0.5f * width + 5.0f;
    const float wallTop = base - 5.0f + height + 50.0f;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-half, 0.0f, half, 0.0f));
    addFixture(s, edgeFixture(-half, 0.0f, -half, wallTop));
    addFixture(s, edgeFixture(half, 0.0f, half, wallTop));
    for (auto& f : s.fixtures) f.friction = 0.6f;

    Uniform U(seed);
    const int cols = (int)ceil(sqrt((double)n * (double)width / (double)height));
    const int rows = (n + cols - 1) / cols;
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        const int c = i % cols, r = i / cols;
        const float x = (float)(-0.5 * width + (c + 0.5) * (double)width / cols);
        const float y = (float)((double)base + (r + 0.5) * (double)height / rows);
        addBody(s, defaultBody(2, x, y));
        FixtureDef f;
        if (i & 1) {
            f = defaultFixture(kCircle);
            f.radius = 0.25f + 0.5f * U();
        } else {
            f = defaultFixture(kPolygon);
            const int k = 3 + std::min(5, (int)(6.0f * U()));
            const float R = 0.3f + 0.5f * U();
            const double phase = 2.0 * pi * (double)U();
            f.polyMode = kPolySet; f.n = k;
            for (int j = 0; j < k; ++j) {
                const double t = phase + 2.0 * pi * j / k;
                f.pts[2 * j + 0] = (float)((double)R * cos(t));
                f.pts[2 * j + 1] = (float)((double)R * sin(t));
            }
        }
        f.density = 1.0f; f.friction = 0.6f; f.restitution = 0.0f;
        addFixture(s, f);
    }
    return s;
}

// Bodies to add to a running world (no static body): the pile's shape mix
// on a grid `width` metres wide, centred on x = 0, from height `base` up, at
// the batched worlds' spacing.
inline Scene rain(int n, uint32_t seed, float width = 12.0f, float base = 25.0f) {
    Scene p = pile(n, seed, width, 1.3416f * (float)((n + 8) / 9), base);
    Scene s;
    for (size_t i = 1; i < p.bodies.size(); ++i) {
        BodyDef b = p.bodies[i];
        b.firstFixture = (int)s.fixtures.size();
        for (int k = 0; k < p.bodies[i].fixtureCount; ++k) s.fixtures.push_back(p.fixtures[p.bodies[i].firstFixture + k]);
        s.bodies.push_back(b);
    }
    return s;
}

// Body count of world w in the batched scene: floor(16 * 64^U), U from seed.
inline std::vector<int> worldSizes(int nWorlds, uint32_t seed, int lo = 16, int hi = 1024) {
    Uniform U(seed);
    std::vector<int> sizes(nWorlds);
    const double ratio = (double)hi / (double)lo;
    for (int w = 0; w < nWorlds; ++w) {
        int v = (int)floor((double)lo * pow(ratio, (double)U()));
        sizes[w] = std::min(hi, std::max(lo, v));
    }
    return sizes;
}

// One world of the batched scene: a pile whose grid is sized so bodies start
// at roughly the same spacing as the large pile (about 1.34 m).
inline Scene batchedWorld(int n, uint32_t seed) {
    const float spacing = 1.3416f;
    const int cols = std::max(4, (int)ceil(sqrt((double)n * 0.5)));
    const int rows = (n + cols - 1) / cols;
    return pile(n, seed, cols * spacing, rows * spacing);
}

// Revolute joint at a world anchor between bodies at angle zero
// (b2RevoluteJointDef::Initialize: local anchor = anchor - body origin,
// reference angle = angleB - angleA = 0).
inline JointDef hinge(const Scene& s, int a, int b, float x, float y) {
    JointDef j{};
    j.bodyA = a; j.bodyB = b;
    j.ax = x - s.bodies[a].px; j.ay = y - s.bodies[a].py;
    j.bx = x - s.bodies[b].px; j.by = y - s.bodies[b].py;
    j.referenceAngle = 0.0f;
    j.lowerAngle = j.upperAngle = 0.0f;
    j.maxMotorTorque = j.motorSpeed = 0.0f;
    j.enableLimit = j.enableMotor = j.collideConnected = 0;
    return j;
}

inline Scene pendulums(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 20.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 5.0f * (float)i;
        const float len = 1.0f + 0.05f * (float)(i % 20);
        BodyDef b = defaultBody(2, x, 20.0f - len);
        b.vx = 0.5f + 0.25f * (float)(i % 13);
        addBody(s, b);
        addFixture(s, boxFixture(0.25f, 0.25f, 1.0f, 0.2f));
        JointDef j = hinge(s, 0, (int)s.bodies.size() - 1, x, 20.0f);
        if (i % 3 == 1) {  // some with a limit, some with a motor
            j.enableLimit = 1; j.lowerAngle = -0.4f; j.upperAngle = 0.3f;
        } else if (i % 3 == 2) {
            j.enableMotor = 1; j.motorSpeed = 1.5f; j.maxMotorTorque = 2.0f;
        }
        s.joints.push_back(j);
    }
    return s;
}

inline Scene chains(int n, int links = 32) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-40.0f, 0.0f, 4.0f * (float)n + 40.0f, 0.0f));
    for (int c = 0; c < n; ++c) {
        const float x0 = 4.0f * (float)c, y = 10.0f + 0.25f * (float)(c % 4);
        addBody(s, defaultBody(0, x0, y));  // anchor
        int prev = (int)s.bodies.size() - 1;
        for (int i = 0; i < links; ++i) {
            addBody(s, defaultBody(2, x0 + 0.25f + 0.5f * (float)i, y));
            addFixture(s, boxFixture(0.25f, 0.05f, 1.0f, 0.2f));
            const int cur = (int)s.bodies.size() - 1;
            s.joints.push_back(hinge(s, prev, cur, x0 + 0.5f * (float)i, y));
            prev = cur;
        }
    }
    return s;
}

// Stress world: the random pile of n bodies (seed) dropped onto a mat of
// `nChains` chains of `links` boxes 0.5 x 0.1 m (density 1, friction 0.6)
// joined end to end by revolute joints that keep neighbours from colliding.
// The mat lies on the container floor in rows 0.25 m apart (at most 18
// chains of 16 m per row across the 300 m floor); the pile's drop grid
// starts 5 m above the mat. The links are bodies in addition to the pile's n.
inline Scene stress(int n, uint32_t seed, int nChains, int links = 32) {
    const float len = 0.5f * (float)links;
    const int perRow = std::max(1, (int)(300.0f / (len + 0.5f)));
    const int rows = (nChains + perRow - 1) / perRow;
    const float matTop = 0.3f + 0.25f * (float)rows;
    Scene s = pile(n, seed, 300.0f, 600.0f, matTop + 5.0f);
    for (int c = 0; c < nChains; ++c) {
        const int r = c / perRow, k = c % perRow;
        const float x0 = -150.0f + 0.25f + (float)k * (len + 0.5f);
        const float y = 0.3f + 0.25f * (float)r;
        int prev = -1;
        for (int i = 0; i < links; ++i) {
            addBody(s, defaultBody(2, x0 + 0.25f + 0.5f * (float)i, y));
            addFixture(s, boxFixture(0.25f, 0.05f, 1.0f, 0.6f));
            const int cur = (int)s.bodies.size() - 1;
            if (prev >= 0) s.joints.push_back(hinge(s, prev, cur, x0 + 0.5f * (float)i, y));
            prev = cur;
        }
    }
    return s;
}

// A plank (a dynamic box 12 x 0.5 m) on the ground carrying n small circles
// (r = 0.05 m) in a row: the plank takes part in n + 1 contacts, more than
// the 64 colours of a body's mask when n > 63.
inline Scene plank(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-20.0f, 0.0f, 20.0f, 0.0f));
    addBody(s, defaultBody(2, 0.0f, 0.25f));
    addFixture(s, boxFixture(6.0f, 0.25f, 1.0f, 0.6f));
    const float step = 11.8f / (float)std::max(n, 1);
    for (int i = 0; i < n; ++i) {
        addBody(s, defaultBody(2, -5.9f + step * ((float)i + 0.5f), 0.55f));
        FixtureDef f = defaultFixture(kCircle);
        f.radius = 0.05f;
        f.density = 1.0f; f.friction = 0.6f;
        addFixture(s, f);
    }
    return s;
}

// One-sided platform (after the reference testbed's OneSidedPlatform): a
// ground edge, a static platform 20 x 0.5 m centred at y = 5, and n boxes
// 0.5 x 0.5 m (fixed rotation, 1.6 m apart) launched upward at 10.5 m/s
// from y = 1 under it. A listener that disables the platform's contacts while a
// box is below its top lets the boxes pass up through it and land on it.
inline Scene oneway(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-20.0f, 0.0f, 20.0f, 0.0f));
    addBody(s, defaultBody(0, 0.0f, 5.0f));
    addFixture(s, boxFixture(10.0f, 0.25f, 0.0f, 0.6f));
    for (int i = 0; i < n; ++i) {
        BodyDef b = defaultBody(2, -8.0f + 1.6f * (float)i, 1.0f);
        b.vy = 10.5f;
        b.fixedRotation = 1;
        addBody(s, b);
        addFixture(s, boxFixture(0.25f, 0.25f, 1.0f, 0.6f));
    }
    return s;
}

// Sensors: n boxes and circles (seeded sizes and spin) dropped through a
// static sensor slab onto a ground edge; every third body carries a second,
// sensor, circle of radius 1.2 m (density 0), so sensors meet sensors,
// dynamic bodies and the ground. A sensor only reports touching
// (b2Contact::Update's sensor branch) and never joins a solve. Seed 0: one
// row of bodies 3 m apart, of fixed rotation, that land apart (one solid
// contact per island).
inline Scene sensors(int n, uint32_t seed) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-30.0f, 0.0f, 30.0f, 0.0f));
    addBody(s, defaultBody(0, 0.0f, 8.0f));
    FixtureDef slab = boxFixture(14.0f, 1.5f, 0.0f, 0.6f);
    slab.sensor = 1;
    addFixture(s, slab);
    Uniform U(seed);
    for (int i = 0; i < n; ++i) {
        BodyDef b = defaultBody(2, -11.0f + 2.4f * (float)(i % 10) + 0.4f * U(), 14.0f + 2.4f * (float)(i / 10));
        b.w = -2.0f + 4.0f * U();
        if (seed == 0) {
            b.px = -13.5f + 3.0f * (float)i;
            b.py = 14.0f;
            b.w = 0.0f;
            b.fixedRotation = 1;
        }
        addBody(s, b);
        FixtureDef f;
        if (i & 1) {
            f = defaultFixture(kCircle);
            f.radius = 0.35f + 0.3f * U();
        } else {
            const float h = 0.35f + 0.3f * U();
            f = boxFixture(h, h, 1.0f, 0.6f);
        }
        f.density = 1.0f; f.friction = 0.6f;
        addFixture(s, f);
        if (i % 3 == 0) {
            FixtureDef halo = defaultFixture(kCircle);
            halo.radius = 1.2f;
            halo.sensor = 1;
            addFixture(s, halo);
        }
    }
    return s;
}

// A conveyor belt: a static platform 20 m long at y = 5 (fixture 1) under n
// boxes of fixed rotation dropped 2.5 m apart, which the belt (a listener
// setting the contacts' tangent speed) carries off its end onto the ground.
inline Scene conveyor(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-40.0f, 0.0f, 40.0f, 0.0f));
    addBody(s, defaultBody(0, 0.0f, 5.0f));
    addFixture(s, boxFixture(10.0f, 0.25f, 0.0f, 0.8f));
    for (int i = 0; i < n; ++i) {
        BodyDef b = defaultBody(2, -9.0f + 2.5f * (float)i, 6.0f);
        b.fixedRotation = 1;
        addBody(s, b);
        addFixture(s, boxFixture(0.4f, 0.4f, 1.0f, 0.8f));
    }
    return s;
}

// Fast bodies for continuous collision: n small circles and boxes fired at
// 80-200 m/s from the left, over a ground edge, through a column of five
// dynamic boxes, at a static wall 0.2 m thick (3 m of travel per step would
// tunnel through it without TOI). Two bodies in three are bullets, so TOI
// events pair bullets with dynamic boxes as well as bodies with the wall.
inline Scene bullets(int n, uint32_t seed) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-40.0f, 0.0f, 40.0f, 0.0f));
    addBody(s, defaultBody(0, 12.0f, 5.0f));
    addFixture(s, boxFixture(0.1f, 5.0f, 0.0f, 0.6f));
    for (int k = 0; k < 5; ++k) {
        addBody(s, defaultBody(2, 8.0f, 0.5f + 1.0f * (float)k));
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.6f));
    }
    Uniform U(seed);
    for (int i = 0; i < n; ++i) {
        BodyDef b = defaultBody(2, -15.0f + 0.5f * (float)(i % 4), 0.8f + 0.7f * (float)(i / 4 % 12));
        b.vx = 80.0f + 120.0f * U();
        b.vy = -5.0f + 10.0f * U();
        b.w = -10.0f + 20.0f * U();
        b.bullet = (i % 3 != 2) ? 1 : 0;
        addBody(s, b);
        FixtureDef f;
        if (i & 1) {
            f = defaultFixture(kCircle);
            f.radius = 0.1f + 0.1f * U();
        } else {
            const float h = 0.1f + 0.1f * U();
            f = boxFixture(h, h, 1.0f, 0.6f);
        }
        f.density = 1.0f; f.friction = 0.6f; f.restitution = 0.1f;
        addFixture(s, f);
    }
    return s;
}

// b2ChainShape of n <= 8 points: CreateLoop when `loop`, else CreateChain,
// with SetPrevVertex / SetNextVertex when the ghosts are given.
inline FixtureDef chainFixture(const float* xy, int n, bool loop, const float* prev = nullptr,
                               const float* next = nullptr) {
    FixtureDef f = defaultFixture(kChain);
    f.n = n;
    f.polyMode = loop ? 1 : 0;
    for (int i = 0; i < 2 * n; ++i) f.pts[i] = xy[i];
    if (prev) { f.e[0] = prev[0]; f.e[1] = prev[1]; f.has0 = 1; }
    if (next) { f.e[6] = next[0]; f.e[7] = next[1]; f.has3 = 1; }
    f.friction = 0.6f;
    return f;
}

// Bodies on chain shapes: a bumpy open chain of 8 points as the floor (with
// ghost vertices beyond both ends), a second open chain without ghosts as a
// ramp, and a hexagonal loop around both; n circles and polygons dropped on
// them from a grid, seeded.
inline Scene terrain(int n, uint32_t seed) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    float floor[16];
    for (int i = 0; i < 8; ++i) {
        floor[2 * i] = -14.0f + 4.0f * (float)i;
        floor[2 * i + 1] = (i & 1) ? 0.8f : 0.0f;
    }
    const float prev[2] = {-18.0f, 1.5f}, next[2] = {18.0f, 1.5f};
    addFixture(s, chainFixture(floor, 8, false, prev, next));
    const float ramp[8] = {6.0f, 6.0f, 9.0f, 5.0f, 12.0f, 4.5f, 15.0f, 4.6f};
    addFixture(s, chainFixture(ramp, 4, false));
    float hex[12];
    for (int i = 0; i < 6; ++i) {
        const double a = 3.14159265358979323846 / 3.0 * (double)i + 0.5235987755982988;
        hex[2 * i] = (float)(24.0 * cos(a));
        hex[2 * i + 1] = (float)(10.0 + 24.0 * sin(a));
    }
    addFixture(s, chainFixture(hex, 6, true));
    Uniform U(seed);
    for (int i = 0; i < n; ++i) {
        BodyDef b = defaultBody(2, -12.0f + 1.5f * (float)(i % 17), 3.0f + 1.5f * (float)(i / 17));
        b.vx = -2.0f + 4.0f * U();
        b.w = -1.0f + 2.0f * U();
        addBody(s, b);
        FixtureDef f;
        if (i & 1) {
            f = defaultFixture(kCircle);
            f.radius = 0.3f + 0.3f * U();
        } else {
            const float h = 0.3f + 0.3f * U();
            f = boxFixture(h, 0.8f * h, 1.0f, 0.6f);
        }
        f.density = 1.0f; f.friction = 0.6f;
        addFixture(s, f);
    }
    return s;
}

// Distance joints: n boxes hung from the ground by rods or springs (every
// other one soft, at 1.5-2.5 Hz with damping 0-0.6), each (when `chained`)
// with a smaller box hanging from it on a rod, over a ground edge. Lengths
// are the anchors' distances at creation (b2DistanceJointDef::Initialize).
inline Scene distanceJoints(int n, bool chained = true) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-200.0f, 0.0f, 200.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 6.0f * (float)i - 3.0f * (float)n;
        const float ax = x + 1.0f + 0.2f * (float)(i % 5), ay = 14.0f;
        addBody(s, defaultBody(2, ax, ay));
        addFixture(s, boxFixture(0.4f, 0.4f, 1.0f, 0.6f));
        const int a = (int)s.bodies.size() - 1;
        JointDef up{};
        up.type = 1;
        up.bodyA = 0; up.bodyB = a;
        up.ax = x; up.ay = 20.0f;         // ground anchor (ground body at the origin)
        up.bx = 0.3f; up.by = 0.3f;       // a corner of box A
        const float dx = (ax + 0.3f) - x, dy = (ay + 0.3f) - 20.0f;
        up.length = sqrtf(dx * dx + dy * dy);
        up.frequencyHz = (i & 1) ? 1.5f + 0.5f * (float)(i % 3) : 0.0f;
        up.dampingRatio = 0.2f * (float)(i % 4);
        s.joints.push_back(up);
        if (!chained) continue;
        const float bx = x + 2.5f, by = 12.5f;
        addBody(s, defaultBody(2, bx, by));
        addFixture(s, boxFixture(0.3f, 0.3f, 1.0f, 0.6f));
        const int b = a + 1;
        JointDef rod{};
        rod.type = 1;
        rod.bodyA = a; rod.bodyB = b;
        const float ex = bx - ax, ey = by - ay;
        rod.length = sqrtf(ex * ex + ey * ey);
        s.joints.push_back(rod);
    }
    return s;
}

// One box on a soft distance joint to a ground anchor 4 m above it, of rest
// length 3 m, so it starts stretched by 1 m: a spring of `hz` Hz and
// damping ratio `ratio` (for runs without gravity).
inline Scene spring(float hz, float ratio) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addBody(s, defaultBody(2, 0.0f, 6.0f));
    addFixture(s, boxFixture(0.25f, 0.25f, 1.0f, 0.6f));
    JointDef j{};
    j.type = 1;
    j.bodyA = 0; j.bodyB = 1;
    j.ax = 0.0f; j.ay = 10.0f;
    j.length = 3.0f;
    j.frequencyHz = hz;
    j.dampingRatio = ratio;
    s.joints.push_back(j);
    return s;
}

// Rope joints: n boxes on ropes from anchors 16 m up, over a ground edge,
// started with sideways speeds and 0-20% slack so the ropes go slack and
// taut; every fourth box starts on its anchor (a rope of length zero), and
// (when `chained`) a smaller box hangs from each on a second rope.
inline Scene ropes(int n, bool chained = true) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-200.0f, 0.0f, 200.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 6.0f * (float)i - 3.0f * (float)n;
        const bool atAnchor = (i % 4) == 3;
        const float ax = atAnchor ? x : x + 1.5f + 0.3f * (float)(i % 4), ay = atAnchor ? 16.0f : 12.0f;
        BodyDef bd = defaultBody(2, ax, ay);
        bd.vx = 3.0f * (float)(i % 3 - 1);
        addBody(s, bd);
        addFixture(s, boxFixture(0.4f, 0.4f, 1.0f, 0.6f));
        const int a = (int)s.bodies.size() - 1;
        JointDef up{};
        up.type = 2;
        up.bodyA = 0; up.bodyB = a;
        up.ax = x; up.ay = 16.0f;
        up.bx = atAnchor ? 0.0f : 0.4f; up.by = atAnchor ? 0.0f : 0.4f;
        const float dx = (ax + up.bx) - x, dy = (ay + up.by) - 16.0f;
        up.length = atAnchor ? 4.0f : sqrtf(dx * dx + dy * dy) * (1.0f + 0.1f * (float)(i % 3));
        s.joints.push_back(up);
        if (!chained) continue;
        const float bx = ax + 0.5f, by = ay - 3.0f;
        addBody(s, defaultBody(2, bx, by));
        addFixture(s, boxFixture(0.3f, 0.3f, 1.0f, 0.6f));
        JointDef low{};
        low.type = 2;
        low.bodyA = a; low.bodyB = a + 1;
        low.length = 1.1f * sqrtf(0.25f + 9.0f);
        s.joints.push_back(low);
    }
    return s;
}

// Weld joints: n cantilever beams of 4-7 boxes (1 x 0.25 m) welded end to
// end from an anchor on the ground body, a box dropped on each (when
// `drop`). Every third
// beam is rigid, the others have soft welds (4 or 8 Hz, damping 0-0.6);
// beams 4, 9, 14, ... start with a fixed-rotation link (the weld then has
// no angular mass when rigid) and odd beams carry a small reference angle.
// `single` makes every beam one link (one weld per island).
inline Scene welds(int n, bool drop = true, bool single = false) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-200.0f, 0.0f, 200.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 10.0f * (float)i - 5.0f * (float)n;
        const int links = single ? 1 : 4 + i % 4;
        const float hz = (i % 3 == 0) ? 0.0f : 4.0f + 4.0f * (float)(i % 2);
        const float ratio = 0.3f * (float)(i % 3);
        const float ref = (i & 1) ? 0.02f : 0.0f;
        int prev = 0;
        for (int k = 0; k < links; ++k) {
            BodyDef bd = defaultBody(2, x + 0.5f + (float)k, 10.0f);
            bd.fixedRotation = (k == 0 && i % 5 == 4) ? 1 : 0;
            addBody(s, bd);
            addFixture(s, boxFixture(0.5f, 0.125f, 1.0f, 0.6f));
            const int b = (int)s.bodies.size() - 1;
            JointDef j{};
            j.type = 3;
            j.bodyA = prev; j.bodyB = b;
            if (prev == 0) { j.ax = x; j.ay = 10.0f; }
            else { j.ax = 0.5f; j.ay = 0.0f; }
            j.bx = -0.5f; j.by = 0.0f;
            j.referenceAngle = k == 0 ? 0.0f : ref;
            j.frequencyHz = hz;
            j.dampingRatio = ratio;
            s.joints.push_back(j);
            prev = b;
        }
        if (!drop) continue;
        addBody(s, defaultBody(2, x + (float)links - 0.5f, 13.0f));
        addFixture(s, boxFixture(0.3f, 0.3f, 1.0f, 0.6f));
    }
    return s;
}

// Friction joints: n boxes (0.5 m, 1 kg) tied to the ground body by
// friction joints at their centres (maxForce 2-6 N, maxTorque 0.5-1.5 N m),
// launched with speeds of 2-6 m/s and spins, over a ground edge 2 m below;
// pairs of neighbours are thrown at each other. Run without gravity it is a
// top-down world.
inline Scene frictionJoints(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        BodyDef bd = defaultBody(2, 4.0f * (float)i - 2.0f * (float)n, 2.5f);
        bd.vx = ((i & 1) ? -1.0f : 1.0f) * (2.0f + (float)(i % 5));
        bd.vy = 0.5f * (float)(i % 3);
        bd.w = 1.0f + (float)(i % 4);
        addBody(s, bd);
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.4f));
        JointDef j{};
        j.type = 4;
        j.bodyA = 0; j.bodyB = (int)s.bodies.size() - 1;
        j.ax = bd.px; j.ay = bd.py;   // the box centre, on the ground body and the box
        j.maxForce = 2.0f + (float)(i % 5);
        j.maxTorque = 0.5f + 0.5f * (float)(i % 3);
        s.joints.push_back(j);
    }
    return s;
}

// Motor joints: n boxes driven by motor joints from the ground body towards
// targets 1.5-3 m away (linear offsets), turned by 0.5-1.5 rad (angular
// offsets), with maxForce 20-40 N (above the 10 N weight), maxTorque 5-15
// N m and correction factors 0.1-0.5; when `chained`, every third box is
// driven from the previous box instead of the ground.
inline Scene motorJoints(int n, bool chained = true) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 6.0f * (float)i - 3.0f * (float)n, y = 4.0f;
        addBody(s, defaultBody(2, x, y));
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.4f));
        const int b = (int)s.bodies.size() - 1;
        JointDef j{};
        j.type = 5;
        const bool fromPrev = chained && (i % 3) == 2;
        j.bodyA = fromPrev ? b - 1 : 0;
        j.bodyB = b;
        const float dx = 1.5f * (float)(1 + i % 2), dy = 1.0f + 0.5f * (float)(i % 3);
        j.ax = fromPrev ? 6.0f : x + dx;          // offset in body A's frame
        j.ay = fromPrev ? 0.5f : y + dy;
        j.referenceAngle = 0.5f + 0.5f * (float)(i % 3);
        j.maxForce = 20.0f + 10.0f * (float)(i % 3);
        j.maxTorque = 5.0f + 5.0f * (float)(i % 3);
        j.correctionFactor = 0.1f + 0.2f * (float)(i % 3);
        s.joints.push_back(j);
    }
    return s;
}

// Pulley joints: n pulleys, each a box A (1 kg) and a box B (0.64-1.69 kg)
// hanging 10 m and 8 m below ground anchors 4 m apart at y = 20, with
// ratios 1, 1.5 and 2; odd pulleys hold their boxes by a top corner (the
// boxes turn), even ones by their centres, over a ground edge.
inline Scene pulleys(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 10.0f * (float)i - 5.0f * (float)n;
        const float hb = 0.4f + 0.15f * (float)(i % 4);
        addBody(s, defaultBody(2, x - 2.0f, 10.0f));
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.4f));
        addBody(s, defaultBody(2, x + 2.0f, 12.0f));
        addFixture(s, boxFixture(hb, hb, 1.0f, 0.4f));
        const int b = (int)s.bodies.size() - 1;
        JointDef j{};
        j.type = 6;
        j.bodyA = b - 1; j.bodyB = b;
        const bool corner = (i & 1) != 0;
        j.ax = corner ? 0.5f : 0.0f; j.ay = corner ? 0.5f : 0.0f;
        j.bx = corner ? -hb : 0.0f; j.by = corner ? hb : 0.0f;
        j.gax = x - 2.0f; j.gay = 20.0f;
        j.gbx = x + 2.0f; j.gby = 20.0f;
        const float dax = x - 2.0f + j.ax - j.gax, day = 10.0f + j.ay - j.gay;
        const float dbx = x + 2.0f + j.bx - j.gbx, dby = 12.0f + j.by - j.gby;
        j.length = sqrtf(dax * dax + day * day);
        j.lengthB = sqrtf(dbx * dbx + dby * dby);
        j.ratio = 1.0f + 0.5f * (float)(i % 3);
        s.joints.push_back(j);
    }
    return s;
}

// Mouse joints: n boxes (1 kg) on the ground edge, each grabbed near a
// corner by a mouse joint (body A the ground; created with its target on the
// grab point (bx, by) of body B, as the testbed's mouse-down does) and then
// pulled by SetTarget to (ax, ay), 3-4 m above; maxForce 1000 N, 2-6 Hz,
// damping 0.7.
inline Scene mouseJoints(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 4.0f * (float)i - 2.0f * (float)n;
        addBody(s, defaultBody(2, x, 0.5f));
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.6f));
        JointDef j{};
        j.type = 7;
        j.bodyA = 0; j.bodyB = (int)s.bodies.size() - 1;
        j.ax = x + 0.4f; j.ay = 3.5f + 0.5f * (float)(i % 3);   // target
        j.maxForce = 1000.0f;
        j.frequencyHz = 2.0f + 2.0f * (float)(i % 3);
        j.dampingRatio = 0.7f;
        j.bx = 0.4f; j.by = 0.4f;   // grab point on body B
        s.joints.push_back(j);
    }
    return s;
}

// Prismatic joints: n boxes (1 kg) sliding on prismatic joints from the
// ground body along axes at 0-75 degrees (some above 45, so they slide down
// under gravity), started at rest or moving; every third has a translation
// limit of [-1, 2] m, every fourth a motor (2 m/s, 15-30 N), one an equal
// limit, over a ground edge 6 m below.
inline Scene prismatics(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 8.0f * (float)i - 4.0f * (float)n, y = 8.0f;
        BodyDef bd = defaultBody(2, x, y);
        bd.vx = 0.5f * (float)(i % 3);
        bd.w = 0.3f * (float)(i % 2);
        addBody(s, bd);
        addFixture(s, boxFixture(0.5f, 0.5f, 1.0f, 0.4f));
        JointDef j{};
        j.type = 8;
        j.bodyA = 0; j.bodyB = (int)s.bodies.size() - 1;
        j.ax = x; j.ay = y;
        const float ang = 0.2618f * (float)(i % 6);
        j.axisX = 2.0f * cosf(ang); j.axisY = 2.0f * sinf(ang);   // normalized on creation
        j.enableLimit = (i % 3) == 0;
        j.lowerAngle = -1.0f; j.upperAngle = (i == 6) ? -1.0f : 2.0f;
        j.enableMotor = (i % 4) == 1;
        j.motorSpeed = 2.0f;
        j.maxMotorTorque = 15.0f + 15.0f * (float)(i % 2);
        s.joints.push_back(j);
    }
    return s;
}

// Wheel joints: n carts, each a chassis (2 x 0.5 m box) on two circle wheels
// (r 0.4) held by wheel joints along the vertical axis (4 Hz, damping 0.7);
// the rear wheel of every other cart is driven (-10 rad/s, 20 N m) over a
// ground edge with bumps. `single` gives each chassis one wheel on a flat
// ground (one joint and one contact per wheel).
inline Scene wheels(int n, bool single = false) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = -10; i < 10 && !single; ++i) {   // bumps
        FixtureDef f = boxFixture(0.3f, 0.1f, 0.0f, 0.6f);
        f.polyMode = kPolyBoxRotated;
        f.pts[2] = 6.0f * (float)i + 1.0f; f.pts[3] = 0.05f; f.pts[4] = 0.0f;
        addFixture(s, f);
    }
    for (int i = 0; i < n; ++i) {
        const float x = 12.0f * (float)i - 6.0f * (float)n, y = 1.4f;
        addBody(s, defaultBody(2, x, y));
        addFixture(s, boxFixture(1.0f, 0.25f, 1.0f, 0.6f));
        const int chassis = (int)s.bodies.size() - 1;
        for (int k = 0; k < (single ? 1 : 2); ++k) {
            const float wx = x + (k == 0 ? -0.7f : 0.7f), wy = 0.6f;
            addBody(s, defaultBody(2, wx, wy));
            FixtureDef wf = defaultFixture(kCircle);
            wf.radius = 0.4f;
            wf.density = 1.0f; wf.friction = 0.9f;
            addFixture(s, wf);
            JointDef j{};
            j.type = 9;
            j.bodyA = chassis; j.bodyB = (int)s.bodies.size() - 1;
            j.ax = wx - x; j.ay = wy - y;
            j.axisX = 0.0f; j.axisY = 1.0f;
            j.frequencyHz = 4.0f;
            j.dampingRatio = 0.7f;
            j.enableMotor = (k == 0 && (i & 1) == 0);
            j.motorSpeed = -10.0f;
            j.maxMotorTorque = 20.0f;
            s.joints.push_back(j);
        }
    }
    return s;
}

// Gear joints: n gear trains (as the reference testbed's gears): a disc of
// radius 1 and one of radius 2 on revolute joints to the ground, geared at
// ratio 1/2, and a rack (0.5 x 3 m box) on a vertical prismatic joint
// (limits +-3 m) geared to the large disc at ratio -1/2; the rack's weight
// drives the train. Odd trains brake the small disc with a motor.
inline Scene gears(int n) {
    Scene s;
    addBody(s, defaultBody(0, 0.0f, 0.0f));
    addFixture(s, edgeFixture(-400.0f, 0.0f, 400.0f, 0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = 14.0f * (float)i - 7.0f * (float)n, y = 8.0f;
        const int j0 = (int)s.joints.size();
        const float cx[2] = {x, x + 3.5f}, rad[2] = {1.0f, 2.0f};
        for (int k = 0; k < 2; ++k) {
            addBody(s, defaultBody(2, cx[k], y));
            FixtureDef f = defaultFixture(kCircle);
            f.radius = rad[k];
            f.density = 5.0f;
            addFixture(s, f);
            JointDef r{};
            r.bodyA = 0; r.bodyB = (int)s.bodies.size() - 1;
            r.ax = cx[k]; r.ay = y;
            r.enableMotor = (k == 0 && (i & 1)) ? 1 : 0;
            r.motorSpeed = 0.0f;
            r.maxMotorTorque = 20.0f;
            s.joints.push_back(r);
        }
        const int disc1 = (int)s.bodies.size() - 2, disc2 = disc1 + 1;
        JointDef g{};
        g.type = 10;
        g.bodyA = disc1; g.bodyB = disc2;
        g.joint1 = j0; g.joint2 = j0 + 1;
        g.ratio = 1.0f / 2.0f;
        s.joints.push_back(g);
        addBody(s, defaultBody(2, x + 6.0f, y));
        addFixture(s, boxFixture(0.5f, 3.0f, 1.0f, 0.4f));
        JointDef p{};
        p.type = 8;
        p.bodyA = 0; p.bodyB = (int)s.bodies.size() - 1;
        p.ax = x + 6.0f; p.ay = y;
        p.axisX = 0.0f; p.axisY = 1.0f;
        p.enableLimit = 1;
        p.lowerAngle = -3.0f; p.upperAngle = 3.0f;
        s.joints.push_back(p);
        JointDef g2{};
        g2.type = 10;
        g2.bodyA = disc2; g2.bodyB = p.bodyB;
        g2.joint1 = j0 + 1; g2.joint2 = (int)s.joints.size() - 1;
        g2.ratio = -1.0f / 2.0f;
        s.joints.push_back(g2);
    }
    return s;
}

// Scene by name: "hello", "pyramid" (n = rows, 0 -> 20), "pile" (n bodies,
// seed), "world" (one batched-scene world of n bodies, seed), "pendulums"
// (n pendulums), "chains" (n chains of `seed` links), "stress" (pile of n
// bodies, seed, on a mat of n / 50 chains of 32 links), "plank" (n circles
// on a plank), "oneway" (n boxes through a one-sided platform), "bullets"
// (n fast bodies, seed), "terrain" (n bodies on chain shapes, seed),
// "distance" (n rods and springs with boxes; seed 0 hangs a second box on
// each), "spring" (one box on a spring of n / 10 Hz and damping ratio
// seed / 10), "ropes" (n boxes on ropes; seed 0 hangs a second box on each),
// "welds" (n welded beams; seed 0 drops a box on each, seed 2 makes the
// beams single links), "friction" (n boxes on friction joints), "motor" (n
// boxes on motor joints; seed 0 drives every third from the previous box),
// "pulley" (n pulleys), "mouse" (n boxes on mouse joints), "prismatic" (n
// boxes on prismatic joints), "wheels" (n carts; seed 1 gives each one
// wheel on flat ground), "gears" (n gear trains), "sensors" (n bodies through
// a sensor slab, seed), "conveyor" (n boxes on a belt). Unknown names give an
// empty scene.
inline Scene byName(const char* kind, int n, uint32_t seed) {
    const std::string k(kind ? kind : "");
    if (k == "hello") return hello();
    if (k == "pyramid") return pyramid(n > 0 ? n : 20);
    if (k == "pile") return pile(n, seed);
    if (k == "world") return batchedWorld(n, seed);
    if (k == "rain") return rain(n > 0 ? n : 20, seed);
    if (k == "pendulums") return pendulums(n > 0 ? n : 16);
    if (k == "chains") return chains(n > 0 ? n : 4, seed > 0 ? (int)seed : 32);
    if (k == "stress") return stress(n, seed, n / 50);
    if (k == "plank") return plank(n > 0 ? n : 100);
    if (k == "oneway") return oneway(n > 0 ? n : 8);
    if (k == "bullets") return bullets(n > 0 ? n : 24, seed);
    if (k == "terrain") return terrain(n > 0 ? n : 60, seed);
    if (k == "distance") return distanceJoints(n > 0 ? n : 16, seed == 0);
    if (k == "spring") return spring(0.1f * (float)(n > 0 ? n : 10), 0.1f * (float)seed);
    if (k == "ropes") return ropes(n > 0 ? n : 16, seed == 0);
    if (k == "welds") return welds(n > 0 ? n : 12, seed == 0, seed == 2);
    if (k == "friction") return frictionJoints(n > 0 ? n : 16);
    if (k == "motor") return motorJoints(n > 0 ? n : 12, seed == 0);
    if (k == "pulley") return pulleys(n > 0 ? n : 8);
    if (k == "mouse") return mouseJoints(n > 0 ? n : 8);
    if (k == "prismatic") return prismatics(n > 0 ? n : 12);
    if (k == "wheels") return wheels(n > 0 ? n : 6, seed == 1);
    if (k == "gears") return gears(n > 0 ? n : 4);
    if (k == "sensors") return sensors(n > 0 ? n : 30, seed);
    if (k == "conveyor") return conveyor(n > 0 ? n : 6);
    return Scene();
}

}  // namespace b2scene
