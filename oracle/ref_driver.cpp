// ORACLE TEST INFRASTRUCTURE - not part of the product.
//
// C-ABI driver over the upstream Box2D 2.3.2 sources that the reference
// vendors (Regression/box2d-ref/Box2DRef, namespace box2dref). oracle/Makefile
// compiles those sources where they lie, together with this file, into
// oracle/_ref/libb2ref.so. Tests use it to pin the CPU restatement
// (oracle/cpu_ref.cpp), to write golden fixtures, and bench.py times it as the
// CPU baseline ("kind": "reference"). Only tests/, the smoke check and the
// benchmark's CPU-baseline leg may load it.
//
// Every entry point mirrors one exported by the restatement with the prefix
// ref_ instead of orc_, so tests treat both checkers alike.
// The driver reads a few private members (proxy ids, sleep time) for dumps;
// it opens class access for the reference headers only.
#define private public
#define protected public
#include <Box2DRef/Box2D.h>
#undef private
#undef protected
#include <stdint.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include <unordered_map>

#include "b2g.h"
#include "scenes/scenes.h"

using namespace box2dref;

namespace {

// Records the listener calls (kind 1 BeginContact, 2 EndContact, 3 PreSolve,
// 4 PostSolve; fixture indices; for PreSolve whether the contact stays
// enabled, for PostSolve the bits of the first normal impulse) and disables the contacts of a one-sided platform fixture while
// the other body's origin is less than `clear` above the platform's top (a
// one-sided platform after the testbed's OneSidedPlatform,
// Testbed/Tests/OneSidedPlatform.h).
// With a conveyor fixture set, PreSolve also gives its contacts the belt's
// tangent speed (b2Contact::SetTangentSpeed).
struct Recorder : public b2ContactListener {
    std::vector<b2Fixture*>* fixtures = nullptr;
    int platform = -1;
    float top = 0.0f, clear = 0.0f;
    int conveyor = -1;  // fixture whose contacts get the belt speed
    float belt = 0.0f;
    int step = 0;
    std::vector<int32_t> events;  // step, kind, fixture A, fixture B, flag
    int indexOf(const b2Fixture* f) const {
        for (size_t i = 0; i < fixtures->size(); ++i)
            if ((*fixtures)[i] == f) return (int)i;
        return -1;
    }
    void record(int kind, b2Contact* c, int flag) {
        events.insert(events.end(),
                      {step, kind, indexOf(c->GetFixtureA()), indexOf(c->GetFixtureB()), flag});
    }
    void BeginContact(b2Contact* c) override { record(1, c, 0); }
    void EndContact(b2Contact* c) override { record(2, c, 0); }
    void PreSolve(b2Contact* c, const b2Manifold*) override {
        const int a = indexOf(c->GetFixtureA()), b = indexOf(c->GetFixtureB());
        if (a == platform || b == platform) {
            const b2Body* other = (a == platform ? c->GetFixtureB() : c->GetFixtureA())->GetBody();
            if (other->GetPosition().y < top + clear) c->SetEnabled(false);
        }
        // a conveyor belt after the testbed's ConveyorBelt: the tangent speed
        // of the belt's contacts, its sign by the belt's side of the contact
        if (a == conveyor) c->SetTangentSpeed(belt);
        if (b == conveyor) c->SetTangentSpeed(-belt);
        record(3, c, c->IsEnabled() ? 1 : 0);
    }
    // PostSolve's flag: the bits of the first point's normal impulse
    void PostSolve(b2Contact* c, const b2ContactImpulse* imp) override {
        int32_t bits = 0;
        if (imp->count > 0) memcpy(&bits, &imp->normalImpulses[0], sizeof(bits));
        record(4, c, bits);
    }
};

struct RefWorld {
    b2World* world;
    std::vector<b2Body*> bodies;
    std::vector<b2Fixture*> fixtures;
    std::vector<b2Joint*> joints;
    Recorder recorder;
};

// Scene joints as b2RevoluteJoints (local anchors and reference angle given
// directly, as b2RevoluteJointDef::Initialize computes them for bodies at
// angle zero).
int loadJoints(RefWorld* w, int base, const b2scene::JointDef* jd, int nj) {
    const int jbase = (int)w->joints.size();
    for (int k = 0; k < nj; ++k) {
        const b2scene::JointDef& d = jd[k];
        if (d.type == 10) {
            b2GearJointDef gd;
            gd.bodyA = w->bodies[base + d.bodyA];
            gd.bodyB = w->bodies[base + d.bodyB];
            gd.collideConnected = d.collideConnected != 0;
            gd.joint1 = w->joints[jbase + d.joint1];
            gd.joint2 = w->joints[jbase + d.joint2];
            gd.ratio = d.ratio;
            w->joints.push_back(w->world->CreateJoint(&gd));
            continue;
        }
        if (d.type == 1) {
            b2DistanceJointDef dd;
            dd.bodyA = w->bodies[base + d.bodyA];
            dd.bodyB = w->bodies[base + d.bodyB];
            dd.collideConnected = d.collideConnected != 0;
            dd.localAnchorA.Set(d.ax, d.ay);
            dd.localAnchorB.Set(d.bx, d.by);
            dd.length = d.length;
            dd.frequencyHz = d.frequencyHz;
            dd.dampingRatio = d.dampingRatio;
            w->joints.push_back(w->world->CreateJoint(&dd));
            continue;
        }
        if (d.type == 2) {
            b2RopeJointDef rd;
            rd.bodyA = w->bodies[base + d.bodyA];
            rd.bodyB = w->bodies[base + d.bodyB];
            rd.collideConnected = d.collideConnected != 0;
            rd.localAnchorA.Set(d.ax, d.ay);
            rd.localAnchorB.Set(d.bx, d.by);
            rd.maxLength = d.length;
            w->joints.push_back(w->world->CreateJoint(&rd));
            continue;
        }
        if (d.type == 8) {
            b2PrismaticJointDef pd;
            pd.bodyA = w->bodies[base + d.bodyA];
            pd.bodyB = w->bodies[base + d.bodyB];
            pd.collideConnected = d.collideConnected != 0;
            pd.localAnchorA.Set(d.ax, d.ay);
            pd.localAnchorB.Set(d.bx, d.by);
            pd.localAxisA.Set(d.axisX, d.axisY);
            pd.referenceAngle = d.referenceAngle;
            pd.enableLimit = d.enableLimit != 0;
            pd.lowerTranslation = d.lowerAngle;
            pd.upperTranslation = d.upperAngle;
            pd.enableMotor = d.enableMotor != 0;
            pd.maxMotorForce = d.maxMotorTorque;
            pd.motorSpeed = d.motorSpeed;
            w->joints.push_back(w->world->CreateJoint(&pd));
            continue;
        }
        if (d.type == 9) {
            b2WheelJointDef wd;
            wd.bodyA = w->bodies[base + d.bodyA];
            wd.bodyB = w->bodies[base + d.bodyB];
            wd.collideConnected = d.collideConnected != 0;
            wd.localAnchorA.Set(d.ax, d.ay);
            wd.localAnchorB.Set(d.bx, d.by);
            wd.localAxisA.Set(d.axisX, d.axisY);
            wd.enableMotor = d.enableMotor != 0;
            wd.maxMotorTorque = d.maxMotorTorque;
            wd.motorSpeed = d.motorSpeed;
            wd.frequencyHz = d.frequencyHz;
            wd.dampingRatio = d.dampingRatio;
            w->joints.push_back(w->world->CreateJoint(&wd));
            continue;
        }
        if (d.type == 6) {
            b2PulleyJointDef pd;
            pd.bodyA = w->bodies[base + d.bodyA];
            pd.bodyB = w->bodies[base + d.bodyB];
            pd.collideConnected = d.collideConnected != 0;
            pd.groundAnchorA.Set(d.gax, d.gay);
            pd.groundAnchorB.Set(d.gbx, d.gby);
            pd.localAnchorA.Set(d.ax, d.ay);
            pd.localAnchorB.Set(d.bx, d.by);
            pd.lengthA = d.length;
            pd.lengthB = d.lengthB;
            pd.ratio = d.ratio;
            w->joints.push_back(w->world->CreateJoint(&pd));
            continue;
        }
        if (d.type == 7) {
            // created on the grab point, then pulled to the target
            b2MouseJointDef md;
            md.bodyA = w->bodies[base + d.bodyA];
            md.bodyB = w->bodies[base + d.bodyB];
            md.collideConnected = d.collideConnected != 0;
            md.target = md.bodyB->GetWorldPoint(b2Vec2(d.bx, d.by));
            md.maxForce = d.maxForce;
            md.frequencyHz = d.frequencyHz;
            md.dampingRatio = d.dampingRatio;
            b2MouseJoint* j = static_cast<b2MouseJoint*>(w->world->CreateJoint(&md));
            j->SetTarget(b2Vec2(d.ax, d.ay));
            w->joints.push_back(j);
            continue;
        }
        if (d.type == 4) {
            b2FrictionJointDef fd;
            fd.bodyA = w->bodies[base + d.bodyA];
            fd.bodyB = w->bodies[base + d.bodyB];
            fd.collideConnected = d.collideConnected != 0;
            fd.localAnchorA.Set(d.ax, d.ay);
            fd.localAnchorB.Set(d.bx, d.by);
            fd.maxForce = d.maxForce;
            fd.maxTorque = d.maxTorque;
            w->joints.push_back(w->world->CreateJoint(&fd));
            continue;
        }
        if (d.type == 5) {
            b2MotorJointDef md;
            md.bodyA = w->bodies[base + d.bodyA];
            md.bodyB = w->bodies[base + d.bodyB];
            md.collideConnected = d.collideConnected != 0;
            md.linearOffset.Set(d.ax, d.ay);
            md.angularOffset = d.referenceAngle;
            md.maxForce = d.maxForce;
            md.maxTorque = d.maxTorque;
            md.correctionFactor = d.correctionFactor;
            w->joints.push_back(w->world->CreateJoint(&md));
            continue;
        }
        if (d.type == 3) {
            b2WeldJointDef wd;
            wd.bodyA = w->bodies[base + d.bodyA];
            wd.bodyB = w->bodies[base + d.bodyB];
            wd.collideConnected = d.collideConnected != 0;
            wd.localAnchorA.Set(d.ax, d.ay);
            wd.localAnchorB.Set(d.bx, d.by);
            wd.referenceAngle = d.referenceAngle;
            wd.frequencyHz = d.frequencyHz;
            wd.dampingRatio = d.dampingRatio;
            w->joints.push_back(w->world->CreateJoint(&wd));
            continue;
        }
        b2RevoluteJointDef def;
        def.bodyA = w->bodies[base + d.bodyA];
        def.bodyB = w->bodies[base + d.bodyB];
        def.collideConnected = d.collideConnected != 0;
        def.localAnchorA.Set(d.ax, d.ay);
        def.localAnchorB.Set(d.bx, d.by);
        def.referenceAngle = d.referenceAngle;
        def.enableLimit = d.enableLimit != 0;
        def.lowerAngle = d.lowerAngle;
        def.upperAngle = d.upperAngle;
        def.enableMotor = d.enableMotor != 0;
        def.motorSpeed = d.motorSpeed;
        def.maxMotorTorque = d.maxMotorTorque;
        w->joints.push_back(w->world->CreateJoint(&def));
    }
    return 0;
}

int fixtureIndex(const b2Fixture* f) { return (int)(intptr_t)f->GetUserData(); }

void shapeToRec(const b2Shape* s, b2g_shape* out) {
    memset(out, 0, sizeof(*out));
    out->type = (int32_t)s->GetType();
    out->radius = s->m_radius;
    if (s->GetType() == b2Shape::e_polygon) {
        const b2PolygonShape* p = (const b2PolygonShape*)s;
        out->count = p->m_count;
        out->centroid = {p->m_centroid.x, p->m_centroid.y};
        for (int i = 0; i < p->m_count; ++i) {
            out->verts[i] = {p->m_vertices[i].x, p->m_vertices[i].y};
            out->normals[i] = {p->m_normals[i].x, p->m_normals[i].y};
        }
    } else if (s->GetType() == b2Shape::e_circle) {
        const b2CircleShape* c = (const b2CircleShape*)s;
        out->centroid = {c->m_p.x, c->m_p.y};
    } else if (s->GetType() == b2Shape::e_edge) {
        const b2EdgeShape* e = (const b2EdgeShape*)s;
        out->verts[0] = {e->m_vertex1.x, e->m_vertex1.y};
        out->verts[1] = {e->m_vertex2.x, e->m_vertex2.y};
        out->v0 = {e->m_vertex0.x, e->m_vertex0.y};
        out->v3 = {e->m_vertex3.x, e->m_vertex3.y};
        out->flags = (e->m_hasVertex0 ? 1 : 0) | (e->m_hasVertex3 ? 2 : 0);
        out->count = 2;
    }
}

void recToPolygon(const b2g_shape* r, b2PolygonShape* p) {
    p->m_count = r->count;
    p->m_radius = r->radius;
    p->m_centroid.Set(r->centroid.x, r->centroid.y);
    for (int i = 0; i < r->count; ++i) {
        p->m_vertices[i].Set(r->verts[i].x, r->verts[i].y);
        p->m_normals[i].Set(r->normals[i].x, r->normals[i].y);
    }
}

void recToCircle(const b2g_shape* r, b2CircleShape* c) {
    c->m_radius = r->radius;
    c->m_p.Set(r->centroid.x, r->centroid.y);
}

void recToEdge(const b2g_shape* r, b2EdgeShape* e) {
    e->m_radius = r->radius;
    e->m_vertex1.Set(r->verts[0].x, r->verts[0].y);
    e->m_vertex2.Set(r->verts[1].x, r->verts[1].y);
    e->m_vertex0.Set(r->v0.x, r->v0.y);
    e->m_vertex3.Set(r->v3.x, r->v3.y);
    e->m_hasVertex0 = (r->flags & 1) != 0;
    e->m_hasVertex3 = (r->flags & 2) != 0;
}

b2Transform toXf(const b2g_transform* t) {
    b2Transform xf;
    xf.p.Set(t->px, t->py);
    xf.q.s = t->s;
    xf.q.c = t->c;
    return xf;
}

void manifoldOut(const b2Manifold& m, b2g_manifold* out) {
    memset(out, 0, sizeof(*out));
    out->pointCount = m.pointCount;
    if (m.pointCount == 0) return;  // other fields are stale for empty manifolds
    out->type = (int32_t)m.type;
    out->localNormal = {m.localNormal.x, m.localNormal.y};
    out->localPoint = {m.localPoint.x, m.localPoint.y};
    for (int i = 0; i < m.pointCount; ++i) {
        out->points[i].localPoint = {m.points[i].localPoint.x, m.points[i].localPoint.y};
        out->points[i].normalImpulse = m.points[i].normalImpulse;
        out->points[i].tangentImpulse = m.points[i].tangentImpulse;
        out->points[i].id = m.points[i].id.key;
    }
}

// Builds one fixture's shape from a scene definition through the public API.
void buildShape(const b2scene::FixtureDef& fd, b2PolygonShape* poly, b2CircleShape* circle,
                b2EdgeShape* edge, b2ChainShape* chain, const b2Shape** shape) {
    if (fd.shape == b2scene::kChain) {
        b2Vec2 pts[8];
        for (int i = 0; i < fd.n && i < 8; ++i) pts[i].Set(fd.pts[2 * i], fd.pts[2 * i + 1]);
        if (fd.polyMode == 1) {
            chain->CreateLoop(pts, fd.n);
        } else {
            chain->CreateChain(pts, fd.n);
            if (fd.has0) chain->SetPrevVertex(b2Vec2(fd.e[0], fd.e[1]));
            if (fd.has3) chain->SetNextVertex(b2Vec2(fd.e[6], fd.e[7]));
        }
        *shape = chain;
    } else if (fd.shape == b2scene::kPolygon) {
        if (fd.polyMode == b2scene::kPolyBox) {
            poly->SetAsBox(fd.pts[0], fd.pts[1]);
        } else if (fd.polyMode == b2scene::kPolyBoxRotated) {
            poly->SetAsBox(fd.pts[0], fd.pts[1], b2Vec2(fd.pts[2], fd.pts[3]), fd.pts[4]);
        } else {
            b2Vec2 pts[b2_maxPolygonVertices];
            for (int i = 0; i < fd.n; ++i) pts[i].Set(fd.pts[2 * i], fd.pts[2 * i + 1]);
            poly->Set(pts, fd.n);
        }
        *shape = poly;
    } else if (fd.shape == b2scene::kCircle) {
        circle->m_radius = fd.radius;
        circle->m_p.Set(fd.cx, fd.cy);
        *shape = circle;
    } else {
        edge->Set(b2Vec2(fd.e[2], fd.e[3]), b2Vec2(fd.e[4], fd.e[5]));
        edge->m_vertex0.Set(fd.e[0], fd.e[1]);
        edge->m_vertex3.Set(fd.e[6], fd.e[7]);
        edge->m_hasVertex0 = fd.has0 != 0;
        edge->m_hasVertex3 = fd.has3 != 0;
        *shape = edge;
    }
}

struct PairCollector {
    std::vector<int32_t> pairs;
    void AddPair(void* a, void* b) {
        pairs.push_back((int32_t)(intptr_t)a - 1);
        pairs.push_back((int32_t)(intptr_t)b - 1);
    }
};

}  // namespace

extern "C" {

void* ref_world_create(float gx, float gy) {
    RefWorld* w = new RefWorld;
    w->world = new b2World(b2Vec2(gx, gy));
    return w;
}

void ref_world_destroy(void* h) {
    RefWorld* w = (RefWorld*)h;
    delete w->world;
    delete w;
}

void ref_world_set_flags(void* h, int allowSleep, int warmStarting, int continuous, int subStepping) {
    RefWorld* w = (RefWorld*)h;
    w->world->SetAllowSleeping(allowSleep != 0);
    w->world->SetWarmStarting(warmStarting != 0);
    w->world->SetContinuousPhysics(continuous != 0);
    w->world->SetSubStepping(subStepping != 0);
}

int ref_world_load(void* h, const b2scene::BodyDef* bodies, int nb, const b2scene::FixtureDef* fx, int nf) {
    RefWorld* w = (RefWorld*)h;
    for (int i = 0; i < nb; ++i) {
        const b2scene::BodyDef& bd = bodies[i];
        b2BodyDef def;
        def.type = (b2BodyType)bd.type;
        def.position.Set(bd.px, bd.py);
        def.angle = bd.angle;
        def.linearVelocity.Set(bd.vx, bd.vy);
        def.angularVelocity = bd.w;
        def.linearDamping = bd.linearDamping;
        def.angularDamping = bd.angularDamping;
        def.gravityScale = bd.gravityScale;
        def.allowSleep = bd.allowSleep != 0;
        def.awake = bd.awake != 0;
        def.fixedRotation = bd.fixedRotation != 0;
        def.bullet = bd.bullet != 0;
        b2Body* body = w->world->CreateBody(&def);
        w->bodies.push_back(body);
        for (int k = 0; k < bd.fixtureCount; ++k) {
            const int fi = bd.firstFixture + k;
            if (fi < 0 || fi >= nf) return -1;
            const b2scene::FixtureDef& fd = fx[fi];
            b2PolygonShape poly;
            b2CircleShape circle;
            b2EdgeShape edge;
            b2ChainShape chain;
            const b2Shape* shape = nullptr;
            buildShape(fd, &poly, &circle, &edge, &chain, &shape);
            b2FixtureDef fdef;
            fdef.shape = shape;
            fdef.density = fd.density;
            fdef.friction = fd.friction;
            fdef.restitution = fd.restitution;
            fdef.isSensor = fd.sensor != 0;
            fdef.filter.categoryBits = (uint16)fd.categoryBits;
            fdef.filter.maskBits = (uint16)fd.maskBits;
            fdef.filter.groupIndex = (int16)fd.groupIndex;
            fdef.userData = (void*)(intptr_t)w->fixtures.size();
            b2Fixture* fixture = body->CreateFixture(&fdef);
            // one entry per proxy (a chain has one per child edge)
            for (int c = 0; c < fixture->GetShape()->GetChildCount(); ++c) w->fixtures.push_back(fixture);
        }
    }
    return 0;
}

int ref_world_load_scene(void* h, const char* kind, int n, unsigned seed) {
    b2scene::Scene s = b2scene::byName(kind, n, seed);
    RefWorld* w = (RefWorld*)h;
    const int base = (int)w->bodies.size();
    int rc = ref_world_load(h, s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size());
    if (rc != 0) return rc;
    return loadJoints(w, base, s.joints.data(), (int)s.joints.size());
}

// A named scene with every body started from a given state: rows of 12
// floats as *_world_get_bodies writes them (origin x, y at 0-1, angle at 6,
// linear velocity at 7-8, angular velocity at 9), one per body in creation
// order. Contacts are created by the first Step, without carried impulses.
// Used to run the reference from another engine's state (the benchmark's CPU
// leg resumes it where the device run stands).
int ref_world_load_scene_state(void* h, const char* kind, int n, unsigned seed, const float* state, int nstate) {
    b2scene::Scene s = b2scene::byName(kind, n, seed);
    if (nstate != (int)s.bodies.size()) return -1;
    for (int i = 0; i < nstate; ++i) {
        const float* o = state + 12 * i;
        b2scene::BodyDef& b = s.bodies[i];
        b.px = o[0]; b.py = o[1]; b.angle = o[6];
        b.vx = o[7]; b.vy = o[8]; b.w = o[9];
    }
    RefWorld* w = (RefWorld*)h;
    const int base = (int)w->bodies.size();
    int rc = ref_world_load(h, s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size());
    if (rc != 0) return rc;
    return loadJoints(w, base, s.joints.data(), (int)s.joints.size());
}

int ref_world_joint_count(void* h) { return (int)((RefWorld*)h)->joints.size(); }

// Installs the recording listener with the one-sided platform rule for
// fixture `platform` (its top and the clearance as given); -1: record only.
void ref_world_set_recorder(void* h, int platform, float top, float clear) {
    RefWorld* w = (RefWorld*)h;
    w->recorder.fixtures = &w->fixtures;
    w->recorder.platform = platform;
    w->recorder.top = top;
    w->recorder.clear = clear;
    w->world->SetContactListener(&w->recorder);
}

// Installs the recording listener (record only, unless set before) with
// `fixture` as a conveyor belt of tangent speed `speed`.
void ref_world_set_conveyor(void* h, int fixture, float speed) {
    RefWorld* w = (RefWorld*)h;
    w->recorder.fixtures = &w->fixtures;
    w->recorder.conveyor = fixture;
    w->recorder.belt = speed;
    w->world->SetContactListener(&w->recorder);
}

// Recorded events, 5 ints each; returns the event count.
int ref_world_events(void* h, int32_t* out, int cap) {
    RefWorld* w = (RefWorld*)h;
    const int n = (int)w->recorder.events.size() / 5;
    for (int i = 0; i < n && i < cap; ++i) memcpy(out + 5 * i, &w->recorder.events[5 * i], 5 * sizeof(int32_t));
    return n;
}

int ref_world_set_joint(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor, float speed,
                        float maxTorque) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_revoluteJoint) return -1;
    b2RevoluteJoint* j = static_cast<b2RevoluteJoint*>(w->joints[joint]);
    j->EnableLimit(enableLimit != 0);
    j->SetLimits(lower, upper);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorTorque(maxTorque);
    return 0;
}

int ref_world_set_distance(void* h, int joint, float length, float hz, float ratio) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_distanceJoint) return -1;
    b2DistanceJoint* j = static_cast<b2DistanceJoint*>(w->joints[joint]);
    j->SetLength(length);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

int ref_world_set_prismatic(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor,
                            float speed, float maxForce) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_prismaticJoint) return -1;
    b2PrismaticJoint* j = static_cast<b2PrismaticJoint*>(w->joints[joint]);
    j->EnableLimit(enableLimit != 0);
    j->SetLimits(lower, upper);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorForce(maxForce);
    return 0;
}

int ref_world_set_wheel(void* h, int joint, int enableMotor, float speed, float maxTorque, float hz, float ratio) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_wheelJoint) return -1;
    b2WheelJoint* j = static_cast<b2WheelJoint*>(w->joints[joint]);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorTorque(maxTorque);
    j->SetSpringFrequencyHz(hz);
    j->SetSpringDampingRatio(ratio);
    return 0;
}

int ref_world_set_mouse(void* h, int joint, float tx, float ty, float maxForce, float hz, float ratio) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_mouseJoint) return -1;
    b2MouseJoint* j = static_cast<b2MouseJoint*>(w->joints[joint]);
    j->SetTarget(b2Vec2(tx, ty));
    j->SetMaxForce(maxForce);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

int ref_world_set_friction(void* h, int joint, float maxForce, float maxTorque) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_frictionJoint) return -1;
    b2FrictionJoint* j = static_cast<b2FrictionJoint*>(w->joints[joint]);
    j->SetMaxForce(maxForce);
    j->SetMaxTorque(maxTorque);
    return 0;
}

int ref_world_set_motor(void* h, int joint, float ox, float oy, float angle, float maxForce, float maxTorque,
                        float factor) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_motorJoint) return -1;
    b2MotorJoint* j = static_cast<b2MotorJoint*>(w->joints[joint]);
    j->SetLinearOffset(b2Vec2(ox, oy));
    j->SetAngularOffset(angle);
    j->SetMaxForce(maxForce);
    j->SetMaxTorque(maxTorque);
    j->SetCorrectionFactor(factor);
    return 0;
}

int ref_world_set_rope(void* h, int joint, float maxLength) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_ropeJoint) return -1;
    static_cast<b2RopeJoint*>(w->joints[joint])->SetMaxLength(maxLength);
    return 0;
}

int ref_world_set_weld(void* h, int joint, float hz, float ratio) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    if (w->joints[joint]->GetType() != e_weldJoint) return -1;
    b2WeldJoint* j = static_cast<b2WeldJoint*>(w->joints[joint]);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

// b2World::DestroyBody; the driver forgets the body, its fixtures and its
// joints (their slots stay, empty, so indices keep their creation order).
int ref_world_destroy_body(void* h, int body) {
    RefWorld* w = (RefWorld*)h;
    if (body < 0 || body >= (int)w->bodies.size() || w->bodies[body] == nullptr) return -1;
    b2Body* b = w->bodies[body];
    for (auto& f : w->fixtures)
        if (f && f->GetBody() == b) f = nullptr;
    for (auto& j : w->joints)
        if (j && (j->GetBodyA() == b || j->GetBodyB() == b)) j = nullptr;
    w->world->DestroyBody(b);
    w->bodies[body] = nullptr;
    return 0;
}

int ref_world_query_aabb(void* h, const float* box, int32_t* out, int cap) {
    struct Q : b2QueryCallback {
        std::vector<int32_t> ids;
        bool ReportFixture(b2Fixture* f) override { ids.push_back(fixtureIndex(f)); return true; }
    } q;
    b2AABB a;
    a.lowerBound.Set(box[0], box[1]);
    a.upperBound.Set(box[2], box[3]);
    ((RefWorld*)h)->world->QueryAABB(&q, a);
    for (size_t i = 0; i < q.ids.size() && (int)i < cap; ++i) out[i] = q.ids[i];
    return (int)q.ids.size();
}

int ref_world_raycast_closest(void* h, float x1, float y1, float x2, float y2, float* out) {
    struct C : b2RayCastCallback {
        int32_t id = -1;
        float o[5] = {0, 0, 0, 0, 0};
        float32 ReportFixture(b2Fixture* f, const b2Vec2& p, const b2Vec2& n, float32 fraction) override {
            id = fixtureIndex(f);
            o[0] = fraction; o[1] = n.x; o[2] = n.y; o[3] = p.x; o[4] = p.y;
            return fraction;
        }
    } c;
    ((RefWorld*)h)->world->RayCast(&c, b2Vec2(x1, y1), b2Vec2(x2, y2));
    for (int k = 0; k < 5; ++k) out[k] = c.o[k];
    return c.id;
}

int ref_world_set_active(void* h, int body, int flag) {
    RefWorld* w = (RefWorld*)h;
    if (body < 0 || body >= (int)w->bodies.size() || w->bodies[body] == nullptr) return -1;
    w->bodies[body]->SetActive(flag != 0);
    return 0;
}

int ref_world_set_type(void* h, int body, int type) {
    RefWorld* w = (RefWorld*)h;
    if (body < 0 || body >= (int)w->bodies.size() || w->bodies[body] == nullptr || type < 0 || type > 2) return -1;
    w->bodies[body]->SetType((b2BodyType)type);
    return 0;
}

// b2Body::DestroyFixture of the fixture whose first proxy slot is `fixture`.
int ref_world_destroy_fixture(void* h, int fixture) {
    RefWorld* w = (RefWorld*)h;
    if (fixture < 0 || fixture >= (int)w->fixtures.size() || w->fixtures[fixture] == nullptr ||
        fixtureIndex(w->fixtures[fixture]) != fixture)
        return -1;
    b2Fixture* f = w->fixtures[fixture];
    for (auto& g : w->fixtures)
        if (g == f) g = nullptr;
    f->GetBody()->DestroyFixture(f);
    return 0;
}

// Per fixture proxy in creation order: its id in the broad-phase tree, -1
// once destroyed.
int ref_world_proxy_ids(void* h, int32_t* out, int cap) {
    RefWorld* w = (RefWorld*)h;
    const int n = (int)w->fixtures.size();
    for (int i = 0; i < n && i < cap; ++i) {
        const b2Fixture* f = w->fixtures[i];
        out[i] = f ? f->m_proxies[i - fixtureIndex(f)].proxyId : -1;
    }
    return n;
}

int ref_world_destroy_joint(void* h, int joint) {
    RefWorld* w = (RefWorld*)h;
    if (joint < 0 || joint >= (int)w->joints.size() || w->joints[joint] == nullptr) return -1;
    w->world->DestroyJoint(w->joints[joint]);
    w->joints[joint] = nullptr;
    return 0;
}

// Per joint in creation order: impulse x, y, z and motor impulse.
void ref_world_get_joints(void* h, float* out) {
    RefWorld* w = (RefWorld*)h;
    for (size_t i = 0; i < w->joints.size(); ++i) {
        if (w->joints[i] == nullptr) {  // destroyed
            for (int k = 0; k < 4; ++k) out[4 * i + k] = 0.0f;
            continue;
        }
        const b2JointType type = w->joints[i]->GetType();
        if (type == e_distanceJoint || type == e_ropeJoint || type == e_weldJoint || type == e_frictionJoint ||
            type == e_motorJoint || type == e_pulleyJoint || type == e_mouseJoint || type == e_prismaticJoint ||
            type == e_wheelJoint || type == e_gearJoint) {
            const b2Vec2 f = w->joints[i]->GetReactionForce(1.0f);
            out[4 * i] = f.x;
            out[4 * i + 1] = f.y;
            out[4 * i + 2] = w->joints[i]->GetReactionTorque(1.0f);
            out[4 * i + 3] = 0.0f;
            continue;
        }
        const b2RevoluteJoint* j = static_cast<const b2RevoluteJoint*>(w->joints[i]);
        out[4 * i] = j->m_impulse.x;
        out[4 * i + 1] = j->m_impulse.y;
        out[4 * i + 2] = j->m_impulse.z;
        out[4 * i + 3] = j->m_motorImpulse;
    }
}

void ref_world_step(void* h, float dt, int vel, int pos, int nsteps) {
    RefWorld* w = (RefWorld*)h;
    for (int i = 0; i < nsteps; ++i) {
        w->world->Step(dt, vel, pos);
        ++w->recorder.step;
    }
}

int ref_world_body_count(void* h) { return (int)((RefWorld*)h)->bodies.size(); }
int ref_world_fixture_count(void* h) { return (int)((RefWorld*)h)->fixtures.size(); }

// 12 floats per body in creation order:
// px py s c cx cy a vx vy w awake sleepTime
void ref_world_get_bodies(void* h, float* out) {
    RefWorld* w = (RefWorld*)h;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        const b2Body* b = w->bodies[i];
        float* o = out + 12 * i;
        if (b == nullptr) {  // destroyed
            for (int k = 0; k < 12; ++k) o[k] = 0.0f;
            continue;
        }
        const b2Transform& xf = b->GetTransform();
        o[0] = xf.p.x; o[1] = xf.p.y; o[2] = xf.q.s; o[3] = xf.q.c;
        o[4] = b->GetWorldCenter().x; o[5] = b->GetWorldCenter().y; o[6] = b->GetAngle();
        o[7] = b->GetLinearVelocity().x; o[8] = b->GetLinearVelocity().y;
        o[9] = b->GetAngularVelocity();
        o[10] = b->IsAwake() ? 1.0f : 0.0f;
        o[11] = b->m_sleepTime;
    }
}

int ref_world_contact_count(void* h) { return ((RefWorld*)h)->world->GetContactCount(); }

// Contacts in world-list order. meta: fixtureA, fixtureB, flags
// (bit0 touching, bit1 enabled), reserved.
int ref_world_get_contacts(void* h, int32_t* meta, b2g_manifold* manifolds, int cap) {
    RefWorld* w = (RefWorld*)h;
    int n = 0;
    for (b2Contact* c = w->world->GetContactList(); c; c = c->GetNext()) {
        if (n < cap) {
            meta[4 * n + 0] = fixtureIndex(c->GetFixtureA()) + c->GetChildIndexA();
            meta[4 * n + 1] = fixtureIndex(c->GetFixtureB()) + c->GetChildIndexB();
            meta[4 * n + 2] = (c->IsTouching() ? 1 : 0) | (c->IsEnabled() ? 2 : 0);
            meta[4 * n + 3] = 0;
            manifoldOut(*c->GetManifold(), manifolds + n);
        }
        ++n;
    }
    return n;
}

// Fat AABB of each fixture's proxy, fixture creation order.
void ref_world_get_fat_aabbs(void* h, b2g_aabb* out) {
    RefWorld* w = (RefWorld*)h;
    const b2BroadPhase& bp = w->world->GetContactManager().m_broadPhase;
    for (size_t i = 0; i < w->fixtures.size(); ++i) {
        if (w->fixtures[i] == nullptr) {  // destroyed
            out[i] = {0.0f, 0.0f, 0.0f, 0.0f};
            continue;
        }
        const int child = (int)i - fixtureIndex(w->fixtures[i]);
        const b2AABB& a = bp.GetFatAABB(w->fixtures[i]->m_proxies[child].proxyId);
        out[i] = {a.lowerBound.x, a.lowerBound.y, a.upperBound.x, a.upperBound.y};
    }
}

void ref_world_get_shapes(void* h, b2g_shape* out) {
    RefWorld* w = (RefWorld*)h;
    for (size_t i = 0; i < w->fixtures.size(); ++i) {
        if (w->fixtures[i] == nullptr) {  // destroyed
            memset(out + i, 0, sizeof(b2g_shape));
            continue;
        }
        const b2Shape* sh = w->fixtures[i]->GetShape();
        if (sh->GetType() == b2Shape::e_chain) {
            b2EdgeShape e;
            static_cast<const b2ChainShape*>(sh)->GetChildEdge(&e, (int)i - fixtureIndex(w->fixtures[i]));
            shapeToRec(&e, out + i);
            out[i].flags |= 4;
        } else {
            shapeToRec(sh, out + i);
        }
    }
}

// Per fixture proxy: owning body index; radius, friction, restitution.
void ref_world_fixture_info(void* h, int32_t* body, float* material) {
    RefWorld* w = (RefWorld*)h;
    std::unordered_map<const b2Body*, int> index;
    for (size_t i = 0; i < w->bodies.size(); ++i) index[w->bodies[i]] = (int)i;
    for (size_t i = 0; i < w->fixtures.size(); ++i) {
        const b2Fixture* f = w->fixtures[i];
        if (f == nullptr) {  // destroyed
            body[i] = -1;
            material[3 * i + 0] = material[3 * i + 1] = material[3 * i + 2] = 0.0f;
            continue;
        }
        body[i] = index[f->GetBody()];
        material[3 * i + 0] = f->GetShape()->m_radius;
        material[3 * i + 1] = f->GetFriction();
        material[3 * i + 2] = f->GetRestitution();
    }
}

// Per body: mass and rotational inertia about the centre of mass (b2Body::m_I
// after ResetMassData).
void ref_world_body_mass(void* h, float* out) {
    RefWorld* w = (RefWorld*)h;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        if (w->bodies[i] == nullptr) {  // destroyed
            out[2 * i + 0] = out[2 * i + 1] = 0.0f;
            continue;
        }
        out[2 * i + 0] = w->bodies[i]->m_mass;
        out[2 * i + 1] = w->bodies[i]->m_I;
    }
}

// Mean and max penetration (metres) over touching manifold points, computed
// from b2WorldManifold separations: mean of max(0, -separation).
void ref_world_penetration(void* h, double* out /* mean, max, points */) {
    RefWorld* w = (RefWorld*)h;
    double sum = 0.0, mx = 0.0;
    long np = 0;
    for (b2Contact* c = w->world->GetContactList(); c; c = c->GetNext()) {
        if (!c->IsTouching()) continue;
        b2WorldManifold wm;
        c->GetWorldManifold(&wm);
        for (int i = 0; i < c->GetManifold()->pointCount; ++i) {
            double p = wm.separations[i] < 0.0f ? -(double)wm.separations[i] : 0.0;
            sum += p;
            mx = std::max(mx, p);
            ++np;
        }
    }
    out[0] = np ? sum / np : 0.0;
    out[1] = mx;
    out[2] = (double)np;
}

// ----------------------------------------------------------- per function
int ref_collide(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                const b2g_transform* xfB, b2g_manifold* out) {
    b2Manifold m;
    m.pointCount = 0;
    b2Transform ta = toXf(xfA), tb = toXf(xfB);
    if (a->type == B2G_POLYGON && b->type == B2G_POLYGON) {
        b2PolygonShape pa, pb;
        recToPolygon(a, &pa);
        recToPolygon(b, &pb);
        b2CollidePolygons(&m, &pa, ta, &pb, tb);
    } else if (a->type == B2G_CIRCLE && b->type == B2G_CIRCLE) {
        b2CircleShape ca, cb;
        recToCircle(a, &ca);
        recToCircle(b, &cb);
        b2CollideCircles(&m, &ca, ta, &cb, tb);
    } else if (a->type == B2G_POLYGON && b->type == B2G_CIRCLE) {
        b2PolygonShape pa;
        b2CircleShape cb;
        recToPolygon(a, &pa);
        recToCircle(b, &cb);
        b2CollidePolygonAndCircle(&m, &pa, ta, &cb, tb);
    } else if (a->type == B2G_EDGE && b->type == B2G_CIRCLE) {
        b2EdgeShape ea;
        b2CircleShape cb;
        recToEdge(a, &ea);
        recToCircle(b, &cb);
        b2CollideEdgeAndCircle(&m, &ea, ta, &cb, tb);
    } else if (a->type == B2G_EDGE && b->type == B2G_POLYGON) {
        b2EdgeShape ea;
        b2PolygonShape pb;
        recToEdge(a, &ea);
        recToPolygon(b, &pb);
        b2CollideEdgeAndPolygon(&m, &ea, ta, &pb, tb);
    } else {
        return -1;
    }
    manifoldOut(m, out);
    return 0;
}

// Shape record -> the reference's shape object, kept alive for a proxy.
struct RefShape {
    b2PolygonShape poly;
    b2CircleShape circle;
    b2EdgeShape edge;
    const b2Shape* set(const b2g_shape* s) {
        if (s->type == B2G_POLYGON) { recToPolygon(s, &poly); return &poly; }
        if (s->type == B2G_CIRCLE) { recToCircle(s, &circle); return &circle; }
        recToEdge(s, &edge);
        return &edge;
    }
};

// b2Distance without radii from an empty cache; out = pointA(2), pointB(2),
// distance, iterations, simplex count.
void ref_distance(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b, const b2g_transform* xfB,
                  float* out) {
    RefShape sa, sb;
    b2DistanceInput in;
    in.proxyA.Set(sa.set(a), 0);
    in.proxyB.Set(sb.set(b), 0);
    in.transformA = toXf(xfA);
    in.transformB = toXf(xfB);
    in.useRadii = false;
    b2SimplexCache cache;
    cache.count = 0;
    b2DistanceOutput o;
    b2Distance(&o, &cache, &in);
    out[0] = o.pointA.x; out[1] = o.pointA.y; out[2] = o.pointB.x; out[3] = o.pointB.y;
    out[4] = o.distance; out[5] = (float)o.iterations; out[6] = (float)cache.count;
}

// b2TimeOfImpact; returns the output state (b2TOIOutput::State order), *t the time.
int ref_time_of_impact(const b2g_shape* a, const b2g_sweep* swA, const b2g_shape* b, const b2g_sweep* swB,
                       float tMax, float* t) {
    RefShape sa, sb;
    b2TOIInput in;
    in.proxyA.Set(sa.set(a), 0);
    in.proxyB.Set(sb.set(b), 0);
    const b2g_sweep* sw[2] = {swA, swB};
    b2Sweep* dst[2] = {&in.sweepA, &in.sweepB};
    for (int k = 0; k < 2; ++k) {
        dst[k]->localCenter.Set(sw[k]->localCenter[0], sw[k]->localCenter[1]);
        dst[k]->c0.Set(sw[k]->c0[0], sw[k]->c0[1]);
        dst[k]->c.Set(sw[k]->c[0], sw[k]->c[1]);
        dst[k]->a0 = sw[k]->a0;
        dst[k]->a = sw[k]->a;
        dst[k]->alpha0 = sw[k]->alpha0;
    }
    in.tMax = tMax;
    b2TOIOutput o;
    b2TimeOfImpact(&o, &in);
    *t = o.t;
    return (int)o.state;
}

// b2WorldManifold::Initialize; out = normal(2), points(4), separations(2).
void ref_world_manifold(const b2g_manifold* mf, const b2g_transform* xfA, float radiusA,
                        const b2g_transform* xfB, float radiusB, float* out) {
    b2Manifold m;
    m.type = (b2Manifold::Type)mf->type;
    m.pointCount = mf->pointCount;
    m.localNormal.Set(mf->localNormal.x, mf->localNormal.y);
    m.localPoint.Set(mf->localPoint.x, mf->localPoint.y);
    for (int i = 0; i < mf->pointCount; ++i) {
        m.points[i].localPoint.Set(mf->points[i].localPoint.x, mf->points[i].localPoint.y);
        m.points[i].id.key = mf->points[i].id;
    }
    b2WorldManifold wm;
    for (int i = 0; i < 8; ++i) out[i] = 0.0f;
    wm.Initialize(&m, toXf(xfA), radiusA, toXf(xfB), radiusB);
    if (m.pointCount == 0) return;
    out[0] = wm.normal.x; out[1] = wm.normal.y;
    for (int i = 0; i < m.pointCount; ++i) {
        out[2 + 2 * i] = wm.points[i].x;
        out[3 + 2 * i] = wm.points[i].y;
        out[6 + i] = wm.separations[i];
    }
}

// b2PolygonShape::Set on raw points (hull, normals, centroid).
void ref_polygon_set(const float* pts, int n, b2g_shape* out) {
    b2PolygonShape p;
    b2Vec2 v[b2_maxPolygonVertices];
    for (int i = 0; i < n; ++i) v[i].Set(pts[2 * i], pts[2 * i + 1]);
    p.Set(v, n);
    shapeToRec(&p, out);
}

void ref_polygon_box(float hx, float hy, float cx, float cy, float angle, int rotated, b2g_shape* out) {
    b2PolygonShape p;
    if (rotated) p.SetAsBox(hx, hy, b2Vec2(cx, cy), angle);
    else p.SetAsBox(hx, hy);
    shapeToRec(&p, out);
}

// b2Shape::ComputeMass; out = mass, center.x, center.y, I.
void ref_shape_mass(const b2g_shape* s, float density, float* out) {
    b2MassData md;
    if (s->type == B2G_POLYGON) {
        b2PolygonShape p;
        recToPolygon(s, &p);
        p.ComputeMass(&md, density);
    } else if (s->type == B2G_CIRCLE) {
        b2CircleShape c;
        recToCircle(s, &c);
        c.ComputeMass(&md, density);
    } else {
        b2EdgeShape e;
        recToEdge(s, &e);
        e.ComputeMass(&md, density);
    }
    out[0] = md.mass; out[1] = md.center.x; out[2] = md.center.y; out[3] = md.I;
}

void ref_shape_aabb(const b2g_shape* s, const b2g_transform* xf, b2g_aabb* out) {
    b2AABB a;
    b2Transform t = toXf(xf);
    if (s->type == B2G_POLYGON) {
        b2PolygonShape p;
        recToPolygon(s, &p);
        p.ComputeAABB(&a, t, 0);
    } else if (s->type == B2G_CIRCLE) {
        b2CircleShape c;
        recToCircle(s, &c);
        c.ComputeAABB(&a, t, 0);
    } else {
        b2EdgeShape e;
        recToEdge(s, &e);
        e.ComputeAABB(&a, t, 0);
    }
    *out = {a.lowerBound.x, a.lowerBound.y, a.upperBound.x, a.upperBound.y};
}

// ------------------------------------------------------------ broad-phase
void* ref_bp_create() { return new b2BroadPhase; }
void ref_bp_destroy(void* h) { delete (b2BroadPhase*)h; }

// Proxy user data is (index in creation order) + 1.
int ref_bp_create_proxy(void* h, const b2g_aabb* a, int userIndex) {
    b2AABB box;
    box.lowerBound.Set(a->lx, a->ly);
    box.upperBound.Set(a->ux, a->uy);
    return ((b2BroadPhase*)h)->CreateProxy(box, (void*)(intptr_t)(userIndex + 1));
}

// b2BroadPhase::MoveProxy; returns whether the tree re-inserted the proxy
// (the value b2DynamicTree::MoveProxy hands to BufferMove).
int ref_bp_move_proxy(void* h, int proxyId, const b2g_aabb* a, float dx, float dy) {
    b2AABB box;
    box.lowerBound.Set(a->lx, a->ly);
    box.upperBound.Set(a->ux, a->uy);
    b2BroadPhase* bp = (b2BroadPhase*)h;
    const bool moved = bp->m_tree.MoveProxy(proxyId, box, b2Vec2(dx, dy));
    if (moved) bp->BufferMove(proxyId);
    return moved ? 1 : 0;
}

// b2FindMaxSeparation (REF/Collision/b2CollidePolygon.cpp:28).
float ref_find_max_separation(const b2g_shape* a, const b2g_transform* xfA, const b2g_shape* b,
                              const b2g_transform* xfB, int* edge) {
    b2PolygonShape pa, pb;
    recToPolygon(a, &pa);
    recToPolygon(b, &pb);
    int32 e = 0;
    const float s = b2FindMaxSeparation(&e, &pa, toXf(xfA), &pb, toXf(xfB));
    *edge = e;
    return s;
}

// b2FindIncidentEdge (REF/Collision/b2CollidePolygon.cpp:69); out: v0.x v0.y v1.x v1.y, ids[2].
void ref_find_incident_edge(const b2g_shape* a, const b2g_transform* xfA, int edge, const b2g_shape* b,
                            const b2g_transform* xfB, float* out, uint32_t* ids) {
    b2PolygonShape pa, pb;
    recToPolygon(a, &pa);
    recToPolygon(b, &pb);
    b2ClipVertex c[2];
    b2FindIncidentEdge(c, &pa, toXf(xfA), edge, &pb, toXf(xfB));
    out[0] = c[0].v.x; out[1] = c[0].v.y; out[2] = c[1].v.x; out[3] = c[1].v.y;
    ids[0] = c[0].id.key;
    ids[1] = c[1].id.key;
}

// b2Transform::Set(position, angle) on the host (glibc sinf/cosf).
void ref_transform(float x, float y, float angle, b2g_transform* out) {
    b2Transform xf;
    xf.Set(b2Vec2(x, y), angle);
    out->px = xf.p.x; out->py = xf.p.y; out->s = xf.q.s; out->c = xf.q.c;
}

void ref_bp_get_fat(void* h, int proxyId, b2g_aabb* out) {
    const b2AABB& a = ((b2BroadPhase*)h)->GetFatAABB(proxyId);
    *out = {a.lowerBound.x, a.lowerBound.y, a.upperBound.x, a.upperBound.y};
}

// Pairs in the order UpdatePairs hands them to AddPair, as user indices.
int ref_bp_update_pairs(void* h, int32_t* out, int cap) {
    PairCollector pc;
    ((b2BroadPhase*)h)->UpdatePairs(&pc);
    const int n = (int)pc.pairs.size() / 2;
    for (int i = 0; i < n && i < cap; ++i) {
        out[2 * i] = pc.pairs[2 * i];
        out[2 * i + 1] = pc.pairs[2 * i + 1];
    }
    return n;
}

}  // extern "C"
