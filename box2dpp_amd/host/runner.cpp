// C entry points over the box2d:: API for the Python tests and benchmark:
// build a scene through b2World/b2Body/b2Fixture, step it, and read bodies,
// contacts and penetration back. The entry points mirror the CPU checkers'
// (orc_* / ref_*) with the prefix b2pp_, so tests compare all three alike.
#include "box2d/Box2D.h"
#include "b2g.h"
#include "scenes/scenes.h"

#include <string.h>
#include <vector>
#include <string>

using namespace box2d;

namespace {

// Records the listener calls (kind 1 BeginContact, 2 EndContact, 3 PreSolve,
// 4 PostSolve; fixture indices; for PreSolve whether the contact stays
// enabled) and disables the contacts of a one-sided platform fixture while
// the other body's origin is less than `clear` above the platform's top (a
// one-sided platform after the reference testbed's OneSidedPlatform).
// With a conveyor fixture set, PreSolve also gives its contacts the belt's
// tangent speed (b2Contact::SetTangentSpeed).
struct Recorder : public b2ContactListener {
    const std::vector<b2Fixture*>* fixtures = nullptr;
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
        events.insert(events.end(), {step, kind, indexOf(c->GetFixtureA()), indexOf(c->GetFixtureB()), flag});
    }
    void BeginContact(b2Contact* c) override { record(1, c, 0); }
    void EndContact(b2Contact* c) override { record(2, c, 0); }
    void PreSolve(b2Contact* c, const b2Manifold*) override {
        const int a = indexOf(c->GetFixtureA()), b = indexOf(c->GetFixtureB());
        if (a == platform || b == platform) {
            const b2Body* other = (a == platform ? c->GetFixtureB() : c->GetFixtureA())->GetBody();
            if (other->GetPosition()[1] < top + clear) c->SetEnabled(false);
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

struct Runner {
    b2World world;
    std::vector<b2Body*> bodies;
    std::vector<uint8_t> destroyed;  // per body: DestroyBody called (the slot stays)
    std::vector<b2Fixture*> fixtures;
    std::vector<b2Joint*> joints;
    Recorder recorder;
    explicit Runner(float gx, float gy) : world(b2Vec2(gx, gy)) {}
};

int loadScene(Runner* r, const b2scene::BodyDef* bodies, int nb, const b2scene::FixtureDef* fx, int nf,
              bool flush = true) {
    for (int i = 0; i < nb; ++i) {
        const b2scene::BodyDef& d = bodies[i];
        b2BodyDef bd;
        bd.type = (b2BodyType)d.type;
        bd.position = b2Vec2(d.px, d.py);
        bd.angle = d.angle;
        bd.linearVelocity = b2Vec2(d.vx, d.vy);
        bd.angularVelocity = d.w;
        bd.linearDamping = d.linearDamping;
        bd.angularDamping = d.angularDamping;
        bd.gravityScale = d.gravityScale;
        bd.allowSleep = d.allowSleep != 0;
        bd.awake = d.awake != 0;
        bd.fixedRotation = d.fixedRotation != 0;
        bd.bullet = d.bullet != 0;
        b2Body* body = r->world.CreateBody(&bd);
        r->bodies.push_back(body);
        for (int k = 0; k < d.fixtureCount; ++k) {
            const int fi = d.firstFixture + k;
            if (fi < 0 || fi >= nf) return B2G_ERR_ARG;
            const b2scene::FixtureDef& f = fx[fi];
            b2PolygonShape poly;
            b2CircleShape circle;
            b2EdgeShape edge;
            b2ChainShape chain;
            const b2Shape* shape = nullptr;
            if (f.shape == b2scene::kChain) {
                b2Vec2 p[8];
                for (int j = 0; j < f.n && j < 8; ++j) p[j] = b2Vec2(f.pts[2 * j], f.pts[2 * j + 1]);
                if (f.polyMode == 1) {
                    chain.CreateLoop(p, f.n);
                } else {
                    chain.CreateChain(p, f.n);
                    if (f.has0) chain.SetPrevVertex(b2Vec2(f.e[0], f.e[1]));
                    if (f.has3) chain.SetNextVertex(b2Vec2(f.e[6], f.e[7]));
                }
                shape = &chain;
            } else if (f.shape == b2scene::kPolygon) {
                if (f.polyMode == b2scene::kPolyBox) poly.SetAsBox(f.pts[0], f.pts[1]);
                else if (f.polyMode == b2scene::kPolyBoxRotated)
                    poly.SetAsBox(f.pts[0], f.pts[1], b2Vec2(f.pts[2], f.pts[3]), f.pts[4]);
                else {
                    b2Vec2 p[MAX_POLYGON_VERTICES];
                    for (int j = 0; j < f.n && j < MAX_POLYGON_VERTICES; ++j) p[j] = b2Vec2(f.pts[2 * j], f.pts[2 * j + 1]);
                    poly.Set(p, f.n);
                }
                shape = &poly;
            } else if (f.shape == b2scene::kCircle) {
                circle.m_radius = f.radius;
                circle.m_p = b2Vec2(f.cx, f.cy);
                shape = &circle;
            } else {
                edge.Set(b2Vec2(f.e[2], f.e[3]), b2Vec2(f.e[4], f.e[5]));
                edge.m_vertex0 = b2Vec2(f.e[0], f.e[1]);
                edge.m_vertex3 = b2Vec2(f.e[6], f.e[7]);
                edge.m_hasVertex0 = f.has0 != 0;
                edge.m_hasVertex3 = f.has3 != 0;
                shape = &edge;
            }
            b2FixtureDef fd;
            fd.shape = shape;
            fd.density = f.density;
            fd.friction = f.friction;
            fd.restitution = f.restitution;
            fd.isSensor = f.sensor != 0;
            fd.filter.categoryBits = (uint16_t)f.categoryBits;
            fd.filter.maskBits = (uint16_t)f.maskBits;
            fd.filter.groupIndex = (int16_t)f.groupIndex;
            b2Fixture* fixture = body->CreateFixture(&fd);
            // one entry per device fixture (a chain has one per child edge)
            for (int c = 0; c < shape->GetChildCount(); ++c) r->fixtures.push_back(fixture);
        }
    }
    return flush ? r->world.Flush() : 0;
}

// Scene joints through b2World::CreateJoint (local anchors and reference
// angle given directly; body indices relative to `base`).
int loadJoints(Runner* r, int base, const b2scene::JointDef* jd, int nj) {
    const int jbase = (int)r->joints.size();
    for (int k = 0; k < nj; ++k) {
        const b2scene::JointDef& d = jd[k];
        if (base + d.bodyA >= (int)r->bodies.size() || base + d.bodyB >= (int)r->bodies.size()) return B2G_ERR_ARG;
        if (d.type == 10) {
            if (jbase + d.joint1 >= (int)r->joints.size() || jbase + d.joint2 >= (int)r->joints.size())
                return B2G_ERR_ARG;
            b2GearJointDef gd;
            gd.bodyA = r->bodies[base + d.bodyA];
            gd.bodyB = r->bodies[base + d.bodyB];
            gd.collideConnected = d.collideConnected != 0;
            gd.joint1 = r->joints[jbase + d.joint1];
            gd.joint2 = r->joints[jbase + d.joint2];
            gd.ratio = d.ratio;
            b2Joint* j = r->world.CreateJoint(&gd);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 1) {
            b2DistanceJointDef dd;
            dd.bodyA = r->bodies[base + d.bodyA];
            dd.bodyB = r->bodies[base + d.bodyB];
            dd.collideConnected = d.collideConnected != 0;
            dd.localAnchorA = b2Vec2(d.ax, d.ay);
            dd.localAnchorB = b2Vec2(d.bx, d.by);
            dd.length = d.length;
            dd.frequencyHz = d.frequencyHz;
            dd.dampingRatio = d.dampingRatio;
            b2Joint* j = r->world.CreateJoint(&dd);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 8) {
            b2PrismaticJointDef pd;
            pd.bodyA = r->bodies[base + d.bodyA];
            pd.bodyB = r->bodies[base + d.bodyB];
            pd.collideConnected = d.collideConnected != 0;
            pd.localAnchorA = b2Vec2(d.ax, d.ay);
            pd.localAnchorB = b2Vec2(d.bx, d.by);
            pd.localAxisA = b2Vec2(d.axisX, d.axisY);
            pd.referenceAngle = d.referenceAngle;
            pd.enableLimit = d.enableLimit != 0;
            pd.lowerTranslation = d.lowerAngle;
            pd.upperTranslation = d.upperAngle;
            pd.enableMotor = d.enableMotor != 0;
            pd.maxMotorForce = d.maxMotorTorque;
            pd.motorSpeed = d.motorSpeed;
            b2Joint* j = r->world.CreateJoint(&pd);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 9) {
            b2WheelJointDef wd;
            wd.bodyA = r->bodies[base + d.bodyA];
            wd.bodyB = r->bodies[base + d.bodyB];
            wd.collideConnected = d.collideConnected != 0;
            wd.localAnchorA = b2Vec2(d.ax, d.ay);
            wd.localAnchorB = b2Vec2(d.bx, d.by);
            wd.localAxisA = b2Vec2(d.axisX, d.axisY);
            wd.enableMotor = d.enableMotor != 0;
            wd.maxMotorTorque = d.maxMotorTorque;
            wd.motorSpeed = d.motorSpeed;
            wd.frequencyHz = d.frequencyHz;
            wd.dampingRatio = d.dampingRatio;
            b2Joint* j = r->world.CreateJoint(&wd);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 6) {
            b2PulleyJointDef pd;
            pd.bodyA = r->bodies[base + d.bodyA];
            pd.bodyB = r->bodies[base + d.bodyB];
            pd.collideConnected = d.collideConnected != 0;
            pd.groundAnchorA = b2Vec2(d.gax, d.gay);
            pd.groundAnchorB = b2Vec2(d.gbx, d.gby);
            pd.localAnchorA = b2Vec2(d.ax, d.ay);
            pd.localAnchorB = b2Vec2(d.bx, d.by);
            pd.lengthA = d.length;
            pd.lengthB = d.lengthB;
            pd.ratio = d.ratio;
            b2Joint* j = r->world.CreateJoint(&pd);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 7) {  // created on the grab point, then pulled to the target
            b2MouseJointDef md;
            md.bodyA = r->bodies[base + d.bodyA];
            md.bodyB = r->bodies[base + d.bodyB];
            md.collideConnected = d.collideConnected != 0;
            md.target = md.bodyB->GetWorldPoint(b2Vec2(d.bx, d.by));
            md.maxForce = d.maxForce;
            md.frequencyHz = d.frequencyHz;
            md.dampingRatio = d.dampingRatio;
            b2Joint* j = r->world.CreateJoint(&md);
            if (j == nullptr) return B2G_ERR_ARG;
            static_cast<b2MouseJoint*>(j)->SetTarget(b2Vec2(d.ax, d.ay));
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 4 || d.type == 5) {
            b2FrictionJointDef fd;
            b2MotorJointDef md;
            b2JointDef* def = &fd;
            if (d.type == 4) {
                fd.localAnchorA = b2Vec2(d.ax, d.ay);
                fd.localAnchorB = b2Vec2(d.bx, d.by);
                fd.maxForce = d.maxForce;
                fd.maxTorque = d.maxTorque;
            } else {
                md.linearOffset = b2Vec2(d.ax, d.ay);
                md.angularOffset = d.referenceAngle;
                md.maxForce = d.maxForce;
                md.maxTorque = d.maxTorque;
                md.correctionFactor = d.correctionFactor;
                def = &md;
            }
            def->bodyA = r->bodies[base + d.bodyA];
            def->bodyB = r->bodies[base + d.bodyB];
            def->collideConnected = d.collideConnected != 0;
            b2Joint* j = r->world.CreateJoint(def);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        if (d.type == 2 || d.type == 3) {
            b2RopeJointDef rd;
            b2WeldJointDef wd;
            b2JointDef* def = &rd;
            if (d.type == 2) {
                rd.localAnchorA = b2Vec2(d.ax, d.ay);
                rd.localAnchorB = b2Vec2(d.bx, d.by);
                rd.maxLength = d.length;
            } else {
                wd.localAnchorA = b2Vec2(d.ax, d.ay);
                wd.localAnchorB = b2Vec2(d.bx, d.by);
                wd.referenceAngle = d.referenceAngle;
                wd.frequencyHz = d.frequencyHz;
                wd.dampingRatio = d.dampingRatio;
                def = &wd;
            }
            def->bodyA = r->bodies[base + d.bodyA];
            def->bodyB = r->bodies[base + d.bodyB];
            def->collideConnected = d.collideConnected != 0;
            b2Joint* j = r->world.CreateJoint(def);
            if (j == nullptr) return B2G_ERR_ARG;
            r->joints.push_back(j);
            continue;
        }
        b2RevoluteJointDef def;
        def.bodyA = r->bodies[base + d.bodyA];
        def.bodyB = r->bodies[base + d.bodyB];
        def.collideConnected = d.collideConnected != 0;
        def.localAnchorA = b2Vec2(d.ax, d.ay);
        def.localAnchorB = b2Vec2(d.bx, d.by);
        def.referenceAngle = d.referenceAngle;
        def.enableLimit = d.enableLimit != 0;
        def.lowerAngle = d.lowerAngle;
        def.upperAngle = d.upperAngle;
        def.enableMotor = d.enableMotor != 0;
        def.motorSpeed = d.motorSpeed;
        def.maxMotorTorque = d.maxMotorTorque;
        b2Joint* j = r->world.CreateJoint(&def);
        if (j == nullptr) return B2G_ERR_ARG;
        r->joints.push_back(j);
    }
    return 0;
}

}  // namespace

extern "C" {

void* b2pp_world_create(float gx, float gy) { return new Runner(gx, gy); }
void b2pp_world_destroy(void* h) { delete (Runner*)h; }

void b2pp_world_set_flags(void* h, int allowSleep, int warmStarting, int continuous, int subStepping) {
    Runner* r = (Runner*)h;
    r->world.SetAllowSleeping(allowSleep != 0);
    r->world.SetWarmStarting(warmStarting != 0);
    r->world.SetContinuousPhysics(continuous != 0);
    r->world.SetSubStepping(subStepping != 0);
}

void b2pp_world_set_mode(void* h, int mode) {
    Runner* r = (Runner*)h;
    b2SolverOptions o = r->world.GetSolverOptions();
    o.mode = mode;
    r->world.SetSolverOptions(o);
}

// Learned warm start on/off and its fallback threshold (b2SolverOptions).
void b2pp_world_set_learned(void* h, int on, float residualThreshold) {
    Runner* r = (Runner*)h;
    b2SolverOptions o = r->world.GetSolverOptions();
    o.learnedWarmStart = on != 0;
    o.residualThreshold = residualThreshold;
    r->world.SetSolverOptions(o);
}

int b2pp_world_load(void* h, const b2scene::BodyDef* bodies, int nb, const b2scene::FixtureDef* fx, int nf) {
    return loadScene((Runner*)h, bodies, nb, fx, nf);
}

int b2pp_world_load_scene(void* h, const char* kind, int n, unsigned seed) {
    Runner* r = (Runner*)h;
    b2scene::Scene s = b2scene::byName(kind, n, seed);
    const int base = (int)r->bodies.size();
    int rc = loadScene(r, s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size(), false);
    if (rc != 0) return rc;
    if ((rc = loadJoints(r, base, s.joints.data(), (int)s.joints.size())) != 0) return rc;
    return r->world.Flush();
}

int b2pp_world_joint_count(void* h) { return (int)((Runner*)h)->joints.size(); }

// Per joint in creation order: impulse x, y, z and motor impulse (the
// reaction force, reaction torque and motor torque at inv_dt = 1).
void b2pp_world_get_joints(void* h, float* out) {
    Runner* r = (Runner*)h;
    for (size_t i = 0; i < r->joints.size(); ++i) {
        const b2Joint* j = r->joints[i];
        if (j == nullptr) {  // destroyed
            for (int k = 0; k < 4; ++k) out[4 * i + k] = 0.0f;
            continue;
        }
        const b2Vec2 f = j->GetReactionForce(1.0f);
        out[4 * i] = f[0];
        out[4 * i + 1] = f[1];
        out[4 * i + 2] = j->GetReactionTorque(1.0f);
        out[4 * i + 3] = j->GetType() == b2JointType::REVOLUTE_JOINT
                             ? static_cast<const b2RevoluteJoint*>(j)->GetMotorTorque(1.0f)
                             : 0.0f;
    }
}

// Distance joint parameters through b2DistanceJoint::SetLength,
// SetFrequency and SetDampingRatio.
int b2pp_world_set_distance(void* h, int joint, float length, float hz, float ratio) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::DISTANCE_JOINT) return B2G_ERR_ARG;
    b2DistanceJoint* j = static_cast<b2DistanceJoint*>(r->joints[joint]);
    j->SetLength(length);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

// b2PrismaticJoint's limit and motor setters.
int b2pp_world_set_prismatic(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor,
                             float speed, float maxForce) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::PRISMATIC_JOINT) return B2G_ERR_ARG;
    b2PrismaticJoint* j = static_cast<b2PrismaticJoint*>(r->joints[joint]);
    j->EnableLimit(enableLimit != 0);
    j->SetLimits(lower, upper);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorForce(maxForce);
    return 0;
}

// b2WheelJoint's motor and spring setters.
int b2pp_world_set_wheel(void* h, int joint, int enableMotor, float speed, float maxTorque, float hz, float ratio) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::WHEEL_JOINT) return B2G_ERR_ARG;
    b2WheelJoint* j = static_cast<b2WheelJoint*>(r->joints[joint]);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorTorque(maxTorque);
    j->SetSpringFrequencyHz(hz);
    j->SetSpringDampingRatio(ratio);
    return 0;
}

// b2MouseJoint::SetTarget, SetMaxForce, SetFrequency, SetDampingRatio.
int b2pp_world_set_mouse(void* h, int joint, float tx, float ty, float maxForce, float hz, float ratio) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::MOUSE_JOINT) return B2G_ERR_ARG;
    b2MouseJoint* j = static_cast<b2MouseJoint*>(r->joints[joint]);
    j->SetTarget(b2Vec2(tx, ty));
    j->SetMaxForce(maxForce);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

// b2FrictionJoint::SetMaxForce / SetMaxTorque.
int b2pp_world_set_friction(void* h, int joint, float maxForce, float maxTorque) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::FRICTION_JOINT) return B2G_ERR_ARG;
    b2FrictionJoint* j = static_cast<b2FrictionJoint*>(r->joints[joint]);
    j->SetMaxForce(maxForce);
    j->SetMaxTorque(maxTorque);
    return 0;
}

// b2MotorJoint's setters.
int b2pp_world_set_motor(void* h, int joint, float ox, float oy, float angle, float maxForce, float maxTorque,
                         float factor) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::MOTOR_JOINT) return B2G_ERR_ARG;
    b2MotorJoint* j = static_cast<b2MotorJoint*>(r->joints[joint]);
    j->SetLinearOffset(b2Vec2(ox, oy));
    j->SetAngularOffset(angle);
    j->SetMaxForce(maxForce);
    j->SetMaxTorque(maxTorque);
    j->SetCorrectionFactor(factor);
    return 0;
}

// b2RopeJoint::SetMaxLength and b2WeldJoint::SetFrequency / SetDampingRatio.
int b2pp_world_set_rope(void* h, int joint, float maxLength) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::ROPE_JOINT) return B2G_ERR_ARG;
    static_cast<b2RopeJoint*>(r->joints[joint])->SetMaxLength(maxLength);
    return 0;
}

int b2pp_world_set_weld(void* h, int joint, float hz, float ratio) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::WELD_JOINT) return B2G_ERR_ARG;
    b2WeldJoint* j = static_cast<b2WeldJoint*>(r->joints[joint]);
    j->SetFrequency(hz);
    j->SetDampingRatio(ratio);
    return 0;
}

// Joint switches through the box2d:: setters (b2RevoluteJoint::EnableMotor,
// SetMotorSpeed, SetMaxMotorTorque, EnableLimit, SetLimits).
int b2pp_world_set_joint(void* h, int joint, int enableLimit, float lower, float upper, int enableMotor, float speed,
                         float maxTorque) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    if (r->joints[joint]->GetType() != b2JointType::REVOLUTE_JOINT) return B2G_ERR_ARG;
    b2RevoluteJoint* j = static_cast<b2RevoluteJoint*>(r->joints[joint]);
    j->EnableLimit(enableLimit != 0);
    j->SetLimits(lower, upper);
    j->EnableMotor(enableMotor != 0);
    j->SetMotorSpeed(speed);
    j->SetMaxMotorTorque(maxTorque);
    return 0;
}

// b2World::DestroyBody; the runner forgets the body's joints (the body's
// slot stays, so indices keep their creation order).
int b2pp_world_destroy_body(void* h, int body) {
    Runner* r = (Runner*)h;
    if (body < 0 || body >= (int)r->bodies.size()) return B2G_ERR_ARG;
    r->destroyed.resize(r->bodies.size(), 0);
    if (r->destroyed[body]) return B2G_ERR_ARG;
    b2Body* b = r->bodies[body];
    for (auto& j : r->joints)
        if (j && (j->GetBodyA() == b || j->GetBodyB() == b)) j = nullptr;
    r->world.DestroyBody(b);
    r->destroyed[body] = 1;
    return 0;
}

int b2pp_world_set_active(void* h, int body, int flag) {
    Runner* r = (Runner*)h;
    if (body < 0 || body >= (int)r->bodies.size()) return B2G_ERR_ARG;
    r->bodies[body]->SetActive(flag != 0);
    return 0;
}

// b2World::QueryAABB: the first proxy slot of each reported fixture, in
// report order; returns the report count.
int b2pp_world_query_aabb(void* h, const float* box, int32_t* out, int cap) {
    struct Q : b2QueryCallback {
        std::vector<int32_t> ids;
        bool ReportFixture(b2Fixture* f) override { ids.push_back(f->GetProxyIndex()); return true; }
    } q;
    b2AABB a;
    a.lowerBound = b2Vec2(box[0], box[1]);
    a.upperBound = b2Vec2(box[2], box[3]);
    ((Runner*)h)->world.QueryAABB(&q, a);
    for (size_t i = 0; i < q.ids.size() && (int)i < cap; ++i) out[i] = q.ids[i];
    return (int)q.ids.size();
}

// b2World::RayCast with the closest-hit callback: out = fraction, normal,
// point; returns the hit fixture's first proxy slot, -1 for none.
int b2pp_world_raycast_closest(void* h, float x1, float y1, float x2, float y2, float* out) {
    struct C : b2RayCastCallback {
        int32_t id = -1;
        float o[5] = {0, 0, 0, 0, 0};
        float ReportFixture(b2Fixture* f, const b2Vec2& p, const b2Vec2& n, float fraction) override {
            id = f->GetProxyIndex();
            o[0] = fraction; o[1] = n[0]; o[2] = n[1]; o[3] = p[0]; o[4] = p[1];
            return fraction;
        }
    } c;
    ((Runner*)h)->world.RayCast(&c, b2Vec2(x1, y1), b2Vec2(x2, y2));
    for (int k = 0; k < 5; ++k) out[k] = c.o[k];
    return c.id;
}

// b2Body::SetType (0 static, 1 kinematic, 2 dynamic).
int b2pp_world_set_type(void* h, int body, int type) {
    Runner* r = (Runner*)h;
    if (body < 0 || body >= (int)r->bodies.size() || type < 0 || type > 2) return B2G_ERR_ARG;
    r->bodies[body]->SetType((b2BodyType)type);
    return 0;
}

// b2Body::DestroyFixture of the fixture whose first proxy slot is `fixture`.
int b2pp_world_destroy_fixture(void* h, int fixture) {
    Runner* r = (Runner*)h;
    if (fixture < 0 || fixture >= (int)r->fixtures.size() || r->fixtures[fixture]->GetProxyIndex() != fixture)
        return B2G_ERR_ARG;
    b2Fixture* f = r->fixtures[fixture];  // stays valid (the world keeps it): the slot is kept
    f->GetBody()->DestroyFixture(f);
    return 0;
}

// Per fixture proxy in creation order: its broad-phase tree proxy id, -1
// once destroyed.
int b2pp_world_proxy_ids(void* h, int32_t* out, int cap) {
    const std::vector<int32_t>& ids = ((Runner*)h)->world.GetProxyIds();
    const int n = (int)ids.size();
    for (int i = 0; i < n && i < cap; ++i) out[i] = ids[i];
    return n;
}

int b2pp_world_destroy_joint(void* h, int joint) {
    Runner* r = (Runner*)h;
    if (joint < 0 || joint >= (int)r->joints.size() || r->joints[joint] == nullptr) return B2G_ERR_ARG;
    r->world.DestroyJoint(r->joints[joint]);
    r->joints[joint] = nullptr;
    return 0;
}

// Batched worlds: nWorlds worlds of the batched scene, each in its own
// sub-world of this b2World (dynamic body counts log-uniform in [lo, hi]
// from the seed, world w built from seed + w); sizes[w] receives world w's
// dynamic body count. World w's bodies are contiguous: its container, then
// its dynamic bodies.
int b2pp_world_load_batch(void* h, int nWorlds, unsigned seed, int lo, int hi, int* sizes) {
    Runner* r = (Runner*)h;
    if (nWorlds <= 0 || lo <= 0 || hi < lo) return B2G_ERR_ARG;
    const std::vector<int> n = b2scene::worldSizes(nWorlds, seed, lo, hi);
    for (int w = 0; w < nWorlds; ++w) {
        if (w > 0) r->world.BeginSubWorld();
        const b2scene::Scene s = b2scene::batchedWorld(n[w], seed + (unsigned)w);
        const int rc = loadScene(r, s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size(),
                                 false);
        if (rc != 0) return rc;
        if (sizes) sizes[w] = n[w];
    }
    return r->world.Flush();  // one upload of every world
}

// nWorlds copies of a named scene (world w built from seed + w), each in its
// own sub-world, joints included; one upload at the end. World w's bodies
// are contiguous, in scene order.
int b2pp_world_load_worlds(void* h, const char* kind, int nWorlds, int n, unsigned seed) {
    Runner* r = (Runner*)h;
    if (nWorlds <= 0) return B2G_ERR_ARG;
    for (int w = 0; w < nWorlds; ++w) {
        if (w > 0 || r->bodies.size() > 0) r->world.BeginSubWorld();
        const b2scene::Scene s = b2scene::byName(kind, n, seed + (unsigned)w);
        const int base = (int)r->bodies.size();
        int rc = loadScene(r, s.bodies.data(), (int)s.bodies.size(), s.fixtures.data(), (int)s.fixtures.size(),
                           false);
        if (rc != 0) return rc;
        if ((rc = loadJoints(r, base, s.joints.data(), (int)s.joints.size())) != 0) return rc;
    }
    return r->world.Flush();
}

int b2pp_world_step(void* h, float dt, int vel, int pos, int nsteps) {
    Runner* r = (Runner*)h;
    for (int i = 0; i < nsteps; ++i) {
        r->world.Step(dt, vel, pos);
        ++r->recorder.step;
        if (r->world.GetLastStatus() != 0) return r->world.GetLastStatus();
    }
    return 0;
}

// Installs the recording listener with the one-sided platform rule for
// fixture `platform` (its top and the clearance as given); -1: record only.
void b2pp_world_set_recorder(void* h, int platform, float top, float clear) {
    Runner* r = (Runner*)h;
    r->recorder.fixtures = &r->fixtures;
    r->recorder.platform = platform;
    r->recorder.top = top;
    r->recorder.clear = clear;
    r->world.SetContactListener(&r->recorder);
}

// Installs the recording listener (record only, unless set before) with
// `fixture` as a conveyor belt of tangent speed `speed`.
void b2pp_world_set_conveyor(void* h, int fixture, float speed) {
    Runner* r = (Runner*)h;
    r->recorder.fixtures = &r->fixtures;
    r->recorder.conveyor = fixture;
    r->recorder.belt = speed;
    r->world.SetContactListener(&r->recorder);
}

// Recorded events, 5 ints each; returns the event count.
int b2pp_world_events(void* h, int32_t* out, int cap) {
    Runner* r = (Runner*)h;
    const int n = (int)r->recorder.events.size() / 5;
    for (int i = 0; i < n && i < cap; ++i)
        for (int k = 0; k < 5; ++k) out[5 * i + k] = r->recorder.events[5 * i + k];
    return n;
}

// Autotunes the world's step schedule on 20 * rounds steps; *sched
// (b2g_schedule) receives the choice when not null.
int b2pp_world_autotune(void* h, float dt, int vel, int pos, int rounds, b2g_schedule* sched) {
    Runner* r = (Runner*)h;
    const int rc = r->world.Autotune(dt, vel, pos, rounds);
    if (rc == 0 && sched) b2g_get_schedule(r->world.GetDeviceContext(), sched);
    return rc;
}

int b2pp_world_set_schedule(void* h, const b2g_schedule* sched) {
    return b2g_set_schedule(((Runner*)h)->world.GetDeviceContext(), sched);
}

int b2pp_world_sync(void* h) { return b2g_sync(((Runner*)h)->world.GetDeviceContext()); }

const char* b2pp_world_last_error(void* h) { return ((Runner*)h)->world.GetLastError().c_str(); }

int b2pp_world_body_count(void* h) { return (int)((Runner*)h)->bodies.size(); }
int b2pp_world_fixture_count(void* h) { return (int)((Runner*)h)->fixtures.size(); }

void b2pp_world_get_bodies(void* h, float* out) {
    Runner* r = (Runner*)h;
    for (size_t i = 0; i < r->bodies.size(); ++i) {
        const b2Body* b = r->bodies[i];
        const b2Transform& xf = b->GetTransform();
        float* o = out + 12 * i;
        o[0] = xf.p[0]; o[1] = xf.p[1]; o[2] = xf.q.s; o[3] = xf.q.c;
        o[4] = b->GetWorldCenter()[0]; o[5] = b->GetWorldCenter()[1]; o[6] = b->GetAngle();
        o[7] = b->GetLinearVelocity()[0]; o[8] = b->GetLinearVelocity()[1];
        o[9] = b->GetAngularVelocity();
        o[10] = b->IsAwake() ? 1.0f : 0.0f;
        o[11] = 0.0f;
    }
}

int b2pp_world_contact_count(void* h) { return ((Runner*)h)->world.GetContactCount(); }

// Device table order (not the reference's list order).
int b2pp_world_get_contacts(void* h, int32_t* meta, b2g_manifold* manifolds, int cap) {
    Runner* r = (Runner*)h;
    int n = 0;
    b2g_get_contacts(r->world.GetDeviceContext(), meta, manifolds, cap, &n);
    return n;
}

void b2pp_world_get_fat_aabbs(void* h, b2g_aabb* out) {
    Runner* r = (Runner*)h;
    b2g_get_fat_aabbs(r->world.GetDeviceContext(), out, (int)r->fixtures.size());
}

// Per fixture proxy: owning body index; radius, friction, restitution.
void b2pp_world_fixture_info(void* h, int32_t* body, float* material) {
    Runner* r = (Runner*)h;
    for (size_t i = 0; i < r->fixtures.size(); ++i) {
        const b2Fixture* f = r->fixtures[i];
        body[i] = f->GetBody()->GetIndex();
        material[3 * i + 0] = f->GetShape()->m_radius;
        material[3 * i + 1] = f->GetFriction();
        material[3 * i + 2] = f->GetRestitution();
    }
}

// Per body: mass and rotational inertia about the centre of mass
// (b2Body::GetInertia is about the origin).
void b2pp_world_body_mass(void* h, float* out) {
    Runner* r = (Runner*)h;
    for (size_t i = 0; i < r->bodies.size(); ++i) {
        const b2Body* b = r->bodies[i];
        const b2Vec2& lc = b->GetLocalCenter();
        const double m = b->GetMass();
        out[2 * i + 0] = (float)m;
        out[2 * i + 1] = (float)((double)b->GetInertia() - m * ((double)lc[0] * lc[0] + (double)lc[1] * lc[1]));
    }
}

void b2pp_world_penetration(void* h, double* out) { ((Runner*)h)->world.GetPenetration(out); }

void* b2pp_world_device_context(void* h) { return ((Runner*)h)->world.GetDeviceContext(); }

int b2pp_world_stats(void* h, b2g_step_stats* out) {
    return b2g_get_stats(((Runner*)h)->world.GetDeviceContext(), out);
}

}  // extern "C"
