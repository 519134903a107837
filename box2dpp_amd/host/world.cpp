// Host side of the box2d:: dynamics API. b2World keeps host mirrors of its
// bodies and fixtures and hands the per-step work to the GPU through the
// C ABI of b2g.h:
//   * bodies and fixtures created since the last Step are appended to the
//     device world in creation order (body index = creation order, proxy
//     index = fixture creation order); the proxies' broad-phase order is the
//     reference tree's proxy ids, modelled here, which grow with creation
//     order until a body is destroyed and are uploaded from then on;
//   * host edits (SetTransform, velocities, forces, SetAwake, mass) mark a
//     body dirty and are uploaded before the next Step;
//   * after Step the mirrors are refreshed from the device on first read,
//     so getters that return references (GetTransform, GetPosition, ...) are
//     current, and contact listeners are replayed.
#include "box2d/b2World.h"
#include "b2g.h"

#include <algorithm>
#include <chrono>
#include <unordered_set>

namespace box2d {

namespace {

b2g_shape toDeviceShape(const b2Shape* s, int child) {
    if (s->GetType() == b2Shape::e_chain) {
        // a chain fixture is one device fixture per child edge
        b2EdgeShape e;
        static_cast<const b2ChainShape*>(s)->GetChildEdge(&e, child);
        b2g_shape r = toDeviceShape(&e, 0);
        r.flags |= 4;
        return r;
    }
    b2g_shape r{};
    r.type = (int32_t)s->GetType();
    r.radius = s->m_radius;
    if (s->GetType() == b2Shape::e_polygon) {
        const b2PolygonShape* p = static_cast<const b2PolygonShape*>(s);
        r.count = p->m_count;
        r.centroid = {p->m_centroid[0], p->m_centroid[1]};
        for (int i = 0; i < p->m_count; ++i) {
            r.verts[i] = {p->m_vertices[i][0], p->m_vertices[i][1]};
            r.normals[i] = {p->m_normals[i][0], p->m_normals[i][1]};
        }
    } else if (s->GetType() == b2Shape::e_circle) {
        const b2CircleShape* c = static_cast<const b2CircleShape*>(s);
        r.centroid = {c->m_p[0], c->m_p[1]};
    } else if (s->GetType() == b2Shape::e_edge) {
        const b2EdgeShape* e = static_cast<const b2EdgeShape*>(s);
        r.count = 2;
        r.verts[0] = {e->m_vertex1[0], e->m_vertex1[1]};
        r.verts[1] = {e->m_vertex2[0], e->m_vertex2[1]};
        r.v0 = {e->m_vertex0[0], e->m_vertex0[1]};
        r.v3 = {e->m_vertex3[0], e->m_vertex3[1]};
        r.flags = (e->m_hasVertex0 ? 1 : 0) | (e->m_hasVertex3 ? 2 : 0);
    }
    return r;
}

uint64_t pairKeyOf(int a, int b) {
    const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)lo << 32) | hi;
}

}  // namespace

// ------------------------------------------------------------ callbacks
// b2ContactFilter::ShouldCollide, REF/Dynamics/b2WorldCallbacks.cpp:26-38
bool b2ContactFilter::ShouldCollide(b2Fixture* fixtureA, b2Fixture* fixtureB) {
    const b2Filter& a = fixtureA->GetFilterData();
    const b2Filter& b = fixtureB->GetFilterData();
    if (a.groupIndex == b.groupIndex && a.groupIndex != 0) return a.groupIndex > 0;
    return (a.maskBits & b.categoryBits) != 0 && (a.categoryBits & b.maskBits) != 0;
}

// ------------------------------------------------------------ fixture
// Device fixtures of this fixture: one, or one per child of a chain.
void b2Fixture::MarkUpdated() {
    for (int32_t c = 0; c < m_shape->GetChildCount(); ++c) m_body->m_world->m_updatedFixtures.push_back(m_index + c);
}
void b2Fixture::SetFilterData(const b2Filter& filter) {
    m_filter = filter;
    MarkUpdated();
}
void b2Fixture::SetFriction(float f) {
    m_friction = f;
    MarkUpdated();
}
void b2Fixture::SetRestitution(float r) {
    m_restitution = r;
    MarkUpdated();
}
// REF/Dynamics/b2Fixture.cpp:223-230
void b2Fixture::SetSensor(bool sensor) {
    if (sensor == m_isSensor) return;
    m_body->SetAwake(true);
    m_isSensor = sensor;
    MarkUpdated();
}

// ------------------------------------------------------------ body
// b2Body::b2Body, REF/Dynamics/b2Body.cpp:25-90
b2Body::b2Body(const b2BodyDef* bd, b2World* world, int32_t index) : m_world(world) {
    m_index = index;
    m_flags = 0;
    if (bd->bullet) m_flags |= e_bulletFlag;
    if (bd->fixedRotation) m_flags |= e_fixedRotationFlag;
    if (bd->allowSleep) m_flags |= e_autoSleepFlag;
    if (bd->awake) m_flags |= e_awakeFlag;
    if (bd->active) m_flags |= e_activeFlag;
    m_xf.p = bd->position;
    m_xf.q.Set(bd->angle);
    m_sweep.localCenter.SetZero();
    m_sweep.c0 = m_xf.p;
    m_sweep.c = m_xf.p;
    m_sweep.a0 = bd->angle;
    m_sweep.a = bd->angle;
    m_sweep.alpha0 = 0.0f;
    m_linearVelocity = bd->linearVelocity;
    m_angularVelocity = bd->angularVelocity;
    m_linearDamping = bd->linearDamping;
    m_angularDamping = bd->angularDamping;
    m_gravityScale = bd->gravityScale;
    m_force.SetZero();
    m_torque = 0.0f;
    m_sleepTime = 0.0f;
    m_type = bd->type;
    if (m_type == b2BodyType::DYNAMIC_BODY) {
        m_mass = 1.0f;
        m_invMass = 1.0f;
    } else {
        m_mass = 0.0f;
        m_invMass = 0.0f;
    }
    m_I = 0.0f;
    m_invI = 0.0f;
    m_userData = bd->userData;
}

void b2Body::Touch() const { m_world->Download(); }

void b2Body::MarkDirty() {
    if (m_index < m_world->m_uploadedBodies) {
        m_dirty = true;
        m_world->m_anyDirty = true;
    }
}

// b2Body::CreateFixture, REF/Dynamics/b2Body.cpp:164-203
b2Fixture* b2Body::CreateFixture(const b2FixtureDef* def) {
    if (m_world->IsLocked() || def == nullptr || def->shape == nullptr) return nullptr;
    Touch();
    b2Fixture* f = new b2Fixture;
    f->m_shape.reset(def->shape->Clone());
    f->m_body = this;
    f->m_userData = def->userData;
    f->m_density = def->density;
    f->m_friction = def->friction;
    f->m_restitution = def->restitution;
    f->m_isSensor = def->isSensor;
    f->m_filter = def->filter;
    f->m_index = (int32_t)m_world->m_fixtures.size();
    for (int32_t c = 0; c < f->m_shape->GetChildCount(); ++c) {
        m_world->m_fixtures.push_back(f);
        // b2Fixture::CreateProxies, active bodies only (REF/Dynamics/b2Body.cpp:182-186)
        m_world->m_proxyIds.push_back(IsActive() ? m_world->CreateProxyId() : -1);
    }
    if (m_world->m_proxyIdsReused) m_world->m_proxyIdsDirty = true;
    f->m_next = m_fixtureList;
    m_fixtureList = f;
    ++m_fixtureCount;
    if (f->m_density > 0.0f) ResetMassData();
    return f;
}

// b2Body::DestroyFixture, REF/Dynamics/b2Body.cpp:216-281: off the body's
// list, its contacts destroyed (on the device: dropped at the next collide,
// their bodies woken when the proxies go), its proxies in child order, the
// mass data reset.
void b2Body::DestroyFixture(b2Fixture* fixture) {
    if (m_world->IsLocked() || fixture == nullptr || fixture->m_body != this || m_destroyed) return;
    Touch();
    b2Fixture** node = &m_fixtureList;
    while (*node && *node != fixture) node = &(*node)->m_next;
    if (*node == nullptr) return;
    *node = fixture->m_next;
    fixture->m_next = nullptr;
    --m_fixtureCount;
    b2World* w = m_world;
    for (int32_t c = 0; c < fixture->m_shape->GetChildCount(); ++c) {
        w->m_removedFixtures.push_back(fixture->m_index + c);
        int32_t& id = w->m_proxyIds[fixture->m_index + c];
        if (id >= 0) {
            w->DestroyProxyId(id);
            id = -1;
            w->m_proxyIdsReused = w->m_proxyIdsDirty = true;
        }
    }
    w->m_destroyedFixtures.push_back(fixture);
    w->m_contactsStale = true;
    ResetMassData();
}

b2Fixture* b2Body::CreateFixture(const b2Shape* shape, float density) {
    b2FixtureDef def;
    def.shape = shape;
    def.density = density;
    return CreateFixture(&def);
}

// b2Body::ResetMassData, REF/Dynamics/b2Body.cpp:285-352
void b2Body::ResetMassData() {
    Touch();
    m_mass = 0.0f;
    m_invMass = 0.0f;
    m_I = 0.0f;
    m_invI = 0.0f;
    m_sweep.localCenter.SetZero();
    if (m_type == b2BodyType::STATIC_BODY || m_type == b2BodyType::KINEMATIC_BODY) {
        m_sweep.c0 = m_xf.p;
        m_sweep.c = m_xf.p;
        m_sweep.a0 = m_sweep.a;
        MarkDirty();
        return;
    }
    b2Vec2 localCenter(0.0f, 0.0f);
    for (b2Fixture* f = m_fixtureList; f; f = f->m_next) {
        if (f->m_density == 0.0f) continue;
        b2MassData md;
        f->GetMassData(&md);
        m_mass += md.mass;
        localCenter += md.mass * md.center;
        m_I += md.I;
    }
    if (m_mass > 0.0f) {
        m_invMass = 1.0f / m_mass;
        localCenter *= m_invMass;
    } else {
        m_mass = 1.0f;
        m_invMass = 1.0f;
    }
    if (m_I > 0.0f && (m_flags & e_fixedRotationFlag) == 0) {
        m_I -= m_mass * b2Dot(localCenter, localCenter);
        m_invI = 1.0f / m_I;
    } else {
        m_I = 0.0f;
        m_invI = 0.0f;
    }
    const b2Vec2 oldCenter = m_sweep.c;
    m_sweep.localCenter = localCenter;
    m_sweep.c0 = m_sweep.c = b2Mul(m_xf, m_sweep.localCenter);
    m_linearVelocity += b2Cross(m_angularVelocity, m_sweep.c - oldCenter);
    MarkDirty();
    // a mouse joint's spring uses the mass of its body B at step time
    for (b2JointEdge* e = m_jointList; e; e = e->next)
        if (e->joint->GetType() == b2JointType::MOUSE_JOINT && e->joint->GetBodyB() == this) e->joint->MarkDirty();
}

const b2Transform& b2Body::GetTransform() const { Touch(); return m_xf; }
const b2Vec2& b2Body::GetPosition() const { Touch(); return m_xf.p; }
float b2Body::GetAngle() const { Touch(); return m_sweep.a; }
const b2Vec2& b2Body::GetWorldCenter() const { Touch(); return m_sweep.c; }
const b2Vec2& b2Body::GetLocalCenter() const { return m_sweep.localCenter; }
const b2Vec2& b2Body::GetLinearVelocity() const { Touch(); return m_linearVelocity; }
float b2Body::GetAngularVelocity() const { Touch(); return m_angularVelocity; }
bool b2Body::IsAwake() const { Touch(); return (m_flags & e_awakeFlag) != 0; }
float b2Body::GetInertia() const { return m_I + m_mass * b2Dot(m_sweep.localCenter, m_sweep.localCenter); }

void b2Body::GetMassData(b2MassData* data) const {
    data->mass = m_mass;
    data->I = GetInertia();
    data->center = m_sweep.localCenter;
}

b2Vec2 b2Body::GetWorldPoint(const b2Vec2& p) const { return b2Mul(GetTransform(), p); }
b2Vec2 b2Body::GetWorldVector(const b2Vec2& v) const { return b2Mul(GetTransform().q, v); }
b2Vec2 b2Body::GetLocalPoint(const b2Vec2& p) const { return b2MulT(GetTransform(), p); }
b2Vec2 b2Body::GetLocalVector(const b2Vec2& v) const { return b2MulT(GetTransform().q, v); }
b2Vec2 b2Body::GetLinearVelocityFromWorldPoint(const b2Vec2& p) const {
    Touch();
    return m_linearVelocity + b2Cross(m_angularVelocity, p - m_sweep.c);
}

// b2Body::SetTransform, REF/Dynamics/b2Body.cpp:419-440 (the proxies are
// re-synchronised on the device at the next Step through the move flags).
void b2Body::SetTransform(const b2Vec2& position, float angle) {
    if (m_world->IsLocked()) return;
    Touch();
    m_xf.q.Set(angle);
    m_xf.p = position;
    m_sweep.c = b2Mul(m_xf, m_sweep.localCenter);
    m_sweep.a = angle;
    m_sweep.c0 = m_sweep.c;
    m_sweep.a0 = angle;
    MarkDirty();
}

void b2Body::SetLinearVelocity(const b2Vec2& v) {
    if (m_type == b2BodyType::STATIC_BODY) return;
    Touch();
    if (b2Dot(v, v) > 0.0f) SetAwake(true);
    m_linearVelocity = v;
    MarkDirty();
}

void b2Body::SetAngularVelocity(float w) {
    if (m_type == b2BodyType::STATIC_BODY) return;
    Touch();
    if (w * w > 0.0f) SetAwake(true);
    m_angularVelocity = w;
    MarkDirty();
}

void b2Body::ApplyForce(const b2Vec2& force, const b2Vec2& point, bool wake) {
    if (m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    if (wake && (m_flags & e_awakeFlag) == 0) SetAwake(true);
    if (m_flags & e_awakeFlag) {
        m_force += force;
        m_torque += b2Cross(point - m_sweep.c, force);
        MarkDirty();
    }
}

void b2Body::ApplyForceToCenter(const b2Vec2& force, bool wake) {
    if (m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    if (wake && (m_flags & e_awakeFlag) == 0) SetAwake(true);
    if (m_flags & e_awakeFlag) {
        m_force += force;
        MarkDirty();
    }
}

void b2Body::ApplyTorque(float torque, bool wake) {
    if (m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    if (wake && (m_flags & e_awakeFlag) == 0) SetAwake(true);
    if (m_flags & e_awakeFlag) {
        m_torque += torque;
        MarkDirty();
    }
}

void b2Body::ApplyLinearImpulse(const b2Vec2& impulse, const b2Vec2& point, bool wake) {
    if (m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    if (wake && (m_flags & e_awakeFlag) == 0) SetAwake(true);
    if (m_flags & e_awakeFlag) {
        m_linearVelocity += m_invMass * impulse;
        m_angularVelocity += m_invI * b2Cross(point - m_sweep.c, impulse);
        MarkDirty();
    }
}

void b2Body::ApplyAngularImpulse(float impulse, bool wake) {
    if (m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    if (wake && (m_flags & e_awakeFlag) == 0) SetAwake(true);
    if (m_flags & e_awakeFlag) {
        m_angularVelocity += m_invI * impulse;
        MarkDirty();
    }
}

void b2Body::SetLinearDamping(float d) { Touch(); m_linearDamping = d; MarkDirty(); }
void b2Body::SetAngularDamping(float d) { Touch(); m_angularDamping = d; MarkDirty(); }
void b2Body::SetGravityScale(float s) { Touch(); m_gravityScale = s; MarkDirty(); }

void b2Body::SetSleepingAllowed(bool flag) {
    Touch();
    if (flag) m_flags |= e_autoSleepFlag;
    else {
        m_flags &= ~e_autoSleepFlag;
        SetAwake(true);
    }
    MarkDirty();
}

// b2Body::SetAwake, REF/Dynamics/b2Body.h:635-654
void b2Body::SetAwake(bool flag) {
    Touch();
    if (flag) {
        if ((m_flags & e_awakeFlag) == 0) {
            m_flags |= e_awakeFlag;
            m_sleepTime = 0.0f;
        }
    } else {
        m_flags &= ~e_awakeFlag;
        m_sleepTime = 0.0f;
        m_linearVelocity.SetZero();
        m_angularVelocity = 0.0f;
        m_force.SetZero();
        m_torque = 0.0f;
    }
    MarkDirty();
}

// ------------------------------------------------------------ world
b2World::b2World(const b2Vec2& gravity) : m_gravity(gravity) {
    if (b2g_create(&m_ctx, 0) != B2G_OK) {
        m_ctx = nullptr;
        m_lastStatus = B2G_ERR_HIP;
        m_lastError = "no HIP device: the GPU step library could not create a context";
    }
}

b2World::~b2World() {
    for (b2Joint* j : m_joints) delete j;
    for (b2Fixture* f : m_destroyedFixtures) delete f;
    for (b2Body* b : m_bodies) {
        if (!b) continue;
        b2Fixture* f = b->m_fixtureList;
        while (f) {
            b2Fixture* n = f->m_next;
            delete f;
            f = n;
        }
        delete b;
    }
    if (m_ctx) b2g_destroy(m_ctx);
}

// b2World::CreateBody, REF/Dynamics/b2World.cpp:109-127: new bodies go to
// the head of the body list.
b2Body* b2World::CreateBody(const b2BodyDef* def) {
    if (IsLocked() || def == nullptr) return nullptr;
    b2Body* b = new b2Body(def, this, (int32_t)m_bodies.size());
    b->m_subWorld = m_currentSubWorld;
    m_bodies.push_back(b);
    b->m_prev = nullptr;
    b->m_next = m_bodyList;
    if (m_bodyList) m_bodyList->m_prev = b;
    m_bodyList = b;
    ++m_bodyCount;
    return b;
}

// Destroyed bodies become inactive static bodies on the device; their
// proxies leave the broad-phase, so their contacts end at the next Step.
void b2World::DestroyBody(b2Body* b) {
    if (IsLocked() || b == nullptr || b->m_world != this || b->m_destroyed) return;
    b->m_destroyed = true;
    // Attached joints go first (REF/Dynamics/b2World.cpp:142-157), each
    // announced to the destruction listener.
    while (b->m_jointList) {
        b2Joint* j = b->m_jointList->joint;
        if (m_destructionListener) m_destructionListener->SayGoodbye(j);
        DestroyJoint(j);
    }
    b->Touch();
    if (b->m_prev) b->m_prev->m_next = b->m_next;
    if (b->m_next) b->m_next->m_prev = b->m_prev;
    if (b == m_bodyList) m_bodyList = b->m_next;
    b->m_flags = 0;
    b->m_type = b2BodyType::STATIC_BODY;
    b->m_linearVelocity.SetZero();
    b->m_angularVelocity = 0.0f;
    b->m_invMass = 0.0f;
    b->m_invI = 0.0f;
    b->MarkDirty();
    // fixtures newest first, each one's proxies in child order
    // (REF/Dynamics/b2World.cpp:170-188, b2Fixture.cpp:141-152)
    for (b2Fixture* f = b->m_fixtureList; f; f = f->m_next) {
        if (m_destructionListener) m_destructionListener->SayGoodbye(f);
        for (int32_t c = 0; c < f->m_shape->GetChildCount(); ++c) {
            m_removedFixtures.push_back(f->m_index + c);
            int32_t& id = m_proxyIds[f->m_index + c];
            if (id >= 0) {
                DestroyProxyId(id);
                id = -1;
                m_proxyIdsReused = m_proxyIdsDirty = true;
            }
        }
    }
    --m_bodyCount;
    m_contactsStale = true;
}

// The reference's broad-phase tree takes node ids from a free list that
// starts as 0..15 and doubles when it runs out (the new ids in ascending
// order), and returns freed nodes to its head (b2DynamicTree::AllocateNode /
// FreeNode, REF/Collision/b2DynamicTree.cpp:25-53, 56-106). A proxy takes a
// leaf node, then a parent node unless the tree was empty (CreateProxy,
// InsertLeaf :108-121, 179-265); destroying one frees its parent unless it
// was the root, then its leaf (DestroyProxy, RemoveLeaf :124-131, 319-376).
// A move frees the leaf's parent and takes the same node straight back, and
// nothing else in a step allocates, so the free list changes only here.
// Which internal node a leaf hangs from depends on the tree's shape; such a
// node is kept as -1 and is only ever taken again as a parent, since a freed
// leaf always lies above its parent in the list.
int32_t b2World::CreateProxyId() {
    auto take = [this]() {
        if (m_treeFree.empty()) {
            const int32_t old = m_treeCapacity;
            m_treeCapacity = old == 0 ? 16 : 2 * old;
            for (int32_t id = m_treeCapacity - 1; id >= old; --id) m_treeFree.push_back(id);
        }
        const int32_t id = m_treeFree.back();
        m_treeFree.pop_back();
        return id;
    };
    const int32_t leaf = take();
    if (m_treeLeaves > 0) (void)take();
    ++m_treeLeaves;
    return leaf;
}

void b2World::DestroyProxyId(int32_t id) {
    if (m_treeLeaves > 1) m_treeFree.push_back(-1);
    m_treeFree.push_back(id);
    --m_treeLeaves;
}

void b2World::QueryAABB(b2QueryCallback* callback, const b2AABB& aabb) {
    if (!callback || !m_ctx) return;
    if (!IsLocked() && Upload() != B2G_OK) return;
    const int n = (int)m_fixtures.size();
    std::vector<b2g_aabb> fat(n);
    if (n == 0 || b2g_get_fat_aabbs(m_ctx, fat.data(), n) != B2G_OK) return;
    for (int i = 0; i < n; ++i) {
        const b2g_aabb& f = fat[i];
        if (!(f.lx <= f.ux && f.ly <= f.uy)) continue;  // no proxy
        // b2TestOverlap (REF/Collision/b2Collision.h:241-252): touching counts
        if (aabb.lowerBound[0] > f.ux || aabb.lowerBound[1] > f.uy || f.lx > aabb.upperBound[0] ||
            f.ly > aabb.upperBound[1])
            continue;
        if (!callback->ReportFixture(m_fixtures[i])) return;
    }
}

// b2World::RayCast over b2DynamicTree::RayCast (REF/Dynamics/b2World.cpp:1026-1040,
// REF/Collision/b2DynamicTree.h:220-290): a proxy is tested when its fat box
// meets the segment cut at the current maximum fraction; a hit's callback
// value clips the segment (> 0), stops the cast (0) or is ignored (< 0).
void b2World::RayCast(b2RayCastCallback* callback, const b2Vec2& point1, const b2Vec2& point2) {
    if (!callback || !m_ctx) return;
    if (!IsLocked() && Upload() != B2G_OK) return;
    const int n = (int)m_fixtures.size();
    std::vector<b2g_aabb> fat(n);
    if (n == 0 || b2g_get_fat_aabbs(m_ctx, fat.data(), n) != B2G_OK) return;
    Download();
    const b2Vec2 r = point2 - point1;
    b2Vec2 v = b2Cross(1.0f, r);
    b2Normalize(v);
    const b2Vec2 absV(fabsf(v[0]), fabsf(v[1]));
    float maxFraction = 1.0f;
    for (int i = 0; i < n; ++i) {
        const b2g_aabb& f = fat[i];
        if (!(f.lx <= f.ux && f.ly <= f.uy)) continue;
        const b2Vec2 t = point1 + maxFraction * r;
        const b2Vec2 lo = b2Min(point1, t), hi = b2Max(point1, t);
        if (lo[0] > f.ux || lo[1] > f.uy || f.lx > hi[0] || f.ly > hi[1]) continue;
        // separating axis along the segment's normal
        const b2Vec2 c(0.5f * (f.lx + f.ux), 0.5f * (f.ly + f.uy));
        const b2Vec2 h(0.5f * (f.ux - f.lx), 0.5f * (f.uy - f.ly));
        if (fabsf(b2Dot(v, point1 - c)) - b2Dot(absV, h) > 0.0f) continue;
        b2Fixture* fixture = m_fixtures[i];
        b2RayCastInput input;
        input.p1 = point1;
        input.p2 = point2;
        input.maxFraction = maxFraction;
        b2RayCastOutput output;
        if (!fixture->m_shape->RayCast(&output, input, fixture->m_body->m_xf, i - fixture->m_index)) continue;
        const float fraction = output.fraction;
        const b2Vec2 point = (1.0f - fraction) * point1 + fraction * point2;
        const float value = callback->ReportFixture(fixture, point, output.normal, fraction);
        if (value == 0.0f) return;
        if (value > 0.0f) maxFraction = value;
    }
}

// ------------------------------------------------------------ joints
b2Joint::b2Joint(const b2JointDef* def, b2World* world, int32_t index)
    : m_type(def->type), m_bodyA(def->bodyA), m_bodyB(def->bodyB), m_world(world), m_index(index),
      m_collideConnected(def->collideConnected), m_userData(def->userData) {}

bool b2Joint::IsActive() const { return m_bodyA->IsActive() && m_bodyB->IsActive(); }

// b2RevoluteJointDef::Initialize, REF/Dynamics/Joints/b2RevoluteJoint.cpp:46-53
void b2RevoluteJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchor);
    localAnchorB = bodyB->GetLocalPoint(anchor);
    referenceAngle = bodyB->GetAngle() - bodyA->GetAngle();
}

b2RevoluteJoint::b2RevoluteJoint(const b2RevoluteJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_referenceAngle(def->referenceAngle), m_enableLimit(def->enableLimit), m_lowerAngle(def->lowerAngle),
      m_upperAngle(def->upperAngle), m_enableMotor(def->enableMotor), m_motorSpeed(def->motorSpeed),
      m_maxMotorTorque(def->maxMotorTorque) {}

b2Vec2 b2RevoluteJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2RevoluteJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }

// REF/Dynamics/Joints/b2RevoluteJoint.cpp:389-412
b2Vec2 b2RevoluteJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    return b2Vec2(inv_dt * m_impulse[0], inv_dt * m_impulse[1]);
}
float b2RevoluteJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
float b2RevoluteJoint::GetMotorTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[3];
}
float b2RevoluteJoint::GetJointAngle() const { return m_bodyB->GetAngle() - m_bodyA->GetAngle() - m_referenceAngle; }
float b2RevoluteJoint::GetJointSpeed() const {
    return m_bodyB->GetAngularVelocity() - m_bodyA->GetAngularVelocity();
}

void b2Joint::MarkDirty() {
    if (m_dirty) return;
    m_dirty = true;
    m_world->m_dirtyJoints.push_back(this);
}

// b2DistanceJointDef::Initialize, REF/Dynamics/Joints/b2DistanceJoint.cpp:40-49
void b2DistanceJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchorA, const b2Vec2& anchorB) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchorA);
    localAnchorB = bodyB->GetLocalPoint(anchorB);
    const b2Vec2 d = anchorB - anchorA;
    length = sqrtf(d[0] * d[0] + d[1] * d[1]);
}

b2DistanceJoint::b2DistanceJoint(const b2DistanceJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_length(def->length), m_frequencyHz(def->frequencyHz), m_dampingRatio(def->dampingRatio) {}

b2Vec2 b2DistanceJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2DistanceJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2DistanceJoint.cpp:233-244: the impulse along the
// direction u of the last step.
b2Vec2 b2DistanceJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    const float s = inv_dt * m_impulse[0];
    return b2Vec2(s * m_impulse[1], s * m_impulse[2]);
}
float b2DistanceJoint::GetReactionTorque(float inv_dt) const {
    B2_NOT_USED(inv_dt);
    return 0.0f;
}
// Plain setters, as the reference's (b2DistanceJoint.h:144-170)
void b2DistanceJoint::SetLength(float length) {
    m_length = length;
    MarkDirty();
}
void b2DistanceJoint::SetFrequency(float hz) {
    m_frequencyHz = hz;
    MarkDirty();
}
void b2DistanceJoint::SetDampingRatio(float ratio) {
    m_dampingRatio = ratio;
    MarkDirty();
}

b2RopeJoint::b2RopeJoint(const b2RopeJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_maxLength(def->maxLength) {}

b2Vec2 b2RopeJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2RopeJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2RopeJoint.cpp:208-218: the impulse along the rope's
// direction u of the last step.
b2Vec2 b2RopeJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    const float s = inv_dt * m_impulse[0];
    return b2Vec2(s * m_impulse[1], s * m_impulse[2]);
}
float b2RopeJoint::GetReactionTorque(float inv_dt) const {
    B2_NOT_USED(inv_dt);
    return 0.0f;
}
// The state the last step's InitVelocityConstraints chose (taut or slack).
b2LimitState b2RopeJoint::GetLimitState() const {
    m_world->DownloadJoints();
    return m_impulse[3] > 0.0f ? b2LimitState::AT_UPPER_LIMIT : b2LimitState::INACTIVE_LIMIT;
}
void b2RopeJoint::SetMaxLength(float length) {
    m_maxLength = length;
    MarkDirty();
}

// b2WeldJointDef::Initialize, REF/Dynamics/Joints/b2WeldJoint.cpp:39-46
void b2WeldJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchor);
    localAnchorB = bodyB->GetLocalPoint(anchor);
    referenceAngle = bodyB->GetAngle() - bodyA->GetAngle();
}

b2WeldJoint::b2WeldJoint(const b2WeldJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_referenceAngle(def->referenceAngle), m_frequencyHz(def->frequencyHz), m_dampingRatio(def->dampingRatio) {}

b2Vec2 b2WeldJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2WeldJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2WeldJoint.cpp:320-329
b2Vec2 b2WeldJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    return b2Vec2(inv_dt * m_impulse[0], inv_dt * m_impulse[1]);
}
float b2WeldJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
void b2WeldJoint::SetFrequency(float hz) {
    m_frequencyHz = hz;
    MarkDirty();
}
void b2WeldJoint::SetDampingRatio(float ratio) {
    m_dampingRatio = ratio;
    MarkDirty();
}

// Reaction force and torque of one joint at inv_dt = 1 from the device.
static b2g_vec2 deviceReaction(b2g_context* ctx, int32_t index, float* torque) {
    float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (ctx) (void)b2g_get_joint_reactions(ctx, index, 1, r);
    if (torque) *torque = r[2];
    b2g_vec2 f;
    f.x = r[0];
    f.y = r[1];
    return f;
}

// b2PrismaticJointDef::Initialize, REF/Dynamics/Joints/b2PrismaticJoint.cpp:90-99
void b2PrismaticJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor, const b2Vec2& axis) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchor);
    localAnchorB = bodyB->GetLocalPoint(anchor);
    localAxisA = bodyA->GetLocalVector(axis);
    referenceAngle = bodyB->GetAngle() - bodyA->GetAngle();
}

// The constructor normalizes the axis (REF b2PrismaticJoint.cpp:107-109).
b2PrismaticJoint::b2PrismaticJoint(const b2PrismaticJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_localXAxisA(def->localAxisA), m_referenceAngle(def->referenceAngle), m_enableLimit(def->enableLimit),
      m_lowerTranslation(def->lowerTranslation), m_upperTranslation(def->upperTranslation),
      m_enableMotor(def->enableMotor), m_maxMotorForce(def->maxMotorForce), m_motorSpeed(def->motorSpeed) {
    const float len = sqrtf(m_localXAxisA[0] * m_localXAxisA[0] + m_localXAxisA[1] * m_localXAxisA[1]);
    if (!(len < FLT_EPSILON)) {
        const float inv = 1.0f / len;
        m_localXAxisA = b2Vec2(m_localXAxisA[0] * inv, m_localXAxisA[1] * inv);
    }
}

b2Vec2 b2PrismaticJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2PrismaticJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2PrismaticJoint.cpp:509-518
float b2PrismaticJoint::GetJointTranslation() const {
    const b2Vec2 d = GetAnchorB() - GetAnchorA();
    const b2Vec2 axis = m_bodyA->GetWorldVector(m_localXAxisA);
    return d[0] * axis[0] + d[1] * axis[1];
}
// REF/Dynamics/Joints/b2PrismaticJoint.cpp:499-507 (the force along the
// last step's axes, from the device), and GetMotorForce.
b2Vec2 b2PrismaticJoint::GetReactionForce(float inv_dt) const {
    if (m_index >= m_world->m_uploadedJoints) return b2Vec2(0.0f, 0.0f);
    m_world->DownloadJoints();  // uploads pending setter edits
    const b2g_vec2 f = deviceReaction(m_world->m_ctx, m_index, nullptr);
    return b2Vec2(inv_dt * f.x, inv_dt * f.y);
}
float b2PrismaticJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[1];
}
float b2PrismaticJoint::GetMotorForce(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[3];
}
// REF/Dynamics/Joints/b2PrismaticJoint.cpp:546-609
void b2PrismaticJoint::EnableLimit(bool flag) {
    if (flag != m_enableLimit) {
        m_bodyA->SetAwake(true);
        m_bodyB->SetAwake(true);
        m_enableLimit = flag;
        MarkDirty();
    }
}
void b2PrismaticJoint::SetLimits(float lower, float upper) {
    if (lower != m_lowerTranslation || upper != m_upperTranslation) {
        m_bodyA->SetAwake(true);
        m_bodyB->SetAwake(true);
        m_lowerTranslation = lower;
        m_upperTranslation = upper;
        MarkDirty();
    }
}
void b2PrismaticJoint::EnableMotor(bool flag) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_enableMotor = flag;
    MarkDirty();
}
void b2PrismaticJoint::SetMotorSpeed(float speed) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_motorSpeed = speed;
    MarkDirty();
}
void b2PrismaticJoint::SetMaxMotorForce(float force) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_maxMotorForce = force;
    MarkDirty();
}

// b2WheelJointDef::Initialize, REF/Dynamics/Joints/b2WheelJoint.cpp:41-48
void b2WheelJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor, const b2Vec2& axis) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchor);
    localAnchorB = bodyB->GetLocalPoint(anchor);
    localAxisA = bodyA->GetLocalVector(axis);
}

b2WheelJoint::b2WheelJoint(const b2WheelJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_localXAxisA(def->localAxisA), m_enableMotor(def->enableMotor), m_maxMotorTorque(def->maxMotorTorque),
      m_motorSpeed(def->motorSpeed), m_frequencyHz(def->frequencyHz), m_dampingRatio(def->dampingRatio) {}

b2Vec2 b2WheelJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2WheelJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2WheelJoint.cpp:341-349
b2Vec2 b2WheelJoint::GetReactionForce(float inv_dt) const {
    if (m_index >= m_world->m_uploadedJoints) return b2Vec2(0.0f, 0.0f);
    m_world->DownloadJoints();  // uploads pending setter edits
    const b2g_vec2 f = deviceReaction(m_world->m_ctx, m_index, nullptr);
    return b2Vec2(inv_dt * f.x, inv_dt * f.y);
}
float b2WheelJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
float b2WheelJoint::GetMotorTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
// REF/Dynamics/Joints/b2WheelJoint.cpp:377-401
void b2WheelJoint::EnableMotor(bool flag) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_enableMotor = flag;
    MarkDirty();
}
void b2WheelJoint::SetMotorSpeed(float speed) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_motorSpeed = speed;
    MarkDirty();
}
void b2WheelJoint::SetMaxMotorTorque(float torque) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_maxMotorTorque = torque;
    MarkDirty();
}
void b2WheelJoint::SetSpringFrequencyHz(float hz) {
    m_frequencyHz = hz;
    MarkDirty();
}
void b2WheelJoint::SetSpringDampingRatio(float ratio) {
    m_dampingRatio = ratio;
    MarkDirty();
}

b2GearJoint::b2GearJoint(const b2GearJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_joint1(def->joint1), m_joint2(def->joint2), m_ratio(def->ratio), m_def{} {}

b2Vec2 b2GearJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2GearJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2GearJoint.cpp:406-416, from the device
b2Vec2 b2GearJoint::GetReactionForce(float inv_dt) const {
    if (m_index >= m_world->m_uploadedJoints) return b2Vec2(0.0f, 0.0f);
    m_world->DownloadJoints();  // uploads pending setter edits
    const b2g_vec2 f = deviceReaction(m_world->m_ctx, m_index, nullptr);
    return b2Vec2(inv_dt * f.x, inv_dt * f.y);
}
float b2GearJoint::GetReactionTorque(float inv_dt) const {
    if (m_index >= m_world->m_uploadedJoints) return 0.0f;
    m_world->DownloadJoints();  // uploads pending setter edits
    float t = 0.0f;
    (void)deviceReaction(m_world->m_ctx, m_index, &t);
    return inv_dt * t;
}
void b2GearJoint::SetRatio(float ratio) {
    m_ratio = ratio;
    m_def.ratio = ratio;
    MarkDirty();
}

// b2PulleyJointDef::Initialize, REF/Dynamics/Joints/b2PulleyJoint.cpp:38-55
void b2PulleyJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& groundA, const b2Vec2& groundB,
                                  const b2Vec2& anchorA, const b2Vec2& anchorB, float r) {
    bodyA = bA;
    bodyB = bB;
    groundAnchorA = groundA;
    groundAnchorB = groundB;
    localAnchorA = bodyA->GetLocalPoint(anchorA);
    localAnchorB = bodyB->GetLocalPoint(anchorB);
    const b2Vec2 dA = anchorA - groundA;
    lengthA = sqrtf(dA[0] * dA[0] + dA[1] * dA[1]);
    const b2Vec2 dB = anchorB - groundB;
    lengthB = sqrtf(dB[0] * dB[0] + dB[1] * dB[1]);
    ratio = r;
}

b2PulleyJoint::b2PulleyJoint(const b2PulleyJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_groundAnchorA(def->groundAnchorA), m_groundAnchorB(def->groundAnchorB),
      m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB), m_lengthA(def->lengthA),
      m_lengthB(def->lengthB), m_ratio(def->ratio) {}

b2Vec2 b2PulleyJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2PulleyJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
float b2PulleyJoint::GetCurrentLengthA() const {
    const b2Vec2 d = GetAnchorA() - m_groundAnchorA;
    return sqrtf(d[0] * d[0] + d[1] * d[1]);
}
float b2PulleyJoint::GetCurrentLengthB() const {
    const b2Vec2 d = GetAnchorB() - m_groundAnchorB;
    return sqrtf(d[0] * d[0] + d[1] * d[1]);
}
// REF/Dynamics/Joints/b2PulleyJoint.cpp:275-285: the impulse along body B's
// rope direction of the last step.
b2Vec2 b2PulleyJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    const b2Vec2 P(m_impulse[0] * m_impulse[1], m_impulse[0] * m_impulse[2]);
    return b2Vec2(inv_dt * P[0], inv_dt * P[1]);
}
float b2PulleyJoint::GetReactionTorque(float inv_dt) const {
    B2_NOT_USED(inv_dt);
    return 0.0f;
}

// b2MouseJoint ctor, REF/Dynamics/Joints/b2MouseJoint.cpp:34-53: the anchor
// on body B is the target at creation.
b2MouseJoint::b2MouseJoint(const b2MouseJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_targetA(def->target),
      m_localAnchorB(b2MulT(def->bodyB->GetTransform(), def->target)), m_maxForce(def->maxForce),
      m_frequencyHz(def->frequencyHz), m_dampingRatio(def->dampingRatio) {}

b2Vec2 b2MouseJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
b2Vec2 b2MouseJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    return b2Vec2(inv_dt * m_impulse[0], inv_dt * m_impulse[1]);
}
float b2MouseJoint::GetReactionTorque(float inv_dt) const { return inv_dt * 0.0f; }
// REF/Dynamics/Joints/b2MouseJoint.cpp:55-97
void b2MouseJoint::SetTarget(const b2Vec2& target) {
    if (!m_bodyB->IsAwake()) m_bodyB->SetAwake(true);
    m_targetA = target;
    MarkDirty();
}
void b2MouseJoint::SetMaxForce(float force) {
    m_maxForce = force;
    MarkDirty();
}
void b2MouseJoint::SetFrequency(float hz) {
    m_frequencyHz = hz;
    MarkDirty();
}
void b2MouseJoint::SetDampingRatio(float ratio) {
    m_dampingRatio = ratio;
    MarkDirty();
}

// b2FrictionJointDef::Initialize, REF/Dynamics/Joints/b2FrictionJoint.cpp:38-44
void b2FrictionJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor) {
    bodyA = bA;
    bodyB = bB;
    localAnchorA = bodyA->GetLocalPoint(anchor);
    localAnchorB = bodyB->GetLocalPoint(anchor);
}

b2FrictionJoint::b2FrictionJoint(const b2FrictionJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_localAnchorA(def->localAnchorA), m_localAnchorB(def->localAnchorB),
      m_maxForce(def->maxForce), m_maxTorque(def->maxTorque) {}

b2Vec2 b2FrictionJoint::GetAnchorA() const { return m_bodyA->GetWorldPoint(m_localAnchorA); }
b2Vec2 b2FrictionJoint::GetAnchorB() const { return m_bodyB->GetWorldPoint(m_localAnchorB); }
// REF/Dynamics/Joints/b2FrictionJoint.cpp:208-216
b2Vec2 b2FrictionJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    return b2Vec2(inv_dt * m_impulse[0], inv_dt * m_impulse[1]);
}
float b2FrictionJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
void b2FrictionJoint::SetMaxForce(float force) {
    m_maxForce = force;
    MarkDirty();
}
void b2FrictionJoint::SetMaxTorque(float torque) {
    m_maxTorque = torque;
    MarkDirty();
}

// b2MotorJointDef::Initialize, REF/Dynamics/Joints/b2MotorJoint.cpp:37-47
void b2MotorJointDef::Initialize(b2Body* bA, b2Body* bB) {
    bodyA = bA;
    bodyB = bB;
    linearOffset = bodyA->GetLocalPoint(bodyB->GetPosition());
    angularOffset = bodyB->GetAngle() - bodyA->GetAngle();
}

b2MotorJoint::b2MotorJoint(const b2MotorJointDef* def, b2World* world, int32_t index)
    : b2Joint(def, world, index), m_linearOffset(def->linearOffset), m_angularOffset(def->angularOffset),
      m_maxForce(def->maxForce), m_maxTorque(def->maxTorque), m_correctionFactor(def->correctionFactor) {}

b2Vec2 b2MotorJoint::GetAnchorA() const { return m_bodyA->GetPosition(); }
b2Vec2 b2MotorJoint::GetAnchorB() const { return m_bodyB->GetPosition(); }
// REF/Dynamics/Joints/b2MotorJoint.cpp:218-226
b2Vec2 b2MotorJoint::GetReactionForce(float inv_dt) const {
    m_world->DownloadJoints();
    return b2Vec2(inv_dt * m_impulse[0], inv_dt * m_impulse[1]);
}
float b2MotorJoint::GetReactionTorque(float inv_dt) const {
    m_world->DownloadJoints();
    return inv_dt * m_impulse[2];
}
// The offset setters wake both bodies when the offset changes
// (REF/Dynamics/Joints/b2MotorJoint.cpp:261-284).
void b2MotorJoint::SetLinearOffset(const b2Vec2& linearOffset) {
    if (linearOffset[0] != m_linearOffset[0] || linearOffset[1] != m_linearOffset[1]) {
        m_bodyA->SetAwake(true);
        m_bodyB->SetAwake(true);
        m_linearOffset = linearOffset;
        MarkDirty();
    }
}
void b2MotorJoint::SetAngularOffset(float angularOffset) {
    if (angularOffset != m_angularOffset) {
        m_bodyA->SetAwake(true);
        m_bodyB->SetAwake(true);
        m_angularOffset = angularOffset;
        MarkDirty();
    }
}
void b2MotorJoint::SetMaxForce(float force) {
    m_maxForce = force;
    MarkDirty();
}
void b2MotorJoint::SetMaxTorque(float torque) {
    m_maxTorque = torque;
    MarkDirty();
}
void b2MotorJoint::SetCorrectionFactor(float factor) {
    m_correctionFactor = factor;
    MarkDirty();
}

// The setters wake both bodies and, where the limit changes, drop the limit
// impulse (REF/Dynamics/Joints/b2RevoluteJoint.cpp:419-471).
void b2RevoluteJoint::EnableMotor(bool flag) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_enableMotor = flag;
    MarkDirty();
}
void b2RevoluteJoint::SetMotorSpeed(float speed) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_motorSpeed = speed;
    MarkDirty();
}
void b2RevoluteJoint::SetMaxMotorTorque(float torque) {
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_maxMotorTorque = torque;
    MarkDirty();
}
void b2RevoluteJoint::EnableLimit(bool flag) {
    if (flag == m_enableLimit) return;
    m_world->DownloadJoints();
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_enableLimit = flag;
    m_impulse[2] = 0.0f;
    MarkDirty();
}
void b2RevoluteJoint::SetLimits(float lower, float upper) {
    if (lower == m_lowerAngle && upper == m_upperAngle) return;
    m_world->DownloadJoints();
    m_bodyA->SetAwake(true);
    m_bodyB->SetAwake(true);
    m_impulse[2] = 0.0f;
    m_lowerAngle = lower;
    m_upperAngle = upper;
    MarkDirty();
}

// b2World::CreateJoint, REF/Dynamics/b2World.cpp:205-262: the joint goes to
// the head of the world's list and of both bodies' joint lists. Contacts
// that a collideConnected = false joint filters out end at the next Step.
b2Joint* b2World::CreateJoint(const b2JointDef* def) {
    if (IsLocked() || def == nullptr || def->bodyA == nullptr || def->bodyB == nullptr) return nullptr;
    if (def->type != b2JointType::REVOLUTE_JOINT && def->type != b2JointType::DISTANCE_JOINT &&
        def->type != b2JointType::ROPE_JOINT && def->type != b2JointType::WELD_JOINT &&
        def->type != b2JointType::FRICTION_JOINT && def->type != b2JointType::MOTOR_JOINT &&
        def->type != b2JointType::PULLEY_JOINT && def->type != b2JointType::MOUSE_JOINT &&
        def->type != b2JointType::PRISMATIC_JOINT && def->type != b2JointType::WHEEL_JOINT &&
        def->type != b2JointType::GEAR_JOINT) {
        m_lastError = "unknown joint type";
        return nullptr;
    }
    if (def->type == b2JointType::GEAR_JOINT) {
        const b2GearJointDef* g = static_cast<const b2GearJointDef*>(def);
        for (const b2Joint* jj : {g->joint1, g->joint2}) {
            if (jj == nullptr || jj->m_world != this ||
                (jj->m_type != b2JointType::REVOLUTE_JOINT && jj->m_type != b2JointType::PRISMATIC_JOINT)) {
                m_lastError = "a gear joint joins two revolute or prismatic joints of this world";
                return nullptr;
            }
        }
    }
    if (def->bodyA == def->bodyB || def->bodyA->m_world != this || def->bodyB->m_world != this) return nullptr;
    if (def->bodyA->m_subWorld != def->bodyB->m_subWorld) {
        m_lastError = "a joint connects bodies of one sub-world";
        return nullptr;
    }
    b2Joint* j = nullptr;
    const int32_t index = (int32_t)m_joints.size();
    if (def->type == b2JointType::DISTANCE_JOINT)
        j = new b2DistanceJoint(static_cast<const b2DistanceJointDef*>(def), this, index);
    else if (def->type == b2JointType::ROPE_JOINT)
        j = new b2RopeJoint(static_cast<const b2RopeJointDef*>(def), this, index);
    else if (def->type == b2JointType::WELD_JOINT)
        j = new b2WeldJoint(static_cast<const b2WeldJointDef*>(def), this, index);
    else if (def->type == b2JointType::FRICTION_JOINT)
        j = new b2FrictionJoint(static_cast<const b2FrictionJointDef*>(def), this, index);
    else if (def->type == b2JointType::MOTOR_JOINT)
        j = new b2MotorJoint(static_cast<const b2MotorJointDef*>(def), this, index);
    else if (def->type == b2JointType::PULLEY_JOINT)
        j = new b2PulleyJoint(static_cast<const b2PulleyJointDef*>(def), this, index);
    else if (def->type == b2JointType::MOUSE_JOINT)
        j = new b2MouseJoint(static_cast<const b2MouseJointDef*>(def), this, index);
    else if (def->type == b2JointType::PRISMATIC_JOINT)
        j = new b2PrismaticJoint(static_cast<const b2PrismaticJointDef*>(def), this, index);
    else if (def->type == b2JointType::WHEEL_JOINT)
        j = new b2WheelJoint(static_cast<const b2WheelJointDef*>(def), this, index);
    else if (def->type == b2JointType::GEAR_JOINT) {
        b2GearJoint* g = new b2GearJoint(static_cast<const b2GearJointDef*>(def), this, index);
        ResolveGear(g);
        j = g;
    }
    else
        j = new b2RevoluteJoint(static_cast<const b2RevoluteJointDef*>(def), this, (int32_t)m_joints.size());
    m_joints.push_back(j);
    j->m_prev = nullptr;
    j->m_next = m_jointList;
    if (m_jointList) m_jointList->m_prev = j;
    m_jointList = j;
    ++m_jointCount;
    j->m_edgeA.joint = j;
    j->m_edgeA.other = j->m_bodyB;
    j->m_edgeA.prev = nullptr;
    j->m_edgeA.next = j->m_bodyA->m_jointList;
    if (j->m_bodyA->m_jointList) j->m_bodyA->m_jointList->prev = &j->m_edgeA;
    j->m_bodyA->m_jointList = &j->m_edgeA;
    j->m_edgeB.joint = j;
    j->m_edgeB.other = j->m_bodyA;
    j->m_edgeB.prev = nullptr;
    j->m_edgeB.next = j->m_bodyB->m_jointList;
    if (j->m_bodyB->m_jointList) j->m_bodyB->m_jointList->prev = &j->m_edgeB;
    j->m_bodyB->m_jointList = &j->m_edgeB;
    m_contactsStale = true;
    return j;
}

// b2World::DestroyJoint, REF/Dynamics/b2World.cpp:267-345: unlinked, both
// bodies woken; on the device the joint stops acting at the next Step.
void b2World::DestroyJoint(b2Joint* j) {
    if (IsLocked() || j == nullptr || j->m_world != this) return;
    if (j->m_prev) j->m_prev->m_next = j->m_next;
    if (j->m_next) j->m_next->m_prev = j->m_prev;
    if (j == m_jointList) m_jointList = j->m_next;
    b2Body* bA = j->m_bodyA;
    b2Body* bB = j->m_bodyB;
    bA->SetAwake(true);
    bB->SetAwake(true);
    if (j->m_edgeA.prev) j->m_edgeA.prev->next = j->m_edgeA.next;
    if (j->m_edgeA.next) j->m_edgeA.next->prev = j->m_edgeA.prev;
    if (&j->m_edgeA == bA->m_jointList) bA->m_jointList = j->m_edgeA.next;
    if (j->m_edgeB.prev) j->m_edgeB.prev->next = j->m_edgeB.next;
    if (j->m_edgeB.next) j->m_edgeB.next->prev = j->m_edgeB.prev;
    if (&j->m_edgeB == bB->m_jointList) bB->m_jointList = j->m_edgeB.next;
    --m_jointCount;
    m_removedJoints.push_back(j->m_index);
    auto& dirty = m_dirtyJoints;
    dirty.erase(std::remove(dirty.begin(), dirty.end(), j), dirty.end());
    // A joint not yet on the device keeps its slot until the upload adds it
    // (device indices are creation order) and removes it again.
    if (j->m_index < m_uploadedJoints) {
        m_joints[j->m_index] = nullptr;
        delete j;
    } else {
        j->m_world = nullptr;
    }
    m_contactsStale = true;
}

void b2World::DownloadJoints() const {
    // setters' edits first, so the impulses read back carry their side
    // effects (a dropped limit impulse) as the reference's members do at once
    if (!m_locked && !m_dirtyJoints.empty()) (void)const_cast<b2World*>(this)->UploadJointEdits();
    if (!m_jointsStale || !m_ctx) return;
    m_jointsStale = false;
    const int n = m_uploadedJoints;
    if (n == 0) return;
    std::vector<float> st(4 * (size_t)n);
    if (b2g_get_joints(m_ctx, 0, n, st.data()) != B2G_OK) return;
    for (int i = 0; i < n; ++i) {
        b2Joint* j = m_joints[i];
        if (j == nullptr) continue;
        if (j->m_dirty) continue;
        for (int k = 0; k < 4; ++k) j->m_impulse[k] = st[4 * (size_t)i + k];
    }
}

static b2g_distance_def distanceDef(const b2DistanceJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_distance_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.length = d->GetLength();
    r.frequencyHz = d->GetFrequency();
    r.dampingRatio = d->GetDampingRatio();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_rope_def ropeDef(const b2RopeJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_rope_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.maxLength = d->GetMaxLength();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_weld_def weldDef(const b2WeldJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_weld_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.referenceAngle = d->GetReferenceAngle();
    r.frequencyHz = d->GetFrequency();
    r.dampingRatio = d->GetDampingRatio();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_friction_def frictionDef(const b2FrictionJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_friction_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.maxForce = d->GetMaxForce();
    r.maxTorque = d->GetMaxTorque();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_motor_def motorDef(const b2MotorJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_motor_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.linearOffset[0] = d->GetLinearOffset()[0];
    r.linearOffset[1] = d->GetLinearOffset()[1];
    r.angularOffset = d->GetAngularOffset();
    r.maxForce = d->GetMaxForce();
    r.maxTorque = d->GetMaxTorque();
    r.correctionFactor = d->GetCorrectionFactor();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_pulley_def pulleyDef(const b2PulleyJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_pulley_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.groundAnchorA[0] = d->GetGroundAnchorA()[0];
    r.groundAnchorA[1] = d->GetGroundAnchorA()[1];
    r.groundAnchorB[0] = d->GetGroundAnchorB()[0];
    r.groundAnchorB[1] = d->GetGroundAnchorB()[1];
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.lengthA = d->GetLengthA();
    r.lengthB = d->GetLengthB();
    r.ratio = d->GetRatio();
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_mouse_def mouseDef(const b2MouseJoint* d, int32_t bodyA, int32_t bodyB, float massB) {
    b2g_mouse_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.target[0] = d->GetTarget()[0];
    r.target[1] = d->GetTarget()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.maxForce = d->GetMaxForce();
    r.frequencyHz = d->GetFrequency();
    r.dampingRatio = d->GetDampingRatio();
    r.massB = massB;
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_prismatic_def prismaticDef(const b2PrismaticJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_prismatic_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.localAxisA[0] = d->GetLocalAxisA()[0];
    r.localAxisA[1] = d->GetLocalAxisA()[1];
    r.referenceAngle = d->GetReferenceAngle();
    r.lowerTranslation = d->GetLowerLimit();
    r.upperTranslation = d->GetUpperLimit();
    r.maxMotorForce = d->GetMaxMotorForce();
    r.motorSpeed = d->GetMotorSpeed();
    r.enableLimit = d->IsLimitEnabled() ? 1 : 0;
    r.enableMotor = d->IsMotorEnabled() ? 1 : 0;
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

static b2g_wheel_def wheelDef(const b2WheelJoint* d, int32_t bodyA, int32_t bodyB) {
    b2g_wheel_def r{};
    r.bodyA = bodyA;
    r.bodyB = bodyB;
    r.localAnchorA[0] = d->GetLocalAnchorA()[0];
    r.localAnchorA[1] = d->GetLocalAnchorA()[1];
    r.localAnchorB[0] = d->GetLocalAnchorB()[0];
    r.localAnchorB[1] = d->GetLocalAnchorB()[1];
    r.localAxisA[0] = d->GetLocalAxisA()[0];
    r.localAxisA[1] = d->GetLocalAxisA()[1];
    r.maxMotorTorque = d->GetMaxMotorTorque();
    r.motorSpeed = d->GetMotorSpeed();
    r.frequencyHz = d->GetSpringFrequencyHz();
    r.dampingRatio = d->GetSpringDampingRatio();
    r.enableMotor = d->IsMotorEnabled() ? 1 : 0;
    r.collideConnected = d->GetCollideConnected() ? 1 : 0;
    return r;
}

// b2GearJoint's constructor, REF/Dynamics/Joints/b2GearJoint.cpp:45-128: the
// four bodies, the two joints' anchors, axes and reference angles, and the
// constant from the bodies' current transforms.
void b2World::ResolveGear(b2GearJoint* g) {
    b2g_gear_def& d = g->m_def;
    d = b2g_gear_def{};
    float coordinate[2];
    for (int side = 0; side < 2; ++side) {
        b2Joint* jj = side == 0 ? g->m_joint1 : g->m_joint2;
        b2Body* other = jj->m_bodyA;   // body C or D
        b2Body* body = jj->m_bodyB;    // body A or B
        b2Vec2 anchorOther, anchorBody, axis(0.0f, 0.0f);
        float ref;
        const bool prismatic = jj->m_type == b2JointType::PRISMATIC_JOINT;
        if (!prismatic) {
            const b2RevoluteJoint* r = static_cast<const b2RevoluteJoint*>(jj);
            anchorOther = r->m_localAnchorA;
            anchorBody = r->m_localAnchorB;
            ref = r->m_referenceAngle;
            coordinate[side] = body->m_sweep.a - other->m_sweep.a - ref;
        } else {
            const b2PrismaticJoint* p = static_cast<const b2PrismaticJoint*>(jj);
            anchorOther = p->m_localAnchorA;
            anchorBody = p->m_localAnchorB;
            ref = p->m_referenceAngle;
            axis = p->m_localXAxisA;
            const b2Transform& xfB = body->GetTransform();
            const b2Transform& xfO = other->GetTransform();
            const b2Vec2 pO = anchorOther;
            const b2Vec2 pB = b2MulT(xfO.q, b2Mul(xfB.q, anchorBody) + (xfB.p - xfO.p));
            const b2Vec2 dd = pB - pO;
            coordinate[side] = dd[0] * axis[0] + dd[1] * axis[1];
        }
        if (side == 0) {
            d.bodyA = body->m_index; d.bodyC = other->m_index; d.prismaticA = prismatic ? 1 : 0;
            d.localAnchorA[0] = anchorBody[0]; d.localAnchorA[1] = anchorBody[1];
            d.localAnchorC[0] = anchorOther[0]; d.localAnchorC[1] = anchorOther[1];
            d.localAxisC[0] = axis[0]; d.localAxisC[1] = axis[1];
            d.referenceAngleA = ref;
            g->m_bodyA = body;
            g->m_localAnchorA = anchorBody;
        } else {
            d.bodyB = body->m_index; d.bodyD = other->m_index; d.prismaticB = prismatic ? 1 : 0;
            d.localAnchorB[0] = anchorBody[0]; d.localAnchorB[1] = anchorBody[1];
            d.localAnchorD[0] = anchorOther[0]; d.localAnchorD[1] = anchorOther[1];
            d.localAxisD[0] = axis[0]; d.localAxisD[1] = axis[1];
            d.referenceAngleB = ref;
            g->m_bodyB = body;
            g->m_localAnchorB = anchorBody;
        }
    }
    d.ratio = g->m_ratio;
    d.constant = coordinate[0] + g->m_ratio * coordinate[1];
    d.collideConnected = g->m_collideConnected ? 1 : 0;
}

static b2g_revolute_def revoluteDef(const b2RevoluteJoint* r, int32_t bodyA, int32_t bodyB) {
    b2g_revolute_def d{};
    d.bodyA = bodyA;
    d.bodyB = bodyB;
    d.localAnchorA[0] = r->GetLocalAnchorA()[0];
    d.localAnchorA[1] = r->GetLocalAnchorA()[1];
    d.localAnchorB[0] = r->GetLocalAnchorB()[0];
    d.localAnchorB[1] = r->GetLocalAnchorB()[1];
    d.referenceAngle = r->GetReferenceAngle();
    d.lowerAngle = r->GetLowerLimit();
    d.upperAngle = r->GetUpperLimit();
    d.maxMotorTorque = r->GetMaxMotorTorque();
    d.motorSpeed = r->GetMotorSpeed();
    d.enableLimit = r->IsLimitEnabled() ? 1 : 0;
    d.enableMotor = r->IsMotorEnabled() ? 1 : 0;
    d.collideConnected = r->GetCollideConnected() ? 1 : 0;
    return d;
}

void b2World::SetGravity(const b2Vec2& g) {
    m_gravity = g;
    m_worldDirty = true;
}

// b2World::SetAllowSleeping, REF/Dynamics/b2World.cpp:370-386
void b2World::SetAllowSleeping(bool flag) {
    if (flag == m_allowSleep) return;
    m_allowSleep = flag;
    m_worldDirty = true;
    if (!flag) {
        for (b2Body* b = m_bodyList; b; b = b->m_next) b->SetAwake(true);
    }
}

void b2World::ClearForces() {
    for (b2Body* b = m_bodyList; b; b = b->m_next) {
        b->Touch();
        b->m_force.SetZero();
        b->m_torque = 0.0f;
        b->MarkDirty();
    }
}

static b2g_body_state stateOf(const b2Body* b, int type, int flags, const b2Transform& xf, const b2Sweep& sw,
                              const b2Vec2& v, float w, const b2Vec2& f, float t, float invMass, float invI,
                              float ld, float ad, float gs, float sleepTime) {
    b2g_body_state s{};
    (void)b;
    s.type = type;
    s.flags = flags;
    s.xf[0] = xf.p[0]; s.xf[1] = xf.p[1]; s.xf[2] = xf.q.s; s.xf[3] = xf.q.c;
    s.lc[0] = sw.localCenter[0]; s.lc[1] = sw.localCenter[1];
    s.c0[0] = sw.c0[0]; s.c0[1] = sw.c0[1]; s.a0 = sw.a0;
    s.c[0] = sw.c[0]; s.c[1] = sw.c[1]; s.a = sw.a;
    s.v[0] = v[0]; s.v[1] = v[1]; s.w = w;
    s.force[0] = f[0]; s.force[1] = f[1]; s.torque = t;
    s.invMass = invMass; s.invI = invI;
    s.linearDamping = ld; s.angularDamping = ad; s.gravityScale = gs; s.sleepTime = sleepTime;
    return s;
}

int32_t b2World::Upload() {
    if (!m_ctx) return B2G_ERR_HIP;
    int rc = B2G_OK;
    if (m_worldDirty) {
        b2g_world_def wd{};
        wd.gravity[0] = m_gravity[0];
        wd.gravity[1] = m_gravity[1];
        wd.allowSleep = m_allowSleep;
        wd.warmStarting = m_warmStarting;
        wd.continuous = m_continuousPhysics ? 1 : 0;
        wd.subStepping = m_subStepping ? 1 : 0;
        wd.autoClearForces = m_autoClearForces;
        rc = b2g_set_world(m_ctx, &wd);
        if (rc != B2G_OK) return rc;
        m_worldDirty = false;
    }
    auto toState = [](const b2Body* b) {
        return stateOf(b, (int)b->m_type, b->m_flags, b->m_xf, b->m_sweep, b->m_linearVelocity,
                       b->m_angularVelocity, b->m_force, b->m_torque, b->m_invMass, b->m_invI, b->m_linearDamping,
                       b->m_angularDamping, b->m_gravityScale, b->m_sleepTime);
    };
    if (m_anyDirty) {
        for (int i = 0; i < m_uploadedBodies; ++i) {
            b2Body* b = m_bodies[i];
            if (!b->m_dirty) continue;
            // contiguous dirty runs in one call
            int j = i;
            while (j < m_uploadedBodies && m_bodies[j]->m_dirty) ++j;
            std::vector<b2g_body_state> st;
            st.reserve(j - i);
            for (int k = i; k < j; ++k) {
                st.push_back(toState(m_bodies[k]));
                m_bodies[k]->m_dirty = false;
            }
            rc = b2g_set_bodies(m_ctx, i, j - i, st.data());
            if (rc != B2G_OK) return rc;
            i = j - 1;
        }
        m_anyDirty = false;
    }
    const int nb = (int)m_bodies.size();
    if (nb > m_uploadedBodies) {
        std::vector<b2g_body_state> st;
        st.reserve(nb - m_uploadedBodies);
        for (int i = m_uploadedBodies; i < nb; ++i) st.push_back(toState(m_bodies[i]));
        int first = 0;
        rc = b2g_add_bodies(m_ctx, st.data(), (int)st.size(), &first);
        if (rc != B2G_OK) return rc;
        if (m_subWorldCount > 0) {
            std::vector<int32_t> ids;
            ids.reserve(nb - m_uploadedBodies);
            for (int i = m_uploadedBodies; i < nb; ++i) ids.push_back(m_bodies[i]->m_subWorld);
            rc = b2g_set_body_worlds(m_ctx, first, (int)ids.size(), ids.data());
            if (rc != B2G_OK) return rc;
        }
        m_uploadedBodies = nb;
    }
    const int nf = (int)m_fixtures.size();
    if (nf > m_uploadedFixtures) {
        std::vector<b2g_fixture_def> defs;
        defs.reserve(nf - m_uploadedFixtures);
        for (int i = m_uploadedFixtures; i < nf; ++i) {
            const b2Fixture* f = m_fixtures[i];
            b2g_fixture_def d{};
            d.body = f->m_body->m_index;
            d.sensor = f->m_isSensor ? 1 : 0;
            d.friction = f->m_friction;
            d.restitution = f->m_restitution;
            d.density = f->m_density;
            d.categoryBits = f->m_filter.categoryBits;
            d.maskBits = f->m_filter.maskBits;
            d.groupIndex = f->m_filter.groupIndex;
            d.shape = toDeviceShape(f->m_shape.get(), i - f->m_index);
            defs.push_back(d);
        }
        int first = 0;
        rc = b2g_add_fixtures(m_ctx, defs.data(), (int)defs.size(), &first);
        if (rc != B2G_OK) return rc;
        m_uploadedFixtures = nf;
    }
    if (!m_updatedFixtures.empty()) {
        std::vector<int32_t> ids;
        std::vector<b2g_fixture_def> defs;
        for (int32_t id : m_updatedFixtures) {
            if (id >= m_uploadedFixtures) continue;
            const b2Fixture* f = m_fixtures[id];
            b2g_fixture_def d{};
            d.body = f->m_body->m_index;
            d.sensor = f->m_isSensor ? 1 : 0;
            d.friction = f->m_friction;
            d.restitution = f->m_restitution;
            d.density = f->m_density;
            d.categoryBits = f->m_filter.categoryBits;
            d.maskBits = f->m_filter.maskBits;
            d.groupIndex = f->m_filter.groupIndex;
            ids.push_back(id);
            defs.push_back(d);
        }
        rc = b2g_update_fixtures(m_ctx, ids.data(), defs.data(), (int)ids.size());
        if (rc != B2G_OK) return rc;
        m_updatedFixtures.clear();
    }
    if (!m_removedFixtures.empty()) {
        std::vector<int32_t> ids;
        for (int32_t id : m_removedFixtures)
            if (id < m_uploadedFixtures) ids.push_back(id);
        rc = b2g_remove_fixtures(m_ctx, ids.data(), (int)ids.size());
        if (rc != B2G_OK) return rc;
        m_removedFixtures.clear();
    }
    if (!m_restoredFixtures.empty()) {
        rc = b2g_restore_fixtures(m_ctx, m_restoredFixtures.data(), (int)m_restoredFixtures.size());
        if (rc != B2G_OK) return rc;
        m_restoredFixtures.clear();
    }
    if (!m_typeChanged.empty()) {
        rc = b2g_drop_body_contacts(m_ctx, m_typeChanged.data(), (int)m_typeChanged.size());
        if (rc != B2G_OK) return rc;
        m_typeChanged.clear();
    }
    if (m_proxyIdsDirty && m_subWorldCount == 0) {
        rc = b2g_set_proxy_ids(m_ctx, m_proxyIds.data(), (int)m_proxyIds.size());
        if (rc != B2G_OK) return rc;
        m_proxyIdsDirty = false;
    }
    const int nj = (int)m_joints.size();
    if (nj > m_uploadedJoints) {
        // runs of one type at a time, so device indices follow creation order
        for (int i = m_uploadedJoints; i < nj;) {
            const b2JointType type = m_joints[i]->m_type;
            int k = i;
            while (k < nj && m_joints[k]->m_type == type) ++k;
            int first = 0;
            if (type == b2JointType::DISTANCE_JOINT) {
                std::vector<b2g_distance_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2DistanceJoint* d = static_cast<const b2DistanceJoint*>(m_joints[q]);
                    defs.push_back(distanceDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_distance_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::ROPE_JOINT) {
                std::vector<b2g_rope_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2RopeJoint* d = static_cast<const b2RopeJoint*>(m_joints[q]);
                    defs.push_back(ropeDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_rope_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::WELD_JOINT) {
                std::vector<b2g_weld_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2WeldJoint* d = static_cast<const b2WeldJoint*>(m_joints[q]);
                    defs.push_back(weldDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_weld_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::FRICTION_JOINT) {
                std::vector<b2g_friction_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2FrictionJoint* d = static_cast<const b2FrictionJoint*>(m_joints[q]);
                    defs.push_back(frictionDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_friction_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::MOTOR_JOINT) {
                std::vector<b2g_motor_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2MotorJoint* d = static_cast<const b2MotorJoint*>(m_joints[q]);
                    defs.push_back(motorDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_motor_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::PULLEY_JOINT) {
                std::vector<b2g_pulley_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2PulleyJoint* d = static_cast<const b2PulleyJoint*>(m_joints[q]);
                    defs.push_back(pulleyDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_pulley_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::MOUSE_JOINT) {
                std::vector<b2g_mouse_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2MouseJoint* d = static_cast<const b2MouseJoint*>(m_joints[q]);
                    defs.push_back(mouseDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index, d->m_bodyB->m_mass));
                }
                rc = b2g_add_mouse_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::PRISMATIC_JOINT) {
                std::vector<b2g_prismatic_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2PrismaticJoint* d = static_cast<const b2PrismaticJoint*>(m_joints[q]);
                    defs.push_back(prismaticDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_prismatic_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::GEAR_JOINT) {
                std::vector<b2g_gear_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2GearJoint* g = static_cast<const b2GearJoint*>(m_joints[q]);
                    b2g_gear_def d = g->m_def;
                    d.bodyA = g->m_joint1->m_bodyB->m_index; d.bodyC = g->m_joint1->m_bodyA->m_index;
                    d.bodyB = g->m_joint2->m_bodyB->m_index; d.bodyD = g->m_joint2->m_bodyA->m_index;
                    defs.push_back(d);
                }
                rc = b2g_add_gear_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else if (type == b2JointType::WHEEL_JOINT) {
                std::vector<b2g_wheel_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2WheelJoint* d = static_cast<const b2WheelJoint*>(m_joints[q]);
                    defs.push_back(wheelDef(d, d->m_bodyA->m_index, d->m_bodyB->m_index));
                }
                rc = b2g_add_wheel_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            } else {
                std::vector<b2g_revolute_def> defs;
                for (int q = i; q < k; ++q) {
                    const b2RevoluteJoint* r = static_cast<const b2RevoluteJoint*>(m_joints[q]);
                    defs.push_back(revoluteDef(r, r->m_bodyA->m_index, r->m_bodyB->m_index));
                }
                rc = b2g_add_joints(m_ctx, defs.data(), (int)defs.size(), &first);
            }
            if (rc != B2G_OK) return rc;
            i = k;
        }
        for (int i = m_uploadedJoints; i < nj; ++i) {
            if (m_joints[i]->m_world == nullptr) {  // destroyed before its upload
                delete m_joints[i];
                m_joints[i] = nullptr;
            }
        }
        m_uploadedJoints = nj;
    }
    if (!m_removedJoints.empty()) {
        rc = b2g_remove_joints(m_ctx, m_removedJoints.data(), (int)m_removedJoints.size());
        if (rc != B2G_OK) return rc;
        m_removedJoints.clear();
    }
    if ((rc = UploadJointEdits()) != B2G_OK) return rc;
    return B2G_OK;
}

int32_t b2World::UploadJointEdits() {
    if (m_dirtyJoints.empty() || !m_ctx) return B2G_OK;
    int rc = B2G_OK;
    for (b2Joint* j : m_dirtyJoints) {
        j->m_dirty = false;
        if (j->m_index >= m_uploadedJoints) continue;
        if (j->m_type == b2JointType::DISTANCE_JOINT) {
            const b2DistanceJoint* d = static_cast<const b2DistanceJoint*>(j);
            rc = b2g_set_distance_joint(m_ctx, j->m_index, d->m_length, d->m_frequencyHz, d->m_dampingRatio);
        } else if (j->m_type == b2JointType::ROPE_JOINT) {
            rc = b2g_set_rope_joint(m_ctx, j->m_index, static_cast<const b2RopeJoint*>(j)->m_maxLength);
        } else if (j->m_type == b2JointType::WELD_JOINT) {
            const b2WeldJoint* d = static_cast<const b2WeldJoint*>(j);
            rc = b2g_set_weld_joint(m_ctx, j->m_index, d->m_frequencyHz, d->m_dampingRatio);
        } else if (j->m_type == b2JointType::FRICTION_JOINT) {
            const b2FrictionJoint* d = static_cast<const b2FrictionJoint*>(j);
            rc = b2g_set_friction_joint(m_ctx, j->m_index, d->m_maxForce, d->m_maxTorque);
        } else if (j->m_type == b2JointType::MOTOR_JOINT) {
            const b2MotorJoint* m = static_cast<const b2MotorJoint*>(j);
            const b2g_motor_def d = motorDef(m, m->m_bodyA->m_index, m->m_bodyB->m_index);
            rc = b2g_set_motor_joint(m_ctx, j->m_index, &d);
        } else if (j->m_type == b2JointType::MOUSE_JOINT) {
            const b2MouseJoint* m = static_cast<const b2MouseJoint*>(j);
            const b2g_mouse_def d = mouseDef(m, m->m_bodyA->m_index, m->m_bodyB->m_index, m->m_bodyB->m_mass);
            rc = b2g_set_mouse_joint(m_ctx, j->m_index, &d);
        } else if (j->m_type == b2JointType::PULLEY_JOINT) {
            rc = B2G_OK;  // no setters
        } else if (j->m_type == b2JointType::PRISMATIC_JOINT) {
            const b2PrismaticJoint* p = static_cast<const b2PrismaticJoint*>(j);
            const b2g_prismatic_def d = prismaticDef(p, p->m_bodyA->m_index, p->m_bodyB->m_index);
            rc = b2g_set_prismatic_joint(m_ctx, j->m_index, &d);
        } else if (j->m_type == b2JointType::GEAR_JOINT) {
            rc = b2g_set_gear_joint(m_ctx, j->m_index, static_cast<const b2GearJoint*>(j)->m_ratio);
        } else if (j->m_type == b2JointType::WHEEL_JOINT) {
            const b2WheelJoint* p = static_cast<const b2WheelJoint*>(j);
            const b2g_wheel_def d = wheelDef(p, p->m_bodyA->m_index, p->m_bodyB->m_index);
            rc = b2g_set_wheel_joint(m_ctx, j->m_index, &d);
        } else {
            const b2RevoluteJoint* r = static_cast<const b2RevoluteJoint*>(j);
            const b2g_revolute_def d = revoluteDef(r, r->m_bodyA->m_index, r->m_bodyB->m_index);
            rc = b2g_set_joint(m_ctx, j->m_index, &d);
        }
        if (rc != B2G_OK) return rc;
    }
    m_dirtyJoints.clear();
    m_jointsStale = true;  // impulses a setter reset on the device
    return B2G_OK;
}

int32_t b2World::Flush() {
    const int rc = Upload();
    if (rc != B2G_OK && m_ctx) m_lastError = b2g_last_error(m_ctx);
    return rc;
}

void b2World::Download() const {
    if (!m_hostStale || !m_ctx) return;
    m_hostStale = false;
    const int n = m_uploadedBodies;
    if (n == 0) return;
    std::vector<b2g_body_state> st(n);
    if (b2g_get_bodies(m_ctx, 0, n, st.data()) != B2G_OK) return;
    for (int i = 0; i < n; ++i) {
        b2Body* b = m_bodies[i];
        if (b->m_dirty) continue;
        const b2g_body_state& s = st[i];
        b->m_flags = (b->m_flags & ~0x2) | (s.flags & 0x2);
        b->m_xf.p = b2Vec2(s.xf[0], s.xf[1]);
        b->m_xf.q.s = s.xf[2];
        b->m_xf.q.c = s.xf[3];
        b->m_sweep.c0 = b2Vec2(s.c0[0], s.c0[1]);
        b->m_sweep.a0 = s.a0;
        b->m_sweep.c = b2Vec2(s.c[0], s.c[1]);
        b->m_sweep.a = s.a;
        b->m_linearVelocity = b2Vec2(s.v[0], s.v[1]);
        b->m_angularVelocity = s.w;
        b->m_force = b2Vec2(s.force[0], s.force[1]);
        b->m_torque = s.torque;
        b->m_sleepTime = s.sleepTime;
    }
}

// b2World::Step, Box2D/Dynamics/b2World.cpp:888-954, on the GPU.
void b2World::Step(float dt, int32_t velocityIterations, int32_t positionIterations) {
    const auto t0 = std::chrono::steady_clock::now();
    m_lastStatus = B2G_OK;
    m_lastError.clear();
    if (!m_ctx) {
        m_lastStatus = B2G_ERR_HIP;
        m_lastError = "no HIP device";
        return;
    }
    m_locked = true;
    int rc = Upload();
    if (rc == B2G_OK) {
        b2g_step_params p{};
        p.dt = dt;
        p.velocityIterations = velocityIterations;
        p.positionIterations = positionIterations;
        p.mode = m_options.mode;
        p.learnedWarmStart = m_options.learnedWarmStart ? 1 : 0;
        p.residualThreshold = m_options.residualThreshold;
        if (m_contactListener) {
            // Begin/End/PreSolve inside the step, as Collide calls them
            rc = CollideWithListener(&p);
            if (rc == B2G_OK) rc = b2g_step_solve(m_ctx);
        } else {
            rc = b2g_step(m_ctx, &p);
        }
    }
    m_locked = false;
    if (rc != B2G_OK) {
        m_lastStatus = rc;
        m_lastError = b2g_last_error(m_ctx);
        return;
    }
    m_hostStale = true;
    m_contactsStale = true;
    m_jointsStale = true;
    b2g_step_stats st{};
    b2g_get_stats(m_ctx, &st);
    m_contactCount = st.contacts;
    if (m_contactListener) ReplayEvents(true);
    const auto t1 = std::chrono::steady_clock::now();
    m_profile.step = std::chrono::duration<float, std::milli>(t1 - t0).count();
}

int32_t b2World::Autotune(float dt, int32_t velocityIterations, int32_t positionIterations, int32_t rounds) {
    if (!m_ctx) return B2G_ERR_HIP;
    if (IsLocked()) return B2G_ERR_STATE;
    int rc = Upload();
    if (rc == B2G_OK) {
        b2g_step_params p{};
        p.dt = dt;
        p.velocityIterations = velocityIterations;
        p.positionIterations = positionIterations;
        p.mode = m_options.mode;
        p.learnedWarmStart = m_options.learnedWarmStart ? 1 : 0;
        p.residualThreshold = m_options.residualThreshold;
        m_locked = true;
        rc = b2g_autotune(m_ctx, &p, rounds, nullptr);
        m_locked = false;
    }
    // the tuning steps are ordinary steps: mirrors and contacts are stale
    m_hostStale = true;
    m_contactsStale = true;
    m_jointsStale = true;
    b2g_step_stats st{};
    if (b2g_get_stats(m_ctx, &st) == B2G_OK) m_contactCount = st.contacts;
    if (rc != B2G_OK) {
        m_lastStatus = rc;
        m_lastError = b2g_last_error(m_ctx);
    }
    return rc;
}

b2Contact* b2World::GetContactList() {
    if (m_contactsStale && m_ctx) {
        int n = 0;
        b2g_get_contacts(m_ctx, nullptr, nullptr, 0, &n);
        std::vector<int32_t> meta(4 * (size_t)n);
        std::vector<b2g_manifold> man(n);
        b2g_get_contacts(m_ctx, meta.data(), man.data(), n, &n);
        std::vector<float> mat(2 * (size_t)n), tangent(n);
        const bool haveMat = n > 0 && b2g_get_contact_materials(m_ctx, mat.data(), n) == B2G_OK;
        const bool haveTangent = n > 0 && b2g_get_contact_tangent_speeds(m_ctx, tangent.data(), n) == B2G_OK;
        m_contacts.assign(n, b2Contact());
        // Newest contacts first, as the reference's contact list.
        for (int i = 0; i < n; ++i) {
            const int src = n - 1 - i;
            b2Contact& c = m_contacts[i];
            c.m_fixtureA = m_fixtures[meta[4 * src]];
            c.m_fixtureB = m_fixtures[meta[4 * src + 1]];
            c.m_indexA = meta[4 * src] - c.m_fixtureA->m_index;
            c.m_indexB = meta[4 * src + 1] - c.m_fixtureB->m_index;
            c.m_touching = (meta[4 * src + 2] & 1) != 0;
            c.m_enabled = (meta[4 * src + 2] & 2) != 0;
            c.m_friction = haveMat ? mat[2 * src] : sqrtf(c.m_fixtureA->m_friction * c.m_fixtureB->m_friction);
            c.m_restitution = haveMat ? mat[2 * src + 1]
                                      : std::max(c.m_fixtureA->m_restitution, c.m_fixtureB->m_restitution);
            c.m_tangentSpeed = haveTangent ? tangent[src] : 0.0f;
            const b2g_manifold& m = man[src];
            c.m_manifold.pointCount = m.pointCount;
            c.m_manifold.type = (b2Manifold::Type)m.type;
            c.m_manifold.localNormal = b2Vec2(m.localNormal.x, m.localNormal.y);
            c.m_manifold.localPoint = b2Vec2(m.localPoint.x, m.localPoint.y);
            for (int k = 0; k < 2; ++k) {
                c.m_manifold.points[k].localPoint = b2Vec2(m.points[k].localPoint.x, m.points[k].localPoint.y);
                c.m_manifold.points[k].normalImpulse = m.points[k].normalImpulse;
                c.m_manifold.points[k].tangentImpulse = m.points[k].tangentImpulse;
                c.m_manifold.points[k].id.key = m.points[k].id;
            }
            c.m_next = i + 1 < n ? &m_contacts[i + 1] : nullptr;
        }
        // Each body's contact list is the world's order restricted to the
        // body (both are newest first in the reference): built oldest first,
        // each edge prepended.
        for (b2Body* b : m_bodies)
            if (b) b->m_contactList = nullptr;
        for (int i = n - 1; i >= 0; --i) {
            b2Contact& c = m_contacts[i];
            b2Body* bA = c.m_fixtureA->m_body;
            b2Body* bB = c.m_fixtureB->m_body;
            b2ContactEdge* edges[2] = {&c.m_nodeA, &c.m_nodeB};
            b2Body* owners[2] = {bA, bB};
            for (int k = 0; k < 2; ++k) {
                b2ContactEdge* e = edges[k];
                e->contact = &c;
                e->other = owners[1 - k];
                e->prev = nullptr;
                e->next = owners[k]->m_contactList;
                if (e->next) e->next->prev = e;
                owners[k]->m_contactList = e;
            }
        }
        m_contactsStale = false;
    }
    return m_contacts.empty() ? nullptr : &m_contacts[0];
}

b2ContactEdge* b2Body::GetContactList() {
    m_world->GetContactList();
    return m_contactList;
}

// REF/Dynamics/b2Body.cpp:115-172. The reference also re-synchronises a
// body made static from its (now equal) sweep ends: that box lies inside
// the fat box its last synchronisation left, so no proxy moves and the
// device has nothing to do for it.
void b2Body::SetType(b2BodyType type) {
    if (m_world->IsLocked() || m_destroyed) return;
    Touch();
    if (m_type == type) return;
    m_type = type;
    ResetMassData();
    if (m_type == b2BodyType::STATIC_BODY) {
        m_linearVelocity.SetZero();
        m_angularVelocity = 0.0f;
        m_sweep.a0 = m_sweep.a;
        m_sweep.c0 = m_sweep.c;
    }
    SetAwake(true);
    m_force.SetZero();
    m_torque = 0.0f;
    MarkDirty();
    m_world->m_typeChanged.push_back(m_index);
    m_world->m_contactsStale = true;
}

// REF/Dynamics/b2Body.cpp:456-496: fixtures in list order (newest first),
// each one's proxies in child order.
void b2Body::SetActive(bool flag) {
    if (m_world->IsLocked() || m_destroyed) return;
    Touch();
    if (flag == IsActive()) return;
    b2World* w = m_world;
    if (flag) m_flags |= e_activeFlag;
    else m_flags &= ~e_activeFlag;
    for (b2Fixture* f = m_fixtureList; f; f = f->m_next)
        for (int32_t c = 0; c < f->m_shape->GetChildCount(); ++c) {
            const int32_t slot = f->m_index + c;
            int32_t& id = w->m_proxyIds[slot];
            if (flag) {
                id = w->CreateProxyId();
                w->m_restoredFixtures.push_back(slot);
            } else if (id >= 0) {
                w->DestroyProxyId(id);
                id = -1;
                w->m_removedFixtures.push_back(slot);
            }
        }
    // ids taken out of creation order from here on
    w->m_proxyIdsReused = w->m_proxyIdsDirty = true;
    w->m_contactsStale = true;
    MarkDirty();
}

void b2Body::SetBullet(bool flag) {
    Touch();
    if (flag) m_flags |= e_bulletFlag;
    else m_flags &= ~e_bulletFlag;
    MarkDirty();
}

// REF/Dynamics/b2Body.cpp:438-457
void b2Body::SetFixedRotation(bool flag) {
    Touch();
    const bool status = (m_flags & e_fixedRotationFlag) != 0;
    if (status == flag) return;
    if (flag) m_flags |= e_fixedRotationFlag;
    else m_flags &= ~e_fixedRotationFlag;
    m_angularVelocity = 0.0f;
    ResetMassData();
}

// REF/Dynamics/b2Body.cpp:354-394
void b2Body::SetMassData(const b2MassData* data) {
    if (m_world->IsLocked() || data == nullptr || m_type != b2BodyType::DYNAMIC_BODY) return;
    Touch();
    m_invMass = 0.0f;
    m_I = 0.0f;
    m_invI = 0.0f;
    m_mass = data->mass;
    if (m_mass <= 0.0f) m_mass = 1.0f;
    m_invMass = 1.0f / m_mass;
    if (data->I > 0.0f && (m_flags & e_fixedRotationFlag) == 0) {
        m_I = data->I - m_mass * b2Dot(data->center, data->center);
        m_invI = 1.0f / m_I;
    }
    const b2Vec2 oldCenter = m_sweep.c;
    m_sweep.localCenter = data->center;
    m_sweep.c0 = m_sweep.c = b2Mul(m_xf, m_sweep.localCenter);
    m_linearVelocity += b2Cross(m_angularVelocity, m_sweep.c - oldCenter);
    MarkDirty();
}

// b2MixFriction, b2MixRestitution (REF/Dynamics/Contacts/b2Contact.h:30-41)
void b2Contact::ResetFriction() { m_friction = sqrtf(m_fixtureA->GetFriction() * m_fixtureB->GetFriction()); }
void b2Contact::ResetRestitution() {
    m_restitution = std::max(m_fixtureA->GetRestitution(), m_fixtureB->GetRestitution());
}

void b2Contact::GetWorldManifold(b2WorldManifold* wm) const {
    const b2Body* bA = m_fixtureA->GetBody();
    const b2Body* bB = m_fixtureB->GetBody();
    wm->Initialize(&m_manifold, bA->GetTransform(), m_fixtureA->GetShape()->m_radius, bB->GetTransform(),
                   m_fixtureB->GetShape()->m_radius);
}

// Contact events after a Step: BeginContact for fixture pairs that started
// touching, EndContact for pairs that stopped (or whose contact ended), and
// PostSolve with each touching contact's stored impulses.
// b2ContactManager::Collide's listener calls (REF/Dynamics/b2ContactManager.cpp:
// 110-175): over the contact list, newest first, a contact destroyed by the
// filter or by fat boxes that stopped overlapping reports EndContact if it was
// touching (b2ContactManager::Destroy, :72-104); an updated contact reports
// BeginContact or EndContact when its touching state changed and PreSolve
// with its old manifold while touching (b2Contact::Update,
// REF/Dynamics/Contacts/b2Contact.cpp:164-256). PreSolve's edits go back
// to the device before the solve.
int32_t b2World::CollideWithListener(const void* params) {
    const b2g_step_params* p = static_cast<const b2g_step_params*>(params);
    int rc = b2g_step_collide(m_ctx, p);
    if (rc != B2G_OK) return rc;
    m_hostStale = true;  // wakes of the narrowphase
    int n = 0;
    b2g_get_contacts(m_ctx, nullptr, nullptr, 0, &n);
    std::vector<int32_t> meta(4 * (size_t)n);
    std::vector<b2g_manifold> man(n), old(n);
    std::vector<float> mat(2 * (size_t)n), tangent(n);
    if (n > 0) {
        if ((rc = b2g_get_contacts(m_ctx, meta.data(), man.data(), n, &n)) != B2G_OK) return rc;
        if ((rc = b2g_get_old_manifolds(m_ctx, old.data(), n, nullptr)) != B2G_OK) return rc;
        if ((rc = b2g_get_contact_materials(m_ctx, mat.data(), n)) != B2G_OK) return rc;
        if ((rc = b2g_get_contact_tangent_speeds(m_ctx, tangent.data(), n)) != B2G_OK) return rc;
    }
    auto toHost = [](const b2g_manifold& m, b2Manifold& out) {
        out.pointCount = m.pointCount;
        out.type = (b2Manifold::Type)m.type;
        out.localNormal = b2Vec2(m.localNormal.x, m.localNormal.y);
        out.localPoint = b2Vec2(m.localPoint.x, m.localPoint.y);
        for (int k = 0; k < 2; ++k) {
            out.points[k].localPoint = b2Vec2(m.points[k].localPoint.x, m.points[k].localPoint.y);
            out.points[k].normalImpulse = m.points[k].normalImpulse;
            out.points[k].tangentImpulse = m.points[k].tangentImpulse;
            out.points[k].id.key = m.points[k].id;
        }
    };
    std::vector<int32_t> ids, enabled, tIds;
    std::vector<float> fr, re, tSpeeds;
    for (int i = n - 1; i >= 0; --i) {
        const int flags = meta[4 * i + 2];
        const bool dead = (flags & 4) != 0, updated = (flags & 8) != 0, touching = (flags & 1) != 0;
        if (!dead && !updated) continue;
        b2Contact c;
        c.m_fixtureA = m_fixtures[meta[4 * i]];
        c.m_fixtureB = m_fixtures[meta[4 * i + 1]];
        c.m_indexA = meta[4 * i] - c.m_fixtureA->m_index;
        c.m_indexB = meta[4 * i + 1] - c.m_fixtureB->m_index;
        c.m_friction = mat[2 * i];
        c.m_restitution = mat[2 * i + 1];
        c.m_tangentSpeed = tangent[i];
        c.m_enabled = true;
        if (dead) {
            c.m_touching = touching;
            toHost(old[i], c.m_manifold);
            if (touching) m_contactListener->EndContact(&c);
            continue;
        }
        const bool wasTouching = (flags & 16) != 0;
        const bool sensor = c.m_fixtureA->m_isSensor || c.m_fixtureB->m_isSensor;
        c.m_touching = touching;
        toHost(man[i], c.m_manifold);
        if (!wasTouching && touching) m_contactListener->BeginContact(&c);
        if (wasTouching && !touching) m_contactListener->EndContact(&c);
        if (touching && !sensor) {
            b2Manifold oldManifold;
            toHost(old[i], oldManifold);
            m_contactListener->PreSolve(&c, &oldManifold);
            if (!c.m_enabled || c.m_friction != mat[2 * i] || c.m_restitution != mat[2 * i + 1]) {
                ids.push_back(i);
                enabled.push_back(c.m_enabled ? 1 : 0);
                fr.push_back(c.m_friction);
                re.push_back(c.m_restitution);
            }
            if (c.m_tangentSpeed != tangent[i]) {
                tIds.push_back(i);
                tSpeeds.push_back(c.m_tangentSpeed);
            }
        }
    }
    if (!ids.empty())
        rc = b2g_override_contacts(m_ctx, ids.data(), enabled.data(), fr.data(), re.data(), (int)ids.size());
    if (rc == B2G_OK && !tIds.empty())
        rc = b2g_set_contact_tangent_speeds(m_ctx, tIds.data(), tSpeeds.data(), (int)tIds.size());
    // The order in which the solve's Report will call PostSolve: the island
    // search over the contacts as the collide and PreSolve left them, with
    // indices as the table has them after the dead contacts leave it.
    {
        std::vector<uint8_t> on(n, 0);
        for (int i = 0; i < n; ++i) on[i] = (meta[4 * i + 2] & 2) != 0;
        for (size_t k = 0; k < ids.size(); ++k) on[ids[k]] = (uint8_t)enabled[k];
        IslandContactOrder(meta, on, &m_postSolveOrder);
        std::vector<int32_t> live(n + 1, 0);
        for (int i = 0; i < n; ++i) live[i + 1] = live[i] + ((meta[4 * i + 2] & 4) == 0 ? 1 : 0);
        for (int32_t& t : m_postSolveOrder) t = live[t];
    }
    // touching set as the listener has seen it; what the TOI phase changes
    // after this is reported once the step is over
    m_touchingKeys.clear();
    for (int i = 0; i < n; ++i) {
        const int flags = meta[4 * i + 2];
        if ((flags & 4) == 0 && (flags & 1) != 0) m_touchingKeys.push_back(pairKeyOf(meta[4 * i], meta[4 * i + 1]));
    }
    std::sort(m_touchingKeys.begin(), m_touchingKeys.end());
    return rc;
}

// b2World::Solve's island search (REF/Dynamics/b2World.cpp:388-565) over the
// state after the narrowphase: seeds in body-list order (newest first) that
// are awake, active and not static; a depth-first walk over each body's
// contact edges (newest first) and joint edges (newest first), taking the
// touching, enabled, non-sensor contacts and the joints whose other body is
// active, never walking on from a static body, whose island flag is cleared
// after its island so that it can join the next. The island contacts in the
// order the walk adds them are the order of b2Island::Report's PostSolve
// calls (REF/Dynamics/b2Island.cpp:517-540), island after island. meta is
// b2g_get_contacts' table (creation order); on[t] whether contact t is
// enabled; order receives table indices.
void b2World::IslandContactOrder(const std::vector<int32_t>& meta, const std::vector<uint8_t>& on,
                                 std::vector<int32_t>* order) const {
    order->clear();
    Download();  // awake flags as the narrowphase left them
    const int n = (int)on.size();
    const int nb = (int)m_bodies.size();
    const int nj = (int)m_joints.size();
    auto bodyOf = [&](int f) { return m_fixtures[f]->m_body->m_index; };
    // contact edges per body, newest first
    std::vector<int32_t> start(nb + 1, 0), edge;
    for (int t = 0; t < n; ++t) {
        if (meta[4 * t + 2] & 4) continue;
        ++start[bodyOf(meta[4 * t]) + 1];
        ++start[bodyOf(meta[4 * t + 1]) + 1];
    }
    for (int b = 0; b < nb; ++b) start[b + 1] += start[b];
    edge.resize(start[nb]);
    std::vector<int32_t> cursor(start.begin(), start.end() - 1);
    for (int t = n - 1; t >= 0; --t) {
        if (meta[4 * t + 2] & 4) continue;
        edge[cursor[bodyOf(meta[4 * t])]++] = t;
        edge[cursor[bodyOf(meta[4 * t + 1])]++] = t;
    }
    std::vector<uint8_t> bodyFlag(nb, 0), contactFlag(n, 0), jointFlag(nj, 0);
    std::vector<int32_t> stack, island;
    for (const b2Body* seed = m_bodyList; seed; seed = seed->m_next) {
        const int s0 = seed->m_index;
        if (bodyFlag[s0] || !(seed->m_flags & b2Body::e_awakeFlag) || !seed->IsActive() ||
            seed->m_type == b2BodyType::STATIC_BODY)
            continue;
        island.clear();
        stack.assign(1, s0);
        bodyFlag[s0] = 1;
        while (!stack.empty()) {
            const int b = stack.back();
            stack.pop_back();
            island.push_back(b);
            const b2Body* body = m_bodies[b];
            if (body->m_type == b2BodyType::STATIC_BODY) continue;
            for (int e = start[b]; e < start[b + 1]; ++e) {
                const int t = edge[e];
                if (contactFlag[t] || !on[t] || !(meta[4 * t + 2] & 1)) continue;
                const b2Fixture* fa = m_fixtures[meta[4 * t]];
                const b2Fixture* fb = m_fixtures[meta[4 * t + 1]];
                if (fa->m_isSensor || fb->m_isSensor) continue;
                order->push_back(t);
                contactFlag[t] = 1;
                const int other = fa->m_body->m_index == b ? fb->m_body->m_index : fa->m_body->m_index;
                if (bodyFlag[other]) continue;
                stack.push_back(other);
                bodyFlag[other] = 1;
            }
            for (const b2JointEdge* je = body->m_jointList; je; je = je->next) {
                const int j = je->joint->m_index;
                if (jointFlag[j] || !je->other->IsActive()) continue;
                jointFlag[j] = 1;
                const int other = je->other->m_index;
                if (bodyFlag[other]) continue;
                stack.push_back(other);
                bodyFlag[other] = 1;
            }
        }
        for (int b : island)
            if (m_bodies[b]->m_type == b2BodyType::STATIC_BODY) bodyFlag[b] = 0;
    }
}

void b2World::ReplayEvents(bool collideEventsDone) {
    b2Contact* list = GetContactList();
    // PostSolve goes to the contacts the step solved (list order is table
    // order reversed)
    const int nc = (int)m_contacts.size();
    std::vector<int32_t> solved(nc, 0);
    if (nc > 0) b2g_get_contact_solved(m_ctx, solved.data(), nc);
    std::vector<uint64_t> now;
    for (b2Contact* c = list; c; c = c->GetNext())
        if (c->IsTouching())
            now.push_back(pairKeyOf(c->m_fixtureA->m_index + c->m_indexA, c->m_fixtureB->m_index + c->m_indexB));
    std::sort(now.begin(), now.end());
    std::unordered_set<uint64_t> before(m_touchingKeys.begin(), m_touchingKeys.end());
    std::unordered_set<uint64_t> after(now.begin(), now.end());
    // PostSolve first, as b2Island::Report calls it at the end of each
    // island's solve (REF/Dynamics/b2Island.cpp:517-540), in the island
    // search's order when the split step computed it, else in list order;
    // then what the TOI phase changed, as SolveTOI runs after the islands.
    auto postSolve = [&](b2Contact* c) {
        b2ContactImpulse imp{};
        imp.count = c->m_manifold.pointCount;
        for (int j = 0; j < imp.count; ++j) {
            imp.normalImpulses[j] = c->m_manifold.points[j].normalImpulse;
            imp.tangentImpulses[j] = c->m_manifold.points[j].tangentImpulse;
        }
        m_contactListener->PostSolve(c, &imp);
    };
    if (collideEventsDone) {
        for (int32_t t : m_postSolveOrder)
            if (t >= 0 && t < nc && solved[t]) postSolve(&m_contacts[nc - 1 - t]);
    } else {
        for (b2Contact* c = list; c; c = c->GetNext())
            if (c->IsTouching() && solved[nc - 1 - (int)(c - &m_contacts[0])]) postSolve(c);
    }
    m_postSolveOrder.clear();
    // Without the split step every touching change of the step is reported
    // here; with it, only those of the TOI phase (the narrowphase's were
    // reported inside the step).
    for (uint64_t k : m_touchingKeys) {
        if (after.count(k)) continue;
        b2Contact ended;
        ended.m_fixtureA = m_fixtures[(int)(k >> 32)];
        ended.m_fixtureB = m_fixtures[(int)(k & 0xFFFFFFFFu)];
        ended.m_indexA = (int)(k >> 32) - ended.m_fixtureA->m_index;
        ended.m_indexB = (int)(k & 0xFFFFFFFFu) - ended.m_fixtureB->m_index;
        ended.m_touching = false;
        m_contactListener->EndContact(&ended);
    }
    for (b2Contact* c = list; c; c = c->GetNext()) {
        if (!c->IsTouching()) continue;
        const uint64_t k = pairKeyOf(c->m_fixtureA->m_index + c->m_indexA, c->m_fixtureB->m_index + c->m_indexB);
        if (!before.count(k)) m_contactListener->BeginContact(c);
    }
    m_touchingKeys.swap(now);
}

void b2World::GetPenetration(double* out) const {
    out[0] = out[1] = out[2] = 0.0;
    if (m_ctx) b2g_penetration(m_ctx, out);
}

}  // namespace box2d
