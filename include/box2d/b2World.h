// box2d:: dynamics API of the host side (the fork's Box2D/Dynamics/b2World.h,
// b2Body.h, b2Fixture.h, b2WorldCallbacks.h, Contacts/b2Contact.h and
// b2TimeStep.h surface). b2World::Step runs on the GPU through the C ABI in
// b2g.h; the objects here are host mirrors that stay current after every
// Step (their getters return references into host members, as the
// reference's do).
//
// Differences from the reference, by design of the GPU path:
//  * continuous collision (SetContinuousPhysics, on by default as in the
//    reference) runs on the GPU after the discrete solve; the contact
//    updates of its TOI events call no listener during the step (their
//    touching changes are reported after it).
//  * with a contact listener, Step splits at the narrowphase: BeginContact,
//    EndContact and PreSolve are called inside the step in list order,
//    PostSolve after the solve for the island contacts, in the reference's
//    island order.
//  * DestroyBody and DestroyFixture keep device slots. The destroyed
//    contacts end at the start of the next Step (EndContact then); the
//    bodies they touched wake at once, as in the reference. Broad-phase
//    proxy ids follow the reference's tree: later fixtures take the freed
//    ids, so new contacts come in the reference's order.
#pragma once
#include <vector>
#include <string>
#include <memory>
#include "b2Math.h"
#include "b2Collision.h"
#include "../b2g.h"

struct b2g_context;

namespace box2d {

class b2World;
class b2Body;
class b2Fixture;
class b2Contact;
class b2Joint;
struct b2JointEdge;
struct b2ContactEdge;

// Box2D/Dynamics/b2Body.h:41-49
enum class b2BodyType { STATIC_BODY = 0, KINEMATIC_BODY, DYNAMIC_BODY };

// Box2D/Dynamics/b2Body.h:53-127
struct b2BodyDef {
    void* userData = nullptr;
    b2Vec2 position{0.0f, 0.0f};
    float angle = 0.0f;
    b2Vec2 linearVelocity{0.0f, 0.0f};
    float angularVelocity = 0.0f;
    float linearDamping = 0.0f;
    float angularDamping = 0.0f;
    bool allowSleep = true;
    bool awake = true;
    bool fixedRotation = false;
    bool bullet = false;
    b2BodyType type = b2BodyType::STATIC_BODY;
    bool active = true;
    float gravityScale = 1.0f;
};

// Box2D/Dynamics/b2Fixture.h:34-57
struct b2Filter {
    uint16_t categoryBits = 0x0001;
    uint16_t maskBits = 0xFFFF;
    int16_t groupIndex = 0;
};

// Box2D/Dynamics/b2Fixture.h:59-94
struct b2FixtureDef {
    const b2Shape* shape = nullptr;
    void* userData = nullptr;
    float friction = 0.2f;
    float restitution = 0.0f;
    float density = 0.0f;
    bool isSensor = false;
    b2Filter filter;
};

// Box2D/Dynamics/b2WorldCallbacks.h:74-79
struct b2ContactImpulse {
    float normalImpulses[MAX_MANIFOLD_POINTS];
    float tangentImpulses[MAX_MANIFOLD_POINTS];
    int32_t count;
};

// Box2D/Dynamics/b2WorldCallbacks.h:59-71
class b2ContactFilter {
public:
    virtual ~b2ContactFilter() {}
    virtual bool ShouldCollide(b2Fixture* fixtureA, b2Fixture* fixtureB);
};

// Box2D/Dynamics/b2WorldCallbacks.h:90-136
class b2ContactListener {
public:
    virtual ~b2ContactListener() {}
    virtual void BeginContact(b2Contact* contact) { B2_NOT_USED(contact); }
    virtual void EndContact(b2Contact* contact) { B2_NOT_USED(contact); }
    virtual void PreSolve(b2Contact* contact, const b2Manifold* oldManifold) {
        B2_NOT_USED(contact);
        B2_NOT_USED(oldManifold);
    }
    virtual void PostSolve(b2Contact* contact, const b2ContactImpulse* impulse) {
        B2_NOT_USED(contact);
        B2_NOT_USED(impulse);
    }
};

// Box2D/Dynamics/b2TimeStep.h:27-37 (milliseconds)
struct b2Profile {
    float step = 0.0f;
    float collide = 0.0f;
    float solve = 0.0f;
    float solveInit = 0.0f;
    float solveVelocity = 0.0f;
    float solvePosition = 0.0f;
    float broadphase = 0.0f;
    float solveTOI = 0.0f;
};

class b2Fixture {
public:
    b2Shape::Type GetType() const { return m_shape->GetType(); }
    b2Shape* GetShape() { return m_shape.get(); }
    const b2Shape* GetShape() const { return m_shape.get(); }
    bool IsSensor() const { return m_isSensor; }
    // b2Fixture::SetSensor (Box2D/Dynamics/b2Fixture.cpp:223-230): wakes the body
    // when the flag changes.
    void SetSensor(bool sensor);
    const b2Filter& GetFilterData() const { return m_filter; }
    void SetFilterData(const b2Filter& filter);
    b2Body* GetBody() { return m_body; }
    const b2Body* GetBody() const { return m_body; }
    b2Fixture* GetNext() { return m_next; }
    const b2Fixture* GetNext() const { return m_next; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }
    float GetDensity() const { return m_density; }
    float GetFriction() const { return m_friction; }
    void SetFriction(float f);
    float GetRestitution() const { return m_restitution; }
    void SetRestitution(float r);
    void GetMassData(b2MassData* md) const { m_shape->ComputeMass(md, m_density); }
    // Broad-phase proxy index of this fixture (its creation order in the
    // world); a chain's children follow it in child order.
    int32_t GetProxyIndex() const { return m_index; }

private:
    friend class b2Body;
    friend class b2World;
    void MarkUpdated();
    std::unique_ptr<b2Shape> m_shape;
    b2Body* m_body = nullptr;
    b2Fixture* m_next = nullptr;
    void* m_userData = nullptr;
    float m_density = 0.0f, m_friction = 0.2f, m_restitution = 0.0f;
    bool m_isSensor = false;
    b2Filter m_filter;
    int32_t m_index = -1;
};

class b2Body {
public:
    b2Fixture* CreateFixture(const b2FixtureDef* def);
    b2Fixture* CreateFixture(const b2Shape* shape, float density);
    // Box2D/Dynamics/b2Body.h:156: the fixture's contacts and broad-phase
    // proxies go (its contacts end at the start of the next Step, the bodies
    // they touched wake at once), the mass is recomputed. The fixture object
    // stays owned by the world until it is destroyed, but must not be used.
    void DestroyFixture(b2Fixture* fixture);

    void SetTransform(const b2Vec2& position, float angle);
    const b2Transform& GetTransform() const;
    const b2Vec2& GetPosition() const;
    float GetAngle() const;
    const b2Vec2& GetWorldCenter() const;
    const b2Vec2& GetLocalCenter() const;
    void SetLinearVelocity(const b2Vec2& v);
    const b2Vec2& GetLinearVelocity() const;
    void SetAngularVelocity(float omega);
    float GetAngularVelocity() const;
    void ApplyForce(const b2Vec2& force, const b2Vec2& point, bool wake);
    void ApplyForceToCenter(const b2Vec2& force, bool wake);
    void ApplyTorque(float torque, bool wake);
    void ApplyLinearImpulse(const b2Vec2& impulse, const b2Vec2& point, bool wake);
    void ApplyAngularImpulse(float impulse, bool wake);
    float GetMass() const { return m_mass; }
    float GetInertia() const;
    void GetMassData(b2MassData* data) const;
    void ResetMassData();
    b2Vec2 GetWorldPoint(const b2Vec2& localPoint) const;
    b2Vec2 GetWorldVector(const b2Vec2& localVector) const;
    b2Vec2 GetLocalPoint(const b2Vec2& worldPoint) const;
    b2Vec2 GetLocalVector(const b2Vec2& worldVector) const;
    b2Vec2 GetLinearVelocityFromWorldPoint(const b2Vec2& worldPoint) const;
    float GetLinearDamping() const { return m_linearDamping; }
    void SetLinearDamping(float d);
    float GetAngularDamping() const { return m_angularDamping; }
    void SetAngularDamping(float d);
    float GetGravityScale() const { return m_gravityScale; }
    void SetGravityScale(float s);
    b2BodyType GetType() const { return m_type; }
    bool IsBullet() const { return (m_flags & e_bulletFlag) != 0; }
    void SetSleepingAllowed(bool flag);
    bool IsSleepingAllowed() const { return (m_flags & e_autoSleepFlag) != 0; }
    void SetAwake(bool flag);
    bool IsAwake() const;
    bool IsActive() const { return (m_flags & e_activeFlag) != 0; }
    bool IsFixedRotation() const { return (m_flags & e_fixedRotationFlag) != 0; }
    b2Fixture* GetFixtureList() { return m_fixtureList; }
    const b2Fixture* GetFixtureList() const { return m_fixtureList; }
    b2JointEdge* GetJointList() { return m_jointList; }
    // Box2D/Dynamics/b2Body.h:301-305: the body's contacts, newest first (the
    // world's contact list order); valid until the next Step or edit.
    b2ContactEdge* GetContactList();
    // Box2D/Dynamics/b2Body.h:260-276, 583-586, 618-628, b2Body.cpp:354-394, 438-457
    b2Vec2 GetLinearVelocityFromLocalPoint(const b2Vec2& localPoint) const {
        return GetLinearVelocityFromWorldPoint(GetWorldPoint(localPoint));
    }
    void SetBullet(bool flag);
    // Box2D/Dynamics/b2Body.h:305: mass data reset, the body woken, its
    // contacts destroyed (ending at the start of the next Step) and its
    // proxies touched so the next Step's broad-phase recreates the pairs.
    void SetType(b2BodyType type);
    // Box2D/Dynamics/b2Body.h:353: proxies created (their pairs found at the
    // end of the next Step) or destroyed together with the contacts.
    void SetActive(bool flag);
    void SetFixedRotation(bool flag);
    void SetMassData(const b2MassData* data);
    const b2JointEdge* GetJointList() const { return m_jointList; }
    b2Body* GetNext() { return m_next; }
    const b2Body* GetNext() const { return m_next; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }
    b2World* GetWorld() { return m_world; }
    const b2World* GetWorld() const { return m_world; }
    // Device index of this body (its creation order in the world).
    int32_t GetIndex() const { return m_index; }
    // Sub-world this body was created in (b2World::BeginSubWorld).
    int32_t GetSubWorld() const { return m_subWorld; }

private:
    friend class b2World;
    friend class b2Fixture;
    enum {
        e_islandFlag = 0x0001, e_awakeFlag = 0x0002, e_autoSleepFlag = 0x0004, e_bulletFlag = 0x0008,
        e_fixedRotationFlag = 0x0010, e_activeFlag = 0x0020,
    };
    b2Body(const b2BodyDef* def, b2World* world, int32_t index);
    void Touch() const;   // reads the device state if the host copy is stale
    void MarkDirty();     // host edits to upload before the next Step
    b2BodyType m_type;
    int32_t m_flags = 0;
    int32_t m_index = -1;
    bool m_destroyed = false;  // by DestroyBody (the device slot stays)
    b2Transform m_xf;
    b2Sweep m_sweep;
    b2Vec2 m_linearVelocity;
    float m_angularVelocity = 0.0f;
    b2Vec2 m_force;
    float m_torque = 0.0f;
    b2World* m_world;
    b2Body* m_prev = nullptr;
    b2Body* m_next = nullptr;
    b2Fixture* m_fixtureList = nullptr;
    int32_t m_fixtureCount = 0;
    b2JointEdge* m_jointList = nullptr;
    b2ContactEdge* m_contactList = nullptr;  // built with the world's contact mirror
    float m_mass = 0.0f, m_invMass = 0.0f;
    float m_I = 0.0f, m_invI = 0.0f;
    float m_linearDamping = 0.0f, m_angularDamping = 0.0f, m_gravityScale = 1.0f;
    float m_sleepTime = 0.0f;
    void* m_userData = nullptr;
    bool m_dirty = false;
    int32_t m_subWorld = 0;
};

// Host view of one device contact, valid until the next Step.
// Box2D/Dynamics/Contacts/b2Contact.h:63-73: a contact's entry in a body's
// contact list.
struct b2ContactEdge {
    b2Body* other = nullptr;
    b2Contact* contact = nullptr;
    b2ContactEdge* prev = nullptr;
    b2ContactEdge* next = nullptr;
};

class b2Contact {
public:
    b2Manifold* GetManifold() { return &m_manifold; }
    const b2Manifold* GetManifold() const { return &m_manifold; }
    void GetWorldManifold(b2WorldManifold* worldManifold) const;
    bool IsTouching() const { return m_touching; }
    bool IsEnabled() const { return m_enabled; }
    b2Contact* GetNext() { return m_next; }
    const b2Contact* GetNext() const { return m_next; }
    b2Fixture* GetFixtureA() { return m_fixtureA; }
    const b2Fixture* GetFixtureA() const { return m_fixtureA; }
    int32_t GetChildIndexA() const { return m_indexA; }
    b2Fixture* GetFixtureB() { return m_fixtureB; }
    const b2Fixture* GetFixtureB() const { return m_fixtureB; }
    int32_t GetChildIndexB() const { return m_indexB; }
    float GetFriction() const { return m_friction; }
    float GetRestitution() const { return m_restitution; }
    float GetTangentSpeed() const { return m_tangentSpeed; }
    // Box2D/Dynamics/Contacts/b2Contact.h:98-142. Edits made inside
    // b2ContactListener::PreSolve reach the step: SetEnabled for this step,
    // friction, restitution and the tangent speed (a conveyor belt) for the
    // rest of the contact's life.
    void SetEnabled(bool flag) { m_enabled = flag; }
    void SetFriction(float friction) { m_friction = friction; }
    void ResetFriction();
    void SetRestitution(float restitution) { m_restitution = restitution; }
    void ResetRestitution();
    void SetTangentSpeed(float speed) { m_tangentSpeed = speed; }

private:
    friend class b2World;
    b2Fixture* m_fixtureA = nullptr;
    b2Fixture* m_fixtureB = nullptr;
    int32_t m_indexA = 0, m_indexB = 0;  // child of a chain fixture
    b2Manifold m_manifold;
    bool m_touching = false, m_enabled = true;
    float m_friction = 0.0f, m_restitution = 0.0f, m_tangentSpeed = 0.0f;
    b2Contact* m_next = nullptr;
    b2ContactEdge m_nodeA, m_nodeB;  // entries in body A's and body B's contact lists
};

// Box2D/Dynamics/Joints/b2Joint.h:31-53
enum class b2JointType {
    UNKNOWN_JOINT, REVOLUTE_JOINT, PRISMATIC_JOINT, DISTANCE_JOINT, PULLEY_JOINT, MOUSE_JOINT, GEAR_JOINT,
    WHEEL_JOINT, WELD_JOINT, FRICTION_JOINT, ROPE_JOINT, MOTOR_JOINT
};
enum class b2LimitState { INACTIVE_LIMIT, AT_LOWER_LIMIT, AT_UPPER_LIMIT, EQUAL_LIMITS };

// Box2D/Dynamics/Joints/b2Joint.h:62-68: a joint's entry in a body's joint list.
struct b2JointEdge {
    b2Body* other = nullptr;
    b2Joint* joint = nullptr;
    b2JointEdge* prev = nullptr;
    b2JointEdge* next = nullptr;
};

// Box2D/Dynamics/Joints/b2Joint.h:71-101
struct b2JointDef {
    b2JointType type = b2JointType::UNKNOWN_JOINT;
    void* userData = nullptr;
    b2Body* bodyA = nullptr;
    b2Body* bodyB = nullptr;
    bool collideConnected = false;
};

// Box2D/Dynamics/Joints/b2RevoluteJoint.h:37-90
struct b2RevoluteJointDef : public b2JointDef {
    b2RevoluteJointDef() { type = b2JointType::REVOLUTE_JOINT; }
    // Bodies, local anchors and reference angle from a world anchor point.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    float referenceAngle = 0.0f;
    bool enableLimit = false;
    float lowerAngle = 0.0f;
    float upperAngle = 0.0f;
    bool enableMotor = false;
    float motorSpeed = 0.0f;
    float maxMotorTorque = 0.0f;
};

// Box2D/Dynamics/Joints/b2Joint.h:105-190. Joints are created by
// b2World::CreateJoint and live on the device; the getters below read the
// device state of the last Step.
class b2Joint {
public:
    virtual ~b2Joint() = default;
    b2JointType GetType() const { return m_type; }
    b2Body* GetBodyA() { return m_bodyA; }
    b2Body* GetBodyB() { return m_bodyB; }
    virtual b2Vec2 GetAnchorA() const = 0;
    virtual b2Vec2 GetAnchorB() const = 0;
    virtual b2Vec2 GetReactionForce(float inv_dt) const = 0;
    virtual float GetReactionTorque(float inv_dt) const = 0;
    b2Joint* GetNext() { return m_next; }
    const b2Joint* GetNext() const { return m_next; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* data) { m_userData = data; }
    bool IsActive() const;
    bool GetCollideConnected() const { return m_collideConnected; }
    // Device index of this joint (its creation order in the world).
    int32_t GetIndex() const { return m_index; }

protected:
    friend class b2World;
    friend class b2Body;
    b2Joint(const b2JointDef* def, b2World* world, int32_t index);
    void MarkDirty();  // parameters to upload before the next Step
    // device state of the last Step: revolute impulse x, y, limit, motor;
    // distance impulse and direction u
    mutable float m_impulse[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool m_dirty = false;
    b2JointType m_type;
    b2Joint* m_prev = nullptr;
    b2Joint* m_next = nullptr;
    b2JointEdge m_edgeA, m_edgeB;
    b2Body* m_bodyA;
    b2Body* m_bodyB;
    b2World* m_world;
    int32_t m_index;
    bool m_collideConnected;
    void* m_userData;
};

// Box2D/Dynamics/Joints/b2RevoluteJoint.h:97-200
class b2RevoluteJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    float GetReferenceAngle() const { return m_referenceAngle; }
    float GetJointAngle() const;
    float GetJointSpeed() const;
    bool IsLimitEnabled() const { return m_enableLimit; }
    void EnableLimit(bool flag);
    float GetLowerLimit() const { return m_lowerAngle; }
    float GetUpperLimit() const { return m_upperAngle; }
    void SetLimits(float lower, float upper);
    bool IsMotorEnabled() const { return m_enableMotor; }
    void EnableMotor(bool flag);
    void SetMotorSpeed(float speed);
    float GetMotorSpeed() const { return m_motorSpeed; }
    void SetMaxMotorTorque(float torque);
    float GetMaxMotorTorque() const { return m_maxMotorTorque; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;
    float GetMotorTorque(float inv_dt) const;

private:
    friend class b2World;
    b2RevoluteJoint(const b2RevoluteJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB;
    float m_referenceAngle;
    bool m_enableLimit;
    float m_lowerAngle, m_upperAngle;
    bool m_enableMotor;
    float m_motorSpeed, m_maxMotorTorque;
};

// Box2D/Dynamics/Joints/b2DistanceJoint.h:32-63
struct b2DistanceJointDef : public b2JointDef {
    b2DistanceJointDef() { type = b2JointType::DISTANCE_JOINT; }
    // Bodies, local anchors and length from two world anchor points.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchorA, const b2Vec2& anchorB);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    float length = 1.0f;
    float frequencyHz = 0.0f;   // 0: a rigid rod
    float dampingRatio = 0.0f;
};

// Box2D/Dynamics/Joints/b2DistanceJoint.h:65-130: keeps two anchor points at
// a distance, rigidly or as a spring-damper.
class b2DistanceJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    void SetLength(float length);
    float GetLength() const { return m_length; }
    void SetFrequency(float hz);
    float GetFrequency() const { return m_frequencyHz; }
    void SetDampingRatio(float ratio);
    float GetDampingRatio() const { return m_dampingRatio; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2DistanceJoint(const b2DistanceJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB;
    float m_length, m_frequencyHz, m_dampingRatio;
};

// Box2D/Dynamics/Joints/b2RopeJoint.h:30-50
struct b2RopeJointDef : public b2JointDef {
    b2RopeJointDef() { type = b2JointType::ROPE_JOINT; }
    b2Vec2 localAnchorA{-1.0f, 0.0f};
    b2Vec2 localAnchorB{1.0f, 0.0f};
    float maxLength = 0.0f;     // must exceed the linear slop to act
};

// Box2D/Dynamics/Joints/b2RopeJoint.h:60-122: bounds the distance between
// two anchor points from above.
class b2RopeJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    void SetMaxLength(float length);
    float GetMaxLength() const { return m_maxLength; }
    b2LimitState GetLimitState() const;
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2RopeJoint(const b2RopeJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB;
    float m_maxLength;
};

// Box2D/Dynamics/Joints/b2WeldJoint.h:29-60
struct b2WeldJointDef : public b2JointDef {
    b2WeldJointDef() { type = b2JointType::WELD_JOINT; }
    // Bodies, local anchors and reference angle from a world anchor point.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    float referenceAngle = 0.0f;
    float frequencyHz = 0.0f;   // 0: rigid; else a soft angular spring
    float dampingRatio = 0.0f;
};

// Box2D/Dynamics/Joints/b2WeldJoint.h:64-146: glues two bodies together.
class b2WeldJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    float GetReferenceAngle() const { return m_referenceAngle; }
    void SetFrequency(float hz);
    float GetFrequency() const { return m_frequencyHz; }
    void SetDampingRatio(float ratio);
    float GetDampingRatio() const { return m_dampingRatio; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2WeldJoint(const b2WeldJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB;
    float m_referenceAngle, m_frequencyHz, m_dampingRatio;
};

// Box2D/Dynamics/Joints/b2FrictionJoint.h:27-53
struct b2FrictionJointDef : public b2JointDef {
    b2FrictionJointDef() { type = b2JointType::FRICTION_JOINT; }
    // Bodies and local anchors from a world anchor point.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    float maxForce = 0.0f;
    float maxTorque = 0.0f;
};

// Box2D/Dynamics/Joints/b2FrictionJoint.h:57-124: top-down friction.
class b2FrictionJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    void SetMaxForce(float force);
    float GetMaxForce() const { return m_maxForce; }
    void SetMaxTorque(float torque);
    float GetMaxTorque() const { return m_maxTorque; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2FrictionJoint(const b2FrictionJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB;
    float m_maxForce, m_maxTorque;
};

// Box2D/Dynamics/Joints/b2MotorJoint.h:27-56
struct b2MotorJointDef : public b2JointDef {
    b2MotorJointDef() { type = b2JointType::MOTOR_JOINT; }
    // Bodies and offsets from the bodies' current transforms.
    void Initialize(b2Body* bodyA, b2Body* bodyB);
    b2Vec2 linearOffset{0.0f, 0.0f};   // position of body B in body A's frame
    float angularOffset = 0.0f;        // angle of body B minus angle of body A
    float maxForce = 1.0f;
    float maxTorque = 1.0f;
    float correctionFactor = 0.3f;     // position correction in [0, 1]
};

// Box2D/Dynamics/Joints/b2MotorJoint.h:61-132: drives body B to an offset
// from body A.
class b2MotorJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    void SetLinearOffset(const b2Vec2& linearOffset);
    const b2Vec2& GetLinearOffset() const { return m_linearOffset; }
    void SetAngularOffset(float angularOffset);
    float GetAngularOffset() const { return m_angularOffset; }
    void SetMaxForce(float force);
    float GetMaxForce() const { return m_maxForce; }
    void SetMaxTorque(float torque);
    float GetMaxTorque() const { return m_maxTorque; }
    void SetCorrectionFactor(float factor);
    float GetCorrectionFactor() const { return m_correctionFactor; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2MotorJoint(const b2MotorJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_linearOffset;
    float m_angularOffset, m_maxForce, m_maxTorque, m_correctionFactor;
};

// Box2D/Dynamics/Joints/b2PulleyJoint.h:30-70
struct b2PulleyJointDef : public b2JointDef {
    b2PulleyJointDef() { type = b2JointType::PULLEY_JOINT; collideConnected = true; }
    // Bodies, anchors, lengths and ratio from ground and world anchor points.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& groundAnchorA, const b2Vec2& groundAnchorB,
                    const b2Vec2& anchorA, const b2Vec2& anchorB, float ratio);
    b2Vec2 groundAnchorA{-1.0f, 1.0f};
    b2Vec2 groundAnchorB{1.0f, 1.0f};
    b2Vec2 localAnchorA{-1.0f, 0.0f};
    b2Vec2 localAnchorB{1.0f, 0.0f};
    float lengthA = 0.0f;
    float lengthB = 0.0f;
    float ratio = 1.0f;
};

// Box2D/Dynamics/Joints/b2PulleyJoint.h:80-150: lengthA + ratio * lengthB
// stays constant.
class b2PulleyJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    b2Vec2 GetGroundAnchorA() const { return m_groundAnchorA; }
    b2Vec2 GetGroundAnchorB() const { return m_groundAnchorB; }
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    float GetLengthA() const { return m_lengthA; }
    float GetLengthB() const { return m_lengthB; }
    float GetRatio() const { return m_ratio; }
    float GetCurrentLengthA() const;
    float GetCurrentLengthB() const;
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2PulleyJoint(const b2PulleyJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_groundAnchorA, m_groundAnchorB, m_localAnchorA, m_localAnchorB;
    float m_lengthA, m_lengthB, m_ratio;
};

// Box2D/Dynamics/Joints/b2MouseJoint.h:28-53
struct b2MouseJointDef : public b2JointDef {
    b2MouseJointDef() { type = b2JointType::MOUSE_JOINT; }
    b2Vec2 target{0.0f, 0.0f};   // world point; also fixes the anchor on body B
    float maxForce = 0.0f;
    float frequencyHz = 5.0f;
    float dampingRatio = 0.7f;
};

// Box2D/Dynamics/Joints/b2MouseJoint.h:62-131: pulls a point of body B to a
// world target with a soft, force-limited spring.
class b2MouseJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override { return m_targetA; }
    b2Vec2 GetAnchorB() const override;
    void SetTarget(const b2Vec2& target);
    const b2Vec2& GetTarget() const { return m_targetA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    void SetMaxForce(float force);
    float GetMaxForce() const { return m_maxForce; }
    void SetFrequency(float hz);
    float GetFrequency() const { return m_frequencyHz; }
    void SetDampingRatio(float ratio);
    float GetDampingRatio() const { return m_dampingRatio; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    friend class b2Body;
    b2MouseJoint(const b2MouseJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_targetA, m_localAnchorB;
    float m_maxForce, m_frequencyHz, m_dampingRatio;
};

// Box2D/Dynamics/Joints/b2PrismaticJoint.h:30-92
struct b2PrismaticJointDef : public b2JointDef {
    b2PrismaticJointDef() { type = b2JointType::PRISMATIC_JOINT; }
    // Bodies, anchors, axis and reference angle from a world anchor and axis.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor, const b2Vec2& axis);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    b2Vec2 localAxisA{1.0f, 0.0f};
    float referenceAngle = 0.0f;
    bool enableLimit = false;
    float lowerTranslation = 0.0f;
    float upperTranslation = 0.0f;
    bool enableMotor = false;
    float maxMotorForce = 0.0f;
    float motorSpeed = 0.0f;
};

// Box2D/Dynamics/Joints/b2PrismaticJoint.h:95-200: body B slides along an
// axis fixed in body A.
class b2PrismaticJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    const b2Vec2& GetLocalAxisA() const { return m_localXAxisA; }
    float GetReferenceAngle() const { return m_referenceAngle; }
    float GetJointTranslation() const;
    bool IsLimitEnabled() const { return m_enableLimit; }
    void EnableLimit(bool flag);
    float GetLowerLimit() const { return m_lowerTranslation; }
    float GetUpperLimit() const { return m_upperTranslation; }
    void SetLimits(float lower, float upper);
    bool IsMotorEnabled() const { return m_enableMotor; }
    void EnableMotor(bool flag);
    void SetMotorSpeed(float speed);
    float GetMotorSpeed() const { return m_motorSpeed; }
    void SetMaxMotorForce(float force);
    float GetMaxMotorForce() const { return m_maxMotorForce; }
    float GetMotorForce(float inv_dt) const;
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2PrismaticJoint(const b2PrismaticJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB, m_localXAxisA;
    float m_referenceAngle;
    bool m_enableLimit;
    float m_lowerTranslation, m_upperTranslation;
    bool m_enableMotor;
    float m_maxMotorForce, m_motorSpeed;
};

// Box2D/Dynamics/Joints/b2WheelJoint.h:30-72
struct b2WheelJointDef : public b2JointDef {
    b2WheelJointDef() { type = b2JointType::WHEEL_JOINT; }
    // Bodies, anchors and axis from a world anchor and axis.
    void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor, const b2Vec2& axis);
    b2Vec2 localAnchorA{0.0f, 0.0f};
    b2Vec2 localAnchorB{0.0f, 0.0f};
    b2Vec2 localAxisA{1.0f, 0.0f};
    bool enableMotor = false;
    float maxMotorTorque = 0.0f;
    float motorSpeed = 0.0f;
    float frequencyHz = 2.0f;
    float dampingRatio = 0.7f;
};

// Box2D/Dynamics/Joints/b2WheelJoint.h:75-190: a wheel on a suspension
// spring along an axis of body A, with a motor.
class b2WheelJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
    const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
    const b2Vec2& GetLocalAxisA() const { return m_localXAxisA; }
    bool IsMotorEnabled() const { return m_enableMotor; }
    void EnableMotor(bool flag);
    void SetMotorSpeed(float speed);
    float GetMotorSpeed() const { return m_motorSpeed; }
    void SetMaxMotorTorque(float torque);
    float GetMaxMotorTorque() const { return m_maxMotorTorque; }
    float GetMotorTorque(float inv_dt) const;
    void SetSpringFrequencyHz(float hz);
    float GetSpringFrequencyHz() const { return m_frequencyHz; }
    void SetSpringDampingRatio(float ratio);
    float GetSpringDampingRatio() const { return m_dampingRatio; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2WheelJoint(const b2WheelJointDef* def, b2World* world, int32_t index);
    b2Vec2 m_localAnchorA, m_localAnchorB, m_localXAxisA;
    bool m_enableMotor;
    float m_maxMotorTorque, m_motorSpeed, m_frequencyHz, m_dampingRatio;
};

// Box2D/Dynamics/Joints/b2GearJoint.h:24-42
struct b2GearJointDef : public b2JointDef {
    b2GearJointDef() { type = b2JointType::GEAR_JOINT; }
    b2Joint* joint1 = nullptr;   // a revolute or prismatic joint
    b2Joint* joint2 = nullptr;   // a revolute or prismatic joint
    float ratio = 1.0f;
};

// Box2D/Dynamics/Joints/b2GearJoint.h:55-120: coordinate1 + ratio *
// coordinate2 = constant, over the four bodies of its two joints.
class b2GearJoint : public b2Joint {
public:
    b2Vec2 GetAnchorA() const override;
    b2Vec2 GetAnchorB() const override;
    b2Joint* GetJoint1() { return m_joint1; }
    b2Joint* GetJoint2() { return m_joint2; }
    void SetRatio(float ratio);
    float GetRatio() const { return m_ratio; }
    b2Vec2 GetReactionForce(float inv_dt) const override;
    float GetReactionTorque(float inv_dt) const override;

private:
    friend class b2World;
    b2GearJoint(const b2GearJointDef* def, b2World* world, int32_t index);
    b2Joint* m_joint1;
    b2Joint* m_joint2;
    float m_ratio;
    b2Vec2 m_localAnchorA, m_localAnchorB;
    b2g_gear_def m_def;   // resolved at creation, as b2GearJoint's constructor does
};

// Per-step solver options of the GPU path (no reference counterpart).
struct b2SolverOptions {
    int32_t mode = 0;                 // 0 coloured, 1 reference order
    bool learnedWarmStart = false;    // fresh contact points start from the warm-start network
    float residualThreshold = 0.0f;   // > 0: re-solve from the plain start when the learned
                                      // start leaves a larger RMS residual (m/s); 0: no check
};

// Box2D/Dynamics/b2WorldCallbacks.h:40-57: told of joints and fixtures that
// DestroyBody destroys implicitly.
class b2DestructionListener {
public:
    virtual ~b2DestructionListener() {}
    virtual void SayGoodbye(b2Joint* joint) = 0;
    virtual void SayGoodbye(b2Fixture* fixture) = 0;
};

// Box2D/Dynamics/b2WorldCallbacks.h:140-162
class b2QueryCallback {
public:
    virtual ~b2QueryCallback() {}
    virtual bool ReportFixture(b2Fixture* fixture) = 0;
};
class b2RayCastCallback {
public:
    virtual ~b2RayCastCallback() {}
    virtual float ReportFixture(b2Fixture* fixture, const b2Vec2& point, const b2Vec2& normal, float fraction) = 0;
};

class b2World {
public:
    explicit b2World(const b2Vec2& gravity);
    ~b2World();
    b2World(const b2World&) = delete;
    b2World& operator=(const b2World&) = delete;

    void SetContactFilter(b2ContactFilter* filter) { m_contactFilter = filter; }
    void SetDestructionListener(b2DestructionListener* listener) { m_destructionListener = listener; }
    void SetContactListener(b2ContactListener* listener) { m_contactListener = listener; }
    b2Body* CreateBody(const b2BodyDef* def);
    void DestroyBody(b2Body* body);
    // Box2D/Dynamics/b2World.h:85-92. Every joint type of the reference
    // (revolute, prismatic, distance, pulley, mouse, gear, wheel, weld,
    // friction, rope, motor).
    b2Joint* CreateJoint(const b2JointDef* def);
    void DestroyJoint(b2Joint* joint);
    b2Joint* GetJointList() { return m_jointList; }
    const b2Joint* GetJointList() const { return m_jointList; }
    void Step(float timeStep, int32_t velocityIterations, int32_t positionIterations);
    void ClearForces();
    b2Body* GetBodyList() { return m_bodyList; }
    const b2Body* GetBodyList() const { return m_bodyList; }
    b2Contact* GetContactList();
    void SetAllowSleeping(bool flag);
    bool GetAllowSleeping() const { return m_allowSleep; }
    void SetWarmStarting(bool flag) { m_warmStarting = flag; }
    bool GetWarmStarting() const { return m_warmStarting; }
    void SetContinuousPhysics(bool flag) { m_continuousPhysics = flag; m_worldDirty = true; }
    bool GetContinuousPhysics() const { return m_continuousPhysics; }
    void SetSubStepping(bool flag) { m_subStepping = flag; m_worldDirty = true; }
    bool GetSubStepping() const { return m_subStepping; }
    int32_t GetProxyCount() const { return (int32_t)m_fixtures.size(); }
    int32_t GetBodyCount() const { return m_bodyCount; }
    int32_t GetJointCount() const { return m_jointCount; }
    int32_t GetContactCount() const { return m_contactCount; }
    void SetGravity(const b2Vec2& gravity);
    b2Vec2 GetGravity() const { return m_gravity; }
    bool IsLocked() const { return m_locked; }
    void SetAutoClearForces(bool flag) { m_autoClearForces = flag; }
    bool GetAutoClearForces() const { return m_autoClearForces; }
    const b2Profile& GetProfile() const { return m_profile; }

    // GPU-path extensions.
    void SetSolverOptions(const b2SolverOptions& o) { m_options = o; }
    const b2SolverOptions& GetSolverOptions() const { return m_options; }
    // Status of the last Step (0 = ok, b2g.h error codes otherwise).
    int32_t GetLastStatus() const { return m_lastStatus; }
    const std::string& GetLastError() const { return m_lastError; }
    b2g_context* GetDeviceContext() { return m_ctx; }
    // Uploads pending bodies and fixtures without stepping.
    int32_t Flush();
    // Mean and max penetration over touching points and the point count,
    // from device world manifolds.
    void GetPenetration(double* meanMax3) const;
    // Box2D/Dynamics/b2World.h:115-123: every proxy whose fat box overlaps
    // the box, and every fixture child the segment hits (the callback's
    // return clips, ignores with -1, or stops with 0). The device's fat
    // boxes are tested by brute force on the host, so the reports are the
    // reference's but in proxy order rather than its tree's order.
    void QueryAABB(b2QueryCallback* callback, const b2AABB& aabb);
    void RayCast(b2RayCastCallback* callback, const b2Vec2& point1, const b2Vec2& point2);
    // Per fixture proxy (chain edges one each) in creation order: the id the
    // reference's broad-phase tree gives it, -1 once destroyed (or on an
    // inactive body).
    const std::vector<int32_t>& GetProxyIds() const { return m_proxyIds; }
    // Batched worlds: bodies created after BeginSubWorld() belong to a new
    // sub-world (ids 1, 2, ...; bodies created before are in sub-world 0).
    // Fixtures of different sub-worlds never collide, so one b2World holds
    // many independent worlds that Step advances together on one device
    // context. Sub-worlds are set before the bodies' first Step.
    int32_t BeginSubWorld() { return ++m_subWorldCount, m_currentSubWorld = m_subWorldCount; }
    int32_t GetSubWorldCount() const { return m_subWorldCount + 1; }
    // Autotuner (b2g_autotune): times the step's schedule choices on
    // 20 * rounds ordinary Steps of this world and keeps the fastest. No
    // choice changes results. Listener events of these steps are not
    // replayed. Returns a b2g.h status.
    int32_t Autotune(float timeStep, int32_t velocityIterations, int32_t positionIterations, int32_t rounds);

private:
    friend class b2Body;
    friend class b2Fixture;
    friend class b2Joint;
    friend class b2RevoluteJoint;
    friend class b2DistanceJoint;
    friend class b2RopeJoint;
    friend class b2WeldJoint;
    friend class b2FrictionJoint;
    friend class b2MotorJoint;
    friend class b2PulleyJoint;
    friend class b2MouseJoint;
    friend class b2PrismaticJoint;
    friend class b2WheelJoint;
    friend class b2GearJoint;
    int32_t Upload();
    // Uploads the setters' edits of joints already on the device (part of
    // Upload); DownloadJoints calls it first, so a reaction read after a
    // setter reflects the setter's side effects, as the reference's members
    // do at once.
    int32_t UploadJointEdits();
    void Download() const;
    void DownloadJoints() const;
    void ResolveGear(b2GearJoint* g);
    int32_t CollideWithListener(const void* params);
    // b2World::Solve's island search over the contact table after the
    // narrowphase: the island contacts in the order Report reaches them.
    void IslandContactOrder(const std::vector<int32_t>& meta, const std::vector<uint8_t>& on,
                            std::vector<int32_t>* order) const;
    // PostSolve after the Step; Begin/End too unless the split step
    // already delivered them
    void ReplayEvents(bool collideEventsDone);
    // Node ids of the reference's broad-phase tree pool for a proxy created
    // or destroyed now (b2DynamicTree::CreateProxy / DestroyProxy).
    int32_t CreateProxyId();
    void DestroyProxyId(int32_t id);

    b2g_context* m_ctx = nullptr;
    b2Vec2 m_gravity;
    b2Body* m_bodyList = nullptr;
    std::vector<b2Body*> m_bodies;        // creation order = device index
    std::vector<b2Fixture*> m_fixtures;   // creation order = proxy index
    int32_t m_bodyCount = 0;
    int32_t m_uploadedBodies = 0;
    int32_t m_uploadedFixtures = 0;
    int32_t m_contactCount = 0;
    bool m_allowSleep = true, m_warmStarting = true, m_continuousPhysics = true, m_subStepping = false;
    bool m_autoClearForces = true;
    bool m_locked = false;
    bool m_worldDirty = true;
    mutable bool m_hostStale = false;
    bool m_anyDirty = false;
    b2ContactFilter* m_contactFilter = nullptr;
    b2DestructionListener* m_destructionListener = nullptr;
    b2ContactListener* m_contactListener = nullptr;
    b2Profile m_profile;
    b2SolverOptions m_options;
    int32_t m_lastStatus = 0;
    std::string m_lastError;
    std::vector<b2Contact> m_contacts;
    bool m_contactsStale = true;
    std::vector<uint64_t> m_touchingKeys;  // touching fixture pairs after the last Step
    std::vector<int32_t> m_postSolveOrder; // PostSolve order of the open step (table indices after it)
    std::vector<int32_t> m_removedFixtures;  // proxies to drop at the next upload
    std::vector<b2Fixture*> m_destroyedFixtures;  // by DestroyFixture, deleted with the world
    std::vector<int32_t> m_typeChanged;    // bodies whose contacts SetType dropped, for the next upload
    std::vector<int32_t> m_restoredFixtures;  // proxies SetActive(true) gave back, for the next upload
    std::vector<int32_t> m_proxyIds;       // tree proxy id per fixture proxy, creation order
    std::vector<int32_t> m_treeFree;       // the tree's free node ids, head at the back (-1: a freed parent)
    int32_t m_treeCapacity = 0, m_treeLeaves = 0;
    bool m_proxyIdsReused = false;         // a proxy was destroyed: ids no longer follow creation order
    bool m_proxyIdsDirty = false;          // ids to upload with the next Upload
    std::vector<int32_t> m_updatedFixtures;  // material/filter changes to upload
    int32_t m_subWorldCount = 0;
    int32_t m_currentSubWorld = 0;
    b2Joint* m_jointList = nullptr;
    int32_t m_jointCount = 0;
    std::vector<b2Joint*> m_joints;        // creation order = device index
    int32_t m_uploadedJoints = 0;
    std::vector<int32_t> m_removedJoints;  // device joints to drop at the next upload
    std::vector<b2Joint*> m_dirtyJoints;
    mutable bool m_jointsStale = false;
};

}  // namespace box2d
