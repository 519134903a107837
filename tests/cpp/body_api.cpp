// b2Body calls beyond HelloWorld's, through the box2d:: headers on the GPU
// path: per-body contact lists (Box2D/Dynamics/b2Body.h:301-305), SetMassData,
// SetFixedRotation, SetBullet and GetLinearVelocityFromLocalPoint
// (b2Body.h:260-276, 583-586, 618-628), and the destruction listener
// (b2WorldCallbacks.h:40-57). Prints "FAIL ..." and exits 1 on
// the first broken fact, "ok" otherwise.
#include <box2d/Box2D.h>

#include <cmath>
#include <cstdio>
#include <vector>

using namespace box2d;

static int fails = 0;

struct Goodbyes : b2DestructionListener {
    int joints = 0, fixtures = 0;
    void SayGoodbye(b2Joint*) override { ++joints; }
    void SayGoodbye(b2Fixture*) override { ++fixtures; }
};
static void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++fails;
    }
}

int main() {
    b2World world(b2Vec2(0.0f, -10.0f));
    b2BodyDef groundDef;
    b2Body* ground = world.CreateBody(&groundDef);
    b2EdgeShape edge;
    edge.Set(b2Vec2(-40.0f, 0.0f), b2Vec2(40.0f, 0.0f));
    ground->CreateFixture(&edge, 0.0f);
    std::vector<b2Body*> boxes;
    b2PolygonShape box;
    box.SetAsBox(0.5f, 0.5f);
    for (int row = 0; row < 6; ++row)
        for (int i = 0; i < 6 - row; ++i) {
            b2BodyDef bd;
            bd.type = b2BodyType::DYNAMIC_BODY;
            bd.position = b2Vec2(-3.0f + row * 0.5f + i * 1.05f, 0.5f + row * 1.0f);
            b2Body* b = world.CreateBody(&bd);
            b->CreateFixture(&box, 1.0f);
            boxes.push_back(b);
        }
    for (int s = 0; s < 60; ++s) world.Step(1.0f / 60.0f, 8, 3);

    // Every body's list is the world's list restricted to it, in order.
    int edges = 0;
    std::vector<b2Body*> all = boxes;
    all.push_back(ground);
    for (b2Body* b : all) {
        std::vector<b2Contact*> want;
        for (b2Contact* c = world.GetContactList(); c; c = c->GetNext())
            if (c->GetFixtureA()->GetBody() == b || c->GetFixtureB()->GetBody() == b) want.push_back(c);
        size_t k = 0;
        b2ContactEdge* prev = nullptr;
        for (b2ContactEdge* e = b->GetContactList(); e; e = e->next, ++k) {
            check(k < want.size() && e->contact == want[k], "contact list order");
            b2Body* other = e->contact->GetFixtureA()->GetBody() == b ? e->contact->GetFixtureB()->GetBody()
                                                                      : e->contact->GetFixtureA()->GetBody();
            check(e->other == other, "edge other body");
            check(e->prev == prev, "edge prev link");
            prev = e;
            ++edges;
        }
        check(k == want.size(), "contact list length");
    }
    check(edges == 2 * world.GetContactCount(), "two edges per contact");
    check(edges > 20, "the pyramid touches itself");

    // SetMassData: mass, centre and rotational inertia about the origin as given.
    b2Body* top = boxes.back();
    const b2Vec2 v0 = top->GetLinearVelocity();
    const float w0 = top->GetAngularVelocity();
    const b2Vec2 c0 = top->GetWorldCenter();
    b2MassData md;
    md.mass = 2.0f;
    md.center = b2Vec2(0.1f, 0.05f);
    md.I = 1.5f;
    top->SetMassData(&md);
    check(top->GetMass() == 2.0f, "SetMassData mass");
    check(std::fabs(top->GetInertia() - 1.5f) < 1e-6f, "SetMassData inertia");
    check(top->GetLocalCenter()[0] == 0.1f && top->GetLocalCenter()[1] == 0.05f, "SetMassData centre");
    const b2Vec2 c1 = top->GetWorldCenter(), p1 = top->GetWorldPoint(md.center);
    check(c1[0] == p1[0] && c1[1] == p1[1], "SetMassData world centre");
    const b2Vec2 dv = top->GetLinearVelocity() - v0;
    const b2Vec2 want = b2Cross(w0, c1 - c0);
    check(std::fabs(dv[0] - want[0]) < 1e-6f && std::fabs(dv[1] - want[1]) < 1e-6f, "SetMassData velocity shift");

    // GetLinearVelocityFromLocalPoint is the world-point velocity of the point.
    const b2Vec2 lp(0.3f, -0.2f);
    const b2Vec2 a = top->GetLinearVelocityFromLocalPoint(lp);
    const b2Vec2 b = top->GetLinearVelocityFromWorldPoint(top->GetWorldPoint(lp));
    check(a[0] == b[0] && a[1] == b[1], "velocity from a local point");

    // SetFixedRotation stops the spin and keeps the angle.
    b2Body* spinner = boxes[boxes.size() - 2];
    spinner->SetAngularVelocity(3.0f);
    spinner->SetFixedRotation(true);
    check(spinner->IsFixedRotation() && spinner->GetAngularVelocity() == 0.0f, "SetFixedRotation");
    const float angle = spinner->GetAngle();
    for (int s = 0; s < 30; ++s) world.Step(1.0f / 60.0f, 8, 3);
    check(spinner->GetAngle() == angle, "a fixed-rotation body keeps its angle");

    boxes[0]->SetBullet(true);
    check(boxes[0]->IsBullet(), "SetBullet on");
    boxes[0]->SetBullet(false);
    check(!boxes[0]->IsBullet(), "SetBullet off");
    for (int s = 0; s < 10; ++s) world.Step(1.0f / 60.0f, 8, 3);
    check(world.GetLastStatus() == 0, "steps after the edits");

    // DestroyBody tells the destruction listener of its joints and fixtures.
    Goodbyes bye;
    world.SetDestructionListener(&bye);
    b2Body* hinged = boxes[1];
    b2RevoluteJointDef jd;
    jd.Initialize(ground, hinged, hinged->GetPosition());
    world.CreateJoint(&jd);
    b2CircleShape extra;
    extra.m_radius = 0.2f;
    hinged->CreateFixture(&extra, 1.0f);
    world.DestroyBody(hinged);
    check(bye.joints == 1 && bye.fixtures == 2, "destruction listener goodbyes");
    for (int s = 0; s < 5; ++s) world.Step(1.0f / 60.0f, 8, 3);
    check(world.GetLastStatus() == 0, "steps after the destruction");

    if (fails) return 1;
    std::printf("ok %d edges\n", edges);
    return 0;
}
