// The reference's HelloWorld (HelloWorld/HelloWorld.cpp:30-101) switched to
// the MI355X step: only the include line differs, plus a status check after
// each Step. Built from the repository root with
// g++ -std=c++14 -Iinclude tests/cpp/hello_world.cpp -Lbox2dpp_amd/lib -lbox2d -lb2g -Wl,-rpath,box2dpp_amd/lib
#include <cstdio>
#include "box2d/Box2D.h"  // was: #include <Box2D/Box2D.h>
using namespace box2d;

#ifndef HELLO_FORMAT
#define HELLO_FORMAT "%4.2f %4.2f %4.2f\n"
#endif

int main() {
    b2World world(b2Vec2(0.0f, -10.0f));

    b2BodyDef groundBodyDef;
    groundBodyDef.position = b2Vec2(0.0f, -10.0f);
    b2Body* groundBody = world.CreateBody(&groundBodyDef);
    b2PolygonShape groundBox;
    groundBox.SetAsBox(50.0f, 10.0f);
    groundBody->CreateFixture(&groundBox, 0.0f);

    b2BodyDef bodyDef;
    bodyDef.type = b2BodyType::DYNAMIC_BODY;
    bodyDef.position = b2Vec2(0.0f, 4.0f);
    b2Body* body = world.CreateBody(&bodyDef);
    b2PolygonShape dynamicBox;
    dynamicBox.SetAsBox(1.0f, 1.0f);
    b2FixtureDef fixtureDef;
    fixtureDef.shape = &dynamicBox;
    fixtureDef.density = 1.0f;
    fixtureDef.friction = 0.3f;
    body->CreateFixture(&fixtureDef);

    float timeStep = 1.0f / 60.0f;
    int32_t velocityIterations = 6;
    int32_t positionIterations = 2;
    for (int32_t i = 0; i < 60; ++i) {
        world.Step(timeStep, velocityIterations, positionIterations);
        if (world.GetLastStatus() != 0) {  // what the GPU path refused
            fprintf(stderr, "%s\n", world.GetLastError().c_str());
            return 1;
        }
        b2Vec<float, 2> position = body->GetPosition();
        float angle = body->GetAngle();
        printf(HELLO_FORMAT, position[b2VecX], position[b2VecY], angle);
    }
    return 0;
}
