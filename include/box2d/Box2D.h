// box2d:: public API of the MI355X Box2D step (the fork's Box2D/Box2D.h
// umbrella, reduced to the types the GPU step path serves).
#pragma once
#include "b2Math.h"
#include "b2Collision.h"
#include "b2World.h"
