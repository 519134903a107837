// ORACLE TEST INFRASTRUCTURE - not part of the product.
//
// CPU restatement of upstream Box2D 2.3.2's per-step hot path (the arithmetic
// the fork keeps; REF = Regression/box2d-ref/Box2DRef in the reference tree).
// It is the checker the GPU step is compared with: FP32 throughout, the same
// operation order, FP contraction off, and the same list orderings (newest-
// first body, fixture, contact and contact-edge lists; pairs sorted by proxy
// id), so its results agree bit for bit with the reference build in
// oracle/_ref on the same inputs.
//
// Only tests/, the smoke check and the CPU-baseline leg of the benchmark may
// load this library (oracle/_build/liborc.so). The product never links it.
//
// Parity pins (tests/test_oracle_*.py):
//   narrowphase  - every b2Collide* against oracle/_ref on seeded corpora and
//                  the reference's known-answer cases; closed forms
//   shapes, mass - b2PolygonShape::Set/SetAsBox/ComputeMass against _ref;
//                  closed-form box mass and inertia
//   broad-phase  - pair sets against _ref's b2BroadPhase::UpdatePairs and an
//                  O(n^2) brute force; the fat-AABB rule against _ref
//   whole step   - HelloWorld, pyramid (sleep on/off) and small piles: every
//                  body pose, velocity and contact manifold bit-exact with
//                  _ref over hundreds of steps
// Not restated (the checker reports B2G_ERR_UNSUPPORTED): continuous
// collision (SolveTOI), sensors (b2TestOverlap/GJK), chain shapes, joints.
#pragma once
#include <stdint.h>
#include "b2g.h"
