// ref_harness.cpp -- C entry points over the reference CPU sources
// (TEST INFRASTRUCTURE). oracle/Makefile compiles this file together with
// QuadTree.cpp, Collision.cpp, Simulation.cpp and Utils.cpp from the
// reference tree into oracle/_ref/libnbref.so, so the tests can run the real
// reference next to the oracle restatement. Nothing here is shipped.
//
// QuadTree keeps its node array private (QuadTree.h:87-107); the harness
// reads it by compiling the header with `private` mapped to `public`, which
// changes no layout.

#include <cstdint>
#include <cstring>
#include <vector>

#define private public
#include "QuadTree.h"
#undef private
#include "Collision.h"
#include "Simulation.h"
#include "Utils.h"
#include "viewer/Scenarios.h"

namespace {

std::vector<Body> gather(uint64_t n, const float *px, const float *py,
                         const float *vx, const float *vy, const float *ax,
                         const float *ay, const float *m, const float *r,
                         const uint32_t *id) {
  std::vector<Body> b;
  b.reserve(n);
  for (uint64_t i = 0; i < n; ++i) {
    Body body(Vec2(px[i], py[i]), Vec2(vx[i], vy[i]), m[i], r[i], id[i]);
    body.acc = Vec2(ax ? ax[i] : 0.0f, ay ? ay[i] : 0.0f);
    b.push_back(body);
  }
  return b;
}

void scatter(const std::vector<Body> &b, float *px, float *py, float *vx,
             float *vy, float *ax, float *ay, float *m, float *r,
             uint32_t *id) {
  for (size_t i = 0; i < b.size(); ++i) {
    px[i] = b[i].pos.x;
    py[i] = b[i].pos.y;
    vx[i] = b[i].vel.x;
    vy[i] = b[i].vel.y;
    if (ax) {
      ax[i] = b[i].acc.x;
    }
    if (ay) {
      ay[i] = b[i].acc.y;
    }
    m[i] = b[i].mass;
    r[i] = b[i].radius;
    id[i] = b[i].id;
  }
}

} // namespace

extern "C" {

// Utils.cpp uniformDisc(n, seed), generator order. Returns the body count
// (the reference always emits the central body, so n = 0 gives 1).
int64_t ref_uniform_disc(uint64_t n, uint32_t seed, float *px, float *py,
                         float *vx, float *vy, float *m, float *r,
                         uint32_t *id) {
  std::vector<Body> b = uniformDisc(n, seed);
  if (b.size() > n) {
    return static_cast<int64_t>(b.size());
  }
  scatter(b, px, py, vx, vy, nullptr, nullptr, m, r, id);
  return static_cast<int64_t>(b.size());
}

// makeScenario(kind, n, seed) (src/viewer/Scenarios.cpp:29-62). Returns the
// number of bodies it made (written only when it fits in n).
int64_t ref_scenario(int kind, uint64_t n, uint32_t seed, float *px, float *py,
                     float *vx, float *vy, float *m, float *r, uint32_t *id) {
  std::vector<Body> b = makeScenario(static_cast<Scenario>(kind), n, seed);
  if (b.size() > n) {
    return static_cast<int64_t>(b.size());
  }
  scatter(b, px, py, vx, vy, nullptr, nullptr, m, r, id);
  return static_cast<int64_t>(b.size());
}

// Body::update on every body.
void ref_iterate(uint64_t n, float dt, float *px, float *py, float *vx,
                 float *vy, const float *ax, const float *ay) {
  for (uint64_t i = 0; i < n; ++i) {
    Body b(Vec2(px[i], py[i]), Vec2(vx[i], vy[i]), 1.0f, 1.0f, 0);
    b.acc = Vec2(ax[i], ay[i]);
    b.update(dt);
    px[i] = b.pos.x;
    py[i] = b.pos.y;
    vx[i] = b.vel.x;
    vy[i] = b.vel.y;
  }
}

// QuadTree::build on the given bodies (which it reorders in place, written
// back to the arrays), then acceleration() for every body. Node arrays get
// min(count, capacity) entries; returns the node count.
int64_t ref_build(uint64_t n, float theta, float epsilon, float *px,
                  float *py, float *vx, float *vy, float *m, float *r,
                  uint32_t *id, int64_t capacity, float *comX, float *comY,
                  float *mass, float *centerX, float *centerY, float *size,
                  int32_t *children, int32_t *next, int32_t *start,
                  int32_t *end, float *ax, float *ay) {
  std::vector<Body> b =
      gather(n, px, py, vx, vy, nullptr, nullptr, m, r, id);
  QuadTree qt(theta, epsilon, 16);
  qt.build(b);
  for (Body &body : b) {
    body.acc = qt.acceleration(body.pos, b);
  }
  scatter(b, px, py, vx, vy, ax, ay, m, r, id);
  const int64_t count = static_cast<int64_t>(qt.nodes.size());
  for (int64_t k = 0; k < count && k < capacity; ++k) {
    const Node &nd = qt.nodes[k];
    comX[k] = nd.pos.x;
    comY[k] = nd.pos.y;
    mass[k] = nd.mass;
    centerX[k] = nd.quad.center.x;
    centerY[k] = nd.quad.center.y;
    size[k] = nd.quad.size;
    children[k] = nd.children;
    next[k] = nd.next;
    start[k] = static_cast<int32_t>(nd.bodyStart);
    end[k] = static_cast<int32_t>(nd.bodyEnd);
  }
  return count;
}

// Collision.cpp collide() in place (array order is unchanged by collide).
void ref_collide(uint64_t n, float *px, float *py, float *vx, float *vy,
                 float *m, float *r, uint32_t *id) {
  std::vector<Body> b =
      gather(n, px, py, vx, vy, nullptr, nullptr, m, r, id);
  collide(b);
  scatter(b, px, py, vx, vy, nullptr, nullptr, m, r, id);
}

// Simulation(n) (uniformDisc seed 0, dt 0.05, eps 1, leaf 16) with the tree's
// theta replaced, advanced `steps` steps (optionally followed by one extra
// iterate(), i.e. the state just before the next collide). State is returned
// in the simulation's array order; ids identify bodies.
void ref_simulation(uint64_t n, uint32_t steps, float theta, int then_iterate,
                    float *px, float *py, float *vx, float *vy, float *ax,
                    float *ay, float *m, float *r, uint32_t *id) {
  Simulation sim(n);
  sim.quadtree = QuadTree(theta, 1.0f, 16);
  for (uint32_t k = 0; k < steps; ++k) {
    sim.step();
  }
  if (then_iterate) {
    sim.iterate();
  }
  scatter(sim.bodies, px, py, vx, vy, ax, ay, m, r, id);
}

} // extern "C"
