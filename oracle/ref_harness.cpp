// ref_harness.cpp -- C entry points over the reference CPU sources
// (TEST INFRASTRUCTURE). oracle/Makefile compiles this file together with
// QuadTree.cpp, Collision.cpp, Simulation.cpp and Utils.cpp from the
// reference tree into oracle/_ref/libnbref.so, so the tests can run the real
// reference next to the oracle restatement. Nothing here is shipped.
//
// QuadTree keeps its node array private (QuadTree.h:87-107); the harness
// reads it by compiling the header with `private` mapped to `public`, which
// changes no layout.

#include <algorithm>
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
#include "cuda/CollisionKernels.cuh"

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

// The reference GPU collide run on the host: the per-body functions of
// src/cuda/CollisionKernels.cuh (compiled here as plain C++, as that header
// intends), driven the way CudaCollision.cu's kernels drive them: keyKernel,
// a stable sort of (key, index), smallBodyKernel for every small box, and the
// default split large-box path (largeBodyPartialKernel with its 128-lane
// fixed tree reduction per 4096-body chunk, then largeBodyCombineKernel).
// The cell is CudaCollider's: 2 x median radius x 1.01.
int64_t ref_collide_jacobi(uint64_t n, float *px, float *py, float *vx,
                           float *vy, const float *m, const float *r,
                           int32_t *contacts, int32_t *aabb, float *cell_out) {
  using namespace collision_detail;
  std::vector<float> radii(r, r + n);
  float cell = 1.0f;
  if (!radii.empty()) {
    std::nth_element(radii.begin(), radii.begin() + radii.size() / 2,
                     radii.end());
    cell = 2.0f * radii[radii.size() / 2] * 1.01f;
  }
  if (!(cell > 0.0f)) {
    cell = 1.0f;
  }
  *cell_out = cell;
  const int N = static_cast<int>(n);
  if (N < 2) {
    return 0;
  }
  const float invCell = 1.0f / cell;
  std::vector<unsigned long long> keys(n);
  int largeCount = 0;
  for (int i = 0; i < N; ++i) {
    keys[i] = boxKey(makeBox(i, px, py, r), cell, invCell);
    largeCount += keys[i] == kLargeKey;
  }
  std::vector<int> order(n);
  for (int i = 0; i < N; ++i) {
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return keys[a] < keys[b]; });
  std::vector<unsigned long long> keysSorted(n);
  for (int i = 0; i < N; ++i) {
    keysSorted[i] = keys[order[i]];
  }
  std::vector<float> opx(n), opy(n), ovx(n), ovy(n);
  std::vector<int> cnt(n), later(n);
  const BodyArrays bodies{px, py, vx, vy, m, r};
  const int nSmall = N - largeCount;
  for (int s = 0; s < nSmall; ++s) {
    const int a = order[s];
    const Accum acc = smallBodyContacts(s, N, nSmall, invCell,
                                        keysSorted.data(), order.data(), bodies);
    writeResult(a, loadState(a, bodies), acc, opx.data(), opy.data(),
                ovx.data(), ovy.data(), cnt.data(), later.data());
  }
  const int lanes = 128, chunk = 4096;
  const int chunks = (N + chunk - 1) / chunk;
  for (int l = 0; l < largeCount; ++l) {
    const int a = order[nSmall + l];
    const Box boxA = makeBox(a, px, py, r);
    const BodyState sa = loadState(a, bodies);
    Accum total;
    for (int c = 0; c < chunks; ++c) {
      const int begin = c * chunk, end = std::min(N, begin + chunk);
      std::vector<Accum> mine(lanes);
      for (int t = 0; t < lanes; ++t) {
        for (int b = begin + t; b < end; b += lanes) {
          if (b != a) {
            accumulatePair(boxA, sa, b, bodies, mine[t]);
          }
        }
      }
      for (int stride = lanes / 2; stride > 0; stride >>= 1) {
        for (int t = 0; t < stride; ++t) {
          Accum &x = mine[t];
          const Accum &y = mine[t + stride];
          x.dpx += y.dpx;
          x.dpy += y.dpy;
          x.dvx += y.dvx;
          x.dvy += y.dvy;
          x.npx += y.npx;
          x.npy += y.npy;
          x.nvx += y.nvx;
          x.nvy += y.nvy;
          x.contacts += y.contacts;
          x.aabbAsLater += y.aabbAsLater;
        }
      }
      const Accum &p = mine[0];
      total.dpx += p.dpx;
      total.dpy += p.dpy;
      total.dvx += p.dvx;
      total.dvy += p.dvy;
      total.npx += p.npx;
      total.npy += p.npy;
      total.nvx += p.nvx;
      total.nvy += p.nvy;
      total.contacts += p.contacts;
      total.aabbAsLater += p.aabbAsLater;
    }
    writeResult(a, sa, total, opx.data(), opy.data(), ovx.data(), ovy.data(),
                cnt.data(), later.data());
  }
  for (int i = 0; i < N; ++i) {
    px[i] = opx[i];
    py[i] = opy[i];
    vx[i] = ovx[i];
    vy[i] = ovy[i];
    if (contacts) {
      contacts[i] = cnt[i];
    }
    if (aabb) {
      aabb[i] = later[i];
    }
  }
  return largeCount;
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
