// nbx_oracle.cpp -- CPU oracle for the Barnes-Hut step (TEST INFRASTRUCTURE).
//
// A restatement, in this project's own code, of the reference CPU algorithm
// of tyreecepaul/n-body-simulation. Each function cites the reference lines
// it follows. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker: the HIP
// product path never calls it.
//
// One deliberate difference from the reference's QuadTree::build: bodies are
// ordered by a stable sort on the float-descent quadrant key instead of by
// in-place Hoare partitions (Partition.h:7-18). The partitions classify with
// the same float comparisons against the same float centres, so both orders
// put every body in the same leaf and give the same node topology; within a
// leaf the reference's order is unstable. partition() performs no swaps on an
// already-partitioned range, so the reference build run on key-sorted input
// keeps that order and then reproduces this oracle's node arrays bit for bit
// (tests/test_oracle_vs_reference.py checks exactly that).
//
// Floating point: compiled with -O2 -ffp-contract=off for baseline x86-64
// (SSE2, no FMA), so every float operation below rounds once, in source order.

#include "nbx_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <utility>
#include <vector>

namespace {

// QuadTree.cpp:11-26 fastInvSqrt: bit-level initial guess + one Newton step.
inline float fisr(float x) {
  const float xhalf = 0.5f * x;
  uint32_t i;
  std::memcpy(&i, &x, sizeof i);
  i = 0x5f3759dfu - (i >> 1);
  float y;
  std::memcpy(&y, &i, sizeof y);
  return y * (1.5f - xhalf * y * y);
}

// std::min(a, FLT_MAX) as written in QuadTree.cpp:290-292 (returns a for NaN).
inline float capMax(float a) {
  const float mx = std::numeric_limits<float>::max();
  return (mx < a) ? mx : a;
}

inline float lawInvSqrt(float dSq, int law) {
  return law == OR_LAW_ACCURATE ? 1.0f / std::sqrt(dSq) : fisr(dSq);
}

inline int digitAt(uint32_t hi, uint32_t lo, int depth) {
  return depth < 16 ? static_cast<int>((hi >> (30 - 2 * depth)) & 3u)
                    : static_cast<int>((lo >> (30 - 2 * (depth - 16))) & 3u);
}

constexpr int kLeafCapacity = 16; // Simulation.cpp:9
constexpr int kMaxDepth = 32;     // 2 key words x 16 levels

// ---- collision helpers (Collision.cpp) ----
struct Box {
  float minX, maxX, minY, maxY;
  int index;
};

// Collision.cpp:88-90
inline bool before(const Box &a, const Box &b) {
  return a.minX < b.minX || (a.minX == b.minX && a.index < b.index);
}

// Collision.cpp:95-98
inline bool overlaps(const Box &earlier, const Box &later) {
  return earlier.maxX >= later.minX && later.minY <= earlier.maxY &&
         earlier.minY <= later.maxY;
}

// Collision.cpp:100-104
inline int32_t cellCoord(float v, float invCell) {
  constexpr double limit = double(1 << 29);
  const double c = std::floor(static_cast<double>(v) * invCell);
  return static_cast<int32_t>(std::clamp(c, -limit, limit));
}

struct Pair {
  float laterMinX;
  int later;
  float earlierMinX;
  int earlier;
};

// Collision.cpp:333-340 pairLess.
inline bool pairLess(const Pair &p, const Pair &q) {
  if (p.later != q.later) {
    return p.laterMinX < q.laterMinX ||
           (p.laterMinX == q.laterMinX && p.later < q.later);
  }
  return p.earlierMinX < q.earlierMinX ||
         (p.earlierMinX == q.earlierMinX && p.earlier < q.earlier);
}

// Collision.cpp:16-84 resolve(bodies, i, j), on SoA arrays.
void resolve(float *px, float *py, float *vx, float *vy, const float *m,
             const float *r, int i, int j) {
  float p1x = px[i], p1y = py[i];
  float p2x = px[j], p2y = py[j];
  const float r1 = r[i], r2 = r[j];

  float dx = p2x - p1x, dy = p2y - p1y;
  const float rr = r1 + r2;
  if (dx * dx + dy * dy > rr * rr) {
    return;
  }

  const float v1x = vx[i], v1y = vy[i];
  const float v2x = vx[j], v2y = vy[j];
  const float wx = v2x - v1x, wy = v2y - v1y;
  float dDotV = dx * wx + dy * wy;

  const float m1 = m[i], m2 = m[j];
  const float weight1 = m2 / (m1 + m2);
  const float weight2 = m1 / (m1 + m2);

  const bool dNonZero = !(dx == 0.0f && dy == 0.0f);
  if (dDotV >= 0.0f && dNonZero) {
    const float mag = std::sqrt(dx * dx + dy * dy);
    const float s = rr / mag - 1.0f;
    const float tx = dx * s, ty = dy * s;
    px[i] -= weight1 * tx;
    py[i] -= weight1 * ty;
    px[j] += weight2 * tx;
    py[j] += weight2 * ty;
    return;
  }

  const float vSqEarly = wx * wx + wy * wy;
  if (vSqEarly < 1e-12f) {
    float dirx = 1.0f, diry = 0.0f;
    if (dNonZero) {
      const float inv = 1.0f / std::sqrt(dx * dx + dy * dy);
      dirx = dx * inv;
      diry = dy * inv;
    }
    const float tx = dirx * rr, ty = diry * rr;
    px[i] -= weight1 * tx;
    py[i] -= weight1 * ty;
    px[j] += weight2 * tx;
    py[j] += weight2 * ty;
    return;
  }

  const float vSq = wx * wx + wy * wy;
  float dSq = dx * dx + dy * dy;
  const float rSq = rr * rr;
  const float t =
      (dDotV + std::sqrt(std::max(dDotV * dDotV - vSq * (dSq - rSq), 0.0f))) /
      vSq;

  px[i] -= v1x * t;
  py[i] -= v1y * t;
  px[j] -= v2x * t;
  py[j] -= v2y * t;

  p1x = px[i];
  p1y = py[i];
  p2x = px[j];
  p2y = py[j];
  dx = p2x - p1x;
  dy = p2y - p1y;
  dDotV = dx * wx + dy * wy;
  dSq = dx * dx + dy * dy;
  const float s = 1.5f * dDotV / dSq;
  const float tx = dx * s, ty = dy * s;

  const float n1x = v1x + tx * weight1, n1y = v1y + ty * weight1;
  const float n2x = v2x - tx * weight2, n2y = v2y - ty * weight2;
  vx[i] = n1x;
  vy[i] = n1y;
  vx[j] = n2x;
  vy[j] = n2y;
  px[i] += n1x * t;
  py[i] += n1y * t;
  px[j] += n2x * t;
  py[j] += n2y * t;
}

template <typename T>
void applyPerm(uint64_t n, const std::vector<uint32_t> &perm, T *a) {
  std::vector<T> tmp(a, a + n);
  for (uint64_t i = 0; i < n; ++i) {
    a[i] = tmp[perm[i]];
  }
}

} // namespace

extern "C" {

void or_fast_inv_sqrt_bits(uint64_t count, const uint32_t *in, uint32_t *out) {
  for (uint64_t k = 0; k < count; ++k) {
    float x;
    std::memcpy(&x, &in[k], sizeof x);
    const float y = fisr(x);
    std::memcpy(&out[k], &y, sizeof y);
  }
}

// Utils.cpp:6-54.
void or_uniform_disc(uint64_t n, uint32_t seed, float *px, float *py,
                     float *vx, float *vy, float *m, float *r, uint32_t *id) {
  if (n == 0) {
    return;
  }
  struct B {
    float px, py, vx, vy, m, r;
    uint32_t id;
  };
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> unit(0.0f, 1.0f);
  const float innerRadius = 25.0f;
  const float outerRadius = std::sqrt(static_cast<float>(n)) * 5.0f;

  std::vector<B> bodies;
  bodies.reserve(n);
  bodies.push_back({0.0f, 0.0f, 0.0f, 0.0f, 1e6f, innerRadius, 0u});
  uint32_t nextId = 1;
  while (bodies.size() < n) {
    const float a = unit(rng) * static_cast<float>(M_PI) * 2.0f;
    const float sn = std::sin(a);
    const float cs = std::cos(a);
    const float t = innerRadius / outerRadius;
    const float rr = unit(rng) * (1.0f - t * t) + t * t;
    const float sq = std::sqrt(rr);
    const float x = cs * outerRadius * sq;
    const float y = sn * outerRadius * sq;
    const float mass = 1.0f;
    bodies.push_back({x, y, sn, -cs, mass, std::cbrt(mass), nextId++});
  }
  std::sort(bodies.begin(), bodies.end(), [](const B &a, const B &b) {
    return a.px * a.px + a.py * a.py < b.px * b.px + b.py * b.py;
  });
  float mass = 0.0f;
  for (uint64_t i = 0; i < n; ++i) {
    B &b = bodies[i];
    mass += b.m;
    if (b.px == 0.0f && b.py == 0.0f) {
      continue;
    }
    const float v = std::sqrt(mass / std::sqrt(b.px * b.px + b.py * b.py));
    b.vx *= v;
    b.vy *= v;
  }
  for (uint64_t i = 0; i < n; ++i) {
    px[i] = bodies[i].px;
    py[i] = bodies[i].py;
    vx[i] = bodies[i].vx;
    vy[i] = bodies[i].vy;
    m[i] = bodies[i].m;
    r[i] = bodies[i].r;
    id[i] = bodies[i].id;
  }
}

// appendDisc (Scenarios.cpp:8-20): a uniformDisc moved by (x, y) with bulk
// velocity (bvx, bvy), rotation flipped first when `reverse`.
static void or_append_disc(uint64_t count, uint32_t seed, uint64_t at, float x,
                           float y, float bvx, float bvy, int reverse,
                           float *px, float *py, float *vx, float *vy,
                           float *m, float *r) {
  std::vector<float> bx(count), by(count), bvxs(count), bvys(count);
  std::vector<uint32_t> ids(count);
  or_uniform_disc(count, seed, bx.data(), by.data(), bvxs.data(), bvys.data(),
                  m + at, r + at, ids.data());
  for (uint64_t i = 0; i < count; ++i) {
    float vxi = bvxs[i], vyi = bvys[i];
    if (reverse) {
      vxi = vxi * -1.0f;
      vyi = vyi * -1.0f;
    }
    px[at + i] = bx[i] + x;
    py[at + i] = by[i] + y;
    vx[at + i] = vxi + bvx;
    vy[at + i] = vyi + bvy;
  }
}

int or_scenario(int kind, uint64_t n, uint32_t seed, float *px, float *py,
                float *vx, float *vy, float *m, float *r, uint32_t *id) {
  if (kind < 0 || kind > 3 || ((kind == 1 || kind == 3) && n < 2)) {
    return -1;
  }
  if (kind == 0) {
    or_uniform_disc(n, seed, px, py, vx, vy, m, r, id);
    return 0;
  }
  if (kind == 1) { // Scenarios.cpp:36-49
    const uint64_t half = n / 2;
    const float radius = 5.0f * std::sqrt(static_cast<float>(half));
    const float sep = 2.5f * radius;
    const float totalMass = 2.0f * (1e6f + static_cast<float>(half));
    const float v = 0.3f * std::sqrt(totalMass / sep);
    or_append_disc(half, seed, 0, -sep / 2, -radius / 3, v, 0.0f, 0, px, py,
                   vx, vy, m, r);
    or_append_disc(n - half, seed + 1, half, sep / 2, radius / 3, -v, 0.0f, 1,
                   px, py, vx, vy, m, r);
  } else if (kind == 2) { // Scenarios.cpp:50-57; y is drawn first
    std::mt19937 rng(seed);
    const float half = 3.0f * std::sqrt(static_cast<float>(n));
    std::uniform_real_distribution<float> coord(-half, half);
    for (uint64_t i = 0; i < n; ++i) {
      py[i] = coord(rng);
      px[i] = coord(rng);
      vx[i] = 0.0f;
      vy[i] = 0.0f;
      m[i] = 1.0f;
      r[i] = 1.0f;
    }
  } else { // the precision-study merger
    const uint64_t half = n / 2;
    const float radius = 5.0f * std::sqrt(static_cast<float>(half));
    const float mass = 1e6f + static_cast<float>(half) - 1.0f;
    const float vc = std::sqrt(mass / radius);
    const float angle = 15.0f * static_cast<float>(M_PI) / 180.0f;
    const float ux = std::cos(angle), uy = std::sin(angle);
    const float sp = 0.15f * vc;
    or_append_disc(half, seed, 0, -2.0f * radius, 0.0f, sp * ux, sp * uy, 0,
                   px, py, vx, vy, m, r);
    or_append_disc(n - half, seed + 1, half, 2.0f * radius, 0.0f, -sp * ux,
                   -sp * uy, 0, px, py, vx, vy, m, r);
  }
  for (uint64_t i = 0; i < n; ++i) { // renumber (Scenarios.cpp:22-27)
    id[i] = static_cast<uint32_t>(i);
  }
  return 0;
}

// Body.h:20-23.
void or_iterate(uint64_t n, float dt, float *px, float *py, float *vx,
                float *vy, const float *ax, const float *ay) {
  for (uint64_t i = 0; i < n; ++i) {
    vx[i] += ax[i] * dt;
    vy[i] += ay[i] * dt;
    px[i] += vx[i] * dt;
    py[i] += vy[i] * dt;
  }
}

// QuadTree.cpp:30-51.
void or_root_quad(uint64_t n, const float *px, const float *py,
                  float quad[3]) {
  float minX = std::numeric_limits<float>::max();
  float minY = std::numeric_limits<float>::max();
  float maxX = std::numeric_limits<float>::lowest();
  float maxY = std::numeric_limits<float>::lowest();
  for (uint64_t i = 0; i < n; ++i) {
    minX = std::min(minX, px[i]);
    minY = std::min(minY, py[i]);
    maxX = std::max(maxX, px[i]);
    maxY = std::max(maxY, py[i]);
  }
  quad[0] = (minX + maxX) * 0.5f;
  quad[1] = (minY + maxY) * 0.5f;
  quad[2] = std::max(maxX - minX, maxY - minY);
}

// Descent through Quad::intoQuadrant (QuadTree.h:14-25) with the partition
// predicates of QuadTree::subdivide (QuadTree.cpp:60-78).
void or_keys(uint64_t n, const float *px, const float *py,
             const float quad[3], uint32_t *hi, uint32_t *lo) {
  for (uint64_t i = 0; i < n; ++i) {
    const float x = px[i], y = py[i];
    float cx = quad[0], cy = quad[1], s = quad[2];
    uint32_t h = 0, l = 0;
    for (int d = 0; d < kMaxDepth; ++d) {
      const int bx = (x < cx) ? 0 : 1;
      const int by = (y < cy) ? 0 : 1;
      const uint32_t q = static_cast<uint32_t>(2 * by + bx);
      if (d < 16) {
        h |= q << (30 - 2 * d);
      } else {
        l |= q << (30 - 2 * (d - 16));
      }
      s *= 0.5f;
      cx += (static_cast<float>(bx) - 0.5f) * s;
      cy += (static_cast<float>(by) - 0.5f) * s;
    }
    hi[i] = h;
    lo[i] = l;
  }
}

void or_sort_perm(uint64_t n, const uint32_t *hi, const uint32_t *lo,
                  uint32_t *perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) {
    return hi[a] != hi[b] ? hi[a] < hi[b] : lo[a] < lo[b];
  });
}

// QuadTree.cpp:154-262 (build, subdivide, accumulateLeaf, propagate).
int64_t or_build_tree(uint64_t n, const float *px, const float *py,
                      const float *m, const uint32_t *hi, const uint32_t *lo,
                      const float quad[3], int64_t capacity, float *comX,
                      float *comY, float *mass, float *centerX,
                      float *centerY, float *size, int32_t *children,
                      int32_t *next, int32_t *start, int32_t *end,
                      int32_t *levels_out) {
  if (capacity < 1) {
    return -1;
  }
  // Root (QuadTree.cpp:158-159): next = 0, whole body range.
  centerX[0] = quad[0];
  centerY[0] = quad[1];
  size[0] = quad[2];
  next[0] = 0;
  children[0] = 0;
  start[0] = 0;
  end[0] = static_cast<int32_t>(n);
  comX[0] = comY[0] = mass[0] = 0.0f;

  std::vector<int32_t> parents;
  int64_t levelBegin = 0, levelEnd = 1;
  int depth = 0;
  int levels = 0;
  while (levelBegin < levelEnd) {
    ++levels;
    // Level bookkeeping (QuadTree.cpp:188-204): children of this level's
    // splitting nodes go at levelEnd + 4 * rank.
    int64_t slot = levelEnd;
    std::vector<int64_t> childSlot(static_cast<size_t>(levelEnd - levelBegin));
    for (int64_t node = levelBegin; node < levelEnd; ++node) {
      if (end[node] - start[node] > kLeafCapacity && depth < kMaxDepth) {
        parents.push_back(static_cast<int32_t>(node));
        childSlot[node - levelBegin] = slot;
        slot += 4;
      } else {
        childSlot[node - levelBegin] = 0;
      }
    }
    if (slot > capacity) {
      return -1;
    }
    for (int64_t node = levelBegin; node < levelEnd; ++node) {
      const int64_t c = childSlot[node - levelBegin];
      const int32_t s = start[node], e = end[node];
      if (c != 0) {
        // subdivide (QuadTree.cpp:55-91): bodies are key-sorted, so the
        // quadrant boundaries are where digit `depth` reaches 1, 2, 3.
        int32_t bound[5];
        bound[0] = s;
        bound[4] = e;
        int32_t k = s;
        for (int q = 1; q < 4; ++q) {
          while (k < e && digitAt(hi[k], lo[k], depth) < q) {
            ++k;
          }
          bound[q] = k;
        }
        children[node] = static_cast<int32_t>(c);
        const float childSize = size[node] * 0.5f;
        for (int q = 0; q < 4; ++q) {
          const int64_t ch = c + q;
          // Quad::intoQuadrant (QuadTree.h:14-25).
          centerX[ch] = centerX[node] +
                        (static_cast<float>(q & 1) - 0.5f) * childSize;
          centerY[ch] = centerY[node] +
                        (static_cast<float>(q >> 1) - 0.5f) * childSize;
          size[ch] = childSize;
          next[ch] = q < 3 ? static_cast<int32_t>(ch + 1) : next[node];
          children[ch] = 0;
          start[ch] = bound[q];
          end[ch] = bound[q + 1];
          comX[ch] = comY[ch] = mass[ch] = 0.0f;
        }
      } else {
        // accumulateLeaf (QuadTree.cpp:93-142): four interleaved sums.
        float p0x = 0, p0y = 0, p1x = 0, p1y = 0, p2x = 0, p2y = 0, p3x = 0,
              p3y = 0;
        float m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        int32_t i = s;
        for (; i + 3 < e; i += 4) {
          p0x += px[i] * m[i];
          p0y += py[i] * m[i];
          m0 += m[i];
          p1x += px[i + 1] * m[i + 1];
          p1y += py[i + 1] * m[i + 1];
          m1 += m[i + 1];
          p2x += px[i + 2] * m[i + 2];
          p2y += py[i + 2] * m[i + 2];
          m2 += m[i + 2];
          p3x += px[i + 3] * m[i + 3];
          p3y += py[i + 3] * m[i + 3];
          m3 += m[i + 3];
        }
        for (; i < e; ++i) {
          p0x += px[i] * m[i];
          p0y += py[i] * m[i];
          m0 += m[i];
        }
        comX[node] = p0x + p1x + p2x + p3x;
        comY[node] = p0y + p1y + p2y + p3y;
        mass[node] = m0 + m1 + m2 + m3;
      }
    }
    levelBegin = levelEnd;
    levelEnd = slot;
    ++depth;
  }
  const int64_t count = levelEnd;

  // propagate (QuadTree.cpp:237-262): branch sums deepest-first, then every
  // weighted sum becomes a centre of mass.
  for (auto it = parents.rbegin(); it != parents.rend(); ++it) {
    const int32_t node = *it;
    const int32_t c = children[node];
    comX[node] = comX[c] + comX[c + 1] + comX[c + 2] + comX[c + 3];
    comY[node] = comY[c] + comY[c + 1] + comY[c + 2] + comY[c + 3];
    mass[node] = mass[c] + mass[c + 1] + mass[c + 2] + mass[c + 3];
  }
  for (int64_t node = 0; node < count; ++node) {
    const float d = std::max(mass[node], std::numeric_limits<float>::min());
    comX[node] /= d;
    comY[node] /= d;
  }
  if (levels_out) {
    *levels_out = levels;
  }
  return count;
}

// QuadTree.cpp:264-389 acceleration().
void or_accelerations(uint64_t n, const float *px, const float *py,
                      const float *m, int64_t nodes, const float *comX,
                      const float *comY, const float *mass,
                      const float *size, const int32_t *children,
                      const int32_t *next, const int32_t *start,
                      const int32_t *end, float theta, float epsilon,
                      int law, float *ax, float *ay, uint32_t *far_count,
                      uint32_t *pair_count) {
  const float tSq = theta * theta;
  const float eSq = epsilon * epsilon;
  for (uint64_t b = 0; b < n; ++b) {
    const float x = px[b], y = py[b];
    float accx = 0.0f, accy = 0.0f;
    uint32_t nFar = 0, nPair = 0;
    int32_t node = 0;
    if (nodes <= 0) {
      ax[b] = ay[b] = 0.0f;
      continue;
    }
    while (true) {
      const float dx = comX[node] - x;
      const float dy = comY[node] - y;
      const float dSq = dx * dx + dy * dy;
      if (size[node] * size[node] < dSq * tSq) {
        if (mass[node] > 0.0f) {
          const float inv = lawInvSqrt(dSq, law);
          const float soft = 1.0f / (dSq + eSq);
          const float f = capMax(mass[node] * inv * soft);
          accx += dx * f;
          accy += dy * f;
          ++nFar;
        }
        if (next[node] == 0) {
          break;
        }
        node = next[node];
      } else if (children[node] == 0) {
        float a0x = 0, a0y = 0, a1x = 0, a1y = 0, a2x = 0, a2y = 0, a3x = 0,
              a3y = 0;
        int32_t i = start[node];
        const int32_t e = end[node];
        auto term = [&](int32_t k, float &sx, float &sy) {
          const float bx = px[k] - x;
          const float by = py[k] - y;
          const float bSq = bx * bx + by * by;
          const float inv = lawInvSqrt(bSq, law);
          const float soft = 1.0f / (bSq + eSq);
          const float f = capMax(m[k] * inv * soft);
          sx += bx * f;
          sy += by * f;
        };
        for (; i + 3 < e; i += 4) {
          term(i, a0x, a0y);
          term(i + 1, a1x, a1y);
          term(i + 2, a2x, a2y);
          term(i + 3, a3x, a3y);
        }
        for (; i < e; ++i) {
          term(i, a0x, a0y);
        }
        nPair += static_cast<uint32_t>(end[node] - start[node]);
        accx += a0x + a1x + a2x + a3x;
        accy += a0y + a1y + a2y + a3y;
        if (next[node] == 0) {
          break;
        }
        node = next[node];
      } else {
        node = children[node];
      }
    }
    ax[b] = accx;
    ay[b] = accy;
    if (far_count) {
      far_count[b] = nFar;
    }
    if (pair_count) {
      pair_count[b] = nPair;
    }
  }
}

// Collision.cpp:135-349.
// ---- the reference GPU collide: Jacobi resolution ----
namespace {
struct JBody {
  float px, py, vx, vy, m, r;
};
struct JBox {
  float minX, maxX, minY, maxY;
  int index;
};
struct JAcc {
  float dpx = 0, dpy = 0, dvx = 0, dvy = 0, npx = 0, npy = 0, nvx = 0,
        nvy = 0;
  int contacts = 0, asLater = 0;
};
const uint64_t kJLargeKey = ~0ull;

JBox jbox(int i, const float *px, const float *py, const float *r) {
  return JBox{px[i] - r[i], px[i] + r[i], py[i] - r[i], py[i] + r[i], i};
}
bool jbefore(const JBox &a, const JBox &b) {
  return a.minX < b.minX || (a.minX == b.minX && a.index < b.index);
}
bool joverlaps(const JBox &e, const JBox &l) {
  return e.maxX >= l.minX && l.minY <= e.maxY && e.minY <= l.maxY;
}
// CollisionKernels.cuh:50-55: float floor, clamped to +-2^29 (NaN too).
int jcell(float v, float invCell) {
  return static_cast<int>(
      std::fmin(std::fmax(std::floor(v * invCell), -536870912.0f), 536870912.0f));
}
uint64_t jkey(int cx, int cy) {
  const int64_t bias = 1ll << 30;
  return (static_cast<uint64_t>(cy + bias) << 32) |
         static_cast<uint64_t>(cx + bias);
}

// resolvePair (CollisionKernels.cuh:104-187): body 1 = the later box.
bool jresolve(const JBody &a, const JBody &b, float &p1x, float &p1y,
              float &v1x, float &v1y, float &p2x, float &p2y, float &v2x,
              float &v2y) {
  float dx = b.px - a.px, dy = b.py - a.py;
  const float rr = a.r + b.r;
  if (dx * dx + dy * dy > rr * rr) {
    return false;
  }
  const float vx = b.vx - a.vx, vy = b.vy - a.vy;
  float dv = dx * vx + dy * vy;
  const float w1 = b.m / (a.m + b.m), w2 = a.m / (a.m + b.m);
  const bool nz = dx != 0.0f || dy != 0.0f;
  v1x = a.vx;
  v1y = a.vy;
  v2x = b.vx;
  v2y = b.vy;
  if (dv >= 0.0f && nz) {
    const float sc = rr / std::sqrt(dx * dx + dy * dy) - 1.0f;
    const float tx = dx * sc, ty = dy * sc;
    p1x = a.px - w1 * tx;
    p1y = a.py - w1 * ty;
    p2x = b.px + w2 * tx;
    p2y = b.py + w2 * ty;
    return true;
  }
  const float vSq = vx * vx + vy * vy;
  if (vSq < 1e-12f) {
    float ux = 1.0f, uy = 0.0f;
    if (nz) {
      const float inv = 1.0f / std::sqrt(dx * dx + dy * dy);
      ux = dx * inv;
      uy = dy * inv;
    }
    const float tx = ux * rr, ty = uy * rr;
    p1x = a.px - w1 * tx;
    p1y = a.py - w1 * ty;
    p2x = b.px + w2 * tx;
    p2y = b.py + w2 * ty;
    return true;
  }
  const float dSq = dx * dx + dy * dy, rSq = rr * rr;
  const float t =
      (dv + std::sqrt(std::fmax(dv * dv - vSq * (dSq - rSq), 0.0f))) / vSq;
  const float q1x = a.px - a.vx * t, q1y = a.py - a.vy * t;
  const float q2x = b.px - b.vx * t, q2y = b.py - b.vy * t;
  dx = q2x - q1x;
  dy = q2y - q1y;
  dv = dx * vx + dy * vy;
  const float sc = 1.5f * dv / (dx * dx + dy * dy);
  const float tx = dx * sc, ty = dy * sc;
  v1x = a.vx + tx * w1;
  v1y = a.vy + ty * w1;
  v2x = b.vx - tx * w2;
  v2y = b.vy - ty * w2;
  p1x = q1x + v1x * t;
  p1y = q1y + v1y * t;
  p2x = q2x + v2x * t;
  p2y = q2y + v2y * t;
  return true;
}

// accumulatePair (CollisionKernels.cuh:199-236): a's side of pair (a, b).
void jaccumulate(const JBox &boxA, const JBody &sa, int b, const float *px,
                 const float *py, const float *vx, const float *vy,
                 const float *m, const float *r, JAcc &acc) {
  const JBox boxB = jbox(b, px, py, r);
  const bool aLater = jbefore(boxB, boxA);
  if (!(aLater ? joverlaps(boxB, boxA) : joverlaps(boxA, boxB))) {
    return;
  }
  if (aLater) {
    ++acc.asLater;
  }
  const JBody sb{px[b], py[b], vx[b], vy[b], m[b], r[b]};
  float p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y;
  const bool hit = aLater
                       ? jresolve(sa, sb, p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y)
                       : jresolve(sb, sa, p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y);
  if (!hit) {
    return;
  }
  const float qx = aLater ? p1x : p2x, qy = aLater ? p1y : p2y;
  const float ux = aLater ? v1x : v2x, uy = aLater ? v1y : v2y;
  ++acc.contacts;
  acc.dpx += qx - sa.px;
  acc.dpy += qy - sa.py;
  acc.dvx += ux - sa.vx;
  acc.dvy += uy - sa.vy;
  acc.npx += qx;
  acc.npy += qy;
  acc.nvx += ux;
  acc.nvy += uy;
}

void jadd(JAcc &t, const JAcc &p) {
  t.dpx += p.dpx;
  t.dpy += p.dpy;
  t.dvx += p.dvx;
  t.dvy += p.dvy;
  t.npx += p.npx;
  t.npy += p.npy;
  t.nvx += p.nvx;
  t.nvy += p.nvy;
  t.contacts += p.contacts;
  t.asLater += p.asLater;
}
} // namespace

float or_jacobi_cell(uint64_t n, const float *r) {
  float cell = 1.0f;
  if (n > 0) {
    std::vector<float> radii(r, r + n);
    std::nth_element(radii.begin(), radii.begin() + n / 2, radii.end());
    cell = 2.0f * radii[n / 2] * 1.01f;
  }
  if (!(cell > 0.0f)) {
    cell = 1.0f;
  }
  return cell;
}

int64_t or_collide_jacobi(uint64_t n, float cell, float *px, float *py,
                          float *vx, float *vy, const float *m, const float *r,
                          int32_t *contacts, int32_t *aabb_as_later) {
  if (n < 2) {
    return 0;
  }
  const int N = static_cast<int>(n);
  const float invCell = 1.0f / cell;
  std::vector<uint64_t> key(n);
  int64_t large = 0;
  for (int i = 0; i < N; ++i) {
    const JBox b = jbox(i, px, py, r);
    if (b.maxX - b.minX > cell || b.maxY - b.minY > cell) {
      key[i] = kJLargeKey;
      ++large;
    } else {
      key[i] = jkey(jcell(b.minX, invCell), jcell(b.minY, invCell));
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < N; ++i) {
    idx[i] = i;
  }
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  std::vector<uint64_t> ks(n);
  for (int i = 0; i < N; ++i) {
    ks[i] = key[idx[i]];
  }
  const int nSmall = N - static_cast<int>(large);
  std::vector<float> opx(px, px + n), opy(py, py + n), ovx(vx, vx + n),
      ovy(vy, vy + n);
  std::vector<int32_t> cnt(n, 0), later(n, 0);
  auto write = [&](int a, const JAcc &acc) {
    float qx = px[a], qy = py[a], ux = vx[a], uy = vy[a];
    if (acc.contacts == 1) {
      qx = acc.npx;
      qy = acc.npy;
      ux = acc.nvx;
      uy = acc.nvy;
    } else if (acc.contacts > 1) {
      qx = px[a] + acc.dpx;
      qy = py[a] + acc.dpy;
      ux = vx[a] + acc.dvx;
      uy = vy[a] + acc.dvy;
    }
    opx[a] = qx;
    opy[a] = qy;
    ovx[a] = ux;
    ovy[a] = uy;
    cnt[a] = acc.contacts;
    later[a] = acc.asLater;
  };
  // Small boxes (smallBodyContacts, CollisionKernels.cuh:241-264).
  for (int s = 0; s < nSmall; ++s) {
    const int a = idx[s];
    const JBox boxA = jbox(a, px, py, r);
    const JBody sa{px[a], py[a], vx[a], vy[a], m[a], r[a]};
    const int cx = jcell(boxA.minX, invCell), cy = jcell(boxA.minY, invCell);
    JAcc acc;
    for (int dy = -1; dy <= 1; ++dy) {
      const uint64_t last = jkey(cx + 1, cy + dy);
      const uint64_t first = jkey(cx - 1, cy + dy);
      int e = static_cast<int>(
          std::lower_bound(ks.begin(), ks.begin() + nSmall, first) - ks.begin());
      for (; e < nSmall && ks[e] <= last; ++e) {
        if (idx[e] != a) {
          jaccumulate(boxA, sa, idx[e], px, py, vx, vy, m, r, acc);
        }
      }
    }
    for (int e = nSmall; e < N; ++e) {
      jaccumulate(boxA, sa, idx[e], px, py, vx, vy, m, r, acc);
    }
    write(a, acc);
  }
  // Large boxes: chunks of 4096 bodies, 128 strided lanes, tree-reduced
  // (largeBodyPartialKernel), chunks summed in order (largeBodyCombineKernel).
  const int lanes = 128, chunk = 4096;
  const int chunks = (N + chunk - 1) / chunk;
  for (int l = 0; l < static_cast<int>(large); ++l) {
    const int a = idx[nSmall + l];
    const JBox boxA = jbox(a, px, py, r);
    const JBody sa{px[a], py[a], vx[a], vy[a], m[a], r[a]};
    JAcc total;
    for (int c = 0; c < chunks; ++c) {
      const int begin = c * chunk, end = std::min(N, begin + chunk);
      std::vector<JAcc> part(lanes);
      for (int t = 0; t < lanes; ++t) {
        for (int b = begin + t; b < end; b += lanes) {
          if (b != a) {
            jaccumulate(boxA, sa, b, px, py, vx, vy, m, r, part[t]);
          }
        }
      }
      for (int stride = lanes / 2; stride > 0; stride >>= 1) {
        for (int t = 0; t < stride; ++t) {
          jadd(part[t], part[t + stride]);
        }
      }
      jadd(total, part[0]);
    }
    write(a, total);
  }
  std::copy(opx.begin(), opx.end(), px);
  std::copy(opy.begin(), opy.end(), py);
  std::copy(ovx.begin(), ovx.end(), vx);
  std::copy(ovy.begin(), ovy.end(), vy);
  if (contacts) {
    std::copy(cnt.begin(), cnt.end(), contacts);
  }
  if (aabb_as_later) {
    std::copy(later.begin(), later.end(), aabb_as_later);
  }
  return large;
}

int64_t or_collide(uint64_t n, float *px, float *py, float *vx, float *vy,
                   const float *m, const float *r, int32_t *later,
                   int32_t *earlier, int64_t pair_cap, float *cell_out) {
  if (n < 2) {
    if (cell_out) {
      *cell_out = 0.0f;
    }
    return 0;
  }
  std::vector<Box> boxes(n);
  std::vector<float> widths(n);
  for (uint64_t i = 0; i < n; ++i) {
    const float minx = px[i] - r[i], miny = py[i] - r[i];
    const float maxx = px[i] + r[i], maxy = py[i] + r[i];
    boxes[i] = {minx, maxx, miny, maxy, static_cast<int>(i)};
    widths[i] = std::max(maxx - minx, maxy - miny);
  }
  // Median box width x 1.01 (Collision.cpp:173-183).
  std::nth_element(widths.begin(), widths.begin() + n / 2, widths.end());
  float cell = widths[n / 2] * 1.01f;
  if (!(cell > 0.0f)) {
    cell = 1.0f;
  }
  if (cell_out) {
    *cell_out = cell;
  }
  const float invCell = 1.0f / cell;
  std::vector<char> large(n);
  std::vector<int32_t> cx(n), cy(n);
  std::map<std::pair<int32_t, int32_t>, std::vector<int>> grid;
  std::vector<int> largeList;
  for (uint64_t i = 0; i < n; ++i) {
    const Box &b = boxes[i];
    large[i] = (b.maxX - b.minX > cell || b.maxY - b.minY > cell) ? 1 : 0;
    if (large[i]) {
      largeList.push_back(static_cast<int>(i));
      continue;
    }
    cx[i] = cellCoord(b.minX, invCell);
    cy[i] = cellCoord(b.minY, invCell);
    grid[{cx[i], cy[i]}].push_back(static_cast<int>(i));
  }

  std::vector<Pair> pairs;
  auto test = [&](const Box &a, const Box &b) {
    const Box &e = before(a, b) ? a : b;
    const Box &l = before(a, b) ? b : a;
    if (overlaps(e, l)) {
      pairs.push_back({l.minX, l.index, e.minX, e.index});
    }
  };
  auto cellList = [&](int32_t x, int32_t y) -> const std::vector<int> * {
    auto it = grid.find({x, y});
    return it == grid.end() ? nullptr : &it->second;
  };
  // Half stencil over small boxes (Collision.cpp:264-317): same cell once,
  // then (x+1, y), then row y+1 at x-1..x+1.
  for (const auto &kv : grid) {
    const int32_t x = kv.first.first, y = kv.first.second;
    const std::vector<int> &own = kv.second;
    for (size_t a = 0; a < own.size(); ++a) {
      for (size_t b = a + 1; b < own.size(); ++b) {
        test(boxes[own[a]], boxes[own[b]]);
      }
      const int32_t nbr[4][2] = {{x + 1, y}, {x - 1, y + 1}, {x, y + 1},
                                 {x + 1, y + 1}};
      for (const auto &c : nbr) {
        if (const std::vector<int> *o = cellList(c[0], c[1])) {
          for (int b : *o) {
            test(boxes[own[a]], boxes[b]);
          }
        }
      }
    }
  }
  // Large boxes against every box (Collision.cpp:319-330).
  for (int l : largeList) {
    for (uint64_t x = 0; x < n; ++x) {
      const Box &a = boxes[l];
      const Box &b = boxes[x];
      if (static_cast<int>(x) == l || (large[x] && before(a, b))) {
        continue;
      }
      test(a, b);
    }
  }
  std::sort(pairs.begin(), pairs.end(), pairLess);
  const int64_t count = static_cast<int64_t>(pairs.size());
  for (int64_t k = 0; k < count && k < pair_cap; ++k) {
    later[k] = pairs[k].later;
    earlier[k] = pairs[k].earlier;
  }
  for (const Pair &p : pairs) {
    resolve(px, py, vx, vy, m, r, p.later, p.earlier);
  }
  return count;
}

// Simulation.cpp:11-16 and :18-50.
int or_step(uint64_t n, float dt, float theta, float epsilon, int collisions,
            int law, float *px, float *py, float *vx, float *vy, float *ax,
            float *ay, float *m, float *r, uint32_t *id) {
  or_iterate(n, dt, px, py, vx, vy, ax, ay);
  if (collisions) {
    or_collide(n, px, py, vx, vy, m, r, nullptr, nullptr, 0, nullptr);
  }
  if (n == 0) {
    return 0;
  }
  float quad[3];
  or_root_quad(n, px, py, quad);
  std::vector<uint32_t> hi(n), lo(n), perm(n);
  or_keys(n, px, py, quad, hi.data(), lo.data());
  or_sort_perm(n, hi.data(), lo.data(), perm.data());
  applyPerm(n, perm, px);
  applyPerm(n, perm, py);
  applyPerm(n, perm, vx);
  applyPerm(n, perm, vy);
  applyPerm(n, perm, m);
  applyPerm(n, perm, r);
  applyPerm(n, perm, id);
  applyPerm(n, perm, hi.data());
  applyPerm(n, perm, lo.data());
  const int64_t cap = 8 * static_cast<int64_t>(n) + 8;
  std::vector<float> comX(cap), comY(cap), mass(cap), cX(cap), cY(cap),
      size(cap);
  std::vector<int32_t> ch(cap), nx(cap), st(cap), en(cap);
  int32_t levels = 0;
  const int64_t nodes = or_build_tree(
      n, px, py, m, hi.data(), lo.data(), quad, cap, comX.data(), comY.data(),
      mass.data(), cX.data(), cY.data(), size.data(), ch.data(), nx.data(),
      st.data(), en.data(), &levels);
  if (nodes < 0) {
    return -1;
  }
  or_accelerations(n, px, py, m, nodes, comX.data(), comY.data(), mass.data(),
                   size.data(), ch.data(), nx.data(), st.data(), en.data(),
                   theta, epsilon, law, ax, ay, nullptr, nullptr);
  return 0;
}

// CheckCuda.cu:456-475 exactTreeLawForces, for a target subset. Coincident
// bodies (|d| = 0) contribute nothing, as in the tree law.
void or_exact_forces(uint64_t n, const float *px, const float *py,
                     const float *m, uint64_t n_targets,
                     const uint32_t *targets, double eps2, double *ax,
                     double *ay) {
  for (uint64_t t = 0; t < n_targets; ++t) {
    const uint64_t i = targets[t];
    double sx = 0.0, sy = 0.0;
    for (uint64_t j = 0; j < n; ++j) {
      if (j == i) {
        continue;
      }
      const double dx = double(px[j]) - px[i];
      const double dy = double(py[j]) - py[i];
      const double dSq = dx * dx + dy * dy;
      if (dSq == 0.0) {
        continue;
      }
      const double f = m[j] / (std::sqrt(dSq) * (dSq + eps2));
      sx += dx * f;
      sy += dy * f;
    }
    ax[t] = sx;
    ay[t] = sy;
  }
}

} // extern "C"
