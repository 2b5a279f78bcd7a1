// nbx_disc.cpp -- initial conditions: the rotating disc (src/Utils.cpp:6-54)
// and the scenarios built from it (src/viewer/Scenarios.cpp:29-62).
//
// The generators are kept bit-compatible with the reference: same mt19937
// stream, same uniform_real_distribution<float>, draw order, float expression
// order, std::sort comparator and velocity pass.
#include "nbx_host.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace nbx_host {

namespace {
struct DiscBody {
  float x, y, vx, vy, m, r;
  uint32_t id;
};
} // namespace

void uniformDisc(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                 float *vy, float *m, float *r, uint32_t *id) {
  if (n == 0) {
    return;
  }
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> unit(0.0f, 1.0f);
  const float inner = 25.0f;
  const float outer = std::sqrt(static_cast<float>(n)) * 5.0f;
  const float t = inner / outer;

  std::vector<DiscBody> b;
  b.reserve(n);
  b.push_back(DiscBody{0.0f, 0.0f, 0.0f, 0.0f, 1e6f, inner, 0u});
  uint32_t next = 1;
  while (b.size() < n) {
    const float angle = unit(rng) * static_cast<float>(M_PI) * 2.0f;
    const float s = std::sin(angle);
    const float c = std::cos(angle);
    const float rad = unit(rng) * (1.0f - t * t) + t * t;
    const float root = std::sqrt(rad);
    b.push_back(DiscBody{c * outer * root, s * outer * root, s, -c, 1.0f,
                         std::cbrt(1.0f), next++});
  }
  std::sort(b.begin(), b.end(), [](const DiscBody &p, const DiscBody &q) {
    return p.x * p.x + p.y * p.y < q.x * q.x + q.y * q.y;
  });
  float enclosed = 0.0f;
  for (DiscBody &d : b) {
    enclosed += d.m;
    if (d.x == 0.0f && d.y == 0.0f) {
      continue;
    }
    const float v = std::sqrt(enclosed / std::sqrt(d.x * d.x + d.y * d.y));
    d.vx *= v;
    d.vy *= v;
  }
  for (uint64_t i = 0; i < n; ++i) {
    px[i] = b[i].x;
    py[i] = b[i].y;
    vx[i] = b[i].vx;
    vy[i] = b[i].vy;
    m[i] = b[i].m;
    r[i] = b[i].r;
    id[i] = b[i].id;
  }
}

namespace {
// One disc of `count` bodies written at offset `at`, moved by (x, y) and
// given the bulk velocity (bvx, bvy); `reverse` flips its rotation first
// (appendDisc, src/viewer/Scenarios.cpp:8-20).
void placeDisc(uint64_t count, uint32_t seed, uint64_t at, float x, float y,
               float bvx, float bvy, bool reverse, float *px, float *py,
               float *vx, float *vy, float *m, float *r) {
  std::vector<float> dx(count), dy(count), dvx(count), dvy(count);
  std::vector<uint32_t> ids(count);
  uniformDisc(count, seed, dx.data(), dy.data(), dvx.data(), dvy.data(),
              m + at, r + at, ids.data());
  for (uint64_t i = 0; i < count; ++i) {
    float u = dvx[i], w = dvy[i];
    if (reverse) {
      u = u * -1.0f;
      w = w * -1.0f;
    }
    px[at + i] = dx[i] + x;
    py[at + i] = dy[i] + y;
    vx[at + i] = u + bvx;
    vy[at + i] = w + bvy;
  }
}

void renumber(uint64_t n, uint32_t *id) {
  for (uint64_t i = 0; i < n; ++i) {
    id[i] = static_cast<uint32_t>(i);
  }
}
} // namespace

void twoGalaxies(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                 float *vy, float *m, float *r, uint32_t *id) {
  const uint64_t half = n / 2;
  const float radius = 5.0f * std::sqrt(static_cast<float>(half));
  const float sep = 2.5f * radius;
  const float totalMass = 2.0f * (1e6f + static_cast<float>(half));
  const float v = 0.3f * std::sqrt(totalMass / sep);
  placeDisc(half, seed, 0, -sep / 2, -radius / 3, v, 0.0f, false, px, py, vx,
            vy, m, r);
  placeDisc(n - half, seed + 1, half, sep / 2, radius / 3, -v, 0.0f, true, px,
            py, vx, vy, m, r);
  renumber(n, id);
}

void coldSquare(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                float *vy, float *m, float *r, uint32_t *id) {
  std::mt19937 rng(seed);
  const float half = 3.0f * std::sqrt(static_cast<float>(n));
  std::uniform_real_distribution<float> coord(-half, half);
  for (uint64_t i = 0; i < n; ++i) {
    const float y = coord(rng);
    const float x = coord(rng);
    px[i] = x;
    py[i] = y;
    vx[i] = 0.0f;
    vy[i] = 0.0f;
    m[i] = 1.0f;
    r[i] = 1.0f;
  }
  renumber(n, id);
}

void galaxyMerger(uint64_t n, uint32_t seed, float *px, float *py, float *vx,
                  float *vy, float *m, float *r, uint32_t *id) {
  const uint64_t half = n / 2;
  const float radius = 5.0f * std::sqrt(static_cast<float>(half));
  const float mass = 1e6f + static_cast<float>(half) - 1.0f;
  const float vc = std::sqrt(mass / radius);
  const float angle = 15.0f * static_cast<float>(M_PI) / 180.0f;
  const float ux = std::cos(angle), uy = std::sin(angle);
  const float s = 0.15f * vc;
  placeDisc(half, seed, 0, -2.0f * radius, 0.0f, s * ux, s * uy, false, px,
            py, vx, vy, m, r);
  placeDisc(n - half, seed + 1, half, 2.0f * radius, 0.0f, -s * ux, -s * uy,
            false, px, py, vx, vy, m, r);
  renumber(n, id);
}

} // namespace nbx_host
