// nbx_disc.cpp -- rotating-disc initial conditions (src/Utils.cpp:6-54).
//
// The generator is kept bit-compatible with the reference: same mt19937
// stream, same uniform_real_distribution<float>, draw order (angle, then
// radius), float expression order, std::sort comparator and velocity pass.
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

} // namespace nbx_host
