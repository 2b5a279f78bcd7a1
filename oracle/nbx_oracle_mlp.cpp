// nbx_oracle_mlp.cpp -- CPU oracle of the learned far-field correction
// (TEST INFRASTRUCTURE; part of libnbx_oracle, never linked by the product).
//
// The reference has no learned component. What it does have, and what this
// file restates, is the far-field term the network corrects -- the monopole
// of an accepted node, QuadTree.cpp:276-296 (accept when size^2 < d^2
// theta^2, add m d / (|d| (|d|^2 + eps^2))) -- and the exact sum that judges
// it, exactTreeLawForces (src/cuda/CheckCuda.cu:456-475). On top of those:
//   * or_mlp_pairs       the (target, node) pairs a coarse-theta walk accepts
//                        inside the trained range (the training set's pairs);
//   * or_mlp_features    the network's 10 inputs, its regression target (the
//                        residual exact - monopole of that node's bodies) and
//                        the pair's loss weight, the squared share of its
//                        monopole in the target's monopole sum;
//   * or_mlp_forward /   the 10 -> 64 -> 64 -> 64 -> 2 GELU network and the
//     or_mlp_loss_grad   gradient of its mean squared error, either exactly
//                        (double) or with the BF16 rounding points of the
//                        GPU kernels emulated (operands rounded to BF16,
//                        FP32 accumulation);
//   * or_mlp_accelerations  the learned-law walk: accurate law, in-range far
//                        nodes monopole + network, out-of-range far nodes
//                        opened.
// Pins (tests/test_oracle_mlp.py): finite differences of the exact loss for
// every parameter block; GELU against its closed form; features of a node
// of known moments against a hand computation; the residual of a single
// body (exactly zero) and of two equal bodies against the closed-form
// quadrupole limit; the pair walk reduces to the plain walk's far set when
// the range test is void; the learned walk with a zero network equals a
// plain FP64 walk with the same opening rule.
#include "nbx_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

constexpr int kIn = OR_MLP_IN, kH = OR_MLP_HIDDEN, kOut = OR_MLP_OUT;
// Offsets of the parameter blocks in the flat array (row-major [out][in]).
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH,
              oB4 = oW4 + kOut * kH;
static_assert(oB4 + kOut == OR_MLP_PARAMS, "parameter layout");

float bf16Round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    return x; // NaN stays NaN
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float y;
  std::memcpy(&y, &u, 4);
  return y;
}

// GELU in its tanh form, x * sigmoid(A (x + B x^3)) = 0.5 x (1 + tanh(
// sqrt(2/pi) (x + 0.044715 x^3))), and its derivative.
constexpr double kGeluA = 1.5957691216057308; // 2 sqrt(2 / pi)
constexpr double kGeluB = 0.044715;

template <class T> T gelu(T x) {
  const T z = T(kGeluA) * (x + T(kGeluB) * x * x * x);
  return x / (T(1) + std::exp(-z));
}
template <class T> T geluGrad(T x) {
  const T z = T(kGeluA) * (x + T(kGeluB) * x * x * x);
  const T s = T(1) / (T(1) + std::exp(-z));
  return s + x * s * (T(1) - s) * T(kGeluA) * (T(1) + T(3 * kGeluB) * x * x);
}

// One pair through the network. EMU: BF16 operands, FP32 arithmetic (the
// GPU's rounding points); otherwise exact double.
template <bool EMU> struct Net {
  using T = typename std::conditional<EMU, float, double>::type;
  static T rnd(T v) { return EMU ? T(bf16Round(float(v))) : v; }

  const float *p;
  T w1[kH * kIn], w2[kH * kH], w3[kH * kH], w4[kOut * kH];
  explicit Net(const float *params) : p(params) {
    for (int i = 0; i < kH * kIn; ++i) w1[i] = rnd(T(p[oW1 + i]));
    for (int i = 0; i < kH * kH; ++i) w2[i] = rnd(T(p[oW2 + i]));
    for (int i = 0; i < kH * kH; ++i) w3[i] = rnd(T(p[oW3 + i]));
    for (int i = 0; i < kOut * kH; ++i) w4[i] = rnd(T(p[oW4 + i]));
  }
  // Forward, keeping what the backward pass needs.
  struct Acts {
    T x[kIn], z1[kH], h1[kH], z2[kH], h2[kH], z3[kH], h3[kH], y[kOut];
  };
  void forward(const float *f, Acts &a) const {
    for (int k = 0; k < kIn; ++k) a.x[k] = rnd(T(f[k]));
    layer(w1, p + oB1, a.x, kIn, a.z1, kH);
    for (int i = 0; i < kH; ++i) a.h1[i] = rnd(gelu(a.z1[i]));
    layer(w2, p + oB2, a.h1, kH, a.z2, kH);
    for (int i = 0; i < kH; ++i) a.h2[i] = rnd(gelu(a.z2[i]));
    layer(w3, p + oB3, a.h2, kH, a.z3, kH);
    for (int i = 0; i < kH; ++i) a.h3[i] = rnd(gelu(a.z3[i]));
    layer(w4, p + oB4, a.h3, kH, a.y, kOut);
  }
  static void layer(const T *w, const float *b, const T *in, int nin, T *out,
                    int nout) {
    for (int o = 0; o < nout; ++o) {
      T s = T(b[o]);
      for (int k = 0; k < nin; ++k) s += w[o * nin + k] * in[k];
      out[o] = s;
    }
  }
  // Backward of one pair's (y - t)^2 summed over outputs, times `scale`,
  // accumulated into grad (double).
  void backward(const Acts &a, const float *t, T scale, double *g) const {
    T dz4[kOut], dz3[kH], dz2[kH], dz1[kH];
    for (int o = 0; o < kOut; ++o) dz4[o] = rnd(T(2) * (a.y[o] - T(t[o])) * scale);
    weightGrad(dz4, kOut, a.h3, kH, g + oW4, g + oB4);
    for (int i = 0; i < kH; ++i) {
      T s = 0;
      for (int o = 0; o < kOut; ++o) s += w4[o * kH + i] * dz4[o];
      dz3[i] = rnd(s * geluGrad(a.z3[i]));
    }
    weightGrad(dz3, kH, a.h2, kH, g + oW3, g + oB3);
    for (int i = 0; i < kH; ++i) {
      T s = 0;
      for (int o = 0; o < kH; ++o) s += w3[o * kH + i] * dz3[o];
      dz2[i] = rnd(s * geluGrad(a.z2[i]));
    }
    weightGrad(dz2, kH, a.h1, kH, g + oW2, g + oB2);
    for (int i = 0; i < kH; ++i) {
      T s = 0;
      for (int o = 0; o < kH; ++o) s += w2[o * kH + i] * dz2[o];
      dz1[i] = rnd(s * geluGrad(a.z1[i]));
    }
    weightGrad(dz1, kH, a.x, kIn, g + oW1, g + oB1);
  }
  static void weightGrad(const T *dz, int nout, const T *in, int nin, double *gw,
                         double *gb) {
    for (int o = 0; o < nout; ++o) {
      for (int k = 0; k < nin; ++k) gw[o * nin + k] += double(dz[o] * in[k]);
      gb[o] += double(dz[o]);
    }
  }
};

// Node moments about its (FP32) centre of mass, from its bodies, in double.
struct NodeMoments {
  double sxx = 0, sxy = 0, syy = 0;
};
NodeMoments momentsOf(const float *px, const float *py, const float *m,
                      int32_t s, int32_t e, double cx, double cy) {
  NodeMoments q;
  for (int32_t j = s; j < e; ++j) {
    const double dx = double(px[j]) - cx, dy = double(py[j]) - cy;
    q.sxx += m[j] * dx * dx;
    q.sxy += m[j] * dx * dy;
    q.syy += m[j] * dy * dy;
  }
  return q;
}

// The trained range: a far node (already accepted by size^2 < d^2 theta^2)
// whose size is at most OR_MLP_MAX_RATIO of its distance and whose square
// lies at least OR_MLP_BOX_MARGIN sizes away from the target. FP32, every
// operation exact or rounded once in this order (no square root), so the
// GPU's decisions are these bit for bit.
bool inRange(float x, float y, float cX, float cY, float size, float dSq) {
  const float maxSq = float(OR_MLP_MAX_RATIO) * float(OR_MLP_MAX_RATIO);
  if (!(size * size <= maxSq * dSq)) {
    return false;
  }
  const float hx = std::fmax(std::fabs(x - cX) - 0.5f * size, 0.0f);
  const float hy = std::fmax(std::fabs(y - cY) - 0.5f * size, 0.0f);
  const float g = float(OR_MLP_BOX_MARGIN) * size;
  return hx * hx + hy * hy >= g * g;
}

void features(double x, double y, double comX, double comY, double M,
              double size, const NodeMoments &q, double eps2, float *f) {
  const double dx = comX - x, dy = comY - y;
  const double r2 = dx * dx + dy * dy;
  const double w = 1.0 / (r2 + eps2);
  const double irho = std::sqrt(w);
  const double invR = 1.0 / std::sqrt(r2);
  const double qxx = q.sxx / M, qxy = q.sxy / M, qyy = q.syy / M;
  const double sdx = qxx * dx + qxy * dy, sdy = qxy * dx + qyy * dy;
  const double k = OR_MLP_MOMENT_SCALE * w;
  const double iR2 = invR * invR;
  f[0] = float(std::log2(M) / 16.0);
  f[1] = float(dx * irho);
  f[2] = float(dy * irho);
  f[3] = float(size * invR);
  f[4] = float(qxx * k);
  f[5] = float(qxy * k);
  f[6] = float(qyy * k);
  const double rad = (dx * sdx + dy * sdy) * iR2 * k;
  f[7] = float(rad);
  f[8] = float((dx * sdy - dy * sdx) * iR2 * k);
  f[9] = float((qxx + qyy) * k - rad);
}

} // namespace

extern "C" {

int64_t or_mlp_pairs(uint64_t n, const float *px, const float *py,
                     int64_t nodes, const float *comX, const float *comY,
                     const float *mass, const float *centerX,
                     const float *centerY, const float *size,
                     const int32_t *children, const int32_t *next,
                     uint64_t n_targets, const uint32_t *targets, float theta,
                     int use_range, int64_t cap, uint32_t *out_target,
                     int32_t *out_node) {
  (void)n;
  int64_t count = 0;
  const float tSq = theta * theta;
  for (uint64_t t = 0; t < n_targets; ++t) {
    const float x = px[targets[t]], y = py[targets[t]];
    int32_t node = 0;
    int64_t guard = 0;
    while (guard++ <= nodes) {
      // The acceptance test in FP32, rounded as QuadTree.cpp:276-280.
      const float dx = comX[node] - x, dy = comY[node] - y;
      const float dSq = dx * dx + dy * dy;
      bool far = size[node] * size[node] < dSq * tSq;
      if (far && use_range && mass[node] > 0.0f &&
          !inRange(x, y, centerX[node], centerY[node], size[node], dSq)) {
        far = false; // outside the trained range: opened
      }
      if (far || children[node] == 0) {
        if (far && mass[node] > 0.0f) {
          if (count < cap) {
            out_target[count] = targets[t];
            out_node[count] = node;
          }
          ++count;
        }
        if (next[node] == 0) {
          break;
        }
        node = next[node];
      } else {
        node = children[node];
      }
    }
  }
  return count;
}

void or_mlp_features(const float *px, const float *py, const float *m,
                     const float *comX, const float *comY, const float *mass,
                     const float *size, const int32_t *start,
                     const int32_t *end, uint64_t count,
                     const uint32_t *targets, const int32_t *nodes,
                     double eps2, float *feat, float *label, float *weight) {
  // Per target: the monopole sum of its pairs, |A_t|, the scale of the
  // pairs' loss weights (M / (r^2 + eps^2) / |A_t|)^2.
  std::map<uint32_t, std::pair<double, double>> acc;
  for (uint64_t k = 0; k < count; ++k) {
    const uint32_t t = targets[k];
    const int32_t nd = nodes[k];
    const double dx = double(comX[nd]) - px[t], dy = double(comY[nd]) - py[t];
    const double r2 = dx * dx + dy * dy;
    const double fm = mass[nd] / (std::sqrt(r2) * (r2 + eps2));
    auto &a = acc[t];
    a.first += dx * fm;
    a.second += dy * fm;
  }
  for (uint64_t k = 0; k < count; ++k) {
    const uint32_t t = targets[k];
    const int32_t nd = nodes[k];
    const double x = px[t], y = py[t];
    const double cx = comX[nd], cy = comY[nd], M = mass[nd];
    const NodeMoments q = momentsOf(px, py, m, start[nd], end[nd], cx, cy);
    features(x, y, cx, cy, M, size[nd], q, eps2, feat + k * OR_MLP_IN);
    // Residual of the node's bodies against its monopole, in FP64
    // (exactTreeLawForces' law, CheckCuda.cu:456-475).
    double ex = 0, ey = 0;
    for (int32_t j = start[nd]; j < end[nd]; ++j) {
      if (uint32_t(j) == t) continue;
      const double dx = double(px[j]) - x, dy = double(py[j]) - y;
      const double d2 = dx * dx + dy * dy;
      if (d2 == 0.0) continue;
      const double f = m[j] / (std::sqrt(d2) * (d2 + eps2));
      ex += dx * f;
      ey += dy * f;
    }
    const double dx = cx - x, dy = cy - y, r2 = dx * dx + dy * dy;
    const double fm = M / (std::sqrt(r2) * (r2 + eps2));
    // Output scale: the node's monopole magnitude M / (r^2 + eps^2) times
    // its quadrupole scale mu = tr(S) / (M (r^2 + eps^2)); mu = 0 (all mass
    // at the centre of mass) has a zero residual, target and weight.
    const double mu = (q.sxx + q.syy) / M / (r2 + eps2);
    const double unit = M / (r2 + eps2) * mu;
    const double norm = unit > 0.0 ? OR_MLP_OUT_SCALE / unit : 0.0;
    label[2 * k] = float((ex - dx * fm) * norm);
    label[2 * k + 1] = float((ey - dy * fm) * norm);
    if (weight) {
      const auto &a = acc[t];
      const double at = std::sqrt(a.first * a.first + a.second * a.second);
      const double share = at > 0.0 ? unit / at : 0.0;
      weight[k] = float(share * share);
    }
  }
}

void or_mlp_forward(const float *params, uint64_t count, const float *feat,
                    int bf16, float *out) {
  if (bf16) {
    Net<true> net(params);
    Net<true>::Acts a;
    for (uint64_t k = 0; k < count; ++k) {
      net.forward(feat + k * kIn, a);
      out[2 * k] = a.y[0];
      out[2 * k + 1] = a.y[1];
    }
  } else {
    Net<false> net(params);
    Net<false>::Acts a;
    for (uint64_t k = 0; k < count; ++k) {
      net.forward(feat + k * kIn, a);
      out[2 * k] = float(a.y[0]);
      out[2 * k + 1] = float(a.y[1]);
    }
  }
}

void or_mlp_loss_grad(const float *params, uint64_t count, const float *feat,
                      const float *label, const float *weight, double wnorm,
                      int bf16, double *loss, double *grad) {
  std::memset(grad, 0, sizeof(double) * OR_MLP_PARAMS);
  double l = 0;
  if (count == 0) {
    *loss = 0;
    return;
  }
  // Loss sum_k w_k |y_k - t_k|^2 / (count wnorm): pair k's delta scale is
  // w_k / (count wnorm).
  const double inv = 1.0 / (double(count) * wnorm);
  auto run = [&](auto &net, auto &a, auto one) {
    using T = decltype(one);
    const T s = T(inv);
    for (uint64_t k = 0; k < count; ++k) {
      const double w = weight ? weight[k] : 1.0;
      net.forward(feat + k * kIn, a);
      for (int o = 0; o < kOut; ++o) {
        const double e = double(a.y[o]) - label[2 * k + o];
        l += w * e * e;
      }
      net.backward(a, label + 2 * k, T(w) * s, grad);
    }
  };
  if (bf16) {
    Net<true> net(params);
    Net<true>::Acts a;
    run(net, a, 1.0f);
  } else {
    Net<false> net(params);
    Net<false>::Acts a;
    run(net, a, 1.0);
  }
  *loss = l * inv;
}

void or_mlp_accelerations(uint64_t n, const float *px, const float *py,
                          const float *m, int64_t nodes, const float *comX,
                          const float *comY, const float *mass,
                          const float *centerX, const float *centerY,
                          const float *size, const int32_t *children,
                          const int32_t *next, const int32_t *start,
                          const int32_t *end, const float *params, int bf16,
                          uint64_t n_targets, const uint32_t *targets,
                          float theta, double eps2, double *ax, double *ay,
                          uint32_t *far_count, uint32_t *pair_count) {
  (void)n;
  const float tSq = theta * theta;
  std::vector<float> f, y;
  for (uint64_t t = 0; t < n_targets; ++t) {
    const uint32_t ti = targets[t];
    const float x = px[ti], yy = py[ti];
    double sx = 0, sy = 0;
    uint32_t nFar = 0, nPair = 0;
    std::vector<double> scale;
    f.clear();
    int32_t node = 0;
    int64_t guard = 0;
    while (guard++ <= nodes) {
      const float dx = comX[node] - x, dy = comY[node] - yy;
      const float dSq = dx * dx + dy * dy;
      bool far = size[node] * size[node] < dSq * tSq;
      if (far && mass[node] > 0.0f &&
          !inRange(x, yy, centerX[node], centerY[node], size[node], dSq)) {
        far = false;
      }
      if (far) {
        if (mass[node] > 0.0f) {
          const double ddx = double(comX[node]) - x, ddy = double(comY[node]) - yy;
          const double r2 = ddx * ddx + ddy * ddy;
          const double fm = mass[node] / (std::sqrt(r2) * (r2 + eps2));
          sx += ddx * fm;
          sy += ddy * fm;
          ++nFar;
          if (params) {
            const NodeMoments q = momentsOf(px, py, m, start[node], end[node],
                                            comX[node], comY[node]);
            f.resize(f.size() + kIn);
            features(x, yy, comX[node], comY[node], mass[node], size[node], q,
                     eps2, f.data() + f.size() - kIn);
            const double mu = (q.sxx + q.syy) / mass[node] / (r2 + eps2);
            scale.push_back(mass[node] / (r2 + eps2) * mu / OR_MLP_OUT_SCALE);
          }
        }
      } else if (children[node] == 0) {
        for (int32_t j = start[node]; j < end[node]; ++j) {
          const double bx = double(px[j]) - x, by = double(py[j]) - yy;
          const double d2 = bx * bx + by * by;
          ++nPair;
          if (d2 == 0.0) continue;
          const double fb = m[j] / (std::sqrt(d2) * (d2 + eps2));
          sx += bx * fb;
          sy += by * fb;
        }
      } else {
        node = children[node];
        continue;
      }
      if (next[node] == 0) {
        break;
      }
      node = next[node];
    }
    if (params && !scale.empty()) {
      y.resize(scale.size() * 2);
      or_mlp_forward(params, scale.size(), f.data(), bf16, y.data());
      for (size_t k = 0; k < scale.size(); ++k) {
        sx += scale[k] * y[2 * k];
        sy += scale[k] * y[2 * k + 1];
      }
    }
    ax[t] = sx;
    ay[t] = sy;
    if (far_count) far_count[t] = nFar;
    if (pair_count) pair_count[t] = nPair;
  }
}

float or_bf16_round(float x) { return bf16Round(x); }

} // extern "C"
