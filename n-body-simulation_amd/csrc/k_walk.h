// k_walk.h -- Barnes-Hut force walk (QuadTree::acceleration,
// QuadTree.cpp:264-389) and the FP64 direct sum used to judge accuracy.
//
// Wavefront-coherent traversal. A wave takes 64 consecutive bodies (the
// build leaves them in key order, so they are spatial neighbours) and walks
// the tree once for all of them: the node index is wave-uniform, so node
// records and leaf bodies arrive through scalar loads, and the only
// per-lane state is the opening decision. Each lane keeps the reference's
// own decisions: when it accepts a node (far field or leaf) but another lane
// opened it, the lane parks with `wait` = that node's `next` and sits out
// the subtree; the wave reaches `next` exactly when the stackless preorder
// walk would, so every lane visits the same nodes in the same order as a
// lone QuadTree::acceleration walk. With SUM_REFERENCE the per-lane
// arithmetic is the reference's too (four interleaved leaf accumulators,
// IEEE division), giving bit-identical forces; SUM_COMPENSATED uses the
// hardware reciprocal and Kahan accumulation.
#pragma once

#include "nbx_device.h"

namespace nbx {

constexpr int kSumReference = 0;
constexpr int kSumCompensated = 1;
constexpr int kLawReference = 0;
constexpr int kLawAccurate = 1;

template <int SUM, int LAW>
__device__ inline float lawFactor(float mass, float dSq, float eSq) {
  float inv, soft;
  if (SUM == kSumReference) {
    inv = LAW == kLawReference ? fisr(dSq) : 1.0f / sqrtf(dSq);
    soft = 1.0f / (dSq + eSq);
  } else {
    inv = LAW == kLawReference ? fisr(dSq) : __builtin_amdgcn_rsqf(dSq);
    soft = __builtin_amdgcn_rcpf(dSq + eSq);
  }
  return capMax(mass * inv * soft);
}

__device__ inline void kahanAdd(float &s, float &c, float term) {
  const float y = term - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int SUM, int LAW, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_walk(
    uint32_t n, const float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ pm, float tSq,
    float eSq, float *__restrict__ ax, float *__restrict__ ay,
    Ctl *__restrict__ ctl) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t base = blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1));
  if (base >= n || (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent))) {
    return;
  }
  const uint32_t i = base + lane;
  const bool valid = i < n;
  const float x = valid ? px[i] : 0.0f;
  const float y = valid ? py[i] : 0.0f;
  int wait = valid ? -1 : static_cast<int>(kDone);
  float sx = 0.0f, sy = 0.0f, cx = 0.0f, cy = 0.0f;
  uint32_t nFar = 0, nPair = 0, nVisit = 0, nLeafTerm = 0;
  int node = 0;
  // A preorder walk visits each node at most once: more iterations than
  // nodes, or an index out of range, means a corrupt tree -- stop and flag.
  const int nodeCount = ctl->nodeCount;
  int visits = 0;
  while (true) {
    node = __builtin_amdgcn_readfirstlane(node);
    if (node < 0 || node >= nodeCount || ++visits > nodeCount) {
      if (lane == 0) {
        atomicOr(&ctl->status, kStatusTreeInconsistent);
      }
      break;
    }
    if (wait == node) {
      wait = -1;
    }
    if (COUNT) {
      ++nVisit;
    }
    const bool active = wait == -1;
    const float4 f = nodeF[node];
    const int4 c = nodeI[node];
    const int nextTarget = c.y == 0 ? static_cast<int>(kDone) : c.y;
    const float dx = f.x - x;
    const float dy = f.y - y;
    const float dSq = dx * dx + dy * dy;
    const bool far = f.w < dSq * tSq;
    if (active && far) {
      if (f.z > 0.0f) {
        const float fac = lawFactor<SUM, LAW>(f.z, dSq, eSq);
        if (SUM == kSumReference) {
          sx = sx + dx * fac;
          sy = sy + dy * fac;
        } else {
          kahanAdd(sx, cx, dx * fac);
          kahanAdd(sy, cy, dy * fac);
        }
        if (COUNT) {
          ++nFar;
        }
      }
      wait = nextTarget;
    }
    const bool leaf = c.x == 0;
    const bool needLeaf = active && !far && leaf;
    if (__ballot(needLeaf)) {
      const int s = c.z, e = c.w;
      if (COUNT) {
        nLeafTerm += static_cast<uint32_t>(e - s);
      }
      if (SUM == kSumReference) {
        float a0x = 0, a0y = 0, a1x = 0, a1y = 0, a2x = 0, a2y = 0, a3x = 0,
              a3y = 0;
        int k = s;
        for (; k + 3 < e; k += 4) {
#define NBX_TERM(K, AX, AY)                                                    \
  {                                                                            \
    const float bx = px[K] - x;                                                \
    const float by = py[K] - y;                                                \
    const float bSq = bx * bx + by * by;                                       \
    const float fac = lawFactor<SUM, LAW>(pm[K], bSq, eSq);                    \
    AX = AX + bx * fac;                                                        \
    AY = AY + by * fac;                                                        \
  }
          NBX_TERM(k, a0x, a0y)
          NBX_TERM(k + 1, a1x, a1y)
          NBX_TERM(k + 2, a2x, a2y)
          NBX_TERM(k + 3, a3x, a3y)
        }
        for (; k < e; ++k) {
          NBX_TERM(k, a0x, a0y)
        }
#undef NBX_TERM
        if (needLeaf) {
          sx = sx + (a0x + a1x + a2x + a3x);
          sy = sy + (a0y + a1y + a2y + a3y);
        }
      } else {
        for (int k = s; k < e; ++k) {
          const float bx = px[k] - x;
          const float by = py[k] - y;
          const float bSq = bx * bx + by * by;
          const float fac = lawFactor<SUM, LAW>(pm[k], bSq, eSq);
          if (needLeaf) {
            kahanAdd(sx, cx, bx * fac);
            kahanAdd(sy, cy, by * fac);
          }
        }
      }
      if (needLeaf) {
        if (COUNT) {
          nPair += static_cast<uint32_t>(e - s);
        }
        wait = nextTarget;
      }
    }
    const bool needOpen = active && !far && !leaf;
    if (__ballot(needOpen)) {
      node = c.x;
    } else {
      if (c.y == 0) {
        break;
      }
      node = c.y;
    }
  }
  if (valid) {
    if (SUM == kSumReference) {
      ax[i] = sx;
      ay[i] = sy;
    } else {
      ax[i] = sx - cx;
      ay[i] = sy - cy;
    }
  }
  if (COUNT) {
    // Wave totals; whole waves may have left early, so no block barrier.
    unsigned long long f64 = nFar, p64 = nPair;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      f64 += __shfl_xor(f64, off, kWave);
      p64 += __shfl_xor(p64, off, kWave);
    }
    if (lane == 0) {
      atomicAdd(&ctl->farCount, f64);
      atomicAdd(&ctl->bodyPairCount, p64);
      atomicAdd(&ctl->waveVisits, static_cast<unsigned long long>(nVisit));
      atomicAdd(&ctl->waveLeafTerms, static_cast<unsigned long long>(nLeafTerm));
    }
  }
}

// Default walk (SUM_COMPENSATED): same traversal and per-lane decisions as
// k_walk, restructured for throughput --
//  * as soon as the next node is chosen, the records of both nodes that can
//    follow it (first child, `next`) and, for a leaf, its first four packed
//    {x, y, m} bodies are requested, so a full iteration of work covers the
//    pointer chase;
//  * leaf bodies are processed four at a time from unconditional loads
//    (the packed array is padded by kBodyPad records), with out-of-range
//    slots masked arithmetically instead of branched around;
//  * terms are accumulated with FMA into a per-lane partial sum that is
//    added to a Kahan-compensated total every 8 node visits and at the end.
constexpr int kFlushEvery = 8;
constexpr int kBodyPad = 4;

template <int LAW>
__device__ inline void leafTerm(const float4 &b, float x, float y, float eSq,
                                float w, float &qx, float &qy) {
  const float bx = b.x - x;
  const float by = b.y - y;
  const float bSq = bx * bx + by * by;
  const float fac = lawFactor<kSumCompensated, LAW>(b.z, bSq, eSq) * w;
  qx = fmaf(bx, fac, qx);
  qy = fmaf(by, fac, qy);
}

template <int LAW, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_walk_fast(
    uint32_t n, const float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, const float4 *__restrict__ body,
    float tSq, float eSq, float *__restrict__ ax, float *__restrict__ ay,
    Ctl *__restrict__ ctl) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t base = blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1));
  if (base >= n ||
      (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent))) {
    return;
  }
  const uint32_t i = base + lane;
  const bool valid = i < n;
  const float4 me = body[valid ? i : base];
  const float x = me.x, y = me.y;
  int wait = valid ? -1 : static_cast<int>(kDone);
  float sx = 0.0f, sy = 0.0f, cx = 0.0f, cy = 0.0f; // compensated total
  float qx = 0.0f, qy = 0.0f;                       // partial since flush
  uint32_t nFar = 0, nPair = 0, nLeafTerm = 0;
  const int nodeCount = ctl->nodeCount;
  int node = 0;
  float4 f = nodeF[0];
  int4 c = nodeI[0];
  float4 fk = nodeF[c.x];
  int4 ck = nodeI[c.x];
  float4 fn = nodeF[c.y];
  int4 cn = nodeI[c.y];
  float4 b0 = body[c.z], b1 = body[c.z + 1], b2 = body[c.z + 2],
         b3 = body[c.z + 3];
  int visits = 0;
  while (true) {
    if (node < 0 || node >= nodeCount || ++visits > nodeCount) {
      if (lane == 0) {
        atomicOr(&ctl->status, kStatusTreeInconsistent);
      }
      break;
    }
    const int kid = c.x, nxt = c.y;
    if (wait == node) {
      wait = -1;
    }
    const bool active = wait == -1;
    const int nextTarget = nxt == 0 ? static_cast<int>(kDone) : nxt;
    const float dx = f.x - x;
    const float dy = f.y - y;
    const float dSq = dx * dx + dy * dy;
    const bool far = f.w < dSq * tSq;
    const bool doFar = active && far && f.z > 0.0f;
    if (__ballot(doFar)) {
      const float fac = lawFactor<kSumCompensated, LAW>(f.z, dSq, eSq);
      if (doFar) {
        qx = fmaf(dx, fac, qx);
        qy = fmaf(dy, fac, qy);
        if (COUNT) {
          ++nFar;
        }
      }
    }
    if (active && far) {
      wait = nextTarget;
    }
    const bool leaf = kid == 0;
    const bool needLeaf = active && !far && leaf;
    if (__ballot(needLeaf)) {
      const int s = c.z, cnt = c.w - c.z;
      if (COUNT) {
        nLeafTerm += static_cast<uint32_t>(cnt);
        if (needLeaf) {
          nPair += static_cast<uint32_t>(cnt);
        }
      }
      const float use = needLeaf ? 1.0f : 0.0f;
      if (cnt > 0) {
        leafTerm<LAW>(b0, x, y, eSq, use, qx, qy);
        leafTerm<LAW>(b1, x, y, eSq, cnt > 1 ? use : 0.0f, qx, qy);
        leafTerm<LAW>(b2, x, y, eSq, cnt > 2 ? use : 0.0f, qx, qy);
        leafTerm<LAW>(b3, x, y, eSq, cnt > 3 ? use : 0.0f, qx, qy);
      }
      for (int k = 4; k < cnt; k += 4) {
        const float4 e0 = body[s + k], e1 = body[s + k + 1],
                     e2 = body[s + k + 2], e3 = body[s + k + 3];
        leafTerm<LAW>(e0, x, y, eSq, use, qx, qy);
        leafTerm<LAW>(e1, x, y, eSq, cnt > k + 1 ? use : 0.0f, qx, qy);
        leafTerm<LAW>(e2, x, y, eSq, cnt > k + 2 ? use : 0.0f, qx, qy);
        leafTerm<LAW>(e3, x, y, eSq, cnt > k + 3 ? use : 0.0f, qx, qy);
      }
      if (needLeaf) {
        wait = nextTarget;
      }
    }
    if ((visits & (kFlushEvery - 1)) == 0) {
      kahanAdd(sx, cx, qx);
      kahanAdd(sy, cy, qy);
      qx = 0.0f;
      qy = 0.0f;
    }
    const bool needOpen = active && !far && !leaf;
    if (__ballot(needOpen)) {
      node = kid;
      f = fk;
      c = ck;
    } else {
      if (nxt == 0) {
        break;
      }
      node = nxt;
      f = fn;
      c = cn;
    }
    node = __builtin_amdgcn_readfirstlane(node);
    // Requests for the iteration after next (see above).
    fk = nodeF[c.x];
    ck = nodeI[c.x];
    fn = nodeF[c.y];
    cn = nodeI[c.y];
    b0 = body[c.z];
    b1 = body[c.z + 1];
    b2 = body[c.z + 2];
    b3 = body[c.z + 3];
  }
  kahanAdd(sx, cx, qx);
  kahanAdd(sy, cy, qy);
  if (valid) {
    ax[i] = sx - cx;
    ay[i] = sy - cy;
  }
  if (COUNT) {
    unsigned long long f64 = nFar, p64 = nPair;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      f64 += __shfl_xor(f64, off, kWave);
      p64 += __shfl_xor(p64, off, kWave);
    }
    if (lane == 0) {
      atomicAdd(&ctl->farCount, f64);
      atomicAdd(&ctl->bodyPairCount, p64);
      atomicAdd(&ctl->waveVisits, static_cast<unsigned long long>(visits));
      atomicAdd(&ctl->waveLeafTerms, static_cast<unsigned long long>(nLeafTerm));
    }
  }
}

// ---- FP64 direct sum, CheckCuda.cu:456-475 (tree law, self excluded,
//      coincident bodies skipped) ----
constexpr int kExactChunks = 64;

// Block (tile of 256 targets, chunk of bodies): partial sums, staged through
// LDS 256 bodies at a time.
__global__ __launch_bounds__(kBlock) void k_exact_partial(
    uint32_t n, uint32_t nt, const uint32_t *__restrict__ targets,
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pm, double eps2, double2 *__restrict__ part) {
  __shared__ float sx[kBlock], sy[kBlock], sm[kBlock];
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = t < nt;
  const uint32_t ti = valid ? targets[t] : 0u;
  const double x = valid ? px[ti] : 0.0, y = valid ? py[ti] : 0.0;
  const uint32_t chunk = (n + kExactChunks - 1) / kExactChunks;
  const uint32_t j0 = blockIdx.y * chunk;
  const uint32_t j1 = min(n, j0 + chunk);
  double ax = 0.0, ay = 0.0;
  for (uint32_t b = j0; b < j1; b += kBlock) {
    const uint32_t j = b + threadIdx.x;
    __syncthreads();
    if (j < j1) {
      sx[threadIdx.x] = px[j];
      sy[threadIdx.x] = py[j];
      sm[threadIdx.x] = pm[j];
    }
    __syncthreads();
    const uint32_t cnt = min<uint32_t>(kBlock, j1 - b);
    for (uint32_t k = 0; k < cnt; ++k) {
      const double dx = static_cast<double>(sx[k]) - x;
      const double dy = static_cast<double>(sy[k]) - y;
      const double dSq = dx * dx + dy * dy;
      if (b + k == ti || dSq == 0.0) {
        continue;
      }
      const double f = static_cast<double>(sm[k]) / (sqrt(dSq) * (dSq + eps2));
      ax += dx * f;
      ay += dy * f;
    }
  }
  if (valid) {
    part[static_cast<size_t>(blockIdx.y) * nt + t] = make_double2(ax, ay);
  }
}

__global__ __launch_bounds__(kBlock) void k_exact_reduce(
    uint32_t nt, const double2 *__restrict__ part, double *__restrict__ ax,
    double *__restrict__ ay) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) {
    return;
  }
  double sx = 0.0, sy = 0.0;
  for (int c = 0; c < kExactChunks; ++c) {
    const double2 v = part[static_cast<size_t>(c) * nt + t];
    sx += v.x;
    sy += v.y;
  }
  ax[t] = sx;
  ay[t] = sy;
}

} // namespace nbx
