// k_walk.h -- Barnes-Hut force walk (QuadTree::acceleration,
// QuadTree.cpp:264-389) and the FP64 direct sum that measures its accuracy.
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

#include "k_mlp.h"
#include "nbx_device.h"

namespace nbx {

constexpr int kSumReference = 0;
constexpr int kSumCompensated = 1;
constexpr int kLawReference = 0;
constexpr int kLawAccurate = 1;
constexpr int kLawQuadrupole = 2; // accurate law + far-node quadrupole term
constexpr int kLawLearned = 3;    // accurate law + learned far-node residual

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
  if (barrierStalled(ctl)) {
    return;
  }
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

// Default walk (SUM_COMPENSATED). Only the *set* of interactions of each
// body must match the reference (its summation order is free in this mode),
// so the wave walks with an explicit stack instead of `next` pointers: an
// entry is a block of four siblings plus the 64-bit mask of lanes that
// opened their parent. Popping an entry evaluates the four siblings for
// exactly those lanes -- far field, leaf bodies, or open -- and pushes each
// opened sibling's children with the mask of lanes that opened it. Every
// lane therefore makes the reference's decisions at the reference's nodes
// (QuadTree.cpp:264-389), visiting each at most once, with the per-node
// bookkeeping of the stackless walk gone.
//  * sibling records are contiguous: one block = two 64-byte scalar loads;
//  * leaf bodies are packed {x, y, m} records read four at a time (the
//    array is padded by kBodyPad records; out-of-range slots are masked);
//  * terms are accumulated with FMA into a per-lane partial sum that is
//    added to a Kahan-compensated total after every other block.
#ifndef NBX_WALK_BLOCK
#define NBX_WALK_BLOCK 128
#endif
constexpr int kWalkBlock = NBX_WALK_BLOCK; // lanes per workgroup of k_walk_fast
constexpr int kBodyPad = 4;
constexpr int kStackDepth = 128; // >= 3 * kMaxDepth + 4 entries per wave

// A 64-bit value every lane holds (read from one LDS address), moved to
// scalar registers. readfirstlane works on int: the halves go through
// uint32_t so the low word is not sign-extended.
__device__ inline uint64_t uniform64(uint64_t v) {
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v))));
  const uint32_t hi = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v >> 32))));
  return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

// The wave's traversal stack: entries 0..63 live in VGPR lanes (entry e in
// lane e of three registers: node, mask low and high words), written by a
// lane-select move and read with v_readlane at the wave-uniform stack
// pointer -- no LDS traffic and no lane-0 stores under an exec mask. Entries
// 64..127 (only trees deeper than ~20 levels reach them) spill to this
// wave's LDS area.
struct LaneStack {
  int node = 0, lo = 0, hi = 0;
  int *spillNode;       // kStackDepth - kWave entries
  uint64_t *spillMask;  // kStackDepth - kWave entries

  __device__ inline void put(int e, int nd, uint64_t m) {
    if (e < kWave) {
      const bool mine = static_cast<int>(threadIdx.x & (kWave - 1)) == e;
      node = mine ? nd : node;
      lo = mine ? static_cast<int>(static_cast<uint32_t>(m)) : lo;
      hi = mine ? static_cast<int>(static_cast<uint32_t>(m >> 32)) : hi;
    } else if ((threadIdx.x & (kWave - 1)) == 0) {
      spillNode[e - kWave] = nd;
      spillMask[e - kWave] = m;
    }
  }
  __device__ inline void get(int e, int &nd, uint64_t &m) const {
    if (e < kWave) {
      nd = __builtin_amdgcn_readlane(node, e);
      const uint32_t mlo = static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, e));
      const uint32_t mhi = static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, e));
      m = static_cast<uint64_t>(mlo) | (static_cast<uint64_t>(mhi) << 32);
    } else {
      nd = __builtin_amdgcn_readfirstlane(spillNode[e - kWave]);
      m = uniform64(spillMask[e - kWave]);
    }
  }
};

// The stack of a tree of at most kRegStackLevels levels: all 64 entries in
// VGPR lanes, no spill branch and no overflow check. Processing a block at
// depth d leaves at most 3 pending blocks for each depth 2..d and pushes at
// most 4, so the stack never holds more than 3 * levels - 5 entries (61 at
// 22 levels). A corrupted tree can only make an entry miss the register (a
// push at e >= 64 selects no lane) -- never a memory access -- and every
// popped block index is validated before use.
constexpr int kRegStackLevels = 21;
struct RegStack {
  int node = 0, lo = 0, hi = 0;

  // v_writelane, one per field, with the lane select in M0 (gfx9 reads one
  // SGPR per VALU instruction besides M0); e is computed on the scalar unit.
  __device__ inline void put(int e, int nd, uint64_t m) {
    // The values are wave-uniform; readfirstlane states it where the
    // compiler could not prove it (a no-op on values already scalar).
    nd = __builtin_amdgcn_readfirstlane(nd);
    e = __builtin_amdgcn_readfirstlane(e);
    const int mlo = __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(m)));
    const int mhi =
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(m >> 32)));
    asm volatile(
        "v_writelane_b32 %0, %4, %3\n\t"
        "v_writelane_b32 %1, %5, %3\n\t"
        "v_writelane_b32 %2, %6, %3"
        : "+v"(node), "+v"(lo), "+v"(hi)
        : "{m0}"(e), "s"(nd), "s"(mlo), "s"(mhi));
  }
  __device__ inline void get(int e, int &nd, uint64_t &m) const {
    const int l = e & (kWave - 1);
    nd = __builtin_amdgcn_readlane(node, l);
    const uint32_t mlo = static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, l));
    const uint32_t mhi = static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, l));
    m = static_cast<uint64_t>(mlo) | (static_cast<uint64_t>(mhi) << 32);
  }
};

// CAP = false drops the min(., FLT_MAX) guard; the host selects it only when
// no factor can overflow (see capFree in the launch code).
// Reference law: m fastInvSqrt(d^2) / (d^2 + eps^2) (QuadTree.cpp:11-26,
// :276-292), with the Newton step y (1.5 - 0.5 d^2 y y) as one fused
// multiply-add (one rounding less than the reference), and computed as
// m (2 fisr) (w / 2): the doubled Newton factor 2b = 3 - (d^2 y) y is b
// doubled exactly (IEEE rounding commutes with scaling by 2) and saves the
// product -0.5 d^2; w / 2 = rcp(2 d^2 + 2 eps^2) is w halved exactly. The
// factors of 2 cancel exactly in the products, so the factor equals
// m * y (1.5 - 0.5 d^2 y y) * w with that fused step, bit for bit, for
// arguments away from overflow and the subnormal range.
template <int LAW, bool CAP>
__device__ inline float fastFactor(float mass, float dSq, float eSq) {
  float f;
  if constexpr (LAW == kLawReference) {
    const float y = __uint_as_float(0x5f3759dfu - (__float_as_uint(dSq) >> 1));
    const float inv2 = y * fmaf(-(dSq * y), y, 3.0f);
    const float w2 = __builtin_amdgcn_rcpf(fmaf(dSq, 2.0f, 2.0f * eSq));
    f = mass * inv2 * w2;
  } else {
    f = mass * __builtin_amdgcn_rsqf(dSq) * __builtin_amdgcn_rcpf(dSq + eSq);
  }
  return CAP ? capMax(f) : f;
}

template <int LAW, bool CAP>
__device__ inline void leafTerm(const float4 &b, float x, float y, float eSq,
                                float &qx, float &qy) {
  const float bx = b.x - x;
  const float by = b.y - y;
  const float fac = fastFactor<LAW, CAP>(b.z, fmaf(bx, bx, by * by), eSq);
  qx = fmaf(bx, fac, qx);
  qy = fmaf(by, fac, qy);
}

// Quadrupole term of a far node (no reference counterpart): the law is
// f(d) = d g(r), g(r) = 1 / (r (r^2 + eps^2)), d = COM - x. Expanding the
// node's bodies about its COM, the first-order term vanishes and the second
// is 1/2 sum_ab S_ab d_a d_b f(d) with S the node's second mass moments
// ({xx, xy, yy, trace}):
//   h1 (S d) + 1/2 (h1 tr S + h2 d^T S d) d,
// h1 = g'/r, h2 = (g'' - g'/r) / r^2. With u = 1/(r (r^2 + eps^2)) and
// a = 3 r^2 + eps^2: g' = -a u^2, g'' = u^2 (2 a^2 u - 6 r). (eps = 0:
// h1 = -3/r^5, h2 = 15/r^7, the Newtonian quadrupole.) invR = 1/r and
// w = 1/(r^2 + eps^2) are the monopole's; this term is free to contract.
__device__ inline void quadTerm(const float4 &sm, float dx, float dy,
                                float dSq, float eSq, float invR, float w,
                                float &qx, float &qy) {
  const float u = invR * w;
  const float a = fmaf(3.0f, dSq, eSq);
  const float u2 = u * u;
  const float h1 = -a * u2 * invR;
  const float r = dSq * invR;
  const float g2 = u2 * fmaf(2.0f * a * a, u, -6.0f * r);
  const float h2 = (g2 - h1) * (invR * invR);
  const float sdx = fmaf(sm.x, dx, sm.y * dy);
  const float sdy = fmaf(sm.y, dx, sm.z * dy);
  const float c = 0.5f * fmaf(h2, fmaf(dx, sdx, dy * sdy), h1 * sm.w);
  qx = fmaf(h1, sdx, fmaf(c, dx, qx));
  qy = fmaf(h1, sdy, fmaf(c, dy, qy));
}

// k = the lanes of `near` (non-zero), logK = ceil(log2 k), on the scalar
// unit (the compiler would test k > 1 as a 64-bit vector compare).
__device__ inline void spreadShape(uint64_t near, uint32_t &k, uint32_t &logK) {
  const uint64_t m = uniform64(near);
  int kk, lk, t;
  asm("s_bcnt1_i32_b64 %0, %3\n\t"
      "s_add_i32 %2, %0, -1\n\t"
      "s_flbit_i32_b32 %2, %2\n\t"
      "s_sub_i32 %1, 32, %2\n\t"
      "s_cmp_gt_u32 %0, 1\n\t"
      "s_cselect_b32 %1, %1, 0"
      : "=&s"(kk), "=&s"(lk), "=&s"(t)
      : "s"(m)
      : "scc");
  (void)t;
  k = static_cast<uint32_t>(kk);
  logK = static_cast<uint32_t>(lk);
}

// Whether spreading a leaf over the wave takes fewer instructions than the
// plain loop (about 18 per record there; 16 + 6 per butterfly step + 20 per
// spread evaluation): 16 + 6 (6 - logK) + 20 iters < 18 cnt. Evaluated with
// scalar instructions (the compiler would fold the multiply-add into a
// vector mad24).
__device__ inline bool spreadPays(uint32_t logK, int iters, int cnt) {
  const int lk = __builtin_amdgcn_readfirstlane(static_cast<int>(logK));
  const int it = __builtin_amdgcn_readfirstlane(iters);
  const int c = __builtin_amdgcn_readfirstlane(cnt);
  int r, t, u;
  asm("s_sub_i32 %1, 6, %4\n\t"
      "s_mul_i32 %1, %1, 6\n\t"
      "s_mul_i32 %2, %5, 20\n\t"
      "s_add_i32 %1, %1, %2\n\t"
      "s_add_i32 %1, %1, 16\n\t"
      "s_mul_i32 %2, %6, 18\n\t"
      "s_cmp_lt_i32 %1, %2\n\t"
      "s_cselect_b32 %0, 1, 0"
      : "=s"(r), "=&s"(t), "=&s"(u)
      : "0"(0), "s"(lk), "s"(it), "s"(c)
      : "scc");
  (void)t;
  (void)u;
  return r != 0;
}

// px += px[lane ^ 2^B], py likewise, when B >= logK (B < 5: ds_swizzle in
// bit-mask mode, and_mask 0x1f, xor_mask 2^B, within each 32-lane half).
template <int B>
__device__ inline void swizzleAdd(uint32_t logK, float &px, float &py) {
  static_assert(B >= 0 && B < 5, "swizzle xor within 32 lanes");
  constexpr int kPattern = 0x1f | ((1 << B) << 10);
  if (static_cast<uint32_t>(B) >= logK) {
    px += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(px), kPattern));
    py += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(py), kPattern));
  }
}

// A leaf's bodies for the lanes of `near` when few lanes need them. The
// (lane, record) pairs are spread over the wave: with k near lanes and K the
// smallest power of two >= k, lane slot r = lane % K takes the r-th near
// lane's body and group g = lane / K takes records g, g + 64/K, ... -- so
// the leaf costs ceil(cnt * K / 64) term evaluations instead of cnt. The
// groups' partial sums are added across lanes (butterfly over the group
// bits) and each near lane takes its slot's total. `slot` is this wave's
// 64-entry LDS area.
template <int LAW, bool CAP>
__device__ inline void leafSpread(uint64_t near, uint32_t k, uint32_t logK,
                                  int s, int cnt, float x, float y, float eSq,
                                  const float4 *__restrict__ body,
                                  float2 *__restrict__ slot, float &qx,
                                  float &qy) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t rank = rankInMask(near);
  const bool isNear = __builtin_amdgcn_inverse_ballot_w64(near);
  if (isNear) {
    slot[rank] = make_float2(x, y);
  }
  const uint32_t r = lane & ((1u << logK) - 1u);
  const int g = static_cast<int>(lane >> logK);
  const int groups = kWave >> logK;
  float px = 0.0f, py = 0.0f;
  if (r < k) {
    const float2 tp = slot[r];
    for (int j = g; j < cnt; j += groups) {
      leafTerm<LAW, CAP>(body[s + j], tp.x, tp.y, eSq, px, py);
    }
  }
  // Butterfly over the group bits: ds_swizzle (xor within 32 lanes, no
  // address arithmetic) for bits 0..4, ds_bpermute for bit 5.
  swizzleAdd<0>(logK, px, py);
  swizzleAdd<1>(logK, px, py);
  swizzleAdd<2>(logK, px, py);
  swizzleAdd<3>(logK, px, py);
  swizzleAdd<4>(logK, px, py);
  if (logK <= 5) {
    const int a = static_cast<int>((lane ^ 32u) << 2);
    px += __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(px)));
    py += __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(py)));
  }
  const int a = static_cast<int>(rank << 2);
  const float ax = __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(px)));
  const float ay = __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(py)));
  if (isNear) {
    qx += ax;
    qy += ay;
  }
}

// ---- Learned far-field correction inside the walk (kLawLearned). ----
// A far node inside the trained range adds its monopole at once and queues
// (features, lane) in this wave's LDS queue; every 32 queued pairs the wave
// runs the network on the matrix cores (mlpForward, weights staged in LDS)
// and adds each output, times the node's M w mu / kMlpOutScale (w = 1 /
// (r^2 + eps^2), mu = w tr(S/M)), to its lane's correction sum. Far nodes outside the range are opened instead.
constexpr int kMlpQueueCap = 96; // < 32 pending + up to 64 new
struct MlpQueue {
  bf16x8 feat[kMlpQueueCap][2];
  float scale[kMlpQueueCap];
  int owner[kMlpQueueCap];
  float2 corr[kWave];
};

struct LearnedArgs {
  const float4 *__restrict__ nodeL = nullptr;
  const float *__restrict__ centerX = nullptr;
  const float *__restrict__ centerY = nullptr;
  const float *__restrict__ size = nullptr;
  MlpQueue *q = nullptr;
  const bf16x8 *pk = nullptr; // forward operands (LDS)
  const float *bc = nullptr;  // biases (LDS)
};

// Runs the network on queue entries 0 .. n-1 (n <= 32) and adds the scaled
// outputs to their lanes' correction sums. Called by the whole wave.
__device__ __forceinline__ void mlpFlush(const LearnedArgs &la, int n) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t j = lane & 31, h = lane >> 5;
  bf16x8 x = {};
  if (static_cast<int>(j) < n) {
    x = la.q->feat[j][h];
  }
  const f32x16 y = mlpForward<false>(la.pk, la.bc, x, lane);
  if (h == 0 && static_cast<int>(j) < n) {
    const int o = la.q->owner[j];
    const float s = la.q->scale[j];
    atomicAdd(&la.q->corr[o].x, s * y[0]);
    atomicAdd(&la.q->corr[o].y, s * y[1]);
  }
}

// Queues the pairs of `farMask` (each lane its own features) and runs full
// tiles of 32. qn: the wave-uniform queue length, < 32 on entry and exit.
__device__ __forceinline__ void mlpEnqueue(const LearnedArgs &la, int &qn,
                                           uint64_t farMask, bf16x8 x0,
                                           bf16x8 x1, float scale) {
  const uint32_t cnt = static_cast<uint32_t>(__popcll(farMask));
  if (cnt == 0) {
    return;
  }
  const uint32_t lane = threadIdx.x & (kWave - 1);
  if (__builtin_amdgcn_inverse_ballot_w64(farMask)) {
    const int slot = qn + __popcll(farMask & laneMaskLt());
    la.q->feat[slot][0] = x0;
    la.q->feat[slot][1] = x1;
    la.q->scale[slot] = scale;
    la.q->owner[slot] = static_cast<int>(lane);
  }
  qn += static_cast<int>(cnt);
  while (qn >= kMlpTile) {
    ldsFence();
    mlpFlush(la, kMlpTile);
    const int rem = qn - kMlpTile;
    bf16x8 a0 = {}, a1 = {};
    float sc = 0.0f;
    int ow = 0;
    ldsFence();
    if (static_cast<int>(lane) < rem) {
      a0 = la.q->feat[kMlpTile + lane][0];
      a1 = la.q->feat[kMlpTile + lane][1];
      sc = la.q->scale[kMlpTile + lane];
      ow = la.q->owner[kMlpTile + lane];
    }
    ldsFence();
    if (static_cast<int>(lane) < rem) {
      la.q->feat[lane][0] = a0;
      la.q->feat[lane][1] = a1;
      la.q->scale[lane] = sc;
      la.q->owner[lane] = ow;
    }
    ldsFence();
    qn = rem;
  }
}

// One node for the lanes of `mask`. Lane membership is carried as wave
// masks (inverse ballots), so the branches below cost scalar instructions
// only. The opening test is rounded exactly as the reference rounds it
// (no contraction): the decisions, and so the interaction set, are the
// reference's.
template <int LAW, bool COUNT, bool CAP, bool MASS, bool POW2 = false>
__device__ inline uint64_t visitNode(const float4 &f, const int4 &c,
                                     const float4 &sm, uint64_t mask, float x,
                                     float y,
                                     float tSq, float eSq,
                                     const float4 *__restrict__ body,
                                     float2 *__restrict__ slot, float &qx,
                                     float &qy, uint32_t &nFar,
                                     uint32_t &nPair, uint32_t &nLeafTerm,
                                     const float4 &geo = float4{},
                                     const LearnedArgs &la = LearnedArgs{},
                                     int *qn = nullptr,
                                     uint32_t *ws = nullptr,
                                     int tShift = 0) {
  const float dx = f.x - x;
  const float dy = f.y - y;
  const float dSq = dx * dx + dy * dy;
  // The reference's test size^2 < d^2 theta^2. POW2 (theta^2 a power of
  // two, node sizes in the normal range): size^2 / theta^2 < d^2, the same
  // decision, both sides scaled exactly -- the scaling an exponent add on
  // the scalar unit (tShift), in place of a vector multiply per lane.
  bool farTest;
  if constexpr (POW2) {
    farTest = __int_as_float(__float_as_int(f.w) + tShift) < dSq;
  } else {
    farTest = f.w < dSq * tSq;
  }
  uint64_t near = mask & ~__builtin_amdgcn_ballot_w64(farTest);
  // f.z > 0 on the scalar unit: positive floats, +inf included, are exactly
  // the bit patterns in (0, 0x7f800000]. MASS = false (every body mass
  // positive and finite) skips the test: a non-empty node's mass is then in
  // (0, inf], and an empty node's far term is exactly 0 (its mass is 0, and
  // a far lane has dSq > 0, so the inverse root and reciprocal are finite).
  const int32_t mb = __float_as_int(f.z);
  uint64_t farMask = (!MASS || (mb > 0 && mb <= 0x7f800000))
                         ? (mask & ~near)
                         : 0ull;
  if constexpr (COUNT) {
    // Wave-level counts, outside the lane branches below (every lane holds
    // the same copy).
    if (ws && farMask) {
      ws[2] += 1;
      ws[3] += static_cast<uint32_t>(__popcll(farMask));
    }
    if (ws && c.x == 0 && c.w == c.z) {
      ws[10] += 1;
    }
  }
  if constexpr (LAW == kLawLearned) {
    // Only nodes of positive mass take the network; far lanes outside the
    // trained range open the node instead (geo = {square centre, size}).
    if (!(mb > 0 && mb <= 0x7f800000)) {
      farMask = 0ull;
    } else if (farMask) {
      const uint64_t inR =
          farMask & __builtin_amdgcn_ballot_w64(
                        mlpInRange(x, y, geo.x, geo.y, geo.z, dSq));
      near |= farMask & ~inR;
      farMask = inR;
    }
  }
  if constexpr (LAW == kLawLearned) {
    bf16x8 x0 = {}, x1 = {};
    float scale = 0.0f;
    if (__builtin_amdgcn_inverse_ballot_w64(farMask)) {
      const float invR = __builtin_amdgcn_rsqf(dSq);
      const float w = __builtin_amdgcn_rcpf(dSq + eSq);
      const float fac = capMax(f.z * invR * w);
      qx = fmaf(dx, fac, qx);
      qy = fmaf(dy, fac, qy);
      float ft[kMlpIn];
      mlpFeatures(dx, dy, dSq, eSq, geo.z, sm, ft);
      x0 = mlpInput(ft, 0);
      x1 = mlpInput(ft, 1);
      // Output unit: monopole magnitude times the quadrupole scale.
      scale = f.z * w * ((sm.x + sm.z) * w) * (1.0f / kMlpOutScale);
      if (COUNT) {
        ++nFar;
      }
    }
    mlpEnqueue(la, *qn, farMask, x0, x1, scale);
  } else if (__builtin_amdgcn_inverse_ballot_w64(farMask)) {
    if constexpr (LAW == kLawQuadrupole) {
      const float invR = __builtin_amdgcn_rsqf(dSq);
      const float w = __builtin_amdgcn_rcpf(dSq + eSq);
      const float fac0 = f.z * invR * w;
      const float fac = CAP ? capMax(fac0) : fac0;
      qx = fmaf(dx, fac, qx);
      qy = fmaf(dy, fac, qy);
      quadTerm(sm, dx, dy, dSq, eSq, invR, w, qx, qy);
    } else {
      const float fac = fastFactor<LAW, CAP>(f.z, dSq, eSq);
      qx = fmaf(dx, fac, qx);
      qy = fmaf(dy, fac, qy);
    }
    if (COUNT) {
      ++nFar;
    }
  }
  if (c.x != 0) {
    return near;
  }
  if (near != 0) {
    const int s = c.z, cnt = c.w - c.z;
    // Spread the pairs over the wave when that takes fewer instructions
    // (about 18 per record in the loop below; 16 + 6 per butterfly step +
    // 20 per spread evaluation).
    uint32_t k, logK;
    spreadShape(near, k, logK);
    const int iters = (cnt + (kWave >> logK) - 1) >> (6 - logK);
    if (spreadPays(logK, iters, cnt)) {
      if (COUNT) {
        if (ws) {
          ws[6] += 1;
          ws[7] += static_cast<uint32_t>(iters);
          ws[8] += 6u - logK;
        }
        nLeafTerm += static_cast<uint32_t>(iters);
        if (__builtin_amdgcn_inverse_ballot_w64(near)) {
          nPair += static_cast<uint32_t>(cnt);
        }
      }
      leafSpread<LAW, CAP>(near, k, logK, s, cnt, x, y, eSq, body, slot, qx,
                           qy);
      return 0;
    }
    if (COUNT) {
      if (ws) {
        ws[4] += 1;
        ws[5] += static_cast<uint32_t>(cnt);
      }
      nLeafTerm += static_cast<uint32_t>(cnt);
    }
    if (__builtin_amdgcn_inverse_ballot_w64(near)) {
      if (COUNT) {
        nPair += static_cast<uint32_t>(cnt);
      }
      // Four records per 64-byte load (the array is padded), terms only
      // for the records inside the leaf.
      for (int k = 0; k < cnt; k += 4) {
        const float4 e0 = body[s + k], e1 = body[s + k + 1],
                     e2 = body[s + k + 2], e3 = body[s + k + 3];
        asm volatile("" ::"s"(e0.x), "s"(e0.y), "s"(e0.z), "s"(e1.x),
                     "s"(e1.y), "s"(e1.z), "s"(e2.x), "s"(e2.y), "s"(e2.z),
                     "s"(e3.x), "s"(e3.y), "s"(e3.z));
        leafTerm<LAW, CAP>(e0, x, y, eSq, qx, qy);
        if (cnt > k + 1) {
          leafTerm<LAW, CAP>(e1, x, y, eSq, qx, qy);
        }
        if (cnt > k + 2) {
          leafTerm<LAW, CAP>(e2, x, y, eSq, qx, qy);
        }
        if (cnt > k + 3) {
          leafTerm<LAW, CAP>(e3, x, y, eSq, qx, qy);
        }
      }
    }
  }
  return 0;
}

// One wave's walk: 64 consecutive bodies from `base`, with stack type STACK
// (RegStack for trees of at most kRegStackLevels levels, else LaneStack).
template <int LAW, bool COUNT, bool CAP, bool MASS, class STACK,
          bool POW2 = false>
__device__ inline void walkWave(uint32_t n, uint32_t base,
                                const float4 *__restrict__ nodeF,
                                const int4 *__restrict__ nodeI,
                                const float4 *__restrict__ nodeQ,
                                const float4 *__restrict__ body, float tSq,
                                float eSq, float *__restrict__ ax,
                                float *__restrict__ ay, Ctl *__restrict__ ctl,
                                STACK &st, float2 *__restrict__ slot,
                                const LearnedArgs &la = LearnedArgs{},
                                int tShift = 0) {
  constexpr bool kSpill = sizeof(STACK) > sizeof(RegStack);
  constexpr bool kMoments = LAW == kLawQuadrupole || LAW == kLawLearned;
  int qn = 0; // learned law: queued pairs
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t i = base + lane;
  const bool valid = i < n;
  const float4 me = body[valid ? i : base];
  const float x = me.x, y = me.y;
  float sx = 0.0f, sy = 0.0f, cx = 0.0f, cy = 0.0f; // compensated total
  float qx = 0.0f, qy = 0.0f;                       // partial since flush
  uint32_t nFar = 0, nPair = 0, nLeafTerm = 0, nVisit = 0;
  uint32_t ws[kWalkStats] = {}; // COUNT only: wave-level work profile
  const int nodeCount = ctl->nodeCount;
  const uint32_t cbLimit =
      nodeCount >= 5 ? static_cast<uint32_t>(nodeCount - 4) : 0u;
  if constexpr (LAW == kLawLearned) {
    la.q->corr[lane] = make_float2(0.0f, 0.0f);
  }

  // The root on its own, then blocks of four siblings from the stack.
  int sp = 0;
  {
    const uint64_t all = __builtin_amdgcn_ballot_w64(valid);
    const float4 q0 = kMoments ? nodeQ[0] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (LAW == kLawLearned) {
      g0 = make_float4(la.centerX[0], la.centerY[0], la.size[0], 0.0f);
    }
    const uint64_t open =
        visitNode<LAW, COUNT, CAP, MASS, POW2>(nodeF[0], nodeI[0], q0, all, x, y, tSq, eSq,
                                   body, slot, qx, qy, nFar, nPair, nLeafTerm,
                                   g0, la, &qn, COUNT ? ws : nullptr, tShift);
    nVisit += 1;
    if (open) {
      st.put(0, nodeI[0].x, open);
      sp = 1;
    }
  }
  int blocks = 0;
  while (sp > 0) {
    --sp;
    int cb;
    uint64_t mask;
    st.get(sp, cb, mask);
    // A valid block lies in [1, nodeCount - 4] and each is popped at most
    // once: one unsigned compare each (nodeCount < 5 leaves no valid block).
    ++blocks;
    if ((static_cast<uint32_t>(cb - 1) >= cbLimit) |
        (static_cast<uint32_t>(blocks) > static_cast<uint32_t>(nodeCount))) {
      if (lane == 0) {
        atomicOr(&ctl->status, kStatusTreeInconsistent);
      }
      break;
    }
    // The four sibling records (two 64-byte blocks) are fetched together and
    // made available at once, rather than field by field where each is used.
    float4 bf[4];
    int4 bi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bf[q] = nodeF[cb + q];
      bi[q] = nodeI[cb + q];
    }
    asm volatile("" ::"s"(bf[0].x), "s"(bf[0].y), "s"(bf[0].z), "s"(bf[0].w),
                 "s"(bf[1].x), "s"(bf[1].y), "s"(bf[1].z), "s"(bf[1].w),
                 "s"(bf[2].x), "s"(bf[2].y), "s"(bf[2].z), "s"(bf[2].w),
                 "s"(bf[3].x), "s"(bf[3].y), "s"(bf[3].z), "s"(bf[3].w));
    asm volatile("" ::"s"(bi[0].x), "s"(bi[0].y), "s"(bi[0].z), "s"(bi[0].w),
                 "s"(bi[1].x), "s"(bi[1].y), "s"(bi[1].z), "s"(bi[1].w),
                 "s"(bi[2].x), "s"(bi[2].y), "s"(bi[2].z), "s"(bi[2].w),
                 "s"(bi[3].x), "s"(bi[3].y), "s"(bi[3].z), "s"(bi[3].w));
    float4 bq[4], bg[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bq[q] = kMoments ? nodeQ[cb + q] : make_float4(0.f, 0.f, 0.f, 0.f);
      bg[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (LAW == kLawLearned) {
        bg[q] = make_float4(la.centerX[cb + q], la.centerY[cb + q], la.size[cb + q], 0.0f);
      }
    }
    uint64_t open[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      open[q] = visitNode<LAW, COUNT, CAP, MASS, POW2>(bf[q], bi[q], bq[q], mask, x, y, tSq,
                                           eSq, body, slot, qx, qy, nFar, nPair,
                                           nLeafTerm, bg[q], la, &qn,
                                           COUNT ? ws : nullptr, tShift);
    }
    if (COUNT) {
      nVisit += 4;
      ws[0] += 1;
      ws[1] += static_cast<uint32_t>(__popcll(mask));
    }
    // Push in reverse so the first child's subtree is walked next.
#pragma unroll
    for (int q = 3; q >= 0; --q) {
      if (open[q]) {
        if (kSpill && sp >= kStackDepth) {
          if (lane == 0) {
            atomicOr(&ctl->status, kStatusTreeInconsistent);
          }
          sp = 0;
          break;
        }
        st.put(sp, bi[q].x, open[q]);
        ++sp;
        if (COUNT) {
          ws[9] += 1;
        }
      }
    }
    if ((blocks & 1) == 0) {
      kahanAdd(sx, cx, qx);
      kahanAdd(sy, cy, qy);
      qx = 0.0f;
      qy = 0.0f;
    }
  }
  kahanAdd(sx, cx, qx);
  kahanAdd(sy, cy, qy);
  if constexpr (LAW == kLawLearned) {
    if (qn > 0) {
      ldsFence();
      mlpFlush(la, qn);
    }
    ldsFence();
    const float2 corr = la.q->corr[lane];
    if (valid) {
      ax[i] = (sx - cx) + corr.x;
      ay[i] = (sy - cy) + corr.y;
    }
  } else if (valid) {
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
      atomicAdd(&ctl->waveVisits, static_cast<unsigned long long>(nVisit));
      atomicAdd(&ctl->waveLeafTerms, static_cast<unsigned long long>(nLeafTerm));
#pragma unroll
      for (int k = 0; k < kWalkStats; ++k) {
        atomicAdd(&ctl->walkStats[k], static_cast<unsigned long long>(ws[k]));
      }
    }
  }
}

// BLOCK: lanes per workgroup (64, 128 or 256), the autotuner's choice
// (nbx_tune_walk); every size gives the same results.
// tShift >= 0: theta^2 = 2^-(tShift >> 23) (a power of two), and the
// reference-law force walk compares size^2 / theta^2 with d^2 when the root
// size puts every node's size^2 and its scaled value in the normal range
// (root size in [2^-31, 2^60]: depth-32 sizes squared >= 2^-126).
template <int LAW, bool COUNT, bool CAP, bool MASS, int BLOCK = kWalkBlock>
__global__ __launch_bounds__(BLOCK) void k_walk_fast(
    uint32_t n, const float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, const float4 *__restrict__ nodeQ,
    const float4 *__restrict__ body,
    float tSq, float eSq, float *__restrict__ ax, float *__restrict__ ay,
    Ctl *__restrict__ ctl, int tShift = -1) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ float2 slots[BLOCK / kWave][kWave];
  __shared__ int spillNode[BLOCK / kWave][kStackDepth - kWave];
  __shared__ uint64_t spillMask[BLOCK / kWave][kStackDepth - kWave];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t base =
      xcdBlock() * BLOCK + (threadIdx.x & ~(kWave - 1));
  if (base >= n ||
      (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent))) {
    return;
  }
  if (ctl->levels <= kRegStackLevels) {
    RegStack st;
    if constexpr (LAW == kLawReference && !COUNT) {
      const float rootSize = ctl->quad[2];
      if (tShift >= 0 && rootSize >= 0x1p-31f && rootSize <= 0x1p60f) {
        walkWave<LAW, COUNT, CAP, MASS, RegStack, true>(
            n, base, nodeF, nodeI, nodeQ, body, tSq, eSq, ax, ay, ctl, st,
            slots[w], LearnedArgs{}, tShift);
        return;
      }
    }
    walkWave<LAW, COUNT, CAP, MASS>(n, base, nodeF, nodeI, nodeQ, body, tSq, eSq, ax, ay,
                              ctl, st, slots[w]);
  } else {
    LaneStack st;
    st.spillNode = spillNode[w];
    st.spillMask = spillMask[w];
    walkWave<LAW, COUNT, CAP, MASS>(n, base, nodeF, nodeI, nodeQ, body, tSq, eSq, ax, ay,
                              ctl, st, slots[w]);
  }
}

// The learned-law walk: k_walk_fast's traversal with the network evaluated
// for in-range far nodes (nodeL = {Sxx/M, Sxy/M, Syy/M, log2(M)/16} per
// node; pkG / bcG the packed forward weights and biases, staged in LDS).
constexpr int kWalkLearnedBlock = 256;
template <bool COUNT, bool MASS>
__global__ __launch_bounds__(kWalkLearnedBlock) void k_walk_learned(
    uint32_t n, const float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, const float4 *__restrict__ nodeL,
    const float *__restrict__ centerX, const float *__restrict__ centerY,
    const float *__restrict__ size, const float4 *__restrict__ body,
    const bf16x8 *__restrict__ pkG, const float *__restrict__ bcG, float tSq,
    float eSq, float *__restrict__ ax, float *__restrict__ ay,
    Ctl *__restrict__ ctl) {
  constexpr int W = kWalkLearnedBlock / kWave;
  __shared__ bf16x8 pk[kPkFwd];
  __shared__ float bc[kBcAll];
  __shared__ MlpQueue queues[W];
  __shared__ float2 slots[W][kWave];
  __shared__ int spillNode[W][kStackDepth - kWave];
  __shared__ uint64_t spillMask[W][kStackDepth - kWave];
  for (int k = threadIdx.x; k < kPkFwd; k += kWalkLearnedBlock) {
    pk[k] = pkG[k];
  }
  for (int k = threadIdx.x; k < kBcAll; k += kWalkLearnedBlock) {
    bc[k] = bcG[k];
  }
  __syncthreads();
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t base =
      xcdBlock() * kWalkLearnedBlock + (threadIdx.x & ~(kWave - 1));
  if (base >= n ||
      (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent))) {
    return;
  }
  LearnedArgs la;
  la.nodeL = nodeL;
  la.centerX = centerX;
  la.centerY = centerY;
  la.size = size;
  la.q = &queues[w];
  la.pk = pk;
  la.bc = bc;
  if (ctl->levels <= kRegStackLevels) {
    RegStack st;
    walkWave<kLawLearned, COUNT, true, MASS>(n, base, nodeF, nodeI, nodeL, body, tSq,
                                             eSq, ax, ay, ctl, st, slots[w], la);
  } else {
    LaneStack st;
    st.spillNode = spillNode[w];
    st.spillMask = spillMask[w];
    walkWave<kLawLearned, COUNT, true, MASS>(n, base, nodeF, nodeI, nodeL, body, tSq,
                                             eSq, ax, ay, ctl, st, slots[w], la);
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

// Per-body FP64 potential of the tree law, the partial sum over one chunk
// of sources: sum over j != i of m_j phi(|x_j - x_i|), with
// phi(r) = -(pi/2 - atan(r / eps)) / eps, whose gradient is the walk's
// force m d / (|d| (|d|^2 + eps^2)); phi(r) = -1/r when eps = 0 (then
// coincident bodies are skipped, as in the direct-sum forces).
__global__ __launch_bounds__(kBlock) void k_energy_partial(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pm, double eps, double *__restrict__ part) {
  __shared__ float sx[kBlock], sy[kBlock], sm[kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  const double x = valid ? px[i] : 0.0, y = valid ? py[i] : 0.0;
  const uint32_t chunk = (n + kExactChunks - 1) / kExactChunks;
  const uint32_t j0 = blockIdx.y * chunk;
  const uint32_t j1 = min(n, j0 + chunk);
  const double halfPi = 1.5707963267948966;
  double phi = 0.0;
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
      if (b + k == i) {
        continue;
      }
      const double dx = static_cast<double>(sx[k]) - x;
      const double dy = static_cast<double>(sy[k]) - y;
      const double r = sqrt(dx * dx + dy * dy);
      if (eps > 0.0) {
        phi -= static_cast<double>(sm[k]) * (halfPi - atan(r / eps)) / eps;
      } else if (r > 0.0) {
        phi -= static_cast<double>(sm[k]) / r;
      }
    }
  }
  if (valid) {
    part[static_cast<size_t>(blockIdx.y) * n + i] = phi;
  }
}

__global__ __launch_bounds__(kBlock) void k_energy_reduce(
    uint32_t n, const double *__restrict__ part, double *__restrict__ phi) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) {
    return;
  }
  double s = 0.0;
  for (int c = 0; c < kExactChunks; ++c) {
    s += part[static_cast<size_t>(c) * n + i];
  }
  phi[i] = s;
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
