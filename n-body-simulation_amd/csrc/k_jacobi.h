// k_jacobi.h -- the reference GPU collide (src/cuda/CudaCollision.cu:205-296,
// src/cuda/CollisionKernels.cuh:16-289), an optional mode beside the
// sequential CPU semantics of k_collide.h.
//
// Every overlapping pair is resolved from the start-of-collide state, with
// resolve()'s math and orientation (body 1 = the later box by (minX, index)),
// and each body combines its own side of all its contacts: none -> unchanged,
// one -> exactly that pair's result, several -> start + sum of the changes.
// The results depend on the order of those float sums, so the kernels keep
// the reference's scan order bit for bit:
//  * small boxes are sorted stably by a row-major 64-bit cell key (row cy,
//    then column cx; boxes wider than a cell get the largest key) with the
//    body index as payload -- here two stable 32-bit LSD sorts, low word
//    then high word;
//  * a small box scans the three cell rows dy = -1, 0, 1 of its stencil, each
//    one contiguous key range [(cx-1, cy+dy), (cx+1, cy+dy)] of the sorted
//    array, then every large box in sorted order;
//  * a large box scans all bodies in chunks of 4096, 128 lanes striding
//    through each chunk, a fixed tree reduction per chunk, chunks summed in
//    order.
#pragma once

#include "nbx_device.h"

namespace nbx {

constexpr uint64_t kJacLargeKey = ~0ull;
constexpr int kJacLanes = 128;
constexpr int kJacChunk = 4096;
constexpr int kJacPartialBytes = 16 << 20; // (large box, chunk) partial sums

struct JacBody {
  float px, py, vx, vy, m, r;
};

// One body's side of its contacts (CollisionKernels.cuh:190-196).
struct JacAcc {
  float dpx, dpy, dvx, dvy, npx, npy, nvx, nvy;
  int contacts, asLater;
};

__device__ inline JacAcc jacZero() {
  return JacAcc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0, 0};
}

__device__ inline JacBody jacLoad(int i, const float *__restrict__ px,
                                  const float *__restrict__ py,
                                  const float *__restrict__ vx,
                                  const float *__restrict__ vy,
                                  const float *__restrict__ m,
                                  const float *__restrict__ r) {
  return JacBody{px[i], py[i], vx[i], vy[i], m[i], r[i]};
}

// Cell coordinate in float, clamped to +-2^29 (NaN maps to -2^29),
// CollisionKernels.cuh:50-55.
__device__ inline int jacCell(float v, float invCell) {
  return static_cast<int>(
      fminf(fmaxf(floorf(v * invCell), -536870912.0f), 536870912.0f));
}

// Row-major key, CollisionKernels.cuh:57-63.
__device__ inline uint64_t jacKey(int cx, int cy) {
  const int64_t bias = 1ll << 30;
  return (static_cast<uint64_t>(cy + bias) << 32) |
         static_cast<uint64_t>(cx + bias);
}

__device__ inline uint64_t jacBoxKey(const Box &b, float cell, float invCell) {
  if (b.maxX - b.minX > cell || b.maxY - b.minY > cell) {
    return kJacLargeKey;
  }
  return jacKey(jacCell(b.minX, invCell), jacCell(b.minY, invCell));
}

// resolve() returning both bodies' new state (CollisionKernels.cuh:104-187);
// body a is the pair's later box. False when the circles do not touch.
__device__ inline bool jacResolve(const JacBody &a, const JacBody &b,
                                  float &p1x, float &p1y, float &v1x,
                                  float &v1y, float &p2x, float &p2y,
                                  float &v2x, float &v2y) {
  float dx = b.px - a.px, dy = b.py - a.py;
  const float rr = a.r + b.r;
  if (dx * dx + dy * dy > rr * rr) {
    return false;
  }
  const float wx = b.vx - a.vx, wy = b.vy - a.vy;
  float dv = dx * wx + dy * wy;
  const float w1 = b.m / (a.m + b.m), w2 = a.m / (a.m + b.m);
  const bool nz = dx != 0.0f || dy != 0.0f;
  v1x = a.vx;
  v1y = a.vy;
  v2x = b.vx;
  v2y = b.vy;
  if (dv >= 0.0f && nz) {
    const float sc = rr / sqrtf(dx * dx + dy * dy) - 1.0f;
    const float tx = dx * sc, ty = dy * sc;
    p1x = a.px - w1 * tx;
    p1y = a.py - w1 * ty;
    p2x = b.px + w2 * tx;
    p2y = b.py + w2 * ty;
    return true;
  }
  const float vSq = wx * wx + wy * wy;
  if (vSq < 1e-12f) {
    float ux = 1.0f, uy = 0.0f;
    if (nz) {
      const float inv = 1.0f / sqrtf(dx * dx + dy * dy);
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
  const float t = (dv + sqrtf(fmaxf(dv * dv - vSq * (dSq - rSq), 0.0f))) / vSq;
  const float q1x = a.px - a.vx * t, q1y = a.py - a.vy * t;
  const float q2x = b.px - b.vx * t, q2y = b.py - b.vy * t;
  dx = q2x - q1x;
  dy = q2y - q1y;
  dv = dx * wx + dy * wy;
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

// Body a's side of the pair (a, b), CollisionKernels.cuh:199-236.
__device__ inline void jacPair(const Box &boxA, int a, const JacBody &sa,
                               int b, const float *__restrict__ px,
                               const float *__restrict__ py,
                               const float *__restrict__ vx,
                               const float *__restrict__ vy,
                               const float *__restrict__ m,
                               const float *__restrict__ r, JacAcc &acc) {
  const Box boxB = bodyBox(px[b], py[b], r[b]);
  const bool aLater = beforeBox(boxB.minX, b, boxA.minX, a);
  if (!(aLater ? overlapsBox(boxB, boxA) : overlapsBox(boxA, boxB))) {
    return;
  }
  if (aLater) {
    ++acc.asLater;
  }
  const JacBody sb = jacLoad(b, px, py, vx, vy, m, r);
  float p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y;
  const bool hit =
      aLater ? jacResolve(sa, sb, p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y)
             : jacResolve(sb, sa, p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y);
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

// writeResult (CollisionKernels.cuh:267-287) into the staging record.
__device__ inline void jacWrite(int a, const JacBody &sa, const JacAcc &acc,
                                float4 *__restrict__ out,
                                int32_t *__restrict__ contacts,
                                int32_t *__restrict__ asLater) {
  float4 o = make_float4(sa.px, sa.py, sa.vx, sa.vy);
  if (acc.contacts == 1) {
    o = make_float4(acc.npx, acc.npy, acc.nvx, acc.nvy);
  } else if (acc.contacts > 1) {
    o = make_float4(sa.px + acc.dpx, sa.py + acc.dpy, sa.vx + acc.dvx,
                    sa.vy + acc.dvy);
  }
  out[a] = o;
  contacts[a] = acc.contacts;
  asLater[a] = acc.asLater;
}

// Keys (keyKernel, CudaCollision.cu:17-33): low and high words, identity
// payload; thread 0 also resets the collide counters.
__global__ __launch_bounds__(kBlock) void k_jac_keys(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ r, float cell, Ctl *__restrict__ ctl,
    uint32_t *__restrict__ keyLo, uint32_t *__restrict__ keyHi,
    uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    ctl->cell = cell;
    ctl->invCell = 1.0f / cell;
    ctl->largeCount = 0;
    ctl->pairCount = 0;
    ctl->compCount = 0;
    ctl->compPairs = 0;
    ctl->bigCount = 0;
    for (int t = 0; t < kTierQueues * kWaveTiers; ++t) {
      ctl->waveCount[t] = 0;
    }
    for (int t = 0; t < kWaveTiers; ++t) {
      ctl->tierNext[t] = 0;
    }
    ctl->largest = 0;
    ctl->status &= ~kStatusJacobiTooManyLarge;
  }
  if (i >= n) {
    return;
  }
  const uint64_t k = jacBoxKey(bodyBox(px[i], py[i], r[i]), cell, 1.0f / cell);
  keyLo[i] = static_cast<uint32_t>(k);
  keyHi[i] = static_cast<uint32_t>(k >> 32);
  vals[i] = i;
}

// Second sort's input: the high key words in the order of the first sort.
__global__ __launch_bounds__(kBlock) void k_jac_gather_hi(
    uint32_t n, const uint32_t *__restrict__ perm,
    const uint32_t *__restrict__ keyHi, uint32_t *__restrict__ out) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k < n) {
    out[k] = keyHi[perm[k]];
  }
}

// Sorted 64-bit keys, and the number of large boxes (they sort last): the
// position where the largest key starts is written by exactly one thread.
__global__ __launch_bounds__(kBlock) void k_jac_sorted(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ r, const uint32_t *__restrict__ perm,
    Ctl *__restrict__ ctl, uint64_t *__restrict__ keysSorted) {
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) {
    return;
  }
  const float cell = ctl->cell, invCell = ctl->invCell;
  const uint32_t i = perm[k];
  const uint64_t key = jacBoxKey(bodyBox(px[i], py[i], r[i]), cell, invCell);
  keysSorted[k] = key;
  const bool large = key == kJacLargeKey;
  bool prevLarge = false;
  if (k > 0) {
    const uint32_t j = perm[k - 1];
    prevLarge = jacBoxKey(bodyBox(px[j], py[j], r[j]), cell, invCell) ==
                kJacLargeKey;
  }
  if (large && !prevLarge) {
    ctl->largeCount = n - k;
  }
}

__device__ inline int jacLowerBound(const uint64_t *__restrict__ keys,
                                    int count, uint64_t key) {
  int lo = 0, hi = count;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < key) {
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  return lo;
}

// smallBodyKernel (CudaCollision.cu:39-60): one thread per small box, in
// sorted order.
__global__ __launch_bounds__(kBlock) void k_jac_small(
    uint32_t n, const Ctl *__restrict__ ctl,
    const uint64_t *__restrict__ keysSorted,
    const uint32_t *__restrict__ indexSorted, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ vx,
    const float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r, float4 *__restrict__ out,
    int32_t *__restrict__ contacts, int32_t *__restrict__ asLater) {
  if (barrierStalled(ctl)) {
    return;
  }
  const int s = static_cast<int>(xcdBlock() * kBlock + threadIdx.x);
  const int nn = static_cast<int>(n);
  const int nSmall = nn - static_cast<int>(ctl->largeCount);
  if (s >= nSmall) {
    return;
  }
  const float invCell = ctl->invCell;
  const int a = static_cast<int>(indexSorted[s]);
  const JacBody sa = jacLoad(a, px, py, vx, vy, m, r);
  const Box boxA = bodyBox(sa.px, sa.py, sa.r);
  const int cx = jacCell(boxA.minX, invCell), cy = jacCell(boxA.minY, invCell);
  JacAcc acc = jacZero();
  for (int dy = -1; dy <= 1; ++dy) {
    const uint64_t last = jacKey(cx + 1, cy + dy);
    for (int e = jacLowerBound(keysSorted, nSmall, jacKey(cx - 1, cy + dy));
         e < nSmall && keysSorted[e] <= last; ++e) {
      const int b = static_cast<int>(indexSorted[e]);
      if (b != a) {
        jacPair(boxA, a, sa, b, px, py, vx, vy, m, r, acc);
      }
    }
  }
  for (int e = nSmall; e < nn; ++e) {
    jacPair(boxA, a, sa, static_cast<int>(indexSorted[e]), px, py, vx, vy, m,
            r, acc);
  }
  jacWrite(a, sa, acc, out, contacts, asLater);
}

// Sum of a 128-lane workgroup's accumulators with the reference's tree
// shape (blockSum, CudaCollision.cu:62-78: stride 64, 32, ..., 1, lane t
// adding lane t + stride), all ten fields at once. The stride-64 step is one
// LDS exchange (the second wave hands its values to the first); strides 32
// to 1 stay inside the first wave: ds_bpermute for 32 (across the wave's
// halves), ds_swizzle xor 16..1 (within 32 lanes, no address arithmetic).
// For t < stride, t ^ stride = t + stride, so every step adds the same two
// partial sums the reference adds and the total is bit-identical. One
// barrier per call; `buf` alternates between two 64-entry areas by call
// parity, so the next call's writes cannot reach an area still being read.
// The total is valid in lane 0 (the reference also uses only thread 0's).
template <int B>
__device__ inline float jacSwz(float v) {
  return __int_as_float(
      __builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1f | ((1 << B) << 10)));
}
template <int B> __device__ inline int jacSwz(int v) {
  return __builtin_amdgcn_ds_swizzle(v, 0x1f | ((1 << B) << 10));
}
__device__ inline void jacAddInto(JacAcc &s, const JacAcc &o) {
  s.dpx += o.dpx;
  s.dpy += o.dpy;
  s.dvx += o.dvx;
  s.dvy += o.dvy;
  s.npx += o.npx;
  s.npy += o.npy;
  s.nvx += o.nvx;
  s.nvy += o.nvy;
  s.contacts += o.contacts;
  s.asLater += o.asLater;
}
template <int B> __device__ inline void jacSwzStep(JacAcc &s) {
  JacAcc o;
  o.dpx = jacSwz<B>(s.dpx);
  o.dpy = jacSwz<B>(s.dpy);
  o.dvx = jacSwz<B>(s.dvx);
  o.dvy = jacSwz<B>(s.dvy);
  o.npx = jacSwz<B>(s.npx);
  o.npy = jacSwz<B>(s.npy);
  o.nvx = jacSwz<B>(s.nvx);
  o.nvy = jacSwz<B>(s.nvy);
  o.contacts = jacSwz<B>(s.contacts);
  o.asLater = jacSwz<B>(s.asLater);
  jacAddInto(s, o);
}
__device__ inline JacAcc jacBlockSum(const JacAcc &v, JacAcc *__restrict__ buf) {
  static_assert(kJacLanes == 2 * kWave, "two waves per workgroup");
  const int t = static_cast<int>(threadIdx.x);
  const int lane = t & (kWave - 1);
  if (t >= kWave) {
    buf[lane] = v;
  }
  __syncthreads();
  JacAcc s = v;
  if (t < kWave) {
    jacAddInto(s, buf[lane]); // stride 64
    const int a = (lane ^ 32) << 2;
    auto perm = [&](float x) {
      return __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(x)));
    };
    JacAcc o;
    o.dpx = perm(s.dpx);
    o.dpy = perm(s.dpy);
    o.dvx = perm(s.dvx);
    o.dvy = perm(s.dvy);
    o.npx = perm(s.npx);
    o.npy = perm(s.npy);
    o.nvx = perm(s.nvx);
    o.nvy = perm(s.nvy);
    o.contacts = __builtin_amdgcn_ds_bpermute(a, s.contacts);
    o.asLater = __builtin_amdgcn_ds_bpermute(a, s.asLater);
    jacAddInto(s, o); // stride 32
    jacSwzStep<4>(s);  // 16
    jacSwzStep<3>(s);  // 8
    jacSwzStep<2>(s);  // 4
    jacSwzStep<1>(s);  // 2
    jacSwzStep<0>(s);  // 1
  }
  return s;
}

// largeBodyPartialKernel (CudaCollision.cu:133-163): workgroup (c, y) sums
// chunk c against large boxes y, y + gridDim.y, ...
__global__ __launch_bounds__(kJacLanes) void k_jac_large_partial(
    uint32_t n, uint32_t chunks, uint32_t maxLarge, Ctl *__restrict__ ctl,
    const uint32_t *__restrict__ indexSorted, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ vx,
    const float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r, JacAcc *__restrict__ partials) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ JacAcc sh[2][kWave];
  const uint32_t large = ctl->largeCount;
  if (large > maxLarge) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      atomicOr(&ctl->status, kStatusJacobiTooManyLarge);
    }
    return;
  }
  const int nn = static_cast<int>(n);
  const int c = static_cast<int>(blockIdx.x);
  const int begin = c * kJacChunk, end = min(nn, begin + kJacChunk);
  int parity = 0;
  for (uint32_t l = blockIdx.y; l < large; l += gridDim.y, parity ^= 1) {
    const int a = static_cast<int>(indexSorted[n - large + l]);
    const JacBody sa = jacLoad(a, px, py, vx, vy, m, r);
    const Box boxA = bodyBox(sa.px, sa.py, sa.r);
    JacAcc mine = jacZero();
    for (int b = begin + static_cast<int>(threadIdx.x); b < end;
         b += kJacLanes) {
      if (b != a) {
        jacPair(boxA, a, sa, b, px, py, vx, vy, m, r, mine);
      }
    }
    const JacAcc t = jacBlockSum(mine, sh[parity]);
    if (threadIdx.x == 0) {
      partials[static_cast<size_t>(l) * chunks + c] = t;
    }
  }
}

// largeBodyCombineKernel (CudaCollision.cu:167-196): chunks in order.
__global__ __launch_bounds__(kBlock) void k_jac_large_combine(
    uint32_t n, uint32_t chunks, uint32_t maxLarge, const Ctl *__restrict__ ctl,
    const uint32_t *__restrict__ indexSorted, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ vx,
    const float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r, const JacAcc *__restrict__ partials,
    float4 *__restrict__ out, int32_t *__restrict__ contacts,
    int32_t *__restrict__ asLater) {
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t large = ctl->largeCount;
  if (large > maxLarge) {
    return;
  }
  for (uint32_t l = blockIdx.x * kBlock + threadIdx.x; l < large;
       l += gridDim.x * kBlock) {
    const int a = static_cast<int>(indexSorted[n - large + l]);
    JacAcc t = jacZero();
    for (uint32_t c = 0; c < chunks; ++c) {
      const JacAcc &p = partials[static_cast<size_t>(l) * chunks + c];
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
    jacWrite(a, jacLoad(a, px, py, vx, vy, m, r), t, out, contacts, asLater);
  }
}

// Every body's output was written by one of the kernels above: it becomes
// the state (the reference swaps its output buffers in).
__global__ __launch_bounds__(kBlock) void k_jac_commit(
    uint32_t n, const Ctl *__restrict__ ctl, const float4 *__restrict__ out,
    float *__restrict__ px, float *__restrict__ py, float *__restrict__ vx,
    float *__restrict__ vy) {
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || (ctl->status & kStatusJacobiTooManyLarge)) {
    return;
  }
  const float4 o = out[i];
  px[i] = o.x;
  py[i] = o.y;
  vx[i] = o.z;
  vy[i] = o.w;
}

} // namespace nbx
