// nbx_device.h -- types and helpers shared by the gfx950 kernels of libnbx.
//
// All kernels are compiled with -ffp-contract=off: every float expression
// rounds once per operation, in source order, like the reference CPU build
// (g++ -O2, baseline x86-64). Where a kernel wants a fused multiply-add it
// calls fmaf() explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace nbx {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kLeafCapacity = 16;  // Simulation.cpp:9 leafCapacity
constexpr int kMaxDepth = 32;      // key depth: 2 words x 16 levels
constexpr int kDeepMax = 4096;     // bodies per equal-hi-key run sortable in LDS
constexpr int kSmallComp = 8;      // contact components replayed by one lane
constexpr int kBigCompMax = 4096;  // pairs per component sorted in LDS at once
// Contact components of kSmallComp+1 .. 4096 pairs are replayed by one wave
// each, in tiers by pair count (k_resolve_wave).
constexpr int kWaveTiers = 4;
constexpr int kTierQueues = 4; // size classes (queues) per wave tier
__host__ __device__ constexpr int waveTierPairs(int tier) {
  return tier == 0 ? 64 : tier == 1 ? 256 : tier == 2 ? 1024 : 4096;
}
constexpr uint32_t kDone = 0x7fffffffu;

// Device status bits (Ctl::status). Any set bit makes the step fail with
// NBX_ERR_CAPACITY; kernels downstream of an overflow skip their writes.
enum : uint32_t {
  kStatusNodeOverflow = 1u,
  kStatusPairOverflow = 2u,
  kStatusDeepTooLarge = 4u,
  kStatusTreeInconsistent = 16u, // internal check failed (never expected)
  kStatusJacobiTooManyLarge = 32u,
  kStatusBarrierStalled = 64u, // a grid barrier of a one-launch kernel timed out
};

// Per-handle control block in device memory: everything the host would
// otherwise have to read back between kernels (counts, sizes, the root quad)
// lives here, so a step needs no host round trip.
// Work profile of the counting walk (wave-level totals, nbx_walk_profile):
// blocks popped, lanes entering them, visits with a far lane, far lanes,
// plain leaf visits, plain leaf records, spread leaf visits, spread term
// evaluations, spread butterfly steps, stack pushes, empty nodes visited.
constexpr int kWalkStats = 11;

struct Ctl {
  uint32_t bbox[4];  // minX, maxX, minY, maxY as order-preserving bits
  float quad[4];     // root quad: cx, cy, size, unused
  // collide
  float cell, invCell;
  uint32_t largeCount, pairCount;
  uint32_t compCount, compPairs, bigCount, largest;
  uint32_t farMoves; // bodies the build's sort moved far (k_deep, k_permute)
  uint32_t largeFused; // large bodies counted by k_iterate_cells (moved to
                       // largeCount by k_grid_scatter, then cleared)
  // tree
  int32_t splits[kMaxDepth];         // splitting nodes at each depth
  int32_t levelStart[kMaxDepth + 2]; // first node index at each depth
  int32_t splitBase[kMaxDepth + 1];  // offset of each depth in the split list
  int32_t totalSplits, nodeCount, levels;
  int32_t subtreeRoots;
  uint32_t status;
  // Components queued per wave tier, in kTierQueues queues a tier by size
  // (queue kTierQueues t + k holds the k-th largest size class), so that a
  // tier's kernel starts its largest components first.
  uint32_t waveCount[kTierQueues * kWaveTiers];
  uint32_t tierNext[kWaveTiers];                // entries of a tier taken so far
  uint32_t prevPairs; // contact pairs the previous collide found
  uint32_t levelDone; // k_level_scan blocks finished (0 at rest)
  unsigned long long farCount, bodyPairCount;
  unsigned long long waveVisits, waveLeafTerms; // counting walk only
  unsigned long long walkStats[kWalkStats];     // counting walk only
};

// A one-launch kernel's grid barrier timed out earlier in this step: the
// kernels after it skip their work (their inputs may be partial), so a
// failed step never indexes outside its buffers.
__device__ inline bool barrierStalled(const Ctl *ctl) {
  return (ctl->status & kStatusBarrierStalled) != 0;
}

// Order-preserving float <-> uint32 map (so exact float min/max can use
// integer atomics).
__host__ __device__ inline uint32_t ordBits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ordFloat(uint32_t u) {
  const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  __builtin_memcpy(&f, &b, 4);
  return f;
}

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2. Renumbering them so that every XCD
// takes one contiguous range of the grid keeps spatially neighbouring work,
// which shares tree nodes and bodies, in one L2.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcdBlock() {
  const uint32_t b = blockIdx.x, g = gridDim.x;
  const uint32_t x = b % kXcds, k = b / kXcds;
  const uint32_t q = g / kXcds, r = g % kXcds;
  return x * q + min(x, r) + k;
}

// Number of lanes of `m` below this lane (v_mbcnt: two instructions).
__device__ inline uint32_t rankInMask(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(
      static_cast<uint32_t>(m >> 32),
      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

__device__ inline uint64_t laneMaskLt() {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of this wave holding the same 8-bit digit (among `valid` lanes).
// Must be called by every lane of the wave.
__device__ inline uint64_t matchDigit8(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Exact bounding box as order-preserving bits {min x, max x, min y, max y};
// NaN coordinates are ignored, as std::min/max ignore them in
// Quad::newContaining (QuadTree.cpp:30-51).
__device__ inline uint4 bboxEmpty() {
  return make_uint4(0xffffffffu, 0u, 0xffffffffu, 0u);
}
__device__ inline uint4 bboxAdd(uint4 b, float x, float y) {
  if (x == x) {
    const uint32_t u = ordBits(x);
    b.x = min(b.x, u);
    b.y = max(b.y, u);
  }
  if (y == y) {
    const uint32_t u = ordBits(y);
    b.z = min(b.z, u);
    b.w = max(b.w, u);
  }
  return b;
}
__device__ inline uint4 bboxMerge(uint4 a, uint4 b) {
  return make_uint4(min(a.x, b.x), max(a.y, b.y), min(a.z, b.z),
                    max(a.w, b.w));
}
__device__ inline uint4 bboxWave(uint4 b) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    b.x = min(b.x, static_cast<uint32_t>(__shfl_xor(b.x, off, kWave)));
    b.y = max(b.y, static_cast<uint32_t>(__shfl_xor(b.y, off, kWave)));
    b.z = min(b.z, static_cast<uint32_t>(__shfl_xor(b.z, off, kWave)));
    b.w = max(b.w, static_cast<uint32_t>(__shfl_xor(b.w, off, kWave)));
  }
  return b;
}

// Root quad from the bounding box (QuadTree.cpp:46-50): centre and size.
__device__ inline void rootQuad(uint4 b, float &qcx, float &qcy, float &qs) {
  const float minX = ordFloat(b.x), maxX = ordFloat(b.y);
  const float minY = ordFloat(b.z), maxY = ordFloat(b.w);
  qcx = (minX + maxX) * 0.5f;
  qcy = (minY + maxY) * 0.5f;
  const float wx = maxX - minX, wy = maxY - minY;
  qs = (wx < wy) ? wy : wx;
}

// Publishes the build's bounding box and root quad and clears its counters
// and status bits (one lane, before any build kernel after the keys).
__device__ inline void publishRoot(Ctl *ctl, uint4 b, float qcx, float qcy,
                                   float qs) {
  ctl->bbox[0] = b.x;
  ctl->bbox[1] = b.y;
  ctl->bbox[2] = b.z;
  ctl->bbox[3] = b.w;
  ctl->quad[0] = qcx;
  ctl->quad[1] = qcy;
  ctl->quad[2] = qs;
  ctl->subtreeRoots = 0;
  atomicExch(&ctl->farMoves, 0u); // counted by atomics after grid barriers
  atomicAnd(&ctl->status, ~(kStatusNodeOverflow | kStatusDeepTooLarge |
                            kStatusTreeInconsistent));
}

// Float-descent quadrant key of a body (QuadTree.cpp:57-90 partitions,
// Quad::intoQuadrant QuadTree.h:14-25): digit 2*(y >= cy) + (x >= cx) at
// each of 32 depths, the centre moved as the reference moves it; depths
// 0-15 in `hi`, 16-31 in `lo`, two bits each from the top. The reference's
// move (b - 0.5) * s is exactly +-(0.5 * s) (a scaling by 0.5, and IEEE
// rounding is symmetric in sign), so the centre takes a select of that
// half-size and its negation: the same bits in fewer instructions.
__device__ inline uint32_t descentWord(float x, float y, float &cx, float &cy,
                                       float &s) {
  uint32_t key = 0;
#pragma unroll 4
  for (int d = 0; d < 16; ++d) {
    const bool bx = !(x < cx);
    const bool by = !(y < cy);
    key = (key << 2) | (by ? 2u : 0u) | (bx ? 1u : 0u);
    s *= 0.5f;
    const float h = 0.5f * s;
    cx += bx ? h : -h;
    cy += by ? h : -h;
  }
  return key;
}

__device__ inline void descentKeys(float x, float y, float cx, float cy,
                                   float s, uint32_t &hi, uint32_t &lo) {
  hi = descentWord(x, y, cx, cy, s);
  lo = descentWord(x, y, cx, cy, s);
}

// QuadTree.cpp:11-26 fastInvSqrt: bit-level guess + one Newton step.
__device__ inline float fisr(float x) {
  const float xhalf = 0.5f * x;
  const uint32_t i = 0x5f3759dfu - (__float_as_uint(x) >> 1);
  const float y = __uint_as_float(i);
  return y * (1.5f - xhalf * y * y);
}

// std::min(a, FLT_MAX) (QuadTree.cpp:290-292): returns a when a is NaN.
__device__ inline float capMax(float a) { return (FLT_MAX < a) ? FLT_MAX : a; }

// Quadrant digit of sorted body j at `depth` (0 = root's split), from the
// sorted high key word and the unsorted low word reached through perm.
__device__ inline int digitAt(const uint32_t *__restrict__ hi,
                              const uint32_t *__restrict__ lo,
                              const uint32_t *__restrict__ perm, int j,
                              int depth) {
  if (depth < 16) {
    return static_cast<int>((hi[j] >> (30 - 2 * depth)) & 3u);
  }
  return static_cast<int>((lo[perm[j]] >> (30 - 2 * (depth - 16))) & 3u);
}

// Number of leading quadrant digits shared by sorted bodies a and b (0..32).
__device__ inline int commonDigits(const uint32_t *__restrict__ hi,
                                   const uint32_t *__restrict__ lo,
                                   const uint32_t *__restrict__ perm, int a,
                                   int b) {
  const uint32_t ha = hi[a], hb = hi[b];
  if (ha != hb) {
    return __clz(ha ^ hb) >> 1;
  }
  const uint32_t la = lo[perm[a]], lb = lo[perm[b]];
  if (la != lb) {
    return 16 + (__clz(la ^ lb) >> 1);
  }
  return 32;
}

} // namespace nbx
