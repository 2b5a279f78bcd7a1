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
constexpr int kSmallComp = 16;     // contact components replayed by one lane
constexpr int kBigCompMax = 4096;  // pairs per component sortable in LDS
constexpr uint32_t kDone = 0x7fffffffu;

// Device status bits (Ctl::status). Any set bit makes the step fail with
// NBX_ERR_CAPACITY; kernels downstream of an overflow skip their writes.
enum : uint32_t {
  kStatusNodeOverflow = 1u,
  kStatusPairOverflow = 2u,
  kStatusDeepTooLarge = 4u,
  kStatusComponentTooLarge = 8u,
  kStatusTreeInconsistent = 16u, // internal check failed (never expected)
  kStatusJacobiTooManyLarge = 32u,
  kStatusSortStalled = 64u, // a grid barrier of the one-launch sort timed out
};

// Per-handle control block in device memory: everything the host would
// otherwise have to read back between kernels (counts, sizes, the root quad)
// lives here, so a step needs no host round trip.
struct Ctl {
  uint32_t bbox[4];  // minX, maxX, minY, maxY as order-preserving bits
  float quad[4];     // root quad: cx, cy, size, unused
  // collide
  float cell, invCell;
  uint32_t largeCount, pairCount;
  uint32_t compCount, compPairs, bigCount, largest;
  // tree
  int32_t splits[kMaxDepth];         // splitting nodes at each depth
  int32_t levelStart[kMaxDepth + 2]; // first node index at each depth
  int32_t splitBase[kMaxDepth + 1];  // offset of each depth in the split list
  int32_t totalSplits, nodeCount, levels, deepCount;
  int32_t subtreeRoots;
  uint32_t status;
  uint32_t pad[2];
  unsigned long long farCount, bodyPairCount;
  unsigned long long waveVisits, waveLeafTerms; // counting walk only
};

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
