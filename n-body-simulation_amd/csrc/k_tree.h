// k_tree.h -- quadtree build on the GPU, reproducing QuadTree::build
// (QuadTree.cpp:154-262) node for node.
//
// Keys. Each body gets a 32-level quadrant key by replaying the reference's
// float descent: at every level the digit is 2*(y >= cy) + (x >= cx) against
// the node centre, and the centre moves by Quad::intoQuadrant
// (QuadTree.h:14-25). The reference partitions bodies with exactly these
// comparisons, so sorting by the key groups bodies exactly as its
// partitions do. Depths 0-15 live in the high word (radix-sorted); depths
// 16-31 in the low word, which matters only inside runs of more than 16
// bodies with equal high words (sorted separately, rarely non-empty).
//
// Topology without a per-level loop. A node at depth d is a run of sorted
// bodies sharing d digits; it splits when it holds more than 16 bodies. With
// lcp(a, b) = number of digits shared by sorted bodies a and b, body j
// starts a splitting run at depth d iff
//     (j == 0 or lcp(j-1, j) < d)  and  lcp(j, j+16) >= d  and  d <= 31,
// a contiguous range of depths per body. The reference places the children
// of level d's k-th splitting node at levelEnd(d) + 4k (QuadTree.cpp:
// 176-204), and nodes within a level are in key order, so ranking the run
// starts per depth gives every node index directly: one counting pass, one
// scan per depth, one emit pass, one node-creation pass.
#pragma once

#include "k_sort.h"
#include "nbx_device.h"

namespace nbx {

// Exclusive scan across a block of BS threads (BS a multiple of 64).
template <int BS>
__device__ inline uint32_t blockScanT(uint32_t v, uint32_t *scratch,
                                      uint32_t *total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) {
      x += y;
    }
  }
  if (lane == kWave - 1) {
    scratch[wave] = x;
  }
  __syncthreads();
  uint32_t waveOff = 0, sum = 0;
  for (int w = 0; w < BS / kWave; ++w) {
    const uint32_t s = scratch[w];
    waveOff += (w < wave) ? s : 0u;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return waveOff + x - v;
}

constexpr int kBboxParts = 256; // workgroups of the bounding-box pass

// Quad::newContaining (QuadTree.cpp:30-51), stage 1: per-workgroup exact
// min/max as order-preserving bits (NaN coordinates are ignored, as
// std::min/max ignore them there).
__global__ __launch_bounds__(kBlock) void k_bbox(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    uint4 *__restrict__ parts) {
  __shared__ uint4 red[kBlock / kWave];
  uint4 b = bboxEmpty();
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n;
       i += gridDim.x * kBlock) {
    b = bboxAdd(b, px[i], py[i]);
  }
  b = bboxWave(b);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; ++w) {
      b = bboxMerge(b, red[w]);
    }
    parts[blockIdx.x] = b;
  }
}

// Root quad (QuadTree.cpp:46-50) and the float-descent key of every body.
// Stage 2 of the bounding box runs in every workgroup here (the partials
// are a few KB in L2), so no atomics and no reset are needed. Thread 0 also
// publishes the root quad and clears the build's counters.
// Body v's low key word (depths 16..31 of the float descent from the root
// quad), stored to keyLo: the one-launch sort computes only the high words,
// and the low words are read only within runs of equal high words.
__device__ inline uint32_t lowKeyWord(const Ctl *__restrict__ ctl,
                                      const float *__restrict__ px,
                                      const float *__restrict__ py,
                                      uint32_t *__restrict__ keyLo, uint32_t v) {
  float cx = ctl->quad[0], cy = ctl->quad[1], s = ctl->quad[2];
  const float x = px[v], y = py[v];
  (void)descentWord(x, y, cx, cy, s); // depths 0..15: the high word
  const uint32_t l = descentWord(x, y, cx, cy, s);
  keyLo[v] = l;
  return l;
}

// Every body's low key word (the keys download; the step computes them only
// for runs of equal high words).
__global__ __launch_bounds__(kBlock) void k_key_low(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const Ctl *__restrict__ ctl, uint32_t *__restrict__ keyLo) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v < n) {
    (void)lowKeyWord(ctl, px, py, keyLo, v);
  }
}

__global__ __launch_bounds__(kBlock) void k_keys(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const uint4 *__restrict__ parts, uint32_t nparts, Ctl *__restrict__ ctl,
    uint32_t *__restrict__ keyHi, uint32_t *__restrict__ keyLo,
    uint32_t *__restrict__ vals) {
  __shared__ uint4 red[kBlock / kWave];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint4 b = bboxWave(threadIdx.x < nparts ? parts[threadIdx.x] : bboxEmpty());
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  b = red[0];
  for (int w = 1; w < kBlock / kWave; ++w) {
    b = bboxMerge(b, red[w]);
  }
  float qcx, qcy, qs;
  rootQuad(b, qcx, qcy, qs);
  if (i == 0) {
    publishRoot(ctl, b, qcx, qcy, qs);
  }
  if (i >= n) {
    return;
  }
  uint32_t hi, lo;
  descentKeys(px[i], py[i], qcx, qcy, qs, hi, lo);
  keyHi[i] = hi;
  keyLo[i] = lo;
  vals[i] = i;
}

// Stable bitonic sort, in LDS, of perm[start, start + len) by low key word
// (len <= kDeepMax), by the whole workgroup.
__device__ inline void deepSortRun(int start, int len, const Ctl *__restrict__ ctl,
                                   const float *__restrict__ px,
                                   const float *__restrict__ py,
                                   uint32_t *__restrict__ keyLo,
                                   uint32_t *__restrict__ perm,
                                   uint32_t *__restrict__ sk,
                                   uint32_t *__restrict__ sv,
                                   int16_t *__restrict__ order) {
  int p2 = 1;
  while (p2 < len) {
    p2 <<= 1;
  }
  for (int k = threadIdx.x; k < p2; k += kBlock) {
    if (k < len) {
      sv[k] = perm[start + k];
      sk[k] = lowKeyWord(ctl, px, py, keyLo, sv[k]);
    }
    order[k] = static_cast<int16_t>(k);
  }
  __syncthreads();
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < p2; k += kBlock) {
        const int partner = k ^ stride;
        if (partner > k) {
          const bool up = (k & size) == 0;
          const int a = order[k], c = order[partner];
          // (key, position) order; padding sorts last.
          bool aGreater;
          if (a >= len || c >= len) {
            aGreater = a > c;
          } else {
            aGreater = sk[a] > sk[c] || (sk[a] == sk[c] && a > c);
          }
          if (aGreater == up) {
            order[k] = static_cast<int16_t>(c);
            order[partner] = static_cast<int16_t>(a);
          }
        }
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < len; k += kBlock) {
    perm[start + k] = sv[order[k]];
  }
  __syncthreads();
}

// Runs of bodies with equal high key words are put in low-word order, so the
// bodies end in full-key order (equal full keys keep their order, as the
// oracle's stable sort keeps them): a run of up to kLeafCapacity bodies by
// the lane at its start (insertion sort), a longer run -- whose order
// decides nodes below depth 16 -- by the whole workgroup whose range holds
// its start, in LDS, after the lanes' runs. Runs are disjoint and only a
// run's start acts on it, so no two lanes or workgroups touch the same
// entries.
// Whether k_permute gathers staged records: the sort moved more than an
// eighth of the bodies far (see kFarMove).
__device__ inline bool permuteStaged(const Ctl *__restrict__ ctl, uint32_t n) {
  return ctl->farMoves > n / 8;
}

// Also stages every body's state as one 32-byte record in its current order
// when the permute will gather those (permuteStaged).
__global__ __launch_bounds__(kBlock) void k_deep(
    uint32_t n, const uint32_t *__restrict__ hi, uint32_t *__restrict__ keyLo,
    uint32_t *__restrict__ perm, Ctl *__restrict__ ctl,
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ vx, const float *__restrict__ vy,
    const float *__restrict__ m, const float *__restrict__ r,
    const uint32_t *__restrict__ id, float4 *__restrict__ stage) {
  if (barrierStalled(ctl)) {
    return;
  }
  {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && permuteStaged(ctl, n)) {
      stage[2 * i] = make_float4(px[i], py[i], vx[i], vy[i]);
      stage[2 * i + 1] = make_float4(m[i], r[i], __uint_as_float(id[i]), 0.0f);
    }
  }
  __shared__ int2 longRuns[kBlock];
  __shared__ int nLong;
  __shared__ uint32_t sk[kDeepMax];
  __shared__ uint32_t sv[kDeepMax];
  __shared__ int16_t order[kDeepMax];
  if (threadIdx.x == 0) {
    nLong = 0;
  }
  __syncthreads();
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t h = j + 1 < n ? hi[j] : 0u;
  const bool start =
      j + 1 < n && !(j > 0 && hi[j - 1] == h) && hi[j + 1] == h;
  if (start && j + kLeafCapacity < n && hi[j + kLeafCapacity] == h) {
    uint32_t a = j + kLeafCapacity + 1, b = n;
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if (hi[mid] == h) {
        a = mid + 1;
      } else {
        b = mid;
      }
    }
    longRuns[atomicAdd(&nLong, 1)] =
        make_int2(static_cast<int>(j), static_cast<int>(a - j));
  } else if (start) {
    uint32_t p[kLeafCapacity], l[kLeafCapacity];
    int len = 0;
    while (len < kLeafCapacity && j + len < n && hi[j + len] == h) {
      const uint32_t v = perm[j + len];
      const uint32_t key = lowKeyWord(ctl, px, py, keyLo, v);
      int q = len;
      while (q > 0 && key < l[q - 1]) {
        p[q] = p[q - 1];
        l[q] = l[q - 1];
        --q;
      }
      p[q] = v;
      l[q] = key;
      ++len;
    }
    for (int k = 0; k < len; ++k) {
      perm[j + k] = p[k];
    }
  }
  __syncthreads();
  const int runs = nLong;
  for (int g = 0; g < runs; ++g) {
    const int2 seg = longRuns[g];
    if (seg.y > kDeepMax) {
      if (threadIdx.x == 0) {
        atomicOr(&ctl->status, kStatusDeepTooLarge);
      }
      continue;
    }
    deepSortRun(seg.x, seg.y, ctl, px, py, keyLo, perm, sk, sv, order);
  }
}

// Applies the sort permutation to every body array (CudaQuadTree.cu:
// 408-414 does this with seven launches; here it is one pass) and writes the
// packed {x, y, m} records the force walk reads.
__global__ __launch_bounds__(kBlock) void k_permute(
    uint32_t n, const uint32_t *__restrict__ perm,
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ vx, const float *__restrict__ vy,
    const float *__restrict__ m, const float *__restrict__ r,
    const uint32_t *__restrict__ id, float *__restrict__ opx,
    float *__restrict__ opy, float *__restrict__ ovx, float *__restrict__ ovy,
    float *__restrict__ om, float *__restrict__ orr,
    uint32_t *__restrict__ oid, float4 *__restrict__ pack,
    const Ctl *__restrict__ ctl, const float4 *__restrict__ stage) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) {
    return;
  }
  const uint32_t s = perm[j];
  float x, y, u, w, mass, rad;
  uint32_t ident;
  if (permuteStaged(ctl, n)) {
    // One 32-byte record a body (k_deep staged them) instead of seven
    // scattered words.
    const float4 a = stage[2 * s], b = stage[2 * s + 1];
    x = a.x;
    y = a.y;
    u = a.z;
    w = a.w;
    mass = b.x;
    rad = b.y;
    ident = __float_as_uint(b.z);
  } else {
    x = px[s];
    y = py[s];
    u = vx[s];
    w = vy[s];
    mass = m[s];
    rad = r[s];
    ident = id[s];
  }
  opx[j] = x;
  opy[j] = y;
  ovx[j] = u;
  ovy[j] = w;
  om[j] = mass;
  orr[j] = rad;
  oid[j] = ident;
  pack[j] = make_float4(x, y, mass, 0.0f);
}

// Depth range [A, B] over which sorted body j starts a splitting run.
__device__ inline void splitRange(uint32_t n, const uint32_t *__restrict__ hi,
                                  const uint32_t *__restrict__ lo,
                                  const uint32_t *__restrict__ perm,
                                  uint32_t j, int &A, int &B) {
  if (j >= n) {
    A = 1;
    B = 0;
    return;
  }
  A = j == 0 ? 0 : commonDigits(hi, lo, perm, j - 1, j) + 1;
  if (j + kLeafCapacity < n) {
    const int c = commonDigits(hi, lo, perm, j, j + kLeafCapacity);
    B = c < kMaxDepth - 1 ? c : kMaxDepth - 1;
  } else {
    B = -1;
  }
}

// Depths [first, last] at which some lane of the wave starts a splitting run
// (first > last when none): the per-depth ballots below only need those.
__device__ inline void waveDepthRange(int A, int B, int &first, int &last) {
  int f = A <= B ? A : kMaxDepth, l = A <= B ? B : -1;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    f = min(f, __shfl_xor(f, off, kWave));
    l = max(l, __shfl_xor(l, off, kWave));
  }
  first = __builtin_amdgcn_readfirstlane(f);
  last = __builtin_amdgcn_readfirstlane(l);
}

// Per-wave counts of splitting runs per depth into waveCnt[wave][*] (zero
// outside the wave's depth range).
__device__ inline void waveDepthCounts(int A, int B, int first, int last,
                                       uint32_t *waveCntRow) {
  const int lane = threadIdx.x & (kWave - 1);
  if (lane < kMaxDepth) {
    waveCntRow[lane] = 0;
  }
  for (int d = first; d <= last; ++d) {
    const uint64_t mask = __ballot(A <= d && d <= B);
    if (lane == 0) {
      waveCntRow[d] = static_cast<uint32_t>(__popcll(mask));
    }
  }
}

// Per 256-body tile: number of splitting runs starting in it, per depth ->
// tileCount[d * tiles + tile].
__global__ __launch_bounds__(kBlock) void k_lcp_count(
    uint32_t n, const uint32_t *__restrict__ hi,
    const uint32_t *__restrict__ lo, const uint32_t *__restrict__ perm,
    uint32_t tiles, uint32_t *__restrict__ tileCount) {
  __shared__ uint32_t waveCnt[kBlock / kWave][kMaxDepth];
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const int wave = threadIdx.x >> 6;
  int A, B;
  splitRange(n, hi, lo, perm, j, A, B);
  int first, last;
  waveDepthRange(A, B, first, last);
  waveDepthCounts(A, B, first, last, waveCnt[wave]);
  __syncthreads();
  if (threadIdx.x < kMaxDepth) {
    uint32_t s = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      s += waveCnt[w][threadIdx.x];
    }
    tileCount[threadIdx.x * tiles + blockIdx.x] = s;
  }
}

// Level bookkeeping (QuadTree.cpp:188-204) for all depths at once, the
// capacity check and the root node; split counts read at agent scope.
__device__ inline void levelFinalize(uint32_t n, int64_t capacity,
                                     Ctl *__restrict__ ctl,
                                     float4 *__restrict__ nodeF,
                                     int4 *__restrict__ nodeI,
                                     float *__restrict__ centerX,
                                     float *__restrict__ centerY,
                                     float *__restrict__ size,
                                     int32_t *__restrict__ nodeOfSplit) {
  int32_t splits[kMaxDepth];
  for (int d = 0; d < kMaxDepth; ++d) {
    splits[d] = static_cast<int32_t>(
        loadAgent(reinterpret_cast<const uint32_t *>(&ctl->splits[d])));
  }
  int32_t base = 0;
  for (int d = 0; d < kMaxDepth; ++d) {
    ctl->splitBase[d] = base;
    base += splits[d];
  }
  ctl->splitBase[kMaxDepth] = base;
  ctl->totalSplits = base;
  int64_t start = 1;
  ctl->levelStart[0] = 0;
  ctl->levelStart[1] = 1;
  int levels = 1;
  for (int d = 1; d <= kMaxDepth; ++d) {
    const int64_t cnt = 4ll * splits[d - 1];
    if (cnt > 0) {
      levels = d + 1;
    }
    start += cnt;
    ctl->levelStart[d + 1] = static_cast<int32_t>(start < 0x7fffffff ? start : 0x7fffffff);
  }
  ctl->levels = levels;
  if (start > capacity) {
    ctl->nodeCount = 0;
    ctl->status |= kStatusNodeOverflow;
    return;
  }
  ctl->nodeCount = static_cast<int32_t>(start);
  const float s = ctl->quad[2];
  nodeI[0] = make_int4(splits[0] > 0 ? 1 : 0, 0, 0, static_cast<int>(n));
  nodeF[0] = make_float4(0.0f, 0.0f, 0.0f, s * s);
  centerX[0] = ctl->quad[0];
  centerY[0] = ctl->quad[1];
  size[0] = s;
  nodeOfSplit[0] = 0;
}

// Block d: exclusive scan over tiles of depth d's counts (in place) and the
// depth total; the last block to finish then does the level bookkeeping
// (levelFinalize), so the build needs no single-thread launch. Totals are
// stored at agent scope and waited for before the block arrives at the
// counter, which the last arrival resets for the next build.
__global__ __launch_bounds__(1024) void k_level_scan(
    uint32_t tiles, uint32_t *__restrict__ tileCount, Ctl *__restrict__ ctl,
    uint32_t n, int64_t capacity, float4 *__restrict__ nodeF,
    int4 *__restrict__ nodeI, float *__restrict__ centerX,
    float *__restrict__ centerY, float *__restrict__ size,
    int32_t *__restrict__ nodeOfSplit) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ uint32_t scratch[1024 / kWave];
  const int d = blockIdx.x;
  uint32_t *row = tileCount + static_cast<size_t>(d) * tiles;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < tiles; base += 1024) {
    const uint32_t t = base + threadIdx.x;
    const uint32_t v = t < tiles ? row[t] : 0u;
    uint32_t total;
    const uint32_t pre = blockScanT<1024>(v, scratch, &total);
    if (t < tiles) {
      row[t] = carry + pre;
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    storeAgent(reinterpret_cast<uint32_t *>(&ctl->splits[d]), carry);
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0); // the total lands before arriving
    const uint32_t arrived = __hip_atomic_fetch_add(
        &ctl->levelDone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      storeAgent(&ctl->levelDone, 0u);
      levelFinalize(n, capacity, ctl, nodeF, nodeI, centerX, centerY, size,
                    nodeOfSplit);
    }
  }
}

// Writes every splitting run start, ranked within its depth, into the split
// list: splitList[splitBase[d] + rank] = j.
__global__ __launch_bounds__(kBlock) void k_emit(
    uint32_t n, const uint32_t *__restrict__ hi,
    const uint32_t *__restrict__ lo, const uint32_t *__restrict__ perm,
    uint32_t tiles, const uint32_t *__restrict__ tileOffset,
    const Ctl *__restrict__ ctl, int32_t *__restrict__ splitList) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ uint32_t waveCnt[kBlock / kWave][kMaxDepth];
  __shared__ uint32_t tileBase[kMaxDepth];
  if (ctl->status & kStatusNodeOverflow) {
    return;
  }
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const int wave = threadIdx.x >> 6;
  int A, B;
  splitRange(n, hi, lo, perm, j, A, B);
  int first, last;
  waveDepthRange(A, B, first, last);
  waveDepthCounts(A, B, first, last, waveCnt[wave]);
  // This tile's output offset per depth, staged once per workgroup.
  if (threadIdx.x < kMaxDepth) {
    tileBase[threadIdx.x] =
        ctl->splitBase[threadIdx.x] +
        tileOffset[static_cast<size_t>(threadIdx.x) * tiles + blockIdx.x];
  }
  __syncthreads();
  const uint64_t lt = laneMaskLt();
  for (int d = first; d <= last; ++d) {
    const bool flag = A <= d && d <= B;
    const uint64_t mask = __ballot(flag);
    if (flag) {
      uint32_t rank = tileBase[d] + static_cast<uint32_t>(__popcll(mask & lt));
      for (int w = 0; w < wave; ++w) {
        rank += waveCnt[w][d];
      }
      splitList[rank] = static_cast<int32_t>(j);
    }
  }
}

// The level tables (split list and node offsets per depth) staged in LDS by
// every workgroup of the kernels that look depths up per lane: a lookup is
// then a few LDS reads instead of a chain of dependent global loads.
struct LevelTables {
  int32_t splitBase[kMaxDepth + 1];
  int32_t levelStart[kMaxDepth + 2];
};
__device__ inline void stageLevels(const Ctl *__restrict__ ctl, LevelTables &lt) {
  for (int t = threadIdx.x; t < kMaxDepth + 2; t += blockDim.x) {
    if (t <= kMaxDepth) {
      lt.splitBase[t] = ctl->splitBase[t];
    }
    lt.levelStart[t] = ctl->levelStart[t];
  }
  __syncthreads();
}

// Creates the 4 children of every splitting node (QuadTree::subdivide,
// QuadTree.cpp:55-91): body ranges split at the digit boundaries, child quads
// by intoQuadrant, and each child's own `children` slot if it splits.
__global__ __launch_bounds__(kBlock) void k_nodes(
    uint32_t n, const uint32_t *__restrict__ hi,
    const uint32_t *__restrict__ lo, const uint32_t *__restrict__ perm,
    const Ctl *__restrict__ ctl, const int32_t *__restrict__ splitList,
    float4 *__restrict__ nodeF, int4 *__restrict__ nodeI,
    float *__restrict__ centerX, float *__restrict__ centerY,
    float *__restrict__ size, int32_t *__restrict__ nodeOfSplit) {
  if (barrierStalled(ctl)) {
    return;
  }
  if (ctl->status & kStatusNodeOverflow) {
    return;
  }
  const int total = ctl->totalSplits;
  if (static_cast<int>(blockIdx.x * kBlock) >= total) {
    return;
  }
  __shared__ LevelTables lt;
  stageLevels(ctl, lt);
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < total;
       s += gridDim.x * kBlock) {
    int d = 0;
    while (s >= lt.splitBase[d + 1]) {
      ++d;
    }
    const int rank = s - lt.splitBase[d];
    const int j = splitList[s];
    int bound[5];
    bound[0] = j;
    // First depth-(d+1) splitting run at or after j: this node's splitting
    // children are the next entries of that level's (sorted) list.
    const int lvBase = lt.splitBase[d + 1];
    const int lvEnd = d + 1 < kMaxDepth ? lt.splitBase[d + 2] : lvBase;
    int firstChildSplit = lvBase;
    if (d < 16) {
      // The run is every body whose high key word starts with j's d-digit
      // prefix, and its children's boundaries are where the next digit
      // reaches 1, 2, 3 and overflows: five independent lower-bound
      // searches (four on the key array, one on the split list), stepped
      // together so their loads overlap.
      const int sh = 32 - 2 * d;
      const uint64_t hj = hi[j];
      const uint64_t base = sh >= 32 ? 0ull : (hj >> sh) << sh;
      uint64_t key[4];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        key[q] = base + (static_cast<uint64_t>(q + 1) << (sh - 2));
      }
      key[3] = base + (1ull << sh);
      // Runs of one depth are disjoint and listed in key order, so this run
      // ends at or before the next depth-d split's first body.
      const int runEnd = s + 1 < lt.splitBase[d + 1]
                             ? splitList[s + 1]
                             : static_cast<int>(n);
      int a[5], b[5];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = j;
        b[q] = runEnd;
      }
      a[4] = lvBase;
      b[4] = lvEnd;
      while (a[0] < b[0] || a[1] < b[1] || a[2] < b[2] || a[3] < b[3] ||
             a[4] < b[4]) {
        // Finished searches probe index 0 (always valid) so that every load
        // is unconditional and they issue together.
        int mid[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          mid[q] = a[q] < b[q] ? (a[q] + b[q]) >> 1 : 0;
        }
        uint32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] = hi[mid[q]];
        }
        const int32_t w = splitList[mid[4]];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (a[q] < b[q]) {
            if (static_cast<uint64_t>(v[q]) < key[q]) {
              a[q] = mid[q] + 1;
            } else {
              b[q] = mid[q];
            }
          }
        }
        if (a[4] < b[4]) {
          if (w < j) {
            a[4] = mid[4] + 1;
          } else {
            b[4] = mid[4];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        bound[q + 1] = a[q];
      }
      firstChildSplit = a[4];
    } else {
      // Below depth 16 the digits come from the low key word (rare deep
      // runs): one search at a time.
      int e = static_cast<int>(n);
      {
        int a = j + kLeafCapacity + 1, b = static_cast<int>(n);
        while (a < b) {
          const int mid = (a + b) >> 1;
          if (commonDigits(hi, lo, perm, j, mid) >= d) {
            a = mid + 1;
          } else {
            b = mid;
          }
        }
        e = a;
      }
      bound[4] = e;
      for (int q = 1; q < 4; ++q) {
        int a = bound[q - 1], b = e;
        while (a < b) {
          const int mid = (a + b) >> 1;
          if (digitAt(hi, lo, perm, mid, d) < q) {
            a = mid + 1;
          } else {
            b = mid;
          }
        }
        bound[q] = a;
      }
      int a = lvBase, b = lvEnd;
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (splitList[mid] < j) {
          a = mid + 1;
        } else {
          b = mid;
        }
      }
      firstChildSplit = a;
    }
    const int e = bound[4];
    // This node's quad: replay intoQuadrant from the root.
    float cx = ctl->quad[0], cy = ctl->quad[1], sz = ctl->quad[2];
    for (int t = 0; t < d; ++t) {
      const int q = digitAt(hi, lo, perm, j, t);
      sz *= 0.5f;
      cx += (static_cast<float>(q & 1) - 0.5f) * sz;
      cy += (static_cast<float>(q >> 1) - 0.5f) * sz;
    }
    (void)e;
    const int c = lt.levelStart[d + 1] + 4 * rank;
    int nextSplit = firstChildSplit;
    const float childSize = sz * 0.5f;
    for (int q = 0; q < 4; ++q) {
      const int ch = c + q;
      const int cs = bound[q], ce = bound[q + 1];
      int grand = 0;
      if (ce - cs > kLeafCapacity && d + 1 < kMaxDepth) {
        // This child splits too: it is the next entry of the depth-(d+1)
        // split list.
        const int a = nextSplit++;
        if (a < lvEnd && splitList[a] == cs) {
          grand = lt.levelStart[d + 2] + 4 * (a - lvBase);
          nodeOfSplit[a] = ch;
        } else {
          atomicOr(&const_cast<Ctl *>(ctl)->status, kStatusTreeInconsistent);
        }
      }
      const float ccx = cx + (static_cast<float>(q & 1) - 0.5f) * childSize;
      const float ccy = cy + (static_cast<float>(q >> 1) - 0.5f) * childSize;
      centerX[ch] = ccx;
      centerY[ch] = ccy;
      size[ch] = childSize;
      nodeF[ch] = make_float4(0.0f, 0.0f, 0.0f, childSize * childSize);
      nodeI[ch] = make_int4(grand, 0, cs, ce);
    }
  }
}

// Depth of node x (x >= 1): largest D with levelStart[D] <= x.
__device__ inline int nodeDepth(const LevelTables &lt, int x) {
  int d = 1;
  while (d < kMaxDepth && lt.levelStart[d + 1] <= x) {
    ++d;
  }
  return d;
}

// Upward pass (QuadTree::propagate, QuadTree.cpp:237-262) in three
// launches instead of one per level:
//   k_node_finish  every node: `next` pointer; leaves: their sums; lists the
//                  subtree roots (branches of <= kSubtreeBodies bodies whose
//                  parent holds more);
//   k_subtree      one workgroup per subtree root: its branches, deepest
//                  level first, synchronised in LDS;
//   k_upward_top   one workgroup: the remaining branches near the root.
// Every branch value is ((c0 + c1) + c2) + c3 of its children's values, the
// reference's association, whichever kernel computes it.
constexpr int kSubtreeBodies = 4096;
constexpr int kSubtreeList = 2048; // branch nodes per subtree held in LDS

struct Sum3 {
  float x, y, m;
};

// accumulateLeaf (QuadTree.cpp:93-142): four interleaved accumulators,
// combined as ((s0 + s1) + s2) + s3.
__device__ inline Sum3 leafSum(const float *__restrict__ px,
                               const float *__restrict__ py,
                               const float *__restrict__ m, int i, int e) {
  float p0x = 0, p0y = 0, p1x = 0, p1y = 0, p2x = 0, p2y = 0, p3x = 0,
        p3y = 0;
  float m0 = 0, m1 = 0, m2 = 0, m3 = 0;
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
  return Sum3{p0x + p1x + p2x + p3x, p0y + p1y + p2y + p3y, m0 + m1 + m2 + m3};
}

__device__ inline void writeSum(int x, const Sum3 &v, float4 *__restrict__ nodeF,
                                float2 *__restrict__ wsum) {
  // propagate's final pass (QuadTree.cpp:256-261): pos /= max(mass, FLT_MIN).
  const float dv = (v.m < FLT_MIN) ? FLT_MIN : v.m;
  wsum[x] = make_float2(v.x, v.y);
  nodeF[x].x = v.x / dv;
  nodeF[x].y = v.y / dv;
  nodeF[x].z = v.m;
}

// Second mass moments of a leaf about its centre of mass (the quadrupole
// law's node data; no reference counterpart): sum m (x - c)(x - c)^T over
// the leaf's bodies, as {xx, xy, yy, xx + yy}.
__device__ inline void leafMoments(int x, const float *__restrict__ px,
                                   const float *__restrict__ py,
                                   const float *__restrict__ m, int i, int e,
                                   const float4 *__restrict__ nodeF,
                                   float4 *__restrict__ nodeQ) {
  const float cx = nodeF[x].x, cy = nodeF[x].y;
  float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
  for (; i < e; ++i) {
    const float dx = px[i] - cx, dy = py[i] - cy, w = m[i];
    sxx += w * dx * dx;
    sxy += w * dx * dy;
    syy += w * dy * dy;
  }
  nodeQ[x] = make_float4(sxx, sxy, syy, sxx + syy);
}

// Branch x from its four children (QuadTree.cpp:243-253); with nodeQ, also
// its second moments from the children's, moved to x's centre of mass
// (parallel axes: S_x = sum over children of S_c + m_c (c_c - c_x)(c_c - c_x)^T).
__device__ inline void sumBranch(int x, int c, float4 *__restrict__ nodeF,
                                 float2 *__restrict__ wsum,
                                 float4 *__restrict__ nodeQ) {
  const float2 w0 = wsum[c], w1 = wsum[c + 1], w2 = wsum[c + 2],
               w3 = wsum[c + 3];
  const Sum3 v{w0.x + w1.x + w2.x + w3.x, w0.y + w1.y + w2.y + w3.y,
               nodeF[c].z + nodeF[c + 1].z + nodeF[c + 2].z + nodeF[c + 3].z};
  writeSum(x, v, nodeF, wsum);
  if (nodeQ) {
    const float cx = nodeF[x].x, cy = nodeF[x].y;
    float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
    for (int q = 0; q < 4; ++q) {
      const float4 f = nodeF[c + q];
      const float4 k = nodeQ[c + q];
      const float dx = f.x - cx, dy = f.y - cy;
      sxx += k.x + f.z * dx * dx;
      sxy += k.y + f.z * dx * dy;
      syy += k.z + f.z * dy * dy;
    }
    nodeQ[x] = make_float4(sxx, sxy, syy, sxx + syy);
  }
}

// Parent of node x >= 1: the splitting node of depth D-1 whose children
// block holds x.
// The parent of node x at depth D >= 1.
__device__ inline int parentAt(const LevelTables &lt,
                               const int32_t *__restrict__ nodeOfSplit, int x,
                               int D) {
  return nodeOfSplit[lt.splitBase[D - 1] + (x - lt.levelStart[D]) / 4];
}

__global__ __launch_bounds__(kBlock) void k_node_finish(
    Ctl *__restrict__ ctl, const float *__restrict__ px,
    const float *__restrict__ py, const float *__restrict__ m,
    const int32_t *__restrict__ nodeOfSplit, float4 *__restrict__ nodeF,
    int4 *__restrict__ nodeI, float2 *__restrict__ wsum,
    uint8_t *__restrict__ done, int32_t *__restrict__ roots,
    float4 *__restrict__ nodeQ) {
  if (barrierStalled(ctl)) {
    return;
  }
  if (ctl->status & kStatusNodeOverflow) {
    return;
  }
  const int count = ctl->nodeCount;
  if (static_cast<int>(blockIdx.x * kBlock) >= count) {
    return;
  }
  __shared__ LevelTables lt;
  stageLevels(ctl, lt);
  for (int x = blockIdx.x * kBlock + threadIdx.x; x < count;
       x += gridDim.x * kBlock) {
    const int D = x == 0 ? 0 : nodeDepth(lt, x);
    // `next` (QuadTree.cpp:87-90): child q < 3 points at q + 1; the last
    // child inherits its parent's next (the parent of a depth-d node is at
    // depth d - 1).
    int y = x, dy = D;
    for (int guard = 0; y != 0 && ((y - 1) & 3) == 3; ++guard) {
      y = dy >= 1 ? parentAt(lt, nodeOfSplit, y, dy) : -1;
      --dy;
      if (guard > kMaxDepth || y < 0 || y >= count) {
        atomicOr(&ctl->status, kStatusTreeInconsistent);
        y = 0;
        break;
      }
    }
    const int4 ni = nodeI[x];
    nodeI[x].y = y == 0 ? 0 : y + 1;
    if (ni.x == 0) {
      writeSum(x, leafSum(px, py, m, ni.z, ni.w), nodeF, wsum);
      if (nodeQ) {
        leafMoments(x, px, py, m, ni.z, ni.w, nodeF, nodeQ);
      }
      done[x] = 1;
      continue;
    }
    done[x] = 0;
    if (ni.w - ni.z <= kSubtreeBodies) {
      bool root = x == 0;
      if (!root) {
        const int4 pi = nodeI[parentAt(lt, nodeOfSplit, x, D)];
        root = pi.w - pi.z > kSubtreeBodies;
      }
      if (root) {
        roots[atomicAdd(&ctl->subtreeRoots, 1)] = x;
      }
    }
  }
}

// One workgroup per subtree root: its branch nodes found top-down into LDS,
// level by level, then summed bottom-up (children first). A subtree with
// more branches than the LDS list holds is left to k_upward_top.
__global__ __launch_bounds__(kBlock) void k_subtree(
    const Ctl *__restrict__ ctl, const int32_t *__restrict__ roots,
    float4 *__restrict__ nodeF, const int4 *__restrict__ nodeI,
    float2 *__restrict__ wsum, uint8_t *__restrict__ done,
    float4 *__restrict__ nodeQ) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ int list[kSubtreeList];
  __shared__ int levelOff[kMaxDepth + 3];
  __shared__ int appended;
  __shared__ int overflow;
  if (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent)) {
    return;
  }
  const int nroots = ctl->subtreeRoots;
  for (int g = blockIdx.x; g < nroots; g += gridDim.x) {
    if (threadIdx.x == 0) {
      list[0] = roots[g];
      levelOff[0] = 0;
      levelOff[1] = 1;
      overflow = 0;
    }
    __syncthreads();
    int levels = 0;
    for (int d = 0; d <= kMaxDepth; ++d) {
      const int lo = levelOff[d], hi = levelOff[d + 1];
      if (lo == hi) {
        break;
      }
      levels = d + 1;
      if (threadIdx.x == 0) {
        appended = 0;
      }
      __syncthreads();
      for (int k = lo + threadIdx.x; k < hi; k += kBlock) {
        const int c = nodeI[list[k]].x;
        for (int q = 0; q < 4; ++q) {
          if (nodeI[c + q].x != 0) {
            const int slot = hi + atomicAdd(&appended, 1);
            if (slot < kSubtreeList) {
              list[slot] = c + q;
            } else {
              overflow = 1;
            }
          }
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        levelOff[d + 2] = min(hi + appended, kSubtreeList);
      }
      __syncthreads();
    }
    if (!overflow) {
      for (int d = levels - 1; d >= 0; --d) {
        for (int k = levelOff[d] + threadIdx.x; k < levelOff[d + 1];
             k += kBlock) {
          const int x = list[k];
          sumBranch(x, nodeI[x].x, nodeF, wsum, nodeQ);
          done[x] = 1;
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }
}

// The top set through global memory (any size): found top-down into `big`
// level by level, then summed bottom-up, one memory round trip per step.
__device__ inline void upwardTopGlobal(const Ctl *__restrict__ ctl,
                                       float4 *__restrict__ nodeF,
                                       const int4 *__restrict__ nodeI,
                                       float2 *__restrict__ wsum,
                                       const uint8_t *__restrict__ done,
                                       int32_t *__restrict__ big,
                                       float4 *__restrict__ nodeQ,
                                       int *__restrict__ levelOff,
                                       int &appended) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    const bool rootOpen = ctl->nodeCount > 0 && nodeI[0].x != 0 && !done[0];
    if (rootOpen) {
      big[0] = 0;
    }
    levelOff[0] = 0;
    levelOff[1] = rootOpen ? 1 : 0;
  }
  __syncthreads();
  int levels = 0;
  for (int d = 0; d <= kMaxDepth; ++d) {
    const int lo = levelOff[d], hi = levelOff[d + 1];
    if (lo == hi) {
      break;
    }
    levels = d + 1;
    if (tid == 0) {
      appended = 0;
    }
    __syncthreads();
    for (int k = lo + tid; k < hi; k += blockDim.x) {
      const int c = nodeI[big[k]].x;
      for (int q = 0; q < 4; ++q) {
        if (nodeI[c + q].x != 0 && !done[c + q]) {
          big[hi + atomicAdd(&appended, 1)] = c + q;
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      levelOff[d + 2] = hi + appended;
    }
    __syncthreads();
  }
  for (int d = levels - 1; d >= 0; --d) {
    for (int k = levelOff[d] + tid; k < levelOff[d + 1]; k += blockDim.x) {
      const int x = big[k];
      sumBranch(x, nodeI[x].x, nodeF, wsum, nodeQ);
    }
    __syncthreads();
  }
}

// One workgroup: every branch not yet summed (the nodes above the subtree
// roots, plus any subtree too large for k_subtree's list). That set is
// closed upward. When it lies among the first kTopPrefetch nodes and has at
// most one node per lane (the usual case: a few hundred nodes in the top
// levels), it is found and summed in LDS: the children index and done flag
// of those nodes are fetched at once, the set is found top-down in LDS (lane
// t owns set member t and keeps its children's set slots in registers),
// each member's finished children are fetched at once into its lane's
// registers, the levels are summed bottom-up through LDS, and every result
// is stored at the end -- a handful of memory round trips instead of a few
// per level. Otherwise upwardTopGlobal. Each value is sumBranch's, in its
// association.
constexpr int kTopBlock = 1024;
constexpr int kTopPrefetch = 4096;
__global__ __launch_bounds__(kTopBlock) void k_upward_top(
    const Ctl *__restrict__ ctl, float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, float2 *__restrict__ wsum,
    const uint8_t *__restrict__ done, int32_t *__restrict__ big,
    float4 *__restrict__ nodeQ) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ int levelOff[kMaxDepth + 3];
  __shared__ int appended;
  __shared__ int overflow;
  __shared__ int32_t pch[kTopPrefetch]; // children index, -1 once summed
  __shared__ int32_t setNode[kTopBlock];
  __shared__ float vwx[kTopBlock], vwy[kTopBlock], vm[kTopBlock];
  __shared__ float vcx[kTopBlock], vcy[kTopBlock];
  __shared__ float vqx[kTopBlock], vqy[kTopBlock], vqz[kTopBlock];
  if (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent)) {
    return;
  }
  const int tid = threadIdx.x;
  const int count = ctl->nodeCount;
  const int pre = min(count, kTopPrefetch);
  // 1. Children index of the first nodes; -1 marks a summed node (a leaf,
  //    or a branch done by k_subtree), which the set never enters.
  for (int x = tid; x < pre; x += kTopBlock) {
    const int c = nodeI[x].x;
    pch[x] = (c == 0 || done[x]) ? -1 : c;
  }
  if (tid == 0) {
    // Lane 0 wrote pch[0] itself above.
    overflow = 0;
    levelOff[0] = 0;
    levelOff[1] = (pre > 0 && pch[0] >= 0) ? 1 : 0;
    setNode[0] = 0;
  }
  __syncthreads();
  // 2. The set, top-down: lane t holds member t and its children's slots.
  int child = 0;            // first child of my member
  int slot[4] = {-1, -1, -1, -1};
  int myLevel = -1, levels = 0;
  for (int d = 0; d <= kMaxDepth; ++d) {
    const int lo = levelOff[d], hi = levelOff[d + 1];
    if (lo == hi || overflow) {
      break;
    }
    levels = d + 1;
    if (tid == 0) {
      appended = 0;
    }
    __syncthreads();
    if (tid >= lo && tid < hi) {
      myLevel = d;
      child = pch[setNode[tid]];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int y = child + q;
        if (y >= pre) {
          overflow = 1;
        } else if (pch[y] >= 0) {
          const int k = hi + atomicAdd(&appended, 1);
          if (k < kTopBlock) {
            setNode[k] = y;
            slot[q] = k;
          } else {
            overflow = 1;
          }
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      levelOff[d + 2] = min(hi + appended, kTopBlock);
    }
    __syncthreads();
  }
  if (overflow) {
    upwardTopGlobal(ctl, nodeF, nodeI, wsum, done, big, nodeQ, levelOff,
                    appended);
    return;
  }
  // 3. My member's finished children, fetched at once.
  float2 w[4] = {};
  float4 f[4] = {}, k4[4] = {};
  if (myLevel >= 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (slot[q] < 0) {
        w[q] = wsum[child + q];
        f[q] = nodeF[child + q];
        if (nodeQ) {
          k4[q] = nodeQ[child + q];
        }
      }
    }
  }
  // 4. Bottom-up through LDS; each member as sumBranch / writeSum.
  for (int d = levels - 1; d >= 0; --d) {
    if (myLevel == d) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = slot[q];
        if (k >= 0) {
          w[q] = make_float2(vwx[k], vwy[k]);
          f[q] = make_float4(vcx[k], vcy[k], vm[k], 0.f);
          k4[q] = make_float4(vqx[k], vqy[k], vqz[k], 0.f);
        }
      }
      const Sum3 v{w[0].x + w[1].x + w[2].x + w[3].x,
                   w[0].y + w[1].y + w[2].y + w[3].y,
                   f[0].z + f[1].z + f[2].z + f[3].z};
      const float dv = (v.m < FLT_MIN) ? FLT_MIN : v.m;
      const float cx = v.x / dv, cy = v.y / dv;
      vwx[tid] = v.x;
      vwy[tid] = v.y;
      vm[tid] = v.m;
      vcx[tid] = cx;
      vcy[tid] = cy;
      if (nodeQ) {
        float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float dx = f[q].x - cx, dy = f[q].y - cy;
          sxx += k4[q].x + f[q].z * dx * dx;
          sxy += k4[q].y + f[q].z * dx * dy;
          syy += k4[q].z + f[q].z * dy * dy;
        }
        vqx[tid] = sxx;
        vqy[tid] = sxy;
        vqz[tid] = syy;
      }
    }
    __syncthreads();
  }
  // 5. Results out.
  if (myLevel >= 0) {
    const int x = setNode[tid];
    wsum[x] = make_float2(vwx[tid], vwy[tid]);
    nodeF[x].x = vcx[tid];
    nodeF[x].y = vcy[tid];
    nodeF[x].z = vm[tid];
    if (nodeQ) {
      nodeQ[x] = make_float4(vqx[tid], vqy[tid], vqz[tid], vqx[tid] + vqz[tid]);
    }
  }
}

} // namespace nbx
