// k_ensemble.h -- many independent simulations stepped together: the step
// of CudaSimulation::step (src/cuda/CudaSimulation.cu:51-58) for every
// instance of an ensemble in two dispatches, whatever the number of
// instances.
//
//  * k_ens_prepare, a persistent kernel: each 1024-lane workgroup takes
//    whole instances from a queue (largest first) and runs, with workgroup
//    barriers only, iterate -> collide -> build for each: the broadphase
//    over a hashed cell grid (the reference's pair set, Collision.cpp:
//    135-330), contact components by union-find, each component's pairs
//    replayed in pairLess order (Collision.cpp:333-348) by one lane, the
//    float-descent keys, a stable LSD radix sort of the 64-bit keys, the
//    permutation, the level-by-level tree (QuadTree.cpp:154-235) and the
//    upward pass (QuadTree.cpp:237-262).
//  * k_ens_walk: one flat grid over every instance's bodies, 64 bodies a
//    wave, each wave walking its own instance's tree with the single
//    simulation's wave walk (walkWave).
//
// Every phase computes what the single-simulation kernels compute, in the
// same arithmetic: the same sequential-replay results, the same stable key
// order, the same node records and sums, the same per-wave walk -- so an
// instance stepped here is bit-identical to the same bodies stepped by a
// handle of its own.
#pragma once

#include "k_collide.h"
#include "k_tree.h"
#include "k_walk.h"
#include "nbx_device.h"

namespace nbx {

constexpr int kEnsBlock = 1024;
constexpr int kEnsWaves = kEnsBlock / kWave;

// One instance's storage (all in one device arena) and its state.
struct EnsInst {
  uint32_t n;          // bodies
  uint32_t cur;        // body buffer holding the state
  uint32_t buckets;    // grid hash buckets (power of two)
  uint32_t pairCap;    // pair capacity
  int32_t nodeCap;     // node capacity
  float cell, invCell; // grid cell: 1.01 x the median box width
  uint32_t waveBase;   // first wave of this instance in k_ens_walk
  float *px[2], *py[2], *vx[2], *vy[2], *m[2], *r[2];
  uint32_t *id[2];
  float *ax, *ay;
  float4 *body;        // {x, y, m, 0} in tree order (+ kBodyPad)
  uint32_t *keyHi[2], *keyLo[2], *val[2];
  uint32_t *bucketCount, *bucketStart, *bucketRank, *bucketOf;
  float4 *grid;        // 2 float4 per small box
  int32_t *largeList;
  float4 *largeBox;
  int4 *pairs, *seg;   // pairs as found; grouped by component
  int32_t *parent;     // union-find
  uint32_t *compCount, *compStart;
  float4 *nodeF;
  int4 *nodeI;
  float *centerX, *centerY, *size;
  float2 *wsum;
  Ctl *ctl;            // nodeCount, levels, levelStart, pairCount, status
};

struct EnsParams {
  float dt, tSq, eSq;
  int collisions;
};

// ---- workgroup helpers ----

__device__ inline uint32_t ensScan(uint32_t v, uint32_t *scratch, uint32_t *total) {
  return blockScanT<kEnsBlock>(v, scratch, total);
}

// Exclusive scan of a[0..count) in place (counts -> starts); a[count] = sum.
__device__ inline void ensScanArray(uint32_t *a, uint32_t count, uint32_t *scratch) {
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) {
    carry = 0;
  }
  __syncthreads();
  for (uint32_t base = 0; base < count; base += kEnsBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < count ? a[i] : 0u;
    uint32_t total;
    const uint32_t pre = ensScan(v, scratch, &total);
    const uint32_t c = carry;
    if (i < count) {
      a[i] = c + pre;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry = c + total;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a[count] = carry;
  }
  __syncthreads();
}

// Grid bucket of a cell: any fixed hash works (entries carry their exact
// cell, so cells sharing a bucket are told apart).
__device__ inline uint32_t ensBucket(int32_t cx, int32_t cy, uint32_t mask) {
  return (static_cast<uint32_t>(cx) * 0x9E3779B1u ^
          static_cast<uint32_t>(cy) * 0x85EBCA77u) & mask;
}

__device__ inline int ensFind(int32_t *parent, int x) {
  while (true) {
    const int p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
    if (p == x) {
      return x;
    }
    const int gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
    if (gp != p) {
      atomicCAS(&parent[x], p, gp); // path halving
    }
    x = p;
  }
}

// Links the components of a and b: the larger root under the smaller.
__device__ inline void ensUnion(int32_t *parent, int a, int b) {
  while (true) {
    int ra = ensFind(parent, a), rb = ensFind(parent, b);
    if (ra == rb) {
      return;
    }
    if (ra < rb) {
      const int t = ra;
      ra = rb;
      rb = t;
    }
    if (atomicCAS(&parent[ra], ra, rb) == ra) {
      return;
    }
  }
}

// In-place heap sort of a component's pairs by pairLess (one lane).
__device__ inline void ensSiftDown(int4 *a, int root, int len) {
  while (true) {
    int c = 2 * root + 1;
    if (c >= len) {
      return;
    }
    if (c + 1 < len && pairLess(a[c], a[c + 1])) {
      ++c;
    }
    if (!pairLess(a[root], a[c])) {
      return;
    }
    const int4 t = a[root];
    a[root] = a[c];
    a[c] = t;
    root = c;
  }
}
__device__ inline void ensSortPairs(int4 *a, int len) {
  if (len <= 16) {
    for (int i = 1; i < len; ++i) {
      const int4 v = a[i];
      int j = i - 1;
      while (j >= 0 && pairLess(v, a[j])) {
        a[j + 1] = a[j];
        --j;
      }
      a[j + 1] = v;
    }
    return;
  }
  for (int i = len / 2 - 1; i >= 0; --i) {
    ensSiftDown(a, i, len);
  }
  for (int e = len - 1; e > 0; --e) {
    const int4 t = a[0];
    a[0] = a[e];
    a[e] = t;
    ensSiftDown(a, 0, e);
  }
}

// ---- phases of one instance (called by the whole workgroup) ----

__device__ inline void ensIterate(const EnsInst &I, float dt) {
  const int c = I.cur;
  for (uint32_t i = threadIdx.x; i < I.n; i += kEnsBlock) {
    const float nvx = I.vx[c][i] + I.ax[i] * dt;
    const float nvy = I.vy[c][i] + I.ay[i] * dt;
    I.vx[c][i] = nvx;
    I.vy[c][i] = nvy;
    I.px[c][i] = I.px[c][i] + nvx * dt;
    I.py[c][i] = I.py[c][i] + nvy * dt;
  }
  __syncthreads();
}

// Collision.cpp:135-349 on this instance: the pair set of the hashed grid
// (small boxes against the 3x3 cells around their minimum corner, large
// boxes against every box; each pair from its `later` side), connected
// components, and each component's pairs replayed in pairLess order.
__device__ inline void ensCollide(const EnsInst &I, uint32_t *scratch) {
  __shared__ uint32_t sLarge, sPairs;
  const int c = I.cur;
  float *px = I.px[c], *py = I.py[c], *vx = I.vx[c], *vy = I.vy[c];
  const float *m = I.m[c], *r = I.r[c];
  const uint32_t n = I.n, mask = I.buckets - 1;
  const float cell = I.cell, invCell = I.invCell;
  if (threadIdx.x == 0) {
    sLarge = 0;
    sPairs = 0;
  }
  for (uint32_t h = threadIdx.x; h <= I.buckets; h += kEnsBlock) {
    I.bucketCount[h] = 0;
  }
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    I.parent[i] = static_cast<int32_t>(i);
    I.compCount[i] = 0;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    const Box b = bodyBox(px[i], py[i], r[i]);
    if (b.maxX - b.minX > cell || b.maxY - b.minY > cell) {
      const uint32_t slot = atomicAdd(&sLarge, 1u);
      I.largeList[slot] = static_cast<int32_t>(i);
      I.largeBox[slot] = make_float4(b.minX, b.maxX, b.minY, b.maxY);
      I.bucketOf[i] = 0xffffffffu;
    } else {
      const uint32_t h =
          ensBucket(cellCoord(b.minX, invCell), cellCoord(b.minY, invCell), mask);
      I.bucketOf[i] = h;
      I.bucketRank[i] = atomicAdd(&I.bucketCount[h], 1u);
    }
  }
  __syncthreads();
  ensScanArray(I.bucketCount, I.buckets, scratch); // counts -> starts
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    const uint32_t h = I.bucketOf[i];
    if (h == 0xffffffffu) {
      continue;
    }
    const Box b = bodyBox(px[i], py[i], r[i]);
    const uint32_t slot = I.bucketCount[h] + I.bucketRank[i];
    I.grid[2 * slot] = make_float4(b.minX, b.maxX, b.minY, b.maxY);
    I.grid[2 * slot + 1] = make_float4(__int_as_float(cellCoord(b.minX, invCell)),
                                       __int_as_float(cellCoord(b.minY, invCell)),
                                       __int_as_float(static_cast<int32_t>(i)), 0.0f);
  }
  __syncthreads();
  const uint32_t large = sLarge;
  auto emit = [&](int later, int earlier, float laterMinX, float earlierMinX) {
    const uint32_t slot = atomicAdd(&sPairs, 1u);
    if (slot < I.pairCap) {
      I.pairs[slot] = make_int4(later, earlier, __float_as_int(laterMinX),
                                __float_as_int(earlierMinX));
    }
  };
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    const int self = static_cast<int>(i);
    const Box me = bodyBox(px[i], py[i], r[i]);
    const bool small = I.bucketOf[i] != 0xffffffffu;
    if (small) {
      const int32_t mx = cellCoord(me.minX, invCell);
      const int32_t my = cellCoord(me.minY, invCell);
      for (int k = 0; k < 9; ++k) {
        const int32_t cx = mx + (k % 3) - 1, cy = my + (k / 3) - 1;
        const uint32_t h = ensBucket(cx, cy, mask);
        const uint32_t e = I.bucketCount[h + 1];
        for (uint32_t o = I.bucketCount[h]; o < e; ++o) {
          const float4 ob = I.grid[2 * o];
          const float4 oc = I.grid[2 * o + 1];
          const int other = __float_as_int(oc.z);
          if (__float_as_int(oc.x) != cx || __float_as_int(oc.y) != cy ||
              other == self) {
            continue;
          }
          if (beforeBox(ob.x, other, me.minX, self) &&
              overlapsBox(Box{ob.x, ob.y, ob.z, ob.w}, me)) {
            emit(self, other, me.minX, ob.x);
          }
        }
      }
    }
    for (uint32_t k = 0; k < large; ++k) {
      const int l = I.largeList[k];
      if (l == self) {
        continue;
      }
      const float4 lbv = I.largeBox[k];
      const Box lb{lbv.x, lbv.y, lbv.z, lbv.w};
      if (beforeBox(lb.minX, l, me.minX, self)) {
        if (small && overlapsBox(lb, me)) {
          emit(self, l, me.minX, lb.minX);
        }
      } else if (overlapsBox(me, lb)) {
        emit(l, self, lb.minX, me.minX);
      }
    }
  }
  __syncthreads();
  const uint32_t found = sPairs;
  if (threadIdx.x == 0) {
    I.ctl->pairCount = found;
    if (found > I.pairCap) {
      I.ctl->status |= kStatusPairOverflow;
    }
  }
  if (found > I.pairCap) {
    __syncthreads();
    return;
  }
  for (uint32_t k = threadIdx.x; k < found; k += kEnsBlock) {
    const int4 p = I.pairs[k];
    ensUnion(I.parent, p.x, p.y);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    I.parent[i] = ensFind(I.parent, static_cast<int>(i));
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < found; k += kEnsBlock) {
    atomicAdd(&I.compCount[I.parent[I.pairs[k].x]], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    I.compStart[i] = I.compCount[i];
  }
  __syncthreads();
  ensScanArray(I.compStart, n, scratch); // component segment starts
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    I.compCount[i] = 0; // fill cursors
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < found; k += kEnsBlock) {
    const int4 p = I.pairs[k];
    const int root = I.parent[p.x];
    I.seg[I.compStart[root] + atomicAdd(&I.compCount[root], 1u)] = p;
  }
  __syncthreads();
  // One lane per component: its pairs in pairLess order, replayed in turn.
  for (uint32_t rt = threadIdx.x; rt < n; rt += kEnsBlock) {
    const uint32_t len = I.compCount[rt];
    if (len == 0) {
      continue;
    }
    int4 *a = I.seg + I.compStart[rt];
    ensSortPairs(a, static_cast<int>(len));
    for (uint32_t k = 0; k < len; ++k) {
      const int4 p = a[k];
      resolvePair(px, py, vx, vy, m, r, p.x, p.y);
    }
  }
  __syncthreads();
}

// Stable LSD radix sort of (hi, lo) with payload, 8 passes of 8 bits, in
// tiles of kEnsBlock items processed in order; each item's rank among equal
// digits comes from wave ballots and a per-digit prefix over the waves.
// The result is in buffer 0.
__device__ inline void ensSort(const EnsInst &I) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t waveCnt[kEnsWaves][256];
  const uint32_t n = I.n;
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
  int src = 0;
  for (int pass = 0; pass < 8; ++pass) {
    const uint32_t *kh = I.keyHi[src], *kl = I.keyLo[src], *kv = I.val[src];
    uint32_t *oh = I.keyHi[src ^ 1], *ol = I.keyLo[src ^ 1], *ov = I.val[src ^ 1];
    const uint32_t *word = pass < 4 ? kl : kh;
    const int shift = 8 * (pass & 3);
    for (uint32_t d = threadIdx.x; d < 256; d += kEnsBlock) {
      hist[d] = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
      atomicAdd(&hist[(word[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    bool uniform = false;
    for (uint32_t d = 0; d < 256; ++d) {
      if (hist[d] == n) {
        uniform = true;
      }
    }
    // Every thread has read the counts before thread 0 turns them into
    // starts (a start can equal n too): all take the same branch.
    __syncthreads();
    if (uniform) {
      continue; // every key has this digit: the order stays
    }
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int d = 0; d < 256; ++d) {
        const uint32_t v = hist[d];
        hist[d] = run;
        run += v;
      }
    }
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kEnsBlock) {
      const uint32_t i = base + threadIdx.x;
      const bool valid = i < n;
      const uint32_t d = valid ? (word[i] >> shift) & 255u : 0u;
      for (uint32_t k = threadIdx.x; k < kEnsWaves * 256; k += kEnsBlock) {
        (&waveCnt[0][0])[k] = 0;
      }
      __syncthreads();
      const uint64_t same = matchDigit8(d, valid);
      const uint32_t rank = static_cast<uint32_t>(__popcll(same & laneMaskLt()));
      if (valid && rank == 0) {
        waveCnt[w][d] = static_cast<uint32_t>(__popcll(same));
      }
      __syncthreads();
      for (uint32_t dd = threadIdx.x; dd < 256; dd += kEnsBlock) {
        uint32_t run = hist[dd];
        for (int ww = 0; ww < kEnsWaves; ++ww) {
          const uint32_t v = waveCnt[ww][dd];
          waveCnt[ww][dd] = run;
          run += v;
        }
        hist[dd] = run;
      }
      __syncthreads();
      if (valid) {
        const uint32_t pos = waveCnt[w][d] + rank;
        oh[pos] = kh[i];
        ol[pos] = kl[i];
        ov[pos] = kv[i];
      }
      __syncthreads();
    }
    (void)lane;
    src ^= 1;
  }
  if (src == 1) {
    for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
      I.keyHi[0][i] = I.keyHi[1][i];
      I.keyLo[0][i] = I.keyLo[1][i];
      I.val[0][i] = I.val[1][i];
    }
  }
  __syncthreads();
}

__device__ inline int ensDigit(const uint32_t *hi, const uint32_t *lo, int j, int d) {
  return d < 16 ? static_cast<int>((hi[j] >> (30 - 2 * d)) & 3u)
                : static_cast<int>((lo[j] >> (30 - 2 * (d - 16))) & 3u);
}

// QuadTree::build on this instance: bounding box and root quad, keys,
// sort, permutation into the other body buffer, level-by-level nodes, and
// the upward pass.
__device__ inline void ensBuild(const EnsInst &I, uint32_t *scratch) {
  __shared__ uint4 red[kEnsWaves];
  __shared__ float quad[3];
  __shared__ int32_t levelStart[kMaxDepth + 2];
  __shared__ int32_t sLevels, sOverflow;
  const uint32_t n = I.n;
  const int c = I.cur;
  // Bounding box and root quad (QuadTree.cpp:30-51).
  uint4 b = bboxEmpty();
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    b = bboxAdd(b, I.px[c][i], I.py[c][i]);
  }
  b = bboxWave(b);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 t = red[0];
    for (int w = 1; w < kEnsWaves; ++w) {
      t = bboxMerge(t, red[w]);
    }
    float qcx, qcy, qs;
    rootQuad(t, qcx, qcy, qs);
    quad[0] = qcx;
    quad[1] = qcy;
    quad[2] = qs;
    I.ctl->quad[0] = qcx;
    I.ctl->quad[1] = qcy;
    I.ctl->quad[2] = qs;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kEnsBlock) {
    uint32_t hi, lo;
    descentKeys(I.px[c][i], I.py[c][i], quad[0], quad[1], quad[2], hi, lo);
    I.keyHi[0][i] = hi;
    I.keyLo[0][i] = lo;
    I.val[0][i] = i;
  }
  __syncthreads();
  ensSort(I);
  // Permutation into the other buffer (k_permute's gather).
  const int o = c ^ 1;
  for (uint32_t j = threadIdx.x; j < n; j += kEnsBlock) {
    const uint32_t s = I.val[0][j];
    const float x = I.px[c][s], y = I.py[c][s], mass = I.m[c][s];
    I.px[o][j] = x;
    I.py[o][j] = y;
    I.vx[o][j] = I.vx[c][s];
    I.vy[o][j] = I.vy[c][s];
    I.m[o][j] = mass;
    I.r[o][j] = I.r[c][s];
    I.id[o][j] = I.id[c][s];
    I.body[j] = make_float4(x, y, mass, 0.0f);
  }
  // Root node.
  if (threadIdx.x == 0) {
    const float s = quad[2];
    I.nodeI[0] = make_int4(0, 0, 0, static_cast<int>(n));
    I.nodeF[0] = make_float4(0.0f, 0.0f, 0.0f, s * s);
    I.centerX[0] = quad[0];
    I.centerY[0] = quad[1];
    I.size[0] = s;
    levelStart[0] = 0;
    levelStart[1] = 1;
    sLevels = 1;
    sOverflow = 0;
  }
  __syncthreads();
  const uint32_t *hi = I.keyHi[0], *lo = I.keyLo[0];
  // Level by level (QuadTree.cpp:176-204): the k-th splitting node of a
  // level puts its children at levelEnd + 4k, in quadrant order.
  for (int d = 0; d < kMaxDepth; ++d) {
    const int ls = levelStart[d], le = levelStart[d + 1];
    if (ls == le || sOverflow) {
      break;
    }
    int splitAfter = 0;
    for (int base = ls; base < le; base += kEnsBlock) {
      const int x = base + static_cast<int>(threadIdx.x);
      bool split = false;
      int4 ni = make_int4(0, 0, 0, 0);
      if (x < le) {
        ni = I.nodeI[x];
        split = ni.w - ni.z > kLeafCapacity;
      }
      uint32_t total;
      const uint32_t rank = ensScan(split ? 1u : 0u, scratch, &total);
      const int first = le + 4 * (splitAfter + static_cast<int>(rank));
      if (split && first + 4 <= I.nodeCap) {
        int bound[5];
        bound[0] = ni.z;
        bound[4] = ni.w;
        for (int q = 1; q < 4; ++q) {
          int a = bound[q - 1], e = ni.w;
          while (a < e) {
            const int mid = (a + e) >> 1;
            if (ensDigit(hi, lo, mid, d) < q) {
              a = mid + 1;
            } else {
              e = mid;
            }
          }
          bound[q] = a;
        }
        const float sz = I.size[x];
        const float cx = I.centerX[x], cy = I.centerY[x];
        const float childSize = sz * 0.5f;
        for (int q = 0; q < 4; ++q) {
          const int ch = first + q;
          I.centerX[ch] = cx + (static_cast<float>(q & 1) - 0.5f) * childSize;
          I.centerY[ch] = cy + (static_cast<float>(q >> 1) - 0.5f) * childSize;
          I.size[ch] = childSize;
          I.nodeF[ch] = make_float4(0.0f, 0.0f, 0.0f, childSize * childSize);
          // `next` (QuadTree.cpp:87-90): the next sibling; the last child
          // inherits its parent's.
          I.nodeI[ch] = make_int4(0, q < 3 ? ch + 1 : ni.y, bound[q], bound[q + 1]);
        }
        I.nodeI[x].x = first;
      } else if (split) {
        sOverflow = 1;
      }
      splitAfter += static_cast<int>(total);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      levelStart[d + 2] = le + 4 * splitAfter;
      if (splitAfter > 0) {
        sLevels = d + 2;
      }
    }
    __syncthreads();
  }
  // (Depth-32 nodes stay leaves: the keys hold 32 levels, and the loop
  // above splits nodes of depth 31 at most.)
  const int levels = sLevels;
  const int count = levelStart[levels];
  if (sOverflow) {
    if (threadIdx.x == 0) {
      I.ctl->status |= kStatusNodeOverflow;
      I.ctl->nodeCount = 0;
    }
    __syncthreads();
    return;
  }
  // Upward pass, deepest level first: leaves sum their bodies
  // (accumulateLeaf), branches their four children (propagate).
  const float *spx = I.px[o], *spy = I.py[o], *sm = I.m[o];
  for (int d = levels - 1; d >= 0; --d) {
    for (int x = levelStart[d] + static_cast<int>(threadIdx.x); x < levelStart[d + 1];
         x += kEnsBlock) {
      const int4 ni = I.nodeI[x];
      if (ni.x == 0) {
        writeSum(x, leafSum(spx, spy, sm, ni.z, ni.w), I.nodeF, I.wsum);
      } else {
        sumBranch(x, ni.x, I.nodeF, I.wsum, nullptr);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    I.ctl->nodeCount = count;
    I.ctl->levels = levels;
  }
  __syncthreads();
}

// Persistent: workgroups take instances from the queue (order[] lists them
// largest first) until none is left.
__global__ __launch_bounds__(kEnsBlock) void k_ens_prepare(
    EnsInst *inst, const uint32_t *__restrict__ order,
    uint32_t count, EnsParams p, uint32_t *__restrict__ queue) {
  __shared__ uint32_t scratch[kEnsWaves];
  __shared__ uint32_t next;
  while (true) {
    if (threadIdx.x == 0) {
      next = atomicAdd(queue, 1u);
    }
    __syncthreads();
    const uint32_t k = next;
    __syncthreads();
    if (k >= count) {
      break;
    }
    const uint32_t idx = order[k];
    const EnsInst &I = inst[idx];
    if (I.n > 0) {
      ensIterate(I, p.dt);
      if (p.collisions) {
        ensCollide(I, scratch);
      }
      ensBuild(I, scratch);
      __syncthreads();
      if (threadIdx.x == 0) {
        inst[idx].cur ^= 1u; // the build gathered the bodies into the other buffer
      }
    }
    __syncthreads();
  }
}

// Resets the queue of k_ens_prepare (a kernel, so a captured step holds no
// memset node).
__global__ void k_ens_reset(uint32_t *__restrict__ queue) { *queue = 0; }

// The walk of every instance in one grid: wave w of the grid takes the
// instance whose waves [waveBase, next waveBase) hold it (waveInst maps a
// wave to its instance).
// BLOCK: lanes per workgroup (64, 128 or 256; nbx_ensemble_set_launch).
constexpr int kEnsWalkBlock = 128;
template <int LAW, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ens_walk(
    const EnsInst *__restrict__ inst, const uint32_t *__restrict__ waveInst,
    uint32_t totalWaves, float tSq, float eSq) {
  __shared__ float2 slots[BLOCK / kWave][kWave];
  __shared__ int spillNode[BLOCK / kWave][kStackDepth - kWave];
  __shared__ uint64_t spillMask[BLOCK / kWave][kStackDepth - kWave];
  const uint32_t w = threadIdx.x >> 6;
  // Wave-uniform (scalar) from here on: the instance's pointers and node
  // records are scalar loads, as in the single-simulation walk.
  const uint32_t gw = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(
      static_cast<int>(xcdBlock() * (BLOCK / kWave) + w)));
  if (gw >= totalWaves) {
    return;
  }
  const uint32_t k = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(waveInst[gw])));
  const EnsInst &I = inst[k];
  // The instance's fields, made scalar (the compiler cannot prove that the
  // walk's stores leave the instance table unchanged).
  auto sp = [](auto *q) {
    return reinterpret_cast<decltype(q)>(uniform64(reinterpret_cast<uint64_t>(q)));
  };
  const uint32_t n = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(I.n)));
  const uint32_t waveBase = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(I.waveBase)));
  const float4 *nodeF = sp(I.nodeF);
  const int4 *nodeI = sp(I.nodeI);
  const float4 *body = sp(I.body);
  float *ax = sp(I.ax), *ay = sp(I.ay);
  Ctl *ctl = sp(I.ctl);
  if (ctl->status & (kStatusNodeOverflow | kStatusTreeInconsistent | kStatusPairOverflow)) {
    return;
  }
  const uint32_t base = (gw - waveBase) * kWave;
  if (base >= n) {
    return;
  }
  if (ctl->levels <= kRegStackLevels) {
    RegStack st;
    walkWave<LAW, false, true, true>(n, base, nodeF, nodeI, nullptr, body, tSq, eSq,
                                     ax, ay, ctl, st, slots[w]);
  } else {
    LaneStack st;
    st.spillNode = spillNode[w];
    st.spillMask = spillMask[w];
    walkWave<LAW, false, true, true>(n, base, nodeF, nodeI, nullptr, body, tSq, eSq,
                                     ax, ay, ctl, st, slots[w]);
  }
}

} // namespace nbx
