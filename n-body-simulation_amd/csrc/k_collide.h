// k_collide.h -- uniform-grid collision broadphase and deterministic pair
// resolution (Collision.cpp:135-349).
//
// Detection finds exactly the reference's pair set: small boxes (no wider
// than a grid cell) pair with small boxes in the 3x3 block of cells around
// them, large boxes pair with every box, and a pair is kept when its boxes
// overlap (Collision.cpp:95-98) -- each unordered pair is emitted once, from
// the side of its `later` box (Collision.cpp:88-90). The reference then
// resolves all pairs sequentially in pairLess order. resolve() touches only
// its two bodies, so pairs in different connected components of the contact
// graph commute: components are found with a lock-free union-find, and each
// component replays its own pairs: in pairLess order on one lane (up to
// kSmallComp pairs), as a topological sort of its per-body pair chains on a
// group of 16-64 lanes (up to 4096 pairs), or in pairLess order on one
// workgroup (beyond, and components a lane group cannot hold). The result
// is bit-identical to the sequential replay.
#pragma once

#include "k_sort.h"
#include "nbx_device.h"

#include <climits>

namespace nbx {

struct Box {
  float minX, maxX, minY, maxY;
};

// Vec2 pos -/+ Vec2::one() * radius (Collision.cpp:156-163).
__device__ inline Box bodyBox(float x, float y, float r) {
  return Box{x - r, x + r, y - r, y + r};
}

// Collision.cpp:88-90 before().
__device__ inline bool beforeBox(float aMinX, int a, float bMinX, int b) {
  return aMinX < bMinX || (aMinX == bMinX && a < b);
}

// Collision.cpp:95-98 overlaps(earlier, later).
__device__ inline bool overlapsBox(const Box &e, const Box &l) {
  return e.maxX >= l.minX && l.minY <= e.maxY && e.minY <= l.maxY;
}

// Collision.cpp:100-104 cellCoord: floor in double, clamped to +-2^29.
__device__ inline int32_t cellCoord(float v, float invCell) {
  const double limit = 536870912.0;
  double c = floor(static_cast<double>(v) * static_cast<double>(invCell));
  c = c < -limit ? -limit : (limit < c ? limit : c);
  return static_cast<int32_t>(c);
}

// Cell -> bucket: cells are grouped in tiles of 8 x 8; a tile's 64 cells
// take 64 consecutive buckets, row by row, and tiles follow each other in
// Z order of their coordinates (low bits), folded into the power-of-two
// bucket count. Grid order is then spatially coherent at tile scale (the
// lanes of a wave of the stencil search cover one tile, and the pairs of a
// clump leave the search together), and the three cells of a stencil row
// are consecutive buckets unless the row crosses a tile edge (then two
// ranges). Cells that share a bucket are told apart by their exact
// coordinates.
constexpr int kTileBits = 3;
__device__ inline uint32_t spreadBits16(uint32_t v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__device__ inline uint32_t tileBucket(int32_t cx, int32_t cy, uint32_t mask) {
  constexpr int32_t kIn = (1 << kTileBits) - 1;
  const uint32_t z = spreadBits16(static_cast<uint32_t>(cx >> kTileBits)) |
                     (spreadBits16(static_cast<uint32_t>(cy >> kTileBits)) << 1);
  return ((z << (2 * kTileBits)) |
          (static_cast<uint32_t>(cy & kIn) << kTileBits) |
          static_cast<uint32_t>(cx & kIn)) &
         mask;
}

// Collision.cpp:333-340 pairLess on (laterMinX, later, earlierMinX, earlier);
// pairs are int4 {later, earlier, bits(laterMinX), bits(earlierMinX)}.
__device__ inline bool pairLess(const int4 &p, const int4 &q) {
  if (p.x != q.x) {
    const float a = __int_as_float(p.z), b = __int_as_float(q.z);
    return a < b || (a == b && p.x < q.x);
  }
  const float a = __int_as_float(p.w), b = __int_as_float(q.w);
  return a < b || (a == b && p.y < q.y);
}

// The two bodies of a pair as resolve() reads and writes them.
struct PairBodies {
  float p1x, p1y, v1x, v1y, p2x, p2y, v2x, v2y;
};

// Collision.cpp:16-84 resolve() on one pair's values, in place; false (and
// nothing changed) when the bodies do not overlap. The reference's three
// contact cases -- moving apart (positions pushed to contact along d),
// nearly equal velocities (placed rr apart along d, or along x when d = 0),
// approaching (moved back to the contact time, velocities reflected) --
// each take one square root and one division first; here they share one
// of each, on operands selected per case, so lanes of a wave in different
// cases do not run three branches one after the other. Every value is the
// reference's operation on the reference's operands.
__device__ inline bool resolveCore(PairBodies &s, float m1, float m2, float r1,
                                   float r2) {
  const float dx = s.p2x - s.p1x, dy = s.p2y - s.p1y;
  const float rr = r1 + r2;
  const float dSq = dx * dx + dy * dy;
  const float rSq = rr * rr;
  if (dSq > rSq) {
    return false;
  }
  const float wx = s.v2x - s.v1x, wy = s.v2y - s.v1y;
  const float dDotV = dx * wx + dy * wy;
  // m2 / (m1 + m2) and m1 / (m1 + m2); for equal positive masses whose sum
  // is finite both quotients are exactly 0.5 (the sum is 2m exactly), so the
  // two divisions are skipped -- the common case, and a uniform branch when
  // every lane's pair has equal masses.
  float weight1 = 0.5f, weight2 = 0.5f;
  if (!(m1 == m2 && m1 > 0.0f && m1 <= 0x1p126f)) {
    weight1 = m2 / (m1 + m2);
    weight2 = m1 / (m1 + m2);
  }
  const bool dNonZero = !(dx == 0.0f && dy == 0.0f);
  const bool apart = dDotV >= 0.0f && dNonZero;
  const float vSq = wx * wx + wy * wy;
  const bool approach = !apart && !(vSq < 1e-12f);
  // apart: rr / |d|; nearly equal velocities: 1 / |d|; approaching:
  // (d.w + sqrt(max(disc, 0))) / |w|^2, the time back to contact.
  const float disc = dDotV * dDotV - vSq * (dSq - rSq);
  const float root = sqrtf(approach ? (disc < 0.0f ? 0.0f : disc) : dSq);
  const float q = (apart ? rr : (approach ? dDotV + root : 1.0f)) / (approach ? vSq : root);
  if (!approach) {
    const float sc = q - 1.0f;
    const float dirx = dNonZero ? dx * q : 1.0f, diry = dNonZero ? dy * q : 0.0f;
    const float tx = apart ? dx * sc : dirx * rr;
    const float ty = apart ? dy * sc : diry * rr;
    s.p1x = s.p1x - weight1 * tx;
    s.p1y = s.p1y - weight1 * ty;
    s.p2x = s.p2x + weight2 * tx;
    s.p2y = s.p2y + weight2 * ty;
    return true;
  }
  const float t = q;
  const float a1x = s.p1x - s.v1x * t, a1y = s.p1y - s.v1y * t;
  const float a2x = s.p2x - s.v2x * t, a2y = s.p2y - s.v2y * t;
  const float ex = a2x - a1x, ey = a2y - a1y;
  const float eDotV = ex * wx + ey * wy;
  const float eSq = ex * ex + ey * ey;
  const float sc = 1.5f * eDotV / eSq;
  const float tx = ex * sc, ty = ey * sc;
  const float n1x = s.v1x + tx * weight1, n1y = s.v1y + ty * weight1;
  const float n2x = s.v2x - tx * weight2, n2y = s.v2y - ty * weight2;
  s.v1x = n1x;
  s.v1y = n1y;
  s.v2x = n2x;
  s.v2y = n2y;
  s.p1x = a1x + n1x * t;
  s.p1y = a1y + n1y * t;
  s.p2x = a2x + n2x * t;
  s.p2y = a2y + n2y * t;
  return true;
}

// Collision.cpp:16-84 resolve(bodies, i, j), on the SoA arrays: every input
// loaded up front (one memory round trip, not two), resolveCore, and the
// pair's eight values written back when the bodies overlapped.
__device__ inline void resolvePair(float *__restrict__ px,
                                   float *__restrict__ py,
                                   float *__restrict__ vx,
                                   float *__restrict__ vy,
                                   const float *__restrict__ m,
                                   const float *__restrict__ r, int i, int j) {
  PairBodies s{px[i], py[i], vx[i], vy[i], px[j], py[j], vx[j], vy[j]};
  if (resolveCore(s, m[i], m[j], r[i], r[j])) {
    px[i] = s.p1x;
    py[i] = s.p1y;
    vx[i] = s.v1x;
    vy[i] = s.v1y;
    px[j] = s.p2x;
    py[j] = s.p2y;
    vx[j] = s.v2x;
    vy[j] = s.v2y;
  }
}

// Grid cell size. The reference takes 1.01 x the median box width
// (Collision.cpp:163-183). The pair set does not depend on the cell (every
// overlapping pair is found for any positive cell: small boxes through the
// 3x3 stencil, boxes wider than a cell directly), only the work does. A
// box's width is 2r up to the rounding of x - r and x + r, and radii never
// change, so the cell is set once per upload on the host: 1.01 x the 5/8
// quantile of 2r (above the median, so that when most widths are equal the
// rounding of x -/+ r cannot turn half the bodies into large boxes).

// Resets the collide's counters (one thread; the step runs it inside the
// iterate kernel, a collide on its own as k_collide_reset). The previous
// collide's pair count was built by atomics in an earlier launch, possibly
// an earlier graph replayed back to back with this one: it is read, and its
// copy for the stencil search written, at agent scope (a plain load here
// has returned a stale zero from a cached line).
__device__ inline void resetCollideCounters(Ctl *__restrict__ ctl) {
  ctl->largeCount = 0;
  storeAgent(&ctl->prevPairs, loadAgent(&ctl->pairCount));
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
  ctl->status &= ~kStatusPairOverflow;
}

__global__ void k_collide_reset(Ctl *__restrict__ ctl) {
  resetCollideCounters(ctl);
}

// Large/small split, grid cell and bucket rank of body i at (x, y); also
// seeds the union-find. Large bodies are counted in *largeCounter.
__device__ inline void cellOf(uint32_t i, float x, float y, float rad,
                              const Ctl *__restrict__ ctl,
                              uint32_t *__restrict__ largeCounter,
                              uint32_t mask, uint32_t *__restrict__ bucketOf,
                              uint32_t *__restrict__ bucketRank,
                              uint32_t *__restrict__ bucketCount,
                              int32_t *__restrict__ largeList,
                              float4 *__restrict__ largeBox,
                              int32_t *__restrict__ ufParent) {
  ufParent[i] = static_cast<int32_t>(i);
  const float cell = ctl->cell, invCell = ctl->invCell;
  const Box b = bodyBox(x, y, rad);
  if (b.maxX - b.minX > cell || b.maxY - b.minY > cell) {
    const uint32_t slot = atomicAdd(largeCounter, 1u);
    largeList[slot] = static_cast<int32_t>(i);
    largeBox[slot] = make_float4(b.minX, b.maxX, b.minY, b.maxY);
    bucketOf[i] = 0xffffffffu;
    return;
  }
  const int32_t cx = cellCoord(b.minX, invCell);
  const int32_t cy = cellCoord(b.minY, invCell);
  const uint32_t h = tileBucket(cx, cy, mask);
  bucketOf[i] = h;
  bucketRank[i] = atomicAdd(&bucketCount[h], 1u);
}

__global__ __launch_bounds__(kBlock) void k_cells(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ r, Ctl *__restrict__ ctl,
    uint32_t mask, uint32_t *__restrict__ bucketOf,
    uint32_t *__restrict__ bucketRank,
    uint32_t *__restrict__ bucketCount, int32_t *__restrict__ largeList,
    float4 *__restrict__ largeBox, int32_t *__restrict__ ufParent) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) {
    return;
  }
  cellOf(i, px[i], py[i], r[i], ctl, &ctl->largeCount, mask, bucketOf,
         bucketRank, bucketCount, largeList, largeBox, ufParent);
}

// Counting-sort scatter of small boxes into their buckets, at the rank
// each body drew in k_cells or k_iterate_cells (the scan already cleared the
// bucket counts for the next step). takeFused: the cells came from
// k_iterate_cells, whose large-body count moves to largeCount here (before
// k_find_small reads it) and is cleared for the next step.
__global__ __launch_bounds__(kBlock) void k_grid_scatter(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ r, Ctl *__restrict__ ctl,
    const uint32_t *__restrict__ bucketOf,
    const uint32_t *__restrict__ bucketRank,
    const uint32_t *__restrict__ bucketStart, float4 *__restrict__ grid,
    uint32_t takeFused) {
  if (takeFused && blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->largeCount = ctl->largeFused;
    ctl->largeFused = 0;
  }
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) {
    return;
  }
  const uint32_t h = bucketOf[i];
  if (h == 0xffffffffu) {
    return;
  }
  const uint32_t slot = bucketStart[h] + bucketRank[i];
  const float invCell = ctl->invCell;
  const Box b = bodyBox(px[i], py[i], r[i]);
  grid[2 * slot] = make_float4(b.minX, b.maxX, b.minY, b.maxY);
  grid[2 * slot + 1] = make_float4(
      __int_as_float(cellCoord(b.minX, invCell)),
      __int_as_float(cellCoord(b.minY, invCell)),
      __int_as_float(static_cast<int32_t>(i)), 0.0f);
}

// Pair output of one workgroup, staged in LDS: lanes append with LDS
// atomics and the workgroup reserves its global range with a single atomic
// (a global counter hit once per pair serialises across the whole chip).
constexpr int kPairStage = 1024;
constexpr int kFindBlock = 1024; // stencil-search workgroup (fewer flushes)
constexpr int kFindStage = 4096;

template <int CAP = kPairStage> struct PairStageT {
  int4 buf[CAP];
  uint32_t count;
  uint32_t base;
};
using PairStage = PairStageT<kPairStage>;

template <int CAP>
__device__ inline void stageInit(PairStageT<CAP> &st) {
  if (threadIdx.x == 0) {
    st.count = 0;
  }
  __syncthreads();
}

// Also hooks the larger-index body of the pair under the smaller one
// (fire-and-forget atomicMin on the union-find parents, all roots at this
// point): every link joins two bodies in contact and points to a smaller
// index, so it is a valid union-find link, and k_union, which processes
// every pair again, restores any link a later atomicMin replaced. The union
// pass then starts from a mostly linked forest instead of singletons.
template <int CAP>
__device__ inline void stagePair(PairStageT<CAP> &st, Ctl *__restrict__ ctl,
                                 int4 *__restrict__ pairs, uint32_t cap,
                                 int32_t *__restrict__ parent, int later,
                                 int earlier, float laterMinX,
                                 float earlierMinX) {
  atomicMin(&parent[max(later, earlier)], min(later, earlier));
  const int4 p = make_int4(later, earlier, __float_as_int(laterMinX),
                           __float_as_int(earlierMinX));
  const uint32_t slot = atomicAdd(&st.count, 1u);
  if (slot < CAP) {
    st.buf[slot] = p;
    return;
  }
  // Staging full (very dense clumps): direct append.
  const uint32_t g = atomicAdd(&ctl->pairCount, 1u);
  if (g < cap) {
    pairs[g] = p;
  } else {
    atomicOr(&ctl->status, kStatusPairOverflow);
  }
}

template <int CAP>
__device__ inline void stageFlush(PairStageT<CAP> &st, Ctl *__restrict__ ctl,
                                  int4 *__restrict__ pairs, uint32_t cap) {
  __syncthreads();
  const uint32_t cnt = min<uint32_t>(st.count, CAP);
  if (threadIdx.x == 0) {
    st.base = cnt ? atomicAdd(&ctl->pairCount, cnt) : 0u;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
    const uint32_t g = st.base + k;
    if (g < cap) {
      pairs[g] = st.buf[k];
    } else {
      atomicOr(&ctl->status, kStatusPairOverflow);
    }
  }
}

// The stencil search: a small box's earlier partners among small boxes of
// the 3x3 neighbouring cells, and every box's pairs with large boxes. Two
// orders of work, chosen by the previous collide's pair count:
//  * few contacts (fewer pairs than half the bodies): one thread per grid
//    entry, in grid order (thread o < n - large takes the small box in grid
//    slot o, the rest the large boxes). A wave then covers one tile of
//    cells, its stencils share records, and each lane's own box, cell and
//    index come from its one 32-byte record;
//  * many contacts: one thread per body, in device order (the previous
//    build's key order), so that the pairs of a clump leave the search
//    together -- the component kernels after it run on that locality.
// The pair set is the same either way.
__global__ __launch_bounds__(kFindBlock) void k_find_small(
    uint32_t n, const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ r, Ctl *__restrict__ ctl, uint32_t mask,
    const uint32_t *__restrict__ bucketOf,
    const uint32_t *__restrict__ bucketStart,
    const float4 *__restrict__ grid,
    const int32_t *__restrict__ largeList,
    const float4 *__restrict__ largeBox, int4 *__restrict__ pairs,
    uint32_t cap, int32_t *__restrict__ parent) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ PairStageT<kFindStage> st;
  stageInit(st);
  const uint32_t o = xcdBlock() * kFindBlock + threadIdx.x;
  if (o < n) {
    const uint32_t large = ctl->largeCount;
    const bool gridOrder = loadAgent(&ctl->prevPairs) < n / 2;
    bool small;
    Box me;
    int self;
    int32_t mx = 0, my = 0;
    if (gridOrder) {
      const uint32_t nSmall = n - min(large, n);
      small = o < nSmall;
      if (small) {
        const float4 rb = grid[2 * o], rc = grid[2 * o + 1];
        me = Box{rb.x, rb.y, rb.z, rb.w};
        mx = __float_as_int(rc.x);
        my = __float_as_int(rc.y);
        self = __float_as_int(rc.z);
      } else {
        const float4 lb = largeBox[o - nSmall];
        me = Box{lb.x, lb.y, lb.z, lb.w};
        self = largeList[o - nSmall];
      }
    } else {
      small = bucketOf[o] != 0xffffffffu;
      me = bodyBox(px[o], py[o], r[o]);
      self = static_cast<int>(o);
      if (small) {
        const float invCell = ctl->invCell;
        mx = cellCoord(me.minX, invCell);
        my = cellCoord(me.minY, invCell);
      }
    }
    if (small) {
      // Per stencil row, the bucket ranges of its three cells: one
      // contiguous range, or two where the row crosses a tile edge (the
      // cell left of a tile's first column, or right of its last, lies in
      // the neighbouring tile) -- six ranges at most, all requested up
      // front. Box and cell of an entry share one 32-byte record: both are
      // requested before either is tested.
      auto test = [&](const float4 &ob, const float4 &ocf, int32_t cy) {
        const int32_t ox = __float_as_int(ocf.x), oy = __float_as_int(ocf.y);
        const int other = __float_as_int(ocf.z);
        if (oy != cy || ox < mx - 1 || ox > mx + 1 || other == self) {
          return;
        }
        if (beforeBox(ob.x, other, me.minX, self) &&
            overlapsBox(Box{ob.x, ob.y, ob.z, ob.w}, me)) {
          stagePair(st, ctl, pairs, cap, parent, self, other, me.minX, ob.x);
        }
      };
      constexpr int32_t kIn = (1 << kTileBits) - 1;
      const int32_t ix = mx & kIn;
      // Cells of the first range: all three, or the one (two) left of the
      // tile edge.
      const int32_t lenA = ix == 0 ? 1 : (ix == kIn ? 2 : 3);
      uint32_t lo[6], hi[6];
#pragma unroll
      for (int row = 0; row < 3; ++row) {
        const int32_t cy = my + row - 1;
        const uint32_t a = tileBucket(mx - 1, cy, mask);
        const uint32_t b = tileBucket(mx - 1 + lenA, cy, mask);
        lo[2 * row] = bucketStart[a];
        hi[2 * row] = bucketStart[a + lenA];
        lo[2 * row + 1] = bucketStart[b];
        hi[2 * row + 1] = bucketStart[b + (3 - lenA)];
      }
      // The six ranges' entries as one list, four entries per iteration
      // (so a wave runs max-over-lanes of the total, not of each range, and
      // four records are in flight per lane; a lane past its end reloads its
      // first entry of the iteration and skips the test).
      uint32_t end[6];
      uint32_t acc = 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        acc += hi[k] - lo[k];
        end[k] = acc;
      }
      const uint32_t total = acc;
      auto at = [&](uint32_t q, int32_t &cy) {
        uint32_t start = 0, o = lo[0] + q;
        int k = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          if (q >= end[j]) {
            start = end[j];
            k = j + 1;
            o = lo[j + 1] + (q - start);
          }
        }
        cy = my + (k >> 1) - 1;
        return o;
      };
      for (uint32_t q = 0; q < total; q += 4) {
        int32_t cy0, cy1, cy2, cy3;
        const uint32_t o0 = at(q, cy0);
        const uint32_t o1 = q + 1 < total ? at(q + 1, cy1) : o0;
        const uint32_t o2 = q + 2 < total ? at(q + 2, cy2) : o0;
        const uint32_t o3 = q + 3 < total ? at(q + 3, cy3) : o0;
        const float4 ob0 = grid[2 * o0], oc0 = grid[2 * o0 + 1];
        const float4 ob1 = grid[2 * o1], oc1 = grid[2 * o1 + 1];
        const float4 ob2 = grid[2 * o2], oc2 = grid[2 * o2 + 1];
        const float4 ob3 = grid[2 * o3], oc3 = grid[2 * o3 + 1];
        asm volatile("" ::"v"(ob0.x), "v"(ob0.y), "v"(ob0.z), "v"(ob0.w),
                     "v"(oc0.x), "v"(oc0.y), "v"(oc0.z), "v"(ob1.x),
                     "v"(ob1.y), "v"(ob1.z), "v"(ob1.w), "v"(oc1.x),
                     "v"(oc1.y), "v"(oc1.z), "v"(ob2.x), "v"(ob2.y),
                     "v"(ob2.z), "v"(ob2.w), "v"(oc2.x), "v"(oc2.y),
                     "v"(oc2.z), "v"(ob3.x), "v"(ob3.y), "v"(ob3.z),
                     "v"(ob3.w), "v"(oc3.x), "v"(oc3.y), "v"(oc3.z));
        test(ob0, oc0, cy0);
        if (q + 1 < total) {
          test(ob1, oc1, cy1);
        }
        if (q + 2 < total) {
          test(ob2, oc2, cy2);
        }
        if (q + 3 < total) {
          test(ob3, oc3, cy3);
        }
      }
    }
    // Large boxes (Collision.cpp:319-330): against every body, each pair
    // once -- a small body emits both orientations, a large one only the
    // pairs with later large boxes (the earlier one emits large-large pairs).
    // Each large box's index and box were recorded together by k_cells: one
    // round trip per large box.
    for (uint32_t k = 0; k < large; ++k) {
      const int l = largeList[k];
      const float4 lbv = largeBox[k];
      if (l == self) {
        continue;
      }
      const Box lb{lbv.x, lbv.y, lbv.z, lbv.w};
      if (beforeBox(lb.minX, l, me.minX, self)) {
        if (small && overlapsBox(lb, me)) {
          stagePair(st, ctl, pairs, cap, parent, self, l, me.minX, lb.minX);
        }
      } else if (overlapsBox(me, lb)) {
        stagePair(st, ctl, pairs, cap, parent, l, self, lb.minX, me.minX);
      }
    }
  }
  stageFlush(st, ctl, pairs, cap);
}

// ---- connected components of the contact graph ----

__device__ inline int ufLoad(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x. Parent links only ever move to a smaller index, so the walk
// terminates and the final root is the component's smallest body index.
// The first kUfReadSteps steps only read; a walk that goes on halves its
// path (each link it passes is pointed at its grandparent). Most finds are
// a few links long, and halving those too makes many lanes store to the
// few cache lines a clump's parents occupy, which cost k_union more than
// the reads (145 vs 84 us at 1.9M pairs). A long chain -- bodies in a line,
// numbered along it by the tree order -- is still cut down by the finds
// that cross it.
constexpr int kUfReadSteps = 2;
__device__ inline int ufFind(int32_t *parent, int x) {
  for (int step = 0;; ++step) {
    const int p = ufLoad(&parent[x]);
    if (p == x) {
      return x;
    }
    const int gp = ufLoad(&parent[p]);
    if (gp == p) {
      return p;
    }
    if (step >= kUfReadSteps) {
      __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    x = gp;
  }
}

// Points hooked body v at the root of its hook tree. The search's hooks
// alone leave chains (each body under its smallest partner, that one under
// its own...); flattened, most pairs then see one common parent at their
// first two loads and leave k_union at once. Links only move to ancestors,
// so the forest stays valid whatever the order of the stores, also while
// other lanes join trees.
__device__ inline void flattenBody(int32_t *parent, uint32_t v) {
  const int p = ufLoad(&parent[v]);
  if (p == static_cast<int>(v)) {
    return;
  }
  const int root = ufFind(parent, p);
  if (root != p) {
    __hip_atomic_store(&parent[v], root, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Each thread flattens its share of the bodies, then joins its share of
// the pairs (one launch: the flattening of one part overlaps the joins of
// another, and a step has one kernel less).
__global__ __launch_bounds__(kBlock) void k_union(
    const Ctl *__restrict__ ctl, uint32_t n, uint32_t cap,
    const int4 *__restrict__ pairs, int32_t *__restrict__ parent) {
  if (barrierStalled(ctl)) {
    return;
  }
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n;
       v += gridDim.x * kBlock) {
    flattenBody(parent, v);
  }
  const uint32_t count = min(ctl->pairCount, cap);
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count;
       k += gridDim.x * kBlock) {
    const int4 p = pairs[k];
    // Already in one tree: a common parent, or one body the other's parent.
    const int pa = ufLoad(&parent[p.x]);
    const int pb = ufLoad(&parent[p.y]);
    if (pa == pb || pa == p.y || pb == p.x) {
      continue;
    }
    int a = ufFind(parent, pa);
    int b = ufFind(parent, pb);
    while (a != b) {
      const int hi = a > b ? a : b;
      const int lo = a > b ? b : a;
      const int old = atomicCAS(&parent[hi], hi, lo);
      if (old == hi) {
        break;
      }
      a = ufFind(parent, old);
      b = ufFind(parent, lo);
    }
  }
}

// Runs of equal roots among a wave's consecutive lanes (valid lanes are a
// prefix of the wave): the run's first lane, its length and, for every lane,
// its run's first lane -- one atomic a run instead of one a pair (the pairs
// of a clump come out of the search next to each other).
struct RootRun {
  bool head;
  uint32_t length;
  int headLane;
};
__device__ inline RootRun rootRun(int root, bool valid) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const int prev = __shfl_up(root, 1, kWave);
  const bool head = valid && (lane == 0 || prev != root);
  const uint64_t heads = __ballot(head);
  const uint32_t nvalid = static_cast<uint32_t>(__popcll(__ballot(valid)));
  RootRun run;
  run.head = head;
  const uint64_t above = lane == kWave - 1 ? 0ull : heads >> (lane + 1);
  run.length = (above ? lane + 1 + static_cast<uint32_t>(__builtin_ctzll(above)) : nvalid) - lane;
  const uint64_t upto = heads & (lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull));
  run.headLane = upto ? 63 - __builtin_clzll(upto) : 0;
  return run;
}

__global__ __launch_bounds__(kBlock) void k_comp_count(
    const Ctl *__restrict__ ctl, uint32_t cap, const int4 *__restrict__ pairs,
    int32_t *__restrict__ parent, int32_t *__restrict__ pairComp,
    uint32_t *__restrict__ compCount) {
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t count = min(ctl->pairCount, cap);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t k0 = blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1u));
       k0 < count; k0 += gridDim.x * kBlock) {
    const uint32_t k = k0 + lane;
    int root = -1;
    if (k < count) {
      root = ufFind(parent, pairs[k].x);
      pairComp[k] = root;
    }
    const RootRun run = rootRun(root, k < count);
    if (run.head) {
      atomicAdd(&compCount[root], run.length);
    }
  }
}

constexpr int kAllocItems = 16; // bodies per thread in k_comp_alloc

// Allocates each component's contiguous pair segment. A workgroup covers
// 4096 bodies (blocked per thread, so offsets follow body order) and
// reserves its ranges with one atomic per counter.
__global__ __launch_bounds__(kBlock) void k_comp_alloc(
    uint32_t n, Ctl *__restrict__ ctl, const uint32_t *__restrict__ compCount,
    uint32_t *__restrict__ compStart, int4 *__restrict__ compInfo) {
  __shared__ uint32_t scratch[kBlock / kWave];
  __shared__ uint32_t wmax[kBlock / kWave];
  __shared__ uint32_t base[2];
  const uint32_t first =
      (blockIdx.x * kBlock + threadIdx.x) * static_cast<uint32_t>(kAllocItems);
  uint32_t c[kAllocItems];
  uint32_t pairs = 0, roots = 0, mx = 0;
#pragma unroll
  for (int k = 0; k < kAllocItems; ++k) {
    c[k] = first + k < n ? compCount[first + k] : 0u;
    pairs += c[k];
    roots += c[k] > 0 ? 1u : 0u;
    mx = max(mx, c[k]);
  }
  uint32_t totalPairs, totalRoots;
  uint32_t prePairs = blockExclusiveScan(pairs, scratch, &totalPairs);
  uint32_t preRoots = blockExclusiveScan(roots, scratch, &totalRoots);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    mx = max(mx, static_cast<uint32_t>(__shfl_xor(mx, off, kWave)));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    wmax[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    base[0] = totalPairs ? atomicAdd(&ctl->compPairs, totalPairs) : 0u;
    base[1] = totalRoots ? atomicAdd(&ctl->compCount, totalRoots) : 0u;
    uint32_t bm = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      bm = max(bm, wmax[w]);
    }
    if (bm) {
      atomicMax(&ctl->largest, bm);
    }
  }
  __syncthreads();
  uint32_t p = base[0] + prePairs, q = base[1] + preRoots;
#pragma unroll
  for (int k = 0; k < kAllocItems; ++k) {
    if (c[k] > 0) {
      compStart[first + k] = p;
      compInfo[q] = make_int4(static_cast<int>(first + k),
                              static_cast<int>(p), static_cast<int>(c[k]), 0);
      p += c[k];
      q += 1;
    }
  }
}

// The same allocation in one launch (G resident 1024-lane workgroups, each
// lane ITEMS consecutive bodies, 16-byte loads): every workgroup publishes
// its pair, component and largest-component totals to tiles[3b..3b+2];
// after one grid barrier it adds the totals of the workgroups before it
// and writes its segments. Segments follow body order, with no atomics;
// the last workgroup publishes the totals to ctl.
template <int ITEMS>
__global__ __launch_bounds__(kPsBlock) void k_comp_alloc_onelaunch(
    uint32_t n, uint32_t chunk, Ctl *ctl, const uint32_t *__restrict__ compCount,
    uint32_t *__restrict__ compStart, int4 *__restrict__ compInfo,
    uint32_t *tiles, uint32_t *bar, uint32_t spinLimit) {
  static_assert(ITEMS % 4 == 0, "16-byte loads");
  __shared__ uint32_t scratch[kPsWaves];
  __shared__ uint32_t wmax[kPsWaves];
  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x, grid = gridDim.x;
  const uint32_t lo = min(n, b * chunk);
  const uint32_t hi = min(n, lo + chunk);
  const uint32_t first = lo + t * ITEMS;
  uint32_t c[ITEMS];
  if (first + ITEMS <= hi) {
    const uint4 *p = reinterpret_cast<const uint4 *>(compCount + first);
#pragma unroll
    for (int k = 0; k < ITEMS / 4; ++k) {
      const uint4 q = p[k];
      c[4 * k] = q.x;
      c[4 * k + 1] = q.y;
      c[4 * k + 2] = q.z;
      c[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      c[k] = first + k < hi ? compCount[first + k] : 0u;
    }
  }
  uint32_t pairs = 0, roots = 0, mx = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    pairs += c[k];
    roots += c[k] > 0 ? 1u : 0u;
    mx = max(mx, c[k]);
  }
  uint32_t totalPairs, totalRoots;
  uint32_t p = blockExclusiveScan<kPsBlock>(pairs, scratch, &totalPairs);
  uint32_t q = blockExclusiveScan<kPsBlock>(roots, scratch, &totalRoots);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    mx = max(mx, static_cast<uint32_t>(__shfl_xor(mx, off, kWave)));
  }
  if ((t & (kWave - 1)) == 0) {
    wmax[t >> 6] = mx;
  }
  __syncthreads();
  if (t == 0) {
    uint32_t bm = 0;
    for (int w = 0; w < kPsWaves; ++w) {
      bm = max(bm, wmax[w]);
    }
    storeAgent(tiles + 3 * b, totalPairs);
    storeAgent(tiles + 3 * b + 1, totalRoots);
    storeAgent(tiles + 3 * b + 2, bm);
  }
  gridBarrier(bar, grid, &ctl->status, spinLimit);
  uint32_t beforePairs, beforeRoots;
  blockExclusiveScan<kPsBlock>(t < b ? loadAgent(tiles + 3 * t) : 0u, scratch,
                               &beforePairs);
  blockExclusiveScan<kPsBlock>(t < b ? loadAgent(tiles + 3 * t + 1) : 0u,
                               scratch, &beforeRoots);
  p += beforePairs;
  q += beforeRoots;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    if (c[k] > 0) {
      compStart[first + k] = p;
      compInfo[q] = make_int4(static_cast<int>(first + k), static_cast<int>(p),
                              static_cast<int>(c[k]), 0);
      p += c[k];
      q += 1;
    }
  }
  if (b == grid - 1) {
    uint32_t m = t < grid ? loadAgent(tiles + 3 * t + 2) : 0u;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      m = max(m, static_cast<uint32_t>(__shfl_xor(m, off, kWave)));
    }
    __syncthreads();
    if ((t & (kWave - 1)) == 0) {
      wmax[t >> 6] = m;
    }
    __syncthreads();
    if (t == 0) {
      for (int w = 0; w < kPsWaves; ++w) {
        m = max(m, wmax[w]);
      }
      ctl->compPairs = beforePairs + totalPairs;
      ctl->compCount = beforeRoots + totalRoots;
      ctl->largest = m;
    }
  }
}

// Moves every pair into its component's segment; consumes the counts back
// to zero.
__global__ __launch_bounds__(kBlock) void k_comp_scatter(
    const Ctl *__restrict__ ctl, uint32_t cap, const int4 *__restrict__ pairs,
    const int32_t *__restrict__ pairComp, uint32_t *__restrict__ compCount,
    const uint32_t *__restrict__ compStart, int4 *__restrict__ segPairs) {
  if (barrierStalled(ctl)) {
    return;
  }
  const uint32_t count = min(ctl->pairCount, cap);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  for (uint32_t k0 = blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1u));
       k0 < count; k0 += gridDim.x * kBlock) {
    const uint32_t k = k0 + lane;
    const bool valid = k < count;
    const int root = valid ? pairComp[k] : -1;
    const RootRun run = rootRun(root, valid);
    uint32_t old = 0;
    if (run.head) {
      old = atomicSub(&compCount[root], run.length);
    }
    // The run takes slots [old - length, old) of its component's segment.
    const uint32_t top = __shfl(old, run.headLane, kWave);
    const uint32_t len = __shfl(run.length, run.headLane, kWave);
    if (valid) {
      segPairs[compStart[root] + top - len + (lane - run.headLane)] = pairs[k];
    }
  }
}

// A component of up to N pairs in registers: loaded together, padded with
// entries that order last, sorted by a fixed odd-even merge network
// (Batcher) -- static indices only, so nothing spills to scratch -- then
// written back in order and replayed.
template <int N>
__device__ __forceinline__ void resolveComponent(int start, int size,
                                        int4 *__restrict__ segPairs,
                                        float *__restrict__ px,
                                        float *__restrict__ py,
                                        float *__restrict__ vx,
                                        float *__restrict__ vy,
                                        const float *__restrict__ m,
                                        const float *__restrict__ r) {
  // Loads from a clamped index and selects values (a select of pointers
  // would load through a private copy of `pad`); inlined, the body
  // accesses below are global ones, not flat.
  const int4 pad = make_int4(INT_MAX, INT_MAX, 0x7f800000, 0x7f800000);
  int4 p[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int4 q = segPairs[start + min(k, size - 1)];
    p[k] = k < size ? q : pad;
  }
#pragma unroll
  for (int w = 1; w < N; w <<= 1) {
#pragma unroll
    for (int k = w; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % w; j + k < N; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i + j + k < N && (i + j) / (2 * w) == (i + j + k) / (2 * w)) {
            const int a = i + j, b = i + j + k;
            if (pairLess(p[b], p[a])) {
              const int4 t = p[a];
              p[a] = p[b];
              p[b] = t;
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k < size) {
      segPairs[start + k] = p[k];
      resolvePair(px, py, vx, vy, m, r, p[k].x, p[k].y);
    }
  }
}

// One lane per component of up to kSmallComp pairs: pairLess order, then the
// sequential replay. Larger components are queued by size for the wave
// tiers (tiers 0..2 in k_resolve_tiers, tier 3 in k_resolve_large), beyond those for the
// workgroup path (k_resolve_large).
__global__ __launch_bounds__(kBlock) void k_resolve_small(
    Ctl *__restrict__ ctl, const int4 *__restrict__ compInfo,
    int4 *__restrict__ segPairs, int4 *__restrict__ tierList,
    uint32_t listCap, int32_t *__restrict__ bigList, float *__restrict__ px,
    float *__restrict__ py, float *__restrict__ vx, float *__restrict__ vy,
    const float *__restrict__ m, const float *__restrict__ r) {
  if (barrierStalled(ctl) || (ctl->status & kStatusPairOverflow)) {
    return;
  }
  // The grid covers the bodies (the most components there can be); a
  // workgroup past the components has nothing to queue or replay.
  const uint32_t comps = ctl->compCount;
  if (blockIdx.x * kBlock >= comps) {
    return;
  }
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const bool live = j < comps;
  int4 info = make_int4(0, 0, 0, 0);
  if (live) {
    info = compInfo[j];
  }
  const int start = info.y, size = info.z;
  // Queue class: kTierQueues t + k for a component of wave tier t in the
  // tier's k-th size class counted from the largest (the tier's pair range
  // (lo, P] cut in kTierQueues equal parts), kTierQueues kWaveTiers the
  // workgroup path, -1 none.
  int cls = -1;
  if (size > kSmallComp) {
    int t = 0;
    while (t < kWaveTiers && size > waveTierPairs(t)) {
      ++t;
    }
    if (t < kWaveTiers) {
      const int lo = t == 0 ? kSmallComp : waveTierPairs(t - 1);
      const int hi = waveTierPairs(t);
      const int k = (size - lo - 1) * kTierQueues / (hi - lo);
      cls = kTierQueues * t + (kTierQueues - 1 - k);
    } else {
      cls = kTierQueues * kWaveTiers;
    }
  }
  // Queue slots: the workgroup's count per class in LDS, then one atomic
  // per class and workgroup, the classes' atomics issued together (a wave
  // waiting on each class's contended counter in turn cost more than the
  // replays).
  constexpr int kClasses = kTierQueues * kWaveTiers + 1;
  constexpr int kWaves = kBlock / kWave;
  __shared__ uint32_t waveCnt[kWaves][kClasses];
  __shared__ uint32_t waveBase[kWaves][kClasses];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const int wave = static_cast<int>(threadIdx.x / kWave);
  uint64_t mine = 0;
#pragma unroll
  for (int c = 0; c < kClasses; ++c) {
    const uint64_t mask = __ballot(cls == c);
    if (cls == c) {
      mine = mask;
    }
    if (lane == static_cast<uint32_t>(c)) {
      waveCnt[wave][c] = static_cast<uint32_t>(__popcll(mask));
    }
  }
  __syncthreads();
  if (threadIdx.x < static_cast<uint32_t>(kClasses)) {
    const int c = static_cast<int>(threadIdx.x);
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      total += waveCnt[w][c];
    }
    uint32_t base = 0;
    if (total > 0) {
      base = atomicAdd(c < kTierQueues * kWaveTiers ? &ctl->waveCount[c] : &ctl->bigCount, total);
    }
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      waveBase[w][c] = base;
      base += waveCnt[w][c];
    }
  }
  __syncthreads();
  if (cls >= 0) {
    const uint32_t at = waveBase[wave][cls] +
                        static_cast<uint32_t>(__popcll(mine & laneMaskLt()));
    if (cls < kTierQueues * kWaveTiers) {
      tierList[static_cast<size_t>(cls) * listCap + at] =
          make_int4(start, size, static_cast<int>(j), 0);
    } else {
      bigList[at] = static_cast<int32_t>(j);
    }
  }
  if (!live || size > kSmallComp) {
    return;
  }
  static_assert(kSmallComp == 4 || kSmallComp == 8 || kSmallComp == 16,
                "size classes below cover the one-lane path");
  if (size == 1) {
    const int4 q = segPairs[start];
    resolvePair(px, py, vx, vy, m, r, q.x, q.y);
  } else if (kSmallComp == 4 || size <= 4) {
    resolveComponent<4>(start, size, segPairs, px, py, vx, vy, m, r);
  } else if (kSmallComp == 8 || size <= 8) {
    resolveComponent<8>(start, size, segPairs, px, py, vx, vy, m, r);
  } else {
    resolveComponent<16>(start, size, segPairs, px, py, vx, vy, m, r);
  }
}

// ---- contact components of kSmallComp+1 .. 4096 pairs: 16-64 lanes each ----
//
// resolve() touches only its two bodies, so the sequential replay of a
// component is fixed by the order of the pairs on each body: any order that
// keeps, for every body, its pairs in pairLess order gives every pair
// exactly the inputs the sequential replay gives it -- a bit-identical
// result. Pairs are oriented by before() (Collision.cpp:88-90): the later
// body has the larger (minX, index), so a body's pairs in pairLess order are
// its pairs ordered by the other body's (minX, index). The replay is then a
// topological sort of the per-body chains: rounds of pairs whose earlier
// pairs on both bodies are done (Kahn's algorithm), the pairs of a round in
// parallel -- as many rounds as the longest chain (~0.3 x the pairs in the
// packed clumps a long collisional run builds) instead of one step a pair.
//
// A group of 16, 32 or 64 lanes per component (by tier), all state in the
// group's own LDS region so that many components are in flight on each CU
// (the rounds are latency-bound): an open-addressing table gives the bodies
// dense slots, counting sort builds each body's list of pairs, one lane a
// body sorts its list (at most kMaxDegree pairs, in registers) and links
// each pair to the next pair on that body, the bodies are staged in LDS and
// the rounds replay the pairs. The pairs are written back in the order they
// were replayed. A component with more than BMAX bodies goes to tier 3's
// queue (from tier 3: the workgroup path), one with a body of more than
// kMaxDegree pairs to the workgroup path (resolveBigComponent, in
// k_resolve_large).
constexpr int kMaxDegree = 16;
constexpr uint32_t kEmptyKey = 0xffffffffu;

template <int PMAX, int BMAX> struct WaveLds {
  static_assert(PMAX < 65535 && BMAX <= 32768 && BMAX % 2 == 0, "16-bit fields");
  static_assert(2 * BMAX >= PMAX + 2, "the body table always has a free entry");
  static_assert(PMAX >= BMAX && (PMAX & (PMAX - 1)) == 0,
                "the ready ring (P / 2 records) holds B / 2, a power of two");
  static constexpr int kHash = 2 * BMAX;            // table entries
  static constexpr int kInfo = 0;                   // u64[P] pair records: slot A |
                                                    // slot B << 16 | next pair on A
                                                    // << 32 | next pair on B << 48
  static constexpr int kIndeg = 2 * PMAX;           // u32[P]: earlier pairs not done
  static constexpr int kInc = 3 * PMAX;             // u16[2P]: lists, then the ready
                                                    // ring (u64[P / 2] records)
  static constexpr uint32_t kRing = PMAX / 2;
  static constexpr int kBodyIdx = 4 * PMAX;         // i32[B]
  static constexpr int kMinX = 4 * PMAX + BMAX;     // u32[B]: bits of the box minX
  static constexpr int kU = 4 * PMAX + 2 * BMAX;    // 6B words, in turn: table
  static constexpr int kWords = kU + 6 * BMAX;      // (3B), lists (2B+1), state (6B)
};

// before() order (Collision.cpp:88-90) as one integer: minX in
// order-preserving bits (-0 made +0, as float < treats them alike), then the
// index.
__device__ inline uint64_t beforeKey(uint32_t minXBits, int32_t index) {
  const float f = __uint_as_float(minXBits) + 0.0f;
  return (static_cast<uint64_t>(ordBits(f)) << 32) | static_cast<uint32_t>(index);
}

// Inserts key e into the table; h is where it lives. True when this call
// added it. Terminates: the table never fills (more entries than bodies).
__device__ inline bool hashInsert(uint32_t *__restrict__ key, uint32_t size,
                                  uint32_t e, uint32_t &h) {
  h = static_cast<uint32_t>((static_cast<uint64_t>(e * 0x9E3779B1u) * size) >> 32);
  while (true) {
    const uint32_t old = atomicCAS(&key[h], kEmptyKey, e);
    if (old == kEmptyKey) {
      return true;
    }
    if (old == e) {
      return false;
    }
    h = (h + 1 == size) ? 0u : h + 1;
  }
}

// Sorts (key, pair) by key with Batcher's odd-even merge network (static
// indices only, so everything stays in registers).
template <int N>
__device__ inline void sortIncident(uint64_t (&key)[N], uint32_t (&pr)[N]) {
#pragma unroll
  for (int w = 1; w < N; w <<= 1) {
#pragma unroll
    for (int k = w; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % w; j + k < N; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i + j + k < N && (i + j) / (2 * w) == (i + j + k) / (2 * w)) {
            const int a = i + j, b = i + j + k;
            if (key[b] < key[a]) {
              const uint64_t tk = key[a];
              key[a] = key[b];
              key[b] = tk;
              const uint32_t tp = pr[a];
              pr[a] = pr[b];
              pr[b] = tp;
            }
          }
        }
      }
    }
  }
}

// Sorts body s's list (d <= N entries at inc[b0..]) into pairLess order and
// links each pair to the next one on s -- the last one to itself -- and
// every pair after the first gets one more earlier pair to wait for.
template <int N>
__device__ inline void linkBody(int s, uint32_t b0, uint32_t d,
                                const uint16_t *__restrict__ inc,
                                uint32_t *__restrict__ info,
                                const uint32_t *__restrict__ minX,
                                const int32_t *__restrict__ bodyIdx,
                                uint32_t *__restrict__ indeg) {
  // info[2k]: the pair's slots; 16-bit halves 4k + 2 and 4k + 3: its next
  // pair on its first and on its second body.
  uint16_t *succ = reinterpret_cast<uint16_t *>(info);
  uint64_t key[N];
  uint32_t pr[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    key[i] = ~0ull;
    pr[i] = 0;
    if (static_cast<uint32_t>(i) < d) {
      const uint32_t k = inc[b0 + i];
      const uint32_t sl = info[2 * k];
      const uint32_t a = sl & 0xffffu, b = sl >> 16;
      const uint32_t o = a == static_cast<uint32_t>(s) ? b : a;
      key[i] = beforeKey(minX[o], bodyIdx[o]);
      pr[i] = k | (a == static_cast<uint32_t>(s) ? 0u : 0x10000u); // side of s
    }
  }
  sortIncident<N>(key, pr);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (static_cast<uint32_t>(i) < d) {
      const uint32_t k = pr[i] & 0xffffu, side = pr[i] >> 16;
      const uint16_t next = static_cast<uint16_t>(
          static_cast<uint32_t>(i + 1) < d ? pr[i + 1 < N ? i + 1 : i] & 0xffffu : k);
      succ[4 * k + 2 + side] = next;
      if (i > 0) {
        atomicAdd(&indeg[k], 1u);
      }
    }
  }
}

// Orders the wave's LDS accesses around this point: a workgroup of one wave
// needs no barrier (a wave's LDS operations complete in order), and no wait
// for its global stores, which nothing in the kernel reads back.
__device__ inline void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Queue entry of a component: {first pair, pair count, component index, 0}.
template <int PMAX, int BMAX, int G> struct WaveCtx {
  static_assert(G == 16 || G == 32 || G == 64, "lanes per component");
  // Pairs and bodies per lane held in registers (prefetched pairs of the
  // next component, staged bodies) where the tier's LDS already limits the
  // occupancy more than these registers do.
  static constexpr int kPf = PMAX <= 8 * G && PMAX % G == 0 ? PMAX / G : 0;
  static constexpr int kSf = BMAX <= 6 * G && BMAX % G == 0 ? BMAX / G : 0;
  struct Pairs {
    int4 v[kPf > 0 ? kPf : 1];
  };
};

template <int PMAX, int BMAX, int G>
__device__ __forceinline__ void resolveWaveComponent(
    uint32_t *__restrict__ lds, Ctl *__restrict__ ctl, int tier, int4 info,
    typename WaveCtx<PMAX, BMAX, G>::Pairs pre,
    int4 *__restrict__ tierList, uint32_t listCap,
    int32_t *__restrict__ bigList, int4 *__restrict__ segPairs,
    float *__restrict__ px, float *__restrict__ py, float *__restrict__ vx,
    float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r, int *localBig = nullptr) {
  using L = WaveLds<PMAX, BMAX>;
  using W = WaveCtx<PMAX, BMAX, G>;
  // A component the wave path cannot hold goes to the next tier, or to the
  // workgroup path: appended to bigList, or (localBig) handed back to the
  // caller, which runs it at once.
  auto toBig = [&]() {
    if (localBig) {
      *localBig = info.z;
    } else {
      bigList[atomicAdd(&ctl->bigCount, 1u)] = info.z;
    }
  };
  uint32_t *slots = lds + L::kInfo; // the slots of pair k at slots[2k]
  const uint64_t *rec = reinterpret_cast<const uint64_t *>(lds + L::kInfo);
  uint32_t *indeg = lds + L::kIndeg;
  uint16_t *inc = reinterpret_cast<uint16_t *>(lds + L::kInc);
  int32_t *bodyIdx = reinterpret_cast<int32_t *>(lds + L::kBodyIdx);
  uint32_t *minX = lds + L::kMinX;
  uint32_t *hkey = lds + L::kU;
  uint16_t *hslot = reinterpret_cast<uint16_t *>(lds + L::kU + 2 * BMAX);
  uint32_t *start = lds + L::kU;
  uint32_t *cursor = lds + L::kU + BMAX + 1;
  float *lx = reinterpret_cast<float *>(lds + L::kU);
  float *ly = lx + BMAX, *lvx = ly + BMAX, *lvy = lvx + BMAX;
  float *lm = lvy + BMAX, *lr = lm + BMAX;
  // This component's G lanes of the wave.
  const int lane = static_cast<int>(threadIdx.x) % G;
  const uint64_t gmask =
      G == kWave ? ~0ull : (((1ull << (G % kWave)) - 1ull) << (threadIdx.x - lane));
  const uint64_t lt = laneMaskLt() & gmask;
  auto gballot = [&](bool p) { return __ballot(p) & gmask; };
  const int size = info.y; // kSmallComp < size <= PMAX
  int4 *seg = segPairs + info.x;

  // 1. A dense slot per body; the records' slot words hold table
  //    positions meanwhile.
  for (int t = lane; t < L::kHash; t += G) {
    hkey[t] = kEmptyKey;
  }
  waveSync();
  uint32_t nb = 0;
  auto insertPair = [&](int k, int4 q) __attribute__((always_inline)) {
    uint32_t h0 = 0, h1 = 0;
    bool w0 = false, w1 = false;
    if (k < size) {
      w0 = hashInsert(hkey, L::kHash, static_cast<uint32_t>(q.x), h0);
      w1 = hashInsert(hkey, L::kHash, static_cast<uint32_t>(q.y), h1);
      slots[2 * k] = h0 | (h1 << 16);
    }
    const uint64_t b0 = gballot(w0), b1 = gballot(w1);
    const uint32_t n0 = static_cast<uint32_t>(__popcll(b0));
    if (w0) {
      const uint32_t slot = nb + static_cast<uint32_t>(__popcll(b0 & lt));
      hslot[h0] = static_cast<uint16_t>(slot);
      if (slot < BMAX) {
        bodyIdx[slot] = q.x;
        minX[slot] = static_cast<uint32_t>(q.z);
      }
    }
    if (w1) {
      const uint32_t slot = nb + n0 + static_cast<uint32_t>(__popcll(b1 & lt));
      hslot[h1] = static_cast<uint16_t>(slot);
      if (slot < BMAX) {
        bodyIdx[slot] = q.y;
        minX[slot] = static_cast<uint32_t>(q.w);
      }
    }
    nb += n0 + static_cast<uint32_t>(__popcll(b1));
  };
  if constexpr (W::kPf > 0) {
#pragma unroll
    for (int j = 0; j < W::kPf; ++j) {
      insertPair(lane + j * G, pre.v[j]);
    }
  } else {
    for (int k0 = 0; k0 < size; k0 += G) {
      const int k = k0 + lane;
      insertPair(k, k < size ? seg[k] : make_int4(0, 0, 0, 0));
    }
  }
  waveSync();
  if (nb > BMAX) {
    if (lane == 0) {
      if (tier + 1 < kWaveTiers) {
        // Tier 3's smallest class: k_resolve_large takes its queue after
        // every smaller tier is done.
        const int q = kTierQueues * (kWaveTiers - 1) + kTierQueues - 1;
        tierList[static_cast<size_t>(q) * listCap +
                 atomicAdd(&ctl->waveCount[q], 1u)] = info;
      } else {
        toBig();
      }
    }
    return;
  }
  // The bodies' state is fetched now and lands in LDS after the links.
  float sx[W::kSf > 0 ? W::kSf : 1], sy[W::kSf > 0 ? W::kSf : 1];
  float svx[W::kSf > 0 ? W::kSf : 1], svy[W::kSf > 0 ? W::kSf : 1];
  float sm[W::kSf > 0 ? W::kSf : 1], sr[W::kSf > 0 ? W::kSf : 1];
  if constexpr (W::kSf > 0) {
#pragma unroll
    for (int j = 0; j < W::kSf; ++j) {
      const int t = lane + j * G;
      const int i = t < static_cast<int>(nb) ? bodyIdx[t] : bodyIdx[0];
      sx[j] = px[i];
      sy[j] = py[i];
      svx[j] = vx[i];
      svy[j] = vy[i];
      sm[j] = m[i];
      sr[j] = r[i];
    }
  }
  for (int k = lane; k < size; k += G) {
    const uint32_t h = slots[2 * k];
    slots[2 * k] = static_cast<uint32_t>(hslot[h & 0xffffu]) |
                  (static_cast<uint32_t>(hslot[h >> 16]) << 16);
  }
  waveSync(); // the table is dead from here on

  // 2. Each body's list of pairs (counting sort by slot).
  for (int t = lane; t <= static_cast<int>(nb); t += G) {
    start[t] = 0;
  }
  for (int k = lane; k < size; k += G) {
    indeg[k] = 0;
  }
  waveSync();
  for (int k = lane; k < size; k += G) {
    const uint32_t sl = slots[2 * k];
    atomicAdd(&start[sl & 0xffffu], 1u);
    atomicAdd(&start[sl >> 16], 1u);
  }
  waveSync();
  uint32_t carry = 0, maxDeg = 0;
  for (uint32_t s0 = 0; s0 < nb; s0 += G) {
    const uint32_t t = s0 + static_cast<uint32_t>(lane);
    const uint32_t own = t < nb ? start[t] : 0u;
    maxDeg = max(maxDeg, own);
    uint32_t v = own;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t u = __shfl_up(v, off, G);
      if (lane >= off) {
        v += u;
      }
    }
    if (t < nb) {
      start[t] = carry + v - own;
      cursor[t] = carry + v - own;
    }
    carry += __shfl(v, G - 1, G);
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    maxDeg = max(maxDeg, static_cast<uint32_t>(__shfl_xor(maxDeg, off, G)));
  }
  if (lane == 0) {
    start[nb] = carry;
  }
  if (maxDeg > static_cast<uint32_t>(kMaxDegree)) {
    if (lane == 0) {
      toBig();
    }
    waveSync();
    return;
  }
  waveSync();
  for (int k = lane; k < size; k += G) {
    const uint32_t sl = slots[2 * k];
    inc[atomicAdd(&cursor[sl & 0xffffu], 1u)] = static_cast<uint16_t>(k);
    inc[atomicAdd(&cursor[sl >> 16], 1u)] = static_cast<uint16_t>(k);
  }
  waveSync();

  // 3. Lists into pairLess order; links to the next pair on each body.
  for (int t = lane; t < static_cast<int>(nb); t += G) {
    const uint32_t b0 = start[t], d = start[t + 1] - b0;
    if (d <= 4) {
      linkBody<4>(t, b0, d, inc, slots, minX, bodyIdx, indeg);
    } else if (d <= 8) {
      linkBody<8>(t, b0, d, inc, slots, minX, bodyIdx, indeg);
    } else {
      linkBody<kMaxDegree>(t, b0, d, inc, slots, minX, bodyIdx, indeg);
    }
  }
  waveSync(); // the lists are dead from here on

  // 4. Bodies into LDS; the first round: pairs with no earlier pair.
  if constexpr (W::kSf > 0) {
#pragma unroll
    for (int j = 0; j < W::kSf; ++j) {
      const int t = lane + j * G;
      if (t < static_cast<int>(nb)) {
        lx[t] = sx[j];
        ly[t] = sy[j];
        lvx[t] = svx[j];
        lvy[t] = svy[j];
        lm[t] = sm[j];
        lr[t] = sr[j];
      }
    }
  } else {
    for (int t = lane; t < static_cast<int>(nb); t += G) {
      const int i = bodyIdx[t];
      lx[t] = px[i];
      ly[t] = py[i];
      lvx[t] = vx[i];
      lvy[t] = vy[i];
      lm[t] = m[i];
      lr[t] = r[i];
    }
  }
  // The ready pairs' records go to a ring (at most B / 2 pairs are ready
  // at once: they share no body), the first round's to its start.
  uint64_t *ring = reinterpret_cast<uint64_t *>(inc);
  constexpr uint32_t kRingMask = L::kRing - 1;
  uint32_t head = 0, tail = 0;
  for (int k0 = 0; k0 < size; k0 += G) {
    const int k = k0 + lane;
    const bool ready = k < size && indeg[k] == 0u;
    const uint64_t bm = gballot(ready);
    if (ready) {
      ring[(tail + static_cast<uint32_t>(__popcll(bm & lt))) & kRingMask] = rec[k];
    }
    tail += static_cast<uint32_t>(__popcll(bm));
  }
  waveSync();

  // 5. Rounds: every lane with a ready pair resolves it, decrements its
  //    successors' counts and reads their records in the same LDS round
  //    trip as the bodies; a successor whose count reaches zero is the
  //    lane's next pair (ready: its earlier pairs are all done), a second
  //    one goes to the ring, and lanes without a pair take one from the
  //    ring. Pairs are written back in the order they are replayed.
  uint32_t done = 0;
  uint64_t cur = 0;
  bool has = false;
  while (true) {
    if (head != tail) {
      const uint64_t need = gballot(!has);
      const uint32_t avail = tail - head;
      const uint32_t rank = static_cast<uint32_t>(__popcll(need & lt));
      if (!has && rank < avail) {
        cur = ring[(head + rank) & kRingMask];
        has = true;
      }
      head += min(avail, static_cast<uint32_t>(__popcll(need)));
    }
    const uint64_t act = gballot(has);
    if (act == 0) {
      break;
    }
    const uint32_t sl = static_cast<uint32_t>(cur);
    const uint32_t nx0 = static_cast<uint32_t>(cur >> 32) & 0xffffu;
    const uint32_t nx1 = static_cast<uint32_t>(cur >> 48);
    uint32_t r0 = 0, r1 = 0;
    uint64_t i0 = 0, i1 = 0;
    if (has) {
      // A pair's last link on a body points to itself: its own count is
      // zero (it is ready) and nothing else counts it down any more, so the
      // decrement returns 0 or ~0, never 1, and releases nothing.
      r0 = atomicSub(&indeg[nx0], 1u);
      r1 = atomicSub(&indeg[nx1], 1u);
      i0 = rec[nx0];
      i1 = rec[nx1];
      const int a = static_cast<int>(sl & 0xffffu), b = static_cast<int>(sl >> 16);
      PairBodies s{lx[a], ly[a], lvx[a], lvy[a], lx[b], ly[b], lvx[b], lvy[b]};
      const float m1 = lm[a], m2 = lm[b], r1f = lr[a], r2f = lr[b];
      const int4 out = make_int4(bodyIdx[a], bodyIdx[b], static_cast<int>(minX[a]),
                                 static_cast<int>(minX[b]));
      // Every load of the pair and its successors in one LDS round trip:
      // the operands are consumed here, so none is sunk into a branch.
      asm volatile("" ::"v"(s.p1x), "v"(s.p1y), "v"(s.v1x), "v"(s.v1y), "v"(s.p2x),
                   "v"(s.p2y), "v"(s.v2x), "v"(s.v2y), "v"(m1), "v"(m2), "v"(r1f),
                   "v"(r2f), "v"(out.x), "v"(out.y), "v"(out.z), "v"(out.w), "v"(r0),
                   "v"(r1), "v"(i0), "v"(i1));
      if (resolveCore(s, m1, m2, r1f, r2f)) {
        lx[a] = s.p1x;
        ly[a] = s.p1y;
        lvx[a] = s.v1x;
        lvy[a] = s.v1y;
        lx[b] = s.p2x;
        ly[b] = s.p2y;
        lvx[b] = s.v2x;
        lvy[b] = s.v2y;
      }
      const uint32_t at = done + static_cast<uint32_t>(__popcll(act & lt));
      if (at < static_cast<uint32_t>(size)) {
        seg[at] = out;
      }
    }
    done += static_cast<uint32_t>(__popcll(act));
    const bool go0 = has && r0 == 1u;
    const bool go1 = has && r1 == 1u;
    has = go0 || go1;
    cur = go0 ? i0 : i1;
    const bool push = go0 && go1;
    const uint64_t pm = gballot(push);
    if (pm != 0) {
      if (push) {
        ring[(tail + static_cast<uint32_t>(__popcll(pm & lt))) & kRingMask] = i1;
      }
      tail += static_cast<uint32_t>(__popcll(pm));
    }
    waveSync();
  }

  // 6. Bodies back.
  for (int t = lane; t < static_cast<int>(nb); t += G) {
    const int i = bodyIdx[t];
    px[i] = lx[t];
    py[i] = ly[t];
    vx[i] = lvx[t];
    vy[i] = lvy[t];
  }
  waveSync();
}

// A tier's queues read as one list: entry g of the concatenation (largest
// size class first) at its place in tierList, relative to the tier's first
// queue (queue k starts k listCap further).
struct TierQueues {
  uint32_t end[kTierQueues]; // running totals
  uint32_t count;
  __device__ TierQueues(const Ctl *__restrict__ ctl, int tier) {
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kTierQueues; ++k) {
      run += ctl->waveCount[kTierQueues * tier + k];
      end[k] = run;
    }
    count = run;
  }
  __device__ inline uint32_t slot(uint32_t g, uint32_t listCap) const {
    uint32_t k = 0, before = 0;
#pragma unroll
    for (int q = 0; q + 1 < kTierQueues; ++q) {
      if (g >= end[q]) {
        k = q + 1;
        before = end[q];
      }
    }
    return k * listCap + (g - before);
  }
};

// One wave's share of a tier: the wave takes 64 / G consecutive queue
// entries at a time (the larger components come first), G lanes a
// component -- the first `waves` waves (those resident from the start)
// take their first batch by their index, every later batch comes from the
// tier's counter (so that a queue one residency covers takes no atomic at
// all); the next entries and, for the small tiers, their pairs are loaded
// while the current ones are resolved.
template <int PMAX, int BMAX, int G>
__device__ __forceinline__ void drainTier(
    Ctl *__restrict__ ctl, int tier, uint32_t waveId, uint32_t waves,
    uint32_t *__restrict__ lds,
    int4 *__restrict__ tierList, uint32_t listCap,
    int32_t *__restrict__ bigList, int4 *__restrict__ segPairs,
    float *__restrict__ px, float *__restrict__ py, float *__restrict__ vx,
    float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r) {
  using L = WaveLds<PMAX, BMAX>;
  using W = WaveCtx<PMAX, BMAX, G>;
  constexpr uint32_t kGroups = kWave / G;
  const TierQueues tq(ctl, tier);
  const uint32_t count = tq.count;
  if (count == 0) {
    return;
  }
  const int group = static_cast<int>(threadIdx.x % kWave) / G;
  const int lane = static_cast<int>(threadIdx.x) % G;
  const int4 *list = tierList + static_cast<size_t>(kTierQueues * tier) * listCap;
  uint32_t *next = &ctl->tierNext[tier];
  const uint32_t firstRound = waves * kGroups;
  auto grab = [&]() {
    if (firstRound >= count) {
      return count;
    }
    uint32_t g = 0;
    if (threadIdx.x % kWave == 0) {
      g = atomicAdd(next, kGroups);
    }
    return firstRound + static_cast<uint32_t>(__shfl(static_cast<int>(g), 0));
  };
  const int4 none = make_int4(0, 0, -1, 0);
  // Loads from a clamped index and selects values (a select of pointers
  // would turn the load into a flat one through a private copy of `none`).
  auto entry = [&](uint32_t g) {
    const int4 e = list[tq.slot(min(g, count - 1), listCap)];
    return g < count ? e : none;
  };
  constexpr int kP = W::kPf > 0 ? W::kPf : 1;
  auto pairs = [&](typename W::Pairs &q, const int4 &e) {
#pragma unroll
    for (int j = 0; j < kP; ++j) {
      const int k = lane + j * G;
      q.v[j] = (W::kPf > 0 && k < e.y) ? segPairs[e.x + k] : make_int4(0, 0, 0, 0);
    }
  };
  uint32_t g0 = waveId < waves ? waveId * kGroups : grab();
  if (g0 >= count) {
    return;
  }
  uint32_t g1 = grab();
  int4 info0 = entry(g0 + group), info1 = entry(g1 + group);
  typename W::Pairs q0, q1;
  pairs(q0, info0);
  while (g0 < count) {
    const uint32_t g2 = g1 < count ? grab() : count;
    const int4 info2 = entry(g2 + group);
    pairs(q1, info1);
    if (g0 + group < count) {
      resolveWaveComponent<PMAX, BMAX, G>(lds + group * L::kWords, ctl, tier,
                                          info0, q0, tierList, listCap,
                                          bigList, segPairs, px, py, vx, vy, m,
                                          r);
    }
    g0 = g1;
    g1 = g2;
    info0 = info1;
    info1 = info2;
    q0 = q1;
  }
}

// Tiers 0..2 of the wave path (17..1024 pairs) in one launch, so that the
// long chains of tier 2 run beside tiers 1 and 0 rather than after them.
// A workgroup is two waves over one LDS area of a tier-2 component: the
// first workgroups run tier 2 (one component each, on wave 0, the larger
// first); in every other workgroup each wave holds half of the area and
// drains tier 1 (two components side by side), then tier 0 (four). The grid
// is more than the `resident` workgroups the chip holds at once, so the
// slots tier 2 frees are taken up; a workgroup that finds the queues empty
// returns at once.
constexpr int kTierWaveWords = 2 * WaveLds<256, 192>::kWords;
static_assert(4 * WaveLds<64, 64>::kWords <= kTierWaveWords &&
                  WaveLds<1024, 640>::kWords <= 2 * kTierWaveWords,
              "a wave's half holds tier 1 or 0, the whole area tier 2");
__global__ __launch_bounds__(2 * kWave) void k_resolve_tiers(
    Ctl *__restrict__ ctl, int4 *__restrict__ tierList, uint32_t listCap,
    int32_t *__restrict__ bigList, int4 *__restrict__ segPairs,
    float *__restrict__ px, float *__restrict__ py, float *__restrict__ vx,
    float *__restrict__ vy, const float *__restrict__ m,
    const float *__restrict__ r, uint32_t resident) {
  static_assert(waveTierPairs(0) == 64 && waveTierPairs(1) == 256 &&
                    waveTierPairs(2) == 1024 && kWaveTiers == 4,
                "tier shapes below");
  if (barrierStalled(ctl) || (ctl->status & kStatusPairOverflow)) {
    return;
  }
  __shared__ __align__(16) uint32_t lds[2 * kTierWaveWords];
  const uint32_t wave = threadIdx.x / kWave;
  // Tier 2 on the first workgroups, at most a quarter of the grid.
  const TierQueues tq2(ctl, 2);
  const uint32_t groups2 = min(tq2.count, gridDim.x / 4);
  if (blockIdx.x < groups2) {
    if (wave == 0) {
      const int4 *list = tierList + static_cast<size_t>(kTierQueues * 2) * listCap;
      typename WaveCtx<1024, 640, kWave>::Pairs none{};
      for (uint32_t g = blockIdx.x; g < tq2.count; g += groups2) {
        resolveWaveComponent<1024, 640, kWave>(lds, ctl, 2, list[tq2.slot(g, listCap)],
                                               none, tierList, listCap, bigList,
                                               segPairs, px, py, vx, vy, m, r);
      }
    }
    return;
  }
  uint32_t *half = lds + wave * kTierWaveWords;
  const uint32_t waveId = (blockIdx.x - groups2) * 2 + wave;
  const uint32_t waves = (resident - groups2) * 2;
  drainTier<256, 192, 32>(ctl, 1, waveId, waves, half, tierList, listCap,
                          bigList, segPairs, px, py, vx, vy, m, r);
  drainTier<64, 64, 16>(ctl, 0, waveId, waves, half, tierList, listCap,
                        bigList, segPairs, px, py, vx, vy, m, r);
}

// ---- contact components of more than 4096 pairs, or that a wave tier
//      cannot hold ----
//
// One workgroup per component. Sort into pairLess order: up to kBigCompMax
// pairs by a bitonic sort in LDS; longer components as runs of kBigCompMax
// sorted in LDS, then merge-path rounds in global memory, ping-ponging with
// `tmp` (the same offsets of the detection's pair array, no longer needed
// once the segments are filled). Replay: every body of the component claims
// a local slot in `bodySlot` (all -1 at rest; the claims are cleared again),
// and when they fit, the bodies are staged in LDS and one lane replays the
// pairs on the staged copies -- the same resolvePair arithmetic in the same
// order -- a few hundred cycles per pair instead of a memory round trip.
// Components with more than kReplayBodies bodies are replayed in global
// memory.
constexpr int kReplayBodies = 5120;
constexpr int kReplayChunk = 1024; // pairs whose slots are staged at once
constexpr int kMergeItems = 8;     // merge outputs per lane per round step
constexpr int kResolveLds =        // words: sort area or replay area
    (kReplayBodies * 7 + kReplayChunk * 2 > kBigCompMax * 4 + kBigCompMax / 2)
        ? kReplayBodies * 7 + kReplayChunk * 2
        : kBigCompMax * 4 + kBigCompMax / 2;

// Sorts run[0, len) (len <= kBigCompMax) into pairLess order in LDS.
__device__ inline void sortRunInLds(int4 *__restrict__ run, int len,
                                    int4 *__restrict__ sp,
                                    int16_t *__restrict__ order) {
  int p2 = 1;
  while (p2 < len) {
    p2 <<= 1;
  }
  for (int k = threadIdx.x; k < p2; k += kBlock) {
    if (k < len) {
      sp[k] = run[k];
    }
    order[k] = static_cast<int16_t>(k);
  }
  __syncthreads();
  for (int width = 2; width <= p2; width <<= 1) {
    for (int stride = width >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < p2; k += kBlock) {
        const int partner = k ^ stride;
        if (partner > k) {
          const bool up = (k & width) == 0;
          const int a = order[k], c = order[partner];
          // Padding (index >= len) sorts after every real pair.
          const bool aGreater = a >= len ? (c < len || a > c)
                                         : (c < len && pairLess(sp[c], sp[a]));
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
    run[k] = sp[order[k]];
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
}

// Merge path: how many of the first `diag` merged outputs come from A.
__device__ inline int mergeSplit(const int4 *__restrict__ a, int na,
                                 const int4 *__restrict__ b, int nb, int diag) {
  int lo = max(0, diag - nb), hi = min(diag, na);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pairLess(b[diag - 1 - mid], a[mid])) {
      hi = mid;
    } else {
      lo = mid + 1;
    }
  }
  return lo;
}

// One contact component by the whole workgroup (info = its compInfo
// entry): pairLess order, then the replay (the two steps described above).
// lds: kResolveLds words.
__device__ inline void resolveBigComponent(
    int4 info, uint32_t *__restrict__ lds, int &nBodies,
    int4 *__restrict__ segPairs, int4 *__restrict__ tmpPairs,
    int32_t *__restrict__ bodySlot, float *__restrict__ px,
    float *__restrict__ py, float *__restrict__ vx, float *__restrict__ vy,
    const float *__restrict__ m, const float *__restrict__ r) {
  // Sort area.
  int4 *sp = reinterpret_cast<int4 *>(lds);
  int16_t *order = reinterpret_cast<int16_t *>(lds + 4 * kBigCompMax);
  // Replay area.
  float *lx = reinterpret_cast<float *>(lds);
  float *ly = lx + kReplayBodies, *lvx = ly + kReplayBodies;
  float *lvy = lvx + kReplayBodies, *lm = lvy + kReplayBodies;
  float *lr = lm + kReplayBodies;
  int32_t *bodyIdx = reinterpret_cast<int32_t *>(lr + kReplayBodies);
  int2 *chunk = reinterpret_cast<int2 *>(bodyIdx + kReplayBodies);
  const int start = info.y, size = info.z;
  int4 *seg = segPairs + start;

  // 1. pairLess order.
  if (size <= kBigCompMax) {
    sortRunInLds(seg, size, sp, order);
  } else {
    for (int c0 = 0; c0 < size; c0 += kBigCompMax) {
      sortRunInLds(seg + c0, min(kBigCompMax, size - c0), sp, order);
    }
    int4 *src = seg, *dst = tmpPairs + start;
    for (int w = kBigCompMax; w < size; w <<= 1) {
      for (int base = 0; base < size; base += 2 * w) {
        const int na = min(w, size - base);
        const int nb = max(0, min(w, size - base - na));
        const int4 *a = src + base, *b = src + base + na;
        const int total = na + nb;
        for (int o = threadIdx.x * kMergeItems; o < total;
             o += kBlock * kMergeItems) {
          int ia = mergeSplit(a, na, b, nb, o);
          int ib = o - ia;
          const int cnt = min(kMergeItems, total - o);
          for (int q = 0; q < cnt; ++q) {
            const bool takeA =
                ib >= nb || (ia < na && !pairLess(b[ib], a[ia]));
            dst[base + o + q] = takeA ? a[ia++] : b[ib++];
          }
        }
      }
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __syncthreads();
      int4 *t = src;
      src = dst;
      dst = t;
    }
    if (src != seg) {
      for (int k = threadIdx.x; k < size; k += kBlock) {
        seg[k] = src[k];
      }
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      __syncthreads();
    }
  }

  // 2. Local slots for the component's bodies.
  if (threadIdx.x == 0) {
    nBodies = 0;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < size; k += kBlock) {
    const int4 q = seg[k];
    const int ends[2] = {q.x, q.y};
    for (int e = 0; e < 2; ++e) {
      if (atomicCAS(&bodySlot[ends[e]], -1, -2) == -1) {
        const int slot = atomicAdd(&nBodies, 1);
        storeAgent(reinterpret_cast<uint32_t *>(&bodySlot[ends[e]]),
                   static_cast<uint32_t>(slot));
        if (slot < kReplayBodies) {
          bodyIdx[slot] = ends[e];
        }
      }
    }
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  const int nb = nBodies;

  // 3. Replay in pairLess order.
  if (nb <= kReplayBodies) {
    for (int s = threadIdx.x; s < nb; s += kBlock) {
      const int i = bodyIdx[s];
      lx[s] = px[i];
      ly[s] = py[i];
      lvx[s] = vx[i];
      lvy[s] = vy[i];
      lm[s] = m[i];
      lr[s] = r[i];
    }
    for (int c0 = 0; c0 < size; c0 += kReplayChunk) {
      const int cnt = min(kReplayChunk, size - c0);
      __syncthreads();
      for (int k = threadIdx.x; k < cnt; k += kBlock) {
        const int4 q = seg[c0 + k];
        chunk[k] = make_int2(
            static_cast<int>(loadAgent(reinterpret_cast<const uint32_t *>(&bodySlot[q.x]))),
            static_cast<int>(loadAgent(reinterpret_cast<const uint32_t *>(&bodySlot[q.y]))));
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int k = 0; k < cnt; ++k) {
          resolvePair(lx, ly, lvx, lvy, lm, lr, chunk[k].x, chunk[k].y);
        }
      }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nb; s += kBlock) {
      const int i = bodyIdx[s];
      px[i] = lx[s];
      py[i] = ly[s];
      vx[i] = lvx[s];
      vy[i] = lvy[s];
      bodySlot[i] = -1;
    }
  } else {
    if (threadIdx.x == 0) {
      for (int k = 0; k < size; ++k) {
        const int4 q = seg[k];
        resolvePair(px, py, vx, vy, m, r, q.x, q.y);
      }
    }
    for (int k = threadIdx.x; k < size; k += kBlock) {
      const int4 q = seg[k];
      bodySlot[q.x] = -1;
      bodySlot[q.y] = -1;
    }
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
}

// The largest contact components, in one launch: first tier 3 of the wave
// path (1025..4096 pairs; wave 0 of each workgroup, 64 lanes a component),
// where a component that the tier cannot hold is run by the workgroup path
// at once; then the components earlier kernels sent to the workgroup path
// (bigList, complete when this kernel starts). LDS holds either layout.
constexpr int kLargeLds = WaveLds<4096, 2560>::kWords > kResolveLds
                              ? WaveLds<4096, 2560>::kWords
                              : kResolveLds;
__global__ __launch_bounds__(kBlock) void k_resolve_large(
    Ctl *__restrict__ ctl, int4 *__restrict__ tierList, uint32_t listCap,
    const int4 *__restrict__ compInfo, int32_t *__restrict__ bigList,
    int4 *__restrict__ segPairs, int4 *__restrict__ tmpPairs,
    int32_t *__restrict__ bodySlot, float *__restrict__ px,
    float *__restrict__ py, float *__restrict__ vx, float *__restrict__ vy,
    const float *__restrict__ m, const float *__restrict__ r) {
  if (barrierStalled(ctl)) {
    return;
  }
  __shared__ __align__(16) uint32_t lds[kLargeLds];
  __shared__ int nBodies;
  __shared__ int local;
  constexpr int kTier = kWaveTiers - 1;
  static_assert(waveTierPairs(kTier) == 4096, "tier 3 is 1025..4096 pairs");
  using W = WaveCtx<4096, 2560, kWave>;
  static_assert(W::kPf == 0, "tier 3 loads its pairs itself");
  if (!(ctl->status & kStatusPairOverflow)) {
    TierQueues tq(ctl, kTier);
    const uint32_t count = tq.count;
    const int4 *list = tierList + static_cast<size_t>(kTierQueues * kTier) * listCap;
    for (uint32_t g = blockIdx.x; g < count; g += gridDim.x) {
      if (threadIdx.x == 0) {
        local = -1;
      }
      __syncthreads();
      if (threadIdx.x < kWave) {
        typename W::Pairs none{};
        resolveWaveComponent<4096, 2560, kWave>(
            lds, ctl, kTier, list[tq.slot(g, listCap)], none, tierList,
            listCap, bigList,
            segPairs, px, py, vx, vy, m, r, &local);
      }
      __syncthreads();
      const int c = local;
      if (c >= 0) {
        resolveBigComponent(compInfo[c], lds, nBodies, segPairs, tmpPairs,
                            bodySlot, px, py, vx, vy, m, r);
      }
      __syncthreads();
    }
  }
  const uint32_t big = ctl->bigCount;
  for (uint32_t g = blockIdx.x; g < big; g += gridDim.x) {
    resolveBigComponent(compInfo[bigList[g]], lds, nBodies, segPairs,
                        tmpPairs, bodySlot, px, py, vx, vy, m, r);
  }
}

} // namespace nbx
