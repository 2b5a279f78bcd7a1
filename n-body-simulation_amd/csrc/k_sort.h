// k_sort.h -- device-wide exclusive scan, stable LSD radix sort (8-bit
// digits, LDS-staged scatter).
#pragma once

#include "nbx_device.h"

namespace nbx {

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems; // 4096 values per tile
#ifndef NBX_SORT_ITEMS
#define NBX_SORT_ITEMS 4
#endif
constexpr int kSortItems = NBX_SORT_ITEMS;
constexpr int kSortTile = kBlock * kSortItems; // 4096 keys per tile

// Exclusive scan of one value per thread across a BLOCK-lane workgroup;
// returns the prefix and writes the block total to *total. Uses `scratch`
// (>= BLOCK / 64 words).
template <int BLOCK = kBlock>
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t *scratch,
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
#pragma unroll
  for (int w = 0; w < BLOCK / kWave; ++w) {
    const uint32_t s = scratch[w];
    waveOff += (w < wave) ? s : 0u;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return waveOff + x - v;
}

// ---- exclusive scan over `count` uint32 (in place allowed) ----

__global__ __launch_bounds__(kBlock) void k_scan_reduce(
    uint32_t count, const uint32_t *__restrict__ in,
    uint32_t *__restrict__ tileSums) {
  __shared__ uint32_t scratch[kBlock / kWave];
  // Thread t sums the 16 consecutive values [base + 16 t, ...): four
  // 16-byte loads, all issued before any is used.
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
  if (base + kScanItems <= count) {
    const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
    uint4 q[kScanItems / 4];
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      q[k] = p[k];
    }
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      s += q[k].x + q[k].y + q[k].z + q[k].w;
    }
  } else {
    for (int k = 0; k < kScanItems; ++k) {
      s += base + k < count ? in[base + k] : 0u;
    }
  }
  uint32_t total;
  blockExclusiveScan(s, scratch, &total);
  if (threadIdx.x == 0) {
    tileSums[blockIdx.x] = total;
  }
}

// One workgroup: exclusive scan of `tiles` tile sums, in place.
__global__ __launch_bounds__(kBlock) void k_scan_tiles(
    uint32_t tiles, uint32_t *__restrict__ tileSums) {
  __shared__ uint32_t scratch[kBlock / kWave];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < tiles; base += kBlock) {
    const uint32_t idx = base + threadIdx.x;
    const uint32_t v = idx < tiles ? tileSums[idx] : 0u;
    uint32_t total;
    const uint32_t pre = blockExclusiveScan(v, scratch, &total);
    if (idx < tiles) {
      tileSums[idx] = carry + pre;
    }
    carry += total;
  }
}

// Blocked per-thread scan of a tile: thread t owns values
// [base + t*kScanItems, ...), so the output order is the input order. The
// input (the grid's bucket counts) is cleared once read.
__global__ __launch_bounds__(kBlock) void k_scan_down(
    uint32_t count, uint32_t *__restrict__ in,
    const uint32_t *__restrict__ tileSums, uint32_t *__restrict__ out) {
  __shared__ uint32_t scratch[kBlock / kWave];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const bool full = base + kScanItems <= count;
  uint32_t v[kScanItems];
  if (full) {
    // Four 16-byte loads (the buffers are 16-byte aligned, base is a
    // multiple of 16 values).
    uint4 *p = reinterpret_cast<uint4 *>(in + base);
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      const uint4 q = p[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      p[k] = make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < count ? in[base + k] : 0u;
      if (base + k < count) {
        in[base + k] = 0u;
      }
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    s += v[k];
  }
  uint32_t total;
  uint32_t run = blockExclusiveScan(s, scratch, &total) + tileSums[blockIdx.x];
  uint32_t o[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (full) {
    uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      p[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < count) {
        out[base + k] = o[k];
      }
    }
  }
}

// ---- stable LSD radix sort, one 8-bit digit per pass ----

// Per-tile digit histogram -> hist[digit * tiles + tile].
__global__ __launch_bounds__(kBlock) void k_rs_hist(
    uint32_t n, const uint32_t *__restrict__ keys, int shift,
    uint32_t *__restrict__ hist, uint32_t tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t idx = base + k * kBlock + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
    const uint64_t peers = matchDigit8(d, valid);
    if (valid && (peers & laneMaskLt()) == 0) {
      atomicAdd(&h[d], static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();
  hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Block d: exclusive scan, in place, of digit d's per-tile counts (one row
// of the digit-major table) and the row total. The scatter adds the digit's
// base (scan of the 256 row totals) itself.
__global__ __launch_bounds__(kBlock) void k_rs_scan(
    uint32_t tiles, uint32_t *__restrict__ hist,
    uint32_t *__restrict__ rowTotal) {
  __shared__ uint32_t scratch[kBlock / kWave];
  uint32_t *row = hist + static_cast<size_t>(blockIdx.x) * tiles;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < tiles; base += kBlock) {
    const uint32_t t = base + threadIdx.x;
    const uint32_t v = t < tiles ? row[t] : 0u;
    uint32_t total;
    const uint32_t pre = blockExclusiveScan(v, scratch, &total);
    if (t < tiles) {
      row[t] = carry + pre;
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    rowTotal[blockIdx.x] = carry;
  }
}

// Stable scatter of one tile: ranks each key among equal digits in tile
// order, stages the tile digit-sorted in LDS, then writes runs coalesced.
__global__ __launch_bounds__(kBlock) void k_rs_scatter(
    uint32_t n, const uint32_t *__restrict__ keysIn,
    const uint32_t *__restrict__ valsIn, int shift,
    const uint32_t *__restrict__ offsets,
    const uint32_t *__restrict__ rowTotal, uint32_t tiles,
    uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut) {
  __shared__ uint32_t sk[kSortTile];
  __shared__ uint32_t sv[kSortTile];
  __shared__ uint32_t digitStart[256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t waveCnt[kBlock / kWave][256];
  __shared__ uint32_t scratch[kBlock / kWave];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const uint32_t base = blockIdx.x * kSortTile;
  const uint32_t tileCount = min<uint32_t>(kSortTile, n - base);

  // Tile histogram -> tile-local digit starts.
  run[tid] = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    waveCnt[w][tid] = 0;
  }
  __syncthreads();
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t idx = base + k * kBlock + tid;
    const bool valid = idx < n;
    const uint32_t d = valid ? (keysIn[idx] >> shift) & 255u : 0u;
    const uint64_t peers = matchDigit8(d, valid);
    if (valid && (peers & laneMaskLt()) == 0) {
      atomicAdd(&run[d], static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();
  uint32_t total;
  digitStart[tid] = blockExclusiveScan(run[tid], scratch, &total);
  const uint32_t digitBase = blockExclusiveScan(rowTotal[tid], scratch, &total);
  goff[tid] = digitBase + offsets[tid * tiles + blockIdx.x];
  __syncthreads();
  run[tid] = 0;
  __syncthreads();

  // Rank in tile order (k, tid) and stage.
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t idx = base + k * kBlock + tid;
    const bool valid = idx < n;
    const uint32_t key = valid ? keysIn[idx] : 0u;
    const uint32_t val = valid ? valsIn[idx] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    const uint64_t peers = matchDigit8(d, valid);
    const uint32_t wrank = static_cast<uint32_t>(__popcll(peers & laneMaskLt()));
    if (valid && wrank == 0) {
      waveCnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();
    if (valid) {
      uint32_t pre = run[d];
      for (int w = 0; w < wave; ++w) {
        pre += waveCnt[w][d];
      }
      const uint32_t pos = digitStart[d] + pre + wrank;
      sk[pos] = key;
      sv[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      add += waveCnt[w][tid];
      waveCnt[w][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }

  for (uint32_t p = tid; p < tileCount; p += kBlock) {
    const uint32_t key = sk[p];
    const uint32_t d = (key >> shift) & 255u;
    const uint32_t dst = goff[d] + (p - digitStart[d]);
    keysOut[dst] = key;
    valsOut[dst] = sv[p];
  }
}

// ---- stable LSD radix sort, all four passes in one launch ----
//
// G workgroups of 1024 lanes, all resident at once (G is at most the CU
// count and one such workgroup fits a CU), meeting at grid barriers.
// Workgroup b owns the contiguous chunk [b*C, b*C + C) of every pass's input,
// C <= ITEMS * 1024. One pass:
//   A. load the chunk, rank every key among equal digits in input order
//      (per wave, then across waves), stage the chunk digit-sorted in LDS,
//      publish its digit counts to table[digit * G + b]              barrier
//   B. workgroup d (d < 256, stepping by G): exclusive scan of table row d
//      in place, and the row total                                    barrier
//   C. write each staged digit run at its global offset                barrier
// so a pass costs three grid barriers instead of three launches. Data that
// crosses workgroups inside the launch (the keys and payloads between
// passes, the table, the row totals) is read and written with agent-scope
// (sc1) loads and stores, which are coherent across the XCDs' L2s; the
// barrier then only has to order them, not write back or invalidate L2.
constexpr int kPsBlock = 1024;
constexpr int kPsWaves = kPsBlock / kWave;
constexpr uint32_t kPsMaxGrid = 1024; // phase B scans one row entry per lane
constexpr uint32_t kBarrierSpins = 1u << 20; // default poll bound, ~1 s
// Barrier words: arrival counter [0], generation [32], one arrival counter
// per XCD at [64 + 32 x] (128 bytes apart).
constexpr int kBarrierWords = 64 + 32 * kXcds;

// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt fields at their
// maxima, so only vector memory operations are waited for).
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ inline uint32_t loadAgent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void storeAgent(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over agent-scope data. Every wave waits for its own memory
// operations to complete, the workgroup meets, and one lane arrives: first
// at its XCD's counter (workgroups b with b % 8 == x), the last arrival of
// each XCD then at the top counter -- 32 + 8 serialised atomics instead of
// 256 on one address. The last top arrival resets the counters and bumps
// the generation; the others poll it. Counters are zero at rest. The poll is
// bounded: a barrier that does not complete (workgroups not co-resident)
// sets kStatusBarrierStalled and returns, so the kernel always ends; the host
// reports the error and clears the barrier words.
__device__ inline void gridBarrier(uint32_t *__restrict__ bar, uint32_t blocks,
                                   uint32_t *__restrict__ status,
                                   uint32_t spinLimit) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t *gen = bar + 32;
    const uint32_t g = loadAgent(gen);
    const uint32_t x = blockIdx.x % kXcds;
    const uint32_t members = blocks / kXcds + (x < blocks % kXcds ? 1u : 0u);
    const uint32_t groups = min(blocks, kXcds);
    uint32_t *sub = bar + 64 + 32 * x;
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0); // g read before arriving
    bool last = false;
    if (__hip_atomic_fetch_add(sub, 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
      storeAgent(sub, 0u);
      last = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == groups - 1;
    }
    if (last) {
      storeAgent(bar, 0u);
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0); // resets land before release
      storeAgent(gen, g + 1u);
    } else {
      uint32_t spins = 0;
      while (loadAgent(gen) == g) {
        if (++spins > spinLimit || (loadAgent(status) & kStatusBarrierStalled)) {
          __hip_atomic_fetch_or(status, kStatusBarrierStalled, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  __syncthreads();
}

// Largest bucket the MSD path sorts in LDS: 8 keys a lane (a top-digit
// bucket of a uniform 1M-body disc holds about 5100).
constexpr uint32_t kMsdCap = 8 * kPsBlock;
constexpr int kMsdItems = kMsdCap / kPsBlock;

// One stable counting pass over the first `cnt` keys in LDS, by digit
// (key >> shift) & 255, from (srcK, srcV) to (dstK, dstV): `it` keys a
// lane at positions wave * 64 * it + j * 64 + lane (j < it), ranked among
// equal digits per wave (ballots), then across waves (wh: per-wave digit
// counts; dstart: digit starts).
__device__ inline void countingPassLds(const uint32_t *__restrict__ srcK,
                                       const uint32_t *__restrict__ srcV,
                                       uint32_t *__restrict__ dstK,
                                       uint32_t *__restrict__ dstV, int it,
                                       uint32_t cnt, int shift,
                                       uint32_t (*wh)[256],
                                       uint32_t *__restrict__ dstart,
                                       uint32_t *__restrict__ scratch) {
  const uint32_t t = threadIdx.x, lane = t & (kWave - 1), wave = t >> 6;
  const uint32_t first = wave * kWave * static_cast<uint32_t>(it) + lane;
  for (uint32_t i = t; i < kPsWaves * 256; i += kPsBlock) {
    (&wh[0][0])[i] = 0;
  }
  __syncthreads();
  uint32_t rank[kMsdItems], dg[kMsdItems];
#pragma unroll
  for (int j = 0; j < kMsdItems; ++j) {
    const uint32_t p = first + j * kWave;
    dg[j] = (j < it && p < cnt) ? (srcK[p] >> shift) & 255u : 0u;
  }
#pragma unroll
  for (int j = 0; j < kMsdItems; ++j) {
    rank[j] = 0;
    if (j < it) {
      const uint32_t p = first + j * kWave;
      const bool valid = p < cnt;
      const uint32_t dig = dg[j];
      const uint64_t peers = matchDigit8(dig, valid);
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & laneMaskLt()));
      const uint32_t base = wh[wave][dig];
      rank[j] = base + below;
      if (valid && below == 0) {
        wh[wave][dig] = base + static_cast<uint32_t>(__popcll(peers));
      }
    }
  }
  __syncthreads();
  // Per digit, the exclusive prefix of the 16 waves' counts: lane t takes
  // digit t / 4 and waves 4 (t % 4) .. 4 (t % 4) + 3, the four quarters are
  // scanned across four neighbouring lanes, and the digit totals (on the
  // quarter-3 lanes) are scanned across the workgroup in lane order, which
  // is digit order.
  {
    const uint32_t d = t >> 2, q = t & 3u;
    const uint32_t c0 = wh[4 * q][d], c1 = wh[4 * q + 1][d];
    const uint32_t c2 = wh[4 * q + 2][d], c3 = wh[4 * q + 3][d];
    const uint32_t sum = c0 + c1 + c2 + c3;
    uint32_t x = sum;
    uint32_t y = __shfl_up(x, 1, 4);
    x += q >= 1 ? y : 0u;
    y = __shfl_up(x, 2, 4);
    x += q >= 2 ? y : 0u;
    const uint32_t ex = x - sum;
    wh[4 * q][d] = ex;
    wh[4 * q + 1][d] = ex + c0;
    wh[4 * q + 2][d] = ex + c0 + c1;
    wh[4 * q + 3][d] = ex + c0 + c1 + c2;
    uint32_t total;
    const uint32_t pre =
        blockExclusiveScan<kPsBlock>(q == 3 ? x : 0u, scratch, &total);
    if (q == 0) {
      dstart[d] = pre;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kMsdItems; ++j) {
    const uint32_t p = first + j * kWave;
    if (j < it && p < cnt) {
      const uint32_t k = srcK[p];
      const uint32_t dig = (k >> shift) & 255u;
      const uint32_t pos = dstart[dig] + wh[wave][dig] + rank[j];
      dstK[pos] = k;
      dstV[pos] = srcV[p];
    }
  }
  __syncthreads();
}

// After the MSD pass: bucket d (the keys whose top digit is d) lies at
// [bstart[d], bstart[d] + bsize[d]) of (bucketKeys, bucketVals); each is
// sorted stably by the three lower digits in LDS (two buffers of kMsdCap
// keys and payloads, sk and sv holding 2 kMsdCap words each) and written to
// (outKeys, outVals) at the same place. The non-empty buckets go to the
// workgroups in digit order, one each for workgroups 0..G-1, then snaking
// back from G-1 (a uniform disc fills about 224 of the 256 top cells, so at
// 1M bodies with G = 245 every workgroup sorts at most one). Every address
// is bounds-checked (a stalled barrier leaves values in range).
// A body the sort places this many positions or more from its old place
// counts as moved far; when many do (the first build from an unsorted
// order), k_permute gathers whole records staged by k_deep instead of seven
// scattered words a body.
constexpr uint32_t kFarMove = 1024;
__device__ inline void msdBuckets(uint32_t n, uint32_t grid,
                                  const uint32_t *bucketKeys,
                                  const uint32_t *bucketVals,
                                  uint32_t *outKeys, uint32_t *outVals,
                                  const uint32_t *__restrict__ bsize,
                                  const uint32_t *__restrict__ bstart,
                                  uint32_t *__restrict__ sk,
                                  uint32_t *__restrict__ sv,
                                  uint32_t (*wh)[256],
                                  uint32_t *__restrict__ dstart,
                                  uint32_t *__restrict__ list,
                                  uint32_t *__restrict__ scratch,
                                  uint32_t *__restrict__ farMoves) {
  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x;
  uint32_t far = 0; // bodies this workgroup placed kFarMove or more from
                    // where they were (an order far from the last one's)
  // The non-empty buckets in digit order (in `list`): one each for
  // workgroups 0..G-1, then snaking back from G-1.
  __shared__ uint32_t nonEmpty;
  {
    const bool nz = t < 256 && bsize[t] > 0;
    uint32_t total;
    const uint32_t r = blockExclusiveScan<kPsBlock>(nz ? 1u : 0u, scratch, &total);
    if (nz) {
      list[r] = t;
    }
    if (t == 0) {
      nonEmpty = total;
    }
  }
  __syncthreads();
  const uint32_t count = nonEmpty;
  uint32_t *k0 = sk, *k1 = sk + kMsdCap, *v0 = sv, *v1 = sv + kMsdCap;
  for (uint32_t round = 0; round * grid < count; ++round) {
    const uint32_t slot = round * grid + ((round & 1u) ? grid - 1u - b : b);
    if (slot >= count) {
      continue;
    }
    const uint32_t bk = list[slot];
    const uint32_t cnt = min(bsize[bk], kMsdCap);
    const uint32_t st = min(bstart[bk], n);
    const uint32_t len = min(cnt, n - st);
    const int it = static_cast<int>((len + kPsBlock - 1) / kPsBlock);
    for (uint32_t p = t; p < len; p += kPsBlock) {
      k0[p] = loadAgent(bucketKeys + st + p);
      v0[p] = loadAgent(bucketVals + st + p);
    }
    __syncthreads();
    countingPassLds(k0, v0, k1, v1, it, len, 0, wh, dstart, scratch);
    countingPassLds(k1, v1, k0, v0, it, len, 8, wh, dstart, scratch);
    countingPassLds(k0, v0, k1, v1, it, len, 16, wh, dstart, scratch);
    for (uint32_t p = t; p < len; p += kPsBlock) {
      outKeys[st + p] = k1[p];
      outVals[st + p] = v1[p];
      const uint32_t to = st + p, from = v1[p];
      far += (from > to ? from - to : to - from) >= kFarMove ? 1u : 0u;
    }
    __syncthreads();
  }
  uint32_t total;
  (void)blockExclusiveScan<kPsBlock>(far, scratch, &total);
  if (t == 0 && total > 0) {
    atomicAdd(farMoves, total);
  }
}

// Sorts (keys0, vals0) by key, stably, through (keys1, vals1); the result
// ends in (keys0, vals0). table holds 256 * gridDim.x words, rowTotal 256;
// chunk <= ITEMS * 1024. Global writes are bounds-checked, so even a stalled
// barrier leaves every value in range.
//
// BUILD: the tree build's sort, with the bounding box and keys folded in
// (what k_bbox and k_keys do in two launches): each workgroup reduces its
// chunk's bounding box and publishes it to parts[b]; after a grid barrier
// every workgroup reduces all G partials to the root quad, computes its
// bodies' high key words by the float descent (indices as payloads) and
// ranks them straight from registers. Workgroup 0 publishes the root quad
// and clears the build's counters. The low words (depths 16..31) are
// computed only where they are read: by k_deep, for the runs of equal high
// words.
template <int ITEMS, bool BUILD>
__global__ __launch_bounds__(kPsBlock) void k_sort_onelaunch(
    uint32_t n, uint32_t chunk, uint32_t *keys0, uint32_t *vals0,
    uint32_t *keys1, uint32_t *vals1, uint32_t *table, uint32_t *rowTotal,
    uint32_t *bar, uint32_t spinLimit, Ctl *ctl, const float *px,
    const float *py, uint32_t *parts) {
  constexpr int kTile = ITEMS * kPsBlock;
  constexpr int kStage = (BUILD && 2 * kMsdCap > kTile) ? 2 * kMsdCap : kTile;
  __shared__ uint32_t sk[kStage];
  __shared__ uint32_t sv[kStage];
  __shared__ uint32_t wh[kPsWaves][256]; // per wave: digit counts, then offsets
  __shared__ uint32_t dstart[256], goff[256];
  __shared__ uint32_t scratch[kPsWaves];
  __shared__ uint4 red[kPsWaves];
  __shared__ uint32_t bsize[BUILD ? 256 : 1], bstart[BUILD ? 256 : 1];

  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x, grid = gridDim.x;
  const uint32_t lane = t & (kWave - 1), wave = t >> 6;
  const uint32_t lo = min(n, b * chunk);
  const uint32_t hi = min(n, lo + chunk);
  const uint32_t cnt = hi - lo;
  uint32_t *status = &ctl->status;
  // Wave w owns chunk positions [w * 64 * ITEMS, (w + 1) * 64 * ITEMS):
  // item j of a lane sits at w * 64 * ITEMS + j * 64 + lane, so (wave, item,
  // lane) is input order.
  const uint32_t first = wave * kWave * ITEMS + lane;
  // Row-scan segment: the smallest power of two >= the grid.
  uint32_t seg = 1;
  while (seg < grid) {
    seg <<= 1;
  }

  uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
  if (BUILD) {
    // Positions are read twice (bounding box, then keys after the barrier;
    // the second read hits L2) rather than held across the barrier.
    uint4 bb = bboxEmpty();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t p = first + j * kWave;
      if (p < cnt) {
        bb = bboxAdd(bb, px[lo + p], py[lo + p]);
      }
    }
    bb = bboxWave(bb);
    if (lane == 0) {
      red[wave] = bb;
    }
    __syncthreads();
    if (t == 0) {
      for (int w = 1; w < kPsWaves; ++w) {
        bb = bboxMerge(bb, red[w]);
      }
      storeAgent(parts + 4 * b, bb.x);
      storeAgent(parts + 4 * b + 1, bb.y);
      storeAgent(parts + 4 * b + 2, bb.z);
      storeAgent(parts + 4 * b + 3, bb.w);
    }
    gridBarrier(bar, grid, status, spinLimit);
    bb = bboxEmpty();
    if (t < grid) {
      bb = make_uint4(loadAgent(parts + 4 * t), loadAgent(parts + 4 * t + 1),
                      loadAgent(parts + 4 * t + 2),
                      loadAgent(parts + 4 * t + 3));
    }
    bb = bboxWave(bb);
    if (lane == 0) {
      red[wave] = bb;
    }
    __syncthreads();
    bb = red[0];
    for (int w = 1; w < kPsWaves; ++w) {
      bb = bboxMerge(bb, red[w]);
    }
    float qcx, qcy, qs;
    rootQuad(bb, qcx, qcy, qs);
    if (b == 0 && t == 0) {
      publishRoot(ctl, bb, qcx, qcy, qs);
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t p = first + j * kWave;
      uint32_t h = 0;
      if (p < cnt) {
        float dcx = qcx, dcy = qcy, ds = qs;
        h = descentWord(px[lo + p], py[lo + p], dcx, dcy, ds);
      }
      key[j] = h;
      val[j] = lo + p;
    }
  }

  // BUILD: pass -1 distributes the keys by their top digit (MSD) into
  // buckets in keys1/vals1, and every workgroup then sorts whole buckets by
  // the three lower digits in LDS (msdBuckets) -- four grid barriers instead
  // of twelve. When a bucket is larger than the LDS holds, every workgroup
  // sees it after the row scans and the four LSD passes run instead, from
  // the keys still in registers.
  bool msdDone = false;
  for (int pass = BUILD ? -1 : 0; pass < 4; ++pass) {
    const bool msd = pass < 0;
    const int shift = msd ? 24 : 8 * pass;
    const uint32_t *kin = (pass & 1) ? keys1 : keys0;
    const uint32_t *vin = (pass & 1) ? vals1 : vals0;
    uint32_t *kout = (msd || (pass & 1) == 0) ? keys1 : keys0;
    uint32_t *vout = (msd || (pass & 1) == 0) ? vals1 : vals0;

    // A. load, rank, stage, publish.
    for (uint32_t i = t; i < kPsWaves * 256; i += kPsBlock) {
      (&wh[0][0])[i] = 0;
    }
    if (!BUILD || pass > 0) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t p = first + j * kWave;
        key[j] = p < cnt ? loadAgent(kin + lo + p) : 0u;
        val[j] = p < cnt ? loadAgent(vin + lo + p) : 0u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const bool valid = first + j * kWave < cnt;
      const uint32_t dig = (key[j] >> shift) & 255u;
      const uint64_t peers = matchDigit8(dig, valid);
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & laneMaskLt()));
      const uint32_t base = wh[wave][dig];
      rank[j] = base + below;
      if (valid && below == 0) {
        wh[wave][dig] = base + static_cast<uint32_t>(__popcll(peers));
      }
    }
    __syncthreads();
    // Per digit: wave counts -> offsets of each wave's run; chunk count.
    uint32_t r = 0;
    if (t < 256) {
#pragma unroll
      for (int w = 0; w < kPsWaves; ++w) {
        const uint32_t c = wh[w][t];
        wh[w][t] = r;
        r += c;
      }
      storeAgent(table + t * grid + b, r);
    }
    {
      uint32_t total;
      const uint32_t pre = blockExclusiveScan<kPsBlock>(r, scratch, &total);
      if (t < 256) {
        dstart[t] = pre;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (first + j * kWave < cnt) {
        const uint32_t dig = (key[j] >> shift) & 255u;
        const uint32_t pos = dstart[dig] + wh[wave][dig] + rank[j];
        sk[pos] = key[j];
        sv[pos] = val[j];
      }
    }
    gridBarrier(bar, grid, status, spinLimit);

    // B. row scans, all of this workgroup's rows (b, b + G, ...) at once:
    // lane t takes entry t % S of row b + G * (t / S), S = segment, and a
    // segmented scan per row (S is a power of two >= G, so S <= 2G - 1 and
    // 1024 / S >= 256 / G rows fit).
    {
      const uint32_t row = b + grid * (t / seg);
      const uint32_t e = t & (seg - 1);
      const bool inRow = row < 256 && t / seg < kPsBlock / seg;
      const uint32_t v =
          inRow && e < grid ? loadAgent(table + row * grid + e) : 0u;
      uint32_t x = v;
#pragma unroll
      for (uint32_t off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (off < seg && (lane & (seg - 1)) >= off) {
          x += y;
        }
      }
      if (seg > kWave) {
        // Segments span whole waves: add the totals of the segment's
        // earlier waves.
        if (lane == kWave - 1) {
          scratch[wave] = x;
        }
        __syncthreads();
        const uint32_t w0 = (t / seg) * (seg / kWave);
        for (uint32_t w = w0; w < wave; ++w) {
          x += scratch[w];
        }
        __syncthreads();
      }
      if (inRow && e < grid) {
        storeAgent(table + row * grid + e, x - v);
      }
      if (inRow && e == seg - 1) {
        storeAgent(rowTotal + row, x);
      }
    }
    gridBarrier(bar, grid, status, spinLimit);

    // C. write the staged runs.
    {
      const uint32_t v = t < 256 ? loadAgent(rowTotal + t) : 0u;
      const uint32_t own = t < 256 ? loadAgent(table + t * grid + b) : 0u;
      uint32_t total;
      const uint32_t base = blockExclusiveScan<kPsBlock>(v, scratch, &total);
      if (t < 256) {
        goff[t] = base + own;
      }
      if (msd) {
        if (t < 256) {
          bsize[t] = v;
          bstart[t] = base;
        }
        // The largest bucket, the same in every workgroup.
        uint32_t mx = v;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
          mx = max(mx, static_cast<uint32_t>(__shfl_xor(mx, off, kWave)));
        }
        if (lane == 0) {
          red[wave].x = mx;
        }
        __syncthreads();
        mx = 0;
        for (uint32_t w = 0; w < kPsWaves; ++w) {
          mx = max(mx, red[w].x);
        }
        __syncthreads();
        if (mx > kMsdCap) {
          continue; // LSD passes 0..3 from the registers
        }
      }
    }
    __syncthreads();
    for (uint32_t p = t; p < cnt; p += kPsBlock) {
      const uint32_t kk = sk[p];
      const uint32_t dig = (kk >> shift) & 255u;
      const uint32_t dst = goff[dig] + (p - dstart[dig]);
      if (dst < n) {
        storeAgent(kout + dst, kk);
        storeAgent(vout + dst, sv[p]);
      }
    }
    if (pass < 3) {
      gridBarrier(bar, grid, status, spinLimit);
    }
    if (msd) {
      msdDone = true;
      break;
    }
  }
  if (BUILD && msdDone) {
    msdBuckets(n, grid, keys1, vals1, keys0, vals0, bsize, bstart, sk, sv,
               wh, dstart, goff, scratch, &ctl->farMoves);
  }
}

// ---- exclusive scan in one launch ----
//
// G resident workgroups (as for k_sort_onelaunch); workgroup b owns the
// chunk [b*C, b*C + C), C <= ITEMS * 1024 and a multiple of 4, each lane
// ITEMS consecutive values (16-byte loads). The chunk is reduced and its
// total published; after one grid barrier each workgroup adds the totals of
// the workgroups before it and writes its chunk's prefixes from registers.
// Replaces the reduce / tile scan / down-sweep launches. The input (the
// grid's bucket counts) is cleared once read, ready for the next count.
template <int ITEMS>
__global__ __launch_bounds__(kPsBlock) void k_scan_onelaunch(
    uint32_t count, uint32_t chunk, uint32_t *__restrict__ in,
    uint32_t *__restrict__ out, uint32_t *tileSums, uint32_t *bar,
    uint32_t spinLimit, Ctl *ctl) {
  static_assert(ITEMS % 4 == 0, "16-byte loads");
  __shared__ uint32_t scratch[kPsWaves];
  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x, grid = gridDim.x;
  const uint32_t lo = min(count, b * chunk);
  const uint32_t hi = min(count, lo + chunk);
  const uint32_t base = lo + t * ITEMS;
  const bool full = base + ITEMS <= hi;
  uint32_t v[ITEMS];
  if (full) {
    uint4 *p = reinterpret_cast<uint4 *>(in + base);
#pragma unroll
    for (int k = 0; k < ITEMS / 4; ++k) {
      const uint4 q = p[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < ITEMS / 4; ++k) {
      p[k] = make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      v[k] = base + k < hi ? in[base + k] : 0u;
      if (base + k < hi) {
        in[base + k] = 0u;
      }
    }
  }
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = blockExclusiveScan<kPsBlock>(sum, scratch, &total);
  if (t == 0) {
    storeAgent(tileSums + b, total);
  }
  gridBarrier(bar, grid, &ctl->status, spinLimit);
  // Totals of the workgroups before this one (b < G <= 1024 lanes).
  {
    uint32_t before;
    blockExclusiveScan<kPsBlock>(t < b ? loadAgent(tileSums + t) : 0u,
                                 scratch, &before);
    run += before;
  }
  uint32_t o[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (full) {
    uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
    for (int k = 0; k < ITEMS / 4; ++k) {
      p[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (base + k < hi) {
        out[base + k] = o[k];
      }
    }
  }
}

} // namespace nbx
