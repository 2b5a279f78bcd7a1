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

// Exclusive scan of one value per thread across the block; returns the
// prefix and writes the block total to *total. Uses `scratch` (>= 4 words).
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
  for (int w = 0; w < kBlock / kWave; ++w) {
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
// [base + t*kScanItems, ...), so the output order is the input order.
__global__ __launch_bounds__(kBlock) void k_scan_down(
    uint32_t count, const uint32_t *__restrict__ in,
    const uint32_t *__restrict__ tileSums, uint32_t *__restrict__ out) {
  __shared__ uint32_t scratch[kBlock / kWave];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const bool full = base + kScanItems <= count;
  uint32_t v[kScanItems];
  if (full) {
    // Four 16-byte loads (the buffers are 16-byte aligned, base is a
    // multiple of 16 values).
    const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      const uint4 q = p[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < count ? in[base + k] : 0u;
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

} // namespace nbx
