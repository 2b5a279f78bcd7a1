// k_mlp.h -- the learned far-field correction: a 10 -> 64 -> 64 -> 64 -> 2
// GELU network that predicts, for an accepted far node and a target body,
// the residual between the node's monopole (QuadTree.cpp:276-296) and the
// exact force of the node's bodies (the law of CheckCuda.cu:456-475). No
// reference counterpart. BF16 weights and activations, FP32 accumulation,
// on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Tiles. A wave evaluates 32 (target, node) pairs at once with activations
// held transposed, features x pairs: the pairs are the MFMA's N dimension.
// In a 32x32 accumulator, lane l holds column l % 32 (one pair) and rows
// 8 (v / 4) + 4 (l / 32) + v % 4, v = 0..15 -- so after a layer each lane
// holds 32 of its pair's 64 hidden values (16 per 32-row tile), which is
// exactly what the next layer's B operand needs (lane l: 8 consecutive K
// entries 8 (l / 32) .. + 7 of column l % 32) once the K order of each
// 16-wide K step is taken as the order the accumulator rows arrive in. That
// permutation is applied to the weights when they are packed (mlpVirtualCol),
// so no activation moves between lanes from the input to the output.
//
// Packed weights (bf16, "A-operand order"): for every (M tile, K step) the
// 64 lanes' 8-element operands, one 16-byte load each. Forward: W1 (2 x 1),
// W2, W3 (2 x 4), W4 (1 x 4, rows 2..31 zero); backward, transposed: W2^T,
// W3^T (2 x 4), W4^T (2 x 1). Biases are kept in FP32 in accumulator order
// and start the accumulators.
#pragma once

#include "nbx_device.h"

namespace nbx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMlpIn = 10;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpParams = 9154;
constexpr int kMlpTile = 32; // pairs per wave evaluation
// Flat parameter layout, row-major [out][in].
constexpr int kOW1 = 0, kOB1 = 640, kOW2 = 704, kOB2 = 4800, kOW3 = 4864,
              kOB3 = 8960, kOW4 = 9024, kOB4 = 9152;
// Packed operand offsets, in bf16x8 units.
constexpr int kPkW1 = 0;
constexpr int kPkW2 = kPkW1 + 2 * 64;
constexpr int kPkW3 = kPkW2 + 8 * 64;
constexpr int kPkW4 = kPkW3 + 8 * 64;
constexpr int kPkFwd = kPkW4 + 4 * 64; // forward operands end here
constexpr int kPkW2T = kPkFwd;
constexpr int kPkW3T = kPkW2T + 8 * 64;
constexpr int kPkW4T = kPkW3T + 8 * 64;
constexpr int kPkAll = kPkW4T + 2 * 64;
// Biases in accumulator order (floats): [M tile][lane half][16].
constexpr int kBcB1 = 0, kBcB2 = 64, kBcB3 = 128, kBcB4 = 192, kBcAll = 224;

// The trained range and the input/output scales (the oracle's OR_MLP_*).
constexpr float kMlpMaxRatio = 1.2f;
constexpr float kMlpBoxMargin = 0.25f;
constexpr float kMlpMomentScale = 8.0f;
constexpr float kMlpOutScale = 0.5f;

// GELU, tanh form: x * sigmoid(A (x + B x^3)).
constexpr float kGeluA = 1.5957691216057308f;
constexpr float kGeluB = 0.044715f;
constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Accumulator row of element v held by lane half h in M tile mt.
__host__ __device__ constexpr int mlpCRow(int v, int h, int mt) {
  return 32 * mt + 8 * (v / 4) + 4 * h + (v % 4);
}
// Feature index that K step ks, lane half h, operand element q stands for
// when the operand comes from a 64-row activation held in accumulator
// order: tile ks / 2, elements v = 8 (ks % 2) + q.
__host__ __device__ constexpr int mlpVirtualCol(int ks, int h, int q) {
  return 32 * (ks / 2) + 16 * (ks % 2) + 8 * (q / 4) + 4 * h + (q % 4);
}

__device__ __forceinline__ float geluSig(float x) {
  const float z = kGeluA * (x + kGeluB * x * x * x);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * kLog2e));
}
__device__ __forceinline__ float gelu(float x) { return x * geluSig(x); }
__device__ __forceinline__ float geluGrad(float x) {
  const float s = geluSig(x);
  return s + x * s * (1.0f - s) * kGeluA * (1.0f + 3.0f * kGeluB * x * x);
}

__device__ __forceinline__ f32x16 biasTile(const float *__restrict__ bc,
                                           int off, uint32_t h) {
  const float4 *p = reinterpret_cast<const float4 *>(bc + off + 16 * h);
  const float4 a = p[0], b = p[1], c = p[2], d = p[3];
  f32x16 r = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w,
              c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
  return r;
}

// GELU of a 64-row activation (two accumulator tiles) packed to BF16 as the
// next layer's four K-step operands.
__device__ __forceinline__ void geluPack(const f32x16 z[2], bf16x8 hb[4]) {
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      hb[ks][q] = static_cast<__bf16>(gelu(z[ks >> 1][8 * (ks & 1) + q]));
    }
  }
}

// One hidden layer: out[mt] = bias + W (packed at `w`) x act (4 K steps).
__device__ __forceinline__ void mlpLayer64(const bf16x8 *__restrict__ w,
                                           const float *__restrict__ bc,
                                           int boff, const bf16x8 act[4],
                                           f32x16 out[2], uint32_t lane) {
  const uint32_t h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16 acc = biasTile(bc, boff + 32 * mt, h);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      acc = mfma32(w[(mt * 4 + ks) * 64 + lane], act[ks], acc);
    }
    out[mt] = acc;
  }
}

// Writes a 64-row activation held in accumulator order (packed operands)
// feature-major: row f, column = this lane's pair.
__device__ __forceinline__ void storeRows(__bf16 (*rows)[kMlpTile],
                                          const bf16x8 a[4], uint32_t lane) {
  const uint32_t j = lane & 31, h = lane >> 5;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      rows[mlpVirtualCol(ks, static_cast<int>(h), q)][j] = a[ks][q];
    }
  }
}

// Forward pass of 32 pairs. x: this lane's operand of its pair's inputs
// (features 8h .. 8h + 7, zero beyond the 10th). Returns the output tile:
// lanes 0..31 hold their pair's two outputs in elements 0 and 1. With KEEP
// (training) the pre-activations go to zs and the BF16 activations, feature
// major, to rows[layer].
template <bool KEEP>
__device__ __forceinline__ f32x16 mlpForward(const bf16x8 *__restrict__ pk,
                                             const float *__restrict__ bc,
                                             bf16x8 x, uint32_t lane,
                                             f32x16 (*zs)[2] = nullptr,
                                             __bf16 (*rows)[64][kMlpTile] = nullptr) {
  const uint32_t h = lane >> 5;
  f32x16 z[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    z[mt] = mfma32(pk[kPkW1 + mt * 64 + lane], x, biasTile(bc, kBcB1 + 32 * mt, h));
  }
  bf16x8 a[4];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    if (l > 0) {
      mlpLayer64(pk + (l == 1 ? kPkW2 : kPkW3), bc, l == 1 ? kBcB2 : kBcB3, a, z,
                 lane);
    }
    geluPack(z, a);
    if (KEEP) {
      zs[l][0] = z[0];
      zs[l][1] = z[1];
      storeRows(rows[l], a, lane);
    }
  }
  f32x16 y = biasTile(bc, kBcB4, h);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    y = mfma32(pk[kPkW4 + ks * 64 + lane], a[ks], y);
  }
  return y;
}

// The network's 10 inputs for target offset d = COM - x (dSq = |d|^2), the
// node's size and `q` = {Sxx/M, Sxy/M, Syy/M, log2(M)/16} (nodeLearn).
// Identical arithmetic in data generation and in the walk.
__device__ __forceinline__ void mlpFeatures(float dx, float dy, float dSq,
                                            float eSq, float size, float4 q,
                                            float f[kMlpIn]) {
  const float w = __builtin_amdgcn_rcpf(dSq + eSq);
  const float irho = __builtin_amdgcn_rsqf(dSq + eSq);
  const float invR = __builtin_amdgcn_rsqf(dSq);
  const float k = kMlpMomentScale * w;
  const float sdx = q.x * dx + q.y * dy;
  const float sdy = q.y * dx + q.z * dy;
  const float iR2k = invR * invR * k;
  const float rad = (dx * sdx + dy * sdy) * iR2k;
  f[0] = q.w;
  f[1] = dx * irho;
  f[2] = dy * irho;
  f[3] = size * invR;
  f[4] = q.x * k;
  f[5] = q.y * k;
  f[6] = q.z * k;
  f[7] = rad;
  f[8] = (dx * sdy - dy * sdx) * iR2k;
  f[9] = (q.x + q.z) * k - rad;
}

// Lane half h's operand of a pair's inputs: features 8h .. 8h + 7.
__device__ __forceinline__ bf16x8 mlpInput(const float f[kMlpIn], uint32_t h) {
  const float hi[8] = {f[8], f[9], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8 x;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    x[q] = static_cast<__bf16>(h ? hi[q] : f[q]);
  }
  return x;
}

// The trained range, in FP32 exactly as the oracle tests it: the node's
// size at most kMlpMaxRatio of its distance and its square at least
// kMlpBoxMargin sizes away from the target.
__device__ __forceinline__ bool mlpInRange(float x, float y, float cX, float cY,
                                           float size, float dSq) {
  const float maxSq = kMlpMaxRatio * kMlpMaxRatio;
  const float hx = fmaxf(fabsf(x - cX) - 0.5f * size, 0.0f);
  const float hy = fmaxf(fabsf(y - cY) - 0.5f * size, 0.0f);
  const float g = kMlpBoxMargin * size;
  return (size * size <= maxSq * dSq) && (hx * hx + hy * hy >= g * g);
}

// ---- Packing: master FP32 parameters -> BF16 operands + FP32 biases. ----
__global__ __launch_bounds__(kBlock) void k_mlp_pack(
    const float *__restrict__ p, __bf16 *__restrict__ pk,
    float *__restrict__ bc) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < kPkAll * 8) {
    const int u = e >> 3, q = e & 7;
    float v = 0.0f;
    if (u < kPkW2) { // W1: [mt][lane]
      const int mt = u / 64, lane = u % 64, h = lane / 32;
      const int row = lane % 32 + 32 * mt, col = 8 * h + q;
      v = col < kMlpIn ? p[kOW1 + row * kMlpIn + col] : 0.0f;
    } else if (u < kPkW4) { // W2, W3: [mt][ks][lane]
      const bool second = u >= kPkW3;
      const int r = u - (second ? kPkW3 : kPkW2);
      const int mt = r / 256, ks = (r / 64) % 4, lane = r % 64, h = lane / 32;
      const int row = lane % 32 + 32 * mt, col = mlpVirtualCol(ks, h, q);
      v = p[(second ? kOW3 : kOW2) + row * kMlpHidden + col];
    } else if (u < kPkFwd) { // W4: [ks][lane], rows >= 2 zero
      const int r = u - kPkW4;
      const int ks = r / 64, lane = r % 64, h = lane / 32;
      const int row = lane % 32, col = mlpVirtualCol(ks, h, q);
      v = row < kMlpOut ? p[kOW4 + row * kMlpHidden + col] : 0.0f;
    } else if (u < kPkW4T) { // W2^T, W3^T: rows = inputs, K = outputs
      const bool second = u >= kPkW3T;
      const int r = u - (second ? kPkW3T : kPkW2T);
      const int mt = r / 256, ks = (r / 64) % 4, lane = r % 64, h = lane / 32;
      const int in = lane % 32 + 32 * mt, out = mlpVirtualCol(ks, h, q);
      v = p[(second ? kOW3 : kOW2) + out * kMlpHidden + in];
    } else { // W4^T: [mt][lane], K = the 2 outputs (then zeros)
      const int r = u - kPkW4T;
      const int mt = r / 64, lane = r % 64, h = lane / 32;
      const int in = lane % 32 + 32 * mt, out = 8 * h + q;
      v = out < kMlpOut ? p[kOW4 + out * kMlpHidden + in] : 0.0f;
    }
    pk[e] = static_cast<__bf16>(v);
  }
  if (e < kBcAll) {
    float v = 0.0f;
    if (e < kBcB4) {
      const int layer = e / 64, idx = e % 64;
      const int mt = idx / 32, h = (idx / 16) % 2, vv = idx % 16;
      const int boff = layer == 0 ? kOB1 : layer == 1 ? kOB2 : kOB3;
      v = p[boff + mlpCRow(vv, h, mt)];
    } else {
      const int idx = e - kBcB4, h = idx / 16, vv = idx % 16;
      const int row = mlpCRow(vv, h, 0);
      v = row < kMlpOut ? p[kOB4 + row] : 0.0f;
    }
    bc[e] = v;
  }
}

// ---- Training: gradient of the mean squared error over a batch. ----
//
// Batch position k (0 <= k < batch) reads dataset pair (a (k0 + k) + c) mod
// count, a permutation of the dataset when gcd(a, count) = 1. Each
// workgroup (4 waves) takes `per` consecutive positions, 32 per wave and
// pass; weight gradients need the pairs in the K dimension, so activations
// and deltas are also written feature-major to LDS (the transpose) and the
// dW products read their operands from there. Per-workgroup sums are
// accumulated in LDS and written to partial[blockIdx.x][0 .. kMlpParams]
// (the last entry: the summed squared error).
constexpr int kMlpTrainBlock = 256;
constexpr int kMlpTrainWaves = kMlpTrainBlock / kWave;

struct MlpWaveLds {
  alignas(16) __bf16 x[16][kMlpTile];    // inputs, feature-major
  alignas(16) __bf16 h[3][64][kMlpTile]; // hidden activations H1..H3
  alignas(16) __bf16 dz[64][kMlpTile];   // the current layer's delta
};

__device__ __forceinline__ void ldsFence() { asm volatile("" ::: "memory"); }

// dW += delta x act^T over this wave's 32 pairs: rows `dzRows` (M = outputs,
// `mtiles` tiles), columns `actRows` (N = inputs, `ntiles` tiles, `nin`
// valid), summed into the LDS accumulator acc[out * nin + in].
__device__ __forceinline__ void weightGrad(const __bf16 (*dzRows)[kMlpTile],
                                           int nout,
                                           const __bf16 (*actRows)[kMlpTile],
                                           int nin, int mtiles, int ntiles,
                                           float *acc, uint32_t lane) {
  const uint32_t i32 = lane & 31, h = lane >> 5;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    if (mt >= mtiles) break;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      if (nt >= ntiles) break;
      f32x16 c = {};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const int o = static_cast<int>(i32) + 32 * mt;
        const int in = static_cast<int>(i32) + 32 * nt;
        bf16x8 a = {}, b = {};
        if (o < nout) {
          a = *reinterpret_cast<const bf16x8 *>(&dzRows[o][16 * ks + 8 * h]);
        }
        if (in < nin) {
          b = *reinterpret_cast<const bf16x8 *>(&actRows[in][16 * ks + 8 * h]);
        }
        c = mfma32(a, b, c);
      }
      const int in = static_cast<int>(i32) + 32 * nt;
      if (in < nin) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int o = mlpCRow(v, static_cast<int>(h), mt);
          if (o < nout) {
            atomicAdd(&acc[o * nin + in], c[v]);
          }
        }
      }
    }
  }
}

// Bias gradient: lane o sums row o of the delta.
__device__ __forceinline__ void biasGrad(const __bf16 (*dzRows)[kMlpTile],
                                         int nout, float *accB, uint32_t lane) {
  if (static_cast<int>(lane) < nout) {
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < kMlpTile; ++j) {
      s += static_cast<float>(dzRows[lane][j]);
    }
    atomicAdd(&accB[lane], s);
  }
}

// Delta of a hidden layer: dH (accumulator order) times GELU'(Z), rounded to
// BF16: packed operands for the next backward product and feature-major rows.
__device__ __forceinline__ void hiddenDelta(const f32x16 dh[2],
                                            const f32x16 z[2], bf16x8 d[4],
                                            __bf16 (*rows)[kMlpTile],
                                            uint32_t lane) {
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int v = 8 * (ks & 1) + q;
      d[ks][q] = static_cast<__bf16>(dh[ks >> 1][v] * geluGrad(z[ks >> 1][v]));
    }
  }
  storeRows(rows, d, lane);
}

// dH = W^T delta for a 64 x 64 layer (packed transposed operands at wt).
__device__ __forceinline__ void backLayer64(const bf16x8 *__restrict__ wt,
                                            const bf16x8 d[4], f32x16 dh[2],
                                            uint32_t lane) {
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16 acc = {};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      acc = mfma32(wt[(mt * 4 + ks) * 64 + lane], d[ks], acc);
    }
    dh[mt] = acc;
  }
}

__global__ __launch_bounds__(kMlpTrainBlock) void k_mlp_grad(
    const bf16x8 *__restrict__ pk, const float *__restrict__ bc,
    const float4 *__restrict__ data, uint64_t count, uint64_t mulA,
    uint64_t addC, uint64_t k0, uint32_t batch, uint32_t per, float invBW,
    float *__restrict__ partial) {
  __shared__ float acc[kMlpParams + 1];
  __shared__ MlpWaveLds wl[kMlpTrainWaves];
  for (int i = threadIdx.x; i < kMlpParams + 1; i += kMlpTrainBlock) {
    acc[i] = 0.0f;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t j = lane & 31, h = lane >> 5;
  MlpWaveLds &L = wl[wave];
  const uint32_t first = blockIdx.x * per;
  const uint32_t last = min(batch, first + per);
  float lossSum = 0.0f;
  for (uint32_t t = first + wave * kMlpTile; t < last;
       t += kMlpTrainWaves * kMlpTile) {
    const uint32_t pos = t + j;
    const bool valid = pos < last;
    float f[kMlpIn] = {};
    float t0 = 0.0f, t1 = 0.0f, wt = 0.0f;
    if (valid) {
      const uint64_t idx = (mulA * ((k0 + pos) % count) + addC) % count;
      const float4 a = data[4 * idx], b = data[4 * idx + 1], c = data[4 * idx + 2];
      f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
      f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
      f[8] = c.x; f[9] = c.y;
      t0 = c.z;
      t1 = c.w;
      wt = data[4 * idx + 3].x;
    }
    const bf16x8 x = mlpInput(f, h);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      L.x[8 * h + q][j] = x[q];
    }
    f32x16 zs[3][2];
    const f32x16 y = mlpForward<true>(pk, bc, x, lane, zs, L.h);
    // Output delta of the weighted loss sum w |y - t|^2 / (batch W):
    // 2 (y - t) w / (batch W), lanes of half 0 (their pair's outputs).
    float e0 = 0.0f, e1 = 0.0f;
    const float sc = wt * invBW;
    if (h == 0 && valid) {
      e0 = y[0] - t0;
      e1 = y[1] - t1;
      lossSum += (e0 * e0 + e1 * e1) * sc;
    }
    bf16x8 d4 = {};
    d4[0] = static_cast<__bf16>(2.0f * e0 * sc);
    d4[1] = static_cast<__bf16>(2.0f * e1 * sc);
    if (h == 0) {
      L.dz[0][j] = d4[0];
      L.dz[1][j] = d4[1];
    }
    ldsFence();
    weightGrad(L.dz, kMlpOut, L.h[2], kMlpHidden, 1, 2, acc + kOW4, lane);
    biasGrad(L.dz, kMlpOut, acc + kOB4, lane);
    // dH3 = W4^T d4 (K = the two outputs).
    f32x16 dh[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x16 z0 = {};
      dh[mt] = mfma32(pk[kPkW4T + mt * 64 + lane], d4, z0);
    }
    bf16x8 d[4];
    ldsFence();
    hiddenDelta(dh, zs[2], d, L.dz, lane);
    ldsFence();
    weightGrad(L.dz, kMlpHidden, L.h[1], kMlpHidden, 2, 2, acc + kOW3, lane);
    biasGrad(L.dz, kMlpHidden, acc + kOB3, lane);
    backLayer64(pk + kPkW3T, d, dh, lane);
    ldsFence();
    hiddenDelta(dh, zs[1], d, L.dz, lane);
    ldsFence();
    weightGrad(L.dz, kMlpHidden, L.h[0], kMlpHidden, 2, 2, acc + kOW2, lane);
    biasGrad(L.dz, kMlpHidden, acc + kOB2, lane);
    backLayer64(pk + kPkW2T, d, dh, lane);
    ldsFence();
    hiddenDelta(dh, zs[0], d, L.dz, lane);
    ldsFence();
    weightGrad(L.dz, kMlpHidden, L.x, kMlpIn, 2, 1, acc + kOW1, lane);
    biasGrad(L.dz, kMlpHidden, acc + kOB1, lane);
    ldsFence();
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lossSum += __shfl_xor(lossSum, off, kWave);
  }
  if (lane == 0) {
    atomicAdd(&acc[kMlpParams], lossSum);
  }
  __syncthreads();
  float *out = partial + static_cast<size_t>(blockIdx.x) * (kMlpParams + 1);
  for (int i = threadIdx.x; i < kMlpParams + 1; i += kMlpTrainBlock) {
    out[i] = acc[i];
  }
}

// Sums the workgroups' partials in order; with `update` applies one Adam
// step to the master parameters (m, v moments; bc1 = 1 / (1 - beta1^t),
// bc2 = 1 / (1 - beta2^t)), else writes the gradient to `grad`. The loss
// (weighted mean squared error of the batch) goes to *loss.
__global__ __launch_bounds__(kBlock) void k_mlp_adam(
    const float *__restrict__ partial, uint32_t groups,
    int update, float lr, float bc1, float bc2, float *__restrict__ p,
    float *__restrict__ m, float *__restrict__ v, float *__restrict__ grad,
    float *__restrict__ loss) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i > kMlpParams) {
    return;
  }
  float g = 0.0f;
  for (uint32_t w = 0; w < groups; ++w) {
    g += partial[static_cast<size_t>(w) * (kMlpParams + 1) + i];
  }
  if (i == kMlpParams) {
    *loss = g;
    return;
  }
  if (!update) {
    grad[i] = g;
    return;
  }
  constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * bc1) / (sqrtf(vi * bc2) + eps);
}

// Sum of the loss weights of pairs [first, first + count) (FP64).
__global__ __launch_bounds__(kBlock) void k_mlp_wsum(
    const float4 *__restrict__ data, uint64_t first, uint64_t count,
    double *__restrict__ out) {
  __shared__ double part[kBlock / kWave];
  double s = 0.0;
  for (uint64_t k = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; k < count;
       k += static_cast<uint64_t>(gridDim.x) * kBlock) {
    s += data[4 * (first + k) + 3].x;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, kWave);
  }
  if ((threadIdx.x & 63) == 0) {
    part[threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t += part[w];
    atomicAdd(out, t);
  }
}

// Network outputs of `count` pairs (inputs, 10 floats each).
__global__ __launch_bounds__(kBlock) void k_mlp_predict(
    const bf16x8 *__restrict__ pk, const float *__restrict__ bc,
    const float *__restrict__ feat, uint32_t count, float *__restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t j = lane & 31, h = lane >> 5;
  const uint32_t pos = (blockIdx.x * kBlock + (threadIdx.x & ~63u)) / 2 + j;
  float f[kMlpIn] = {};
  if (pos < count) {
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) f[k] = feat[static_cast<size_t>(pos) * kMlpIn + k];
  }
  const f32x16 y = mlpForward<false>(pk, bc, mlpInput(f, h), lane);
  if (h == 0 && pos < count) {
    out[2 * pos] = y[0];
    out[2 * pos + 1] = y[1];
  }
}

// ---- Data generation on the tree of a step. ----

// Per node: {Sxx/M, Sxy/M, Syy/M, log2(M)/16} from the second moments the
// build wrote (nodeQ) and the node's mass; zero for an empty node.
__global__ __launch_bounds__(kBlock) void k_node_learn(
    const float4 *__restrict__ nodeF, const float4 *__restrict__ nodeQ,
    const Ctl *__restrict__ ctl, float4 *__restrict__ nodeL) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= ctl->nodeCount) {
    return;
  }
  const float mass = nodeF[k].z;
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  if (mass > 0.0f) {
    const float4 q = nodeQ[k];
    const float im = 1.0f / mass;
    o = make_float4(q.x * im, q.y * im, q.z * im, __log2f(mass) * (1.0f / 16.0f));
  }
  nodeL[k] = o;
}

// One thread per target: the stackless preorder walk (next pointers, as
// QuadTree.cpp:264-389 walks) at the coarse theta, recording every accepted
// node of positive mass inside the trained range; accepted nodes outside it
// are opened instead. Pairs {target body, node, target's list index} go to
// `pairs` (at most cap; the counter keeps counting past it); norm[t] gets
// |sum of the recorded nodes' monopoles| (FP64), the scale of the pairs'
// loss weights.
__global__ __launch_bounds__(kBlock) void k_mlp_gen_pairs(
    uint32_t nt, const uint32_t *__restrict__ targets,
    const float4 *__restrict__ body, const float4 *__restrict__ nodeF,
    const int4 *__restrict__ nodeI, const float *__restrict__ centerX,
    const float *__restrict__ centerY, const float *__restrict__ size,
    const Ctl *__restrict__ ctl, float tSq, double eps2, int4 *__restrict__ pairs,
    uint32_t cap, uint32_t *__restrict__ counter, double *__restrict__ norm) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) {
    return;
  }
  const uint32_t ti = targets[t];
  const float x = body[ti].x, y = body[ti].y;
  const int nodeCount = ctl->nodeCount;
  int node = 0;
  double ax = 0.0, ay = 0.0;
  for (int guard = 0; guard <= nodeCount; ++guard) {
    const float4 f = nodeF[node];
    const int4 c = nodeI[node];
    const float dx = f.x - x, dy = f.y - y;
    const float dSq = dx * dx + dy * dy;
    bool far = f.w < dSq * tSq;
    if (far && f.z > 0.0f &&
        !mlpInRange(x, y, centerX[node], centerY[node], size[node], dSq)) {
      far = false;
    }
    if (far || c.x == 0) {
      if (far && f.z > 0.0f) {
        const uint32_t slot = atomicAdd(counter, 1u);
        if (slot < cap) {
          pairs[slot] = make_int4(static_cast<int>(ti), node, static_cast<int>(t), 0);
        }
        const double ddx = static_cast<double>(f.x) - x;
        const double ddy = static_cast<double>(f.y) - y;
        const double r2 = ddx * ddx + ddy * ddy;
        const double fm = static_cast<double>(f.z) / (sqrt(r2) * (r2 + eps2));
        ax += ddx * fm;
        ay += ddy * fm;
      }
      if (c.y == 0) {
        break;
      }
      node = c.y;
    } else {
      node = c.x;
    }
    if (node <= 0 || node >= nodeCount) {
      break;
    }
  }
  norm[t] = sqrt(ax * ax + ay * ay);
}

// One wave per pair: inputs (the walk's own arithmetic), the regression
// target -- the FP64 residual of the node's bodies against its monopole, in
// units of u = M w mu (w = 1 / (r^2 + eps^2), mu = w tr(S/M)) times
// kMlpOutScale, zero when u = 0 -- and the loss weight (u / norm[target])^2:
// what an output error costs that target's force, squared and relative.
// Written as four float4 per pair: {f0..f3}, {f4..f7}, {f8, f9, t0, t1},
// {weight, 0, 0, 0}.
__global__ __launch_bounds__(kBlock) void k_mlp_label(
    uint32_t count, const int4 *__restrict__ pairs,
    const double *__restrict__ norm, const float4 *__restrict__ body,
    const float4 *__restrict__ nodeF, const int4 *__restrict__ nodeI,
    const float4 *__restrict__ nodeL, const float *__restrict__ size,
    float eSq, double eps2, float4 *__restrict__ out) {
  const uint32_t k = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  if (k >= count) {
    return;
  }
  const int4 pr = pairs[k];
  const float4 me = body[pr.x];
  const float4 f = nodeF[pr.y];
  const int4 c = nodeI[pr.y];
  double sx = 0.0, sy = 0.0;
  for (int b = c.z + static_cast<int>(lane); b < c.w; b += 64) {
    if (b == pr.x) {
      continue;
    }
    const float4 o = body[b];
    const double dx = static_cast<double>(o.x) - me.x;
    const double dy = static_cast<double>(o.y) - me.y;
    const double d2 = dx * dx + dy * dy;
    if (d2 == 0.0) {
      continue;
    }
    const double fac = static_cast<double>(o.z) / (sqrt(d2) * (d2 + eps2));
    sx += dx * fac;
    sy += dy * fac;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, kWave);
    sy += __shfl_xor(sy, off, kWave);
  }
  if (lane == 0) {
    const float dx = f.x - me.x, dy = f.y - me.y;
    const float dSq = dx * dx + dy * dy;
    float ft[kMlpIn];
    mlpFeatures(dx, dy, dSq, eSq, size[pr.y], nodeL[pr.y], ft);
    const double ddx = static_cast<double>(f.x) - me.x;
    const double ddy = static_cast<double>(f.y) - me.y;
    const double r2 = ddx * ddx + ddy * ddy;
    const double fm = static_cast<double>(f.z) / (sqrt(r2) * (r2 + eps2));
    // Output unit: monopole magnitude times the quadrupole scale tr(S/M) w.
    const float4 q = nodeL[pr.y];
    const double mu = (static_cast<double>(q.x) + q.z) / (r2 + eps2);
    const double unit = static_cast<double>(f.z) / (r2 + eps2) * mu;
    const double scale = unit > 0.0 ? kMlpOutScale / unit : 0.0;
    const float t0 = static_cast<float>((sx - ddx * fm) * scale);
    const float t1 = static_cast<float>((sy - ddy * fm) * scale);
    const double at = norm[pr.z];
    const double share = at > 0.0 ? unit / at : 0.0;
    out[4 * k] = make_float4(ft[0], ft[1], ft[2], ft[3]);
    out[4 * k + 1] = make_float4(ft[4], ft[5], ft[6], ft[7]);
    out[4 * k + 2] = make_float4(ft[8], ft[9], t0, t1);
    out[4 * k + 3] = make_float4(static_cast<float>(share * share), 0.f, 0.f, 0.f);
  }
}

} // namespace nbx
