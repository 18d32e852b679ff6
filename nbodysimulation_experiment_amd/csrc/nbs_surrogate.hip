// Learned relaxation kernels (BF16 MFMA, FP32 accumulation) for gfx950.
//
// v_mfma_f32_32x32x8_bf16 operand layouts (lane l, g = l / 32, c = l % 32):
//   A (32 x 8):  lane holds A[c][4g + t], t = 0..3
//   B (8 x 32):  lane holds B[4g + t][c]
//   C (32 x 32): register r holds C[8(r/4) + 4g + (r%4)][c]
// The MFMA self-test checks this layout against a host matrix product.
//
// Pair layer: one wave serves 64 particles, one per lane.  Each lane walks its
// own particle's candidates (the three row ranges of its 3x3 cells) and stops
// at its next in-range pair; the wave's 64 current pairs are two 32-row MFMA
// tiles (rows = lanes 0..31 and lanes 32..63).  A lane without a pair left
// contributes a zero feature row, whose messages are relu(0) = 0 because the
// bias is a feature.  The relu'd pre-activations accumulate per row in the C
// layout, which after the last pair holds H for the wave's 64 particles.
#include "nbs_device.hpp"
#include "nbs_surrogate.hpp"

#include <algorithm>
#include <cmath>

namespace nbs {

typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float floatx16 __attribute__((__vector_size__(16 * sizeof(float))));

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}
__device__ __forceinline__ bf16x4 words_to_bf4(uint32_t w0, uint32_t w1) {
  bf16x4 v;
  v[0] = (short)(w0 & 0xffffu);
  v[1] = (short)(w0 >> 16);
  v[2] = (short)(w1 & 0xffffu);
  v[3] = (short)(w1 >> 16);
  return v;
}

__device__ __forceinline__ floatx16 mfma(bf16x4 a, bf16x4 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ floatx16 zero16() {
  floatx16 z;
#pragma unroll
  for (int k = 0; k < 16; ++k) z[k] = 0.0f;
  return z;
}

// B operand of a row-major BF16 weight matrix W (rows = outputs, K inputs):
// B[k][n] = W[n0 + n][k0 + k].
__device__ __forceinline__ bf16x4 load_wb(const uint16_t *W, int K, int n0, int k0) {
  const int c = threadIdx.x & 31, g = (threadIdx.x >> 5) & 1;
  const uint2 w = *reinterpret_cast<const uint2 *>(W + (size_t)(n0 + c) * K + k0 + 4 * g);
  return words_to_bf4(w.x, w.y);
}

// C-layout row of register r for lane group g.
__device__ __forceinline__ int c_row(int r, int g) { return 8 * (r >> 2) + 4 * g + (r & 3); }

constexpr int kWaveRows = 64;    // particles per wave in the pair kernels
constexpr int kHbStride = 68;    // BF16 per row of the head operand buffer (padded)
constexpr int kGStride = kHid + 1;

// Per-particle inputs of the pair layer.
struct PartInfo {
  DevScene sc;
  int cx, cy;
  float2 xi;
  float vix, viy;
  float fr, fn;  // normalised density, near density
};

template <bool MULTI>
__device__ __forceinline__ PartInfo part_info(const Buffers &b, const DevParams &p, const DevScene &s0,
                                              const float *norm, int i) {
  PartInfo q;
  const uint32_t key = b.keys[i];
  q.sc = scene_of_slot<MULTI>(b, s0, key);
  const uint32_t local = key - q.sc.cellBase;
  q.cy = (int)(local / (uint32_t)q.sc.nx);
  q.cx = (int)local - q.cy * q.sc.nx;
  q.xi = b.x1[i];
  const float2 pv = b.prev[i];
  q.vix = (q.xi.x - pv.x) * p.invDt;
  q.viy = (q.xi.y - pv.y) * p.invDt;
  const float2 d = b.dens[i];
  q.fr = (d.x - norm[1]) * norm[2];
  q.fn = (d.y - norm[3]) * norm[4];
  return q;
}

// A lane's walk over its particle's candidates: the rows dy = -1..1 of the
// 3x3 cells, each a contiguous slot range (demo4.cpp:188-203 order).
struct Cursor {
  uint32_t j, jend;
  int d;
  bool done;
};

__device__ __forceinline__ bool cursor_next_row(Cursor &cu, const uint32_t *start, const PartInfo &q) {
  while (cu.d < 3) {
    const RowRange rr = row_range(start, q.sc, q.cx, q.cy, cu.d - 1);
    ++cu.d;
    if (rr.j0 < rr.j1) {
      cu.j = rr.j0;
      cu.jend = rr.j1;
      return true;
    }
  }
  return false;
}

// Advances to the lane's next pair (j != i, r^2 < h^2); false when exhausted.
__device__ __forceinline__ bool next_pair(Cursor &cu, const Buffers &b, const DevParams &p, const PartInfo &q, int i,
                                          uint32_t &jOut, float2 &xj, float &Rx, float &Ry, float &r2) {
  while (!cu.done) {
    if (cu.j >= cu.jend && !cursor_next_row(cu, b.startNext, q)) {
      cu.done = true;
      break;
    }
    const uint32_t j = cu.j++;
    xj = b.x1[j];
    Rx = xj.x - q.xi.x;
    Ry = xj.y - q.xi.y;
    r2 = Rx * Rx + Ry * Ry;
    if (r2 < p.h2 && j != (uint32_t)i) {
      jOut = j;
      return true;
    }
  }
  return false;
}

// The pair's 16 BF16 features as 8 words (features 2k, 2k+1 in word k).
__device__ __forceinline__ void pair_words(uint32_t w[8], const Buffers &b, const DevParams &p, const float *norm,
                                           const PartInfo &q, uint32_t j, float2 xj, float Rx, float Ry, float r2) {
  const float2 pj = b.prev[j];
  const float2 dj = b.dens[j];
  const float vjx = (xj.x - pj.x) * p.invDt, vjy = (xj.y - pj.y) * p.invDt;
  const float r = sqrt_rn(r2);
  const float qq = 1.0f - r * p.invH;
  w[0] = pack2(Rx * p.invH, Ry * p.invH);
  w[1] = pack2((vjx - q.vix) * norm[0], (vjy - q.viy) * norm[0]);
  w[2] = pack2(q.fr, q.fn);
  w[3] = pack2((dj.x - norm[1]) * norm[2], (dj.y - norm[3]) * norm[4]);
  w[4] = pack2(1.0f, qq);
  w[5] = pack2(qq * qq, 0.0f);
  w[6] = 0u;
  w[7] = 0u;
}

struct W1Operands {
  bf16x4 lo0, lo1, hi0, hi1;  // hidden 0..31 / 32..63, K steps 0 / 1
};

__device__ __forceinline__ W1Operands load_w1(const uint16_t *w1b) {
  return {load_wb(w1b, kFeat, 0, 0), load_wb(w1b, kFeat, 0, 8), load_wb(w1b, kFeat, 32, 0),
          load_wb(w1b, kFeat, 32, 8)};
}

// A operands of the two row tiles from each lane's own feature words: tile 0
// rows are lanes 0..31, tile 1 rows lanes 32..63; a lane of group g supplies
// features 8s + 4g .. 8s + 4g + 3, so half of the words cross the lane halves.
struct PairA {
  bf16x4 t0s0, t0s1, t1s0, t1s1;  // tile, K step
};

__device__ __forceinline__ PairA pair_operands(const uint32_t w[8]) {
  const int g = (threadIdx.x >> 5) & 1;
  // step 0: features 0..7 (words 0..3)
  const uint32_t s0 = g ? w[0] : w[2], s1 = g ? w[1] : w[3];
  const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 32, 64), r1 = (uint32_t)__shfl_xor((int)s1, 32, 64);
  // step 1: features 8..15 (words 4..7; words 6, 7 are zero)
  const uint32_t q0 = (uint32_t)__shfl_xor((int)w[4], 32, 64), q1 = (uint32_t)__shfl_xor((int)w[5], 32, 64);
  PairA a;
  a.t0s0 = g ? words_to_bf4(r0, r1) : words_to_bf4(w[0], w[1]);
  a.t1s0 = g ? words_to_bf4(w[2], w[3]) : words_to_bf4(r0, r1);
  a.t0s1 = g ? words_to_bf4(0u, 0u) : words_to_bf4(w[4], w[5]);
  a.t1s1 = g ? words_to_bf4(0u, 0u) : words_to_bf4(q0, q1);
  return a;
}

// Pre-activations z = F W1^T of one row tile (C layout, lane = hidden unit).
__device__ __forceinline__ void pair_layer(bf16x4 as0, bf16x4 as1, const W1Operands &w, floatx16 &zlo,
                                           floatx16 &zhi) {
  zlo = mfma(as1, w.lo1, mfma(as0, w.lo0, zero16()));
  zhi = mfma(as1, w.hi1, mfma(as0, w.hi0, zero16()));
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(64) void k_sur_forward(Buffers b, DevParams p, DevScene s0, int n, SurrogateBufs m,
                                                    int mode) {
  __shared__ uint16_t hb[kSurRows][kHbStride];    // H, then M, of one row tile as BF16 operand rows
  __shared__ float gbuf[kSurRows][kGStride];      // G (FP32) of one row tile
  __shared__ float shCnt[kWaveRows];
  const int lane = threadIdx.x, c = lane & 31, g = lane >> 5;
  const int base = (int)blockIdx.x * kWaveRows;
  const int me = base + lane;
  const float *prm = m.params;
  const float *norm = prm + kOffNorm;
  const W1Operands w1 = load_w1(m.w1b);

  // 1. pair layer: H_i = sum_j relu(W1 f_ij)
  PartInfo q{};
  Cursor cu{0u, 0u, 0, true};
  if (me < n) {
    q = part_info<MULTI>(b, p, s0, norm, me);
    cu.done = false;
  }
  floatx16 h00 = zero16(), h01 = zero16(), h10 = zero16(), h11 = zero16();  // tile, hidden half
  uint32_t np = 0;
  for (;;) {
    uint32_t j = 0;
    float2 xj = make_float2(0.0f, 0.0f);
    float Rx = 0.0f, Ry = 0.0f, r2 = 0.0f;
    const bool have = next_pair(cu, b, p, q, me, j, xj, Rx, Ry, r2);
    if (__ballot(have) == 0ull) break;  // wave-uniform
    uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (have) {
      pair_words(w, b, p, norm, q, j, xj, Rx, Ry, r2);
      ++np;
    }
    const PairA a = pair_operands(w);
    floatx16 zlo, zhi;
    pair_layer(a.t0s0, a.t0s1, w1, zlo, zhi);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      h00[k] += fmaxf(zlo[k], 0.0f);
      h01[k] += fmaxf(zhi[k], 0.0f);
    }
    pair_layer(a.t1s0, a.t1s1, w1, zlo, zhi);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      h10[k] += fmaxf(zlo[k], 0.0f);
      h11[k] += fmaxf(zhi[k], 0.0f);
    }
  }
  shCnt[lane] = (float)np;

  // 2. head, one 32-row tile at a time: C = A W^T with A in hb (BF16)
  auto dense = [&](const uint16_t *W, floatx16 &olo, floatx16 &ohi) {
    bf16x4 a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const uint16_t *src = &hb[c][8 * s + 4 * g];
      a[s] = words_to_bf4((uint32_t)src[0] | ((uint32_t)src[1] << 16), (uint32_t)src[2] | ((uint32_t)src[3] << 16));
    }
    olo = zero16();
    ohi = zero16();
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      olo = mfma(a[s], load_wb(W, kHid, 0, 8 * s), olo);
      ohi = mfma(a[s], load_wb(W, kHid, 32, 8 * s), ohi);
    }
  };
  const bool train = (mode == 2);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const floatx16 &hlo = t ? h10 : h00;
    const floatx16 &hhi = t ? h11 : h01;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = c_row(k, g);
      hb[row][c] = f2bf(hlo[k]);
      hb[row][32 + c] = f2bf(hhi[k]);
    }
    __syncthreads();
    const int i = base + 32 * t + c;  // A-operand row c of this tile
    if (train && g == 0 && i < n) {
      float *act = m.act + (size_t)i * kActStride;
      for (int k = 0; k < kHid; ++k) act[k] = bf2f(hb[c][k]);
      m.cnt[i] = shCnt[32 * t + c];
    }
    // M = W2 H + n b2, kept as BF16 in hb
    floatx16 olo, ohi;
    dense(m.w2b, olo, ohi);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = c_row(k, g);
      const float nc = shCnt[32 * t + row];
      hb[row][c] = f2bf(olo[k] + nc * prm[kOffB2 + c]);
      hb[row][32 + c] = f2bf(ohi[k] + nc * prm[kOffB2 + 32 + c]);
    }
    __syncthreads();
    if (train && g == 0 && i < n) {
      float *act = m.act + (size_t)i * kActStride + kHid;
      for (int k = 0; k < kHid; ++k) act[k] = bf2f(hb[c][k]);
    }
    // G = relu(W3 M + b3), FP32
    dense(m.w3b, olo, ohi);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = c_row(k, g);
      gbuf[row][c] = fmaxf(olo[k] + prm[kOffB3 + c], 0.0f);
      gbuf[row][32 + c] = fmaxf(ohi[k] + prm[kOffB3 + 32 + c], 0.0f);
    }
    __syncthreads();
    if (train && g == 0 && i < n) {
      float *act = m.act + (size_t)i * kActStride + 2 * kHid;
      for (int k = 0; k < kHid; ++k) act[k] = gbuf[c][k];
    }
    // 3. o = W4 G + b4 (FP32): lane (c, g) computes output g of particle i
    float o = prm[kOffB4 + g];
    for (int k = 0; k < kHid; ++k) o = o + prm[kOffW4 + g * kHid + k] * gbuf[c][k];
    const float oy = __shfl_down(o, 32, 64);
    if (g == 0 && i < n) {
      if (m.out) {
        m.out[2 * (size_t)i] = o;
        m.out[2 * (size_t)i + 1] = oy;
      }
      if (mode == 1) {
        const uint32_t key = b.keys[i];
        const DevScene sc = scene_of_slot<MULTI>(b, s0, key);
        const float2 xi = b.x1[i];
        float x = xi.x + o * norm[5];
        float y = xi.y + oy * norm[5];
        collide(b, p, sc, x, y);                                               // demo4.cpp:411-444
        const float2 pv = b.prev[i];
        b.x0[i] = make_float2(x, y);
        b.v[i] = make_float2((x - pv.x) * p.invDt, (y - pv.y) * p.invDt);      // demo4.cpp:446-450
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------
// Per particle (FP32): dO, dG (pre-activation), dM, dH and the squared error.
__global__ __launch_bounds__(kBlock) void k_sur_back_head(int n, SurrogateBufs m, float invN) {
  __shared__ float red[kBlock / 64];
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  float err = 0.0f;
  if (i < n) {
    const float *prm = m.params;
    const float os = prm[kOffNorm + 5];
    const float *G = m.act + (size_t)i * kActStride + 2 * kHid;
    const float2 t = m.target[i];
    const float e0 = m.out[2 * (size_t)i] - t.x / os;
    const float e1 = m.out[2 * (size_t)i + 1] - t.y / os;
    err = e0 * e0 + e1 * e1;
    const float d0 = 2.0f * invN * e0, d1 = 2.0f * invN * e1;
    m.dO[2 * (size_t)i] = d0;
    m.dO[2 * (size_t)i + 1] = d1;
    float gpre[kHid];
    for (int k = 0; k < kHid; ++k) {
      const float v = prm[kOffW4 + k] * d0 + prm[kOffW4 + kHid + k] * d1;
      gpre[k] = G[k] > 0.0f ? v : 0.0f;
      m.dG[(size_t)i * kHid + k] = gpre[k];
    }
    float dm[kHid];
    for (int k = 0; k < kHid; ++k) {
      float s = 0.0f;
      for (int r = 0; r < kHid; ++r) s += prm[kOffW3 + r * kHid + k] * gpre[r];
      dm[k] = s;
      m.dM[(size_t)i * kHid + k] = s;
    }
    for (int k = 0; k < kHid; ++k) {
      float s = 0.0f;
      for (int r = 0; r < kHid; ++r) s += prm[kOffW2 + r * kHid + k] * dm[r];
      m.dH[(size_t)i * kHid + k] = s;
    }
  }
  // squared error, summed per workgroup
  for (int off = 32; off > 0; off >>= 1) err += __shfl_xor(err, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = err;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    m.lossPart[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_loss_final(const float *part, int nb, float invN, float *loss) {
  __shared__ float red[kBlock];
  float s = 0.0f;
  for (int k = threadIdx.x; k < nb; k += kBlock) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * invN;
}

// part[blockIdx][a * nb + b] = sum over the workgroup's rows of A[i][a] B[i][b];
// part[blockIdx][na * nb + a] = sum of A[i][a] * scale_i (scale 1 if null).
// Workgroups stride over rows in tiles of 32.
__global__ __launch_bounds__(kBlock) void k_outer_partial(int n, const float *A, int lda, int na, const float *B,
                                                          int ldb, int nb, const float *scale, float *part) {
  __shared__ float sa[32][kHid], sb[32][kHid], ss[32];
  const int t = threadIdx.x;
  const int a0 = 4 * (t / 16), b0 = 4 * (t % 16);
  float acc[4][4] = {};
  float accb[4] = {};
  for (int base = (int)blockIdx.x * 32; base < n; base += (int)gridDim.x * 32) {
    for (int e = t; e < 32 * kHid; e += kBlock) {
      const int rr = e / kHid, k = e % kHid, i = base + rr;
      sa[rr][k] = (i < n && k < na) ? A[(size_t)i * lda + k] : 0.0f;
      sb[rr][k] = (i < n && k < nb) ? B[(size_t)i * ldb + k] : 0.0f;
      if (k == 0) ss[rr] = (i < n) ? (scale ? scale[i] : 1.0f) : 0.0f;
    }
    __syncthreads();
    for (int rr = 0; rr < 32; ++rr) {
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const float av = sa[rr][a0 + x];
        accb[x] += av * ss[rr];
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] += av * sb[rr][b0 + y];
      }
    }
    __syncthreads();
  }
  float *out = part + (size_t)blockIdx.x * (size_t)(na * nb + na);
  for (int x = 0; x < 4; ++x) {
    if (a0 + x >= na) continue;
    for (int y = 0; y < 4; ++y)
      if (b0 + y < nb) out[(a0 + x) * nb + b0 + y] = acc[x][y];
    if (b0 == 0) out[na * nb + a0 + x] = accb[x];
  }
}

// out[e] = sum over parts q of part[q * width + e] (fixed order: deterministic).
__global__ __launch_bounds__(kBlock) void k_sum_parts(const float *part, int nparts, int width, float *out) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e >= width) return;
  float s = 0.0f;
  for (int q = 0; q < nparts; ++q) s += part[(size_t)q * width + e];
  out[e] = s;
}

// Copies a reduced [na * nb | na] block into the weight and bias slots.
__global__ __launch_bounds__(kBlock) void k_scatter_grad(const float *sum, int na, int nb, float *w, float *bias) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e < na * nb) w[e] = sum[e];
  else if (e < na * nb + na) bias[e - na * nb] = sum[e];
}

// dW1 = sum_i sum_j (dH_i . [z_ij > 0]) f_ij^T, with the pair walk of the
// forward pass.  For row tile t the pair layer's accumulator (C layout: lane =
// hidden unit, register 4k+u = tile row 8k+4g+u), masked and scaled by the
// rows' dH, is the A operand (hidden x 8 rows) of the k-th 8-row block; the
// rows' features, staged in LDS, are the B operand (8 rows x 32 features, 16
// used).  Each wave accumulates over its particles and writes one 64 x 16
// partial.
template <bool MULTI>
__global__ __launch_bounds__(64) void k_sur_back_pairs(Buffers b, DevParams p, DevScene s0, int n, SurrogateBufs m) {
  __shared__ uint32_t fw[kWaveRows][kFeat / 2];
  const int lane = threadIdx.x, c = lane & 31, g = lane >> 5;
  const float *norm = m.params + kOffNorm;
  const W1Operands w1 = load_w1(m.w1b);
  floatx16 glo = zero16(), ghi = zero16();
  for (int base = (int)blockIdx.x * kWaveRows; base < n; base += (int)gridDim.x * kWaveRows) {
    const int me = base + lane;
    // dH of the rows this lane's accumulator registers hold
    floatx16 d00, d01, d10, d11;  // tile, hidden half
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int r0 = base + c_row(k, g), r1 = base + 32 + c_row(k, g);
      d00[k] = r0 < n ? m.dH[(size_t)r0 * kHid + c] : 0.0f;
      d01[k] = r0 < n ? m.dH[(size_t)r0 * kHid + 32 + c] : 0.0f;
      d10[k] = r1 < n ? m.dH[(size_t)r1 * kHid + c] : 0.0f;
      d11[k] = r1 < n ? m.dH[(size_t)r1 * kHid + 32 + c] : 0.0f;
    }
    PartInfo q{};
    Cursor cu{0u, 0u, 0, true};
    if (me < n) {
      q = part_info<MULTI>(b, p, s0, norm, me);
      cu.done = false;
    }
    for (;;) {
      uint32_t j = 0;
      float2 xj = make_float2(0.0f, 0.0f);
      float Rx = 0.0f, Ry = 0.0f, r2 = 0.0f;
      const bool have = next_pair(cu, b, p, q, me, j, xj, Rx, Ry, r2);
      if (__ballot(have) == 0ull) break;  // wave-uniform
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      if (have) pair_words(w, b, p, norm, q, j, xj, Rx, Ry, r2);
#pragma unroll
      for (int k = 0; k < 8; ++k) fw[lane][k] = w[k];
      const PairA a = pair_operands(w);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        floatx16 zlo, zhi;
        pair_layer(t ? a.t1s0 : a.t0s0, t ? a.t1s1 : a.t0s1, w1, zlo, zhi);
        const floatx16 &dlo = t ? d10 : d00;
        const floatx16 &dhi = t ? d11 : d01;
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
          bf16x4 bf, alo, ahi;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t wd = fw[32 * t + 8 * blk + 4 * g + u][(c & 15) >> 1];
            const uint16_t hv = (c & 1) ? (uint16_t)(wd >> 16) : (uint16_t)(wd & 0xffffu);
            bf[u] = (short)(c < kFeat ? hv : 0);
            alo[u] = (short)f2bf(zlo[4 * blk + u] > 0.0f ? dlo[4 * blk + u] : 0.0f);
            ahi[u] = (short)f2bf(zhi[4 * blk + u] > 0.0f ? dhi[4 * blk + u] : 0.0f);
          }
          glo = mfma(alo, bf, glo);
          ghi = mfma(ahi, bf, ghi);
        }
      }
      __syncthreads();
    }
  }
  float *out = m.part + (size_t)blockIdx.x * (kHid * kFeat);
  if (c < kFeat) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int hid = c_row(k, g);
      out[hid * kFeat + c] = glo[k];
      out[(32 + hid) * kFeat + c] = ghi[k];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_bf16_weights(const float *params, uint16_t *w1b, uint16_t *w2b,
                                                         uint16_t *w3b) {
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (t < kHid * kFeat) w1b[t] = f2bf(params[kOffW1 + t]);
  if (t < kHid * kHid) {
    w2b[t] = f2bf(params[kOffW2 + t]);
    w3b[t] = f2bf(params[kOffW3 + t]);
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; c1, c2 are the bias
// corrections 1 - b1^t, 1 - b2^t.
__global__ __launch_bounds__(kBlock) void k_adam(float *params, const float *grad, float *mo, float *ve, int count,
                                                 float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (t >= count) return;
  const float gr = grad[t];
  const float m1 = b1 * mo[t] + (1.0f - b1) * gr;
  const float v1 = b2 * ve[t] + (1.0f - b2) * gr * gr;
  mo[t] = m1;
  ve[t] = v1;
  params[t] = params[t] - lr * (m1 / c1) / (sqrtf(v1 / c2) + eps);
}

// FP64 sums for the density check: sum |a.x - b.x| and sum b.x.
constexpr int kErrItems = 8;
__global__ __launch_bounds__(kBlock) void k_dens_err_partial(const float2 *a, const float2 *bb, int n,
                                                             double *partials) {
  __shared__ double red[2][kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  const int base = (int)blockIdx.x * kBlock * kErrItems;
  for (int k = 0; k < kErrItems; ++k) {
    const int i = base + k * kBlock + (int)threadIdx.x;
    if (i >= n) break;
    const double x = a[i].x, y = bb[i].x;
    s0 += fabs(x - y);
    s1 += y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s0;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t0 += red[0][w], t1 += red[1][w];
    partials[2 * blockIdx.x] = t0;
    partials[2 * blockIdx.x + 1] = t1;
  }
}

__global__ void k_dens_err_final(const double *partials, int nb, double *out) {
  if (threadIdx.x != 0) return;
  double t0 = 0.0, t1 = 0.0;
  for (int k = 0; k < nb; ++k) t0 += partials[2 * k], t1 += partials[2 * k + 1];
  out[0] = t0;
  out[1] = t1;
}

// MFMA layout self-test (see the file header).
__global__ __launch_bounds__(64) void k_mfma_probe(const float *A, const float *B, float *C) {
  const int lane = threadIdx.x & 63, c = lane & 31, g = lane >> 5;
  bf16x4 a, bb;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a[t] = (short)f2bf(A[c * 8 + 4 * g + t]);
    bb[t] = (short)f2bf(B[(4 * g + t) * 32 + c]);
  }
  const floatx16 r = mfma(a, bb, zero16());
#pragma unroll
  for (int k = 0; k < 16; ++k) C[c_row(k, g) * 32 + c] = r[k];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static unsigned waves_for(int n) { return (unsigned)((n + kWaveRows - 1) / kWaveRows); }
static unsigned back_pair_waves(int n) { return std::min(waves_for(n), 4096u); }
static unsigned outer_blocks(int n) { return (unsigned)std::min((n + kOuterRows - 1) / kOuterRows, 1024); }

size_t surrogate_part_floats(int n) {
  const size_t a = (size_t)back_pair_waves(n) * kHid * kFeat;
  const size_t o = (size_t)outer_blocks(n) * (kHid * kHid + kHid);
  return std::max<size_t>(std::max(a, o), 1) + (kHid * kHid + kHid);  // + the reduced block
}

int surrogate_head_blocks(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

void launch_surrogate_forward(const Buffers &b, const DevParams &p, const DevScene &s0, int n,
                              const SurrogateBufs &m, int mode, hipStream_t s) {
  if (n <= 0) return;
  if (b.nScenes > 1)
    hipLaunchKernelGGL(k_sur_forward<true>, dim3(waves_for(n)), dim3(64), 0, s, b, p, s0, n, m, mode);
  else
    hipLaunchKernelGGL(k_sur_forward<false>, dim3(waves_for(n)), dim3(64), 0, s, b, p, s0, n, m, mode);
}

// One outer-product gradient block: partials, then the fixed-order sum into
// the weight (na x nb) and bias (na) slots.
static void outer_grad(int n, const float *A, int lda, int na, const float *B, int ldb, int nb, const float *scale,
                       float *part, float *w, float *bias, hipStream_t s) {
  const unsigned blocks = outer_blocks(n);
  const int width = na * nb + na;
  float *sum = part + (size_t)blocks * width;
  hipLaunchKernelGGL(k_outer_partial, dim3(blocks), dim3(kBlock), 0, s, n, A, lda, na, B, ldb, nb, scale, part);
  hipLaunchKernelGGL(k_sum_parts, dim3((width + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (const float *)part,
                     (int)blocks, width, sum);
  hipLaunchKernelGGL(k_scatter_grad, dim3((width + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (const float *)sum, na,
                     nb, w, bias);
}

void launch_surrogate_backward(const Buffers &b, const DevParams &p, const DevScene &s0, int n,
                               const SurrogateBufs &m, hipStream_t s) {
  if (n <= 0) return;
  const int hb = surrogate_head_blocks(n);
  const float invN = 1.0f / (float)n;
  hipLaunchKernelGGL(k_sur_back_head, dim3(hb), dim3(kBlock), 0, s, n, m, invN);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kBlock), 0, s, (const float *)m.lossPart, hb, invN,
                     m.loss + m.lossIndex);
  // dW4 = sum dO G^T, db4 = sum dO
  outer_grad(n, m.dO, 2, 2, m.act + 2 * kHid, kActStride, kHid, nullptr, m.part, m.grad + kOffW4, m.grad + kOffB4, s);
  // dW3 = sum dG M^T, db3 = sum dG
  outer_grad(n, m.dG, kHid, kHid, m.act + kHid, kActStride, kHid, nullptr, m.part, m.grad + kOffW3, m.grad + kOffB3,
             s);
  // dW2 = sum dM H^T, db2 = sum n_i dM
  outer_grad(n, m.dM, kHid, kHid, m.act, kActStride, kHid, m.cnt, m.part, m.grad + kOffW2, m.grad + kOffB2, s);
  // dW1
  const unsigned waves = back_pair_waves(n);
  if (b.nScenes > 1)
    hipLaunchKernelGGL(k_sur_back_pairs<true>, dim3(waves), dim3(64), 0, s, b, p, s0, n, m);
  else
    hipLaunchKernelGGL(k_sur_back_pairs<false>, dim3(waves), dim3(64), 0, s, b, p, s0, n, m);
  hipLaunchKernelGGL(k_sum_parts, dim3((kHid * kFeat + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     (const float *)m.part, (int)waves, kHid * kFeat, m.grad + kOffW1);
}

void launch_bf16_weights(const float *params, uint16_t *w1b, uint16_t *w2b, uint16_t *w3b, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_weights, dim3((kHid * kHid + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, w1b, w2b,
                     w3b);
}

void launch_adam(float *params, const float *grad, float *mo, float *ve, int count, float lr, float b1, float b2,
                 float eps, int t, hipStream_t s) {
  const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, grad, mo, ve, count, lr,
                     b1, b2, eps, c1, c2);
}

void launch_density_error(const float2 *a, const float2 *bb, int n, double *partials, double *out, hipStream_t s) {
  const int nb = std::max(1, (n + kBlock * kErrItems - 1) / (kBlock * kErrItems));
  if (n > 0) hipLaunchKernelGGL(k_dens_err_partial, dim3(nb), dim3(kBlock), 0, s, a, bb, n, partials);
  hipLaunchKernelGGL(k_dens_err_final, dim3(1), dim3(64), 0, s, (const double *)partials, n > 0 ? nb : 0, out);
}

void launch_mfma_probe(const float *a, const float *b, float *c, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, a, b, c);
}

}  // namespace nbs
