// Learned relaxation kernels (BF16 MFMA, FP32 accumulation) for gfx950.
//
// v_mfma_f32_32x32x16_bf16 operand layouts (lane l, g = l / 32, c = l % 32):
//   A (32 x 16): lane holds A[c][8g + t], t = 0..7
//   B (16 x 32): lane holds B[8g + t][c]
//   C (32 x 32): register r holds C[8(r/4) + 4g + (r%4)][c]
// (v_mfma_f32_32x32x8_bf16, the layout self-test's other form: A[c][4g + t],
// B[4g + t][c], t = 0..3, the same C.)  The MFMA self-tests check both
// layouts against host matrix products.
//
// Pair layer: one wave serves 64 particles, one per lane.  Each lane walks its
// own particle's candidates (the three row ranges of its 3x3 cells) and stops
// at its next in-range pair; the wave's 64 current pairs are two 32-row MFMA
// tiles (rows = lanes 0..31 and lanes 32..63), one K = 16 MFMA per tile and
// hidden half gives the pre-activations z.  A lane without a pair left
// contributes a zero feature row, whose messages are relu(0) = 0 because the
// bias is a feature.  relu(z) is taken on BF16: two values are converted with
// one v_cvt_pk_bf16_f32 and clamped at 0 with one v_pk_max_i16 (a negative
// BF16 is a negative 16-bit integer).  The per-particle sum H_i += relu(z_ij)
// runs on the matrix core: in the C layout, registers 8s .. 8s + 7 of a lane
// hold exactly the B operand (K = 16 rows of z, the lane's hidden column) of
// a K = 16 MFMA, and a 0/1 selection matrix as A sends row k of z to its
// particle's row of H; with one nonzero product per output, the MFMA adds
// each BF16 value to the FP32 accumulator exactly as an FP32 add would.
#include "nbs_device.hpp"
#include "nbs_surrogate.hpp"

#include <algorithm>
#include <cmath>

namespace nbs {

typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float floatx16 __attribute__((__vector_size__(16 * sizeof(float))));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even: one v_cvt_pk_bf16_f32 for both values
// (converting the elements one at a time costs two conversions and a permute
// when the inputs are MFMA results).
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  const f32x2 f = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
}
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
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

// relu(z) as an integer maximum of the bit pattern with 0: negative values
// (and -0) become +0, others are unchanged.  One instruction; fmaxf costs two
// here (the IEEE-mode canonicalisation of its operand).
__device__ __forceinline__ float relu(float z) { return __int_as_float(max(__float_as_int(z), 0)); }

// K = 16 operands: eight BF16 values in four words.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x8 bf8(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
  const u32x4 u = {w0, w1, w2, w3};
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ floatx16 mfma16(bf16x8 a, bf16x8 b, floatx16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// relu of two values rounded to BF16 (round to nearest even), packed: one
// v_cvt_pk_bf16_f32, then one v_pk_max_i16 with 0 (negative values and -0 are
// negative 16-bit integers and become +0).
__device__ __forceinline__ uint32_t relu2_bf16(float lo, float hi) {
  const s16x2 v = __builtin_bit_cast(s16x2, pack2(lo, hi));
  const s16x2 z = {0, 0};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, z));
}

// K = 16 B operand of a row-major BF16 matrix W (rows = outputs, K inputs):
// B[k][n] = W[n0 + n][k0 + k].  K and k0 are multiples of 8 (16-byte loads).
__device__ __forceinline__ bf16x8 load_wb16(const uint16_t *W, int K, int n0, int k0) {
  const int c = threadIdx.x & 31, g = (threadIdx.x >> 5) & 1;
  const uint4 w = *reinterpret_cast<const uint4 *>(W + (size_t)(n0 + c) * K + k0 + 8 * g);
  return bf8(w.x, w.y, w.z, w.w);
}

// Eight BF16 values at p (a row of an LDS buffer, 4-byte aligned).
__device__ __forceinline__ bf16x8 lds_bf8(const uint16_t *p) {
  const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
  return bf8(w[0], w[1], w[2], w[3]);
}

// B operand of step s (K = 16 rows of a 32-row C-layout tile z): this lane's
// registers 8s .. 8s + 7, relu'd in BF16.
__device__ __forceinline__ bf16x8 relu_rows(const floatx16 &z, int s) {
  return bf8(relu2_bf16(z[8 * s + 0], z[8 * s + 1]), relu2_bf16(z[8 * s + 2], z[8 * s + 3]),
             relu2_bf16(z[8 * s + 4], z[8 * s + 5]), relu2_bf16(z[8 * s + 6], z[8 * s + 7]));
}

// C-layout row of register r for lane group g.
__device__ __forceinline__ int c_row(int r, int g) { return 8 * (r >> 2) + 4 * g + (r & 3); }

// A operand of step s of the pair sum: 32 x 16, A[p][k] = 1 when the K index k
// (K = 16 rows of z as relu_rows supplies them: lane group g', register
// 8s + t for k = 8g' + t) is row p of the tile.  Lane (c, g) holds row c.
__device__ __forceinline__ bf16x8 selection(int s) {
  const int c = threadIdx.x & 31, g = (threadIdx.x >> 5) & 1;
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t lo = c_row(8 * s + 2 * q, g) == c ? 0x3F80u : 0u;  // BF16 1.0
    const uint32_t hi = c_row(8 * s + 2 * q + 1, g) == c ? 0x3F80u : 0u;
    w[q] = lo | (hi << 16);
  }
  return bf8(w[0], w[1], w[2], w[3]);
}

constexpr int kWaveRows = 64;    // particles per wave in the pair kernels
constexpr int kHbStride = 68;    // BF16 per row of the head operand buffer (padded)
constexpr int kGStride = kHid + 1;
constexpr int kHbBytes = kSurRows * kHbStride * 2;
constexpr int kHeadBytes = kHbBytes + kSurRows * kGStride * 4;
constexpr int kTileCap = 512;    // staged candidates (positions, previous positions, densities)
constexpr int kStageBytes = 3 * kTileCap * 8;
constexpr int kPoolBytes = kHeadBytes > kStageBytes ? kHeadBytes : kStageBytes;

// Per-particle inputs of the pair layer.
struct PartInfo {
  DevScene sc;
  int cx, cy;
  float2 xi;
  float vix, viy;
  float fr, fn;  // normalised density, near density
};

// A particle's own loads for part_info, requested before the tile is staged
// (the staging's waits then cover them) and turned into PartInfo after.
struct PartRaw {
  uint32_t key;
  float2 x, pv, d;
};
__device__ __forceinline__ PartRaw part_raw(const Buffers &b, int i) { return {b.keys[i], b.x1[i], b.prev[i], b.dens[i]}; }

template <bool MULTI>
__device__ __forceinline__ PartInfo part_info(const Buffers &b, const DevParams &p, const DevScene &s0,
                                              const float *norm, const PartRaw &r) {
  PartInfo q;
  q.sc = scene_of_slot<MULTI>(b, s0, r.key);
  const uint32_t local = r.key - q.sc.cellBase;
  q.cy = (int)(local / (uint32_t)q.sc.nx);
  q.cx = (int)local - q.cy * q.sc.nx;
  q.xi = r.x;
  q.vix = (q.xi.x - r.pv.x) * p.invDt;
  q.viy = (q.xi.y - r.pv.y) * p.invDt;
  q.fr = (r.d.x - norm[1]) * norm[2];
  q.fn = (r.d.y - norm[3]) * norm[4];
  return q;
}

// A lane's walk over its particle's candidates: the rows dy = -1..1 of the
// 3x3 cells, each a contiguous slot range (demo4.cpp:188-203 order).  tb maps
// a slot of the current row to its index in the neighbour source.
struct Cursor {
  uint32_t j, jend, tb;
  int d;
  bool done;
  uint32_t mask, mbase;  // in-range candidates of the current chunk: bit b is slot mbase + b
  // rows tested up front (cursor_rows): the masks and source bases of the
  // rows after the current one
  uint32_t m1, b1, m2, b2;
};

// Neighbour data of a wave's candidates: an LDS copy of the three candidate
// slot ranges of the wave's 64 particles (the layout of TilePlan) when they
// fit, else the slot arrays.

struct PairSrc {
  const float2 *x, *pv, *d;
  uint32_t rowStart0, rowStart1, rowStart2, off0, off1, off2;
  bool tiled;
  // slot j of candidate row r sits at j + base(r)
  __device__ __forceinline__ uint32_t base(int r) const {
    if (!tiled) return 0u;
    return r == 0 ? off0 - rowStart0 : (r == 1 ? off1 - rowStart1 : off2 - rowStart2);
  }
};

__device__ __forceinline__ bool cursor_next_row(Cursor &cu, const uint32_t *start, const PartInfo &q,
                                                const PairSrc &src) {
  while (cu.d < 3) {
    const RowRange rr = row_range(start, q.sc, q.cx, q.cy, cu.d - 1);
    const int r = cu.d++;
    if (rr.j0 < rr.j1) {
      cu.j = rr.j0;
      cu.jend = rr.j1;
      cu.tb = src.base(r);
      return true;
    }
  }
  return false;
}

// Plans the wave's tile and stages it when it fits (every lane must call it).
template <bool MULTI>
__device__ __forceinline__ PairSrc stage_pairs(const Buffers &b, const DevScene &s0, int base, int n,
                                               unsigned char *pool, TilePlan &plan) {
  if (threadIdx.x == 0) {
    const uint32_t kA = b.keys[base], kB = b.keys[min(base + kWaveRows, n) - 1];
    const DevScene sa = scene_of_slot<MULTI>(b, s0, kA);
    plan_tile<kTileCap>(b.startNext, sa, kA, kB, plan);
    if (MULTI && sa.cellBase != scene_of_slot<MULTI>(b, s0, kB).cellBase) plan.use = 0;
  }
  __syncthreads();
  PairSrc src;
  src.tiled = plan.use != 0;
  src.rowStart0 = plan.rowStart[0], src.rowStart1 = plan.rowStart[1], src.rowStart2 = plan.rowStart[2];
  src.off0 = plan.off[0], src.off1 = plan.off[1], src.off2 = plan.off[2];
  if (src.tiled) {
    float2 *tx = reinterpret_cast<float2 *>(pool);
    float2 *tp = tx + kTileCap, *td = tp + kTileCap;
    const float2 *const from[3] = {b.x1, b.prev, b.dens};
    float2 *const to[3] = {tx, tp, td};
    stage_arrays<3>(plan, from, to);  // the three arrays' loads in flight together
    src.x = tx;
    src.pv = tp;
    src.d = td;
  } else {
    src.x = b.x1;
    src.pv = b.prev;
    src.d = b.dens;
  }
  __syncthreads();
  return src;
}

// The source with its pointers re-derived from the LDS pool (tiled) or the
// slot arrays, for code that handles one kind only.
__device__ __forceinline__ PairSrc tile_src(PairSrc s, unsigned char *pool) {
  s.x = reinterpret_cast<const float2 *>(pool);
  s.pv = s.x + kTileCap;
  s.d = s.pv + kTileCap;
  return s;
}
__device__ __forceinline__ PairSrc slot_src(PairSrc s, const Buffers &b) {
  s.x = b.x1;
  s.pv = b.prev;
  s.d = b.dens;
  return s;
}

// Cut-off mask of the cnt <= 32 candidates at xs[0 .. cnt): bit k is set when
// candidate k lies within h of xi.
__device__ __forceinline__ uint32_t chunk_bits(const float2 *xs, uint32_t cnt, float2 xi, float h2) {
  // m = 2 * m + (d2 < h2), last candidate first, so candidate k ends at bit
  // k: the compare's lane mask is the add's carry-in
  uint32_t m = 0;
  auto push = [&](float d2) {
    asm("v_cmp_gt_f32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc" : "+v"(m) : "s"(h2), "v"(d2) : "vcc");
  };
  auto dist2 = [&](uint32_t k) {
    const float2 x = xs[k];
    const float dx = x.x - xi.x, dy = x.y - xi.y;
    return dx * dx + dy * dy;
  };
  uint32_t u = cnt;
  for (; u >= 8u; u -= 8u) {  // eight position loads in flight
    float d2[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) d2[k] = dist2(u - 1 - k);
#pragma unroll
    for (int k = 0; k < 8; ++k) push(d2[k]);
  }
  for (; u >= 4u; u -= 4u) {
    const float a0 = dist2(u - 1), a1 = dist2(u - 2), a2 = dist2(u - 3), a3 = dist2(u - 4);
    push(a0);
    push(a1);
    push(a2);
    push(a3);
  }
  for (; u > 0u; --u) push(dist2(u - 1));
  return m;
}

// Rows of at most kWalkRowMax candidates can be tested up front: when every
// lane's three rows fit (rows_fit, wave-uniform), cursor_rows tests all of
// them with the wave's lanes together and next_pair then only moves from one
// row's mask to the next.  Otherwise a lane that runs out of marked
// candidates tests its next chunk of kWalkChunk while the rest of the wave
// waits, once per chunk of each lane.
constexpr uint32_t kWalkChunk = 16;
constexpr uint32_t kWalkRowMax = 32;

// rows: the lane's three candidate rows (row_ranges; ignored for a lane past
// n).
__device__ __forceinline__ bool rows_fit(const RowRange (&rows)[3], bool live) {
  bool fit = true;
  if (live)
    for (int d = 0; d < 3; ++d) fit = fit && rows[d].j1 - rows[d].j0 <= kWalkRowMax;
  return __ballot(!fit) == 0ull;
}

// Tests the lane's three rows (every row at most kWalkRowMax candidates) and
// loads their masks into the cursor; the particle itself is no pair.
__device__ __forceinline__ void cursor_rows(Cursor &cu, const RowRange (&rows)[3], const DevParams &p,
                                            const PairSrc &src, const PartInfo &q, int i) {
  uint32_t m[3], base[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const RowRange rr = rows[r];
    const uint32_t cnt = rr.j1 - rr.j0;
    // the row's first index in the source: j0 + base(r) wraps modulo 2^32
    // (off - rowStart on the LDS tile), so it is formed in 32 bits first
    base[r] = (uint32_t)(rr.j0 + src.base(r));
    m[r] = chunk_bits(src.x + base[r], cnt, q.xi, p.h2);
    const uint32_t self = (uint32_t)i - rr.j0;
    if (self < cnt) m[r] &= ~(1u << self);
  }
  cu.mask = m[0], cu.mbase = base[0], cu.tb = 0u;
  cu.m1 = m[1], cu.b1 = base[1], cu.m2 = m[2], cu.b2 = base[2];
}

// Advances to the lane's next pair (j != i, r^2 < h^2); false when exhausted.
// t is the pair's index in src.  Candidates are tested kWalkChunk at a time
// into a mask, whose bits are then taken lowest first (candidate order); with
// rows (the cursor loaded by cursor_rows) the masks of all three rows are
// there already.
__device__ __forceinline__ bool next_pair(Cursor &cu, const Buffers &b, const DevParams &p, const PairSrc &src,
                                          const PartInfo &q, int i, bool rows, uint32_t &t, float2 &xj, float &Rx,
                                          float &Ry, float &r2) {
  for (;;) {
    if (cu.mask) {
      const uint32_t bit = (uint32_t)__builtin_ctz(cu.mask);
      cu.mask &= cu.mask - 1u;
      t = cu.mbase + bit + cu.tb;
      xj = src.x[t];
      Rx = xj.x - q.xi.x;
      Ry = xj.y - q.xi.y;
      r2 = Rx * Rx + Ry * Ry;
      return true;
    }
    if (cu.done) return false;
    if (rows) {
      if ((cu.m1 | cu.m2) == 0u) {
        cu.done = true;
        return false;
      }
      cu.mask = cu.m1, cu.mbase = cu.b1;
      cu.m1 = cu.m2, cu.b1 = cu.b2;
      cu.m2 = 0u;
      continue;
    }
    if (cu.j >= cu.jend && !cursor_next_row(cu, b.startNext, q, src)) {
      cu.done = true;
      return false;
    }
    const uint32_t cnt = min(kWalkChunk, cu.jend - cu.j);
    // the chunk's first index in the source: j + tb wraps modulo 2^32 (tb is
    // off - rowStart on the LDS tile), so it is formed in 32 bits first
    uint32_t m = chunk_bits(src.x + (uint32_t)(cu.j + cu.tb), cnt, q.xi, p.h2);
    const uint32_t self = (uint32_t)i - cu.j;  // the particle itself is no pair
    if (self < cnt) m &= ~(1u << self);
    cu.mbase = cu.j;
    cu.mask = m;
    cu.j += cnt;
  }
}

// The pair's 16 BF16 features as 8 words (features 2k, 2k+1 in word k).
__device__ __forceinline__ void pair_words(uint32_t w[8], const PairSrc &src, const DevParams &p, const float *norm,
                                           const PartInfo &q, uint32_t t, float2 xj, float Rx, float Ry, float r2) {
  const float2 pj = src.pv[t];
  const float2 dj = src.d[t];
  const float vjx = (xj.x - pj.x) * p.invDt, vjy = (xj.y - pj.y) * p.invDt;
  const float r = pair_sqrt(r2);
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
  bf16x8 lo, hi;  // hidden 0..31 / 32..63
};

__device__ __forceinline__ W1Operands load_w1(const uint16_t *w1b) {
  return {load_wb16(w1b, kFeat, 0, 0), load_wb16(w1b, kFeat, 32, 0)};
}

// A operands (K = 16) of the two row tiles from each lane's own feature words:
// tile 0 rows are lanes 0..31, tile 1 rows lanes 32..63; lane (c, g) supplies
// features 8g .. 8g + 7 of its row, so one half of each tile's words crosses
// the lane halves (words 6, 7 are zero).
struct PairA {
  bf16x8 t0, t1;
};

__device__ __forceinline__ PairA pair_operands(const uint32_t w[8]) {
  const int g = (threadIdx.x >> 5) & 1;
  const uint32_t s0 = g ? w[0] : w[4], s1 = g ? w[1] : w[5], s2 = g ? w[2] : 0u, s3 = g ? w[3] : 0u;
  const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 32, 64), r1 = (uint32_t)__shfl_xor((int)s1, 32, 64);
  const uint32_t r2 = (uint32_t)__shfl_xor((int)s2, 32, 64), r3 = (uint32_t)__shfl_xor((int)s3, 32, 64);
  PairA a;
  a.t0 = g ? bf8(r0, r1, r2, r3) : bf8(w[0], w[1], w[2], w[3]);
  a.t1 = g ? bf8(w[4], w[5], 0u, 0u) : bf8(r0, r1, r2, r3);
  return a;
}

// Pre-activations z = F W1^T of one row tile (C layout, lane = hidden unit).
__device__ __forceinline__ void pair_layer(bf16x8 a, const W1Operands &w, floatx16 &zlo, floatx16 &zhi) {
  zlo = mfma16(a, w.lo, zero16());
  zhi = mfma16(a, w.hi, zero16());
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
// Coalesced copy of a 32-row head tile into one block of the training
// activations (rows row0 .. row0 + 31, columns col .. col + 63 of the
// kActStride-wide rows): 16 lanes store one row's 64 values as float4, 256
// contiguous bytes, so one store instruction covers four rows.
template <typename Load4>
__device__ __forceinline__ void store_act_rows(float *act, int col, int row0, int n, Load4 &&load4) {
  for (int q = (int)threadIdx.x; q < 32 * 16; q += 64) {
    const int r = q >> 4, c4 = (q & 15) * 4;
    if (row0 + r < n) *reinterpret_cast<float4 *>(act + (size_t)(row0 + r) * kActStride + col + c4) = load4(r, c4);
  }
}

template <bool MULTI>
__global__ __launch_bounds__(64, 3) void k_sur_forward(Buffers b, DevParams p, DevScene s0, int n, SurrogateBufs m,
                                                    int mode) {
  // LDS: the staged candidates during the pair layer, then the head's operand
  // rows (hb: H, then M, of one row tile as BF16) and G (FP32)
  __shared__ __align__(16) unsigned char pool[kPoolBytes];
  __shared__ TilePlan plan;
  __shared__ float shCnt[kWaveRows];
  uint16_t(*hb)[kHbStride] = reinterpret_cast<uint16_t(*)[kHbStride]>(pool);
  float(*gbuf)[kGStride] = reinterpret_cast<float(*)[kGStride]>(pool + kHbBytes);
  const int lane = threadIdx.x, c = lane & 31, g = lane >> 5;
  const int base = (int)blockIdx.x * kWaveRows;
  const int me = base + lane;
  const float *prm = m.params;
  const float *norm = prm + kOffNorm;
  const W1Operands w1 = load_w1(m.w1b);
  const bf16x8 sel[2] = {selection(0), selection(1)};
  const PartRaw raw = part_raw(b, min(me, n - 1));  // a lane past n reads the last particle, unused
  const PairSrc src = stage_pairs<MULTI>(b, s0, base, n, pool, plan);

  // 1. pair layer: H_i = sum_j relu(W1 f_ij), each term rounded to BF16
  PartInfo q{};
  Cursor cu{0u, 0u, 0u, 0, true, 0u, 0u};
  if (me < n) {
    q = part_info<MULTI>(b, p, s0, norm, raw);
    cu.done = false;
  }
  floatx16 h00 = zero16(), h01 = zero16(), h10 = zero16(), h11 = zero16();  // tile, hidden half
  uint32_t np = 0;
  auto pair_pass = [&](const PairSrc &sp) {
    // rows up front except for the multi-scene training forward, where they
    // measured slower
    bool rows = false;
    if (!(MULTI && (mode & kFwdTrain))) {
      RowRange rr[3] = {};
      if (me < n) row_ranges(b.startNext, q.sc, q.cx, q.cy, rr);  // the six cell starts in flight together
      rows = rows_fit(rr, me < n);
      if (rows && me < n) cursor_rows(cu, rr, p, sp, q, me);
    }
    for (;;) {
      uint32_t t = 0;
      float2 xj = make_float2(0.0f, 0.0f);
      float Rx = 0.0f, Ry = 0.0f, r2 = 0.0f;
      const bool have = next_pair(cu, b, p, sp, q, me, rows, t, xj, Rx, Ry, r2);
      if (__ballot(have) == 0ull) break;  // wave-uniform
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      if (have) {
        pair_words(w, sp, p, norm, q, t, xj, Rx, Ry, r2);
        ++np;
      }
      const PairA a = pair_operands(w);
      floatx16 zlo, zhi;
      // K steps 0 and 1 of each tile and hidden half
      pair_layer(a.t0, w1, zlo, zhi);
      h00 = mfma16(sel[0], relu_rows(zlo, 0), h00);
      h01 = mfma16(sel[0], relu_rows(zhi, 0), h01);
      h00 = mfma16(sel[1], relu_rows(zlo, 1), h00);
      h01 = mfma16(sel[1], relu_rows(zhi, 1), h01);
      pair_layer(a.t1, w1, zlo, zhi);
      h10 = mfma16(sel[0], relu_rows(zlo, 0), h10);
      h11 = mfma16(sel[0], relu_rows(zhi, 0), h11);
      h10 = mfma16(sel[1], relu_rows(zlo, 1), h10);
      h11 = mfma16(sel[1], relu_rows(zhi, 1), h11);
    }
  };
  // one copy of the pass per source, so that the tile's loads are LDS loads
  // (a pointer that may be either is a flat address)
  if (src.tiled)
    pair_pass(tile_src(src, pool));
  else
    pair_pass(slot_src(src, b));
  shCnt[lane] = (float)np;
  if (m.pairPart) {
    uint32_t tot = np;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += (uint32_t)__shfl_xor((int)tot, off, 64);
    if (lane == 0) m.pairPart[blockIdx.x] = tot;
  }

  // 2. head, one 32-row tile at a time: C = A W^T with A in hb (BF16)
  auto dense = [&](const uint16_t *W, floatx16 &olo, floatx16 &ohi) {
    bf16x8 a[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = lds_bf8(&hb[c][16 * s + 8 * g]);
    olo = zero16();
    ohi = zero16();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      olo = mfma16(a[s], load_wb16(W, kHid, 0, 16 * s), olo);
      ohi = mfma16(a[s], load_wb16(W, kHid, 32, 16 * s), ohi);
    }
  };
  const bool train = (mode & kFwdTrain) != 0;
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
    auto hb4 = [&](int r, int c4) {
      return make_float4(bf2f(hb[r][c4]), bf2f(hb[r][c4 + 1]), bf2f(hb[r][c4 + 2]), bf2f(hb[r][c4 + 3]));
    };
    if (train) {
      store_act_rows(m.act, 0, base + 32 * t, n, hb4);
      if (g == 0 && i < n) m.cnt[i] = shCnt[32 * t + c];
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
    if (train) store_act_rows(m.act, kHid, base + 32 * t, n, hb4);
    // G = relu(W3 M + b3), FP32
    dense(m.w3b, olo, ohi);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int row = c_row(k, g);
      gbuf[row][c] = relu(olo[k] + prm[kOffB3 + c]);
      gbuf[row][32 + c] = relu(ohi[k] + prm[kOffB3 + 32 + c]);
    }
    __syncthreads();
    if (train)
      store_act_rows(m.act, 2 * kHid, base + 32 * t, n, [&](int r, int c4) {
        return make_float4(gbuf[r][c4], gbuf[r][c4 + 1], gbuf[r][c4 + 2], gbuf[r][c4 + 3]);
      });
    // 3. o = W4 G + b4 (FP32): lane (c, g) computes output g of particle i
    float o = prm[kOffB4 + g];
    for (int k = 0; k < kHid; ++k) o = __builtin_fmaf(prm[kOffW4 + g * kHid + k], gbuf[c][k], o);
    const float oy = __shfl_down(o, 32, 64);
    const bool head = g == 0 && i < n;
    if (head && m.out) {
      m.out[2 * (size_t)i] = o;
      m.out[2 * (size_t)i + 1] = oy;
    }
    if (mode & kFwdState) {
      DevScene sc = s0;
      float x = 0.0f, y = 0.0f;
      if (head) {
        sc = scene_of_slot<MULTI>(b, s0, b.keys[i]);
        const float2 xi = b.x1[i];
        x = xi.x + o * norm[5];
        y = xi.y + oy * norm[5];
      }
      collide_wave<MULTI>(b, p, sc, x, y, head);                               // demo4.cpp:411-444
      if (head) {
        if (mode & kFwdPressure) {
          const float2 d = b.dens[i];
          b.pres[i] = make_float2(p.k * (d.x - p.rho0), p.kn * d.y);           // sph.h:478-481
        }
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
// Head backward, one wave per 32 particles (the rows of one tile), on the
// matrix core: dO (FP32), dG = (dO W4) . [G > 0] (FP32), dM = dG W3 and
// dH = dM W2 as K = 16 BF16 MFMAs (dG, dM and the weights rounded to BF16,
// FP32 accumulation; the B operands read the transposed BF16 weight copies),
// and the wave's squared error and squared normalised target.  dG, dM and dH
// are written in FP32 for the weight gradients and the pair backward.
constexpr int kHeadRows = 32;
__global__ __launch_bounds__(64) void k_sur_back_head(int n, SurrogateBufs m, float invN) {
  __shared__ __align__(16) uint16_t tb[kHeadRows][kHbStride];  // dG, then dM, of the tile, BF16 (A operands)
  __shared__ float shD[kHeadRows][2];                          // dO of the tile's rows
  const int lane = threadIdx.x, c = lane & 31, g = lane >> 5;
  const int i = (int)blockIdx.x * kHeadRows + c;  // A-operand row c
  const bool live = i < n;
  const float *prm = m.params;
  const float os = prm[kOffNorm + 5];
  float d0 = 0.0f, d1 = 0.0f, err = 0.0f, ref = 0.0f;
  if (live) {
    const float2 t = m.target[i];
    const float t0 = t.x / os, t1 = t.y / os;
    const float e0 = m.out[2 * (size_t)i] - t0;
    const float e1 = m.out[2 * (size_t)i + 1] - t1;
    d0 = 2.0f * invN * e0;
    d1 = 2.0f * invN * e1;
    if (g == 0) {
      err = e0 * e0 + e1 * e1;
      ref = t0 * t0 + t1 * t1;
      m.dO[2 * (size_t)i] = d0;
      m.dO[2 * (size_t)i + 1] = d1;
    }
  }
  if (g == 0) {
    shD[c][0] = d0;
    shD[c][1] = d1;
  }
  __syncthreads();
  // dG = (dO W4) . [G > 0], row by row: 16 lanes take one row's 64 columns as
  // float4 (columns c4 .. c4 + 3 in every row this lane visits), so the G
  // reads and dG writes are 256 contiguous bytes per row; the BF16 copy in tb
  // is the A operand of dM = dG W3.
  {
    const int c4 = (lane & 15) * 4;
    float wa[4], wb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      wa[j] = prm[kOffW4 + c4 + j];
      wb[j] = prm[kOffW4 + kHid + c4 + j];
    }
    for (int q = lane; q < kHeadRows * 16; q += 64) {
      const int r = q >> 4, pi = (int)blockIdx.x * kHeadRows + r;
      float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (pi < n) {
        const float4 G4 = *reinterpret_cast<const float4 *>(m.act + (size_t)pi * kActStride + 2 * kHid + c4);
        const float gv[4] = {G4.x, G4.y, G4.z, G4.w};
        const float e0 = shD[r][0], e1 = shD[r][1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float x = __builtin_fmaf(wa[j], e0, wb[j] * e1);
          v[j] = gv[j] > 0.0f ? x : 0.0f;
        }
        *reinterpret_cast<float4 *>(m.dG + (size_t)pi * kHid + c4) = make_float4(v[0], v[1], v[2], v[3]);
      }
      *reinterpret_cast<uint32_t *>(&tb[r][c4]) = pack2(v[0], v[1]);
      *reinterpret_cast<uint32_t *>(&tb[r][c4 + 2]) = pack2(v[2], v[3]);
    }
  }
  __syncthreads();
  // A operands of dM: this lane's K indices k = 16s + 8g + u of row c
  bf16x8 ag[4];
#pragma unroll
  for (int st = 0; st < 4; ++st) ag[st] = lds_bf8(&tb[c][16 * st + 8 * g]);
  __syncthreads();  // tb is rewritten with dM below
  // dM = dG W3: B[k][n] = W3[k][n] = W3t[n][k]
  floatx16 mlo = zero16(), mhi = zero16();
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    mlo = mfma16(ag[st], load_wb16(m.w3t, kHid, 0, 16 * st), mlo);
    mhi = mfma16(ag[st], load_wb16(m.w3t, kHid, 32, 16 * st), mhi);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = c_row(r, g);
    const int pi = (int)blockIdx.x * kHeadRows + row;
    if (pi < n) {
      m.dM[(size_t)pi * kHid + c] = mlo[r];
      m.dM[(size_t)pi * kHid + 32 + c] = mhi[r];
    }
    tb[row][c] = f2bf(mlo[r]);
    tb[row][32 + c] = f2bf(mhi[r]);
  }
  __syncthreads();
  // dH = dM W2
  floatx16 hlo = zero16(), hhi = zero16();
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const bf16x8 a = lds_bf8(&tb[c][16 * st + 8 * g]);
    hlo = mfma16(a, load_wb16(m.w2t, kHid, 0, 16 * st), hlo);
    hhi = mfma16(a, load_wb16(m.w2t, kHid, 32, 16 * st), hhi);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int pi = (int)blockIdx.x * kHeadRows + c_row(r, g);
    if (pi < n) {
      m.dH[(size_t)pi * kHid + c] = hlo[r];
      m.dH[(size_t)pi * kHid + 32 + c] = hhi[r];
    }
  }
  // squared error and squared target of the wave's particles
  for (int off = 32; off > 0; off >>= 1) {
    err += __shfl_xor(err, off, 64);
    ref += __shfl_xor(ref, off, 64);
  }
  if (lane == 0) {
    m.lossPart[2 * blockIdx.x] = err;
    m.lossPart[2 * blockIdx.x + 1] = ref;
  }
}

// loss[0] = mean squared error, loss[1] = sum |e|^2 / sum |target/os|^2.
constexpr int kLossBlock = 1024;
__global__ __launch_bounds__(kLossBlock) void k_loss_final(const float *part, int nb, float invN, float *loss) {
  __shared__ float red[2][kLossBlock];
  // four partials in flight per thread, summed in a fixed order
  const float2 *pp = reinterpret_cast<const float2 *>(part);
  float2 acc[4] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
  int k = (int)threadIdx.x;
  for (; k + 3 * kLossBlock < nb; k += 4 * kLossBlock) {
    float2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = pp[k + u * kLossBlock];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u].x += v[u].x, acc[u].y += v[u].y;
  }
  for (; k < nb; k += kLossBlock) acc[0].x += pp[k].x, acc[0].y += pp[k].y;
  red[0][threadIdx.x] = (acc[0].x + acc[1].x) + (acc[2].x + acc[3].x);
  red[1][threadIdx.x] = (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y);
  __syncthreads();
  for (int w = kLossBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    loss[0] = red[0][0] * invN;
    loss[1] = red[1][0] > 0.0f ? red[0][0] / red[1][0] : 0.0f;
  }
}

// Weight-gradient partials on the matrix core: for the particles of this
// wave's K = 16 blocks, part[wave][a * 64 + b] = sum_i bf16(A[i][a]) bf16(B[i][b])
// (FP32 accumulation) and part[wave][na * 64 + a] = sum_i A[i][a] scale_i
// (FP32; scale 1 if null), for na = 2 or 64 rows of A and 64 columns of B.
// A K = 16 step takes particles k0 + 8g + u (u = 0..7) in lane group g: the A
// operand holds A[i][a] of row a = c (+32) and the B operand B[i][b] of
// column b = c (+32), 32-lane coalesced row reads.
constexpr int kOuterWaves = 2048;
template <bool SCALE>
__global__ __launch_bounds__(64) void k_outer_partial(int n, const float *A, int lda, int na, const float *B, int ldb,
                                                      const float *scale, float *part) {
  const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int aHalves = na > 32 ? 2 : 1;
  floatx16 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = zero16();
  float bias[2] = {0.0f, 0.0f};
  for (int k0 = (int)blockIdx.x * 16; k0 < n; k0 += (int)gridDim.x * 16) {
    // every load reads an in-range address (row and column clamped) and
    // out-of-range values are zeroed afterwards: with no branch around them,
    // the step's loads are all in flight together
    float va[2][8], vb[2][8], sc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = min(k0 + 8 * g + u, n - 1);
      sc[u] = SCALE ? scale[i] : 1.0f;
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        va[x][u] = A[(size_t)i * lda + min(32 * x + c, na - 1)];
        vb[x][u] = B[(size_t)i * ldb + 32 * x + c];
      }
    }
    bf16x8 ao[2], bo[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool in = k0 + 8 * g + u < n;
        va[x][u] = (in && 32 * x + c < na) ? va[x][u] : 0.0f;
        vb[x][u] = in ? vb[x][u] : 0.0f;
        if (in) bias[x] = __builtin_fmaf(va[x][u], sc[u], bias[x]);
      }
      ao[x] = bf8(pack2(va[x][0], va[x][1]), pack2(va[x][2], va[x][3]), pack2(va[x][4], va[x][5]),
                  pack2(va[x][6], va[x][7]));
      bo[x] = bf8(pack2(vb[x][0], vb[x][1]), pack2(vb[x][2], vb[x][3]), pack2(vb[x][4], vb[x][5]),
                  pack2(vb[x][6], vb[x][7]));
    }
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      if (x >= aHalves) break;
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[x][y] = mfma16(ao[x], bo[y], acc[x][y]);
    }
  }
  float *out = part + (size_t)blockIdx.x * (size_t)(na * kHid + na);
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    if (x >= aHalves) break;
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int a = 32 * x + c_row(r, g);
        if (a < na) out[a * kHid + 32 * y + c] = acc[x][y][r];
      }
    const float tot = bias[x] + __shfl_xor(bias[x], 32, 64);
    if (g == 0 && 32 * x + c < na) out[na * kHid + 32 * x + c] = tot;
  }
}

// out[e] = sum over parts q of part[q * width + e] (fixed order: deterministic).
__global__ __launch_bounds__(kBlock) void k_sum_parts(const float *part, int nparts, int width, float *out) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e >= width) return;
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  int q = 0;
  for (; q + 32 <= nparts; q += 32) {  // 32 loads in flight, added in the order of the loop below
    float v[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) v[u] = part[(size_t)(q + u) * width + e];
#pragma unroll
    for (int u = 0; u < 32; ++u) acc[u & 7] += v[u];
  }
  for (; q + 8 <= nparts; q += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += part[(size_t)(q + u) * width + e];
  }
  for (; q < nparts; ++q) acc[0] += part[(size_t)q * width + e];
  out[e] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Copies a reduced [na * nb | na] block into the weight and bias slots.
__global__ __launch_bounds__(kBlock) void k_scatter_grad(const float *sum, int na, int nb, float *w, float *bias) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e < na * nb) w[e] = sum[e];
  else if (e < na * nb + na) bias[e - na * nb] = sum[e];
}

// dW1 = sum_i sum_j (dH_i . [z_ij > 0]) f_ij^T, with the pair walk of the
// forward pass.  For row tile t the pair layer's pre-activations (C layout:
// lane = hidden unit, register r = tile row c_row(r, g)), masked and scaled by
// the rows' dH, are the A operand (hidden x 16 rows) of K step s (registers
// 8s .. 8s + 7); the same rows' features, staged in LDS, are the B operand
// (16 rows x 32 features, 16 used).  Each wave accumulates over its particles
// and writes one 64 x 16 partial.
template <bool MULTI>
__global__ __launch_bounds__(64) void k_sur_back_pairs(Buffers b, DevParams p, DevScene s0, int n, SurrogateBufs m) {
  __shared__ uint32_t fw[kWaveRows][kFeat / 2];
  __shared__ __align__(16) unsigned char pool[kStageBytes];
  __shared__ TilePlan plan;
  const int lane = threadIdx.x, c = lane & 31, g = lane >> 5;
  const float *norm = m.params + kOffNorm;
  const W1Operands w1 = load_w1(m.w1b);
  floatx16 glo = zero16(), ghi = zero16();
  for (int base = (int)blockIdx.x * kWaveRows; base < n; base += (int)gridDim.x * kWaveRows) {
    const int me = base + lane;
    __syncthreads();  // the previous group's tile is no longer read
    const PartRaw raw = part_raw(b, min(me, n - 1));  // a lane past n reads the last particle, unused
    const PairSrc src = stage_pairs<MULTI>(b, s0, base, n, pool, plan);
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
    Cursor cu{0u, 0u, 0u, 0, true, 0u, 0u};
    if (me < n) {
      q = part_info<MULTI>(b, p, s0, norm, raw);
      cu.done = false;
    }
    for (;;) {
      uint32_t t = 0;
      float2 xj = make_float2(0.0f, 0.0f);
      float Rx = 0.0f, Ry = 0.0f, r2 = 0.0f;
      const bool have = next_pair(cu, b, p, src, q, me, false, t, xj, Rx, Ry, r2);
      if (__ballot(have) == 0ull) break;  // wave-uniform
      uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      if (have) pair_words(w, src, p, norm, q, t, xj, Rx, Ry, r2);
#pragma unroll
      for (int k = 0; k < 8; ++k) fw[lane][k] = w[k];
      const PairA a = pair_operands(w);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        floatx16 zlo, zhi;
        pair_layer(t ? a.t1 : a.t0, w1, zlo, zhi);
        const floatx16 &dlo = t ? d10 : d00;
        const floatx16 &dhi = t ? d11 : d01;
        // K step st: K index 8g + u is tile row c_row(8st + u, g), this lane's
        // register 8st + u (A: masked dH of hidden c; B: that row's feature c)
#pragma unroll
        for (int st = 0; st < 2; ++st) {
          uint32_t bw[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            uint32_t hv[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
              const uint32_t wd = fw[32 * t + c_row(8 * st + 2 * q + e, g)][(c & 15) >> 1];
              hv[e] = c < kFeat ? ((c & 1) ? (wd >> 16) : (wd & 0xffffu)) : 0u;
            }
            bw[q] = hv[0] | (hv[1] << 16);
          }
          float ml[8], mh[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            ml[u] = zlo[8 * st + u] > 0.0f ? dlo[8 * st + u] : 0.0f;
            mh[u] = zhi[8 * st + u] > 0.0f ? dhi[8 * st + u] : 0.0f;
          }
          const bf16x8 bf = bf8(bw[0], bw[1], bw[2], bw[3]);
          glo = mfma16(bf8(pack2(ml[0], ml[1]), pack2(ml[2], ml[3]), pack2(ml[4], ml[5]), pack2(ml[6], ml[7])), bf,
                       glo);
          ghi = mfma16(bf8(pack2(mh[0], mh[1]), pack2(mh[2], mh[3]), pack2(mh[4], mh[5]), pack2(mh[6], mh[7])), bf,
                       ghi);
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

// BF16 copies of W1..W3 and the transposes of W2, W3 (w2t[n][k] = W2[k][n]) for
// the backward pass's B operands.
__global__ __launch_bounds__(kBlock) void k_bf16_weights(const float *params, uint16_t *w1b, uint16_t *w2b,
                                                         uint16_t *w3b, uint16_t *w2t, uint16_t *w3t) {
  const int t = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (t < kHid * kFeat) w1b[t] = f2bf(params[kOffW1 + t]);
  if (t < kHid * kHid) {
    w2b[t] = f2bf(params[kOffW2 + t]);
    w3b[t] = f2bf(params[kOffW3 + t]);
    const int r = t / kHid, k = t % kHid;
    w2t[k * kHid + r] = f2bf(params[kOffW2 + t]);
    w3t[k * kHid + r] = f2bf(params[kOffW3 + t]);
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

// Per-scene sums of |a.x - b.x| and b.x into out[2 s], out[2 s + 1] (zeroed
// beforehand).  A scene's particles occupy consecutive slots.  A workgroup
// covers kErrItems * kBlock consecutive slots; when they all belong to one
// scene (the first and last slot share it) it reduces and adds once, else each
// wave adds per 64-slot group (once when the group shares a scene, per lane
// otherwise).
__global__ __launch_bounds__(kBlock) void k_dens_err_scenes(const float2 *a, const float2 *bb, const uint32_t *keys,
                                                            const uint32_t *sceneBase, int nScenes, int n,
                                                            double *out) {
  __shared__ double red[2][kBlock / 64];
  const int base = (int)blockIdx.x * kBlock * kErrItems;
  const int last = min(n, base + kBlock * kErrItems) - 1;
  const int s0 = scene_of_key(sceneBase, nScenes, keys[base]);
  const bool one = scene_of_key(sceneBase, nScenes, keys[last]) == s0;
  double e = 0.0, r = 0.0;
  for (int k = 0; k < kErrItems; ++k) {
    const int i = base + k * kBlock + (int)threadIdx.x;
    const bool valid = i < n;
    double ei = 0.0, ri = 0.0;
    if (valid) {
      ei = fabs((double)a[i].x - (double)bb[i].x);
      ri = (double)bb[i].x;
    }
    if (one) {
      e += ei;
      r += ri;
      continue;
    }
    const int sc = valid ? scene_of_key(sceneBase, nScenes, keys[i]) : -1;
    const int first = __shfl(sc, 0, 64);
    if (__ballot(valid && sc != first) == 0ull) {
      for (int off = 32; off > 0; off >>= 1) {
        ei += __shfl_xor(ei, off, 64);
        ri += __shfl_xor(ri, off, 64);
      }
      if ((threadIdx.x & 63) == 0 && first >= 0) {
        atomicAdd(&out[2 * first], ei);
        atomicAdd(&out[2 * first + 1], ri);
      }
    } else if (valid) {
      atomicAdd(&out[2 * sc], ei);
      atomicAdd(&out[2 * sc + 1], ri);
    }
  }
  if (!one) return;  // uniform across the workgroup
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    r += __shfl_xor(r, off, 64);
  }
  if ((threadIdx.x & 63) == 0) red[0][threadIdx.x >> 6] = e, red[1][threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t0 += red[0][w], t1 += red[1][w];
    atomicAdd(&out[2 * s0], t0);
    atomicAdd(&out[2 * s0 + 1], t1);
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double *p, int count) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i < count) p[i] = 0.0;
}

// Sums the partials in a fixed order: thread t adds partials t, t + 1024, ...,
// then a tree over the threads.
__global__ __launch_bounds__(1024) void k_dens_err_final(const double *partials, int nb, double *out) {
  __shared__ double red[2][1024 / 64];
  double t0 = 0.0, t1 = 0.0;
  for (int k = (int)threadIdx.x; k < nb; k += 1024) t0 += partials[2 * k], t1 += partials[2 * k + 1];
  for (int off = 32; off > 0; off >>= 1) {
    t0 += __shfl_xor(t0, off, 64);
    t1 += __shfl_xor(t1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) red[0][threadIdx.x >> 6] = t0, red[1][threadIdx.x >> 6] = t1;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < 1024 / 64; ++w) s0 += red[0][w], s1 += red[1][w];
    out[0] = s0;
    out[1] = s1;
  }
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

// K = 16 layout self-test: C (32 x 32) = A (32 x 16) B (16 x 32).
__global__ __launch_bounds__(64) void k_mfma16_probe(const float *A, const float *B, float *C) {
  const int lane = threadIdx.x & 63, c = lane & 31, g = lane >> 5;
  float a[8], bb[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    a[t] = A[c * 16 + 8 * g + t];
    bb[t] = B[(8 * g + t) * 32 + c];
  }
  const floatx16 r = mfma16(bf8(pack2(a[0], a[1]), pack2(a[2], a[3]), pack2(a[4], a[5]), pack2(a[6], a[7])),
                            bf8(pack2(bb[0], bb[1]), pack2(bb[2], bb[3]), pack2(bb[4], bb[5]), pack2(bb[6], bb[7])),
                            zero16());
#pragma unroll
  for (int k = 0; k < 16; ++k) C[c_row(k, g) * 32 + c] = r[k];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static unsigned waves_for(int n) { return (unsigned)((n + kWaveRows - 1) / kWaveRows); }
static unsigned back_pair_waves(int n) { return std::min(waves_for(n), 4096u); }
static unsigned outer_blocks(int n) { return (unsigned)std::max(1, std::min((n + 15) / 16, kOuterWaves)); }

size_t surrogate_part_floats(int n) {
  const size_t a = (size_t)back_pair_waves(n) * kHid * kFeat;
  const size_t o = (size_t)outer_blocks(n) * (kHid * kHid + kHid);
  return std::max<size_t>(std::max(a, o), 1) + (kHid * kHid + kHid);  // + the reduced block
}

int surrogate_head_blocks(int n) { return std::max(1, (n + kHeadRows - 1) / kHeadRows); }
int surrogate_forward_waves(int n) { return std::max(1, (int)waves_for(n)); }

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
  if (scale)
    hipLaunchKernelGGL(k_outer_partial<true>, dim3(blocks), dim3(64), 0, s, n, A, lda, na, B, ldb, scale, part);
  else
    hipLaunchKernelGGL(k_outer_partial<false>, dim3(blocks), dim3(64), 0, s, n, A, lda, na, B, ldb, scale, part);
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
  hipLaunchKernelGGL(k_sur_back_head, dim3(hb), dim3(64), 0, s, n, m, invN);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kLossBlock), 0, s, (const float *)m.lossPart, hb, invN,
                     m.loss + 2 * (size_t)m.lossIndex);
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

void launch_bf16_weights(const float *params, uint16_t *w1b, uint16_t *w2b, uint16_t *w3b, uint16_t *w2t,
                         uint16_t *w3t, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_weights, dim3((kHid * kHid + kBlock - 1) / kBlock), dim3(kBlock), 0, s, params, w1b, w2b,
                     w3b, w2t, w3t);
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
  hipLaunchKernelGGL(k_dens_err_final, dim3(1), dim3(1024), 0, s, (const double *)partials, n > 0 ? nb : 0, out);
}

void launch_density_error_scenes(const float2 *a, const float2 *bb, const uint32_t *keys, const uint32_t *sceneBase,
                                 int nScenes, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f64, dim3((2 * nScenes + kBlock - 1) / kBlock), dim3(kBlock), 0, s, out, 2 * nScenes);
  if (n > 0)
    hipLaunchKernelGGL(k_dens_err_scenes, dim3((n + kBlock * kErrItems - 1) / (kBlock * kErrItems)), dim3(kBlock), 0,
                       s, a, bb, keys, sceneBase, nScenes, n, out);
}

void launch_mfma_probe(const float *a, const float *b, float *c, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, a, b, c);
}

void launch_mfma16_probe(const float *a, const float *b, float *c, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma16_probe, dim3(1), dim3(64), 0, s, a, b, c);
}

}  // namespace nbs
