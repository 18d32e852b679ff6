// HIP kernels of the SPH step for gfx950 (MI355X).
//
// One step is five kernel groups on one stream:
//   1. k_visc_predict   integrate, viscosity (Jacobi gather over the previous
//                       binning), predict, cell key of the predicted position and
//                       the per-cell arrival count (demo4.cpp:301-356)
//   2. k_scan_reduce + k_scan_final (k_scan_small for small grids)
//                       exclusive scan of the per-cell counts -> cell starts
//   3. k_place (grids above kSmallGridCells cells once every particle is
//      binned), else k_scatter + k_rank_reorder
//                       bucket the particles by cell, order each cell by the
//                       particles' previous slots (a stable counting sort of the
//                       slot order), and move the state into the new slot order
//   4. k_density        density, near density, pressure (demo4.cpp:208-221)
//   5. k_delta          double-density relaxation (Jacobi gather), plane /
//                       circle / segment / polygon collisions, velocity
//                       (demo4.cpp:239-255, 411-450)
// Small steps use the wide variants of 1, 4 and 5 (one wave per particle;
// 1 is then followed by k_count_fill, or for the smallest steps the counting
// and the bucketing move into the one-workgroup scan, k_count_scan_small,
// whose records k_rank_reorder reads).
//
// Every FP32 expression keeps the reference's operand order and the library
// is compiled with -ffp-contract=off and correctly rounded sqrt / divide, so the
// arithmetic of each pair term is the reference's (sph.h:465-542).
#include <algorithm>

#include "nbs_device.hpp"

namespace nbs {

// For every candidate slot j of the particle at xi in cell (cx, cy) whose
// r2 < h2 (sph.h:470, 488, 502), in candidate order, applies the pair's
// contribution: apply(term(j, Rx, Ry, r2, true)).  term(..., false) must return
// a contribution that leaves the accumulators unchanged bit for bit.
// Positions are read from X at idx(d, j) (the LDS tile or the slot array);
// returns the number of candidates (the reference's neighbour count).  Each
// row range is taken in chunks of 32: a cheap pass marks the in-kernel
// candidates in a 32-bit mask, then the pair work runs only for the marked
// ones, lowest first, so the accumulation order is the candidate order and
// the expensive work is not issued for candidates outside the kernel radius.
constexpr int kPairsPerIter = 4;

template <typename Idx, typename Term, typename Apply>
__device__ __forceinline__ uint32_t for_each_pair(const uint32_t *start, const DevScene &sc, int cx, int cy,
                                                  const float2 *X, Idx idx, float2 xi, float h2, Term &&term,
                                                  Apply &&apply) {
  uint32_t cand = 0;
  RowRange rows[3];
  row_ranges(start, sc, cx, cy, rows);
  for (int d = 0; d < 3; ++d) {
    const RowRange rr = rows[d];
    cand += rr.j1 - rr.j0;
    for (uint32_t base = rr.j0; base < rr.j1; base += 32u) {
      const uint32_t cnt = min(32u, rr.j1 - base);
      const uint32_t q0 = idx(d, base);
      // mask = 2 * mask + (r2 < h2), last candidate first, so candidate t
      // ends at bit t: the compare's lane mask is the add's carry-in.
      uint32_t mask = 0;
      auto push = [&](float r2) {
        asm("v_cmp_gt_f32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
            : "+v"(mask)
            : "s"(h2), "v"(r2)
            : "vcc");
      };
      auto dist2 = [&](uint32_t t) {
        const float2 xj = X[q0 + t];
        const float Rx = xj.x - xi.x, Ry = xj.y - xi.y;
        return Rx * Rx + Ry * Ry;
      };
      uint32_t u = cnt;
      for (; u >= 8u; u -= 8u) {  // eight position loads in flight
        float r2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r2[k] = dist2(u - 1 - k);
#pragma unroll
        for (int k = 0; k < 8; ++k) push(r2[k]);
      }
      for (; u >= 4u; u -= 4u) {
        const float a = dist2(u - 1), b = dist2(u - 2), c = dist2(u - 3), e = dist2(u - 4);
        push(a);
        push(b);
        push(c);
        push(e);
      }
      for (; u > 0u; --u) push(dist2(u - 1));
      // Four pairs per iteration: their loads and arithmetic are independent,
      // which shortens the serial chain of a crowded particle's walk; the
      // contributions are applied lowest first.  Slots past the last marked
      // bit repeat the first pair and contribute the identity.
      while (mask) {
        uint32_t t[kPairsPerIter];
        bool use[kPairsPerIter];
        t[0] = (uint32_t)__builtin_ctz(mask);
        mask &= mask - 1u;
        use[0] = true;
#pragma unroll
        for (int k = 1; k < kPairsPerIter; ++k) {
          use[k] = mask != 0u;
          t[k] = use[k] ? (uint32_t)__builtin_ctz(mask) : t[0];
          mask &= use[k] ? mask - 1u : ~0u;
        }
        float2 xs[kPairsPerIter];
#pragma unroll
        for (int k = 0; k < kPairsPerIter; ++k) xs[k] = X[q0 + t[k]];
        decltype(term(0u, 0.0f, 0.0f, 0.0f, true)) c[kPairsPerIter];
#pragma unroll
        for (int k = 0; k < kPairsPerIter; ++k) {
          const float Rx = xs[k].x - xi.x, Ry = xs[k].y - xi.y;
          c[k] = term(base + t[k], Rx, Ry, Rx * Rx + Ry * Ry, use[k]);
        }
#pragma unroll
        for (int k = 0; k < kPairsPerIter; ++k) apply(c[k]);
      }
    }
  }
  return cand;
}

// Wide pair walk for small particle counts: one wave per particle.  Lane t
// tests candidate base + t of a 64-candidate chunk and, when the pair is in
// range, term(j, Rx, Ry, r2, G[j], t) computes the pair's K terms and says
// whether it contributes.  The contributions are then applied in candidate
// order to two accumulators held in lanes 0 and 1: for each contributing pair
// and each of STAGES stages s, lane c in {0, 1} does acc -= t[(2s + c) % K]
// (SUB) or acc += t[(2s + c) % K], reading the terms back from the wave's LDS
// rows (lds: 65 * K floats; row 64 holds the neutral value +0 for SUB, -0
// otherwise, which leaves an accumulator unchanged bit for bit and pads groups
// of four pairs).  So the accumulation performs the same FP32 operations in
// the same order as the one-lane-per-particle walk; only the per-pair
// arithmetic runs in parallel across the wave.  acc0 and acc1 are the
// accumulators' initial and final values (the same in every lane).
// The chunks are the 64-candidate pieces of the rows dy = -1, 0, 1 in order;
// the row ranges are the wave's particle's, so the chunk sequence is
// wave-uniform.  The position (and G value) loads of the first kWideAhead
// chunks are issued together before the first is used (a particle of a
// crowded scene has several chunks, and one chunk ahead left a memory latency
// per chunk on the wave's path); later chunks are loaded one ahead.  idx(1, .)
// of the particle's own row, which always holds the particle, serves as the
// in-range address of the lanes past a row's end.  Returns the candidate count.
constexpr uint32_t kWideAhead = 8;

template <int K, int STAGES, bool SUB, typename Idx, typename Term>
__device__ __forceinline__ uint32_t for_each_pair_wide(const uint32_t *start, const DevScene &sc, int cx, int cy,
                                                       const float2 *X, Idx idx, float2 xi, float h2, const float2 *G,
                                                       float *lds, float &acc0, float &acc1, Term &&term) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t c = lane & 1u;
  if (lane < (uint32_t)K) lds[64 * K + lane] = SUB ? 0.0f : -0.0f;
  float acc = c ? acc1 : acc0;
  RowRange rr[3];
  row_ranges(start, sc, cx, cy, rr);
  uint32_t j0[3], j1[3], nc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {  // wave-uniform
    j0[d] = (uint32_t)__builtin_amdgcn_readfirstlane((int)rr[d].j0);
    j1[d] = (uint32_t)__builtin_amdgcn_readfirstlane((int)rr[d].j1);
    nc[d] = (j1[d] - j0[d] + 63u) / 64u;
  }
  const uint32_t cand = (j1[0] - j0[0]) + (j1[1] - j0[1]) + (j1[2] - j0[2]);
  const uint32_t nch = nc[0] + nc[1] + nc[2];
  const uint32_t own = idx(1, j0[1]);  // an in-range slot (the own row holds the particle)
  // chunk k: row d, first slot base, row end
  auto chunk = [&](uint32_t k, int &d, uint32_t &base, uint32_t &end) {
    if (k < nc[0]) {
      d = 0;
      base = j0[0] + 64u * k;
    } else if (k < nc[0] + nc[1]) {
      d = 1;
      base = j0[1] + 64u * (k - nc[0]);
    } else {
      d = 2;
      base = j0[2] + 64u * (k - nc[0] - nc[1]);
    }
    end = j1[d];
  };
  auto load = [&](uint32_t k, float2 &x, float2 &g, bool &in, uint32_t &j) {
    int d;
    uint32_t base, end;
    chunk(k, d, base, end);
    j = base + lane;
    in = j < end;
    x = X[in ? idx(d, base) + lane : own];
    if (G) g = G[in ? j : own];
  };
  auto process = [&](float2 xj, float2 gj, bool in, uint32_t j) {
    float t[K];
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] = 0.0f;
    bool ok = false;
    if (in) {
      const float Rx = xj.x - xi.x, Ry = xj.y - xi.y;
      const float r2 = Rx * Rx + Ry * Ry;
      if (r2 < h2) ok = term(j, Rx, Ry, r2, gj, t);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) lds[lane * K + k] = t[k];
    uint64_t mask = __ballot(ok);
    while (mask) {
      uint32_t q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        q[u] = mask ? (uint32_t)__builtin_ctzll(mask) : 64u;
        mask &= mask - 1u;  // no-op once empty
      }
      float v[4][STAGES];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int st = 0; st < STAGES; ++st) v[u][st] = lds[q[u] * K + (2 * st + c) % K];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int st = 0; st < STAGES; ++st) acc = SUB ? acc - v[u][st] : acc + v[u][st];
    }
  };
  float2 xa[kWideAhead], ga[kWideAhead];
  bool ia[kWideAhead];
  uint32_t ja[kWideAhead];
#pragma unroll
  for (uint32_t k = 0; k < kWideAhead; ++k) {
    if (k >= nch) break;  // wave-uniform
    ga[k] = make_float2(0.0f, 0.0f);
    load(k, xa[k], ga[k], ia[k], ja[k]);
  }
#pragma unroll
  for (uint32_t k = 0; k < kWideAhead; ++k) {
    if (k >= nch) break;
    process(xa[k], ga[k], ia[k], ja[k]);
  }
  if (nch > kWideAhead) {  // the rest, one chunk ahead
    float2 xn, gn = make_float2(0.0f, 0.0f);
    bool inN;
    uint32_t jn;
    load(kWideAhead, xn, gn, inN, jn);
    for (uint32_t k = kWideAhead; k < nch; ++k) {
      const float2 xj = xn, gj = gn;
      const bool in = inN;
      const uint32_t j = jn;
      if (k + 1u < nch) load(k + 1u, xn, gn, inN, jn);
      process(xj, gj, in, j);
    }
  }
  acc0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), 0));
  acc1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), 1));
  return cand;
}

// Arrival index of this lane's particle in cell `key`: lanes of the wave with
// equal keys are grouped, the lowest lane of each group adds the group size to
// the cell counter with one atomic, and each lane takes base + its rank in the
// group.  Consecutive sorted slots mostly share a cell, so this issues a few
// atomics per wave instead of one per particle.  All active lanes must call it.
__device__ __forceinline__ uint32_t wave_aggregated_fill(uint32_t *count, uint32_t key) {
  const int lane = (int)(threadIdx.x & 63u);
  uint64_t todo = __ballot(1);
  int leader = lane;
  uint64_t group = 0;
  while (todo) {  // one pass per distinct key; the key and masks are wave-uniform
    const int l = __builtin_ctzll(todo);
    const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, l);
    const uint64_t m = __ballot(key == k) & todo;
    if (key == k) {
      leader = l;
      group = m;
    }
    todo &= ~m;
  }
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(group >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)group, 0u));
  uint32_t base = 0;
  if (leader == lane) base = atomicAdd(&count[key], (uint32_t)__popcll(group));
  base = (uint32_t)__shfl((int)base, leader, 64);
  return base + rank;
}

// Appends `value` of every lane with `flag` to list[*counter ...] with one
// atomic per wave; entries at cap or beyond are dropped.  All active lanes must
// call it.
__device__ __forceinline__ void wave_append(uint32_t *list, uint32_t *counter, uint32_t cap, bool flag,
                                            uint32_t value) {
  const uint64_t m = __ballot(flag);
  if (m == 0ull) return;  // wave-uniform
  const int lane = (int)(threadIdx.x & 63u);
  const int leader = __builtin_ctzll(m);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, leader, 64);
  const uint32_t at = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  if (flag && at < cap) list[at] = value;
}

constexpr uint32_t kFarFlag = 0x80000000u;  // bit 31 of keyp in placement mode

// ---------------------------------------------------------------------------
// 1. integrate + viscosity + predict + key + count
// ---------------------------------------------------------------------------
// demo4.cpp:223-237 in Jacobi order: both halves of each pair's impulse from
// the integrated velocities (neighbours have no own acceleration: v_j + g dt).
template <typename Idx>
__device__ __forceinline__ void viscosity_sum(const uint32_t *start, const DevScene &sc, int cx, int cy,
                                              const float2 *X, const float2 *V, Idx idx, const DevParams &p,
                                              float2 xi, float vx, float vy, float &accx, float &accy) {
  for_each_pair(
      start, sc, cx, cy, X, idx, xi, p.h2,
      [&](uint32_t j, float Rx, float Ry, float r2, bool use) {
        const float2 vj = V[j];
        const float vjx = p.gdtx + vj.x, vjy = p.gdty + vj.y;
        float r, inv;
        pair_length(r2, r, inv);                          // sph.h:503, 505 Vec2Normalize (length 0 -> 1)
        const float qq = r * p.invH;                      // sph.h:504
        const float nX = Rx * inv, nY = Ry * inv;
        const float u = (vx - vjx) * nX + (vy - vjy) * nY;  // sph.h:506
        // no impulse unless u > 0 (sph.h:507): a zero force times the finite
        // normal, 0.5 and dt gives +0 or -0, and subtracting either leaves the
        // sums unchanged (they start at +0 and never hold -0)
        const float f = (use && u > 0.0f) ? (1.0f - qq) * (p.sigma * u + p.beta * (u * u)) : 0.0f;  // sph.h:508
        const float ix = ((f * nX) * 0.5f) * p.dt;       // demo4.cpp:233 force * 0.5f * deltaTime
        const float iy = ((f * nY) * 0.5f) * p.dt;
        return make_float2(ix, iy);
      },
      [&](float2 im) {
        accx = accx - im.x;                               // own visit
        accy = accy - im.y;
        accx = accx - im.x;                               // neighbour's visit (f_ji = -f_ij)
        accy = accy - im.y;
      });
}

template <bool MULTI, int CAP>
__global__ __launch_bounds__(kBlock) void k_visc_predict(Buffers b, DevParams p, DevScene s0, int n, int nBinned,
                                                         int lists) {
  __shared__ float2 tx[CAP > 0 ? CAP : 1];
  __shared__ TilePlan plan;
  const int i0 = (int)(blockIdx.x * kBlock);
  const int i = i0 + (int)threadIdx.x;
  // Viscosity runs for slots of the previous binning when lists exist.
  const bool tileVisc = lists && i0 < nBinned;
  if (tileVisc) {
    if (threadIdx.x == 0) {
      const uint32_t kA = b.keys[i0], kB = b.keys[min(i0 + kBlock, nBinned) - 1];
      const DevScene sa = scene_of_slot<MULTI>(b, s0, kA);
      plan_tile<CAP>(b.startCur, sa, kA, kB, plan);
      if (MULTI && sa.cellBase != scene_of_slot<MULTI>(b, s0, kB).cellBase) plan.use = 0;
    }
    __syncthreads();
    if (CAP > 0 && plan.use) {
      stage(plan, b.x0, tx);
    }
    __syncthreads();
  }
  if (i >= n) return;
  const float2 xi = b.x0[i];
  const float2 vi = b.v[i];
  const bool binned = i < nBinned;
  float aX, aY;
  DevScene sc;
  if (binned) {
    aX = p.gex + 0.0f;                                   // demo4.cpp:306 (own acceleration is zero)
    aY = p.gey + 0.0f;
    sc = scene_of_slot<MULTI>(b, s0, b.keys[i]);
  } else {
    const float2 a = b.acc[i];
    aX = p.gex + a.x;
    aY = p.gey + a.y;
    sc = MULTI ? b.scenes[b.scn[i]] : s0;
  }
  float vx = (aX * p.dt) + vi.x;                         // demo4.cpp:307
  float vy = (aY * p.dt) + vi.y;

  int pcx = -2, pcy = -2;  // cell of the last binning (none: not adjacent to any cell)
  if (lists && binned) {
    const uint32_t local = b.keys[i] - sc.cellBase;
    pcy = (int)(local / (uint32_t)sc.nx);
    pcx = (int)local - pcy * sc.nx;
    float accx = 0.0f, accy = 0.0f;
    if (CAP > 0 && plan.use)
      viscosity_sum(b.startCur, sc, pcx, pcy, tx, b.v, TileIndex(plan), p, xi, vx, vy, accx, accy);
    else
      viscosity_sum(b.startCur, sc, pcx, pcy, b.x0, b.v, SlotIndex(), p, xi, vx, vy, accx, accy);
    vx = accx + vx;
    vy = accy + vy;
  }

  const float xpx = (vx * p.dt) + xi.x;                  // demo4.cpp:334-335
  const float xpy = (vy * p.dt) + xi.y;
  b.xp[i] = make_float2(xpx, xpy);
  const int cx = cell_coord(xpx, sc.halfW, sc.cellSize, sc.nx);
  const int cy = cell_coord(xpy, sc.halfH, sc.cellSize, sc.ny);
  const uint32_t key = sc.cellBase + (uint32_t)(cy * sc.nx + cx);
  if (b.farList) {
    // placement mode: counts only (k_place orders the cells); a particle that
    // left the 3x3 cells around its last cell is flagged and listed
    const bool far = abs(cx - pcx) > 1 || abs(cy - pcy) > 1;
    b.keyp[i] = key | (far ? kFarFlag : 0u);
    (void)wave_aggregated_fill(b.count, key);
    wave_append(b.farList, b.farCount, (uint32_t)n, far, (uint32_t)i);
  } else {
    b.keyp[i] = key;
    b.fill[i] = wave_aggregated_fill(b.count, key);
  }
}

// Restores the per-slot scene and zero acceleration of particles that lose
// their binning (a new scene changes the cell tables; a restored state starts
// without neighbour lists).
__global__ __launch_bounds__(kBlock) void k_unbin(Buffers b, int nBinned) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= nBinned) return;
  b.scn[i] = (uint32_t)scene_index(b, b.keys[i]);
  b.acc[i] = make_float2(0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// 2. exclusive scan of the per-cell counts
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across the workgroup.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t waveTot[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) waveTot[wave] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const uint32_t t = waveTot[w];
    if (w < wave) off += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Workgroup-wide min and max (result valid in every thread).
__device__ __forceinline__ void block_min_max(uint32_t &mn, uint32_t &mx) {
  __shared__ uint32_t shMin[kBlock / 64], shMax[kBlock / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, d, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    shMin[wave] = mn;
    shMax[wave] = mx;
  }
  __syncthreads();
  mn = shMin[0];
  mx = shMax[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) {
    mn = min(mn, shMin[w]);
    mx = max(mx, shMax[w]);
  }
}

// The kScanItems = 4 counts of a scan thread's cells [base, base + 4), zero
// past the last cell: one 16-byte load with no branch around it (the counter
// array holds whole scan chunks, so the load is always in range).
static_assert(kScanItems == 4, "a scan thread's cells are one 16-byte vector");
__device__ __forceinline__ void load_counts4(const uint32_t *count, uint64_t base, uint32_t cells, uint32_t (&v)[4]) {
  const uint4 q = *reinterpret_cast<const uint4 *>(count + base);
  v[0] = base < cells ? q.x : 0u;
  v[1] = base + 1u < cells ? q.y : 0u;
  v[2] = base + 2u < cells ? q.z : 0u;
  v[3] = base + 3u < cells ? q.w : 0u;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(Buffers b, uint32_t cells) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint32_t s = 0, mx = 0, mn = 0xFFFFFFFFu;
  uint32_t cv[kScanItems];
  load_counts4(b.count, base, cells, cv);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t v = cv[k];
    s += v;
    mx = max(mx, v);
    if (v > 0u) mn = min(mn, v);
  }
  uint32_t tot;
  block_excl_scan(s, &tot);
  // occupied-cell min / max, one pair per workgroup
  block_min_max(mn, mx);
  if (threadIdx.x == 0) {
    b.cellStats[2 * blockIdx.x] = mn;
    b.cellStats[2 * blockIdx.x + 1] = mx;
  }
  if (threadIdx.x == 0) b.blockSum[blockIdx.x] = tot;
  // the next step's far-mover counter (its previous user, k_place of the last
  // step, has finished)
  if (blockIdx.x == 0 && threadIdx.x == 0 && b.farCountClear) *b.farCountClear = 0u;
}

__global__ __launch_bounds__(kBlock) void k_scan_final(Buffers b, uint32_t cells, uint32_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  load_counts4(b.count, base, cells, v);
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += v[k];
  // offset of this workgroup: the sum of the preceding workgroups' totals
  __shared__ uint32_t shOff[kBlock / 64];
  uint32_t pre = 0;
  for (uint32_t q = threadIdx.x; q < blockIdx.x; q += kBlock) pre += b.blockSum[q];
  for (int d = 32; d >= 1; d >>= 1) pre += (uint32_t)__shfl_xor((int)pre, d, 64);
  if ((threadIdx.x & 63u) == 0) shOff[threadIdx.x >> 6] = pre;
  __syncthreads();
  uint32_t blockOff = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) blockOff += shOff[w];
  uint32_t tot;
  uint32_t run = blockOff + block_excl_scan(s, &tot);
  uint32_t st[kScanItems];  // the cells' starts
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    st[k] = run;
    run += v[k];
  }
  if (base + kScanItems <= cells) {
    *reinterpret_cast<uint4 *>(b.startNext + base) = make_uint4(st[0], st[1], st[2], st[3]);
    *reinterpret_cast<uint4 *>(b.count + base) = make_uint4(0u, 0u, 0u, 0u);  // ready for the next step's counting
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t c = base + k;
    if (c < cells) {
      if (base + kScanItems > cells) {
        b.startNext[c] = st[k];
        b.count[c] = 0u;
      }
      if (b.chunkCells && v[k] > 0u) {
        // placement chunks whose first or last slot lies in this cell
        const uint32_t first = st[k], last = st[k] + v[k] - 1u, lastChunk = (n - 1u) / kPlaceChunk;
        for (uint32_t g = (first + kPlaceChunk - 1u) / kPlaceChunk; g * kPlaceChunk <= last; ++g)
          b.chunkCells[2 * g] = (uint32_t)c;
        for (uint32_t g = (first + 1u + kPlaceChunk - 1u) / kPlaceChunk; g <= (last + 1u) / kPlaceChunk; ++g)
          if (g >= 1u && g - 1u < lastChunk) b.chunkCells[2 * (g - 1u) + 1] = (uint32_t)c;
        if (last == n - 1u) b.chunkCells[2 * lastChunk + 1] = (uint32_t)c;
      }
      if (c == (uint64_t)cells - 1) b.startNext[cells] = st[k] + v[k];
    }
  }
}

// ---------------------------------------------------------------------------
// 3. bucket by cell, order each cell by particle index, move the state
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_scatter(Buffers b, int n) {
  const int s = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (s >= n) return;
  const uint32_t key = b.keyp[s];
  const uint32_t pos = b.startNext[key] + b.fill[s];
  b.tmpSlot[pos] = (uint32_t)s;
}

// A workgroup's slots are consecutive in cell order, so the cells they belong
// to span [c0 of its first slot, c1 of its last slot); when that span fits
// kRankTile, its source slots are staged in LDS for the ranking loops.
constexpr uint32_t kRankTile = 2048;

// Rank of source slot `mine` among the source slots of its new cell, the
// entries [c0, c1) of the bucketed slots: the canonical order within a cell is
// the order of the slots (a stable counting sort of the slot order).
template <typename Slots>
__device__ __forceinline__ uint32_t rank_in_cell(Slots slots, uint32_t c0, uint32_t c1, uint32_t mine) {
  uint32_t rank = 0, q = c0;
  for (; q + 4u <= c1; q += 4u) {  // four independent loads in flight
    const uint32_t a0 = slots(q), a1 = slots(q + 1), a2 = slots(q + 2), a3 = slots(q + 3);
    rank += (a0 < mine ? 1u : 0u) + (a1 < mine ? 1u : 0u) + (a2 < mine ? 1u : 0u) + (a3 < mine ? 1u : 0u);
  }
  for (; q < c1; ++q) rank += (slots(q) < mine) ? 1u : 0u;
  return rank;
}

// The fields a move carries: particle s of the previous slot order goes to
// slot d of the new one with key k.  The destination arrays are distinct from
// the sources, so every field can be loaded before the first store, which
// keeps the loads in flight together (interleaved, each load would wait for
// the store before it).
struct Moved {
  float2 xp, x0, dn;
  uint32_t id;
};
__device__ __forceinline__ Moved load_moved(const Buffers &b, uint32_t s) {
  Moved m;
  m.xp = b.xp[s];
  m.x0 = b.x0[s];
  m.id = b.id[s];
  // the carried density, or (not stored) any in-range float2 slot array: an
  // unconditional load, which a branch would make wait at once
  m.dn = (b.densTo ? b.densFrom : b.xp)[s];
  return m;
}
__device__ __forceinline__ void store_moved(const Buffers &b, uint32_t d, uint32_t k, const Moved &m) {
  b.x1[d] = m.xp;
  b.prev[d] = m.x0;  // demo4.cpp:334 prevPosition = curPosition
  b.id2[d] = m.id;
  b.keys[d] = k;
  if (b.densTo) b.densTo[d] = m.dn;
}
__device__ __forceinline__ void place_move(const Buffers &b, uint32_t s, uint32_t d, uint32_t k) {
  store_moved(b, d, k, load_moved(b, s));
}

// K moves at once: every load of the K particles is issued before the first
// store.  Move u goes from slot s[u] to slot d[u] with key k[u]; d[u] =
// 0xFFFFFFFF moves nothing (its loads read slot s[u] = 0, always in range).
template <uint32_t K>
__device__ __forceinline__ void place_moves(const Buffers &b, const uint32_t (&s)[K], const uint32_t (&d)[K],
                                            const uint32_t (&k)[K]) {
  Moved m[K];
#pragma unroll
  for (uint32_t u = 0; u < K; ++u) m[u] = load_moved(b, s[u]);
#pragma unroll
  for (uint32_t u = 0; u < K; ++u)
    if (d[u] != 0xFFFFFFFFu) store_moved(b, d[u], k[u], m[u]);
}

// REC: the bucketing wrote a record per new slot (binRec), so each thread
// starts from one load instead of three dependent ones.
template <bool REC>
__global__ __launch_bounds__(kBlock) void k_rank_reorder(Buffers b, int n) {
  __shared__ uint32_t tileSlots[kRankTile];
  __shared__ uint32_t span[2];
  const int p0 = (int)(blockIdx.x * kBlock);
  const int p = p0 + (int)threadIdx.x;
  const bool live = p < n;
  uint32_t s = 0, key = 0, c0 = 0, c1 = 0;
  if (live) {
    if (REC) {
      const uint4 r = b.binRec[p];
      s = r.x, key = r.y, c0 = r.z, c1 = r.w;
    } else {
      s = b.tmpSlot[p];
      key = b.keyp[s];
      c0 = b.startNext[key];
      c1 = b.startNext[key + 1];
    }
  }
  // the moved fields depend only on the source slot: requested now, stored
  // after the rank (a thread past n loads slot 0 and stores nothing)
  const Moved mv = load_moved(b, live ? s : 0u);
  if (p == p0) span[0] = c0;
  if (p == min(p0 + kBlock, n) - 1) span[1] = c1;
  __syncthreads();
  const uint32_t lo = span[0], hi = span[1];
  const bool tiled = hi - lo <= kRankTile;
  if (tiled)
    for (uint32_t q = lo + threadIdx.x; q < hi; q += kBlock) tileSlots[q - lo] = b.tmpSlot[q];
  __syncthreads();
  if (!live) return;
  const uint32_t rank = tiled ? rank_in_cell([&](uint32_t q) { return tileSlots[q - lo]; }, c0, c1, s)
                              : rank_in_cell([&](uint32_t q) { return b.tmpSlot[q]; }, c0, c1, s);
  store_moved(b, c0 + rank, key, mv);
}

// ---------------------------------------------------------------------------
// 3'. Placement: bucket, order and move in one pass (large grids, every
// particle binned by the last step).
//
// Workgroup g produces the new slots [g*kPlaceChunk, (g+1)*kPlaceChunk): the
// cells [cf, cl] that cover them (the chunk table the cell scan writes).  Its
// sources are the particles whose new cell lies in [cf, cl]: those whose cell
// of the last binning is 3x3-adjacent to it sit in three slot windows of the
// last binning (the rows below, at and above [cf, cl], one cell wider on each
// side), the others are in the far list.  They are bucketed by new cell in an
// LDS histogram (the bucket of cell c starts at its new cell start), each is
// ranked by its slot among its cell's bucket (the canonical order: a stable
// counting sort of the slot order), and the ones whose rank lands in the
// workgroup's range are moved.  A chunk whose cells hold more than kPlaceCap
// particles ranks by scanning its sources instead (correct, slow; only for
// cells of thousands of particles).
// ---------------------------------------------------------------------------

// Slot ranges, in the last binning, of the rows below, at and above the cells
// [cf, cl] of scene sc (one cell wider on each side), merged in key order.
// The six bounds are loaded together; absent windows are empty.
struct PlaceWindows {
  uint32_t ws[3], we[3];
};

__device__ __forceinline__ PlaceWindows place_windows(const Buffers &b, const DevScene &sc, uint32_t cf,
                                                      uint32_t cl) {
  const int64_t base = sc.cellBase, last = base + (int64_t)sc.nx * sc.ny - 1;
  const int64_t a = max((int64_t)cf, base), e = min((int64_t)cl, last);
  int64_t wa[3] = {a, a, a}, wb[3] = {e, e, e};
  int m = 0;
  for (int dy = -1; dy <= 1; ++dy) {
    const int64_t lo = max(a + (int64_t)dy * sc.nx - 1, base), hi = min(e + (int64_t)dy * sc.nx + 1, last);
    if (lo > hi) continue;
    if (m > 0 && lo <= wb[m - 1] + 1) {
      wb[m - 1] = max(wb[m - 1], hi);
    } else {
      wa[m] = lo;
      wb[m] = hi;
      ++m;
    }
  }
  PlaceWindows pw;
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    pw.ws[w] = b.startCur[wa[w]];
    pw.we[w] = b.startCur[wb[w] + 1];
  }
#pragma unroll
  for (int w = 0; w < 3; ++w)
    if (w >= m) pw.we[w] = pw.ws[w];
  return pw;
}

// Calls f(s, key) for every slot s of the windows whose key is in [cf, cl]
// (far movers carry the flag and fall outside).  The windows are walked as one
// range, each thread taking a strided share with kU key loads in flight.
template <uint32_t kU, typename F>
__device__ __forceinline__ void for_each_window_source(const Buffers &b, const PlaceWindows &pw, uint32_t cf,
                                                       uint32_t cl, F &&f) {
  const uint32_t o1 = pw.we[0] - pw.ws[0], o2 = o1 + (pw.we[1] - pw.ws[1]), total = o2 + (pw.we[2] - pw.ws[2]);
  for (uint32_t q0 = threadIdx.x; q0 < total; q0 += kU * kBlock) {
    uint32_t sl[kU], k[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      const uint32_t at = min(q0 + u * kBlock, total - 1u);
      sl[u] = at >= o2 ? pw.ws[2] + (at - o2) : (at >= o1 ? pw.ws[1] + (at - o1) : pw.ws[0] + at);
      k[u] = b.keyp[sl[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u)
      if (q0 + u * kBlock < total && k[u] >= cf && k[u] <= cl) f(sl[u], k[u]);
  }
}

// Calls f(s, key) for every particle whose new cell key is in [cf, cl]; each
// thread takes a strided share.  Window slots of far movers are skipped (they
// come from the far list).  pw0 holds the windows of a one-scene launch.
template <bool MULTI, uint32_t kU, typename F>
__device__ __forceinline__ void for_each_place_source(const Buffers &b, const PlaceWindows &pw0, uint32_t cf,
                                                      uint32_t cl, uint32_t farN, F &&f) {
  if (MULTI) {
    const int sFirst = scene_index(b, cf), sLast = scene_index(b, cl);
    for (int si = sFirst; si <= sLast; ++si)
      for_each_window_source<kU>(b, place_windows(b, b.scenes[si], cf, cl), cf, cl, f);
  } else {
    for_each_window_source<kU>(b, pw0, cf, cl, f);
  }
  for (uint32_t q = threadIdx.x; q < farN; q += kBlock) {
    const uint32_t s = b.farList[q];
    const uint32_t k = b.keyp[s] & ~kFarFlag;
    if (k >= cf && k <= cl) f(s, k);
  }
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_place(Buffers b, DevScene s0, int n, uint32_t cells) {
  __shared__ uint32_t bucket[kPlaceCap];   // source slot per new slot of [lo, hi)
  __shared__ uint32_t bkey[kPlaceCap];     // its new cell
  __shared__ uint32_t fillc[kPlaceCap];    // arrivals per cell, at the cell's first bucket entry
  __shared__ uint32_t cstart[kPlaceCap + 1];  // new cell starts of [cf, cl + 1]
  const uint32_t g0 = blockIdx.x * kPlaceChunk, g1 = min(g0 + kPlaceChunk, (uint32_t)n);
  const uint32_t cf = b.chunkCells[2 * blockIdx.x], cl = b.chunkCells[2 * blockIdx.x + 1];
  // every bound the workgroup needs is requested before the first is used
  PlaceWindows pw0{};
  if (!MULTI) pw0 = place_windows(b, s0, cf, cl);
  const uint32_t lo = b.startNext[cf], hi = b.startNext[cl + 1];
  const uint32_t farN = min(*b.farCount, (uint32_t)n);
  const bool staged = cl - cf + 2u <= kPlaceCap + 1u;
  if (staged)
    for (uint32_t q = threadIdx.x; q < cl - cf + 2u; q += kBlock) cstart[q] = b.startNext[cf + q];
  if (hi - lo <= kPlaceCap) {
    for (uint32_t q = threadIdx.x; q < hi - lo; q += kBlock) {
      fillc[q] = 0u;
      bucket[q] = 0xFFFFFFFFu;
    }
    __syncthreads();
    auto start_of = [&](uint32_t k) { return staged ? cstart[k - cf] : b.startNext[k]; };
    for_each_place_source<MULTI, 16>(b, pw0, cf, cl, farN, [&](uint32_t s, uint32_t k) {
      const uint32_t c0 = start_of(k), count = start_of(k + 1) - c0;
      const uint32_t a = atomicAdd(&fillc[c0 - lo], 1u);  // LDS histogram: arrival in the cell's bucket
      if (a < count) {
        bucket[c0 - lo + a] = s;
        bkey[c0 - lo + a] = k;
      }
    });
    __syncthreads();
    // kMoves moves per thread at a time: their loads are issued before their
    // stores (a slot that moves nothing loads slot 0 and stores nothing)
    constexpr uint32_t kMoves = 4;
    for (uint32_t q0 = threadIdx.x; q0 < hi - lo; q0 += kMoves * kBlock) {
      uint32_t ms[kMoves], md[kMoves], mk[kMoves];
#pragma unroll
      for (uint32_t u = 0; u < kMoves; ++u) {
        const uint32_t q = q0 + u * kBlock;
        ms[u] = 0u;
        md[u] = 0xFFFFFFFFu;  // no move
        mk[u] = 0u;
        if (q >= hi - lo) continue;
        const uint32_t s = bucket[q];
        if (s >= (uint32_t)n) continue;  // never, when every particle is found once
        const uint32_t k = bkey[q];
        const uint32_t c0 = start_of(k), c1 = start_of(k + 1);
        if (c1 <= g0 || c0 >= g1) continue;  // a cell the neighbouring workgroup moves
        const uint32_t d = c0 + rank_in_cell([&](uint32_t e) { return bucket[e - lo]; }, c0, c1, s);
        if (d < g0 || d >= g1) continue;
        ms[u] = s;
        md[u] = d;
        mk[u] = k;
      }
      place_moves<kMoves>(b, ms, md, mk);
    }
    return;
  }
  // Oversized cells: rank each source against every source of its cell.
  for_each_place_source<MULTI, 1>(b, pw0, cf, cl, farN, [&](uint32_t s, uint32_t k) {
    const uint32_t c0 = b.startNext[k], c1 = b.startNext[k + 1];
    if (c1 <= g0 || c0 >= g1) return;
    uint32_t rank = 0;
    const int si = MULTI ? scene_index(b, k) : 0;
    const DevScene sc = MULTI ? b.scenes[si] : s0;
    const int64_t base = sc.cellBase, last = base + (int64_t)sc.nx * sc.ny - 1;
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t wlo = max((int64_t)k + (int64_t)dy * sc.nx - 1, base);
      const int64_t whi = min((int64_t)k + (int64_t)dy * sc.nx + 1, last);
      if (wlo > whi) continue;
      const uint32_t t1 = min(b.startCur[whi + 1], s);
      for (uint32_t t = b.startCur[wlo]; t < t1; ++t) rank += (b.keyp[t] == k) ? 1u : 0u;
    }
    for (uint32_t q = 0; q < farN; ++q) {
      const uint32_t t = b.farList[q];
      rank += (t < s && (b.keyp[t] & ~kFarFlag) == k) ? 1u : 0u;
    }
    const uint32_t d = c0 + rank;
    if (d >= g0 && d < g1) place_move(b, s, d, k);
  });
}

// Exclusive scan of the cell counts in one workgroup, for grids of at most
// kSmallGridCells cells: one launch instead of k_scan_reduce + k_scan_final,
// with the same cell starts, count reset and occupied-cell statistics
// (in cellStats[0..1]).
// The one-workgroup scan body: counts from LDS (hist, when the workgroup has
// just counted there) or from the global counters (reset for the next step).
// With ldsStarts the cell starts [0, cells] also go to LDS (which may be hist:
// every count is read before the first start is written).  Thread t takes the
// kItems consecutive cells from kItems * t: in LDS they sit at lds_cell(c),
// one padding word per 32 cells, so that the 32 lanes of a half-wave reading
// their k-th cell hit 32 different banks; in global memory a thread's cells are
// moved as 16-byte vectors.
__device__ __forceinline__ uint32_t lds_cell(uint32_t c) { return c + (c >> 5); }
constexpr uint32_t kSmallHistWords = kSmallGridCells + 1 + ((kSmallGridCells + 1) >> 5) + 1;

template <bool LDS_COUNTS>
__device__ __forceinline__ void scan_small_body(Buffers &b, uint32_t cells, const uint32_t *hist,
                                                uint32_t *ldsStarts = nullptr) {
  constexpr int kWaves = kSmallGridBlock / 64;
  constexpr uint32_t kItems = kSmallGridCells / kSmallGridBlock;  // consecutive cells per thread
  static_assert(kItems % 4 == 0, "cells per thread in 16-byte vectors");
  __shared__ uint32_t waveTot[kWaves], waveMin[kWaves], waveMax[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c0 = threadIdx.x * kItems;
  const bool whole = c0 + kItems <= cells;  // every cell of the thread exists
  uint32_t v[kItems];
  if (!LDS_COUNTS && whole) {
#pragma unroll
    for (uint32_t k = 0; k < kItems; k += 4) {
      const uint4 q = *reinterpret_cast<const uint4 *>(b.count + c0 + k);
      v[k] = q.x, v[k + 1] = q.y, v[k + 2] = q.z, v[k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k)
      v[k] = (c0 + k < cells) ? (LDS_COUNTS ? hist[lds_cell(c0 + k)] : b.count[c0 + k]) : 0u;
  }
  uint32_t sum = 0, mn = 0xFFFFFFFFu, mx = 0u;
#pragma unroll
  for (uint32_t k = 0; k < kItems; ++k) {
    sum += v[k];
    mx = max(mx, v[k]);
    if (v[k] > 0u) mn = min(mn, v[k]);
  }
  const uint32_t incl = wave_incl_scan(sum);
  if (lane == 63) waveTot[wave] = incl;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, d, 64));
  }
  if (lane == 0) waveMin[wave] = mn, waveMax[wave] = mx;
  __syncthreads();
  uint32_t run = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t t = waveTot[w];
    if (w < wave) run += t;
    tot += t;
  }
  run += incl - sum;
  uint32_t st[kItems];  // exclusive starts of the thread's cells
#pragma unroll
  for (uint32_t k = 0; k < kItems; ++k) {
    st[k] = run;
    run += v[k];
  }
  if (whole) {
#pragma unroll
    for (uint32_t k = 0; k < kItems; k += 4) {
      *reinterpret_cast<uint4 *>(b.startNext + c0 + k) = make_uint4(st[k], st[k + 1], st[k + 2], st[k + 3]);
      if (!LDS_COUNTS) *reinterpret_cast<uint4 *>(b.count + c0 + k) = make_uint4(0u, 0u, 0u, 0u);  // next step
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k) {
      if (c0 + k < cells) {
        b.startNext[c0 + k] = st[k];
        if (!LDS_COUNTS) b.count[c0 + k] = 0u;  // ready for the next step's counting
      }
    }
  }
  if (ldsStarts) {
#pragma unroll
    for (uint32_t k = 0; k < kItems; ++k)
      if (c0 + k < cells) ldsStarts[lds_cell(c0 + k)] = st[k];
  }
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; ++w) mn = min(mn, waveMin[w]), mx = max(mx, waveMax[w]);
    b.cellStats[0] = mn;
    b.cellStats[1] = mx;
    b.startNext[cells] = tot;
    if (ldsStarts) ldsStarts[lds_cell(cells)] = tot;
  }
}

__global__ __launch_bounds__(kSmallGridBlock) void k_scan_small(Buffers b, uint32_t cells) {
  scan_small_body<false>(b, cells, nullptr);
}

// Counting, scan and scatter of a small step in one workgroup: the cell
// counts are an LDS histogram (kSmallGridCells counters), each particle's
// arrival index in its cell is the LDS atomic's old value, the scan runs from
// LDS and leaves the cell starts there, and each particle is then bucketed at
// its cell start + arrival index: its source slot in tmpSlot, and in binRec
// the record (source slot, key, cell start, cell end) the rank pass starts
// from.  Used after the wide viscosity kernel (which leaves the counting to a
// later kernel) for steps of at most kSmallCountParticles particles: one
// launch instead of k_count_fill + k_scan_small + k_scatter, and no global
// atomics.
constexpr int kSmallItems = kSmallCountParticles / kSmallGridBlock;  // particles per thread, at most
__global__ __launch_bounds__(kSmallGridBlock) void k_count_scan_small(Buffers b, int n, uint32_t cells) {
  __shared__ uint32_t hist[kSmallHistWords];  // counts, then cell starts [0, cells], at lds_cell(c)
  for (uint32_t c = threadIdx.x; c < cells; c += kSmallGridBlock) hist[lds_cell(c)] = 0u;
  uint32_t key[kSmallItems], fill[kSmallItems];
#pragma unroll
  for (int u = 0; u < kSmallItems; ++u) {  // the key loads in flight together
    if (u * kSmallGridBlock >= n) break;    // workgroup-uniform
    key[u] = b.keyp[min((int)threadIdx.x + u * kSmallGridBlock, n - 1)];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kSmallItems; ++u) {
    if (u * kSmallGridBlock >= n) break;
    fill[u] = (int)threadIdx.x + u * kSmallGridBlock < n ? atomicAdd(&hist[lds_cell(key[u])], 1u) : 0u;  // arrival
  }
  __syncthreads();
  scan_small_body<true>(b, cells, hist, hist);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kSmallItems; ++u) {
    if (u * kSmallGridBlock >= n) break;
    const int i = (int)threadIdx.x + u * kSmallGridBlock;
    if (i < n) {
      const uint32_t c0 = hist[lds_cell(key[u])], c1 = hist[lds_cell(key[u] + 1)], pos = c0 + fill[u];
      b.tmpSlot[pos] = (uint32_t)i;
      b.binRec[pos] = make_uint4((uint32_t)i, key[u], c0, c1);
    }
  }
}

// ---------------------------------------------------------------------------
// 4. density + pressure
// ---------------------------------------------------------------------------
template <typename Idx>
__device__ __forceinline__ uint32_t density_sum(const uint32_t *start, const DevScene &sc, int cx, int cy,
                                                const float2 *X, Idx idx, const DevParams &p, float2 xi, float &d0,
                                                float &d1) {
  return for_each_pair(
      start, sc, cx, cy, X, idx, xi, p.h2,
      [&](uint32_t, float, float, float r2, bool use) {
        const float r = pair_sqrt(r2);                    // sph.h:471
        // term +0 gives terms +0: adding them leaves the sums unchanged (they
        // start at +0 and never hold -0)
        const float term = use ? 1.0f - r * p.invH : 0.0f;  // sph.h:472
        const float t2 = term * term, t3 = (term * term) * term;
        return make_float2(t2, t3);
      },
      [&](float2 w) {
        d0 = d0 + w.x;                                    // sph.h:473
        d1 = d1 + w.y;                                    // sph.h:474
      });
}

// Plans the tile of the workgroup at slots [i0, i0 + kBlock) of the new binning.
template <bool MULTI, int CAP>
__device__ __forceinline__ void plan_new(const Buffers &b, const DevScene &s0, int i0, int n, TilePlan &plan) {
  if (threadIdx.x == 0) {
    const uint32_t kA = b.keys[i0], kB = b.keys[min(i0 + kBlock, n) - 1];
    const DevScene sa = scene_of_slot<MULTI>(b, s0, kA);
    plan_tile<CAP>(b.startNext, sa, kA, kB, plan);
    if (MULTI && sa.cellBase != scene_of_slot<MULTI>(b, s0, kB).cellBase) plan.use = 0;
  }
  __syncthreads();
}

template <bool MULTI, int CAP>
__global__ __launch_bounds__(kBlock) void k_density(Buffers b, DevParams p, DevScene s0, int n) {
  __shared__ float2 tx[CAP > 0 ? CAP : 1];
  __shared__ TilePlan plan;
  const int i0 = (int)(blockIdx.x * kBlock);
  const int i = i0 + (int)threadIdx.x;
  plan_new<MULTI, CAP>(b, s0, i0, n, plan);
  if (CAP > 0 && plan.use) stage(plan, b.x1, tx);
  __syncthreads();
  uint32_t cand = 0;
  if (i < n) {
    const uint32_t key = b.keys[i];
    const DevScene sc = scene_of_slot<MULTI>(b, s0, key);
    const uint32_t local = key - sc.cellBase;
    const int cy = (int)(local / (uint32_t)sc.nx);
    const int cx = (int)local - cy * sc.nx;
    const float2 xi = b.x1[i];
    float d0 = 0.0f, d1 = 0.0f;
    if (CAP > 0 && plan.use)
      cand = density_sum(b.startNext, sc, cx, cy, tx, TileIndex(plan), p, xi, d0, d1);
    else
      cand = density_sum(b.startNext, sc, cx, cy, b.x1, SlotIndex(), p, xi, d0, d1);
    b.dens[i] = make_float2(d0, d1);
    b.pres[i] = make_float2(p.k * (d0 - p.rho0), p.kn * d1);  // sph.h:478-481
  }
  // candidate-count statistics (demo4.cpp:369-376): one min/max pair per
  // workgroup, reduced when the statistics are read.
  uint32_t mn = (i < n) ? cand : 0xFFFFFFFFu, mx = cand;
  block_min_max(mn, mx);
  if (threadIdx.x == 0) {
    b.candStats[2 * blockIdx.x] = mn;
    b.candStats[2 * blockIdx.x + 1] = mx;
  }
}

// ---------------------------------------------------------------------------
// 5. relaxation + collisions + velocity
// ---------------------------------------------------------------------------
// demo4.cpp:239-255 in Jacobi order: -delta_ij/2 from the particle's own visit
// and +delta_ji/2 from the neighbour's visit (n_ji = -n_ij).
template <typename Idx>
__device__ __forceinline__ void delta_sum(const uint32_t *start, const DevScene &sc, int cx, int cy, const float2 *X,
                                          const float2 *P, Idx idx, const DevParams &p, float2 xi, float2 pi,
                                          float &accx, float &accy) {
  for_each_pair(
      start, sc, cx, cy, X, idx, xi, p.h2,
      [&](uint32_t j, float Rx, float Ry, float r2, bool use) {
        const float2 pj = P[j];
        float r, inv;
        pair_length(r2, r, inv);                          // sph.h:489-490 Vec2Normalize (length 0 -> 1)
        const float nX = Rx * inv, nY = Ry * inv;
        const float term = 1.0f - r * p.invH;             // sph.h:491
        const float tt = term * term;
        // a zero displacement times the finite normal and 0.5 gives +0 or -0,
        // and subtracting either leaves the sums unchanged (they start at +0
        // and never hold -0)
        const float dij = use ? p.dt2 * (pi.x * term + pi.y * tt) : 0.0f;  // sph.h:492 with P_i
        const float dji = use ? p.dt2 * (pj.x * term + pj.y * tt) : 0.0f;  // sph.h:492 with P_j (j's visit)
        return make_float4((dij * nX) * 0.5f, (dij * nY) * 0.5f, (dji * nX) * 0.5f, (dji * nY) * 0.5f);
      },
      [&](float4 d) {
        accx = accx - d.x;                                // demo4.cpp:251 dx -= delta * 0.5f
        accy = accy - d.y;
        accx = accx - d.z;                                // demo4.cpp:250 (j's visit)
        accy = accy - d.w;
      });
}

template <bool MULTI, int CAP>
__global__ __launch_bounds__(kBlock) void k_delta(Buffers b, DevParams p, DevScene s0, int n) {
  __shared__ float2 tx[CAP > 0 ? CAP : 1];
  __shared__ TilePlan plan;
  const int i0 = (int)(blockIdx.x * kBlock);
  const int i = i0 + (int)threadIdx.x;
  plan_new<MULTI, CAP>(b, s0, i0, n, plan);
  if (CAP > 0 && plan.use) {
    stage(plan, b.x1, tx);
  }
  __syncthreads();
  // Lanes past n stay active (collide_wave reduces over the whole wave) but
  // read and write nothing.
  const bool live = i < n;
  DevScene sc = s0;
  float x = 0.0f, y = 0.0f;
  if (live) {
    const uint32_t key = b.keys[i];
    sc = scene_of_slot<MULTI>(b, s0, key);
    const uint32_t local = key - sc.cellBase;
    const int cy = (int)(local / (uint32_t)sc.nx);
    const int cx = (int)local - cy * sc.nx;
    const float2 xi = b.x1[i];
    const float2 pi = b.pres[i];
    float accx = 0.0f, accy = 0.0f;
    if (CAP > 0 && plan.use)
      delta_sum(b.startNext, sc, cx, cy, tx, b.pres, TileIndex(plan), p, xi, pi, accx, accy);
    else
      delta_sum(b.startNext, sc, cx, cy, b.x1, b.pres, SlotIndex(), p, xi, pi, accx, accy);
    if (b.dxOut) b.dxOut[i] = make_float2(accx, accy);
    x = accx + xi.x;                                      // demo4.cpp:253
    y = accy + xi.y;
  }
  collide_wave<MULTI>(b, p, sc, x, y, live);              // demo4.cpp:411-444
  if (live) {
    const float2 pv = b.prev[i];
    b.x0[i] = make_float2(x, y);
    b.v[i] = make_float2((x - pv.x) * p.invDt, (y - pv.y) * p.invDt);  // demo4.cpp:449
  }
}

// ---------------------------------------------------------------------------
// Wide variants for small particle counts: one wave per particle, the pair
// terms computed across the wave and accumulated in candidate order
// (for_each_pair_wide), so results equal the one-lane-per-particle kernels'
// bit for bit.  Used when a step has at most nbs_tuning.wide_max_particles
// particles: there, one lane per particle leaves most of the chip idle and
// each particle's serial candidate walk sets the kernel time.
// ---------------------------------------------------------------------------
constexpr int kWidePerBlock = kBlock / 64;

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_visc_predict_wide(Buffers b, DevParams p, DevScene s0, int n, int nBinned,
                                                              int lists) {
  const int i = (int)(blockIdx.x * kWidePerBlock + (threadIdx.x >> 6));
  if (i >= n) return;  // whole wave
  const float2 xi = b.x0[i];
  const float2 vi = b.v[i];
  const bool binned = i < nBinned;
  float aX, aY;
  DevScene sc;
  if (binned) {
    aX = p.gex + 0.0f;                                   // demo4.cpp:306 (own acceleration is zero)
    aY = p.gey + 0.0f;
    sc = scene_of_slot<MULTI>(b, s0, b.keys[i]);
  } else {
    const float2 a = b.acc[i];
    aX = p.gex + a.x;
    aY = p.gey + a.y;
    sc = MULTI ? b.scenes[b.scn[i]] : s0;
  }
  float vx = (aX * p.dt) + vi.x;                         // demo4.cpp:307
  float vy = (aY * p.dt) + vi.y;
  if (lists && binned) {
    const uint32_t local = b.keys[i] - sc.cellBase;
    const int cy = (int)(local / (uint32_t)sc.nx);
    const int cx = (int)local - cy * sc.nx;
    __shared__ float lds[kWidePerBlock][65 * 2];
    float accx = 0.0f, accy = 0.0f;
    // per pair: accx -= ix, accy -= iy (own visit), then again (neighbour's visit)
    for_each_pair_wide<2, 2, true>(
        b.startCur, sc, cx, cy, b.x0, SlotIndex(), xi, p.h2, b.v, lds[threadIdx.x >> 6], accx, accy,
        [&](uint32_t, float Rx, float Ry, float r2, float2 vj, float *t) {
          const float vjx = p.gdtx + vj.x, vjy = p.gdty + vj.y;
          float r, inv;
          pair_length(r2, r, inv);                              // sph.h:503, 505
          const float qq = r * p.invH;                          // sph.h:504
          const float nX = Rx * inv, nY = Ry * inv;
          const float u = (vx - vjx) * nX + (vy - vjy) * nY;    // sph.h:506
          if (!(u > 0.0f)) return false;
          const float f = (1.0f - qq) * (p.sigma * u + p.beta * (u * u));  // sph.h:508
          t[0] = ((f * nX) * 0.5f) * p.dt;                      // demo4.cpp:233
          t[1] = ((f * nY) * 0.5f) * p.dt;
          return true;
        });
    vx = accx + vx;
    vy = accy + vy;
  }
  if ((threadIdx.x & 63) == 0) {
    const float xpx = (vx * p.dt) + xi.x;                // demo4.cpp:334-335
    const float xpy = (vy * p.dt) + xi.y;
    b.xp[i] = make_float2(xpx, xpy);
    const int cx = cell_coord(xpx, sc.halfW, sc.cellSize, sc.nx);
    const int cy = cell_coord(xpy, sc.halfH, sc.cellSize, sc.ny);
    const uint32_t key = sc.cellBase + (uint32_t)(cy * sc.nx + cx);
    b.keyp[i] = key;  // counted by k_count_fill
  }
}

// Cell counting after the wide viscosity kernel: one lane per particle, so
// particles of one cell share wave-aggregated atomics instead of one atomic
// per particle on the same counter.
__global__ __launch_bounds__(kBlock) void k_count_fill(Buffers b, int n) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n) return;
  b.fill[i] = wave_aggregated_fill(b.count, b.keyp[i]);  // arrival slot; the rank pass orders the cell
}

// Candidate-count statistics: one min/max pair per workgroup (kWidePerBlock
// particles), slot blockIdx.x.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_density_wide(Buffers b, DevParams p, DevScene s0, int n) {
  __shared__ uint32_t shCand[kWidePerBlock];
  __shared__ float lds[kWidePerBlock][65 * 2];
  const int w = (int)(threadIdx.x >> 6);
  const int i = (int)(blockIdx.x * kWidePerBlock + w);
  if (i < n) {  // whole wave
    const uint32_t key = b.keys[i];
    const DevScene sc = scene_of_slot<MULTI>(b, s0, key);
    const uint32_t local = key - sc.cellBase;
    const int cy = (int)(local / (uint32_t)sc.nx);
    const int cx = (int)local - cy * sc.nx;
    const float2 xi = b.x1[i];
    float d0 = 0.0f, d1 = 0.0f;
    // per pair: d0 += term^2, d1 += term^3 (sph.h:473-474)
    const uint32_t cand = for_each_pair_wide<2, 1, false>(
        b.startNext, sc, cx, cy, b.x1, SlotIndex(), xi, p.h2, nullptr, lds[w], d0, d1,
        [&](uint32_t, float, float, float r2, float2, float *t) {
          const float r = pair_sqrt(r2);                        // sph.h:471
          const float term = 1.0f - r * p.invH;                 // sph.h:472
          t[0] = term * term;
          t[1] = (term * term) * term;
          return true;
        });
    if ((threadIdx.x & 63) == 0) {
      b.dens[i] = make_float2(d0, d1);
      b.pres[i] = make_float2(p.k * (d0 - p.rho0), p.kn * d1); // sph.h:478-481
      shCand[w] = cand;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    for (int k = 0; k < kWidePerBlock && (int)(blockIdx.x * kWidePerBlock) + k < n; ++k) {
      mn = min(mn, shCand[k]);
      mx = max(mx, shCand[k]);
    }
    b.candStats[2 * blockIdx.x] = mn;
    b.candStats[2 * blockIdx.x + 1] = mx;
  }
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_delta_wide(Buffers b, DevParams p, DevScene s0, int n) {
  const int i = (int)(blockIdx.x * kWidePerBlock + (threadIdx.x >> 6));
  if (i >= n) return;  // whole wave
  const uint32_t key = b.keys[i];
  const DevScene sc = scene_of_slot<MULTI>(b, s0, key);
  const uint32_t local = key - sc.cellBase;
  const int cy = (int)(local / (uint32_t)sc.nx);
  const int cx = (int)local - cy * sc.nx;
  const float2 xi = b.x1[i];
  const float2 pi = b.pres[i];
  __shared__ float lds[kWidePerBlock][65 * 4];
  float accx = 0.0f, accy = 0.0f;
  // per pair: accx -= t0, accy -= t1 (demo4.cpp:251), then accx -= t2, accy -= t3
  // (demo4.cpp:250, j's visit)
  for_each_pair_wide<4, 2, true>(
      b.startNext, sc, cx, cy, b.x1, SlotIndex(), xi, p.h2, b.pres, lds[threadIdx.x >> 6], accx, accy,
      [&](uint32_t, float Rx, float Ry, float r2, float2 pj, float *t) {
        float r, inv;
        pair_length(r2, r, inv);                                // sph.h:489-490
        const float nX = Rx * inv, nY = Ry * inv;
        const float term = 1.0f - r * p.invH;                   // sph.h:491
        const float tt = term * term;
        const float dij = p.dt2 * (pi.x * term + pi.y * tt);    // sph.h:492 with P_i
        const float dji = p.dt2 * (pj.x * term + pj.y * tt);    // sph.h:492 with P_j
        t[0] = (dij * nX) * 0.5f;
        t[1] = (dij * nY) * 0.5f;
        t[2] = (dji * nX) * 0.5f;
        t[3] = (dji * nY) * 0.5f;
        return true;
      });
  // Every lane holds the same sums: the wave walks the bodies once for lane 0.
  float x = accx + xi.x;                                        // demo4.cpp:253
  float y = accy + xi.y;
  collide_wave<MULTI>(b, p, sc, x, y, (threadIdx.x & 63) == 0);  // demo4.cpp:411-444
  if ((threadIdx.x & 63) == 0) {
    if (b.dxOut) b.dxOut[i] = make_float2(accx, accy);
    const float2 pv = b.prev[i];
    b.x0[i] = make_float2(x, y);
    b.v[i] = make_float2((x - pv.x) * p.invDt, (y - pv.y) * p.invDt);  // demo4.cpp:449
  }
}

// ---------------------------------------------------------------------------
// State readback in particle-index order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_gather_state(Buffers b, DevScene s0, int n, int nBinned, float2 *x,
                                                         float2 *prev, float2 *v, float2 *dens, float2 *pres,
                                                         uint32_t *key, int32_t *scene) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n) return;
  const uint32_t id = b.id[i];
  if (x) x[id] = b.x0[i];
  if (prev) prev[id] = b.prev[i];
  if (v) v[id] = b.v[i];
  if (dens) dens[id] = b.dens[i];
  if (pres) pres[id] = b.pres[i];
  if (key) key[id] = b.keys[i];
  if (scene) {
    int s = 0;
    if (b.nScenes > 1) s = (i < nBinned) ? scene_index(b, b.keys[i]) : (int)b.scn[i];
    scene[id] = s;
  }
}

// sph.h:683-695 SPHGetParticleColor, in particle-index order.  The clamps
// keep std::min / std::max's comparison order (NaN passes through).
__global__ __launch_bounds__(kBlock) void k_colors(Buffers b, float rho0, int n, float4 *rgba) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n) return;
  const float2 d = b.dens[i], pr = b.pres[i], v = b.v[i];
  const float r = pr.x / (-10.0f);
  const float g = d.x / rho0;
  const float bl = sqrtf(v.x * v.x + v.y * v.y) / 10.0f;  // vecmath.h:267-270 Vec2Length
  auto clamp01 = [](float c) {
    const float m = (1.0f < c) ? 1.0f : c;  // std::min(c, 1.0f)
    return (m < 0.0f) ? 0.0f : m;           // std::max(m, 0.0f)
  };
  rgba[b.id[i]] = make_float4(clamp01(r), clamp01(g), clamp01(bl), 1.0f);
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *p, uint32_t value, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = value;
}

// ---------------------------------------------------------------------------
// FP64 diagnostics: per-workgroup partial sums over fixed slot ranges, then one
// workgroup combines them in workgroup order.
// ---------------------------------------------------------------------------
constexpr int kDiag = kDiagWidth;
constexpr int kDiagItems = kDiagPerBlock / kBlock;

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
  return t;
}
__device__ __forceinline__ double block_max_d(double v, double *sh) {
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = sh[0];
  for (int w = 1; w < kBlock / 64; ++w) t = fmax(t, sh[w]);
  return t;
}

__global__ __launch_bounds__(kBlock) void k_diag_partial(Buffers b, DevParams p, int n, double *partials) {
  __shared__ double sh[kBlock / 64];
  double acc[kDiag];
#pragma unroll
  for (int k = 0; k < kDiag; ++k) acc[k] = 0.0;
  acc[4] = -1.0e300;
  const double rho0 = (double)p.rho0;
  const int base = blockIdx.x * kBlock * kDiagItems;
  for (int k = 0; k < kDiagItems; ++k) {
    const int i = base + k * kBlock + threadIdx.x;
    if (i >= n) break;
    const float2 v = b.v[i], x = b.x0[i], d = b.dens[i];
    const bool finite = isfinite(v.x) && isfinite(v.y) && isfinite(x.x) && isfinite(x.y);
    if (!finite) { acc[9] += 1.0; continue; }
    const double vx = v.x, vy = v.y, rho = d.x;
    acc[0] += vx;
    acc[1] += vy;
    acc[2] += vx * vx + vy * vy;
    acc[3] += rho;
    acc[4] = fmax(acc[4], rho);
    acc[5] += fabs(rho - rho0);
    acc[6] += (rho - rho0) * (rho - rho0);
    acc[7] += (double)x.x;
    acc[8] += (double)x.y;
    acc[10] += 1.0;
  }
  double out[kDiag];
#pragma unroll
  for (int k = 0; k < kDiag; ++k) out[k] = (k == 4) ? block_max_d(acc[k], sh) : block_sum_d(acc[k], sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kDiag; ++k) partials[(size_t)blockIdx.x * kDiag + k] = out[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_diag_final(int nblocks, const double *partials, double *out) {
  if (threadIdx.x != 0) return;
  double r[kDiag];
  for (int k = 0; k < kDiag; ++k) r[k] = 0.0;
  r[4] = -1.0e300;
  for (int bI = 0; bI < nblocks; ++bI) {
    for (int k = 0; k < kDiag; ++k) {
      const double v = partials[(size_t)bI * kDiag + k];
      r[k] = (k == 4) ? fmax(r[k], v) : r[k] + v;
    }
  }
  for (int k = 0; k < kDiag; ++k) out[k] = r[k];
}

// Combines the partials like k_diag_final and writes the record.
__global__ __launch_bounds__(kBlock) void k_diag_record(int nblocks, const double *partials, uint64_t n, double rho0,
                                                        nbs_diagnostics *out) {
  if (threadIdx.x != 0) return;
  double r[kDiag];
  for (int k = 0; k < kDiag; ++k) r[k] = 0.0;
  r[4] = -1.0e300;
  for (int bI = 0; bI < nblocks; ++bI) {
    for (int k = 0; k < kDiag; ++k) {
      const double v = partials[(size_t)bI * kDiag + k];
      r[k] = (k == 4) ? fmax(r[k], v) : r[k] + v;
    }
  }
  nbs_diagnostics d;
  diagnostics_record(r, n, rho0, &d);
  *out = d;
}

int diagnostics_blocks(int n) { return (n + kBlock * kDiagItems - 1) / (kBlock * kDiagItems); }

// ---------------------------------------------------------------------------
// Arithmetic self-test
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_selftest_math(const float *a, const float *b, uint64_t n, float *out) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  out[6 * i + 0] = sqrtf(x);
  out[6 * i + 1] = 1.0f / x;
  out[6 * i + 2] = x / y;
  out[6 * i + 3] = (float)cell_coord(x, y, 1.0f, 1 << 20);
  float r, inv;
  pair_length(x, r, inv);
  out[6 * i + 4] = pair_sqrt(x);
  out[6 * i + 5] = inv;
}

// Plain device copy (the bandwidth probe and counter calibration).
__global__ __launch_bounds__(kBlock) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_unbin(const Buffers &b, int nBinned, hipStream_t s) {
  if (nBinned <= 0) return;
  hipLaunchKernelGGL(k_unbin, dim3(grid_for(nBinned)), dim3(kBlock), 0, s, b, nBinned);
}

template <int CAP>
static void visc_cap(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int nBinned, int lists,
                     hipStream_t s) {
  if (b.nScenes > 1)
    hipLaunchKernelGGL((k_visc_predict<true, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n, nBinned, lists);
  else
    hipLaunchKernelGGL((k_visc_predict<false, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n, nBinned,
                       lists);
}

static unsigned wide_grid(int n) { return (unsigned)((n + kWidePerBlock - 1) / kWidePerBlock); }

uint32_t density_stat_slots(int n, int cap) { return cap == kWideCap ? wide_grid(n) : grid_for(n); }

bool count_in_grid(int n, uint32_t cells, int viscCap) {
  return viscCap == kWideCap && cells <= kSmallGridCells && n <= kSmallCountParticles;
}

void launch_visc_predict(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int nBinned, int lists,
                         int cap, uint32_t cells, hipStream_t s) {
  if (n <= 0) return;
  if (cap == kWideCap) {
    if (b.nScenes > 1)
      hipLaunchKernelGGL(k_visc_predict_wide<true>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n, nBinned, lists);
    else
      hipLaunchKernelGGL(k_visc_predict_wide<false>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n, nBinned,
                         lists);
    if (!count_in_grid(n, cells, cap)) hipLaunchKernelGGL(k_count_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, b, n);
    return;
  }
  switch (cap) {
    case 0: visc_cap<0>(b, p, s0, n, nBinned, lists, s); break;
    case 1024: visc_cap<1024>(b, p, s0, n, nBinned, lists, s); break;
    case 2048: visc_cap<2048>(b, p, s0, n, nBinned, lists, s); break;
    default: visc_cap<4096>(b, p, s0, n, nBinned, lists, s); break;
  }
}

void launch_scan(const Buffers &b, uint32_t cells, int n, hipStream_t s) {
  const uint32_t nb = (cells + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, b, cells);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kBlock), 0, s, b, cells, (uint32_t)n);
}

void launch_scatter_reorder(const Buffers &b, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, b, n);
  hipLaunchKernelGGL(k_rank_reorder<false>, dim3(grid_for(n)), dim3(kBlock), 0, s, b, n);
}

bool place_binning(uint32_t cells, int viscCap, bool allBinned) {
  return cells > kSmallGridCells && viscCap != kWideCap && allBinned;
}

uint32_t launch_grid(const Buffers &b, const DevScene &s0, uint32_t cells, int n, int viscCap, bool place,
                     hipStream_t s) {
  uint32_t slots = 1u;
  if (place) {
    launch_scan(b, cells, n, s);
    if (n > 0) {
      const unsigned chunks = (unsigned)(((uint64_t)n + kPlaceChunk - 1) / kPlaceChunk);
      if (b.nScenes > 1)
        hipLaunchKernelGGL(k_place<true>, dim3(chunks), dim3(kBlock), 0, s, b, s0, n, cells);
      else
        hipLaunchKernelGGL(k_place<false>, dim3(chunks), dim3(kBlock), 0, s, b, s0, n, cells);
    }
    return (cells + kScanChunk - 1) / kScanChunk;
  }
  if (count_in_grid(n, cells, viscCap)) {
    // counted, scanned and bucketed by one workgroup; the rank pass reads its records
    hipLaunchKernelGGL(k_count_scan_small, dim3(1), dim3(kSmallGridBlock), 0, s, b, n, cells);
    if (n > 0) hipLaunchKernelGGL(k_rank_reorder<true>, dim3(grid_for(n)), dim3(kBlock), 0, s, b, n);
    return slots;
  } else if (cells <= kSmallGridCells) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallGridBlock), 0, s, b, cells);
  } else {
    Buffers bs = b;
    bs.chunkCells = nullptr;
    launch_scan(bs, cells, n, s);
    slots = (cells + kScanChunk - 1) / kScanChunk;
  }
  launch_scatter_reorder(b, n, s);
  return slots;
}

template <int CAP>
static void density_cap(const Buffers &b, const DevParams &p, const DevScene &s0, int n, hipStream_t s) {
  if (b.nScenes > 1)
    hipLaunchKernelGGL((k_density<true, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n);
  else
    hipLaunchKernelGGL((k_density<false, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n);
}

void launch_density(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int cap, hipStream_t s) {
  if (n <= 0) return;
  if (cap == kWideCap) {
    if (b.nScenes > 1)
      hipLaunchKernelGGL(k_density_wide<true>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n);
    else
      hipLaunchKernelGGL(k_density_wide<false>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n);
    return;
  }
  switch (cap) {
    case 0: density_cap<0>(b, p, s0, n, s); break;
    case 1024: density_cap<1024>(b, p, s0, n, s); break;
    case 2048: density_cap<2048>(b, p, s0, n, s); break;
    default: density_cap<4096>(b, p, s0, n, s); break;
  }
}

template <int CAP>
static void delta_cap(const Buffers &b, const DevParams &p, const DevScene &s0, int n, hipStream_t s) {
  if (b.nScenes > 1)
    hipLaunchKernelGGL((k_delta<true, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n);
  else
    hipLaunchKernelGGL((k_delta<false, CAP>), dim3(grid_for(n)), dim3(kBlock), 0, s, b, p, s0, n);
}

void launch_delta(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int cap, hipStream_t s) {
  if (n <= 0) return;
  if (cap == kWideCap) {
    if (b.nScenes > 1)
      hipLaunchKernelGGL(k_delta_wide<true>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n);
    else
      hipLaunchKernelGGL(k_delta_wide<false>, dim3(wide_grid(n)), dim3(kBlock), 0, s, b, p, s0, n);
    return;
  }
  switch (cap) {
    case 0: delta_cap<0>(b, p, s0, n, s); break;
    case 1024: delta_cap<1024>(b, p, s0, n, s); break;
    case 2048: delta_cap<2048>(b, p, s0, n, s); break;
    default: delta_cap<4096>(b, p, s0, n, s); break;
  }
}

void launch_gather_state(const Buffers &b, const DevScene &s0, int n, int nBinned, float2 *x, float2 *prev, float2 *v,
                         float2 *dens, float2 *pres, uint32_t *key, int32_t *scene, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_state, dim3(grid_for(n)), dim3(kBlock), 0, s, b, s0, n, nBinned, x, prev, v, dens, pres,
                     key, scene);
}

void launch_colors(const Buffers &b, float rho0, int n, float4 *rgba, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_colors, dim3(grid_for(n)), dim3(kBlock), 0, s, b, rho0, n, rgba);
}

void launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, value, n);
}

void launch_diagnostics(const Buffers &b, const DevParams &p, int n, double *partials, double *out, hipStream_t s) {
  const int nb = diagnostics_blocks(n);
  if (nb > 0) hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(kBlock), 0, s, b, p, n, partials);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, s, nb, partials, out);
}

void launch_diagnostics_record(const float2 *x, const float2 *v, const float2 *dens, int n, float rho0,
                               double *partials, nbs_diagnostics *out, hipStream_t s) {
  Buffers b{};
  b.x0 = const_cast<float2 *>(x);
  b.v = const_cast<float2 *>(v);
  b.dens = const_cast<float2 *>(dens);
  DevParams p{};
  p.rho0 = rho0;
  const int nb = diagnostics_blocks(n);
  if (nb > 0) hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(kBlock), 0, s, b, p, n, partials);
  hipLaunchKernelGGL(k_diag_record, dim3(1), dim3(kBlock), 0, s, nb, (const double *)partials, (uint64_t)n,
                     (double)rho0, out);
}

void launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / 16;
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<uint64_t>((n + kBlock - 1) / kBlock, 256ull * 64);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, s, (const float4 *)src, (float4 *)dst, n);
}

void launch_selftest_math(const float *a, const float *b, uint64_t n, float *out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_selftest_math, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, n, out);
}

}  // namespace nbs
