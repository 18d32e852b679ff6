// Device helpers shared by the SPH step kernels and the learned relaxation:
// cell keys, scene lookup, correctly rounded arithmetic, boundary collisions,
// candidate row ranges.
#pragma once

#include <climits>

#include "nbs_kernels.hpp"

namespace nbs {

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------

// sph.h:450-463 per axis.  Out-of-range values are clamped in floating point so
// NaN and values beyond the int range are defined (NaN -> 0).
__device__ __forceinline__ int cell_coord(float p, float half, float cs, int n) {
  const float f = (p + half) / cs;
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)n) return n - 1;
  return (int)f;
}

__device__ __forceinline__ int scene_of_key(const uint32_t *base, int n, uint32_t key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Scene of cell `key`: one load from the per-cell table when the launch has
// one, else a binary search of the scenes' first cells.
__device__ __forceinline__ int scene_index(const Buffers &b, uint32_t key) {
  return b.cellScene ? (int)b.cellScene[key] : scene_of_key(b.sceneBase, b.nScenes, key);
}

template <bool MULTI>
__device__ __forceinline__ DevScene scene_of_slot(const Buffers &b, const DevScene &s0, uint32_t key) {
  if (!MULTI) return s0;
  return b.scenes[scene_index(b, key)];
}

// Correctly rounded FP32 square root and reciprocal of the pair kernels'
// distances, in fewer instructions than the general IEEE lowering and without
// branches.  Operands: 0 <= r2 < h^2 <= 2^80 (kernel_height <= 2^40 is enforced
// when parameters are set, and only pairs with r2 < h^2 reach these).
//
// v_sqrt_f32 and v_rcp_f32 are within 1 ulp for normal operands; the result is
// then rounded correctly:
//  - sqrt_near: of s and its neighbouring floats s - ulp, s + ulp, the residual
//    x - s'*s (one FMA each) picks the correctly rounded one; exact for x == 0
//    and x in [2^-96, 2^100];
//  - rcp_near: one Newton step with FMA (Markstein), e = 1 - x*y exactly,
//    y' = y + e*y; exact for x in [2^-60, 2^60].
// r2 below 2^-96 (subnormals included) is scaled by 2^64 first.  Scaling by a
// power of two is exact and commutes with correct rounding, so
//   sqrt(r2) = sqrt(r2 * 2^64) * 2^-32  and  1/sqrt(r2) = (1/sqrt(r2 * 2^64)) * 2^32
// with every intermediate inside the ranges above.  The device arithmetic
// self-test compares both with the correctly rounded results for every
// mantissa of the subnormals and of binades across the whole operand range.
__device__ __forceinline__ float sqrt_near(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  const float rdn = __builtin_fmaf(-dn, s, x);
  const float rup = __builtin_fmaf(-up, s, x);
  s = (rdn <= 0.0f) ? dn : s;
  s = (rup > 0.0f) ? up : s;
  return s;
}

__device__ __forceinline__ float rcp_near(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(e, y, y);
}

// r = sqrt(r2) correctly rounded.  The scaled form runs only in waves where
// some active lane has r2 < 2^-96 (a uniform branch).
__device__ __forceinline__ float pair_sqrt(float r2) {
  if (__builtin_expect(__ballot(r2 < 0x1p-96f) != 0, 0)) {
    const bool tiny = r2 < 0x1p-96f;
    const float s = sqrt_near(tiny ? r2 * 0x1p64f : r2);
    return tiny ? s * 0x1p-32f : s;
  }
  return sqrt_near(r2);
}

// r = sqrt(r2) and inv = 1/r, each correctly rounded; inv = 1 for r == 0
// (vecmath.h Vec2Normalize's zero-length case).  r2 >= 2^-96 gives r >= 2^-48,
// so outside the scaled form r is never 0.
__device__ __forceinline__ void pair_length(float r2, float &r, float &inv) {
  if (__builtin_expect(__ballot(r2 < 0x1p-96f) != 0, 0)) {
    const bool tiny = r2 < 0x1p-96f;
    const float s = sqrt_near(tiny ? r2 * 0x1p64f : r2);
    r = tiny ? s * 0x1p-32f : s;
    const float y = rcp_near(s);
    inv = (s == 0.0f) ? 1.0f : (tiny ? y * 0x1p32f : y);
    return;
  }
  r = sqrt_near(r2);
  inv = rcp_near(r);
}

// ---------------------------------------------------------------------------
// Boundary collisions (sph.h:514-681), bodies of the scene in insertion order
// (demo4.cpp:414-441).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void normalize2(float vx, float vy, float &nx, float &ny) {
  float l = sqrtf(vx * vx + vy * vy);  // vecmath.h:267-280
  if (l == 0.0f) l = 1.0f;
  const float inv = 1.0f / l;
  nx = vx * inv;
  ny = vy * inv;
}

__device__ __forceinline__ void solve_plane(const DevParams &p, float &x, float &y, float nx, float ny, float d) {
  const float ppx = nx * d, ppy = ny * d;                 // sph.h:515
  const float dx = x - ppx, dy = y - ppy;                 // sph.h:517
  const float proj = dx * nx + dy * ny;                   // sph.h:518
  if (proj <= p.R) {                                      // sph.h:519
    const float pen = p.R - proj;                         // sph.h:520
    x = (nx * pen) + x;                                   // sph.h:521
    y = (ny * pen) + y;
  }
}

__device__ __forceinline__ void solve_circle(const DevParams &p, float &x, float &y, float cx, float cy, float r) {
  const float both = r + p.R;                             // sph.h:527
  const float dx = x - cx, dy = y - cy;                   // sph.h:529
  const float d2 = dx * dx + dy * dy;                     // sph.h:530
  if (d2 <= both * both) {                                // sph.h:531
    if (fabsf(d2) > 0.0f) {                               // sph.h:534
      const float dist = sqrtf(d2);
      const float inv = 1.0f / dist;
      const float nx = dx * inv, ny = dy * inv;           // sph.h:536
      const float pen = both - dist;                      // sph.h:537
      x = (nx * pen) + x;                                 // sph.h:538
      y = (ny * pen) + y;
    }
  }
}

__device__ __forceinline__ float dot2(float ax, float ay, float bx, float by) { return ax * bx + ay * by; }

// The segment and polygon solvers are out of line (they are large and rarely
// taken) and take and return the position by value, so the caller's x, y stay
// in registers.
__device__ __noinline__ float2 segment_push(float both, float x, float y, float ax, float ay, float bx, float by) {
  const float ex = bx - ax, ey = by - ay;                 // sph.h:548
  const float u = dot2(ex, ey, bx - x, by - y);           // sph.h:549
  const float v = dot2(ex, ey, x - ax, y - ay);           // sph.h:550
  float clx, cly, nx, ny;
  if (v <= 0.0f) {                                        // sph.h:555-563
    clx = ax; cly = ay;
    const float dx = x - clx, dy = y - cly;
    if (dot2(dx, dy, dx, dy) > both * both) return make_float2(x, y);
    normalize2(x - clx, y - cly, nx, ny);
  } else if (u <= 0.0f) {                                 // sph.h:564-572
    clx = bx; cly = by;
    const float dx = x - clx, dy = y - cly;
    if (dot2(dx, dy, dx, dy) > both * both) return make_float2(x, y);
    normalize2(x - clx, y - cly, nx, ny);
  } else {                                                // sph.h:573-591
    const float den = dot2(ex, ey, ex, ey);
    const float inv = 1.0f / den;
    clx = ((ax * u) + (bx * v)) * inv;
    cly = ((ay * u) + (by * v)) * inv;
    const float dx = x - clx, dy = y - cly;
    if (dot2(dx, dy, dx, dy) > both * both) return make_float2(x, y);
    float px = -ey, py = ex;
    if (dot2(px, py, x - ax, y - ay) < 0.0f) { px = -px; py = -py; }
    normalize2(px, py, nx, ny);
  }
  const float dist = dot2(nx, ny, x - clx, y - cly);      // sph.h:593-594
  const float pen = both - dist;                          // sph.h:595
  return make_float2((nx * pen) + x, (ny * pen) + y);     // sph.h:596
}

__device__ __forceinline__ void solve_segment(const DevParams &p, float &x, float &y, float ax, float ay, float bx,
                                              float by) {
  const float2 r = segment_push(p.margin + p.R, x, y, ax, ay, bx, by);  // sph.h:545 both
  x = r.x;
  y = r.y;
}

__device__ __noinline__ float2 polygon_push(float radius, float x, float y, int nv, const float2 *verts) {
  // sph.h:600-672 FindMTVCirclePolygon, sph.h:674-681 SPHSolvePolygonCollision
  int edge = 0;
  float nX = 0.0f, nY = 0.0f, sep = -3.402823466e+38f;
  for (int k = 0; k < nv; ++k) {
    const float2 a = verts[k];
    const float2 bb = verts[(k + 1) % nv];
    const float ex = bb.x - a.x, ey = bb.y - a.y;
    float cx, cy;
    normalize2(1.0f * ey, -1.0f * ex, cx, cy);  // vecmath.h:299-301 Vec2Cross(v, 1)
    const float s = dot2(cx, cy, x - a.x, y - a.y);
    if (s > radius) return make_float2(x, y);
    if (s > sep) { nX = cx; nY = cy; sep = s; edge = k; }
  }
  const float2 v1 = verts[edge];
  const float2 v2 = verts[(edge + 1) % nv];
  float mx, my;
  if (sep < 1.19209290e-07f) {                            // sph.h:626 FLT_EPSILON
    const float pen = radius - sep;
    mx = nX * pen; my = nY * pen;
  } else {
    const float u1 = dot2(x - v1.x, y - v1.y, v2.x - v1.x, v2.y - v1.y);
    const float u2 = dot2(x - v2.x, y - v2.y, v1.x - v2.x, v1.y - v2.y);
    if (u1 <= 0.0f || u2 <= 0.0f) {
      const float2 c = (u1 <= 0.0f) ? v1 : v2;
      // vecmath.h:292-296: the reference's "distance squared" is ((bx-ax)*(by-ay))^2
      const float f = (c.x - x) * (c.y - y);
      if (f * f > radius * radius) return make_float2(x, y);
      const float dx = x - c.x, dy = y - c.y;
      float ux, uy;
      normalize2(dx, dy, ux, uy);
      const float pen = radius - dot2(ux, uy, dx, dy);
      mx = ux * pen; my = uy * pen;
    } else {
      // vecmath.h:228-231 ScalarLerp(a, 0.5, b) = (1 - 0.5) * a + 0.5 * b
      const float fx = (1.0f - 0.5f) * v1.x + 0.5f * v2.x;
      const float fy = (1.0f - 0.5f) * v1.y + 0.5f * v2.y;
      const float s = dot2(x - fx, y - fy, nX, nY);
      if (s > radius) return make_float2(x, y);
      const float pen = radius - s;
      mx = nX * pen; my = nY * pen;
    }
  }
  return make_float2(x + mx, y + my);
}

__device__ __forceinline__ void solve_polygon(const DevParams &p, float &x, float &y, int nv, const float2 *verts) {
  const float2 r = polygon_push(p.margin + p.R, x, y, nv, verts);
  x = r.x;
  y = r.y;
}

// Body records do not change during a step: they are read through the
// constant address space, so a wave-uniform index loads with scalar loads.
__device__ __forceinline__ DevBody load_body(const DevBody *bodies, int k) {
  const __attribute__((address_space(4))) DevBody &c = ((const __attribute__((address_space(4))) DevBody *)bodies)[k];
  DevBody d;
  d.type = c.type;
  d.nverts = c.nverts;
  d.vertOffset = c.vertOffset;
  d.pad = c.pad;
  d.a0 = c.a0, d.a1 = c.a1, d.a2 = c.a2, d.a3 = c.a3;
  for (int q = 0; q < 4; ++q) d.cull[q] = c.cull[q];
  return d;
}

__device__ __forceinline__ void collide(const Buffers &b, const DevParams &p, const DevScene &sc, float &x, float &y) {
  for (int k = sc.bodyBegin; k < sc.bodyEnd; ++k) {
    const DevBody bd = load_body(b.bodies, k);
    switch (bd.type) {
      case BODY_PLANE: solve_plane(p, x, y, bd.a0, bd.a1, bd.a2); break;
      case BODY_CIRCLE: solve_circle(p, x, y, bd.a0, bd.a1, bd.a2); break;
      case BODY_SEGMENT: solve_segment(p, x, y, bd.a0, bd.a1, bd.a2, bd.a3); break;
      case BODY_POLYGON: solve_polygon(p, x, y, bd.nverts, b.verts + bd.vertOffset); break;
      default: break;
    }
  }
}

// Bounding box of the live lanes' positions as float_order images, the same
// in every lane.  All 64 lanes of the wave must be active.
struct WaveBox {
  int x0, y0, x1, y1;
};

// Minimum / maximum over the 16 lanes of each DPP row (quad swaps, half-row
// and row mirrors), then over the four rows' lanes 0, 16, 32, 48 in scalars.
__device__ __forceinline__ int row_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
  return v;
}
__device__ __forceinline__ int row_max(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
  v = row_min(v);
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_max(int v) {
  v = row_max(v);
  return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ __forceinline__ WaveBox wave_box(float x, float y, bool live) {
  const int ox = float_order(__float_as_int(x)), oy = float_order(__float_as_int(y));
  return {wave_min(live ? ox : INT_MAX), wave_min(live ? oy : INT_MAX), wave_max(live ? ox : INT_MIN),
          wave_max(live ? oy : INT_MIN)};
}

// Bodies among [c0, c0 + cnt) (cnt <= 64) that may move a position inside
// `box`: bit t for body c0 + t.  A circle is excluded when the box lies
// outside its cull box.  A plane is excluded when its projection, computed as
// solve_plane computes it, exceeds R at the box corner on the plane's side
// (the lowest x for nx >= 0, else the highest, and the same for y): each step
// of that computation (x - nx d, times nx, the sum) is a rounded subtraction,
// product or sum, monotone in the position for a fixed normal, so the
// corner's projection is at most any live position's, and every live
// position's test proj <= R fails as well (a NaN anywhere keeps the plane).
// Every other body is included.
__device__ __forceinline__ uint64_t body_candidates(const DevBody *bodies, int c0, int cnt, const WaveBox &box,
                                                    float R) {
  const int t = threadIdx.x & 63;
  bool may = false;
  if (t < cnt) {
    // the type, the parameters and the whole cull box (the record's 48 bytes
    // as three 16-byte loads) are loaded together and tested without
    // short-circuiting, so the test costs one memory round trip
    const DevBody &bd = bodies[c0 + t];
    const int type = bd.type;
    const float4 a = *reinterpret_cast<const float4 *>(&bd.a0);
    const int4 cl = *reinterpret_cast<const int4 *>(bd.cull);
    const bool outside = (cl.x > box.x1) | (cl.z < box.x0) | (cl.y > box.y1) | (cl.w < box.y0);
    const float cx = __int_as_float(float_order(a.x >= 0.0f ? box.x0 : box.x1));
    const float cy = __int_as_float(float_order(a.y >= 0.0f ? box.y0 : box.y1));
    const float ppx = a.x * a.z, ppy = a.y * a.z;  // solve_plane's operations, in its order
    const float proj = (cx - ppx) * a.x + (cy - ppy) * a.y;
    const bool beyond = proj > R;  // false for NaN
    may = ((type != BODY_CIRCLE) | !outside) & ((type != BODY_PLANE) | !beyond);
  }
  return __ballot(may);
}

// collide() for the live lanes of a wave, with the same results bit for bit.
// When the live lanes share one body list, the wave walks it once: the lanes
// test 64 bodies at a time against the wave's bounding box and only the
// bodies that may touch it are applied, in body order.  A circle is skipped
// when the box lies outside its cull box: every live position is then farther
// than both * (1 + 2^-8) from the centre along one axis, so its FP32 test
// d2 <= both^2 fails and the circle would leave it unchanged.  A plane is
// skipped when its test fails at the box's nearest corner (body_candidates).  After a body
// moves a live position the box and the remaining candidates are recomputed.
// Lanes that are not live compute on whatever x, y hold and must not store
// the result.  All 64 lanes of the wave must be active.
template <bool MULTI>
__device__ __forceinline__ void collide_wave(const Buffers &b, const DevParams &p, const DevScene &sc, float &x,
                                             float &y, bool live) {
  const uint64_t liveMask = __ballot(live);
  if (!liveMask) return;
  int k0 = sc.bodyBegin, k1 = sc.bodyEnd;
  if (MULTI) {
    const int src = __builtin_ctzll(liveMask);
    k0 = __builtin_amdgcn_readfirstlane(__shfl(sc.bodyBegin, src, 64));
    k1 = __builtin_amdgcn_readfirstlane(__shfl(sc.bodyEnd, src, 64));
    if (__ballot(live && (sc.bodyBegin != k0 || sc.bodyEnd != k1))) {
      if (live) collide(b, p, sc, x, y);
      return;
    }
  }
  if (k0 >= k1) return;
  WaveBox box = wave_box(x, y, live);
  for (int c0 = k0; c0 < k1; c0 += 64) {
    const int cnt = min(64, k1 - c0);
    uint64_t cand = body_candidates(b.bodies, c0, cnt, box, p.R);
    while (cand) {
      const int t = __builtin_amdgcn_readfirstlane(__builtin_ctzll(cand));
      cand &= cand - 1u;
      const DevBody bd = load_body(b.bodies, c0 + t);
      const float px = x, py = y;
      switch (bd.type) {
        case BODY_PLANE: solve_plane(p, x, y, bd.a0, bd.a1, bd.a2); break;
        case BODY_CIRCLE: solve_circle(p, x, y, bd.a0, bd.a1, bd.a2); break;
        case BODY_SEGMENT: solve_segment(p, x, y, bd.a0, bd.a1, bd.a2, bd.a3); break;
        case BODY_POLYGON: solve_polygon(p, x, y, bd.nverts, b.verts + bd.vertOffset); break;
        default: break;
      }
      const bool moved = (__float_as_uint(x) != __float_as_uint(px)) || (__float_as_uint(y) != __float_as_uint(py));
      if (__ballot(live && moved)) {
        box = wave_box(x, y, live);
        cand = body_candidates(b.bodies, c0, cnt, box, p.R) & ~((2ull << t) - 1u);  // bodies after t
      }
    }
  }
}

// Candidate ranges of a particle in cell (cx, cy): the 3x3 cells dy = -1..1
// (outer), dx = -1..1 (inner), restricted to the grid (demo4.cpp:188-203).  The
// three cells of one row have consecutive keys, so each row is one contiguous
// range of sorted slots.
struct RowRange {
  uint32_t j0, j1;
};

__device__ __forceinline__ RowRange row_range(const uint32_t *start, const DevScene &sc, int cx, int cy, int dy) {
  const int yy = cy + dy;
  if (yy < 0 || yy >= sc.ny) return {0u, 0u};
  const int xlo = max(cx - 1, 0), xhi = min(cx + 1, sc.nx - 1);
  const uint32_t row = sc.cellBase + (uint32_t)(yy * sc.nx);
  return {start[row + xlo], start[row + xhi + 1]};
}

// The three row ranges dy = -1, 0, 1 with their six loads in flight together:
// a row outside the grid reads the particle's own row and is emptied after.
__device__ __forceinline__ void row_ranges(const uint32_t *start, const DevScene &sc, int cx, int cy, RowRange rr[3]) {
  const int xlo = max(cx - 1, 0), xhi = min(cx + 1, sc.nx - 1);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int yy = cy + d - 1;
    const bool in = yy >= 0 && yy < sc.ny;
    const uint32_t row = sc.cellBase + (uint32_t)((in ? yy : cy) * sc.nx);
    rr[d].j0 = start[row + xlo];
    rr[d].j1 = start[row + xhi + 1];
    if (!in) rr[d].j1 = rr[d].j0;
  }
}

// LDS tile of a workgroup (positions only: every candidate's position is read
// by the cheap pass, the other per-particle data only for in-kernel pairs, from
// the slot arrays).  The workgroup's particles occupy consecutive
// sorted slots, so their keys span [kA, kB] and the candidates of all of them
// lie in three slot ranges, one per row offset dy: the cells with keys in
// [kA + dy*nx - 1, kB + dy*nx + 1] (clamped to the scene).  Each range is
// staged at tile index off[d]; a slot j of range d sits at off[d] + j - rowStart[d].
struct TilePlan {
  uint32_t rowStart[3];
  uint32_t off[4];  // off[3] = staged particles
  int use;          // 1 when the ranges fit the tile
};

template <int CAP>
__device__ __forceinline__ void plan_tile(const uint32_t *start, const DevScene &sc, uint32_t kA, uint32_t kB,
                                          TilePlan &t) {
  const int64_t first = sc.cellBase, last = (int64_t)sc.cellBase + (int64_t)sc.nx * sc.ny - 1;
  // all six cell starts are loaded together (an empty range reads the first
  // cell and is discarded), so the plan costs one memory round trip
  uint32_t a[3], e[3];
  bool ok[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t lo = max((int64_t)kA + (int64_t)(d - 1) * sc.nx - 1, first);
    const int64_t hi = min((int64_t)kB + (int64_t)(d - 1) * sc.nx + 1, last);
    ok[d] = lo <= hi;
    a[d] = start[ok[d] ? lo : first];
    e[d] = start[(ok[d] ? hi : first) + 1];
  }
  uint32_t total = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint32_t a0 = ok[d] ? a[d] : 0u, e0 = ok[d] ? e[d] : 0u;
    t.rowStart[d] = a0;
    t.off[d] = total;
    total += e0 - a0;
  }
  t.off[3] = total;
  t.use = (CAP > 0 && total <= (uint32_t)CAP) ? 1 : 0;
}

// Cooperative, coalesced copy of the planned ranges of N source arrays into N
// tiles.  Each thread keeps four loads per array in flight, all arrays'
// together.  Tile entry q of range d is slot q + (rowStart[d] - off[d]),
// formed in 32 bits (the difference wraps).  Entries past the end are clamped
// to the last one, which is then loaded and stored more than once with the
// same value.
template <int N>
__device__ __forceinline__ void stage_arrays(const TilePlan &t, const float2 *const (&src)[N],
                                             float2 *const (&dst)[N]) {
  constexpr uint32_t kU = 4;
  const uint32_t total = t.off[3], off1 = t.off[1], off2 = t.off[2];
  const uint32_t b0 = t.rowStart[0] - t.off[0], b1 = t.rowStart[1] - off1, b2 = t.rowStart[2] - off2;
  for (uint32_t q0 = threadIdx.x; q0 < total; q0 += kU * blockDim.x) {
    float2 v[N][kU];
    uint32_t at[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      at[u] = min(q0 + u * blockDim.x, total - 1u);
      const uint32_t b = at[u] >= off2 ? b2 : (at[u] >= off1 ? b1 : b0);
#pragma unroll
      for (int a = 0; a < N; ++a) v[a][u] = src[a][(uint32_t)(at[u] + b)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u)
#pragma unroll
      for (int a = 0; a < N; ++a) dst[a][at[u]] = v[a][u];
  }
}

__device__ __forceinline__ void stage(const TilePlan &t, const float2 *src, float2 *dst) {
  const float2 *const s[1] = {src};
  float2 *const d[1] = {dst};
  stage_arrays<1>(t, s, d);
}

// Index of slot j of row range d in the neighbour data source.
struct TileIndex {
  uint32_t rowStart[3], off[3];
  __device__ __forceinline__ explicit TileIndex(const TilePlan &t) {
    for (int d = 0; d < 3; ++d) rowStart[d] = t.rowStart[d], off[d] = t.off[d];
  }
  __device__ __forceinline__ uint32_t operator()(int d, uint32_t j) const { return off[d] + (j - rowStart[d]); }
};
struct SlotIndex {
  __device__ __forceinline__ uint32_t operator()(int, uint32_t j) const { return j; }
};

}  // namespace nbs
