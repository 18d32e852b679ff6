// CPU restatement of the reference's data-oriented ("Demo 4") SPH step.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the GPU library is checked
// against; nothing in the shipped library links or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Every arithmetic expression keeps the operand order of the reference line it
// cites (paths relative to NBodySimulation/ of f1nalspace/nbodysimulation_experiment).
// Build with -ffp-contract=off and without fast-math so FP32 results are the
// IEEE results of exactly these expressions (the reference's MSVC x64
// /O2 /fp:precise scalar-SSE semantics).
//
// Three update orders are provided:
//   * GS    - the reference's sequential Gauss-Seidel order: incremental
//             swap-remove grid, materialised neighbour lists, in-place
//             viscosity and position updates (demo4.cpp:183-255, 286-451).
//   * GS-MT - the same with the reference's pooled phases and contiguous range
//             partitioning (threading.h:111-129); shared particle fields are
//             read and written with relaxed atomics, so the outcome of the
//             reference's cross-partition races is nondeterministic.  Used as the
//             CPU throughput baseline only.
//   * JACOBI - the order the GPU uses: every neighbour pass reads the state
//             before the pass (gather), viscosity and relaxation add both halves
//             of each pair's symmetric update to the particle being updated, and
//             the grid is a canonical counting sort by (cell key, particle id).
//
// Templated on the arithmetic type: float reproduces the reference's FP32
// arithmetic, double is the FP64 oracle for the same update order.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Planted-mistake switch.  0 = faithful.  Non-zero values plant one plausible
// mistake each so tests can show the pinning checks catch it.
// ---------------------------------------------------------------------------
enum Mutation : int {
  MUT_NONE = 0,
  MUT_DROP_NEAR_DENSITY = 1,   // density[1] never accumulated
  MUT_DELTA_SIGN = 2,          // relaxation pushes toward the neighbour
  MUT_KEY_SWAP_XY = 3,         // cell key uses x*ny + y
  MUT_CUTOFF_LE = 4,           // r^2 <= h^2 instead of r^2 < h^2
  MUT_DROP_VISC_QUAD = 5,      // quadratic viscosity term dropped
  MUT_PLANE_SIGN = 6,          // plane projection pushes the wrong way
  MUT_PRESSURE_NO_REST = 7,    // pressure ignores the rest density
  MUT_VELOCITY_NO_INV_DT = 8,  // velocity = x - prev (no 1/dt)
  MUT_CIRCLE_NO_RADIUS = 9,    // circle contact radius ignores the particle radius
  MUT_SKIP_SELF = 10,          // candidate list drops the particle itself
  MUT_GRID_BY_INDEX = 11,      // Jacobi grid sorts cells by particle index, not by slot order
  MUT_COUNT
};
extern int g_mutation;

template <typename T>
struct V2 {
  T x, y;
};

// vecmath.h:233-236  operator*(Vec2f, float)
template <typename T> inline V2<T> vmul(V2<T> a, T b) { return {a.x * b, a.y * b}; }
// vecmath.h:245-252  operator+ ; operator+= evaluates b + a (IEEE add is commutative)
template <typename T> inline V2<T> vadd(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
// vecmath.h:253-260  operator- ; operator-= evaluates a - b
template <typename T> inline V2<T> vsub(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
// vecmath.h:241-244  unary minus
template <typename T> inline V2<T> vneg(V2<T> a) { return {-a.x, -a.y}; }
// vecmath.h:262-265
template <typename T> inline T vdot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
// vecmath.h:267-270
template <typename T> inline T vlen(V2<T> v) { return std::sqrt(v.x * v.x + v.y * v.y); }
// vecmath.h:272-280  (zero-length vectors map to the zero vector)
template <typename T> inline V2<T> vnormalize(V2<T> v) {
  T l = vlen(v);
  if (l == T(0)) l = T(1);
  T invL = T(1) / l;
  return vmul(v, invL);
}
// vecmath.h:282-285  Vec2Hadamard(Vec2f(d), n)
template <typename T> inline V2<T> vhadamard(T d, V2<T> n) { return {d * n.x, d * n.y}; }
// vecmath.h:292-296  NOTE: the reference's "distance squared" is ((bx-ax)*(by-ay))^2
template <typename T> inline T vdistsq_quirk(V2<T> a, V2<T> b) {
  T f = (b.x - a.x) * (b.y - a.y);
  return f * f;
}
// vecmath.h:299-301  right perpendicular
template <typename T> inline V2<T> vcross_vs(V2<T> a, T s) { return {s * a.y, -s * a.x}; }
// vecmath.h:228-231, 324-329
template <typename T> inline T slerp(T a, T t, T b) { return (T(1) - t) * a + t * b; }
template <typename T> inline V2<T> vlerp(V2<T> a, T t, V2<T> b) { return {slerp(a.x, t, b.x), slerp(a.y, t, b.y)}; }

// ---------------------------------------------------------------------------
// Parameters (sph.h:77-123).  The FP64 oracle receives the same float-rounded
// constants promoted to double, so both precisions model the same problem.
// ---------------------------------------------------------------------------
struct ParamsF {
  float h, rho0, k, k_near, sigma, beta, collision_radius, collision_margin;
};

template <typename T>
struct Params {
  T h, invH, rho0, k, kn, sigma, beta, R, margin;
  static Params from(const ParamsF &p) {
    Params r;
    r.h = T(p.h);
    r.invH = T(1) / r.h;  // sph.h:92, 102, 116  invKernelHeight = 1.0f / kernelHeight
    r.rho0 = T(p.rho0);
    r.k = T(p.k);
    r.kn = T(p.k_near);
    r.sigma = T(p.sigma);
    r.beta = T(p.beta);
    r.R = T(p.collision_radius);   // sph.h:38 kSPHParticleCollisionRadius
    r.margin = T(p.collision_margin);  // sph.h:54 kSPHCollisionMargin
    return r;
  }
};

// ---------------------------------------------------------------------------
// Pair kernels (sph.h:465-512)
// ---------------------------------------------------------------------------
template <typename T>
inline bool inKernel(const Params<T> &p, T r2) {
  if (g_mutation == MUT_CUTOFF_LE) return r2 <= (p.h * p.h);
  return r2 < (p.h * p.h);  // sph.h:470, 488, 502
}

// sph.h:465-476 SPHComputeDensity
template <typename T>
inline void computeDensity(const Params<T> &p, V2<T> pos, V2<T> npos, T out[2]) {
  V2<T> Rij = vsub(npos, pos);
  T r2 = vdot(Rij, Rij);
  if (inKernel(p, r2)) {
    T rij = std::sqrt(r2);
    T term = T(1) - rij * p.invH;
    out[0] += (term * term);
    if (g_mutation != MUT_DROP_NEAR_DENSITY) out[1] += (term * term * term);
  }
}

// sph.h:478-481 SPHComputePressure
template <typename T>
inline void computePressure(const Params<T> &p, const T d[2], T out[2]) {
  if (g_mutation == MUT_PRESSURE_NO_REST)
    out[0] = p.k * d[0];
  else
    out[0] = p.k * (d[0] - p.rho0);
  out[1] = p.kn * d[1];
}

// sph.h:483-495 SPHComputeDelta
template <typename T>
inline void computeDelta(const Params<T> &p, V2<T> pos, V2<T> npos, const T pres[2], T dt, V2<T> *out) {
  V2<T> Rij = vsub(npos, pos);
  T r2 = vdot(Rij, Rij);
  if (inKernel(p, r2)) {
    T rij = std::sqrt(r2);
    V2<T> n = vnormalize(Rij);
    if (g_mutation == MUT_DELTA_SIGN) n = vneg(n);
    T term = T(1) - rij * p.invH;
    T d = (dt * dt) * (pres[0] * term + pres[1] * (term * term));
    *out = vhadamard(d, n);
  }
}

// sph.h:497-512 SPHComputeViscosityForce
template <typename T>
inline void computeViscosity(const Params<T> &p, V2<T> pos, V2<T> npos, V2<T> vel, V2<T> nvel, V2<T> *out) {
  V2<T> Rij = vsub(npos, pos);
  T r2 = vdot(Rij, Rij);
  if (inKernel(p, r2)) {
    T rij = std::sqrt(r2);
    T q = rij * p.invH;
    V2<T> n = vnormalize(Rij);
    T u = vdot(vsub(vel, nvel), n);
    if (u > T(0)) {
      T f;
      if (g_mutation == MUT_DROP_VISC_QUAD)
        f = (T(1) - q) * (p.sigma * u);
      else
        f = (T(1) - q) * (p.sigma * u + p.beta * (u * u));
      *out = vhadamard(f, n);
    }
  }
}

// ---------------------------------------------------------------------------
// Boundary collisions (sph.h:514-681)
// ---------------------------------------------------------------------------
// sph.h:514-524 SPHSolvePlaneCollision
template <typename T>
inline void solvePlane(const Params<T> &p, V2<T> *position, V2<T> normal, T distance) {
  V2<T> pp = vmul(normal, distance);
  V2<T> particlePos = *position;
  V2<T> delta = vsub(particlePos, pp);
  T proj = vdot(delta, normal);
  if (proj <= p.R) {
    T penetration = p.R - proj;
    if (g_mutation == MUT_PLANE_SIGN) penetration = -penetration;
    particlePos = vadd(vmul(normal, penetration), particlePos);
    *position = particlePos;
  }
}

// sph.h:526-542 SPHSolveCircleCollision.  `abs(distanceSquared) > 0` is read as
// the floating-point absolute value (distanceSquared is never negative).
template <typename T>
inline void solveCircle(const Params<T> &p, V2<T> *position, V2<T> c, T radius) {
  T bothRadius = (g_mutation == MUT_CIRCLE_NO_RADIUS) ? radius : radius + p.R;
  V2<T> particlePos = *position;
  V2<T> deltaPos = vsub(particlePos, c);
  T distanceSquared = vdot(deltaPos, deltaPos);
  if (distanceSquared <= bothRadius * bothRadius) {
    if (std::fabs(distanceSquared) > T(0)) {
      T distance = std::sqrt(distanceSquared);
      V2<T> normal = vmul(deltaPos, T(1) / distance);
      T penetration = bothRadius - distance;
      particlePos = vadd(vmul(normal, penetration), particlePos);
      *position = particlePos;
    }
  }
}

// sph.h:544-598 SPHSolveLineSegmentCollision
template <typename T>
inline void solveSegment(const Params<T> &p, V2<T> *position, V2<T> a, V2<T> b) {
  T bothRadius = p.margin + p.R;
  V2<T> particlePos = *position;
  V2<T> e = vsub(b, a);
  T u = vdot(e, vsub(b, particlePos));
  T v = vdot(e, vsub(particlePos, a));
  V2<T> closest{T(0), T(0)};
  V2<T> normal{T(0), T(0)};
  if (v <= T(0)) {
    closest = a;
    V2<T> d = vsub(particlePos, closest);
    T dd = vdot(d, d);
    if (dd > bothRadius * bothRadius) return;
    normal = vnormalize(vsub(particlePos, closest));
  } else if (u <= T(0)) {
    closest = b;
    V2<T> d = vsub(particlePos, closest);
    T dd = vdot(d, d);
    if (dd > bothRadius * bothRadius) return;
    normal = vnormalize(vsub(particlePos, closest));
  } else {
    T den = vdot(e, e);
    closest = vmul(vadd(vmul(a, u), vmul(b, v)), T(1) / den);
    V2<T> d = vsub(particlePos, closest);
    T dd = vdot(d, d);
    if (dd > bothRadius * bothRadius) return;
    normal = V2<T>{-e.y, e.x};
    if (vdot(normal, vsub(particlePos, a)) < T(0)) normal = vneg(normal);
    normal = vnormalize(normal);
  }
  V2<T> deltaPos = vsub(particlePos, closest);
  T distance = vdot(normal, deltaPos);
  T penetration = bothRadius - distance;
  particlePos = vadd(vmul(normal, penetration), particlePos);
  *position = particlePos;
}

// sph.h:600-672 FindMTVCirclePolygon (including the Vec2DistanceSquared quirk)
template <typename T>
inline bool findMTVCirclePolygon(const Params<T> &p, V2<T> circlePosition, int vertexCount, const V2<T> *verts,
                                 V2<T> *mtv) {
  int edgeIndex = 0;
  V2<T> normal{T(0), T(0)};
  T separation = T(-FLT_MAX);
  T radius = p.margin + p.R;
  for (int vi = 0; vi < vertexCount; ++vi) {
    V2<T> a = verts[vi];
    V2<T> b = verts[(vi + 1) % vertexCount];
    V2<T> n = vnormalize(vcross_vs(vsub(b, a), T(1)));
    T s = vdot(n, vsub(circlePosition, a));
    if (s > radius) return false;
    if (s > separation) {
      normal = n;
      separation = s;
      edgeIndex = vi;
    }
  }
  V2<T> v1 = verts[edgeIndex];
  V2<T> v2 = verts[(edgeIndex + 1) % vertexCount];
  if (separation < T(FLT_EPSILON)) {
    T penetration = radius - separation;
    *mtv = vmul(normal, penetration);
    return true;
  }
  T u1 = vdot(vsub(circlePosition, v1), vsub(v2, v1));
  T u2 = vdot(vsub(circlePosition, v2), vsub(v1, v2));
  if (u1 <= T(0)) {
    if (vdistsq_quirk(circlePosition, v1) > radius * radius) return false;
    V2<T> distanceToEdge = vsub(circlePosition, v1);
    normal = vnormalize(distanceToEdge);
    T penetration = radius - vdot(normal, distanceToEdge);
    *mtv = vmul(normal, penetration);
    return true;
  } else if (u2 <= T(0)) {
    if (vdistsq_quirk(circlePosition, v2) > radius * radius) return false;
    V2<T> distanceToEdge = vsub(circlePosition, v2);
    normal = vnormalize(distanceToEdge);
    T penetration = radius - vdot(normal, distanceToEdge);
    *mtv = vmul(normal, penetration);
    return true;
  } else {
    V2<T> faceCenter = vlerp(v1, T(0.5f), v2);
    V2<T> distanceToFace = vsub(circlePosition, faceCenter);
    T s = vdot(distanceToFace, normal);
    if (s > radius) return false;
    T penetration = radius - s;
    *mtv = vmul(normal, penetration);
    return true;
  }
}

// sph.h:674-681 SPHSolvePolygonCollision
template <typename T>
inline void solvePolygon(const Params<T> &p, V2<T> *position, int vertexCount, const V2<T> *verts) {
  V2<T> particlePos = *position;
  V2<T> mtv{T(0), T(0)};
  if (findMTVCirclePolygon(p, particlePos, vertexCount, verts, &mtv)) {
    particlePos = vadd(particlePos, mtv);
    *position = particlePos;
  }
}

// ---------------------------------------------------------------------------
// Scenes, bodies, grid keys
// ---------------------------------------------------------------------------
enum BodyType : int { BODY_PLANE = 1, BODY_CIRCLE = 2, BODY_SEGMENT = 3, BODY_POLYGON = 4 };  // demo4.h:18-27
constexpr int kMaxPolygonVerts = 8;  // sph.h:161 kMaxScenarioPolygonCount

template <typename T>
struct Body {
  int type;
  V2<T> a, b;  // plane: a = normal, b.x = distance; circle: a = centre, b.x = radius; segment: a, b
  int vertexCount;
  V2<T> verts[kMaxPolygonVerts];
};

template <typename T>
struct Scene {
  T width, height, halfW, halfH, cellSize;
  int nx, ny;
  uint32_t cellBase;
  std::vector<Body<T>> bodies;
};

// sph.h:450-463 SPHComputeCellPos + SPHComputeCellIndex, per axis.  The clamp
// is applied in floating point first so values the reference's (int) cast leaves
// undefined (NaN, beyond INT_MAX) are defined: NaN and negatives give 0, values
// at or beyond n give n-1.  For every finite value the reference defines, the
// result is the reference's.
template <typename T>
inline int cellCoord(T p, T half, T cellSize, int n) {
  T f = (p + half) / cellSize;
  if (!(f >= T(0))) return 0;
  if (f >= T(n)) return n - 1;
  return (int)f;
}

template <typename T>
inline uint32_t cellKey(const Scene<T> &s, V2<T> p, int *cx_out = nullptr, int *cy_out = nullptr) {
  int cx = cellCoord(p.x, s.halfW, s.cellSize, s.nx);
  int cy = cellCoord(p.y, s.halfH, s.cellSize, s.ny);
  if (cx_out) *cx_out = cx;
  if (cy_out) *cy_out = cy;
  if (g_mutation == MUT_KEY_SWAP_XY) return s.cellBase + (uint32_t)(cx * s.ny + cy);
  return s.cellBase + (uint32_t)(cy * s.nx + cx);  // sph.h:444-448 y * countX + x
}

struct Stats {
  uint64_t minParticleNeighborCount, maxParticleNeighborCount;
  uint64_t minCellParticleCount, maxCellParticleCount;
  // sph.h:131-141 per-phase milliseconds
  float time[9];  // emitters, integration, viscosityForces, predict, updateGrid,
                  // neighborSearch, densityAndPressure, deltaPositions, collisions
};

// ---------------------------------------------------------------------------
// Minimal thread pool with the reference's task partitioning
// (threading.h:101-129): chunk = max(1, n / threads), inclusive end indices,
// tasks pulled FIFO by the workers, the caller waits until all are done.
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int threads) : nthreads_(std::max(1, threads)) {
    for (int t = 0; t < nthreads_; ++t) workers_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &w : workers_) w.join();
  }
  int threads() const { return nthreads_; }
  // threading.h:111-129 CreateTasks + WaitUntilDone
  void run(int64_t itemCount, const std::function<void(int64_t, int64_t)> &fn) {
    if (itemCount <= 0) return;
    int64_t per = std::max<int64_t>(1, itemCount / nthreads_);
    std::vector<std::pair<int64_t, int64_t>> tasks;
    for (int64_t s = 0; s < itemCount; s += per) tasks.push_back({s, std::min(s + per - 1, itemCount - 1)});
    {
      std::lock_guard<std::mutex> g(m_);
      tasks_ = std::move(tasks);
      fn_ = &fn;
      next_.store(0);
      pending_.store((int64_t)tasks_.size());
      ++generation_;
    }
    cv_.notify_all();
    while (pending_.load(std::memory_order_acquire) > 0) std::this_thread::yield();
    // Every worker that joined this generation must leave its task loop before
    // the task list and the function are replaced.
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int64_t, int64_t)> *fn;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
        fn = fn_;
        ++active_;
      }
      if (fn) {
        for (;;) {
          int64_t t = next_.fetch_add(1);
          if (t >= (int64_t)tasks_.size()) break;
          (*fn)(tasks_[t].first, tasks_[t].second);
          pending_.fetch_sub(1, std::memory_order_release);
        }
      }
      {
        std::lock_guard<std::mutex> g(m_);
        if (--active_ == 0) done_.notify_all();
      }
    }
  }
  int nthreads_;
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  bool stop_ = false;
  uint64_t generation_ = 0;
  int active_ = 0;
  std::vector<std::pair<int64_t, int64_t>> tasks_;
  const std::function<void(int64_t, int64_t)> *fn_ = nullptr;
  std::atomic<int64_t> next_{0};
  std::atomic<int64_t> pending_{0};
};

}  // namespace orc
