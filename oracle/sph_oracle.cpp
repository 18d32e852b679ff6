// CPU restatement of the reference's "Demo 4" SPH step: simulations and the
// C API used by the tests through ctypes.  TEST INFRASTRUCTURE ONLY (see
// sph_oracle.hpp for what each update order means).
#include "sph_oracle.hpp"

namespace orc {

int g_mutation = MUT_NONE;

// sph.h:68-70, 144-147: initial statistics values
static constexpr uint64_t kMaxCellParticleCount = 500;
static constexpr uint64_t kMaxParticleNeighborCount = 1000;

using Clock = std::chrono::steady_clock;
static inline float msSince(Clock::time_point t0) {
  return (float)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count() * (1.0f / 1000000);
}

// SplitMix64 uniform in [-1, 1) with 24-bit resolution: the seeded stand-in
// for the reference's Vec2RandomDirection() jitter (vecmath.h:317-322).
static inline float splitmixUniform(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (float)(uint32_t)(z >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

struct Emitter {  // demo4.h:123-135
  int scene;
  float px, py, dx, dy, radius, speed, rate, duration, elapsed, totalElapsed;
  bool active;
  uint64_t seed, draws;
};

// ---------------------------------------------------------------------------
// Abstract simulation used by the C API.
// ---------------------------------------------------------------------------
struct SimBase {
  virtual ~SimBase() {}
  virtual int addScene(float width, float height, float cellSize) = 0;
  virtual int addBody(int scene, int type, const float *data, int nverts) = 0;
  virtual int64_t addParticles(int scene, const float *pos, const float *force, int64_t n) = 0;
  virtual void setGravity(float gx, float gy) = 0;
  virtual void setExternalForce(float fx, float fy) = 0;
  virtual void addEmitter(const Emitter &e) = 0;
  virtual void setSpacing(float s) = 0;
  virtual void step(float dt, int nsteps) = 0;
  virtual int64_t count() const = 0;
  virtual int64_t totalCells() const = 0;
  virtual void getState(double *x, double *prev, double *v, double *dens, double *pres, uint32_t *key) const = 0;
  virtual void setState(const double *x, const double *v) = 0;
  virtual int64_t neighbors(int64_t i, uint32_t *out, int64_t cap) const = 0;
  virtual void grid(uint32_t *order, uint32_t *cellStart) const = 0;
  virtual Stats stats() const = 0;
  virtual void resetStats() = 0;
};

// Shared scene/body bookkeeping.
template <typename T>
struct World {
  Params<T> p;
  std::vector<Scene<T>> scenes;
  V2<T> gravity{T(0), T(0)}, externalForce{T(0), T(0)};
  uint32_t cells = 0;

  int addScene(float width, float height, float cellSize) {
    Scene<T> s;
    // sph.h:18-23, 60-63: half extents and cell counts derive from the width,
    // height and cell size in float, as the reference's constants do.
    float halfW = width * 0.5f, halfH = height * 0.5f;
    s.width = T(width);
    s.height = T(height);
    s.halfW = T(halfW);
    s.halfH = T(halfH);
    s.cellSize = T(cellSize);
    s.nx = std::max(1, (int)(width / cellSize));
    s.ny = std::max(1, (int)(height / cellSize));
    s.cellBase = cells;
    cells += (uint32_t)(s.nx * s.ny);
    scenes.push_back(s);
    return (int)scenes.size() - 1;
  }
  int addBody(int scene, int type, const float *d, int nverts) {
    if (scene < 0 || scene >= (int)scenes.size()) return -1;
    Body<T> b{};
    b.type = type;
    if (type == BODY_PLANE || type == BODY_CIRCLE) {
      b.a = {T(d[0]), T(d[1])};
      b.b = {T(d[2]), T(0)};
    } else if (type == BODY_SEGMENT) {
      b.a = {T(d[0]), T(d[1])};
      b.b = {T(d[2]), T(d[3])};
    } else if (type == BODY_POLYGON) {
      if (nverts < 3 || nverts > kMaxPolygonVerts) return -1;
      b.vertexCount = nverts;
      for (int k = 0; k < nverts; ++k) b.verts[k] = {T(d[2 * k]), T(d[2 * k + 1])};
    } else {
      return -1;
    }
    scenes[scene].bodies.push_back(b);
    return 0;
  }
  // demo4.cpp:411-441: every body of the particle's scene, in insertion order.
  void collide(int scene, V2<T> *x) const {
    for (const Body<T> &b : scenes[scene].bodies) {
      switch (b.type) {
        case BODY_PLANE: solvePlane(p, x, b.a, b.b.x); break;
        case BODY_CIRCLE: solveCircle(p, x, b.a, b.b.x); break;
        case BODY_SEGMENT: solveSegment(p, x, b.a, b.b); break;
        case BODY_POLYGON: solvePolygon(p, x, b.vertexCount, b.verts); break;
      }
    }
  }
  int sceneOfKey(uint32_t key) const {
    int lo = 0, hi = (int)scenes.size() - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) / 2;
      if (scenes[mid].cellBase <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
  }
  // demo4.cpp:188-203: candidate cells dy = -1..1 (outer), dx = -1..1 (inner),
  // restricted to the grid (sph.h:439-442).
  template <typename F>
  void forCandidateCells(uint32_t key, int scene, F &&f) const {
    const Scene<T> &s = scenes[scene];
    uint32_t local = key - s.cellBase;
    int cx = (int)(local % (uint32_t)s.nx), cy = (int)(local / (uint32_t)s.nx);
    if (g_mutation == MUT_KEY_SWAP_XY) { cx = (int)(local / (uint32_t)s.ny); cy = (int)(local % (uint32_t)s.ny); }
    for (int y = -1; y <= 1; ++y) {
      for (int x = -1; x <= 1; ++x) {
        int px = cx + x, py = cy + y;
        if (px >= 0 && px < s.nx && py >= 0 && py < s.ny) {
          uint32_t k = (g_mutation == MUT_KEY_SWAP_XY) ? s.cellBase + (uint32_t)(px * s.ny + py)
                                                        : s.cellBase + (uint32_t)(py * s.nx + px);
          f(k);
        }
      }
    }
  }
};

// demo4.cpp:257-284 UpdateEmitter, in FP32 like the reference; emitted particles
// are added through the simulation's AddParticle.
template <typename AddFn>
static void runEmitter(Emitter &e, float deltaTime, float spacing, float h, AddFn &&addParticle) {
  const float invDeltaTime = 1.0f / deltaTime;
  if (!e.active) return;
  const float rate = 1.0f / e.rate;
  e.elapsed += deltaTime;
  e.totalElapsed += deltaTime;
  if (e.elapsed >= rate) {
    e.elapsed = 0;
    const float ax = (e.dx * e.speed) * invDeltaTime, ay = (e.dy * e.speed) * invDeltaTime;
    const float dirx = -1.0f * e.dy, diry = 1.0f * e.dx;  // vecmath.h:304-306 Vec2Cross(1, d)
    const int count = (int)std::floor(e.radius / spacing);
    const float offx = ((dirx * (float)count) * spacing) * 0.5f;
    const float offy = ((diry * (float)count) * spacing) * 0.5f;
    for (int index = 0; index < count; ++index) {
      float px = (dirx * (float)index) * spacing, py = (diry * (float)index) * spacing;
      px = ((dirx * spacing) * 0.5f) + px;
      py = ((diry * spacing) * 0.5f) + py;
      px = (e.px - offx) + px;
      py = (e.py - offy) + py;
      const float jx = (splitmixUniform(e.seed, e.draws) * h) * 0.01f;
      const float jy = (splitmixUniform(e.seed, e.draws + 1) * h) * 0.01f;
      e.draws += 2;
      const float pos[2] = {jx + px, jy + py}, force[2] = {ax, ay};
      addParticle(e.scene, pos, force);
    }
  }
  if (e.totalElapsed >= e.duration) e.active = false;
}

// ---------------------------------------------------------------------------
// Reference order (demo4.cpp), serial or pooled.
// ---------------------------------------------------------------------------
template <typename T>
struct ParticleData {  // demo4.h:81-109 (48 bytes for T = float)
  V2<T> cur, prev, acc, vel;
  T dens[2];
  T pres[2];
};

struct ParticleIndex {  // demo4.h:111-116, neighbour storage made dynamic
  uint32_t key = 0;
  int scene = 0;
  std::vector<size_t> neighbors;
  size_t indexInCell = 0;
};

// Field access: plain for the serial order, relaxed atomics for the pooled
// order (the reference's threads read and write each other's particles).
template <typename T, bool ATOMIC>
struct Acc {
  static inline T ld(const T &v) {
    if constexpr (ATOMIC) return std::atomic_ref<T>(const_cast<T &>(v)).load(std::memory_order_relaxed);
    else return v;
  }
  static inline void st(T &v, T x) {
    if constexpr (ATOMIC) std::atomic_ref<T>(v).store(x, std::memory_order_relaxed);
    else v = x;
  }
  static inline V2<T> ld2(const V2<T> &v) { return {ld(v.x), ld(v.y)}; }
  static inline void st2(V2<T> &v, V2<T> x) { st(v.x, x.x); st(v.y, x.y); }
};

template <typename T>
struct GSSim : SimBase {
  World<T> w;
  std::vector<ParticleData<T>> pd;
  std::vector<ParticleIndex> pi;
  std::vector<std::vector<size_t>> cells;  // demo4.h:118-121, capacity made dynamic
  Stats st;
  std::unique_ptr<Pool> pool;  // null: serial order
  std::vector<Emitter> emitters;
  float spacing;

  void setSpacing(float s) override { spacing = s; }
  GSSim(const ParamsF &pf, int threads) : spacing(pf.h * 0.5f) {
    w.p = Params<T>::from(pf);
    if (threads > 0) pool.reset(new Pool(threads));
    resetStats();
  }
  void resetStats() override {  // sph.h:143-149
    st = Stats{};
    st.minParticleNeighborCount = kMaxCellParticleCount;
    st.maxParticleNeighborCount = 0;
    st.minCellParticleCount = kMaxCellParticleCount;
    st.maxCellParticleCount = 0;
  }
  int addScene(float width, float height, float cs) override {
    int s = w.addScene(width, height, cs);
    cells.resize(w.cells);
    return s;
  }
  int addBody(int scene, int type, const float *d, int nv) override { return w.addBody(scene, type, d, nv); }
  void setGravity(float gx, float gy) override { w.gravity = {T(gx), T(gy)}; }
  void setExternalForce(float fx, float fy) override { w.externalForce = {T(fx), T(fy)}; }
  void addEmitter(const Emitter &e) override { emitters.push_back(e); }
  int64_t count() const override { return (int64_t)pd.size(); }
  int64_t totalCells() const override { return w.cells; }

  // demo4.cpp:37-54 InsertParticleIntoGrid
  void insertIntoGrid(size_t i) {
    uint32_t key = cellKey(w.scenes[pi[i].scene], pd[i].cur);
    std::vector<size_t> &cell = cells[key];
    size_t indexInCell = cell.size();
    cell.push_back(i);
    pi[i].key = key;
    pi[i].indexInCell = indexInCell;
    uint64_t c = cell.size();
    st.minCellParticleCount = std::min(c, st.minCellParticleCount);
    st.maxCellParticleCount = std::max(c, st.maxCellParticleCount);
  }
  // demo4.cpp:56-76 RemoveParticleFromGrid (swap-remove)
  void removeFromGrid(size_t i) {
    std::vector<size_t> &cell = cells[pi[i].key];
    size_t removalIndex = pi[i].indexInCell;
    size_t lastIndex = cell.size() - 1;
    if (removalIndex != lastIndex) {
      cell[removalIndex] = cell[lastIndex];
      cell[lastIndex] = i;
      pi[cell[removalIndex]].indexInCell = removalIndex;
    }
    cell.pop_back();
    uint64_t c = cell.size();
    st.minCellParticleCount = std::min(c, st.minCellParticleCount);
    st.maxCellParticleCount = std::max(c, st.maxCellParticleCount);
  }
  // demo4.cpp:142-153 AddParticle
  int64_t addParticles(int scene, const float *pos, const float *force, int64_t n) override {
    if (scene < 0 || scene >= (int)w.scenes.size()) return -1;
    int64_t first = (int64_t)pd.size();
    for (int64_t k = 0; k < n; ++k) {
      ParticleData<T> d{};
      d.cur = d.prev = {T(pos[2 * k]), T(pos[2 * k + 1])};
      d.acc = force ? V2<T>{T(force[2 * k]), T(force[2 * k + 1])} : V2<T>{T(0), T(0)};
      pd.push_back(d);
      ParticleIndex x;
      x.scene = scene;
      pi.push_back(x);
      insertIntoGrid(pd.size() - 1);
    }
    return first;
  }
  void setState(const double *x, const double *v) override {
    for (size_t i = 0; i < pd.size(); ++i) {
      pd[i].cur = pd[i].prev = {T(x[2 * i]), T(x[2 * i + 1])};
      pd[i].vel = {T(v[2 * i]), T(v[2 * i + 1])};
      pd[i].acc = {T(0), T(0)};
      removeFromGrid(i);
      insertIntoGrid(i);
      pi[i].neighbors.clear();
    }
  }

  template <typename F>
  void phase(F &&fn) {
    int64_t n = (int64_t)pd.size();
    if (pool) pool->run(n, fn);
    else if (n > 0) fn(0, n - 1);  // demo4.cpp:323 (0, particleCount - 1)
  }

  // demo4.cpp:223-237 ViscosityForces
  template <bool ATOMIC>
  void viscosityRange(int64_t s, int64_t e, T dt) {
    using A = Acc<T, ATOMIC>;
    for (int64_t i = s; i <= e; ++i) {
      ParticleData<T> &a = pd[i];
      const std::vector<size_t> &nb = pi[i].neighbors;
      for (size_t idx = 0; idx < nb.size(); ++idx) {
        ParticleData<T> &b = pd[nb[idx]];
        V2<T> force{T(0), T(0)};
        computeViscosity(w.p, A::ld2(a.cur), A::ld2(b.cur), A::ld2(a.vel), A::ld2(b.vel), &force);
        A::st2(a.vel, vsub(A::ld2(a.vel), vmul(vmul(force, T(0.5f)), dt)));
        A::st2(b.vel, vadd(vmul(vmul(force, T(0.5f)), dt), A::ld2(b.vel)));
      }
    }
  }
  // demo4.cpp:183-206 NeighborSearch
  void neighborRange(int64_t s, int64_t e) {
    for (int64_t i = s; i <= e; ++i) {
      ParticleIndex &x = pi[i];
      x.neighbors.clear();
      w.forCandidateCells(x.key, x.scene, [&](uint32_t k) {
        const std::vector<size_t> &cell = cells[k];
        for (size_t c = 0; c < cell.size(); ++c) {
          if (g_mutation == MUT_SKIP_SELF && cell[c] == (size_t)i) continue;
          x.neighbors.push_back(cell[c]);
        }
      });
    }
  }
  // demo4.cpp:208-221 DensityAndPressure
  void densityRange(int64_t s, int64_t e) {
    for (int64_t i = s; i <= e; ++i) {
      ParticleData<T> &a = pd[i];
      a.dens[0] = a.dens[1] = T(0);
      for (size_t j : pi[i].neighbors) computeDensity(w.p, a.cur, pd[j].cur, a.dens);
      computePressure(w.p, a.dens, a.pres);
    }
  }
  // demo4.cpp:239-255 DeltaPositions
  template <bool ATOMIC>
  void deltaRange(int64_t s, int64_t e, T dt) {
    using A = Acc<T, ATOMIC>;
    for (int64_t i = s; i <= e; ++i) {
      ParticleData<T> &a = pd[i];
      V2<T> dx{T(0), T(0)};
      const std::vector<size_t> &nb = pi[i].neighbors;
      for (size_t idx = 0; idx < nb.size(); ++idx) {
        ParticleData<T> &b = pd[nb[idx]];
        V2<T> delta{T(0), T(0)};
        computeDelta(w.p, A::ld2(a.cur), A::ld2(b.cur), a.pres, dt, &delta);
        A::st2(b.cur, vadd(vmul(delta, T(0.5f)), A::ld2(b.cur)));
        dx = vsub(dx, vmul(delta, T(0.5f)));
      }
      A::st2(a.cur, vadd(dx, A::ld2(a.cur)));
    }
  }

  // demo4.cpp:286-451 Update
  void update(T dt) {
    const T invDt = T(1) / dt;
    Clock::time_point t0 = Clock::now();  // demo4.cpp:290-299 emitters
    for (Emitter &e : emitters)
      runEmitter(e, (float)dt, spacing, (float)w.p.h,
                 [&](int sc, const float *pos, const float *force) { addParticles(sc, pos, force, 1); });
    const int64_t n = (int64_t)pd.size();
    st.time[0] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:301-312 integrate
    for (int64_t i = 0; i < n; ++i) {
      ParticleData<T> &d = pd[i];
      d.acc = vadd(vadd(w.gravity, w.externalForce), d.acc);
      d.vel = vadd(vmul(d.acc, dt), d.vel);
      d.acc = {T(0), T(0)};
    }
    st.time[1] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:314-327 viscosity
    if (pool) phase([&](int64_t s, int64_t e) { viscosityRange<true>(s, e, dt); });
    else phase([&](int64_t s, int64_t e) { viscosityRange<false>(s, e, dt); });
    st.time[2] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:329-339 predict
    for (int64_t i = 0; i < n; ++i) {
      pd[i].prev = pd[i].cur;
      pd[i].cur = vadd(vmul(pd[i].vel, dt), pd[i].cur);
    }
    st.time[3] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:341-356 update grid
    for (int64_t i = 0; i < n; ++i) {
      uint32_t k = cellKey(w.scenes[pi[i].scene], pd[i].cur);
      if (k != pi[i].key) {
        removeFromGrid((size_t)i);
        insertIntoGrid((size_t)i);
      }
    }
    st.time[4] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:358-379 neighbour search + stats
    phase([&](int64_t s, int64_t e) { neighborRange(s, e); });
    st.minParticleNeighborCount = kMaxParticleNeighborCount;
    st.maxParticleNeighborCount = 0;
    for (int64_t i = 0; i < n; ++i) {
      uint64_t c = pi[i].neighbors.size();
      st.minParticleNeighborCount = std::min(c, st.minParticleNeighborCount);
      st.maxParticleNeighborCount = std::max(c, st.maxParticleNeighborCount);
    }
    st.time[5] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:381-394 density and pressure
    phase([&](int64_t s, int64_t e) { densityRange(s, e); });
    st.time[6] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:396-409 delta positions
    if (pool) phase([&](int64_t s, int64_t e) { deltaRange<true>(s, e, dt); });
    else phase([&](int64_t s, int64_t e) { deltaRange<false>(s, e, dt); });
    st.time[7] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:411-444 collisions
    for (int64_t i = 0; i < n; ++i) w.collide(pi[i].scene, &pd[i].cur);
    st.time[8] = msSince(t0);

    for (int64_t i = 0; i < n; ++i) {  // demo4.cpp:446-450 velocity
      if (g_mutation == MUT_VELOCITY_NO_INV_DT) pd[i].vel = vsub(pd[i].cur, pd[i].prev);
      else pd[i].vel = vmul(vsub(pd[i].cur, pd[i].prev), invDt);
    }
  }
  void step(float dt, int nsteps) override {
    for (int s = 0; s < nsteps; ++s) update(T(dt));
  }
  void getState(double *x, double *prev, double *v, double *dens, double *pres, uint32_t *key) const override {
    for (size_t i = 0; i < pd.size(); ++i) {
      if (x) { x[2 * i] = pd[i].cur.x; x[2 * i + 1] = pd[i].cur.y; }
      if (prev) { prev[2 * i] = pd[i].prev.x; prev[2 * i + 1] = pd[i].prev.y; }
      if (v) { v[2 * i] = pd[i].vel.x; v[2 * i + 1] = pd[i].vel.y; }
      if (dens) { dens[2 * i] = pd[i].dens[0]; dens[2 * i + 1] = pd[i].dens[1]; }
      if (pres) { pres[2 * i] = pd[i].pres[0]; pres[2 * i + 1] = pd[i].pres[1]; }
      if (key) key[i] = pi[i].key;
    }
  }
  int64_t neighbors(int64_t i, uint32_t *out, int64_t cap) const override {
    const std::vector<size_t> &nb = pi[i].neighbors;
    for (int64_t k = 0; k < (int64_t)nb.size() && k < cap; ++k) out[k] = (uint32_t)nb[k];
    return (int64_t)nb.size();
  }
  // Cell contents in the reference's per-cell (swap-remove) order.
  void grid(uint32_t *order, uint32_t *cellStart) const override {
    uint32_t pos = 0;
    for (uint32_t k = 0; k < w.cells; ++k) {
      if (cellStart) cellStart[k] = pos;
      for (size_t j : cells[k]) { if (order) order[pos] = (uint32_t)j; ++pos; }
    }
    if (cellStart) cellStart[w.cells] = pos;
  }
  Stats stats() const override { return st; }
};

// ---------------------------------------------------------------------------
// Jacobi order with the canonical grid: a stable counting sort of the slot
// order by cell key (see bin()).
// ---------------------------------------------------------------------------
template <typename T>
struct JacobiSim : SimBase {
  World<T> w;
  std::vector<V2<T>> x, prev, v, a, dens, pres, tmp;
  // order: the slot order, a permutation of the particle indices.  Its first
  // nBinned entries are the last binning's cell lists; particles added since
  // follow in index order.
  std::vector<uint32_t> scene, key, order, cellStart;
  int64_t nBinned = 0;  // particles present at the last binning
  Stats st;
  std::unique_ptr<Pool> pool;
  std::vector<Emitter> emitters;
  float spacing;

  void setSpacing(float s) override { spacing = s; }
  JacobiSim(const ParamsF &pf, int threads) : spacing(pf.h * 0.5f) {
    w.p = Params<T>::from(pf);
    if (threads > 1) pool.reset(new Pool(threads));
    resetStats();
  }
  void resetStats() override {
    st = Stats{};
    st.minParticleNeighborCount = kMaxCellParticleCount;
    st.minCellParticleCount = kMaxCellParticleCount;
  }
  int addScene(float width, float height, float cs) override {
    int s = w.addScene(width, height, cs);
    cellStart.assign((size_t)w.cells + 1, 0u);
    nBinned = 0;  // tables change shape: the next step starts without lists
    return s;
  }
  int addBody(int sc, int type, const float *d, int nv) override { return w.addBody(sc, type, d, nv); }
  void setGravity(float gx, float gy) override { w.gravity = {T(gx), T(gy)}; }
  void setExternalForce(float fx, float fy) override { w.externalForce = {T(fx), T(fy)}; }
  void addEmitter(const Emitter &e) override { emitters.push_back(e); }
  int64_t count() const override { return (int64_t)x.size(); }
  int64_t totalCells() const override { return w.cells; }

  int64_t addParticles(int sc, const float *pos, const float *force, int64_t n) override {
    if (sc < 0 || sc >= (int)w.scenes.size()) return -1;
    int64_t first = (int64_t)x.size();
    for (int64_t k = 0; k < n; ++k) {
      V2<T> p{T(pos[2 * k]), T(pos[2 * k + 1])};
      x.push_back(p);
      prev.push_back(p);
      v.push_back({T(0), T(0)});
      a.push_back(force ? V2<T>{T(force[2 * k]), T(force[2 * k + 1])} : V2<T>{T(0), T(0)});
      dens.push_back({T(0), T(0)});
      pres.push_back({T(0), T(0)});
      scene.push_back((uint32_t)sc);
      key.push_back(cellKey(w.scenes[sc], p));
      order.push_back((uint32_t)(first + k));
    }
    return first;
  }
  void setState(const double *xs, const double *vs) override {
    for (size_t i = 0; i < x.size(); ++i) {
      x[i] = prev[i] = {T(xs[2 * i]), T(xs[2 * i + 1])};
      v[i] = {T(vs[2 * i]), T(vs[2 * i + 1])};
      a[i] = {T(0), T(0)};
      key[i] = cellKey(w.scenes[scene[i]], x[i]);
      order[i] = (uint32_t)i;  // the slot order restarts in index order
    }
    nBinned = 0;  // a restored state starts without neighbour lists
  }

  template <typename F>
  void parallel(int64_t n, F &&fn) {
    if (pool) pool->run(n, [&](int64_t s, int64_t e) { for (int64_t i = s; i <= e; ++i) fn(i); });
    else for (int64_t i = 0; i < n; ++i) fn(i);
  }
  template <typename F>
  void forCandidates(int64_t i, F &&f) const {
    w.forCandidateCells(key[i], (int)scene[i], [&](uint32_t k) {
      for (uint32_t q = cellStart[k]; q < cellStart[k + 1]; ++q) {
        uint32_t j = order[q];
        if (g_mutation == MUT_SKIP_SELF && j == (uint32_t)i) continue;
        f(j);
      }
    });
  }
  // Stable counting sort of the slot order by cell key: each cell lists its
  // particles in slot order, i.e. by their cell of the last binning and their
  // place in it, then the particles added since by index.  The slot order
  // starts in index order (and restarts there after setState), so a first
  // binning lists each cell by ascending index, the reference's order after
  // AddParticle (demo4.cpp:46-47).  The reference's later per-cell order is
  // its swap-remove history (demo4.cpp:56-76), which no parallel re-binning
  // reproduces; this order only keeps the candidate sets.
  void bin() {
    const int64_t n = (int64_t)x.size();
    std::fill(cellStart.begin(), cellStart.end(), 0u);
    for (int64_t i = 0; i < n; ++i) cellStart[key[i] + 1]++;
    uint64_t mx = 0, mn = kMaxCellParticleCount;
    for (uint32_t k = 0; k < w.cells; ++k) {
      uint64_t c = cellStart[k + 1];
      if (c > 0) { mx = std::max(mx, c); mn = std::min(mn, c); }
    }
    st.maxCellParticleCount = mx;
    st.minCellParticleCount = mn;
    for (uint32_t k = 0; k < w.cells; ++k) cellStart[k + 1] += cellStart[k];
    std::vector<uint32_t> next((size_t)n, 0u);
    std::vector<uint32_t> fill(cellStart.begin(), cellStart.end() - 1);
    for (int64_t q = 0; q < n; ++q) {
      const uint32_t i = g_mutation == MUT_GRID_BY_INDEX ? (uint32_t)q : order[q];
      next[fill[key[i]]++] = i;
    }
    order.swap(next);
    nBinned = n;
  }

  void update(T dt) {
    const T invDt = T(1) / dt;
    Clock::time_point t0 = Clock::now();  // demo4.cpp:290-299 emitters
    for (Emitter &e : emitters)
      runEmitter(e, (float)dt, spacing, (float)w.p.h,
                 [&](int sc, const float *pos, const float *force) { addParticles(sc, pos, force, 1); });
    const int64_t n = (int64_t)x.size();
    st.time[0] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:301-312 integrate
    const V2<T> gPlusE = vadd(w.gravity, w.externalForce);
    for (int64_t i = 0; i < n; ++i) {
      a[i] = vadd(gPlusE, a[i]);
      v[i] = vadd(vmul(a[i], dt), v[i]);
      a[i] = {T(0), T(0)};
    }
    st.time[1] = msSince(t0);

    // Viscosity, demo4.cpp:223-237 in Jacobi order.  Particle i receives the
    // half impulse of its own visit to j (v_i -= f_ij*0.5*dt) and of j's visit
    // to i (v_i += f_ji*0.5*dt with f_ji = -f_ij exactly), both computed from
    // the velocities after integration.  Lists come from the previous binning
    // (the reference's lists are built after its re-grid, demo4.cpp:358-379,
    // and first used by the next step's viscosity pass).
    t0 = Clock::now();
    tmp = v;
    if (nBinned > 0) {
      parallel(nBinned, [&](int64_t i) {
        V2<T> acc{T(0), T(0)};
        forCandidates(i, [&](uint32_t j) {
          V2<T> fij{T(0), T(0)}, fji{T(0), T(0)};
          computeViscosity(w.p, x[i], x[j], v[i], v[j], &fij);
          computeViscosity(w.p, x[j], x[i], v[j], v[i], &fji);
          acc = vsub(acc, vmul(vmul(fij, T(0.5f)), dt));
          acc = vadd(vmul(vmul(fji, T(0.5f)), dt), acc);
        });
        tmp[i] = vadd(acc, v[i]);
      });
    }
    v.swap(tmp);
    st.time[2] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:329-339 predict
    for (int64_t i = 0; i < n; ++i) {
      prev[i] = x[i];
      x[i] = vadd(vmul(v[i], dt), x[i]);
    }
    st.time[3] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:341-356 re-grid, as a canonical counting sort
    for (int64_t i = 0; i < n; ++i) key[i] = cellKey(w.scenes[scene[i]], x[i]);
    bin();
    st.time[4] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:358-379 candidate counts (lists stay implicit)
    uint64_t mn = kMaxParticleNeighborCount, mx = 0;
    for (int64_t i = 0; i < n; ++i) {
      uint64_t c = 0;
      forCandidates(i, [&](uint32_t) { ++c; });
      mn = std::min(mn, c);
      mx = std::max(mx, c);
    }
    st.minParticleNeighborCount = mn;
    st.maxParticleNeighborCount = mx;
    st.time[5] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:208-221 density and pressure
    parallel(n, [&](int64_t i) {
      T d[2] = {T(0), T(0)};
      forCandidates(i, [&](uint32_t j) { computeDensity(w.p, x[i], x[j], d); });
      T pr[2];
      computePressure(w.p, d, pr);
      dens[i] = {d[0], d[1]};
      pres[i] = {pr[0], pr[1]};
    });
    st.time[6] = msSince(t0);

    // Relaxation, demo4.cpp:239-255 in Jacobi order.  Particle i receives
    // -delta_ij*0.5 from its own visit and +delta_ji*0.5 from j's visit, where
    // delta_ji = d_ji * n_ji and n_ji = -n_ij exactly, so the second term is
    // -(d_ji * n_ij)*0.5.  Both use positions before the pass.
    t0 = Clock::now();
    tmp.resize((size_t)n);
    parallel(n, [&](int64_t i) {
      V2<T> acc{T(0), T(0)};
      const T pi_[2] = {pres[i].x, pres[i].y};
      forCandidates(i, [&](uint32_t j) {
        V2<T> dij{T(0), T(0)}, dji{T(0), T(0)};
        const T pj_[2] = {pres[j].x, pres[j].y};
        computeDelta(w.p, x[i], x[j], pi_, dt, &dij);
        computeDelta(w.p, x[j], x[i], pj_, dt, &dji);
        acc = vsub(acc, vmul(dij, T(0.5f)));
        acc = vadd(vmul(dji, T(0.5f)), acc);
      });
      tmp[i] = vadd(acc, x[i]);
    });
    x.swap(tmp);
    st.time[7] = msSince(t0);

    t0 = Clock::now();  // demo4.cpp:411-444 collisions
    for (int64_t i = 0; i < n; ++i) w.collide((int)scene[i], &x[i]);
    st.time[8] = msSince(t0);

    for (int64_t i = 0; i < n; ++i) {  // demo4.cpp:446-450 velocity
      if (g_mutation == MUT_VELOCITY_NO_INV_DT) v[i] = vsub(x[i], prev[i]);
      else v[i] = vmul(vsub(x[i], prev[i]), invDt);
    }
  }
  void step(float dt, int nsteps) override {
    for (int s = 0; s < nsteps; ++s) update(T(dt));
  }
  void getState(double *xo, double *po, double *vo, double *dn, double *pr, uint32_t *ko) const override {
    for (size_t i = 0; i < x.size(); ++i) {
      if (xo) { xo[2 * i] = x[i].x; xo[2 * i + 1] = x[i].y; }
      if (po) { po[2 * i] = prev[i].x; po[2 * i + 1] = prev[i].y; }
      if (vo) { vo[2 * i] = v[i].x; vo[2 * i + 1] = v[i].y; }
      if (dn) { dn[2 * i] = dens[i].x; dn[2 * i + 1] = dens[i].y; }
      if (pr) { pr[2 * i] = pres[i].x; pr[2 * i + 1] = pres[i].y; }
      if (ko) ko[i] = key[i];
    }
  }
  int64_t neighbors(int64_t i, uint32_t *out, int64_t cap) const override {
    int64_t c = 0;
    if (nBinned == 0 || i >= nBinned) return 0;
    forCandidates(i, [&](uint32_t j) { if (c < cap) out[c] = j; ++c; });
    return c;
  }
  void grid(uint32_t *o, uint32_t *cs) const override {
    if (o) for (size_t q = 0; q < order.size(); ++q) o[q] = order[q];
    if (cs) for (size_t k = 0; k < cellStart.size(); ++k) cs[k] = cellStart[k];
  }
  Stats stats() const override { return st; }
};

}  // namespace orc

using namespace orc;

extern "C" {

// mode: 0 = Jacobi (GPU order), 1 = reference Gauss-Seidel order (serial),
//       2 = reference order with the pooled phases (CPU baseline).
// precision: 32 or 64.  threads: worker threads for modes 0 and 2.
void *orc_create(int precision, int mode, int threads, const ParamsF *p) {
  if (mode == 0) {
    if (precision == 64) return new JacobiSim<double>(*p, threads);
    return new JacobiSim<float>(*p, threads);
  }
  int t = (mode == 2) ? std::max(1, threads) : 0;
  if (precision == 64) return new GSSim<double>(*p, t);
  return new GSSim<float>(*p, t);
}
void orc_destroy(void *h) { delete (SimBase *)h; }
int orc_add_scene(void *h, float width, float height, float cellSize) {
  return ((SimBase *)h)->addScene(width, height, cellSize);
}
int orc_add_body(void *h, int scene, int type, const float *data, int nverts) {
  return ((SimBase *)h)->addBody(scene, type, data, nverts);
}
int64_t orc_add_particles(void *h, int scene, const float *pos, const float *force, int64_t n) {
  return ((SimBase *)h)->addParticles(scene, pos, force, n);
}
void orc_set_gravity(void *h, float gx, float gy) { ((SimBase *)h)->setGravity(gx, gy); }
void orc_set_external_force(void *h, float fx, float fy) { ((SimBase *)h)->setExternalForce(fx, fy); }
void orc_step(void *h, float dt, int nsteps) { ((SimBase *)h)->step(dt, nsteps); }
// spacing: SPHParameters::particleSpacing (demo4.cpp:258); the constructor uses h/2.
void orc_set_spacing(void *h, float spacing) { ((SimBase *)h)->setSpacing(spacing); }
void orc_add_emitter(void *h, int scene, float px, float py, float dx, float dy, float radius, float speed,
                     float rate, float duration, uint64_t seed) {
  ((SimBase *)h)->addEmitter(Emitter{scene, px, py, dx, dy, radius, speed, rate, duration, 0.0f, 0.0f, true, seed, 0});
}
int64_t orc_count(void *h) { return ((SimBase *)h)->count(); }
int64_t orc_total_cells(void *h) { return ((SimBase *)h)->totalCells(); }
void orc_get_state(void *h, double *x, double *prev, double *v, double *dens, double *pres, uint32_t *key) {
  ((SimBase *)h)->getState(x, prev, v, dens, pres, key);
}
void orc_set_state(void *h, const double *x, const double *v) { ((SimBase *)h)->setState(x, v); }
int64_t orc_neighbors(void *h, int64_t i, uint32_t *out, int64_t cap) { return ((SimBase *)h)->neighbors(i, out, cap); }
void orc_grid(void *h, uint32_t *order, uint32_t *cellStart) { ((SimBase *)h)->grid(order, cellStart); }
void orc_get_stats(void *h, uint64_t counts[4], float times[9]) {
  Stats s = ((SimBase *)h)->stats();
  counts[0] = s.minParticleNeighborCount;
  counts[1] = s.maxParticleNeighborCount;
  counts[2] = s.minCellParticleCount;
  counts[3] = s.maxCellParticleCount;
  for (int k = 0; k < 9; ++k) times[k] = s.time[k];
}
void orc_reset_stats(void *h) { ((SimBase *)h)->resetStats(); }
void orc_set_mutation(int m) { g_mutation = m; }
int orc_get_mutation(void) { return g_mutation; }

// Single-function entry points for the known-answer checks.  They run the
// same templated functions the simulations use.
#define ORC_DISPATCH(prec, ...) \
  do {                          \
    if ((prec) == 64) {         \
      using T = double;         \
      __VA_ARGS__;              \
    } else {                    \
      using T = float;          \
      __VA_ARGS__;              \
    }                           \
  } while (0)

void orc_kat_density(int prec, const ParamsF *pf, const float *xi, const float *xj, int64_t n, double *out) {
  ORC_DISPATCH(prec, {
    Params<T> p = Params<T>::from(*pf);
    for (int64_t k = 0; k < n; ++k) {
      T d[2] = {T(0), T(0)};
      computeDensity(p, V2<T>{T(xi[2 * k]), T(xi[2 * k + 1])}, V2<T>{T(xj[2 * k]), T(xj[2 * k + 1])}, d);
      out[2 * k] = d[0];
      out[2 * k + 1] = d[1];
    }
  });
}
void orc_kat_pressure(int prec, const ParamsF *pf, const float *dens, int64_t n, double *out) {
  ORC_DISPATCH(prec, {
    Params<T> p = Params<T>::from(*pf);
    for (int64_t k = 0; k < n; ++k) {
      T d[2] = {T(dens[2 * k]), T(dens[2 * k + 1])}, o[2];
      computePressure(p, d, o);
      out[2 * k] = o[0];
      out[2 * k + 1] = o[1];
    }
  });
}
void orc_kat_delta(int prec, const ParamsF *pf, const float *xi, const float *xj, const float *pres, float dt,
                   int64_t n, double *out) {
  ORC_DISPATCH(prec, {
    Params<T> p = Params<T>::from(*pf);
    for (int64_t k = 0; k < n; ++k) {
      V2<T> o{T(0), T(0)};
      T pr[2] = {T(pres[2 * k]), T(pres[2 * k + 1])};
      computeDelta(p, V2<T>{T(xi[2 * k]), T(xi[2 * k + 1])}, V2<T>{T(xj[2 * k]), T(xj[2 * k + 1])}, pr, T(dt), &o);
      out[2 * k] = o.x;
      out[2 * k + 1] = o.y;
    }
  });
}
void orc_kat_viscosity(int prec, const ParamsF *pf, const float *xi, const float *xj, const float *vi,
                       const float *vj, int64_t n, double *out) {
  ORC_DISPATCH(prec, {
    Params<T> p = Params<T>::from(*pf);
    for (int64_t k = 0; k < n; ++k) {
      V2<T> o{T(0), T(0)};
      computeViscosity(p, V2<T>{T(xi[2 * k]), T(xi[2 * k + 1])}, V2<T>{T(xj[2 * k]), T(xj[2 * k + 1])},
                       V2<T>{T(vi[2 * k]), T(vi[2 * k + 1])}, V2<T>{T(vj[2 * k]), T(vj[2 * k + 1])}, &o);
      out[2 * k] = o.x;
      out[2 * k + 1] = o.y;
    }
  });
}
// type: 1 plane (data = nx, ny, d), 2 circle (cx, cy, r), 3 segment (ax, ay, bx, by),
// 4 polygon (nverts pairs).
void orc_kat_collide(int prec, const ParamsF *pf, int type, const float *data, int nverts, const float *pos,
                     int64_t n, double *out) {
  ORC_DISPATCH(prec, {
    World<T> w;
    w.p = Params<T>::from(*pf);
    w.addScene(1.0f, 1.0f, 1.0f);
    w.addBody(0, type, data, nverts);
    for (int64_t k = 0; k < n; ++k) {
      V2<T> x{T(pos[2 * k]), T(pos[2 * k + 1])};
      w.collide(0, &x);
      out[2 * k] = x.x;
      out[2 * k + 1] = x.y;
    }
  });
}
// sph.h:683-695 SPHGetParticleColor (FP32, std::min / std::max order).
void orc_kat_color(float restDensity, const float *density, const float *pressure, const float *vel, int64_t n,
                   float *rgba) {
  for (int64_t k = 0; k < n; ++k) {
    float r = pressure[k] / (-10.0f);
    float g = density[k] / (restDensity);
    float b = vlen(V2<float>{vel[2 * k], vel[2 * k + 1]}) / 10.0f;
    rgba[4 * k + 0] = std::max(std::min(r, 1.0f), 0.0f);
    rgba[4 * k + 1] = std::max(std::min(g, 1.0f), 0.0f);
    rgba[4 * k + 2] = std::max(std::min(b, 1.0f), 0.0f);
    rgba[4 * k + 3] = 1.0f;
  }
}

void orc_kat_cell(int prec, float width, float height, float cellSize, const float *pos, int64_t n, int32_t *out) {
  ORC_DISPATCH(prec, {
    World<T> w;
    w.addScene(width, height, cellSize);
    for (int64_t k = 0; k < n; ++k) {
      int cx, cy;
      cellKey(w.scenes[0], V2<T>{T(pos[2 * k]), T(pos[2 * k + 1])}, &cx, &cy);
      out[3 * k] = cx;
      out[3 * k + 1] = cy;
      out[3 * k + 2] = (int32_t)cellKey(w.scenes[0], V2<T>{T(pos[2 * k]), T(pos[2 * k + 1])});
    }
  });
}

}  // extern "C"
