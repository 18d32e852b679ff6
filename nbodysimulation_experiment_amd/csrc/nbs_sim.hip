// Host side of the MI355X SPH step library: device storage, step scheduling,
// HIP-graph replay, timing, and the C ABI declared in include/nbs_sph.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbs_sph.h"
#include "../../include/nbs_surrogate.h"
#include "nbs_kernels.hpp"
#include "nbs_surrogate.hpp"

namespace nbs {

static_assert(NBS_SURROGATE_PARAM_COUNT == kNumParams, "surrogate parameter count");
static_assert(NBS_SURROGATE_WEIGHT_COUNT == kNumWeights, "surrogate weight count");

static thread_local std::string g_lastError;

struct Error : std::runtime_error {
  nbs_status code;
  Error(nbs_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define NBS_HIP(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) throw Error(NBS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(NBS_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    n = count;
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

struct HostBody {
  int type;
  float a[4];
  std::vector<float> verts;
};

struct HostScene {
  nbs_domain dom;
  float halfW, halfH;
  int nx, ny;
  uint32_t cellBase;
  std::vector<HostBody> bodies;
};

struct Emitter {  // demo4.h:123-135
  int scene;
  float px, py, dx, dy, radius, speed, rate, duration, elapsed, totalElapsed;
  bool active;
  uint64_t seed, draws;
};

// SplitMix64 uniform in [-1, 1) with 24-bit resolution (exact in FP32).
static inline float splitmix_uniform(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  const float k = (float)(uint32_t)(z >> 40);
  return k * (1.0f / 8388608.0f) - 1.0f;
}

class Simulation {
 public:
  Simulation(const nbs_params &p, const nbs_domain &d, uint64_t capacity) : params_(p) {
    validateParams(p);
    int dev = 0;
    NBS_HIP(hipGetDevice(&dev));
    device_ = dev;
    NBS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    resetStats();
    addScene(d);
    reserve(std::max<uint64_t>(capacity, 1024));
  }
  ~Simulation() {
    destroyGraph();
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // ---- parameters -------------------------------------------------------
  static void validateParams(const nbs_params &p) {
    // The pair kernels' square root and reciprocal take r^2 < h^2 <= 2^80.
    if (!(p.kernel_height > 0.0f && p.kernel_height <= 0x1p40f))
      throw Error(NBS_ERR_INVALID_ARGUMENT, "kernel_height must be in (0, 2^40]");
    const float v[] = {p.rest_density, p.stiffness, p.near_stiffness, p.linear_viscosity, p.quadratic_viscosity,
                       p.collision_radius, p.collision_margin, p.particle_spacing, p.cell_size};
    for (float x : v)
      if (!std::isfinite(x)) throw Error(NBS_ERR_INVALID_ARGUMENT, "parameters must be finite");
  }
  void setParams(const nbs_params &p) {
    validateParams(p);
    params_ = p;
    bodiesDirty_ = true;  // circle cull boxes depend on the collision radius
  }
  const nbs_params &params() const { return params_; }
  void setGravity(float gx, float gy) { gravity_[0] = gx, gravity_[1] = gy; }
  void addExternalForce(float fx, float fy) {  // demo4.h:189-191 externalForce += force
    ext_[0] = fx + ext_[0];
    ext_[1] = fy + ext_[1];
  }
  void clearExternalForce() { ext_[0] = ext_[1] = 0.0f; }

  // ---- scenes and bodies -------------------------------------------------
  int addScene(const nbs_domain &d) {
    if (!(d.width > 0.0f) || !(d.height > 0.0f) || !(d.cell_size > 0.0f) || !std::isfinite(d.width) ||
        !std::isfinite(d.height) || !std::isfinite(d.cell_size))
      throw Error(NBS_ERR_INVALID_ARGUMENT, "domain sizes must be positive and finite");
    HostScene s;
    s.dom = d;
    s.halfW = d.width * 0.5f;  // sph.h:21-22
    s.halfH = d.height * 0.5f;
    const double fx = std::floor((double)(d.width / d.cell_size)), fy = std::floor((double)(d.height / d.cell_size));
    if (fx > 65536.0 || fy > 65536.0) throw Error(NBS_ERR_CAPACITY, "grid larger than 65536 cells per axis");
    s.nx = std::max(1, (int)(d.width / d.cell_size));  // sph.h:61-62
    s.ny = std::max(1, (int)(d.height / d.cell_size));
    const uint64_t total = (uint64_t)totalCells_ + (uint64_t)s.nx * (uint64_t)s.ny;
    if (total > (1ull << 31)) throw Error(NBS_ERR_CAPACITY, "more than 2^31 cells in total");
    s.cellBase = totalCells_;
    scenes_.push_back(s);
    totalCells_ = (uint32_t)total;
    allocCells();
    bodiesDirty_ = true;
    scenesDirty_ = true;
    // The cell tables change shape: the next step starts without lists.
    unbinAll();
    return (int)scenes_.size() - 1;
  }
  int sceneCount() const { return (int)scenes_.size(); }
  void checkScene(int s) const {
    if (s < 0 || s >= (int)scenes_.size()) throw Error(NBS_ERR_INVALID_ARGUMENT, "scene index out of range");
  }
  void clearBodies() {
    for (auto &s : scenes_) s.bodies.clear();
    bodiesDirty_ = true;
  }
  void addBody(int scene, HostBody b) {
    checkScene(scene);
    scenes_[scene].bodies.push_back(std::move(b));
    bodiesDirty_ = true;
  }

  // ---- particles ----------------------------------------------------------
  uint64_t count() const { return n_; }
  uint64_t addParticles(int scene, const float *pos, const float *force, uint64_t k) {
    checkScene(scene);
    if (k == 0) return n_;
    if (!pos) throw Error(NBS_ERR_INVALID_ARGUMENT, "positions are NULL");
    if (n_ + k > (uint64_t)INT32_MAX) throw Error(NBS_ERR_CAPACITY, "more than 2^31-1 particles");
    reserve(n_ + k);
    const HostScene &sc = scenes_[scene];
    std::vector<float> zero(2 * k, 0.0f), acc(2 * k, 0.0f);
    std::vector<uint32_t> ids(k), keys(k), scn(k, (uint32_t)scene);
    if (force) std::memcpy(acc.data(), force, 2 * k * sizeof(float));
    for (uint64_t q = 0; q < k; ++q) {
      ids[q] = (uint32_t)(n_ + q);
      keys[q] = hostKey(sc, pos[2 * q], pos[2 * q + 1]);  // demo4.cpp:151 InsertParticleIntoGrid
    }
    const size_t o = n_;
    copyIn(buf_.x0 + o, pos, 2 * k);
    copyIn(buf_.prev + o, pos, 2 * k);  // demo4.h:101-106 ParticleData(pos)
    copyIn(buf_.v + o, zero.data(), 2 * k);
    copyIn(buf_.dens + o, zero.data(), 2 * k);
    copyIn(buf_.pres + o, zero.data(), 2 * k);
    copyIn(buf_.acc + o, acc.data(), 2 * k);
    NBS_HIP(hipMemcpyAsync(buf_.id + o, ids.data(), k * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(buf_.keys + o, keys.data(), k * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(buf_.scn + o, scn.data(), k * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    const uint64_t first = n_;
    n_ += k;
    dropCarriedDensity();
    destroyGraph();
    return first;
  }
  void clearParticles() {
    n_ = 0;
    nBinned_ = 0;
    lists_ = false;
    dropCarriedDensity();
    destroyGraph();
  }
  // demo4.cpp:169-181 AddVolume with seeded per-axis jitter.
  void addVolume(int scene, float cx, float cy, float fx, float fy, int countX, int countY, float spacing,
                 uint64_t seed) {
    checkScene(scene);
    if (countX < 0 || countY < 0) throw Error(NBS_ERR_INVALID_ARGUMENT, "negative volume count");
    const uint64_t k = (uint64_t)countX * (uint64_t)countY;
    if (k == 0) return;
    std::vector<float> pos(2 * k), force(2 * k);
    const float offX = ((float)countX * spacing) * 0.5f;  // demo4.cpp:170
    const float offY = ((float)countY * spacing) * 0.5f;
    const float h = params_.kernel_height;
    uint64_t q = 0;
    for (int iy = 0; iy < countY; ++iy) {
      for (int ix = 0; ix < countX; ++ix, ++q) {
        float px = (float)ix * spacing, py = (float)iy * spacing;  // demo4.cpp:173
        const float half = spacing * 0.5f;
        px = half + px;                                           // demo4.cpp:174
        py = half + py;
        px = (cx - offX) + px;                                    // demo4.cpp:175
        py = (cy - offY) + py;
        const float jx = (splitmix_uniform(seed, 2 * q) * h) * 0.01f;  // demo4.cpp:176
        const float jy = (splitmix_uniform(seed, 2 * q + 1) * h) * 0.01f;
        pos[2 * q] = jx + px;                                     // demo4.cpp:177
        pos[2 * q + 1] = jy + py;
        force[2 * q] = fx;
        force[2 * q + 1] = fy;
      }
    }
    addParticles(scene, pos.data(), force.data(), k);
  }
  void addEmitter(const Emitter &e) {
    checkScene(e.scene);
    emitters_.push_back(e);
  }
  void clearEmitters() { emitters_.clear(); }

  // ---- state ------------------------------------------------------------
  void getState(float *x, float *prev, float *v, float *dens, float *pres, uint32_t *key, int32_t *scene) {
    if (n_ == 0) return;
    ensureStaging();
    launch_gather_state(buf_, devScene0(), (int)n_, (int)nBinned_, x ? stX_.p : nullptr, prev ? stPrev_.p : nullptr,
                        v ? stV_.p : nullptr, dens ? stDens_.p : nullptr, pres ? stPres_.p : nullptr,
                        key ? stKey_.p : nullptr, scene ? stScene_.p : nullptr, stream_);
    NBS_HIP(hipGetLastError());
    if (x) copyOut(x, stX_.p, 2 * n_);
    if (prev) copyOut(prev, stPrev_.p, 2 * n_);
    if (v) copyOut(v, stV_.p, 2 * n_);
    if (dens) copyOut(dens, stDens_.p, 2 * n_);
    if (pres) copyOut(pres, stPres_.p, 2 * n_);
    if (key) NBS_HIP(hipMemcpyAsync(key, stKey_.p, n_ * 4, hipMemcpyDeviceToHost, stream_));
    if (scene) NBS_HIP(hipMemcpyAsync(scene, stScene_.p, n_ * 4, hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
  }
  // sph.h:683-695 per-particle colour, as demo4.cpp:521-524 computes it for Render.
  void getColors(float *rgba) {
    if (n_ == 0) return;
    if (stColor_.n < n_) stColor_.alloc(cap_);
    launch_colors(buf_, params_.rest_density, (int)n_, stColor_.p, stream_);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipMemcpyAsync(rgba, stColor_.p, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
  }
  void setState(const float *x, const float *v) {
    if (n_ == 0) return;
    if (!x || !v) throw Error(NBS_ERR_INVALID_ARGUMENT, "x and v are required");
    std::vector<int32_t> scene(n_);
    getState(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, scene.data());
    std::vector<uint32_t> ids(n_), keys(n_), scn(n_);
    std::vector<float> zero(2 * n_, 0.0f);
    for (uint64_t i = 0; i < n_; ++i) {
      ids[i] = (uint32_t)i;
      scn[i] = (uint32_t)scene[i];
      keys[i] = hostKey(scenes_[scene[i]], x[2 * i], x[2 * i + 1]);
    }
    copyIn(buf_.x0, x, 2 * n_);
    copyIn(buf_.prev, x, 2 * n_);
    copyIn(buf_.v, v, 2 * n_);
    copyIn(buf_.acc, zero.data(), 2 * n_);
    copyIn(buf_.dens, zero.data(), 2 * n_);
    copyIn(buf_.pres, zero.data(), 2 * n_);
    NBS_HIP(hipMemcpyAsync(buf_.id, ids.data(), n_ * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(buf_.keys, keys.data(), n_ * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(buf_.scn, scn.data(), n_ * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    nBinned_ = 0;
    lists_ = false;
    dropCarriedDensity();
    destroyGraph();
  }
  void getGrid(uint32_t *sortedIds, uint32_t *cellStart) {
    if (sortedIds && n_) NBS_HIP(hipMemcpyAsync(sortedIds, buf_.id, n_ * 4, hipMemcpyDeviceToHost, stream_));
    if (cellStart)
      NBS_HIP(hipMemcpyAsync(cellStart, buf_.startCur, ((size_t)totalCells_ + 1) * 4, hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
  }
  uint64_t totalCells() const { return totalCells_; }

  // ---- statistics ---------------------------------------------------------
  void resetStats() {
    std::memset(&stats_h_, 0, sizeof(stats_h_));
    stats_h_.min_particle_neighbor_count = 500;  // sph.h:143-147
    stats_h_.min_cell_particle_count = 500;
    for (int k = 0; k < NBS_KERNEL_COUNT; ++k) kernelMs_[k] = 0.0, kernelLaunches_[k] = 0;
  }
  // Reduces the per-workgroup minima and maxima the last step left behind.
  nbs_stats stats() {
    if (statsPending_ && n_ > 0) {
      const size_t nc = 2 * (size_t)statSlots_;
      const size_t ns = 2 * (size_t)cellSlots_;
      std::vector<uint32_t> c(nc), s(ns);
      NBS_HIP(hipMemcpyAsync(c.data(), d_.candStats.p, nc * 4, hipMemcpyDeviceToHost, stream_));
      NBS_HIP(hipMemcpyAsync(s.data(), d_.cellStats.p, ns * 4, hipMemcpyDeviceToHost, stream_));
      NBS_HIP(hipStreamSynchronize(stream_));
      uint32_t cmn = 0xFFFFFFFFu, cmx = 0, smn = 0xFFFFFFFFu, smx = 0;
      for (size_t k = 0; k < nc; k += 2) cmn = std::min(cmn, c[k]), cmx = std::max(cmx, c[k + 1]);
      for (size_t k = 0; k < ns; k += 2) smn = std::min(smn, s[k]), smx = std::max(smx, s[k + 1]);
      stats_h_.min_particle_neighbor_count = cmn;
      stats_h_.max_particle_neighbor_count = cmx;
      stats_h_.min_cell_particle_count = (smn == 0xFFFFFFFFu) ? 0 : smn;
      stats_h_.max_cell_particle_count = smx;
    }
    statsPending_ = false;
    return stats_h_;
  }
  void kernelTimes(double *avg, uint64_t *launches) const {
    for (int k = 0; k < NBS_KERNEL_COUNT; ++k) {
      if (avg) avg[k] = kernelLaunches_[k] ? kernelMs_[k] / (double)kernelLaunches_[k] : 0.0;
      if (launches) launches[k] = kernelLaunches_[k];
    }
  }
  void setTiming(bool on) {
    timing_ = on;
    destroyGraph();
  }
  void setOptions(uint32_t f) {
    options_ = f;
    destroyGraph();
  }

  nbs_diagnostics diagnostics() {
    nbs_diagnostics d{};
    d.particle_count = n_;
    if (n_ == 0) return d;
    const int nb = diagnostics_blocks((int)n_);
    if (diagPartials_.n < (size_t)nb * kDiagWidth) diagPartials_.alloc((size_t)nb * kDiagWidth);
    if (diagOut_.n < kDiagWidth) diagOut_.alloc(kDiagWidth);
    launch_diagnostics(buf_, devParams(0.0f), (int)n_, diagPartials_.p, diagOut_.p, stream_);
    NBS_HIP(hipGetLastError());
    double r[kDiagWidth];
    NBS_HIP(hipMemcpyAsync(r, diagOut_.p, sizeof(r), hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    diagnostics_record(r, n_, (double)params_.rest_density, &d);
    return d;
  }

  // ---- the step -----------------------------------------------------------
  void update(float dt, uint32_t nsteps) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) throw Error(NBS_ERR_INVALID_ARGUMENT, "dt must be positive and finite");
    syncBodies();
    uint32_t s = 0, slots = 0;  // slots: steps with events this call
    while (s < nsteps) {
      if (!emitters_.empty()) runEmitters(dt);  // demo4.cpp:290-299
      if (n_ == 0) {
        ++s;
        continue;
      }
      if ((tuneActive_ || wantTune()) && !surrogateMode()) {
        tuneStep(dt, slots);  // one timed step of the next candidate
        ++slots;
        ++s;
        continue;
      }
      const bool graphable = (options_ & NBS_OPT_HIP_GRAPH) && !timing_ && emitters_.empty() && lists_ &&
                             nBinned_ == n_ && !surrogateMode();
      if (graphable) {
        // Replay pairs of steps (the cell tables and the slot-id arrays
        // alternate between two buffers every step).
        uint32_t pairs = (nsteps - s) / 2;
        if (tuning_.autotune && tuning_.retune_interval > 0 && n_ >= kTuneMinParticles) {
          // stop replaying when the next re-tuning is due
          const uint64_t until = tuning_.retune_interval > stepsSinceTune_ ? tuning_.retune_interval - stepsSinceTune_ : 0;
          pairs = std::min<uint32_t>(pairs, (uint32_t)std::max<uint64_t>(1, until / 2));
        }
        if (pairs > 0) {
          ensureGraph(dt);
          for (uint32_t q = 0; q < pairs; ++q) NBS_HIP(hipGraphLaunch(graphExec_, stream_));
          s += 2 * pairs;
          stepsSinceTune_ += 2 * pairs;
          continue;
        }
      }
      int ev = -1;
      if (timing_) {
        ensureEvents(slots + 1);
        ev = (int)slots++;
      }
      int relax = surrogateMode() ? nextRelax() : RELAX_PLAIN;
      if (relax != RELAX_PLAIN) {
        ensureSurInference();
        ensureSceneErr();
        if (relax == RELAX_CHECK && sur_->onlineLr > 0.0f) ensureSurTraining();
      }
      if (relax == RELAX_SURROGATE && carryReady()) relax = RELAX_SURROGATE_CARRIED;
      enqueueStep(dt, ev, relax);
      swapAfterStep();
      if (relax == RELAX_CHECK) finishCheck();
      ++stepsSinceTune_;
      ++s;
    }
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipStreamSynchronize(stream_));
    if (!tunePending_.empty()) absorbTuneSlots();  // this call's tuning steps (event slots are per call)
    if (timing_ && slots > 0) collectTimes(slots);
    statsPending_ = true;
    if (options_ & NBS_OPT_CHECK_FINITE) {
      nbs_diagnostics d = diagnostics();
      if (d.non_finite_count) throw Error(NBS_ERR_NON_FINITE, "particle state holds NaN or Inf");
    }
  }

  void setTuning(const nbs_tuning &t) {
    for (int k = 0; k < 3; ++k)
      if (t.tile_cap[k] != 0 && t.tile_cap[k] != 1024 && t.tile_cap[k] != 2048 && t.tile_cap[k] != 4096 &&
          t.tile_cap[k] != kWideCap)
        throw Error(NBS_ERR_INVALID_ARGUMENT, "tile capacities are 0, 1024, 2048, 4096 or -1 (wide)");
    if (t.wide_max_particles < 0) throw Error(NBS_ERR_INVALID_ARGUMENT, "wide_max_particles must be >= 0");
    if (t.autotune < 0 || t.autotune > 2) throw Error(NBS_ERR_INVALID_ARGUMENT, "autotune is 0, 1 or 2");
    tuning_ = t;
    // 2: tile_cap holds an earlier tuning's result; only the periodic
    // re-tuning runs.
    tuned_ = t.autotune == 2;
    if (tuning_.autotune == 2) tuning_.autotune = 1;
    tuneActive_ = false;
    tunePending_.clear();
    stepsSinceTune_ = 0;
    destroyGraph();
  }
  nbs_tuning tuning() const { return tuning_; }

  hipStream_t stream() const { return stream_; }

  // ---- learned relaxation ---------------------------------------------------
  void surInit(uint64_t seed) {
    if (!sur_) sur_.reset(new SurrogateState());
    SurrogateState &S = *sur_;
    S.params.alloc(kNumParams);
    S.adamM.alloc(kNumWeights);
    S.adamV.alloc(kNumWeights);
    S.grad.alloc(kNumWeights);
    S.w1b.alloc(kHid * kFeat);
    S.w2b.alloc(kHid * kHid);
    S.w3b.alloc(kHid * kHid);
    S.w2t.alloc(kHid * kHid);
    S.w3t.alloc(kHid * kHid);
    std::vector<float> h(kNumParams, 0.0f);
    uint64_t draw = 0;
    auto fill = [&](int off, int count, float bound) {
      for (int k = 0; k < count; ++k) h[off + k] = bound * splitmix_uniform(seed, draw++);
    };
    fill(kOffW1, kHid * kFeat, std::sqrt(6.0f / 11.0f));      // 11 features in use
    fill(kOffW2, kHid * kHid, std::sqrt(6.0f / kHid) / 8.0f);  // inputs are sums over pairs
    fill(kOffW3, kHid * kHid, std::sqrt(6.0f / kHid));
    fill(kOffW4, 2 * kHid, 0.1f * std::sqrt(6.0f / kHid));
    const float norm[8] = {1.0f, 0.0f, 1.0f, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f};
    std::memcpy(&h[kOffNorm], norm, sizeof(norm));
    uploadParams(h.data());
    std::vector<float> zero(kNumWeights, 0.0f);
    NBS_HIP(hipMemcpyAsync(S.adamM.p, zero.data(), kNumWeights * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(S.adamV.p, zero.data(), kNumWeights * 4, hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    S.adamT = 0;
    S.normSet = false;
    S.rep = nbs_surrogate_report{};
    S.errSum = 0.0;
  }
  void surSetParams(const float *h, uint64_t count) {
    if (!h || count != (uint64_t)kNumParams)
      throw Error(NBS_ERR_INVALID_ARGUMENT, "the parameter vector has NBS_SURROGATE_PARAM_COUNT floats");
    if (!sur_) surInit(0);
    uploadParams(h);
    sur_->normSet = true;
  }
  void surGetParams(float *h, uint64_t count) {
    requireSur();
    if (!h || count != (uint64_t)kNumParams)
      throw Error(NBS_ERR_INVALID_ARGUMENT, "the parameter vector has NBS_SURROGATE_PARAM_COUNT floats");
    NBS_HIP(hipMemcpyAsync(h, sur_->params.p, kNumParams * 4, hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
  }
  void surSetMode(int mode, uint32_t interval, float maxErr) {
    if (mode != NBS_RELAX_PLAIN && mode != NBS_RELAX_SURROGATE && mode != NBS_RELAX_SURROGATE_CARRIED_DENSITY)
      throw Error(NBS_ERR_INVALID_ARGUMENT, "unknown mode");
    if (mode != NBS_RELAX_PLAIN) {
      requireSur();
      if (interval == 0) throw Error(NBS_ERR_INVALID_ARGUMENT, "check_interval must be at least 1");
      if (!(maxErr >= 0.0f) || !std::isfinite(maxErr))
        throw Error(NBS_ERR_INVALID_ARGUMENT, "max_density_error must be finite and non-negative");
    }
    if (!sur_) return;
    SurrogateState &S = *sur_;
    S.mode = mode;
    S.interval = interval;
    S.maxErr = maxErr;
    S.sinceCheck = 0;
    S.plainLeft = 0;
    S.rep.surrogate_steps = S.rep.plain_steps = S.rep.check_steps = S.rep.fallbacks = 0;
    S.rep.last_density_error = S.rep.max_density_error = S.rep.mean_density_error = 0.0;
    S.rep.last_max_scene_error = 0.0;
    S.rep.last_scenes_over_limit = 0;
    S.errSum = 0.0;
    S.carrySrc = nullptr;
    destroyGraph();
  }
  void surSetOnlineTraining(float lr) {
    requireSur();
    if (!(lr >= 0.0f) || !std::isfinite(lr)) throw Error(NBS_ERR_INVALID_ARGUMENT, "learning_rate must be >= 0");
    if (lr > 0.0f && !sur_->normSet)
      throw Error(NBS_ERR_INVALID_ARGUMENT, "online training needs the normalisation (calibrate or set_params)");
    sur_->onlineLr = lr;
  }
  void surSceneErrors(double *out, uint64_t count) {
    requireSur();
    if (!out || count != scenes_.size()) throw Error(NBS_ERR_INVALID_ARGUMENT, "count must be the scene count");
    SurrogateState &S = *sur_;
    if (S.rep.check_steps == 0 || S.sceneErr.n < 2 * count) {
      for (uint64_t k = 0; k < count; ++k) out[k] = 0.0;
      return;
    }
    std::vector<double> v(2 * count);
    NBS_HIP(hipMemcpyAsync(v.data(), S.sceneErr.p, v.size() * 8, hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    for (uint64_t k = 0; k < count; ++k) out[k] = v[2 * k + 1] != 0.0 ? v[2 * k] / std::fabs(v[2 * k + 1]) : 0.0;
  }
  nbs_surrogate_report surReport() {
    if (!sur_) return nbs_surrogate_report{};
    nbs_surrogate_report r = sur_->rep;
    r.mean_density_error = r.check_steps ? sur_->errSum / (double)r.check_steps : 0.0;
    r.last_pair_count = 0;
    if (sur_->pairWaves > 0 && sur_->pairPart.p) {
      std::vector<uint32_t> c((size_t)sur_->pairWaves);
      NBS_HIP(hipMemcpyAsync(c.data(), sur_->pairPart.p, c.size() * 4, hipMemcpyDeviceToHost, stream_));
      NBS_HIP(hipStreamSynchronize(stream_));
      for (uint32_t v : c) r.last_pair_count += v;
    }
    return r;
  }
  void surCalibrate(float dt) {
    requireSur();
    std::vector<float> x, prev, dens, target;
    runCapturedStep(dt, RELAX_CALIBRATE, 0, &x, &prev, &dens, &target, nullptr);
    const uint64_t n = n_;
    double sr = 0, sr2 = 0, sn = 0, sn2 = 0, sv2 = 0, so2 = 0;
    const double invDt = (double)(1.0f / dt);
    for (uint64_t i = 0; i < n; ++i) {
      const double r = dens[2 * i], rn = dens[2 * i + 1];
      sr += r, sr2 += r * r, sn += rn, sn2 += rn * rn;
      const double vx = ((double)x[2 * i] - prev[2 * i]) * invDt, vy = ((double)x[2 * i + 1] - prev[2 * i + 1]) * invDt;
      sv2 += vx * vx + vy * vy;
      so2 += (double)target[2 * i] * target[2 * i] + (double)target[2 * i + 1] * target[2 * i + 1];
    }
    const double N = (double)std::max<uint64_t>(n, 1);
    const double mr = sr / N, mn = sn / N;
    const double vr = std::max(sr2 / N - mr * mr, 0.0), vn = std::max(sn2 / N - mn * mn, 0.0);
    float norm[8];
    norm[0] = (float)(1.0 / std::max(std::sqrt(sv2 / (2.0 * N)), 1e-6));
    norm[1] = (float)mr;
    norm[2] = (float)(1.0 / std::max(std::sqrt(vr), 1e-6));
    norm[3] = (float)mn;
    norm[4] = (float)(1.0 / std::max(std::sqrt(vn), 1e-6));
    norm[5] = (float)std::max(std::sqrt(so2 / (2.0 * N)), 1e-12);
    norm[6] = norm[7] = 0.0f;
    NBS_HIP(hipMemcpyAsync(sur_->params.p + kOffNorm, norm, sizeof(norm), hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    sur_->normSet = true;
  }
  void surTrain(float dt, uint32_t steps, float lr, float *losses, float *relLosses) {
    requireSur();
    if (!(lr > 0.0f) || !std::isfinite(lr)) throw Error(NBS_ERR_INVALID_ARGUMENT, "learning_rate must be positive");
    if (!(dt > 0.0f) || !std::isfinite(dt)) throw Error(NBS_ERR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (!sur_->normSet) surCalibrate(dt);
    SurrogateState &S = *sur_;
    if (S.loss.n < 2 * (size_t)std::max<uint32_t>(steps, 1)) S.loss.alloc(2 * (size_t)std::max<uint32_t>(steps, 1));
    S.lr = lr;
    syncBodies();
    for (uint32_t s = 0; s < steps; ++s) {
      if (!emitters_.empty()) runEmitters(dt);
      if (n_ == 0) continue;
      ensureSurTraining();
      enqueueStep(dt, -1, RELAX_TRAIN, (int)s);
      swapAfterStep();
      ++stepsSinceTune_;
    }
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipStreamSynchronize(stream_));
    if ((losses || relLosses) && steps) {
      std::vector<float> l(2 * (size_t)steps);
      NBS_HIP(hipMemcpy(l.data(), S.loss.p, l.size() * 4, hipMemcpyDeviceToHost));
      for (uint32_t k = 0; k < steps; ++k) {
        if (losses) losses[k] = l[2 * k];
        if (relLosses) relLosses[k] = l[2 * k + 1];
      }
    }
    statsPending_ = true;
    destroyGraph();
  }
  void surFrame(float dt, float *grad, float *loss, float *out, float *x, float *prev, float *dens, float *target,
                int32_t *scene) {
    requireSur();
    std::vector<float> hx, hp, hd, ht, ho;
    std::vector<uint32_t> keys;
    runCapturedStep(dt, RELAX_FRAME, 0, x ? &hx : nullptr, prev ? &hp : nullptr, dens ? &hd : nullptr,
                    target ? &ht : nullptr, scene ? &keys : nullptr);
    const uint64_t n = n_;
    if (grad) NBS_HIP(hipMemcpy(grad, sur_->grad.p, kNumWeights * 4, hipMemcpyDeviceToHost));
    if (loss) NBS_HIP(hipMemcpy(loss, sur_->loss.p, 4, hipMemcpyDeviceToHost));
    if (out && n) NBS_HIP(hipMemcpy(out, sur_->out.p, n * 8, hipMemcpyDeviceToHost));
    if (x) std::memcpy(x, hx.data(), n * 8);
    if (prev) std::memcpy(prev, hp.data(), n * 8);
    if (dens) std::memcpy(dens, hd.data(), n * 8);
    if (target) std::memcpy(target, ht.data(), n * 8);
    if (scene) {
      std::vector<uint32_t> bases;
      for (const HostScene &sc : scenes_) bases.push_back(sc.cellBase);
      for (uint64_t i = 0; i < n; ++i)
        scene[i] = (int32_t)(std::upper_bound(bases.begin(), bases.end(), keys[i]) - bases.begin() - 1);
    }
  }

 private:
  // FP32 derivations in the reference's order.
  DevParams devParams(float dt) const {
    DevParams p;
    p.h = params_.kernel_height;
    p.invH = 1.0f / p.h;                       // sph.h:92
    p.h2 = p.h * p.h;                          // sph.h:470
    p.rho0 = params_.rest_density;
    p.k = params_.stiffness;
    p.kn = params_.near_stiffness;
    p.sigma = params_.linear_viscosity;
    p.beta = params_.quadratic_viscosity;
    p.R = params_.collision_radius;
    p.margin = params_.collision_margin;
    p.dt = dt;
    p.dt2 = dt * dt;                           // sph.h:492
    p.invDt = (dt != 0.0f) ? 1.0f / dt : 0.0f; // demo4.cpp:287
    p.gex = gravity_[0] + ext_[0];             // demo4.cpp:306 gravity + externalForce
    p.gey = gravity_[1] + ext_[1];
    p.gdtx = (p.gex + 0.0f) * dt;              // demo4.cpp:306-307 with a zero own acceleration
    p.gdty = (p.gey + 0.0f) * dt;
    return p;
  }
  DevScene devScene(const HostScene &s, int bodyBegin, int bodyEnd) const {
    DevScene d;
    d.halfW = s.halfW;
    d.halfH = s.halfH;
    d.cellSize = s.dom.cell_size;
    d.nx = s.nx;
    d.ny = s.ny;
    d.cellBase = s.cellBase;
    d.bodyBegin = bodyBegin;
    d.bodyEnd = bodyEnd;
    return d;
  }
  DevScene devScene0() const { return hostScenes_.empty() ? devScene(scenes_[0], 0, 0) : hostScenes_[0]; }

  static int hostCoord(float p, float half, float cs, int n) {  // sph.h:450-463
    const float f = (p + half) / cs;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)n) return n - 1;
    return (int)f;
  }
  static uint32_t hostKey(const HostScene &s, float x, float y) {
    const int cx = hostCoord(x, s.halfW, s.dom.cell_size, s.nx);
    const int cy = hostCoord(y, s.halfH, s.dom.cell_size, s.ny);
    return s.cellBase + (uint32_t)(cy * s.nx + cx);
  }

  void copyIn(float2 *dst, const float *src, size_t floats) {
    NBS_HIP(hipMemcpyAsync(dst, src, floats * sizeof(float), hipMemcpyHostToDevice, stream_));
  }
  void copyOut(float *dst, const float2 *src, size_t floats) {
    NBS_HIP(hipMemcpyAsync(dst, src, floats * sizeof(float), hipMemcpyDeviceToHost, stream_));
  }

  void reserve(uint64_t want) {
    if (want <= cap_) return;
    const uint64_t cap = std::max<uint64_t>(want, cap_ * 2);
    // persistent arrays keep their contents
    auto grow2 = [&](DevBuf<float2> &b) {
      DevBuf<float2> nb;
      nb.alloc(cap);
      if (b.p && n_) NBS_HIP(hipMemcpyAsync(nb.p, b.p, n_ * sizeof(float2), hipMemcpyDeviceToDevice, stream_));
      NBS_HIP(hipStreamSynchronize(stream_));
      b.swap(nb);
    };
    auto grow4 = [&](DevBuf<uint32_t> &b) {
      DevBuf<uint32_t> nb;
      nb.alloc(cap);
      if (b.p && n_) NBS_HIP(hipMemcpyAsync(nb.p, b.p, n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
      NBS_HIP(hipStreamSynchronize(stream_));
      b.swap(nb);
    };
    grow2(d_.x0); grow2(d_.prev); grow2(d_.v); grow2(d_.dens); grow2(d_.pres); grow2(d_.acc);
    grow4(d_.id); grow4(d_.keys); grow4(d_.scn);
    d_.xp.alloc(cap); d_.x1.alloc(cap); d_.id2.alloc(cap); d_.keyp.alloc(cap); d_.fill.alloc(cap);
    d_.tmpSlot.alloc(cap); d_.binRec.alloc(cap); d_.farList.alloc(cap);
    d_.chunkCells.alloc(2 * ((cap + kPlaceChunk - 1) / kPlaceChunk));
    d_.candStats.alloc(2 * (size_t)density_stat_slots((int)cap, kWideCap));
    cap_ = cap;
    rebindBuffers();
    destroyGraph();
  }
  void allocCells() {
    cellSceneCells_ = 0;  // the per-cell scene table no longer matches (rebuilt by syncBodies)
    const size_t c = totalCells_;
    const size_t nb = (c + kScanChunk - 1) / kScanChunk;
    d_.count.alloc(nb * kScanChunk);  // whole scan chunks: a scan thread's four counts are one load
    d_.startCur.alloc(c + 1);
    d_.startNext.alloc(c + 1);
    d_.blockSum.alloc(std::max<size_t>(nb, 1));
    d_.cellStats.alloc(2 * std::max<size_t>(nb, 1));
    if (!d_.farCounts.p) d_.farCounts.alloc(2);
    launch_fill_u32(d_.farCounts.p, 0u, 2, stream_);
    launch_fill_u32(d_.count.p, 0u, nb * kScanChunk, stream_);  // the entries past the last cell stay zero
    launch_fill_u32(d_.startCur.p, 0u, c + 1, stream_);
    launch_fill_u32(d_.startNext.p, 0u, c + 1, stream_);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipStreamSynchronize(stream_));
    rebindBuffers();
    destroyGraph();
  }
  void unbinAll() {
    if (nBinned_ > 0) {
      syncBodies();
      launch_unbin(buf_, (int)nBinned_, stream_);
      NBS_HIP(hipGetLastError());
      NBS_HIP(hipStreamSynchronize(stream_));
    }
    nBinned_ = 0;
    lists_ = false;
    dropCarriedDensity();
    destroyGraph();
  }
  void ensureStaging() {
    if (stX_.n >= n_) return;
    const size_t c = cap_;
    stX_.alloc(c); stPrev_.alloc(c); stV_.alloc(c); stDens_.alloc(c); stPres_.alloc(c);
    stKey_.alloc(c); stScene_.alloc(c);
  }
  void rebindBuffers() {
    buf_.x0 = d_.x0.p; buf_.xp = d_.xp.p; buf_.x1 = d_.x1.p; buf_.prev = d_.prev.p; buf_.v = d_.v.p;
    buf_.dens = d_.dens.p; buf_.pres = d_.pres.p; buf_.acc = d_.acc.p;
    buf_.id = d_.id.p; buf_.id2 = d_.id2.p; buf_.keys = d_.keys.p; buf_.keyp = d_.keyp.p; buf_.fill = d_.fill.p;
    buf_.tmpSlot = d_.tmpSlot.p; buf_.binRec = d_.binRec.p; buf_.scn = d_.scn.p;
    buf_.farList = nullptr; buf_.farCount = buf_.farCountClear = nullptr; buf_.chunkCells = nullptr;
    buf_.count = d_.count.p; buf_.startCur = d_.startCur.p; buf_.startNext = d_.startNext.p;
    buf_.blockSum = d_.blockSum.p;
    buf_.candStats = d_.candStats.p; buf_.cellStats = d_.cellStats.p;
    buf_.scenes = devScenes_.p; buf_.sceneBase = devSceneBase_.p; buf_.nScenes = (int)scenes_.size();
    // the per-cell scene table once syncBodies has built it for these cells
    buf_.cellScene = (scenes_.size() > 1 && cellSceneCells_ == totalCells_) ? devCellScene_.p : nullptr;
    buf_.bodies = devBodies_.p; buf_.verts = devVerts_.p;
  }
  // Cull box of a circle (collide_wave): positions x < cx - reach or
  // x > cx + reach (likewise y), reach = both * (1 + 2^-8) + 2^-40 with
  // both = r + R as the kernel forms it, fail the kernel's FP32 test
  // d2 <= both * both: |fl(x - cx)| >= reach * (1 - 2^-24), so d2 >= reach^2 *
  // (1 - 2^-24)^3 > both^2 * (1 + 2^-24) >= fl(both * both).  The bounds are
  // rounded outwards.  Other bodies get the full range (never skipped).
  void setCullBox(DevBody &d) const {
    d.cull[0] = d.cull[1] = INT_MIN;
    d.cull[2] = d.cull[3] = INT_MAX;
    if (d.type != BODY_CIRCLE) return;
    const float both = d.a2 + params_.collision_radius;  // sph.h:527
    if (!std::isfinite(both)) return;
    const double reach = (double)both * (1.0 + 0x1p-8) + 0x1p-40;
    auto down = [](double v) {
      float f = (float)v;
      if ((double)f > v) f = std::nextafter(f, -INFINITY);
      return f;
    };
    auto up = [](double v) {
      float f = (float)v;
      if ((double)f < v) f = std::nextafter(f, INFINITY);
      return f;
    };
    auto order = [](float f) {
      int bits;
      std::memcpy(&bits, &f, 4);
      return float_order(bits);
    };
    d.cull[0] = order(down((double)d.a0 - reach));
    d.cull[1] = order(down((double)d.a1 - reach));
    d.cull[2] = order(up((double)d.a0 + reach));
    d.cull[3] = order(up((double)d.a1 + reach));
  }
  // Uploads scene descriptors and the body table when they changed.
  void syncBodies() {
    if (!bodiesDirty_ && !scenesDirty_) return;
    std::vector<DevBody> bodies;
    std::vector<float2> verts;
    hostScenes_.clear();
    std::vector<uint32_t> bases;
    for (const HostScene &s : scenes_) {
      const int b0 = (int)bodies.size();
      for (const HostBody &hb : s.bodies) {
        DevBody d{};
        d.type = hb.type;
        d.a0 = hb.a[0]; d.a1 = hb.a[1]; d.a2 = hb.a[2]; d.a3 = hb.a[3];
        setCullBox(d);
        if (hb.type == BODY_POLYGON) {
          d.nverts = (int)(hb.verts.size() / 2);
          d.vertOffset = (int)verts.size();
          for (size_t k = 0; k < hb.verts.size(); k += 2) verts.push_back(make_float2(hb.verts[k], hb.verts[k + 1]));
        }
        bodies.push_back(d);
      }
      hostScenes_.push_back(devScene(s, b0, (int)bodies.size()));
      bases.push_back(s.cellBase);
    }
    if (devBodies_.n < std::max<size_t>(bodies.size(), 1)) devBodies_.alloc(std::max<size_t>(bodies.size(), 16));
    if (devVerts_.n < std::max<size_t>(verts.size(), 1)) devVerts_.alloc(std::max<size_t>(verts.size(), 16));
    if (devScenes_.n < hostScenes_.size()) devScenes_.alloc(hostScenes_.size() * 2);
    if (devSceneBase_.n < bases.size()) devSceneBase_.alloc(bases.size() * 2);
    if (!bodies.empty())
      NBS_HIP(hipMemcpyAsync(devBodies_.p, bodies.data(), bodies.size() * sizeof(DevBody), hipMemcpyHostToDevice, stream_));
    if (!verts.empty())
      NBS_HIP(hipMemcpyAsync(devVerts_.p, verts.data(), verts.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(devScenes_.p, hostScenes_.data(), hostScenes_.size() * sizeof(DevScene),
                           hipMemcpyHostToDevice, stream_));
    NBS_HIP(hipMemcpyAsync(devSceneBase_.p, bases.data(), bases.size() * 4, hipMemcpyHostToDevice, stream_));
    // scene of every cell: the multi-scene kernels' scene lookup in one load
    // (up to 65,536 scenes; beyond that they search the scenes' first cells)
    std::vector<uint16_t> cellScene;
    if (scenes_.size() > 1 && scenes_.size() <= 65536) {
      cellScene.resize(totalCells_);
      for (size_t si = 0; si < scenes_.size(); ++si) {
        const HostScene &hs = scenes_[si];
        std::fill(cellScene.begin() + hs.cellBase, cellScene.begin() + hs.cellBase + (size_t)hs.nx * hs.ny,
                  (uint16_t)si);
      }
      if (devCellScene_.n < cellScene.size()) devCellScene_.alloc(cellScene.size());
      NBS_HIP(hipMemcpyAsync(devCellScene_.p, cellScene.data(), cellScene.size() * sizeof(uint16_t),
                             hipMemcpyHostToDevice, stream_));
    }
    cellSceneCells_ = cellScene.empty() ? 0u : totalCells_;
    NBS_HIP(hipStreamSynchronize(stream_));
    bodiesDirty_ = scenesDirty_ = false;
    rebindBuffers();
    destroyGraph();
  }

  // Pair-kernel variant per phase (viscosity, density, relaxation): an LDS
  // tile capacity, or kWideCap for small steps.
  // Wide kernels for every phase: below the autotuner's range (a few hundred
  // particles, where one lane per particle fills a single workgroup) or up to
  // nbs_tuning.wide_max_particles.
  bool wideStep() const {
    return (tuning_.autotune && n_ < kTuneMinParticles) ||
           (tuning_.wide_max_particles > 0 && n_ <= (uint64_t)tuning_.wide_max_particles);
  }
  void pairCaps(int cap[3]) const {
    for (int k = 0; k < 3; ++k) cap[k] = wideStep() ? kWideCap : tuning_.tile_cap[k];
  }

  // One step on stream_ (ev >= 0: record timing events for step slot ev);
  // relax selects how the relaxation phase runs (Relax).
  void enqueueStep(float dt, int ev, int relax = RELAX_PLAIN, int lossIndex = 0) {
    // every step that does not move the carried density leaves it in an old
    // slot order (a check step sets it afresh when it finishes)
    if (relax != RELAX_SURROGATE_CARRIED) dropCarriedDensity();
    const DevParams p = devParams(dt);
    const DevScene s0 = hostScenes_[0];
    const int n = (int)n_;
    int cap[3];
    pairCaps(cap);
    // Placement binning: the viscosity kernel lists the far movers, k_place
    // orders and moves; on large grids every step's scan clears the counter
    // of the next step's parity.
    const bool place = place_binning(totalCells_, cap[0], lists_ && nBinned_ == n_);
    Buffers bg = buf_;
    bg.farList = place ? d_.farList.p : nullptr;
    bg.chunkCells = place ? d_.chunkCells.p : nullptr;
    bg.farCount = d_.farCounts.p + parity_;
    bg.farCountClear = d_.farCounts.p + (parity_ ^ 1u);
    const bool carried = relax == RELAX_SURROGATE_CARRIED;
    if (carried) {
      bg.densFrom = sur_->carrySrc;
      bg.densTo = sur_->carry.p;
    }
    if (ev >= 0) NBS_HIP(hipEventRecord(events_[5 * ev + 0], stream_));
    launch_visc_predict(bg, p, s0, n, (int)nBinned_, lists_ ? 1 : 0, cap[0], totalCells_, stream_);
    if (ev >= 0) NBS_HIP(hipEventRecord(events_[5 * ev + 1], stream_));
    cellSlots_ = launch_grid(bg, s0, totalCells_, n, cap[0], place, stream_);
    if (ev >= 0) NBS_HIP(hipEventRecord(events_[5 * ev + 2], stream_));
    if (!carried) {
      launch_density(buf_, p, s0, n, cap[1], stream_);
      statSlots_ = density_stat_slots(n, cap[1]);
    }
    if (ev >= 0) NBS_HIP(hipEventRecord(events_[5 * ev + 3], stream_));
    if (relax == RELAX_PLAIN) {
      launch_delta(buf_, p, s0, n, cap[2], stream_);
    } else {
      enqueueLearnedRelaxation(p, s0, n, relax, lossIndex);
    }
    if (ev >= 0) NBS_HIP(hipEventRecord(events_[5 * ev + 4], stream_));
  }

  // Autotuner: every tile capacity gives bit-identical results, so the
  // candidates are timed on real steps with HIP events.  The first tuning tries
  // every capacity; later ones (every retune_interval steps, the fluid's
  // packing changes as it moves) try the current choice and its neighbours.
  // Each of kTuneReps rounds runs one step per candidate; each kernel keeps
  // its fastest candidate (minimum over rounds).
  static constexpr uint32_t kTuneReps = 2;
  static constexpr uint64_t kTuneMinParticles = 1024;
  // The wide kernels (one wave per particle) are a candidate only for steps
  // small enough that one lane per particle leaves the chip idle.
  static constexpr uint64_t kWideTuneMax = 131072;
  int numCaps() const { return n_ <= kWideTuneMax ? kNumTileCaps : kNumTileCaps - 1; }
  uint32_t tuneVariants() const { return tuned_ ? 3u : (uint32_t)numCaps(); }
  bool wantTune() const {
    if (!tuning_.autotune || n_ < kTuneMinParticles || !lists_ || nBinned_ != n_ || surrogateMode() || wideStep())
      return false;
    return !tuned_ || (tuning_.retune_interval > 0 && stepsSinceTune_ >= tuning_.retune_interval);
  }
  int capIndex(int cap) const {
    for (int v = 0; v < numCaps(); ++v)
      if (kTileCaps[v] == cap) return v;
    return numCaps() - 1;
  }
  // A tuning run is kTuneReps rounds of one step per candidate, one step per
  // iteration of the update loop (so emitters still run before every step), and
  // it may span several update calls: each call reads the times of its own
  // tuning steps, whose event slots it reuses afterwards.
  void beginTune() {
    tuneV_ = tuneVariants();
    for (int k = 0; k < 3; ++k) {
      if (!tuned_) {
        for (uint32_t v = 0; v < tuneV_; ++v) tuneCand_[k][v] = kTileCaps[v];
      } else {
        const int c = capIndex(tuning_.tile_cap[k]);
        tuneCand_[k][0] = kTileCaps[std::max(c - 1, 0)];
        tuneCand_[k][1] = kTileCaps[c];
        tuneCand_[k][2] = kTileCaps[std::min(c + 1, numCaps() - 1)];
      }
      for (int v = 0; v < kNumTileCaps; ++v) tuneBest_[k][v] = 3.0e38f;
    }
    tuneDone_ = 0;
    tunePending_.clear();
    tuneActive_ = true;
  }
  void tuneStep(float dt, uint32_t slot) {
    if (!tuneActive_) beginTune();
    const uint32_t v = tuneDone_ % tuneV_;
    for (int k = 0; k < 3; ++k) tuning_.tile_cap[k] = tuneCand_[k][v];
    ensureEvents(slot + 1);
    enqueueStep(dt, (int)slot);
    swapAfterStep();
    tunePending_.push_back({slot, v});
    if (++tuneDone_ < kTuneReps * tuneV_) return;
    NBS_HIP(hipEventSynchronize(events_[5 * slot + 4]));
    absorbTuneSlots();
    for (int k = 0; k < 3; ++k) {  // each kernel keeps its fastest candidate
      float best = 3.0e38f;
      int choice = tuneCand_[k][0];
      for (uint32_t c = 0; c < tuneV_; ++c)
        if (tuneBest_[k][c] < best) best = tuneBest_[k][c], choice = tuneCand_[k][c];
      tuning_.tile_cap[k] = choice;
    }
    tuneActive_ = false;
    tuned_ = true;
    stepsSinceTune_ = 0;
    ++tuneCount_;
    destroyGraph();
  }
  void absorbTuneSlots() {
    const int evPair[3][2] = {{0, 1}, {2, 3}, {3, 4}};  // viscosity, density, relaxation
    for (const auto &sv : tunePending_) {
      for (int k = 0; k < 3; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, events_[5 * sv.first + evPair[k][0]], events_[5 * sv.first + evPair[k][1]]) ==
            hipSuccess)
          tuneBest_[k][sv.second] = std::min(tuneBest_[k][sv.second], ms);
      }
    }
    tunePending_.clear();
  }

  void swapAfterStep() {
    parity_ ^= 1u;
    std::swap(buf_.id, buf_.id2);
    d_.id.swap(d_.id2);
    std::swap(buf_.startCur, buf_.startNext);
    d_.startCur.swap(d_.startNext);
    nBinned_ = n_;
    lists_ = true;
  }

  void ensureEvents(uint32_t slots) {
    const size_t want = 5 * (size_t)std::max<uint32_t>(slots, 1);
    while (events_.size() < want) {
      hipEvent_t e;
      NBS_HIP(hipEventCreate(&e));
      events_.push_back(e);
    }
  }
  void collectTimes(uint32_t slots) {
    float last[4] = {0, 0, 0, 0};
    for (uint32_t s = 0; s < slots; ++s) {
      for (int k = 0; k < 4; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, events_[5 * s + k], events_[5 * s + k + 1]) != hipSuccess) continue;
        kernelMs_[k] += ms;
        kernelLaunches_[k] += 1;
        last[k] = ms;
      }
    }
    stats_h_.time_emitters = 0.0f;
    stats_h_.time_integration = 0.0f;
    stats_h_.time_viscosity_forces = last[0];
    stats_h_.time_predict = 0.0f;
    stats_h_.time_update_grid = last[1];
    stats_h_.time_neighbor_search = 0.0f;
    stats_h_.time_density_and_pressure = last[2];
    stats_h_.time_delta_positions = last[3];
    stats_h_.time_collisions = 0.0f;
  }

  // Captures two steps (both buffer parities) as one graph.
  void ensureGraph(float dt) {
    const DevParams p = devParams(dt);
    if (graphExec_ && std::memcmp(&p, &graphParams_, sizeof(p)) == 0 && graphN_ == n_) return;  // caps: destroyGraph on change
    destroyGraph();
    const Buffers saved = buf_;
    NBS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    enqueueStep(dt, -1);
    swapAfterStep();
    enqueueStep(dt, -1);
    swapAfterStep();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess) throw Error(NBS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&graphExec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      graphExec_ = nullptr;
      throw Error(NBS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    // Two swaps restore the buffer roles.
    if (buf_.id != saved.id || buf_.startCur != saved.startCur) throw Error(NBS_ERR_HIP, "graph capture parity");
    graphParams_ = p;
    graphN_ = n_;
  }
  void destroyGraph() {
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    graphExec_ = nullptr;
  }

  // demo4.cpp:257-284 UpdateEmitter
  void runEmitters(float dt) {
    const float spacing = params_.particle_spacing;
    const float invDt = 1.0f / dt;
    for (Emitter &e : emitters_) {
      if (!e.active) continue;
      const float rate = 1.0f / e.rate;
      e.elapsed += dt;
      e.totalElapsed += dt;
      if (e.elapsed >= rate) {
        e.elapsed = 0.0f;
        const float ax = (e.dx * e.speed) * invDt, ay = (e.dy * e.speed) * invDt;
        const float dirx = -1.0f * e.dy, diry = 1.0f * e.dx;  // vecmath.h:304-306 Vec2Cross(1, d)
        const int count = (int)std::floor(e.radius / spacing);
        if (count > 0) {
          const float offx = ((dirx * (float)count) * spacing) * 0.5f;
          const float offy = ((diry * (float)count) * spacing) * 0.5f;
          std::vector<float> pos(2 * (size_t)count), force(2 * (size_t)count);
          const float h = params_.kernel_height;
          for (int q = 0; q < count; ++q) {
            float px = (dirx * (float)q) * spacing, py = (diry * (float)q) * spacing;
            px = ((dirx * spacing) * 0.5f) + px;
            py = ((diry * spacing) * 0.5f) + py;
            px = (e.px - offx) + px;
            py = (e.py - offy) + py;
            const float jx = (splitmix_uniform(e.seed, e.draws) * h) * 0.01f;
            const float jy = (splitmix_uniform(e.seed, e.draws + 1) * h) * 0.01f;
            e.draws += 2;
            pos[2 * q] = jx + px;
            pos[2 * q + 1] = jy + py;
            force[2 * q] = ax;
            force[2 * q + 1] = ay;
          }
          addParticles(e.scene, pos.data(), force.data(), (uint64_t)count);
        }
      }
      if (e.totalElapsed >= e.duration) e.active = false;
    }
  }

  // ---- learned relaxation (internals) ----------------------------------------
  enum Relax { RELAX_PLAIN = 0, RELAX_SURROGATE = 1, RELAX_CHECK = 2, RELAX_TRAIN = 3, RELAX_FRAME = 4,
               RELAX_CALIBRATE = 5, RELAX_SURROGATE_CARRIED = 6 };

  struct SurrogateState {
    DevBuf<float> params, adamM, adamV, grad, out, act, cnt, dH, dG, dM, dO, part, lossPart, loss;
    DevBuf<uint16_t> w1b, w2b, w3b, w2t, w3t;
    DevBuf<float2> target, shadowX, shadowV, densA, densB, presScratch;
    DevBuf<uint32_t> candScratch, pairPart;
    DevBuf<double> errPartials, errOut, sceneErr;
    size_t inferCap = 0, trainCap = 0;
    float onlineLr = 0.0f;  // > 0: check steps also train on their plain frame
    int pairWaves = 0;  // pair counts written by the last network evaluation
    bool normSet = false;
    int mode = NBS_RELAX_PLAIN;
    uint32_t interval = 8;
    float maxErr = 0.01f;
    float lr = 1.0e-3f;
    uint64_t adamT = 0;
    uint64_t sinceCheck = 0;
    uint32_t plainLeft = 0;
    nbs_surrogate_report rep{};
    double errSum = 0.0;
    // Carried density (NBS_RELAX_SURROGATE_CARRIED_DENSITY): carrySrc is the
    // density the last check accepted (then the state's density), in the
    // previous step's slot order; the binning of each network step moves it
    // into `carry`, which then becomes the state's density array.
    DevBuf<float2> carry;
    const float2 *carrySrc = nullptr;
    uint64_t carryN = 0;
  };

  bool surrogateMode() const { return sur_ && sur_->mode != NBS_RELAX_PLAIN; }
  bool carryReady() const {
    return sur_->mode == NBS_RELAX_SURROGATE_CARRIED_DENSITY && sur_->carrySrc && sur_->carryN == n_ && lists_ &&
           nBinned_ == n_ && emitters_.empty();
  }
  void dropCarriedDensity() {
    if (sur_) sur_->carrySrc = nullptr;
  }
  void requireSur() const {
    if (!sur_) throw Error(NBS_ERR_INVALID_ARGUMENT, "the surrogate is not initialised (nbs_surrogate_init)");
  }
  void uploadParams(const float *h) {
    SurrogateState &S = *sur_;
    NBS_HIP(hipMemcpyAsync(S.params.p, h, kNumParams * 4, hipMemcpyHostToDevice, stream_));
    launch_bf16_weights(S.params.p, S.w1b.p, S.w2b.p, S.w3b.p, S.w2t.p, S.w3t.p, stream_);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipStreamSynchronize(stream_));
  }
  // Buffers of surrogate inference and the density check, sized to cap_.
  void ensureSurInference() {
    SurrogateState &S = *sur_;
    if (S.inferCap >= cap_) return;
    const size_t c = cap_;
    S.shadowX.alloc(c);
    S.shadowV.alloc(c);
    S.densA.alloc(c);
    S.densB.alloc(c);
    S.presScratch.alloc(c);
    S.candScratch.alloc(2 * (size_t)density_stat_slots((int)c, kWideCap));
    S.errPartials.alloc(2 * std::max<size_t>((c + 2047) / 2048, 1));
    S.errOut.alloc(2);
    S.pairPart.alloc(surrogate_forward_waves((int)c));
    S.carry.alloc(c);
    S.carrySrc = nullptr;
    S.inferCap = c;
  }
  void ensureSceneErr() {
    SurrogateState &S = *sur_;
    if (S.sceneErr.n < 2 * scenes_.size()) S.sceneErr.alloc(2 * scenes_.size());
  }
  void ensureSurTraining() {
    SurrogateState &S = *sur_;
    ensureSurInference();
    if (S.trainCap >= cap_) return;
    const size_t c = cap_;
    S.out.alloc(2 * c);
    S.act.alloc((size_t)kActStride * c);
    S.cnt.alloc(c);
    S.dH.alloc(kHid * c);
    S.dG.alloc(kHid * c);
    S.dM.alloc(kHid * c);
    S.dO.alloc(2 * c);
    S.target.alloc(c);
    S.part.alloc(surrogate_part_floats((int)c));
    S.lossPart.alloc(2 * (size_t)surrogate_head_blocks((int)c));
    if (S.loss.n < 2) S.loss.alloc(2);
    S.trainCap = c;
  }
  SurrogateBufs surBufs(bool training, int lossIndex) const {
    const SurrogateState &S = *sur_;
    SurrogateBufs m{};
    m.params = S.params.p;
    m.w1b = S.w1b.p;
    m.w2b = S.w2b.p;
    m.w3b = S.w3b.p;
    m.w2t = S.w2t.p;
    m.w3t = S.w3t.p;
    m.pairPart = S.pairPart.p;
    if (training) {
      m.out = S.out.p;
      m.act = S.act.p;
      m.cnt = S.cnt.p;
      m.dH = S.dH.p;
      m.dG = S.dG.p;
      m.dM = S.dM.p;
      m.dO = S.dO.p;
      m.target = S.target.p;
      m.part = S.part.p;
      m.grad = S.grad.p;
      m.lossPart = S.lossPart.p;
      m.loss = S.loss.p;
      m.lossIndex = lossIndex;
    }
    return m;
  }
  // The relaxation phase of a step in the learned modes.
  void enqueueLearnedRelaxation(const DevParams &p, const DevScene &s0, int n, int relax, int lossIndex) {
    SurrogateState &S = *sur_;
    if (relax != RELAX_CALIBRATE) S.pairWaves = surrogate_forward_waves(n);
    int cap[3];
    pairCaps(cap);
    switch (relax) {
      case RELAX_SURROGATE:
        launch_surrogate_forward(buf_, p, s0, n, surBufs(false, 0), kFwdState, stream_);
        break;
      case RELAX_SURROGATE_CARRIED: {
        // the binning moved the carried density into S.carry; the network
        // reads it there and writes the matching pressures, and it becomes
        // the state's density array (the old one is the next step's target)
        Buffers bc = buf_;
        bc.dens = S.carry.p;
        launch_surrogate_forward(bc, p, s0, n, surBufs(false, 0), kFwdState | kFwdPressure, stream_);
        d_.dens.swap(S.carry);
        buf_.dens = d_.dens.p;
        S.carrySrc = buf_.dens;
        break;
      }
      case RELAX_CHECK: {
        // plain result into the shadow state, the network's into the state;
        // with online training the plain displacement is also the target
        const bool online = S.onlineLr > 0.0f;
        Buffers bp = buf_;
        bp.x0 = S.shadowX.p;
        bp.v = S.shadowV.p;
        if (online) bp.dxOut = S.target.p;
        launch_delta(bp, p, s0, n, cap[2], stream_);
        const SurrogateBufs m = surBufs(online, 0);
        launch_surrogate_forward(buf_, p, s0, n, m, kFwdState | (online ? kFwdTrain : 0), stream_);
        // density of both results on this step's grid
        Buffers bd = buf_;
        bd.pres = S.presScratch.p;
        bd.candStats = S.candScratch.p;
        bd.x1 = buf_.x0;
        bd.dens = S.densA.p;
        launch_density(bd, p, s0, n, cap[1], stream_);
        bd.x1 = S.shadowX.p;
        bd.dens = S.densB.p;
        launch_density(bd, p, s0, n, cap[1], stream_);
        launch_density_error(S.densA.p, S.densB.p, n, S.errPartials.p, S.errOut.p, stream_);
        launch_density_error_scenes(S.densA.p, S.densB.p, buf_.keys, buf_.sceneBase, buf_.nScenes, n,
                                    S.sceneErr.p, stream_);
        if (online) {
          launch_surrogate_backward(buf_, p, s0, n, m, stream_);
          launch_adam(S.params.p, S.grad.p, S.adamM.p, S.adamV.p, kNumWeights, S.onlineLr, 0.9f, 0.999f, 1.0e-8f,
                      (int)++S.adamT, stream_);
          launch_bf16_weights(S.params.p, S.w1b.p, S.w2b.p, S.w3b.p, S.w2t.p, S.w3t.p, stream_);
          ++S.rep.train_steps;
        }
        break;
      }
      case RELAX_TRAIN:
      case RELAX_FRAME:
      case RELAX_CALIBRATE: {
        Buffers bt = buf_;
        bt.dxOut = S.target.p;
        launch_delta(bt, p, s0, n, cap[2], stream_);
        if (relax == RELAX_CALIBRATE) break;
        const SurrogateBufs m = surBufs(true, lossIndex);
        launch_surrogate_forward(buf_, p, s0, n, m, kFwdTrain, stream_);
        launch_surrogate_backward(buf_, p, s0, n, m, stream_);
        if (relax == RELAX_TRAIN) {
          launch_adam(S.params.p, S.grad.p, S.adamM.p, S.adamV.p, kNumWeights, S.lr, 0.9f, 0.999f, 1.0e-8f,
                      (int)++S.adamT, stream_);
          launch_bf16_weights(S.params.p, S.w1b.p, S.w2b.p, S.w3b.p, S.w2t.p, S.w3t.p, stream_);
          ++S.rep.train_steps;
        }
        break;
      }
      default:
        launch_delta(buf_, p, s0, n, cap[2], stream_);
    }
  }
  // Relaxation of the next step in surrogate mode.
  int nextRelax() {
    SurrogateState &S = *sur_;
    if (S.plainLeft > 0) {
      --S.plainLeft;
      ++S.rep.plain_steps;
      return RELAX_PLAIN;
    }
    if (++S.sinceCheck >= S.interval) {
      S.sinceCheck = 0;
      return RELAX_CHECK;
    }
    ++S.rep.surrogate_steps;
    return RELAX_SURROGATE;
  }
  // The check passes only if every scene's density error is within the limit:
  // the global error is the plain-density-weighted mean of the scene errors,
  // so a small scene far off would pass a global test.
  void finishCheck() {
    SurrogateState &S = *sur_;
    const size_t ns = scenes_.size();
    std::vector<double> e(2 + 2 * ns);
    NBS_HIP(hipMemcpyAsync(e.data(), S.errOut.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipMemcpyAsync(e.data() + 2, S.sceneErr.p, 2 * ns * sizeof(double), hipMemcpyDeviceToHost, stream_));
    NBS_HIP(hipStreamSynchronize(stream_));
    auto ratio = [](double num, double den) { return den != 0.0 ? num / std::fabs(den) : (num != 0.0 ? INFINITY : 0.0); };
    const double err = ratio(e[0], e[1]);
    double worst = 0.0;
    uint64_t over = 0;
    for (size_t s = 0; s < ns; ++s) {
      const double es = ratio(e[2 + 2 * s], e[3 + 2 * s]);
      if (!(es <= worst)) worst = es;  // NaN propagates
      if (!(es <= (double)S.maxErr)) ++over;
    }
    ++S.rep.check_steps;
    S.rep.last_density_error = err;
    S.rep.max_density_error = std::max(S.rep.max_density_error, err);
    S.rep.last_max_scene_error = worst;
    S.rep.last_scenes_over_limit = over;
    S.errSum += err;
    const bool fallback = !(err <= (double)S.maxErr) || over > 0;
    if (fallback) {
      NBS_HIP(hipMemcpyAsync(buf_.x0, S.shadowX.p, n_ * sizeof(float2), hipMemcpyDeviceToDevice, stream_));
      NBS_HIP(hipMemcpyAsync(buf_.v, S.shadowV.p, n_ * sizeof(float2), hipMemcpyDeviceToDevice, stream_));
      ++S.rep.fallbacks;
      ++S.rep.plain_steps;
      S.plainLeft = S.interval;
    } else {
      ++S.rep.surrogate_steps;
    }
    // the density of the accepted result, on this step's slot order
    S.carrySrc = fallback ? S.densB.p : S.densA.p;
    S.carryN = n_;
  }
  // One plain step that also captures the frame (relax: RELAX_CALIBRATE or
  // RELAX_FRAME); copies the requested slot-order arrays to the host.
  void runCapturedStep(float dt, int relax, int lossIndex, std::vector<float> *x, std::vector<float> *prev,
                       std::vector<float> *dens, std::vector<float> *target, std::vector<uint32_t> *keys) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) throw Error(NBS_ERR_INVALID_ARGUMENT, "dt must be positive and finite");
    syncBodies();
    if (!emitters_.empty()) runEmitters(dt);
    if (n_ == 0) throw Error(NBS_ERR_INVALID_ARGUMENT, "no particles");
    ensureSurTraining();
    enqueueStep(dt, -1, relax, lossIndex);
    swapAfterStep();
    ++stepsSinceTune_;
    NBS_HIP(hipGetLastError());
    const size_t n = n_;
    auto get2 = [&](std::vector<float> *v, const float2 *src) {
      if (!v) return;
      v->resize(2 * n);
      NBS_HIP(hipMemcpyAsync(v->data(), src, n * sizeof(float2), hipMemcpyDeviceToHost, stream_));
    };
    get2(x, buf_.x1);
    get2(prev, buf_.prev);
    get2(dens, buf_.dens);
    get2(target, sur_->target.p);
    if (keys) {
      keys->resize(n);
      NBS_HIP(hipMemcpyAsync(keys->data(), buf_.keys, n * 4, hipMemcpyDeviceToHost, stream_));
    }
    NBS_HIP(hipStreamSynchronize(stream_));
    statsPending_ = true;
    destroyGraph();
  }

  struct Storage {
    DevBuf<float2> x0, xp, x1, prev, v, dens, pres, acc;
    DevBuf<uint32_t> id, id2, keys, keyp, fill, tmpSlot, scn, count, startCur, startNext, blockSum;
    DevBuf<uint32_t> farList, farCounts;  // placement binning: far movers, one counter per step parity
    DevBuf<uint32_t> chunkCells;          // placement binning: first and last cell of each chunk
    DevBuf<uint4> binRec;                 // one-workgroup binning: a record per new slot for the rank pass
    DevBuf<uint32_t> candStats, cellStats;
  };

  nbs_params params_;
  float gravity_[2] = {0.0f, 0.0f};
  float ext_[2] = {0.0f, 0.0f};
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::vector<HostScene> scenes_;
  std::vector<DevScene> hostScenes_;
  uint32_t totalCells_ = 0;
  bool bodiesDirty_ = true, scenesDirty_ = true;
  std::vector<Emitter> emitters_;

  Storage d_;
  Buffers buf_{};
  DevBuf<DevScene> devScenes_;
  DevBuf<uint32_t> devSceneBase_;
  DevBuf<uint16_t> devCellScene_;  // scene of each cell (multi-scene)
  uint32_t cellSceneCells_ = 0;     // cells the table covers; 0 when stale or absent
  DevBuf<DevBody> devBodies_;
  DevBuf<float2> devVerts_;
  DevBuf<float2> stX_, stPrev_, stV_, stDens_, stPres_;
  DevBuf<uint32_t> stKey_;
  DevBuf<int32_t> stScene_;
  DevBuf<float4> stColor_;
  DevBuf<double> diagPartials_, diagOut_;

  uint64_t n_ = 0, nBinned_ = 0, cap_ = 0;
  bool lists_ = false;
  uint32_t parity_ = 0;  // step parity: selects the far-mover counter of placement binning

  nbs_stats stats_h_{};
  bool statsPending_ = false;
  bool timing_ = false;
  uint32_t options_ = 0;
  std::vector<hipEvent_t> events_;
  double kernelMs_[NBS_KERNEL_COUNT] = {};
  uint64_t kernelLaunches_[NBS_KERNEL_COUNT] = {};

  nbs_tuning tuning_{1, 256, {2048, 2048, 2048}, 0};
  bool tuned_ = false;
  bool tuneActive_ = false;
  uint32_t statSlots_ = 0;  // candStats pairs written by the last step's density launch
  uint32_t cellSlots_ = 0;  // cellStats pairs written by the last step's grid phase
  uint32_t tuneV_ = 0, tuneDone_ = 0;
  int tuneCand_[3][kNumTileCaps] = {};
  float tuneBest_[3][kNumTileCaps] = {};
  std::vector<std::pair<uint32_t, uint32_t>> tunePending_;  // (event slot, candidate) of this call
  uint64_t stepsSinceTune_ = 0;
  uint64_t tuneCount_ = 0;

  hipGraphExec_t graphExec_ = nullptr;
  DevParams graphParams_{};
  uint64_t graphN_ = 0;

  std::unique_ptr<SurrogateState> sur_;
};

}  // namespace nbs

struct nbs_sim {
  std::unique_ptr<nbs::Simulation> impl;
};

using nbs::Error;

template <typename F>
static nbs_status guarded(F &&f) {
  try {
    nbs::g_lastError.clear();
    f();
    return NBS_OK;
  } catch (const Error &e) {
    nbs::g_lastError = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    nbs::g_lastError = "host out of memory";
    return NBS_ERR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    nbs::g_lastError = e.what();
    return NBS_ERR_HIP;
  } catch (...) {
    nbs::g_lastError = "unknown error";
    return NBS_ERR_HIP;
  }
}

#define NBS_CHECK_SIM(sim) \
  if (!(sim) || !(sim)->impl) return (nbs::g_lastError = "null handle", NBS_ERR_INVALID_ARGUMENT)

extern "C" {

nbs_status nbs_create(const nbs_params *params, const nbs_domain *domain, uint64_t capacity, nbs_sim **out) {
  if (!params || !domain || !out) return (nbs::g_lastError = "null argument", NBS_ERR_INVALID_ARGUMENT);
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    nbs::g_lastError = "no HIP device";
    return NBS_ERR_NO_DEVICE;
  }
  return guarded([&] {
    std::unique_ptr<nbs_sim> s(new nbs_sim);
    s->impl.reset(new nbs::Simulation(*params, *domain, capacity));
    *out = s.release();
  });
}

nbs_status nbs_destroy(nbs_sim *sim) {
  if (!sim) return NBS_OK;
  return guarded([&] { delete sim; });
}

nbs_status nbs_add_scene(nbs_sim *sim, const nbs_domain *domain, int32_t *scene_out) {
  NBS_CHECK_SIM(sim);
  if (!domain) return (nbs::g_lastError = "null domain", NBS_ERR_INVALID_ARGUMENT);
  return guarded([&] {
    int s = sim->impl->addScene(*domain);
    if (scene_out) *scene_out = s;
  });
}

nbs_status nbs_get_scene_count(nbs_sim *sim, int32_t *count) {
  NBS_CHECK_SIM(sim);
  if (!count) return NBS_ERR_INVALID_ARGUMENT;
  *count = sim->impl->sceneCount();
  return NBS_OK;
}

nbs_status nbs_set_params(nbs_sim *sim, const nbs_params *params) {
  NBS_CHECK_SIM(sim);
  if (!params) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { sim->impl->setParams(*params); });
}

nbs_status nbs_get_params(nbs_sim *sim, nbs_params *params) {
  NBS_CHECK_SIM(sim);
  if (!params) return NBS_ERR_INVALID_ARGUMENT;
  *params = sim->impl->params();
  return NBS_OK;
}

nbs_status nbs_set_gravity(nbs_sim *sim, float gx, float gy) {
  NBS_CHECK_SIM(sim);
  sim->impl->setGravity(gx, gy);
  return NBS_OK;
}

nbs_status nbs_add_external_force(nbs_sim *sim, float fx, float fy) {
  NBS_CHECK_SIM(sim);
  sim->impl->addExternalForce(fx, fy);
  return NBS_OK;
}

nbs_status nbs_clear_external_force(nbs_sim *sim) {
  NBS_CHECK_SIM(sim);
  sim->impl->clearExternalForce();
  return NBS_OK;
}

nbs_status nbs_reset_stats(nbs_sim *sim) {
  NBS_CHECK_SIM(sim);
  sim->impl->resetStats();
  return NBS_OK;
}

nbs_status nbs_clear_bodies(nbs_sim *sim) {
  NBS_CHECK_SIM(sim);
  sim->impl->clearBodies();
  return NBS_OK;
}

nbs_status nbs_clear_particles(nbs_sim *sim) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->clearParticles(); });
}

nbs_status nbs_clear_emitters(nbs_sim *sim) {
  NBS_CHECK_SIM(sim);
  sim->impl->clearEmitters();
  return NBS_OK;
}

nbs_status nbs_add_plane(nbs_sim *sim, int32_t scene, float nx, float ny, float distance) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->addBody(scene, nbs::HostBody{nbs::BODY_PLANE, {nx, ny, distance, 0.0f}, {}}); });
}

nbs_status nbs_add_circle(nbs_sim *sim, int32_t scene, float px, float py, float radius) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->addBody(scene, nbs::HostBody{nbs::BODY_CIRCLE, {px, py, radius, 0.0f}, {}}); });
}

nbs_status nbs_add_line_segment(nbs_sim *sim, int32_t scene, float ax, float ay, float bx, float by) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->addBody(scene, nbs::HostBody{nbs::BODY_SEGMENT, {ax, ay, bx, by}, {}}); });
}

nbs_status nbs_add_polygon(nbs_sim *sim, int32_t scene, uint64_t vertex_count, const float *verts_xy) {
  NBS_CHECK_SIM(sim);
  if (!verts_xy || vertex_count < 3 || vertex_count > 8) {
    nbs::g_lastError = "a polygon has 3..8 vertices";
    return NBS_ERR_INVALID_ARGUMENT;
  }
  return guarded([&] {
    nbs::HostBody b{nbs::BODY_POLYGON, {0.0f, 0.0f, 0.0f, 0.0f}, {}};
    b.verts.assign(verts_xy, verts_xy + 2 * vertex_count);
    sim->impl->addBody(scene, std::move(b));
  });
}

nbs_status nbs_add_particles(nbs_sim *sim, int32_t scene, const float *pos_xy, const float *force_xy, uint64_t n,
                             uint64_t *first_index) {
  NBS_CHECK_SIM(sim);
  return guarded([&] {
    uint64_t f = sim->impl->addParticles(scene, pos_xy, force_xy, n);
    if (first_index) *first_index = f;
  });
}

nbs_status nbs_add_volume(nbs_sim *sim, int32_t scene, float center_x, float center_y, float force_x, float force_y,
                          int32_t count_x, int32_t count_y, float spacing, uint64_t jitter_seed) {
  NBS_CHECK_SIM(sim);
  return guarded([&] {
    sim->impl->addVolume(scene, center_x, center_y, force_x, force_y, count_x, count_y, spacing, jitter_seed);
  });
}

nbs_status nbs_add_emitter(nbs_sim *sim, int32_t scene, float pos_x, float pos_y, float dir_x, float dir_y,
                           float radius, float speed, float rate, float duration, uint64_t jitter_seed) {
  NBS_CHECK_SIM(sim);
  if (!(rate > 0.0f)) return (nbs::g_lastError = "rate must be positive", NBS_ERR_INVALID_ARGUMENT);
  return guarded([&] {
    nbs::Emitter e{scene, pos_x, pos_y, dir_x, dir_y, radius, speed, rate, duration, 0.0f, 0.0f, true, jitter_seed, 0};
    sim->impl->addEmitter(e);
  });
}

nbs_status nbs_update(nbs_sim *sim, float dt, uint32_t nsteps) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->update(dt, nsteps); });
}

nbs_status nbs_get_particle_count(nbs_sim *sim, uint64_t *count) {
  NBS_CHECK_SIM(sim);
  if (!count) return NBS_ERR_INVALID_ARGUMENT;
  *count = sim->impl->count();
  return NBS_OK;
}

nbs_status nbs_get_state(nbs_sim *sim, float *x, float *prev, float *v, float *density, float *pressure,
                         uint32_t *cell_key, int32_t *scene) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->getState(x, prev, v, density, pressure, cell_key, scene); });
}

nbs_status nbs_get_colors(nbs_sim *sim, float *rgba) {
  NBS_CHECK_SIM(sim);
  if (!rgba) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { sim->impl->getColors(rgba); });
}

nbs_status nbs_set_state(nbs_sim *sim, const float *x, const float *v) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->setState(x, v); });
}

nbs_status nbs_get_stats(nbs_sim *sim, nbs_stats *stats) {
  NBS_CHECK_SIM(sim);
  if (!stats) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { *stats = sim->impl->stats(); });
}

nbs_status nbs_get_kernel_times(nbs_sim *sim, double *avg_ms, uint64_t *launches) {
  NBS_CHECK_SIM(sim);
  sim->impl->kernelTimes(avg_ms, launches);
  return NBS_OK;
}

nbs_status nbs_set_timing(nbs_sim *sim, int32_t enabled) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->setTiming(enabled != 0); });
}

nbs_status nbs_set_options(nbs_sim *sim, uint32_t flags) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->setOptions(flags); });
}

nbs_status nbs_set_tuning(nbs_sim *sim, const nbs_tuning *tuning) {
  NBS_CHECK_SIM(sim);
  if (!tuning) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { sim->impl->setTuning(*tuning); });
}

nbs_status nbs_get_tuning(nbs_sim *sim, nbs_tuning *tuning) {
  NBS_CHECK_SIM(sim);
  if (!tuning) return NBS_ERR_INVALID_ARGUMENT;
  *tuning = sim->impl->tuning();
  return NBS_OK;
}

nbs_status nbs_get_diagnostics(nbs_sim *sim, nbs_diagnostics *out) {
  NBS_CHECK_SIM(sim);
  if (!out) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { *out = sim->impl->diagnostics(); });
}

// ---------------------------------------------------------------------------
// Diagnostics without allocation (nbs_diagnostics_noalloc).  The partial sums
// live in `scratch`, one fixed array per GPU for the life of the process: one
// row of kDiagWidth doubles per workgroup of kDiagPerBlock particles, for at
// most NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES particles (384 KiB).  Every row
// that a call reads is written by the same call, so nothing needs clearing;
// calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace nbs {
constexpr size_t kNoAllocRows = NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES / kDiagPerBlock;
static_assert(NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES % kDiagPerBlock == 0, "whole workgroups");
static __device__ __align__(256) double scratch[kNoAllocRows * kDiagWidth];

// Device address of `scratch` on each GPU, looked up on the first call with
// that GPU current.  The host-side name is only a stand-in for the symbol.
constexpr int kMaxDevices = 64;
static std::mutex g_scratchMutex;
static double *g_scratchAddr[kMaxDevices] = {};

static double *scratch_on_current_device() {
  int dev = 0;
  NBS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(NBS_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchAddr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error(NBS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchAddr[dev] = static_cast<double *>(ptr);
  }
  return g_scratchAddr[dev];
}
}  // namespace nbs

nbs_status nbs_diagnostics_noalloc(const float *x, const float *v, const float *density, uint64_t n,
                                   float rest_density, nbs_diagnostics *out, void *stream) {
  if (!out || (n > 0 && (!x || !v || !density)))
    return (nbs::g_lastError = "null argument", NBS_ERR_INVALID_ARGUMENT);
  if (n > NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES)
    return (nbs::g_lastError = "n exceeds NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES (16,777,216)", NBS_ERR_CAPACITY);
  if (!std::isfinite(rest_density)) return (nbs::g_lastError = "rest_density must be finite", NBS_ERR_INVALID_ARGUMENT);
  return guarded([&] {
    double *partials = nbs::scratch_on_current_device();
    nbs::launch_diagnostics_record(reinterpret_cast<const float2 *>(x), reinterpret_cast<const float2 *>(v),
                                   reinterpret_cast<const float2 *>(density), (int)n, rest_density, partials, out,
                                   static_cast<hipStream_t>(stream));
    NBS_HIP(hipGetLastError());
  });
}

nbs_status nbs_get_total_cells(nbs_sim *sim, uint64_t *total_cells) {
  NBS_CHECK_SIM(sim);
  if (!total_cells) return NBS_ERR_INVALID_ARGUMENT;
  *total_cells = sim->impl->totalCells();
  return NBS_OK;
}

nbs_status nbs_get_grid(nbs_sim *sim, uint32_t *sorted_ids, uint32_t *cell_start) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->getGrid(sorted_ids, cell_start); });
}

nbs_status nbs_is_multithreading_supported(nbs_sim *sim, int32_t *supported) {
  NBS_CHECK_SIM(sim);
  if (!supported) return NBS_ERR_INVALID_ARGUMENT;
  *supported = 0;
  return NBS_OK;
}

nbs_status nbs_selftest_math(const float *a, const float *b, uint64_t n, float *out) {
  if (!a || !b || !out) return NBS_ERR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NBS_ERR_NO_DEVICE;
  return guarded([&] {
    nbs::DevBuf<float> da, db, dout;
    da.alloc(n);
    db.alloc(n);
    dout.alloc(6 * n);
    NBS_HIP(hipMemcpy(da.p, a, n * 4, hipMemcpyHostToDevice));
    NBS_HIP(hipMemcpy(db.p, b, n * 4, hipMemcpyHostToDevice));
    nbs::launch_selftest_math(da.p, db.p, n, dout.p, nullptr);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipMemcpy(out, dout.p, 24 * n, hipMemcpyDeviceToHost));
  });
}

nbs_status nbs_bench_copy(uint64_t bytes, int32_t reps, double *gbps) {
  if (!gbps || reps <= 0 || bytes < 16) return NBS_ERR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NBS_ERR_NO_DEVICE;
  return guarded([&] {
    nbs::DevBuf<uint8_t> a, b;
    a.alloc(bytes);
    b.alloc(bytes);
    hipEvent_t e0, e1;
    NBS_HIP(hipEventCreate(&e0));
    NBS_HIP(hipEventCreate(&e1));
    nbs::launch_copy(a.p, b.p, bytes, nullptr);  // first touch
    NBS_HIP(hipEventRecord(e0, nullptr));
    for (int32_t r = 0; r < reps; ++r) nbs::launch_copy(a.p, b.p, bytes, nullptr);
    NBS_HIP(hipEventRecord(e1, nullptr));
    NBS_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    NBS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbps = 2.0 * (double)(bytes / 16 * 16) * reps / ((double)ms * 1e-3) / 1e9;
  });
}

// ---- learned relaxation (include/nbs_surrogate.h) ---------------------------
nbs_status nbs_surrogate_init(nbs_sim *sim, uint64_t seed) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surInit(seed); });
}

nbs_status nbs_surrogate_set_params(nbs_sim *sim, const float *params, uint64_t count) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surSetParams(params, count); });
}

nbs_status nbs_surrogate_get_params(nbs_sim *sim, float *params, uint64_t count) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surGetParams(params, count); });
}

nbs_status nbs_surrogate_calibrate(nbs_sim *sim, float dt) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surCalibrate(dt); });
}

nbs_status nbs_surrogate_train(nbs_sim *sim, float dt, uint32_t steps, float learning_rate, float *losses,
                               float *relative_losses) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surTrain(dt, steps, learning_rate, losses, relative_losses); });
}

nbs_status nbs_surrogate_set_mode(nbs_sim *sim, int32_t mode, uint32_t check_interval, float max_density_error) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surSetMode(mode, check_interval, max_density_error); });
}

nbs_status nbs_surrogate_set_online_training(nbs_sim *sim, float learning_rate) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surSetOnlineTraining(learning_rate); });
}

nbs_status nbs_surrogate_get_scene_errors(nbs_sim *sim, double *errors, uint64_t count) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surSceneErrors(errors, count); });
}

nbs_status nbs_surrogate_get_report(nbs_sim *sim, nbs_surrogate_report *report) {
  NBS_CHECK_SIM(sim);
  if (!report) return NBS_ERR_INVALID_ARGUMENT;
  return guarded([&] { *report = sim->impl->surReport(); });
}

nbs_status nbs_surrogate_frame(nbs_sim *sim, float dt, float *grad, float *loss, float *out, float *x, float *prev,
                               float *density, float *target, int32_t *scene) {
  NBS_CHECK_SIM(sim);
  return guarded([&] { sim->impl->surFrame(dt, grad, loss, out, x, prev, density, target, scene); });
}

nbs_status nbs_selftest_mfma(const float *a, const float *b, float *c) {
  if (!a || !b || !c) return NBS_ERR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NBS_ERR_NO_DEVICE;
  return guarded([&] {
    nbs::DevBuf<float> da, db, dc;
    da.alloc(32 * 8);
    db.alloc(8 * 32);
    dc.alloc(32 * 32);
    NBS_HIP(hipMemcpy(da.p, a, 32 * 8 * 4, hipMemcpyHostToDevice));
    NBS_HIP(hipMemcpy(db.p, b, 8 * 32 * 4, hipMemcpyHostToDevice));
    nbs::launch_mfma_probe(da.p, db.p, dc.p, nullptr);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipMemcpy(c, dc.p, 32 * 32 * 4, hipMemcpyDeviceToHost));
  });
}

nbs_status nbs_selftest_mfma16(const float *a, const float *b, float *c) {
  if (!a || !b || !c) return NBS_ERR_INVALID_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NBS_ERR_NO_DEVICE;
  return guarded([&] {
    nbs::DevBuf<float> da, db, dc;
    da.alloc(32 * 16);
    db.alloc(16 * 32);
    dc.alloc(32 * 32);
    NBS_HIP(hipMemcpy(da.p, a, 32 * 16 * 4, hipMemcpyHostToDevice));
    NBS_HIP(hipMemcpy(db.p, b, 16 * 32 * 4, hipMemcpyHostToDevice));
    nbs::launch_mfma16_probe(da.p, db.p, dc.p, nullptr);
    NBS_HIP(hipGetLastError());
    NBS_HIP(hipMemcpy(c, dc.p, 32 * 32 * 4, hipMemcpyDeviceToHost));
  });
}

const char *nbs_status_string(int32_t status) {
  switch (status) {
    case NBS_OK: return "NBS_OK";
    case NBS_ERR_INVALID_ARGUMENT: return "NBS_ERR_INVALID_ARGUMENT";
    case NBS_ERR_OUT_OF_MEMORY: return "NBS_ERR_OUT_OF_MEMORY";
    case NBS_ERR_HIP: return "NBS_ERR_HIP";
    case NBS_ERR_CAPACITY: return "NBS_ERR_CAPACITY";
    case NBS_ERR_NON_FINITE: return "NBS_ERR_NON_FINITE";
    case NBS_ERR_NO_DEVICE: return "NBS_ERR_NO_DEVICE";
    case NBS_ERR_UNSUPPORTED: return "NBS_ERR_UNSUPPORTED";
    default: return "NBS_ERR_UNKNOWN";
  }
}

const char *nbs_last_error(void) { return nbs::g_lastError.c_str(); }

}  // extern "C"
