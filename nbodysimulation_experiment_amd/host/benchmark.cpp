// Built-in benchmark (see benchmark.h).
#include "benchmark.h"

#include <algorithm>
#include <chrono>
#include <limits>

namespace Demo5 {

namespace {

// The nine per-phase times of SPHStatistics (sph.h:131-141) as an array view.
float *phases(SPHStatistics &s) { return &s.time.emitters; }
constexpr int kPhases = 9;
static_assert(sizeof(SPHStatistics::time) == kPhases * sizeof(float), "nine per-phase times");

}  // namespace

BenchmarkSummary RunBenchmark(BaseSimulation &sim, size_t scenarioIndex, size_t frames, size_t iterations,
                              float deltaTime, int substeps) {
  const Scenario &scenario = Scenarios().at(scenarioIndex);
  BenchmarkSummary out;
  out.scenario = scenario.name;
  out.scenarioIndex = scenarioIndex;
  out.frameCount = frames;
  out.iterationCount = iterations;
  out.frames.reserve(frames * iterations);
  for (size_t it = 0; it < iterations; ++it) {
    LoadScenario(sim, scenario);  // app.cpp:409 (first iteration), 271 (next iterations)
    for (size_t f = 0; f < frames; ++f) {
      const auto t0 = std::chrono::high_resolution_clock::now();
      for (int s = 0; s < substeps; ++s) sim.Update(deltaTime);  // app.cpp:230-233
      const auto t1 = std::chrono::high_resolution_clock::now();
      FrameStatistics fs;
      fs.stats = sim.GetStats();
      fs.simulationTime = (float)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count() * 1.0e-6f;
      fs.particleCount = sim.GetParticleCount();
      out.frames.push_back(fs);
    }
  }

  // app.cpp:88-161
  const float inf = std::numeric_limits<float>::infinity();
  out.min = FrameStatistics{};
  out.max = FrameStatistics{};
  out.avg = FrameStatistics{};
  out.min.simulationTime = inf;
  float *mn = phases(out.min.stats), *mx = phases(out.max.stats), *av = phases(out.avg.stats);
  for (int k = 0; k < kPhases; ++k) mn[k] = inf, mx[k] = 0.0f, av[k] = 0.0f;
  out.min.particleCount = std::numeric_limits<size_t>::max();
  for (FrameStatistics &fs : out.frames) {
    const float *p = phases(fs.stats);
    for (int k = 0; k < kPhases; ++k) {
      mn[k] = std::min(mn[k], p[k]);
      mx[k] = std::max(mx[k], p[k]);
      av[k] += p[k];
    }
    out.min.simulationTime = std::min(out.min.simulationTime, fs.simulationTime);
    out.max.simulationTime = std::max(out.max.simulationTime, fs.simulationTime);
    out.avg.simulationTime += fs.simulationTime;
    out.min.particleCount = std::min(out.min.particleCount, fs.particleCount);
    out.max.particleCount = std::max(out.max.particleCount, fs.particleCount);
    out.avg.particleCount += fs.particleCount;
  }
  if (!out.frames.empty()) {
    const float inv = 1.0f / (float)out.frames.size();
    for (int k = 0; k < kPhases; ++k) av[k] *= inv;
    out.avg.simulationTime *= inv;
    out.avg.particleCount /= out.frames.size();
  } else {
    out.min.simulationTime = 0.0f;
    for (int k = 0; k < kPhases; ++k) mn[k] = 0.0f;
    out.min.particleCount = 0;
  }
  return out;
}

}  // namespace Demo5
