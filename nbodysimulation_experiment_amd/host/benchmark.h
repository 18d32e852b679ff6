// The reference application's built-in benchmark (app.cpp:88-161, 215-275;
// app.h:21-30, 72-103) over any BaseSimulation: a scenario is loaded, stepped
// for a number of frames with one Update of the frame's substep dt each, and
// reloaded for the next iteration; every frame records the wall-clock time of
// its Update and the simulation's per-phase statistics, and the frames of all
// iterations are reduced to per-field minimum, maximum and average.
#ifndef NBS_BENCHMARK_H
#define NBS_BENCHMARK_H

#include <string>
#include <vector>

#include "scenarios.h"

namespace Demo5 {

constexpr size_t kBenchmarkFrameCount = 64;      // app.h:24
constexpr size_t kBenchmarkIterationCount = 16;  // app.h:25

struct FrameStatistics {  // app.h:72-85
  SPHStatistics stats;
  float simulationTime;   // ms of the frame's Update calls
  size_t particleCount;   // after the frame
};

struct BenchmarkSummary {  // app.h:95-103
  std::string scenario;
  size_t scenarioIndex;
  size_t frameCount;
  size_t iterationCount;
  FrameStatistics min, max, avg;
  std::vector<FrameStatistics> frames;  // every frame of every iteration
};

// Runs `iterations` x `frames` frames of `scenarioIndex`.  The per-phase
// minimums start from +infinity (the reference starts them from 0,
// sph.h:143-149, so its per-phase minimum always reads 0).
BenchmarkSummary RunBenchmark(BaseSimulation &sim, size_t scenarioIndex, size_t frames = kBenchmarkFrameCount,
                              size_t iterations = kBenchmarkIterationCount, float deltaTime = Constants().deltaTime,
                              int substeps = 1);

}  // namespace Demo5

#endif  // NBS_BENCHMARK_H
