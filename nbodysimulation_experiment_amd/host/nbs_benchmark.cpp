// Command-line front end of the built-in benchmark and scenario table for the
// GPU simulation (Demo5::ParticleSimulation).
//
//   nbs_benchmark [--scenario I|all] [--frames F] [--iterations K] [--seed S] [--json]
//       runs the benchmark (default: every scenario, 64 frames x 16 iterations)
//       and prints per-phase min / max / avg milliseconds per scenario.
//   nbs_benchmark --list
//       prints the constants and the resolved scenario table as JSON.
//   nbs_benchmark --dump-state FILE --scenario I --frames F [--seed S]
//       loads scenario I, runs F frames and writes the particle count followed
//       by the 48-byte ParticleData records.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "benchmark.h"
#include "demo5.h"

using namespace Demo5;

namespace {

void jv(const Vec2f &v) { std::printf("[%.9g, %.9g]", v.x, v.y); }

const char *kindName(BodyKind k) {
  switch (k) {
    case BodyKind::Plane: return "plane";
    case BodyKind::Circle: return "circle";
    case BodyKind::LineSegment: return "segment";
    case BodyKind::Polygon: return "polygon";
  }
  return "?";
}

void listScenarios() {
  const ReferenceConstants &c = Constants();
  std::printf("{\"constants\": {\"boundary_width\": %.9g, \"boundary_height\": %.9g, \"delta_time\": %.9g, "
              "\"kernel_height\": %.9g, \"particle_spacing\": %.9g, \"collision_radius\": %.9g, "
              "\"collision_margin\": %.9g, \"rest_density\": %.9g, \"stiffness\": %.9g, \"near_stiffness\": %.9g, "
              "\"linear_viscosity\": %.9g, \"quadratic_viscosity\": %.9g, \"grid_cell_size\": %.9g, "
              "\"grid_count_x\": %d, \"grid_count_y\": %d},\n \"scenarios\": [",
              c.boundaryWidth, c.boundaryHeight, c.deltaTime, c.kernelHeight, c.particleSpacing, c.collisionRadius,
              c.collisionMargin, c.restDensity, c.stiffness, c.nearStiffness, c.linearViscosity,
              c.quadraticViscosity, c.gridCellSize, c.gridCountX, c.gridCountY);
  const auto &all = Scenarios();
  for (size_t i = 0; i < all.size(); ++i) {
    const Scenario &s = all[i];
    const SPHParameters &p = s.parameters;
    std::printf("%s\n  {\"name\": \"%s\", \"gravity\": ", i ? "," : "", s.name.c_str());
    jv(s.gravity);
    std::printf(", \"params\": {\"kernel_height\": %.9g, \"cell_size\": %.9g, \"particle_spacing\": %.9g, "
                "\"inv_kernel_height\": %.9g, \"rest_density\": %.9g, \"stiffness\": %.9g, \"near_stiffness\": %.9g, "
                "\"linear_viscosity\": %.9g, \"quadratic_viscosity\": %.9g}",
                p.kernelHeight, p.cellSize, p.particleSpacing, p.invKernelHeight, p.restDensity, p.stiffness,
                p.nearStiffness, p.linearViscosity, p.quadraticViscosity);
    std::printf(", \"volumes\": [");
    for (size_t v = 0; v < s.volumes.size(); ++v) {
      int nx = 0, ny = 0;
      VolumeCounts(s.volumes[v], p.particleSpacing, nx, ny);
      std::printf("%s{\"position\": ", v ? ", " : "");
      jv(s.volumes[v].position);
      std::printf(", \"size\": ");
      jv(s.volumes[v].size);
      std::printf(", \"force\": ");
      jv(s.volumes[v].force);
      std::printf(", \"count\": [%d, %d]}", nx, ny);
    }
    std::printf("], \"emitters\": [");
    for (size_t e = 0; e < s.emitters.size(); ++e) {
      const ScenarioEmitter &m = s.emitters[e];
      std::printf("%s{\"position\": ", e ? ", " : "");
      jv(m.position);
      std::printf(", \"direction\": ");
      jv(m.direction);
      std::printf(", \"radius\": %.9g, \"speed\": %.9g, \"rate\": %.9g, \"duration\": %.9g}", m.radius, m.speed,
                  m.rate, m.duration);
    }
    std::printf("], \"bodies\": [");
    for (size_t b = 0; b < s.bodies.size(); ++b) {
      const ResolvedBody r = ResolveBody(s.bodies[b]);
      std::printf("%s{\"kind\": \"%s\"", b ? ", " : "", kindName(r.kind));
      if (r.kind == BodyKind::Plane) {
        std::printf(", \"normal\": ");
        jv(r.normal);
        std::printf(", \"distance\": %.9g", r.distance);
      } else if (r.kind == BodyKind::Circle) {
        std::printf(", \"center\": ");
        jv(r.center);
        std::printf(", \"radius\": %.9g", r.radius);
      } else {
        std::printf(", \"verts\": [");
        for (size_t k = 0; k < r.verts.size(); ++k) {
          if (k) std::printf(", ");
          jv(r.verts[k]);
        }
        std::printf("]");
      }
      std::printf("}");
    }
    std::printf("]}");
  }
  std::printf("\n]}\n");
}

ParticleSimulation *newSimulation(const Scenario &s, uint64_t seed) {
  const ReferenceConstants &c = Constants();
  auto *sim = new ParticleSimulation(s.parameters, c.boundaryWidth, c.boundaryHeight, c.gridCellSize,
                                     c.collisionRadius, c.collisionMargin);
  sim->SetSeed(seed);
  if (nbs_set_timing(sim->Handle(), 1) != NBS_OK) throw Error(NBS_ERR_HIP, "nbs_set_timing failed");
  return sim;
}

const char *kPhaseNames[9] = {"emitters", "integration", "viscosityForces", "predict", "updateGrid",
                              "neighborSearch", "densityAndPressure", "deltaPositions", "collisions"};

void printSummary(const BenchmarkSummary &b, bool json) {
  const float *mn = &b.min.stats.time.emitters, *mx = &b.max.stats.time.emitters, *av = &b.avg.stats.time.emitters;
  if (json) {
    std::printf("{\"scenario\": \"%s\", \"frames\": %zu, \"iterations\": %zu, \"particles\": [%zu, %zu, %zu], "
                "\"simulation_ms\": [%.6g, %.6g, %.6g], \"phases_ms\": {",
                b.scenario.c_str(), b.frameCount, b.iterationCount, b.min.particleCount, b.max.particleCount,
                b.avg.particleCount, b.min.simulationTime, b.max.simulationTime, b.avg.simulationTime);
    for (int k = 0; k < 9; ++k)
      std::printf("%s\"%s\": [%.6g, %.6g, %.6g]", k ? ", " : "", kPhaseNames[k], mn[k], mx[k], av[k]);
    std::printf("}}\n");
    return;
  }
  std::printf("Scenario %zu \"%s\": %zu iterations x %zu frames, particles %zu..%zu\n", b.scenarioIndex,
              b.scenario.c_str(), b.iterationCount, b.frameCount, b.min.particleCount, b.max.particleCount);
  std::printf("  %-20s %10s %10s %10s\n", "ms", "min", "max", "avg");
  std::printf("  %-20s %10.4f %10.4f %10.4f\n", "simulation", b.min.simulationTime, b.max.simulationTime,
              b.avg.simulationTime);
  for (int k = 0; k < 9; ++k) std::printf("  %-20s %10.4f %10.4f %10.4f\n", kPhaseNames[k], mn[k], mx[k], av[k]);
}

int usage() {
  std::fprintf(stderr,
               "usage: nbs_benchmark [--scenario I|all] [--frames F] [--iterations K] [--seed S] [--json]\n"
               "       nbs_benchmark --list\n"
               "       nbs_benchmark --dump-state FILE --scenario I --frames F [--seed S]\n");
  return 2;
}

}  // namespace

int main(int argc, char **argv) {
  std::string scenario = "all", dump;
  size_t frames = kBenchmarkFrameCount, iterations = kBenchmarkIterationCount;
  uint64_t seed = 0;
  bool json = false, list = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    const char *v = nullptr;
    if (a == "--list") list = true;
    else if (a == "--json") json = true;
    else if (a == "--scenario" && (v = next())) scenario = v;
    else if (a == "--frames" && (v = next())) frames = std::strtoull(v, nullptr, 10);
    else if (a == "--iterations" && (v = next())) iterations = std::strtoull(v, nullptr, 10);
    else if (a == "--seed" && (v = next())) seed = std::strtoull(v, nullptr, 10);
    else if (a == "--dump-state" && (v = next())) dump = v;
    else return usage();
  }
  if (list) {
    listScenarios();
    return 0;
  }
  const size_t count = Scenarios().size();
  size_t first = 0, last = count;
  if (scenario != "all") {
    char *end = nullptr;
    first = std::strtoull(scenario.c_str(), &end, 10);
    if (!end || *end || first >= count) return usage();
    last = first + 1;
  } else if (!dump.empty()) {
    return usage();
  }
  try {
    if (!dump.empty()) {
      const Scenario &s = Scenarios()[first];
      ParticleSimulation *sim = newSimulation(s, seed);
      LoadScenario(*sim, s);
      for (size_t f = 0; f < frames; ++f) sim->Update(Constants().deltaTime);
      const size_t n = sim->GetParticleCount();
      const ParticleData *d = sim->GetParticleDatas();
      FILE *fp = std::fopen(dump.c_str(), "wb");
      if (!fp) {
        delete sim;
        return 3;
      }
      const uint64_t hdr = n;
      std::fwrite(&hdr, sizeof(hdr), 1, fp);
      std::fwrite(d, sizeof(ParticleData), n, fp);
      std::fclose(fp);
      delete sim;
      std::printf("nbs_benchmark: scenario %zu \"%s\", %zu frames, %zu particles\n", first, s.name.c_str(), frames, n);
      return 0;
    }
    for (size_t i = first; i < last; ++i) {
      ParticleSimulation *sim = newSimulation(Scenarios()[i], seed);
      const BenchmarkSummary b = RunBenchmark(*sim, i, frames, iterations);
      delete sim;
      printSummary(b, json);
      std::fflush(stdout);
    }
  } catch (const Error &e) {
    std::fprintf(stderr, "nbs_benchmark: %s\n", e.what());
    return 1;
  }
  return 0;
}
