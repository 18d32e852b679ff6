// Drives Demo5::ParticleSimulation through the reference's BaseSimulation calls
// the way DemoApplication does: LoadScenario (app.cpp:477-534) then the fixed
// dt Update loop (app.cpp:231-233).  Writes the 48-byte ParticleData mirror
// and the statistics for the GPU test to compare with the oracle.
//
//   demo5_driver <out.bin> <steps>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "demo5.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: demo5_driver <out.bin> <steps>\n");
    return 2;
  }
  const int steps = std::atoi(argv[2]);
  try {
    SPHParameters p{};
    p.kernelHeight = 1.0f;
    p.cellSize = 1.0f;
    p.particleSpacing = 0.5f;
    p.restDensity = 10.0f;
    p.stiffness = 0.004f;
    p.nearStiffness = 0.01f;
    p.linearViscosity = 0.0f;
    p.quadraticViscosity = 0.3f;
    Demo5::ParticleSimulation sim(p, 80.0f, 60.0f, 1.0f, 1.0f / 6.0f, 1.0f / 30.0f);
    // LoadScenario order: stats, clears, gravity, params, bodies, volumes
    sim.ResetStats();
    sim.ClearBodies();
    sim.ClearParticles();
    sim.ClearEmitters();
    sim.SetGravity(Vec2f(0.0f, -9.8f));
    sim.SetParams(p);
    const float hw = 40.0f, hh = 30.0f;
    const Vec2f walls[4][2] = {{Vec2f(0, 1), Vec2f(0, -hh)}, {Vec2f(0, -1), Vec2f(0, hh)},
                               {Vec2f(1, 0), Vec2f(-hw, 0)}, {Vec2f(-1, 0), Vec2f(hw, 0)}};
    for (auto &w : walls) sim.AddPlane(w[0], w[0].x * w[1].x + w[0].y * w[1].y);  // app.cpp:493 d = n . p
    sim.SetSeed(0);
    sim.AddVolume(Vec2f(-24.0f, -14.0f), Vec2f(0.0f, 0.0f), 64, 64, 0.5f);
    for (int s = 0; s < steps; ++s) sim.Update(0.005f);
    const Demo5::ParticleData *d = sim.GetParticleDatas();
    const size_t n = sim.GetParticleCount();
    SPHStatistics &st = sim.GetStats();
    FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    const uint64_t hdr[5] = {n, st.minParticleNeighborCount, st.maxParticleNeighborCount, st.minCellParticleCount,
                             st.maxCellParticleCount};
    std::fwrite(hdr, sizeof(hdr), 1, f);
    std::fwrite(d, sizeof(Demo5::ParticleData), n, f);
    std::fclose(f);
    std::printf("demo5_driver: %zu particles, %d steps, multithreading supported %d\n", n, steps,
                (int)sim.IsMultiThreadingSupported());
  } catch (const Demo5::Error &e) {
    std::fprintf(stderr, "demo5_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
