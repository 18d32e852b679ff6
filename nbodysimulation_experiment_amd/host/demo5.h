// Demo5::ParticleSimulation: the reference's BaseSimulation implemented over
// the MI355X library's C ABI (include/nbs_sph.h), a drop-in for
// Demo4::ParticleSimulation (NBodySimulation/demo4.h:137-226).
//
// The particle state lives in HBM.  A host mirror with Demo4's 48-byte
// ParticleData layout (demo4.h:81-109) is refreshed on demand
// (GetParticleDatas, Render), because the reference's renderer reads
// curPosition at offset 0 of that record with a 48-byte stride
// (demo4.cpp:521-531, main.cpp:123-140).
//
// Errors: every C ABI status other than NBS_OK throws Demo5::Error (the
// reference's methods return void and check capacities with assert only).
#ifndef NBS_DEMO5_H
#define NBS_DEMO5_H

#include <stdexcept>
#include <string>
#include <vector>

#include "nbs_sph.h"
#include "sph_interface.h"

namespace Demo5 {

struct ParticleData {  // demo4.h:81-109
  Vec2f curPosition;
  Vec2f prevPosition;
  Vec2f acceleration;
  Vec2f velocity;
  float density, nearDensity;
  float pressure, nearPressure;
};
static_assert(sizeof(ParticleData) == 48, "ParticleData must keep Demo4's 48-byte layout");

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &what) : std::runtime_error(what), status(status) {}
  int status;
};

class ParticleSimulation : public BaseSimulation {
 public:
  // The domain is an origin-centred box of width x height with a grid of
  // cellSize cells (sph.h:18-23, 58-65); params as SPHParameters, plus the
  // collision radius and margin the reference keeps as constants.
  ParticleSimulation(const SPHParameters &params, float width, float height, float cellSize,
                     float collisionRadius, float collisionMargin, size_t capacity = 0);
  ~ParticleSimulation() override;
  ParticleSimulation(const ParticleSimulation &) = delete;
  ParticleSimulation &operator=(const ParticleSimulation &) = delete;

  void ResetStats() override;
  void ClearBodies() override;
  void ClearParticles() override;
  void ClearEmitters() override;

  void AddPlane(const Vec2f &normal, const float distance) override;
  void AddCircle(const Vec2f &pos, const float radius) override;
  void AddLineSegment(const Vec2f &a, const Vec2f &b) override;
  void AddPolygon(const size_t vertexCount, const Vec2f *verts) override;

  size_t AddParticle(const Vec2f &position, const Vec2f &force) override;
  void AddVolume(const Vec2f &center, const Vec2f &force, const int countX, const int countY,
                 const float spacing) override;
  void AddEmitter(const Vec2f &position, const Vec2f &direction, const float radius, const float speed,
                  const float rate, const float duration) override;

  void Update(const float deltaTime) override;
  void Render(Render::CommandBuffer *commandBuffer, const float worldToScreenScale) override;

  void AddExternalForces(const Vec2f &force) override;
  void ClearExternalForce() override;

  size_t GetParticleCount() override;
  void SetGravity(const Vec2f &gravity) override;
  const SPHParameters &GetParams() override;
  SPHStatistics &GetStats() override;
  void SetParams(const SPHParameters &params) override;
  void SetMultiThreading(const bool value) override;
  bool IsMultiThreadingSupported() override;
  bool IsMultiThreading() override;
  size_t GetWorkerThreadCount() override;

  // Extensions.
  void UpdateSteps(float deltaTime, uint32_t steps);  // several steps per call
  const ParticleData *GetParticleDatas();              // refreshed 48-byte mirror
  const float *GetParticleColors();                    // RGBA per particle (demo4.cpp:521-524)
  nbs_sim *Handle() { return sim_; }
  void SetSeed(uint64_t seed) { seed_ = seed; }         // jitter seed of the next AddVolume/AddEmitter

 private:
  void check(int status, const char *what);
  void refreshMirror();

  nbs_sim *sim_ = nullptr;
  SPHParameters params_{};
  SPHStatistics stats_{};
  float collisionRadius_, collisionMargin_;
  std::vector<ParticleData> mirror_;
  std::vector<float> colors_;
  std::vector<float> scratch_;
  bool mirrorDirty_ = true;
  bool multiThreading_ = false;
  uint64_t seed_ = 0;
};

}  // namespace Demo5

#endif  // NBS_DEMO5_H
