// The simulation interface the reference's demos share, declared with the same
// names, member layout and virtual methods as NBodySimulation/base.h:8-39,
// sph.h:77-150 and vecmath.h:11-54, so Demo5::ParticleSimulation can be built
// and tested here.  Inside the reference application the reference's own
// headers declare these types instead (NBS_WITH_REFERENCE_HEADERS).
#ifndef NBS_SPH_INTERFACE_H
#define NBS_SPH_INTERFACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef NBS_WITH_REFERENCE_HEADERS
#include "base.h"
#else

union Vec2i {  // vecmath.h:11-28
  struct {
    int x, y;
  };
  int m[2];
  Vec2i() : x(0), y(0) {}
  Vec2i(int ix, int iy) : x(ix), y(iy) {}
};

union Vec2f {  // vecmath.h:30-54
  struct {
    float x, y;
  };
  float m[2];
  Vec2f() : x(0), y(0) {}
  explicit Vec2f(float v) : x(v), y(v) {}
  Vec2f(float ix, float iy) : x(ix), y(iy) {}
};

struct SPHParameters {  // sph.h:77-123 (fields and order)
  float kernelHeight;
  float cellSize;
  float particleSpacing;
  float invKernelHeight;
  float restDensity;
  float stiffness;
  float nearStiffness;
  float linearViscosity;
  float quadraticViscosity;
};

struct SPHStatistics {  // sph.h:125-150
  size_t minParticleNeighborCount;
  size_t maxParticleNeighborCount;
  size_t minCellParticleCount;
  size_t maxCellParticleCount;
  struct {
    float emitters;
    float integration;
    float viscosityForces;
    float predict;
    float updateGrid;
    float neighborSearch;
    float densityAndPressure;
    float deltaPositions;
    float collisions;
  } time;
};

namespace Render {
struct CommandBuffer;  // render.h:163-177 (drawing is the host application's)
}

class BaseSimulation {  // base.h:8-39
 public:
  virtual ~BaseSimulation() {}
  virtual void ResetStats() = 0;
  virtual void ClearBodies() = 0;
  virtual void ClearParticles() = 0;
  virtual void ClearEmitters() = 0;

  virtual void AddPlane(const Vec2f &normal, const float distance) = 0;
  virtual void AddCircle(const Vec2f &pos, const float radius) = 0;
  virtual void AddLineSegment(const Vec2f &a, const Vec2f &b) = 0;
  virtual void AddPolygon(const size_t vertexCount, const Vec2f *verts) = 0;

  virtual size_t AddParticle(const Vec2f &position, const Vec2f &force) = 0;
  virtual void AddVolume(const Vec2f &center, const Vec2f &force, const int countX, const int countY,
                         const float spacing) = 0;
  virtual void AddEmitter(const Vec2f &position, const Vec2f &direction, const float radius, const float speed,
                          const float rate, const float duration) = 0;

  virtual void Update(const float deltaTime) = 0;
  virtual void Render(Render::CommandBuffer *commandBuffer, const float worldToScreenScale) = 0;

  virtual void AddExternalForces(const Vec2f &force) = 0;
  virtual void ClearExternalForce() = 0;

  virtual size_t GetParticleCount() = 0;
  virtual void SetGravity(const Vec2f &gravity) = 0;
  virtual const SPHParameters &GetParams() = 0;
  virtual SPHStatistics &GetStats() = 0;
  virtual void SetParams(const SPHParameters &params) = 0;
  virtual void SetMultiThreading(const bool value) = 0;
  virtual bool IsMultiThreadingSupported() = 0;
  virtual bool IsMultiThreading() = 0;
  virtual size_t GetWorkerThreadCount() = 0;
};

#endif  // NBS_WITH_REFERENCE_HEADERS
#endif  // NBS_SPH_INTERFACE_H
