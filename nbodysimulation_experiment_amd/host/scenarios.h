// The reference's built-in scenarios and its scenario loader, over any
// BaseSimulation:
//   - the scenario table SPHScenarios (NBodySimulation/sph.h:262-437) with the
//     constants it is built from (sph.h:14-72),
//   - DemoApplication::LoadScenario (app.cpp:477-534).
//
// The scenes live in the reference's fixed domain: a 10 x 5.625 box centred on
// the origin with a 33 x 18 grid of 0.3 cells (sph.h:17-22, 58-63), which does
// not cover the whole box (particles beyond it clamp into the border cells).
#ifndef NBS_SCENARIOS_H
#define NBS_SCENARIOS_H

#include <string>
#include <vector>

#include "sph_interface.h"

namespace Demo5 {

// Compile-time constants of sph.h:17-54, evaluated in float in source order.
struct ReferenceConstants {
  float boundaryWidth, boundaryHeight, boundaryHalfWidth, boundaryHalfHeight;
  float deltaTime;          // kSPHSubstepDeltaTime (one substep per frame)
  float particleRadius;     // kSPHParticleRadius
  float kernelHeight;       // kSPHKernelHeight
  float particleSpacing;    // kSPHParticleSpacing
  float collisionRadius;    // kSPHParticleCollisionRadius
  float collisionMargin;    // kSPHCollisionMargin
  float restDensity, stiffness, nearStiffness, linearViscosity, quadraticViscosity;
  float gridCellSize;
  int gridCountX, gridCountY;
};
const ReferenceConstants &Constants();

enum class BodyKind { Plane, Circle, LineSegment, Polygon };

// SPHScenarioBody (sph.h:162-222): a position, an orientation (the two columns
// of a 2x2 matrix), a radius and local vertices.
struct ScenarioBody {
  BodyKind kind;
  Vec2f position;
  Vec2f col1, col2;
  float radius;
  std::vector<Vec2f> localVerts;
};

struct ScenarioVolume {  // SPHScenarioVolume (sph.h:224-235)
  Vec2f position, size, force;
};

struct ScenarioEmitter {  // SPHScenarioEmitter (sph.h:237-260)
  Vec2f position, direction;
  float radius, speed, rate, duration;
};

struct Scenario {  // SPHScenario (sph.h:265-304)
  std::string name;
  Vec2f gravity;
  std::vector<ScenarioVolume> volumes;
  std::vector<ScenarioEmitter> emitters;
  std::vector<ScenarioBody> bodies;
  SPHParameters parameters;
};

// The eight scenarios in the reference's order: Dambreak, Dambreak x 2, Blob,
// Blob x 2, Liquid, Glass, Fontain, Fun.
const std::vector<Scenario> &Scenarios();

// The world-space form of a body as LoadScenario hands it to the simulation:
// planes as (normal, distance = normal . position), circles as (centre,
// radius), segments and polygons as transformed vertices (app.cpp:487-515).
struct ResolvedBody {
  BodyKind kind;
  Vec2f normal;
  float distance;
  Vec2f center;
  float radius;
  std::vector<Vec2f> verts;
};
ResolvedBody ResolveBody(const ScenarioBody &body);

// Particles per axis of a volume: floor(size / spacing) in float (app.cpp:523-524).
void VolumeCounts(const ScenarioVolume &volume, float spacing, int &countX, int &countY);

// DemoApplication::LoadScenario: reset statistics, clear bodies, particles and
// emitters, set gravity and parameters, then add bodies, volumes and emitters
// in the table's order.
void LoadScenario(BaseSimulation &sim, const Scenario &scenario);

}  // namespace Demo5

#endif  // NBS_SCENARIOS_H
