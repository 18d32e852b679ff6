// Scenario table and loader (see scenarios.h).
#include "scenarios.h"

#include <cmath>

namespace Demo5 {

static ReferenceConstants makeConstants() {
  // sph.h:17-63, each constant in float, in source order.
  ReferenceConstants c{};
  const float aspect = 16.0f / 9.0f;
  c.boundaryWidth = 10.0f;
  c.boundaryHeight = c.boundaryWidth / aspect;
  c.boundaryHalfWidth = c.boundaryWidth * 0.5f;
  c.boundaryHalfHeight = c.boundaryHeight * 0.5f;
  const float frameDt = 1.0f / 60.0f;
  c.deltaTime = frameDt / (float)1;  // one substep (sph.h:31-33)
  c.particleRadius = 0.05f;
  c.kernelHeight = 6.0f * c.particleRadius;
  c.particleSpacing = c.kernelHeight * 0.5f;
  c.collisionRadius = c.particleRadius;
  c.restDensity = 20.0f;
  c.stiffness = 0.6f;
  c.nearStiffness = c.stiffness * 10.0f;
  c.linearViscosity = 0.5f;
  c.quadraticViscosity = 0.3f;
  c.collisionMargin = 0.005f * 2.0f;
  c.gridCellSize = c.kernelHeight;
  c.gridCountX = (int)(c.boundaryWidth / c.gridCellSize);
  c.gridCountY = (int)(c.boundaryHeight / c.gridCellSize);
  return c;
}

const ReferenceConstants &Constants() {
  static const ReferenceConstants c = makeConstants();
  return c;
}

// The eight-argument SPHParameters constructor (sph.h:112-122) keeps the
// compile-time kernel height and rest density whatever is passed, and takes
// 1/h of the argument.
static SPHParameters scenarioParams(float particleSpacing, float stiffness, float nearStiffness) {
  const ReferenceConstants &c = Constants();
  const float kernelHeightArg = c.kernelHeight;
  SPHParameters p;
  p.kernelHeight = c.kernelHeight;
  p.cellSize = c.gridCellSize;
  p.particleSpacing = particleSpacing;
  p.invKernelHeight = 1.0f / kernelHeightArg;
  p.restDensity = c.restDensity;
  p.stiffness = stiffness;
  p.nearStiffness = nearStiffness;
  p.linearViscosity = c.linearViscosity;
  p.quadraticViscosity = c.quadraticViscosity;
  return p;
}

// Body constructors of sph.h:178-222.  sinf/cosf are taken as the float
// rounding of the double-precision sine and cosine.
static ScenarioBody plane(Vec2f position, Vec2f normal) {
  ScenarioBody b;
  b.kind = BodyKind::Plane;
  b.position = position;
  b.col1 = normal;                                  // Mat2FromAxis (vecmath.h:367-372)
  b.col2 = Vec2f(-1.0f * normal.y, 1.0f * normal.x);  // Vec2Cross(1, axis) (vecmath.h:304-306)
  b.radius = 0.0f;
  return b;
}

static ScenarioBody circle(Vec2f position, float radius) {
  ScenarioBody b;
  b.kind = BodyKind::Circle;
  b.position = position;
  b.col1 = Vec2f(1.0f, 0.0f);
  b.col2 = Vec2f(0.0f, 1.0f);
  b.radius = radius;
  return b;
}

static void rotation(float angle, Vec2f &col1, Vec2f &col2) {  // Mat2FromAngle (vecmath.h:358-365)
  const float s = (float)std::sin((double)angle);
  const float c = (float)std::cos((double)angle);
  col1 = Vec2f(c, s);
  col2 = Vec2f(-s, c);
}

static ScenarioBody box(Vec2f position, float angle, Vec2f ext) {
  ScenarioBody b;
  b.kind = BodyKind::Polygon;
  b.position = position;
  rotation(angle, b.col1, b.col2);
  b.radius = 0.0f;
  b.localVerts = {Vec2f(ext.x, ext.y), Vec2f(-ext.x, ext.y), Vec2f(-ext.x, -ext.y), Vec2f(ext.x, -ext.y)};
  return b;
}

static std::vector<ScenarioBody> walls(bool top) {  // bottom, top, left, right (sph.h:323-326)
  const ReferenceConstants &c = Constants();
  std::vector<ScenarioBody> w;
  w.push_back(plane(Vec2f(0.0f, -c.boundaryHalfHeight), Vec2f(0.0f, 1.0f)));
  if (top) w.push_back(plane(Vec2f(0.0f, c.boundaryHalfHeight), Vec2f(0.0f, -1.0f)));
  w.push_back(plane(Vec2f(-c.boundaryHalfWidth, 0.0f), Vec2f(1.0f, 0.0f)));
  w.push_back(plane(Vec2f(c.boundaryHalfWidth, 0.0f), Vec2f(-1.0f, 0.0f)));
  return w;
}

static std::vector<Scenario> makeScenarios() {
  const ReferenceConstants &c = Constants();
  const float h = c.kernelHeight, k = c.stiffness;
  const float wallW = c.boundaryWidth * 0.05f, wallH = c.boundaryHeight * 0.85f;   // sph.h:306-307
  const float volW = c.boundaryWidth * 0.25f, volH = c.boundaryHeight * 0.95f;     // sph.h:308-309
  const float blobW = c.boundaryWidth * 0.5f, blobH = c.boundaryHeight * 0.5f;     // sph.h:311-312
  const float deg2rad = (float)M_PI / 180.0f;                                       // vecmath.h:9
  const Vec2f zero(0.0f, 0.0f);
  std::vector<Scenario> s;

  {  // sph.h:315-328
    Scenario d{"Dambreak", Vec2f(0.0f, -10.0f), {}, {}, walls(true), scenarioParams(h / 6.0f, k, c.nearStiffness)};
    d.volumes.push_back({Vec2f(-c.boundaryHalfWidth + volW * 0.5f, 0.0f), Vec2f(volW, volH), zero});
    d.bodies.push_back(box(Vec2f(-c.boundaryHalfWidth + volW + wallW * 0.5f + c.collisionRadius,
                                 c.boundaryHeight * 0.1f),
                           0.0f, Vec2f(wallW * 0.5f, wallH * 0.5f)));
    s.push_back(d);
  }
  {  // sph.h:330-343
    Scenario d{"Dambreak x 2", Vec2f(0.0f, -10.0f), {}, {}, walls(true), scenarioParams(h / 3.0f, k, k * 20.0f)};
    d.volumes.push_back({Vec2f(-c.boundaryHalfWidth + volW * 0.5f, 0.0f), Vec2f(volW, volH), zero});
    d.volumes.push_back({Vec2f(c.boundaryHalfWidth - volW * 0.5f, 0.0f), Vec2f(volW, volH), zero});
    s.push_back(d);
  }
  {  // sph.h:345-357
    Scenario d{"Blob", Vec2f(0.0f, 0.0f), {}, {}, walls(true), scenarioParams(h / 3.0f, k, c.nearStiffness)};
    d.volumes.push_back({zero, Vec2f(blobW, blobH), zero});
    s.push_back(d);
  }
  {  // sph.h:359-372
    Scenario d{"Blob x 2", Vec2f(0.0f, 0.0f), {}, {}, walls(true), scenarioParams(h / 3.0f, k, c.nearStiffness)};
    const float q = blobH * 0.75f;
    d.volumes.push_back({Vec2f(-q, 0.0f), Vec2f(q, q), Vec2f(10.0f, 0.0f)});
    d.volumes.push_back({Vec2f(q, 0.0f), Vec2f(q, q), Vec2f(-10.0f, 0.0f)});
    s.push_back(d);
  }
  {  // sph.h:374-386
    Scenario d{"Liquid", Vec2f(0.0f, -2.0f), {}, {}, walls(true), scenarioParams(h / 4.0f, k, k * 10.0f)};
    d.emitters.push_back({Vec2f(-3.5f, 0.0f), Vec2f(1.0f, 0.0f), h * 3.0f, 2.5f, 15.0f, 30.0f});
    s.push_back(d);
  }
  {  // sph.h:388-403
    Scenario d{"Glass", Vec2f(0.0f, -10.0f), {}, {}, walls(true), scenarioParams(h / 4.0f, k, k * 6.0f)};
    d.emitters.push_back({Vec2f(-1.5f, 2.0f), Vec2f(1.0f, 0.0f), h * 3.0f, 2.5f, 15.0f, 25.0f});
    d.bodies.push_back(box(Vec2f(0.0f, -2.0f), 0.0f, Vec2f(1.0f, 0.2f)));
    d.bodies.push_back(box(Vec2f(-1.0f, -0.5f), 0.0f, Vec2f(0.2f, 1.5f)));
    d.bodies.push_back(box(Vec2f(1.0f, -0.5f), 0.0f, Vec2f(0.2f, 1.5f)));
    s.push_back(d);
  }
  {  // sph.h:405-417
    Scenario d{"Fontain", Vec2f(0.0f, -10.0f), {}, {}, walls(true), scenarioParams(h / 4.0f, k, k * 2.0f)};
    d.emitters.push_back({Vec2f(0.0f, -c.boundaryHalfHeight + 1.0f), Vec2f(0.0f, 1.0f), h * 4.0f, 8.0f, 15.0f, 25.0f});
    s.push_back(d);
  }
  {  // sph.h:419-436 (no top wall)
    Scenario d{"Fun", Vec2f(0.0f, -10.0f), {}, {}, walls(false), scenarioParams(h / 4.0f, k, k * 6.0f)};
    d.emitters.push_back({Vec2f(-4.0f, 2.0f), Vec2f(1.0f, 0.0f), h * 4.0f, 3.5f, 15.0f, 20.0f});
    d.bodies.push_back(box(Vec2f(-1.5f, 1.0f), deg2rad * -2.5f, Vec2f(3.5f, 0.1f)));
    d.bodies.push_back(box(Vec2f(1.5f, -0.25f), deg2rad * 2.5f, Vec2f(3.5f, 0.1f)));
    d.bodies.push_back(circle(Vec2f(-4.0f, -1.5f), 0.5f));
    d.bodies.push_back(box(Vec2f(0.0f, -c.boundaryHalfHeight + 0.5f), 0.0f, Vec2f(0.3f, 1.0f)));
    s.push_back(d);
  }
  return s;
}

const std::vector<Scenario> &Scenarios() {
  static const std::vector<Scenario> s = makeScenarios();
  return s;
}

// Vec2MultMat2 (vecmath.h:287-290) followed by + position (vecmath.h:245-248).
static Vec2f toWorld(const ScenarioBody &b, const Vec2f &v) {
  const Vec2f r(b.col1.x * v.x + b.col2.x * v.y, b.col1.y * v.x + b.col2.y * v.y);
  return Vec2f(r.x + b.position.x, r.y + b.position.y);
}

ResolvedBody ResolveBody(const ScenarioBody &b) {
  ResolvedBody r{};
  r.kind = b.kind;
  switch (b.kind) {
    case BodyKind::Plane:  // app.cpp:490-494: distance = Vec2Dot(col1, position)
      r.normal = b.col1;
      r.distance = b.col1.x * b.position.x + b.col1.y * b.position.y;
      break;
    case BodyKind::Circle:
      r.center = b.position;
      r.radius = b.radius;
      break;
    case BodyKind::LineSegment:
    case BodyKind::Polygon:
      for (const Vec2f &v : b.localVerts) r.verts.push_back(toWorld(b, v));
      break;
  }
  return r;
}

void VolumeCounts(const ScenarioVolume &v, float spacing, int &countX, int &countY) {
  const float fx = v.size.x / spacing, fy = v.size.y / spacing;
  countX = (int)std::floor(fx);
  countY = (int)std::floor(fy);
}

void LoadScenario(BaseSimulation &sim, const Scenario &s) {
  sim.ResetStats();
  sim.ClearBodies();
  sim.ClearParticles();
  sim.ClearEmitters();
  sim.SetGravity(s.gravity);
  sim.SetParams(s.parameters);
  for (const ScenarioBody &b : s.bodies) {
    const ResolvedBody r = ResolveBody(b);
    switch (r.kind) {
      case BodyKind::Plane: sim.AddPlane(r.normal, r.distance); break;
      case BodyKind::Circle: sim.AddCircle(r.center, r.radius); break;
      case BodyKind::LineSegment: sim.AddLineSegment(r.verts[0], r.verts[1]); break;
      case BodyKind::Polygon: sim.AddPolygon(r.verts.size(), r.verts.data()); break;
    }
  }
  // app.cpp:519-526 (the reference runs this loop to the body count; every
  // scenario has at least as many bodies as volumes, so the extra iterations
  // only add empty volumes)
  const float spacing = sim.GetParams().particleSpacing;
  for (const ScenarioVolume &v : s.volumes) {
    int nx = 0, ny = 0;
    VolumeCounts(v, spacing, nx, ny);
    sim.AddVolume(v.position, v.force, nx, ny, spacing);
  }
  for (const ScenarioEmitter &e : s.emitters)
    sim.AddEmitter(e.position, e.direction, e.radius, e.speed, e.rate, e.duration);
}

}  // namespace Demo5
