// Demo5::ParticleSimulation over the C ABI (see demo5.h).
#include "demo5.h"

#include <cstring>

namespace Demo5 {

static nbs_params toC(const SPHParameters &p, float radius, float margin) {
  nbs_params c;
  c.kernel_height = p.kernelHeight;
  c.cell_size = p.cellSize;
  c.particle_spacing = p.particleSpacing;
  c.rest_density = p.restDensity;
  c.stiffness = p.stiffness;
  c.near_stiffness = p.nearStiffness;
  c.linear_viscosity = p.linearViscosity;
  c.quadratic_viscosity = p.quadraticViscosity;
  c.collision_radius = radius;
  c.collision_margin = margin;
  return c;
}

void ParticleSimulation::check(int status, const char *what) {
  if (status != NBS_OK)
    throw Error(status, std::string(what) + ": " + nbs_status_string(status) + ": " + nbs_last_error());
}

ParticleSimulation::ParticleSimulation(const SPHParameters &params, float width, float height, float cellSize,
                                       float collisionRadius, float collisionMargin, size_t capacity)
    : params_(params), collisionRadius_(collisionRadius), collisionMargin_(collisionMargin) {
  params_.invKernelHeight = 1.0f / params_.kernelHeight;  // sph.h:92
  const nbs_params c = toC(params_, collisionRadius_, collisionMargin_);
  const nbs_domain d = {width, height, cellSize};
  check(nbs_create(&c, &d, capacity, &sim_), "nbs_create");
  ResetStats();
}

ParticleSimulation::~ParticleSimulation() { nbs_destroy(sim_); }

void ParticleSimulation::ResetStats() {
  check(nbs_reset_stats(sim_), "nbs_reset_stats");
  stats_ = SPHStatistics{};
  stats_.minParticleNeighborCount = 500;  // sph.h:143-147
  stats_.minCellParticleCount = 500;
}

void ParticleSimulation::ClearBodies() { check(nbs_clear_bodies(sim_), "nbs_clear_bodies"); }

void ParticleSimulation::ClearParticles() {
  check(nbs_clear_particles(sim_), "nbs_clear_particles");
  mirrorDirty_ = true;
}

void ParticleSimulation::ClearEmitters() { check(nbs_clear_emitters(sim_), "nbs_clear_emitters"); }

void ParticleSimulation::AddPlane(const Vec2f &normal, const float distance) {
  check(nbs_add_plane(sim_, 0, normal.x, normal.y, distance), "nbs_add_plane");
}

void ParticleSimulation::AddCircle(const Vec2f &pos, const float radius) {
  check(nbs_add_circle(sim_, 0, pos.x, pos.y, radius), "nbs_add_circle");
}

void ParticleSimulation::AddLineSegment(const Vec2f &a, const Vec2f &b) {
  check(nbs_add_line_segment(sim_, 0, a.x, a.y, b.x, b.y), "nbs_add_line_segment");
}

void ParticleSimulation::AddPolygon(const size_t vertexCount, const Vec2f *verts) {
  scratch_.resize(2 * vertexCount);
  for (size_t k = 0; k < vertexCount; ++k) scratch_[2 * k] = verts[k].x, scratch_[2 * k + 1] = verts[k].y;
  check(nbs_add_polygon(sim_, 0, vertexCount, scratch_.data()), "nbs_add_polygon");
}

size_t ParticleSimulation::AddParticle(const Vec2f &position, const Vec2f &force) {
  const float pos[2] = {position.x, position.y}, frc[2] = {force.x, force.y};
  uint64_t first = 0;
  check(nbs_add_particles(sim_, 0, pos, frc, 1, &first), "nbs_add_particles");
  mirrorDirty_ = true;
  return (size_t)first;
}

void ParticleSimulation::AddVolume(const Vec2f &center, const Vec2f &force, const int countX, const int countY,
                                   const float spacing) {
  check(nbs_add_volume(sim_, 0, center.x, center.y, force.x, force.y, countX, countY, spacing, seed_),
        "nbs_add_volume");
  mirrorDirty_ = true;
}

void ParticleSimulation::AddEmitter(const Vec2f &position, const Vec2f &direction, const float radius,
                                    const float speed, const float rate, const float duration) {
  check(nbs_add_emitter(sim_, 0, position.x, position.y, direction.x, direction.y, radius, speed, rate, duration,
                        seed_),
        "nbs_add_emitter");
}

void ParticleSimulation::Update(const float deltaTime) { UpdateSteps(deltaTime, 1); }

void ParticleSimulation::UpdateSteps(float deltaTime, uint32_t steps) {
  check(nbs_update(sim_, deltaTime, steps), "nbs_update");
  mirrorDirty_ = true;
}

void ParticleSimulation::refreshMirror() {
  if (!mirrorDirty_) return;
  const size_t n = GetParticleCount();
  mirror_.resize(n);
  std::vector<float> x(2 * n), prev(2 * n), v(2 * n), dens(2 * n), pres(2 * n);
  if (n) check(nbs_get_state(sim_, x.data(), prev.data(), v.data(), dens.data(), pres.data(), nullptr, nullptr),
               "nbs_get_state");
  for (size_t i = 0; i < n; ++i) {
    ParticleData &d = mirror_[i];
    d.curPosition = Vec2f(x[2 * i], x[2 * i + 1]);
    d.prevPosition = Vec2f(prev[2 * i], prev[2 * i + 1]);
    d.acceleration = Vec2f(0.0f, 0.0f);  // zero between steps (demo4.cpp:308)
    d.velocity = Vec2f(v[2 * i], v[2 * i + 1]);
    d.density = dens[2 * i];
    d.nearDensity = dens[2 * i + 1];
    d.pressure = pres[2 * i];
    d.nearPressure = pres[2 * i + 1];
  }
  mirrorDirty_ = false;
}

const ParticleData *ParticleSimulation::GetParticleDatas() {
  refreshMirror();
  return mirror_.data();
}

const float *ParticleSimulation::GetParticleColors() {
  colors_.resize(4 * GetParticleCount());
  if (!colors_.empty()) check(nbs_get_colors(sim_, colors_.data()), "nbs_get_colors");
  return colors_.data();
}

// Drawing belongs to the host application (demo4.cpp:453-532 pushes the grid,
// bodies and the particle vertex array into the command buffer).  This method
// refreshes the 48-byte mirror that vertex array points into and the colours
// computed on the GPU.
void ParticleSimulation::Render(Render::CommandBuffer *, const float) {
  refreshMirror();
  GetParticleColors();
}

void ParticleSimulation::AddExternalForces(const Vec2f &force) {
  check(nbs_add_external_force(sim_, force.x, force.y), "nbs_add_external_force");
}

void ParticleSimulation::ClearExternalForce() { check(nbs_clear_external_force(sim_), "nbs_clear_external_force"); }

size_t ParticleSimulation::GetParticleCount() {
  uint64_t n = 0;
  check(nbs_get_particle_count(sim_, &n), "nbs_get_particle_count");
  return (size_t)n;
}

void ParticleSimulation::SetGravity(const Vec2f &gravity) {
  check(nbs_set_gravity(sim_, gravity.x, gravity.y), "nbs_set_gravity");
}

const SPHParameters &ParticleSimulation::GetParams() { return params_; }

SPHStatistics &ParticleSimulation::GetStats() {
  nbs_stats s;
  check(nbs_get_stats(sim_, &s), "nbs_get_stats");
  stats_.minParticleNeighborCount = (size_t)s.min_particle_neighbor_count;
  stats_.maxParticleNeighborCount = (size_t)s.max_particle_neighbor_count;
  stats_.minCellParticleCount = (size_t)s.min_cell_particle_count;
  stats_.maxCellParticleCount = (size_t)s.max_cell_particle_count;
  stats_.time.emitters = s.time_emitters;
  stats_.time.integration = s.time_integration;
  stats_.time.viscosityForces = s.time_viscosity_forces;
  stats_.time.predict = s.time_predict;
  stats_.time.updateGrid = s.time_update_grid;
  stats_.time.neighborSearch = s.time_neighbor_search;
  stats_.time.densityAndPressure = s.time_density_and_pressure;
  stats_.time.deltaPositions = s.time_delta_positions;
  stats_.time.collisions = s.time_collisions;
  return stats_;
}

void ParticleSimulation::SetParams(const SPHParameters &params) {
  params_ = params;
  params_.invKernelHeight = 1.0f / params_.kernelHeight;  // sph.h:102 copy ctor recomputes it
  const nbs_params c = toC(params_, collisionRadius_, collisionMargin_);
  check(nbs_set_params(sim_, &c), "nbs_set_params");
}

// The GPU is the parallelism: the switch is kept for the interface and ignored.
void ParticleSimulation::SetMultiThreading(const bool value) { multiThreading_ = value; }
bool ParticleSimulation::IsMultiThreadingSupported() { return false; }
bool ParticleSimulation::IsMultiThreading() { return false; }
size_t ParticleSimulation::GetWorkerThreadCount() { return 1; }

}  // namespace Demo5
