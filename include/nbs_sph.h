/*
 * nbs_sph.h - C ABI of the MI355X SPH step library (libnbs_sph.so).
 *
 * The entry points replace the reference's simulation interface
 * `class BaseSimulation` (NBodySimulation/base.h:8-39) as implemented by
 * `Demo4::ParticleSimulation` (NBodySimulation/demo4.h:137-226,
 * demo4.cpp:14-451).  Each declaration names the method it stands in for.
 *
 * Conventions
 *   - Every call returns an nbs_status; no C++ exception crosses the ABI.
 *   - Pointers are plain host pointers; arrays of 2D vectors are interleaved
 *     x,y float pairs.  Device memory is owned by the library.
 *   - Particle arrays exchanged with the caller are in particle-index order
 *     (the order particles were added, like Demo4's particleDatas[]).
 *   - A handle is used from one host thread at a time (the reference calls its
 *     simulation only from the main thread, main.cpp:512-556).
 *   - A handle may hold several independent scenes (batched simulation);
 *     nbs_create makes scene 0.  Single-scene callers pass scene 0.
 *   - The library computes on one GPU: the device current when nbs_create is
 *     called.
 */
#ifndef NBS_SPH_H
#define NBS_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nbs_status {
  NBS_OK = 0,
  NBS_ERR_INVALID_ARGUMENT = 1,
  NBS_ERR_OUT_OF_MEMORY = 2,
  NBS_ERR_HIP = 3,
  NBS_ERR_CAPACITY = 4,   /* a size limit of the library was exceeded */
  NBS_ERR_NON_FINITE = 5, /* NaN/Inf found in the particle state */
  NBS_ERR_NO_DEVICE = 6,
  NBS_ERR_UNSUPPORTED = 7
} nbs_status;

typedef struct nbs_sim nbs_sim;

/* SPHParameters (NBodySimulation/sph.h:77-123) without the derived
 * invKernelHeight, plus the two collision constants the reference keeps at
 * compile time: kSPHParticleCollisionRadius (sph.h:35-38) and
 * kSPHCollisionMargin (sph.h:53-54). */
typedef struct nbs_params {
  float kernel_height;       /* h, in (0, 2^40] */
  float cell_size;           /* informational, as in SPHParameters; the grid uses nbs_domain */
  float particle_spacing;    /* emitter column spacing (demo4.cpp:258) */
  float rest_density;        /* rho0 */
  float stiffness;           /* k */
  float near_stiffness;      /* k_near */
  float linear_viscosity;    /* sigma */
  float quadratic_viscosity; /* beta */
  float collision_radius;    /* R */
  float collision_margin;    /* line-segment and polygon contact margin */
} nbs_params;

/* An origin-centred box and its uniform grid (sph.h:18-23, 58-65):
 * countX = (int)(width / cell_size), countY = (int)(height / cell_size). */
typedef struct nbs_domain {
  float width;
  float height;
  float cell_size;
} nbs_domain;

/* SPHStatistics (sph.h:125-150).  Candidate counts are the 3x3-cell candidate
 * sets of the last step (demo4.cpp:369-376); cell counts are the smallest and
 * largest occupied-cell population of the last binning.  Times are the last
 * step's milliseconds per phase, measured with HIP events when timing is on;
 * fused phases report their time in the first field of the fusion and 0 in
 * the others (integration and predict run inside viscosityForces; collisions
 * run inside deltaPositions; neighbour lists are not materialised). */
typedef struct nbs_stats {
  uint64_t min_particle_neighbor_count;
  uint64_t max_particle_neighbor_count;
  uint64_t min_cell_particle_count;
  uint64_t max_cell_particle_count;
  float time_emitters;
  float time_integration;
  float time_viscosity_forces;
  float time_predict;
  float time_update_grid;
  float time_neighbor_search;
  float time_density_and_pressure;
  float time_delta_positions;
  float time_collisions;
} nbs_stats;

/* FP64 reductions over the current particle state, computed on the GPU. */
typedef struct nbs_diagnostics {
  uint64_t particle_count;     /* mass, in particle units */
  double momentum[2];          /* sum of velocities */
  double kinetic_energy;       /* 0.5 * sum |v|^2 */
  double mean_density;         /* mean of density */
  double max_density;
  double mean_abs_density_error; /* mean |density - rest_density| / rest_density */
  double rms_density_error;      /* sqrt(mean((density - rest_density)^2)) / rest_density */
  double centroid[2];
  uint64_t non_finite_count;   /* particles with a NaN/Inf position or velocity */
} nbs_diagnostics;

/* Per-kernel timing accumulated since the last nbs_reset_stats (timing on). */
enum {
  NBS_KERNEL_VISCOSITY_PREDICT = 0, /* integrate + viscosity + predict + cell key + count */
  NBS_KERNEL_GRID = 1,              /* scan + scatter + rank/reorder */
  NBS_KERNEL_DENSITY = 2,           /* density + pressure */
  NBS_KERNEL_DELTA = 3,             /* relaxation + collisions + velocity */
  NBS_KERNEL_COUNT = 4
};

/* Option flags (nbs_set_options). */
enum {
  NBS_OPT_HIP_GRAPH = 1u << 0, /* replay captured steps as a HIP graph */
  NBS_OPT_CHECK_FINITE = 1u << 1 /* nbs_update returns NBS_ERR_NON_FINITE if the state holds NaN/Inf */
};

/* Demo4::ParticleSimulation ctor (demo4.cpp:14-26) + LoadScenario's SetParams
 * (app.cpp:485).  Creates a handle with one scene; capacity is a hint for the
 * particle count (storage grows when exceeded). */
nbs_status nbs_create(const nbs_params *params, const nbs_domain *domain, uint64_t capacity, nbs_sim **out);
/* Demo4::ParticleSimulation dtor (demo4.cpp:28-35). */
nbs_status nbs_destroy(nbs_sim *sim);

/* Batched simulation: adds an independent scene (its own box, grid and
 * bodies, the same parameters).  No reference counterpart: the reference
 * simulates one scene per object. */
nbs_status nbs_add_scene(nbs_sim *sim, const nbs_domain *domain, int32_t *scene_out);
nbs_status nbs_get_scene_count(nbs_sim *sim, int32_t *count);

/* BaseSimulation::SetParams / GetParams (base.h:32, 34). */
nbs_status nbs_set_params(nbs_sim *sim, const nbs_params *params);
nbs_status nbs_get_params(nbs_sim *sim, nbs_params *params);
/* BaseSimulation::SetGravity (base.h:31). */
nbs_status nbs_set_gravity(nbs_sim *sim, float gx, float gy);
/* BaseSimulation::AddExternalForces / ClearExternalForce (base.h:27-28). */
nbs_status nbs_add_external_force(nbs_sim *sim, float fx, float fy);
nbs_status nbs_clear_external_force(nbs_sim *sim);

/* BaseSimulation::ResetStats / ClearBodies / ClearParticles / ClearEmitters
 * (base.h:10-13).  ClearBodies clears the bodies of every scene. */
nbs_status nbs_reset_stats(nbs_sim *sim);
nbs_status nbs_clear_bodies(nbs_sim *sim);
nbs_status nbs_clear_particles(nbs_sim *sim);
nbs_status nbs_clear_emitters(nbs_sim *sim);

/* BaseSimulation::AddPlane / AddCircle / AddLineSegment / AddPolygon
 * (base.h:15-18).  Bodies are solved per particle in insertion order
 * (demo4.cpp:414-441).  A polygon has 3..8 vertices (sph.h:161). */
nbs_status nbs_add_plane(nbs_sim *sim, int32_t scene, float normal_x, float normal_y, float distance);
nbs_status nbs_add_circle(nbs_sim *sim, int32_t scene, float pos_x, float pos_y, float radius);
nbs_status nbs_add_line_segment(nbs_sim *sim, int32_t scene, float ax, float ay, float bx, float by);
nbs_status nbs_add_polygon(nbs_sim *sim, int32_t scene, uint64_t vertex_count, const float *verts_xy);

/* BaseSimulation::AddParticle (base.h:20), for n particles at once.  force_xy
 * may be NULL (zero force).  *first_index receives the index of the first new
 * particle. */
nbs_status nbs_add_particles(nbs_sim *sim, int32_t scene, const float *pos_xy, const float *force_xy, uint64_t n,
                             uint64_t *first_index);
/* BaseSimulation::AddVolume (base.h:21, demo4.cpp:169-181).  The jitter is
 * per-axis uniform in [-0.01h, 0.01h) from SplitMix64(jitter_seed) instead of
 * the reference's unseeded rand(). */
nbs_status nbs_add_volume(nbs_sim *sim, int32_t scene, float center_x, float center_y, float force_x, float force_y,
                          int32_t count_x, int32_t count_y, float spacing, uint64_t jitter_seed);
/* BaseSimulation::AddEmitter (base.h:22, demo4.cpp:155-167). */
nbs_status nbs_add_emitter(nbs_sim *sim, int32_t scene, float pos_x, float pos_y, float dir_x, float dir_y,
                           float radius, float speed, float rate, float duration, uint64_t jitter_seed);

/* BaseSimulation::Update (base.h:24, demo4.cpp:286-451), nsteps times. */
nbs_status nbs_update(nbs_sim *sim, float dt, uint32_t nsteps);

/* BaseSimulation::GetParticleCount (base.h:30). */
nbs_status nbs_get_particle_count(nbs_sim *sim, uint64_t *count);

/* Particle state in particle-index order (the fields of Demo4::ParticleData,
 * demo4.h:81-109, read by Render at demo4.cpp:521-531).  Any pointer may be
 * NULL.  x, prev, v, density (density, nearDensity) and pressure (pressure,
 * nearPressure) hold 2*n floats; cell_key holds n keys (y * countX + x plus
 * the scene's first cell) of the last binning; scene holds n scene indices. */
nbs_status nbs_get_state(nbs_sim *sim, float *x, float *prev, float *v, float *density, float *pressure,
                         uint32_t *cell_key, int32_t *scene);
/* Particle colours of Demo4::Render (demo4.cpp:521-524, SPHGetParticleColor
 * sph.h:683-695): rgba holds 4*n floats in particle-index order, r = clamp(
 * pressure / -10), g = clamp(density / rest_density), b = clamp(|v| / 10), a = 1. */
nbs_status nbs_get_colors(nbs_sim *sim, float *rgba);

/* Restores positions and velocities (particle-index order, 2*n floats each,
 * n = current particle count).  prev is set to x and the next step runs
 * without viscosity, as the first step after AddParticle does. */
nbs_status nbs_set_state(nbs_sim *sim, const float *x, const float *v);

/* BaseSimulation::GetStats (base.h:33). */
nbs_status nbs_get_stats(nbs_sim *sim, nbs_stats *stats);
/* Average and count of each kernel group's duration since nbs_reset_stats
 * (timing on); avg_ms and launches have NBS_KERNEL_COUNT entries. */
nbs_status nbs_get_kernel_times(nbs_sim *sim, double *avg_ms, uint64_t *launches);
/* Per-phase HIP-event timing on/off (off by default). */
nbs_status nbs_set_timing(nbs_sim *sim, int32_t enabled);
/* NBS_OPT_* flags. */
nbs_status nbs_set_options(nbs_sim *sim, uint32_t flags);

/* Kernel-shape tuning.  The pair kernels stage the positions of each
 * workgroup's candidate rows in LDS tiles of tile_cap particles (0: no staging,
 * read neighbours from HBM/L2).  With autotune on, the library times every capacity on real steps
 * (all capacities give bit-identical results) and keeps the fastest per
 * kernel, again every retune_interval steps (0: once; default 256); steps of
 * fewer than 1,024 particles use the wide kernels.  No reference
 * counterpart: the reference's only tunable is its thread count. */
typedef struct nbs_tuning {
  int32_t autotune;          /* 0: tile_cap fixed; 1: tune now, then every retune_interval steps;
                                2: as 1, with tile_cap taken as an earlier tuning's result (no first tuning;
                                   nbs_get_tuning then reports 1) */
  uint32_t retune_interval;
  int32_t tile_cap[3]; /* viscosity, density, relaxation: LDS tile capacity 0, 1024, 2048 or 4096, or
                          -1 for the wide kernels (one wave per particle); all give identical results */
  int32_t wide_max_particles; /* steps with at most this many particles use the wide kernels for all three
                                 phases whatever tile_cap says; 0 (default) leaves the choice to tile_cap */
} nbs_tuning;
nbs_status nbs_set_tuning(nbs_sim *sim, const nbs_tuning *tuning);
nbs_status nbs_get_tuning(nbs_sim *sim, nbs_tuning *tuning);

/* FP64 diagnostics over the current state. */
nbs_status nbs_get_diagnostics(nbs_sim *sim, nbs_diagnostics *out);

/* The diagnostics of nbs_get_diagnostics over caller-owned device arrays,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).
 *   x, v, density: device pointers (8-byte aligned) to n interleaved float
 *     pairs each: positions, velocities, (density, nearDensity);
 *   out: a device pointer to one nbs_diagnostics, written by the call's last
 *     kernel; the count and mean fields are over the n particles;
 *   stream: a hipStream_t to enqueue on (NULL: the null stream); the call
 *     returns after enqueueing, without synchronising.
 * The partial sums use a fixed 384 KiB device array per GPU, held until the
 * process ends, so n is at most NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES
 * (NBS_ERR_CAPACITY beyond).  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or events). */
#define NBS_DIAGNOSTICS_NOALLOC_MAX_PARTICLES 16777216u
nbs_status nbs_diagnostics_noalloc(const float *x, const float *v, const float *density, uint64_t n,
                                   float rest_density, nbs_diagnostics *out, void *stream);

/* The cell grid of the last binning: sorted_ids[n] lists particle indices in
 * cell order: by cell key, each cell in the order of its particles' slots
 * before that binning (a stable counting sort of the slot order; the slot
 * order starts in particle-index order and restarts there after
 * nbs_set_state, and particles added since the last binning follow in index
 * order), so a first binning lists each cell by ascending index, the
 * reference's order after AddParticle (demo4.cpp:46-47); cell_start[total_cells + 1]
 * holds each cell's first position in sorted_ids. */
nbs_status nbs_get_total_cells(nbs_sim *sim, uint64_t *total_cells);
nbs_status nbs_get_grid(nbs_sim *sim, uint32_t *sorted_ids, uint32_t *cell_start);

/* Multi-threading switches of BaseSimulation (base.h:35-38): the GPU is the
 * parallelism, so support is reported as 0 and the worker count as 1. */
nbs_status nbs_is_multithreading_supported(nbs_sim *sim, int32_t *supported);

/* Device arithmetic self-test: out[6*i .. 6*i+5] = sqrtf(a[i]), 1.0f/a[i],
 * a[i]/b[i], the cell coordinate of a[i] for (half = b[i], size = 1,
 * n = 1<<20), and the pair kernels' square root of a[i] and reciprocal of
 * that root (1 for a[i] == 0), both correctly rounded for 0 <= a[i] < 2^80.
 * Used to show the device's FP32 arithmetic is the correctly rounded IEEE
 * arithmetic the reference's CPU build performs. */
nbs_status nbs_selftest_math(const float *a, const float *b, uint64_t n, float *out);

/* Device copy-bandwidth probe: copies `bytes` between two device buffers
 * `reps` times and returns read+write GB/s.  No reference counterpart: it is
 * the measured HBM ceiling the step's roofline is compared with, and a kernel
 * of known traffic for calibrating the FETCH_SIZE / WRITE_SIZE counters. */
nbs_status nbs_bench_copy(uint64_t bytes, int32_t reps, double *gbps);

/* Human-readable name of a status code. */
const char *nbs_status_string(int32_t status);
/* Message of the last error on this thread (empty if none). */
const char *nbs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NBS_SPH_H */
