// Shared host/device declarations of the MI355X SPH step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/nbs_sph.h"

namespace nbs {

constexpr int kBlock = 256;          // threads per workgroup of the particle kernels
constexpr int kScanItems = 4;        // cells per thread in the scan kernels (1,024 per workgroup)
constexpr int kScanChunk = kBlock * kScanItems;
// One-workgroup cell scan for small grids (k_scan_small).
constexpr int kSmallGridBlock = 1024;
constexpr uint32_t kSmallGridCells = 16384;
constexpr int kSmallCountParticles = 16384;  // most particles counted by the one-workgroup scan

// Step constants, derived on the host in FP32 exactly as the reference derives
// them (sph.h:92 invKernelHeight, sph.h:470 h*h, sph.h:492 dt*dt,
// demo4.cpp:287 invDt, demo4.cpp:306 gravity + externalForce).
struct DevParams {
  float h, invH, h2, rho0, k, kn, sigma, beta, R, margin;
  float dt, dt2, invDt;
  float gex, gey;    // gravity + external force
  float gdtx, gdty;  // ((gravity + external) + 0) * dt: velocity increment of a particle without own force
};

// One scene: an origin-centred box and its grid.  Keys of the scene's cells
// are cellBase + cy * nx + cx (sph.h:444-448).
struct DevScene {
  float halfW, halfH, cellSize;
  int nx, ny;
  uint32_t cellBase;
  int bodyBegin, bodyEnd;
};

enum BodyType : int { BODY_PLANE = 1, BODY_CIRCLE = 2, BODY_SEGMENT = 3, BODY_POLYGON = 4 };

// plane: (a0, a1) normal, a2 distance; circle: (a0, a1) centre, a2 radius;
// segment: (a0, a1)-(a2, a3); polygon: nverts vertices at verts[vertOffset].
struct DevBody {
  int type;
  int nverts;
  int vertOffset;
  int pad;
  float a0, a1, a2, a3;
  // Circles: float_order images of a box outside which no position can pass
  // the circle's collision test (x0, y0, x1, y1); full range for other bodies.
  int cull[4];
};
// body_candidates reads a0..a3 and cull as one 16-byte load each
static_assert(sizeof(DevBody) == 48 && offsetof(DevBody, a0) == 16 && offsetof(DevBody, cull) == 32, "DevBody layout");

// Order-preserving integer image of a float: for non-NaN a, b,
// a < b implies float_order(a) < float_order(b) (and -0 maps below +0).
__host__ __device__ inline int float_order(int bits) { return bits ^ ((bits >> 31) & 0x7fffffff); }

// Device storage.  Particle arrays are indexed by slot: slots [0, nBinned)
// hold the particles of the last binning sorted by cell key, each cell in the
// order of the particles' slots before that binning (a stable counting sort
// of the slot order); slots [nBinned, n) hold particles added since, in index
// order.
struct Buffers {
  float2 *x0;    // positions (after the last step)
  float2 *xp;    // predicted positions, slot order of the previous binning
  float2 *x1;    // predicted positions, slot order of the new binning
  float2 *prev;  // positions before predict (new binning order)
  float2 *v;     // velocities
  float2 *dens;  // (density, nearDensity)
  float2 *pres;  // (pressure, nearPressure)
  float2 *acc;   // own acceleration of particles not yet stepped
  uint32_t *id, *id2;  // particle index per slot (current / next)
  uint32_t *keys;      // cell key of the last binning per slot
  uint32_t *keyp;      // cell key of the predicted position (previous slot order)
  uint32_t *fill;      // arrival index within the new cell
  uint32_t *tmpSlot;   // source slots bucketed by new cell
  uint4 *binRec;       // per new slot of a one-workgroup binning: source slot, key, cell start, cell end
  uint32_t *scn;       // scene of particles not yet binned
  uint32_t *count;     // per cell, zero between steps (whole scan chunks; past the last cell always zero)
  uint32_t *startCur;  // per cell + 1: the last binning
  uint32_t *startNext; // per cell + 1: the binning being built
  uint32_t *blockSum;  // per scan workgroup: its cells' total
  uint32_t *candStats; // per density workgroup: min, max candidate count
  uint32_t *cellStats; // per scan workgroup: min, max occupied-cell count
  const DevScene *scenes;
  const uint32_t *sceneBase;  // cellBase of each scene, ascending
  const uint16_t *cellScene;  // scene of each cell (multi-scene steps), or null: search sceneBase
  int nScenes;
  const DevBody *bodies;
  const float2 *verts;
  float2 *dxOut;       // if set, the relaxation kernel also writes each particle's
                       // relaxation displacement (before collisions) here
  // Placement mode of the grid phase (launch_grid with place = true): the
  // viscosity kernel flags (bit 31 of keyp) and lists here the particles whose
  // new cell is not 3x3-adjacent to their cell of the last binning.
  uint32_t *farList;
  uint32_t *farCount;       // entries of farList this step (zero when the step starts)
  uint32_t *farCountClear;  // the other step parity's counter: cleared by the cell scan
  uint32_t *chunkCells;     // per kPlaceChunk new slots: the cells holding its first and last slot
  // If set, the binning also moves densFrom (previous slot order) into densTo
  // (new slot order): a per-particle value carried along with its particle.
  const float2 *densFrom;
  float2 *densTo;
};

constexpr uint32_t kPlaceChunk = 1024;  // new slots per placement workgroup
constexpr uint32_t kPlaceCap = 2048;    // LDS bucket entries of a placement workgroup

// Launchers (nbs_kernels.hip).  All kernels are enqueued on `stream`.
void launch_unbin(const Buffers &b, int nBinned, hipStream_t stream);
// cap: particle positions staged in LDS per workgroup (0, 1024, 2048 or 4096;
// 0 reads neighbours from global memory).  Every choice gives bit-identical
// results.
// cells: the grid's cell count (the counting may move into the grid phase,
// count_in_grid).
void launch_visc_predict(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int nBinned, int lists,
                         int cap, uint32_t cells, hipStream_t stream);
// Whether the cell counting of a step runs in the grid phase's one-workgroup
// scan (LDS histogram), which then also buckets the particles, instead of
// after the viscosity kernel: wide viscosity (viscCap == kWideCap), at most
// kSmallGridCells cells and kSmallCountParticles particles.
bool count_in_grid(int n, uint32_t cells, int viscCap);
// n: particles (with b.chunkCells set, the scan also fills the chunk table).
void launch_scan(const Buffers &b, uint32_t cells, int n, hipStream_t stream);
void launch_scatter_reorder(const Buffers &b, int n, hipStream_t stream);
// The grid phase: the cell scan (one workgroup for small grids, else
// launch_scan), then launch_scatter_reorder; with count_in_grid the one
// workgroup counts, scans and buckets, and only the rank pass follows (from
// its records, b.binRec); with place = true the cell scan and k_place
// (b.farList etc. set, see place_binning).  Returns the number of
// (min, max) occupied-cell pairs it leaves in cellStats.
uint32_t launch_grid(const Buffers &b, const DevScene &s0, uint32_t cells, int n, int viscCap, bool place,
                     hipStream_t stream);
// Whether a step can bin by placement (k_place): a grid too large for the
// one-workgroup scan, the one-lane-per-particle viscosity kernel (which flags
// the far movers), and every particle binned by the last step.
bool place_binning(uint32_t cells, int viscCap, bool allBinned);
void launch_density(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int cap, hipStream_t stream);
// Workgroups of the density launch for n particles at capacity cap: the number
// of (min, max) candidate-count pairs it leaves in candStats.
uint32_t density_stat_slots(int n, int cap);
void launch_delta(const Buffers &b, const DevParams &p, const DevScene &s0, int n, int cap, hipStream_t stream);
// cap value selecting the wide kernels (one wave per particle) of the three
// pair launchers; results are bit-identical to the other capacities.
constexpr int kWideCap = -1;
// The variants the autotuner orders and times (the wide kernels last).
constexpr int kNumTileCaps = 5;
constexpr int kTileCaps[kNumTileCaps] = {0, 1024, 2048, 4096, kWideCap};
void launch_gather_state(const Buffers &b, const DevScene &s0, int n, int nBinned, float2 *x, float2 *prev, float2 *v,
                         float2 *dens, float2 *pres, uint32_t *key, int32_t *scene, hipStream_t stream);
void launch_colors(const Buffers &b, float rho0, int n, float4 *rgba, hipStream_t stream);
void launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t stream);
// FP64 diagnostics.  The sums (kDiagWidth doubles: sum vx, sum vy, sum |v|^2,
// sum rho, max rho, sum |rho - rho0|, sum (rho - rho0)^2, sum x, sum y,
// non-finite count, finite count, 0) are reduced per workgroup of
// kDiagPerBlock slots into partials[diagnostics_blocks(n) * kDiagWidth], then
// combined in workgroup order.
constexpr int kDiagWidth = 12;
constexpr int kDiagPerBlock = kBlock * 16;
void launch_diagnostics(const Buffers &b, const DevParams &p, int n, double *partials, double *out,
                        hipStream_t stream);
int diagnostics_blocks(int n);
// The same reduction over caller-owned SoA arrays, finished on the device into
// one nbs_diagnostics record at `out` (device memory); n particles.
void launch_diagnostics_record(const float2 *x, const float2 *v, const float2 *dens, int n, float rho0,
                               double *partials, nbs_diagnostics *out, hipStream_t stream);

// The nbs_diagnostics record of n particles from the combined sums r.
__host__ __device__ inline void diagnostics_record(const double *r, uint64_t n, double rho0, nbs_diagnostics *d) {
  const double cnt = r[10] > 0 ? r[10] : 1.0;
  d->particle_count = n;
  d->momentum[0] = r[0];
  d->momentum[1] = r[1];
  d->kinetic_energy = 0.5 * r[2];
  d->mean_density = r[3] / cnt;
  d->max_density = n ? r[4] : 0.0;
  d->mean_abs_density_error = rho0 != 0.0 ? (r[5] / cnt) / fabs(rho0) : 0.0;
  d->rms_density_error = rho0 != 0.0 ? sqrt(r[6] / cnt) / fabs(rho0) : 0.0;
  d->centroid[0] = r[7] / cnt;
  d->centroid[1] = r[8] / cnt;
  d->non_finite_count = (uint64_t)r[9];
}
void launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t stream);
void launch_selftest_math(const float *a, const float *b, uint64_t n, float *out, hipStream_t stream);

}  // namespace nbs
