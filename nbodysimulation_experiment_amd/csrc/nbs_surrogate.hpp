// Learned relaxation: a pairwise-message network standing in for the
// double-density relaxation (demo4.cpp:239-255, sph.h:483-495), evaluated and
// trained with BF16 MFMA (FP32 accumulation) on gfx950.
//
// For every pair (i, j != i, r^2 < h^2) of the step's grid, with R = x_j - x_i
// (predicted positions), v = (x - prev) / dt, q = 1 - r/h:
//   f_ij = ( R.x/h, R.y/h, (v_j - v_i).x * sv, (v_j - v_i).y * sv,
//            (rho_i - mr) * sr, (rhon_i - mn) * sn, (rho_j - mr) * sr, (rhon_j - mn) * sn,
//            1, q, q^2, 0, 0, 0, 0, 0 )                         16 features
// The message net is Linear(16 -> 64), ReLU, Linear(64 -> 64); the messages
// are summed over the particle's pairs.  The second layer is linear, so the
// sum is taken before it:
//   H_i = sum_j relu(W1 f_ij)                                    (per pair, MFMA)
//   M_i = W2 H_i + n_i b2,   n_i = number of pairs               (per particle, MFMA)
// The per-particle head 64 -> 64 -> 2:
//   G_i = relu(W3 M_i + b3)                                      (MFMA)
//   o_i = W4 G_i + b4                                            (FP32)
//   dx_i = o_i * os
// and the step continues as the plain one does after the relaxation: x_i =
// collide(x1_i + dx_i), v_i = (x_i - prev_i) / dt (demo4.cpp:411-450).
//
// MFMA operands (features, W1..W3, H, M) are rounded to BF16 (round to
// nearest even); products accumulate in FP32.  Training minimises
//   L = (1/N) sum_i |o_i - dxplain_i / os|^2
// over frames of the plain method (dxplain = the plain relaxation's
// displacement before collisions) with Adam on FP32 master weights.
#pragma once

#include "nbs_kernels.hpp"

namespace nbs {

constexpr int kHid = 64;
constexpr int kFeat = 16;  // features per pair (one MFMA K step)

// Offsets into the flat FP32 parameter vector.
constexpr int kOffW1 = 0;                       // 64 x 16 (column 8 acts as the bias)
constexpr int kOffW2 = kOffW1 + kHid * kFeat;   // 64 x 64
constexpr int kOffB2 = kOffW2 + kHid * kHid;    // 64
constexpr int kOffW3 = kOffB2 + kHid;           // 64 x 64
constexpr int kOffB3 = kOffW3 + kHid * kHid;    // 64
constexpr int kOffW4 = kOffB3 + kHid;           // 2 x 64
constexpr int kOffB4 = kOffW4 + 2 * kHid;       // 2
constexpr int kNumWeights = kOffB4 + 2;         // trained parameters (9474)
// Normalisation (fixed during training): sv, mr, sr, mn, sn, os, 0, 0.
constexpr int kOffNorm = kNumWeights;
constexpr int kNumParams = kOffNorm + 8;

constexpr int kActStride = 3 * kHid;            // H, M, G per particle
constexpr int kSurRows = 32;                    // particles per wave
constexpr int kOuterRows = 512;                 // particles per outer-product workgroup

struct SurrogateBufs {
  const float *params;       // kNumParams FP32 (master)
  const uint16_t *w1b;       // 64 x 16 BF16
  const uint16_t *w2b;       // 64 x 64 BF16
  const uint16_t *w3b;       // 64 x 64 BF16
  const uint16_t *w2t, *w3t; // transposes of W2, W3 (BF16), for the backward pass
  float *out;                // N x 2 network output o (may be null)
  float *act;                // N x 192: H, M, G (training)
  float *cnt;                // N: pairs per particle (training)
  float *dH, *dG, *dM, *dO;  // N x 64, 64, 64, 2 (training)
  const float2 *target;      // N: plain relaxation displacement (training)
  float *part;               // partial sums of the weight gradient
  float *grad;               // kNumWeights
  float *lossPart;           // per head workgroup: squared error, squared target
  float *loss;               // two values per training step at 2 * lossIndex:
                             // mean squared error, relative error
  int lossIndex;
  uint32_t *pairPart;        // per forward wave: pairs processed (may be null)
};

// Evaluates the network on the step's grid.  mode flags: kFwdState writes the
// relaxation step's result (x0, v, as the plain relaxation does); kFwdTrain
// writes the training activations (out, act, cnt); out is written when set.
constexpr int kFwdState = 1;
constexpr int kFwdTrain = 2;
constexpr int kFwdPressure = 4;  // with kFwdState: also pres = (k (rho - rho0), k_near rho_near) from b.dens
void launch_surrogate_forward(const Buffers &b, const DevParams &p, const DevScene &s0, int n,
                              const SurrogateBufs &m, int mode, hipStream_t stream);
// Gradient of L for the frame of the last training forward into m.grad
// (overwritten) and the losses into m.loss[2 * m.lossIndex ...].
void launch_surrogate_backward(const Buffers &b, const DevParams &p, const DevScene &s0, int n,
                               const SurrogateBufs &m, hipStream_t stream);
// Workgroups of the forward pass (one pair count each) for n particles.
int surrogate_forward_waves(int n);
// Sizes of the partial-sum buffer (floats) and of lossPart (pairs) for n particles.
size_t surrogate_part_floats(int n);
int surrogate_head_blocks(int n);
void launch_bf16_weights(const float *params, uint16_t *w1b, uint16_t *w2b, uint16_t *w3b, uint16_t *w2t,
                         uint16_t *w3t, hipStream_t stream);
void launch_adam(float *params, const float *grad, float *m, float *v, int count, float lr, float b1, float b2,
                 float eps, int t, hipStream_t stream);
// out[0] = sum |a.x - b.x|, out[1] = sum b.x over n (FP64), partials sized
// diagnostics_blocks(n) x 2.
void launch_density_error(const float2 *a, const float2 *b, int n, double *partials, double *out,
                          hipStream_t stream);
// Per-scene FP64 sums: out[2 s] = sum |a.x - b.x|, out[2 s + 1] = sum b.x over
// the slots of scene s (found from the slot's cell key).
void launch_density_error_scenes(const float2 *a, const float2 *b, const uint32_t *keys, const uint32_t *sceneBase,
                                 int nScenes, int n, double *out, hipStream_t stream);
// C (32 x 32) = A (32 x 8) B (8 x 32), row-major FP32 inputs rounded to BF16.
void launch_mfma_probe(const float *a, const float *b, float *c, hipStream_t stream);
// C (32 x 32) = A (32 x 16) B (16 x 32), one v_mfma_f32_32x32x16_bf16.
void launch_mfma16_probe(const float *a, const float *b, float *c, hipStream_t stream);

}  // namespace nbs
