/* Learned relaxation for the MI355X SPH step library (libnbs_sph.so).
 *
 * None of these entry points has a counterpart in the reference: they add a
 * pairwise-message network that can stand in for the double-density
 * relaxation (NBodySimulation/demo4.cpp:239-255, sph.h:483-495), trained on
 * the GPU from frames of the plain method, and checked against it while it
 * runs.
 *
 * Network (BF16 MFMA operands, FP32 accumulation), per pair (i, j != i,
 * r^2 < h^2) with R = x_j - x_i, v = (x - prev) / dt, q = 1 - r/h:
 *   f_ij = (R/h, (v_j - v_i) sv, (rho_i - mr) sr, (rhon_i - mn) sn,
 *           (rho_j - mr) sr, (rhon_j - mn) sn, 1, q, q^2, 0 x 5)    16 features
 *   H_i = sum_j bf16(relu(W1 f_ij)); M_i = W2 H_i + n_i b2       message net 16-64-64
 *   G_i = relu(W3 M_i + b3);    o_i = W4 G_i + b4                head 64-64-2
 *   x_i = collide(x1_i + o_i os),  v_i = (x_i - prev_i) / dt
 * Parameter vector (FP32, NBS_SURROGATE_PARAM_COUNT floats): W1 (64 x 16,
 * row-major, rows = outputs), W2 (64 x 64), b2 (64), W3 (64 x 64), b3 (64),
 * W4 (2 x 64), b4 (2), then the normalisation sv, mr, sr, mn, sn, os, 0, 0.
 * The first NBS_SURROGATE_WEIGHT_COUNT entries are trained.
 *
 * Relaxation modes (nbs_surrogate_set_mode):
 *   NBS_RELAX_PLAIN       the plain method (default);
 *   NBS_RELAX_SURROGATE   the network relaxes every step; every
 *       check_interval-th step is a check step: the plain relaxation is
 *       computed as well from the same input, the density of both results is
 *       evaluated on the step's grid, and the density error
 *       sum_i |rho_i(surrogate) - rho_i(plain)| / sum_i rho_i(plain)
 *       (FP64), in total and per scene, is compared with
 *       max_density_error.  If any scene's error is above it, the plain
 *       result replaces the surrogate's for that step (in every scene) and
 *       the next check_interval steps run the plain method (fallback).
 *   NBS_RELAX_SURROGATE_CARRIED_DENSITY   as NBS_RELAX_SURROGATE, but the
 *       network steps between checks run no density pass: their features take
 *       each particle's density in the result the last check accepted,
 *       carried with the particle through the binning.  Between checks the
 *       state's density is that carried density, its pressure is computed
 *       from it (sph.h:478-481), and the candidate-count statistics are those
 *       of the last density pass.  Steps that change the particle set or the
 *       state, and plain steps, drop the carried density; the next check
 *       sets it again.
 */
#ifndef NBS_SURROGATE_H
#define NBS_SURROGATE_H

#include "nbs_sph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NBS_SURROGATE_WEIGHT_COUNT 9474
#define NBS_SURROGATE_PARAM_COUNT 9482

enum { NBS_RELAX_PLAIN = 0, NBS_RELAX_SURROGATE = 1, NBS_RELAX_SURROGATE_CARRIED_DENSITY = 2 };

typedef struct nbs_surrogate_report {
  uint64_t surrogate_steps;   /* steps relaxed by the network (check steps that passed included) */
  uint64_t plain_steps;       /* steps relaxed by the plain method while in surrogate mode */
  uint64_t check_steps;
  uint64_t fallbacks;         /* checks with a scene above max_density_error */
  uint64_t train_steps;       /* Adam updates since nbs_surrogate_init */
  double last_density_error;
  double max_density_error;
  double mean_density_error;  /* over the checks since the last nbs_surrogate_set_mode */
  uint64_t last_pair_count;   /* in-range pairs the last network evaluation processed */
  double last_max_scene_error;     /* largest per-scene density error of the last check */
  uint64_t last_scenes_over_limit; /* scenes above max_density_error at the last check */
} nbs_surrogate_report;

/* Seeded initial weights (He-uniform, zero biases) and neutral normalisation;
 * resets the Adam state and the report. */
nbs_status nbs_surrogate_init(nbs_sim *sim, uint64_t seed);
/* Parameter vector in and out (count must be NBS_SURROGATE_PARAM_COUNT). */
nbs_status nbs_surrogate_set_params(nbs_sim *sim, const float *params, uint64_t count);
nbs_status nbs_surrogate_get_params(nbs_sim *sim, float *params, uint64_t count);
/* Runs one plain step and sets the normalisation from its frame: mean and
 * 1/std of density and near density, 1/RMS of velocity, RMS of the plain
 * relaxation displacement. */
nbs_status nbs_surrogate_calibrate(nbs_sim *sim, float dt);
/* Runs `steps` plain steps; after each, one Adam step (learning_rate,
 * betas 0.9 / 0.999, eps 1e-8) on that frame.  losses[steps] and
 * relative_losses[steps] (either may be NULL) receive each frame's loss
 * before its update: the mean squared error, and the squared error relative
 * to the squared normalised target, sum |o - t/os|^2 / sum |t/os|^2.
 * Calibrates first if the normalisation was never set. */
nbs_status nbs_surrogate_train(nbs_sim *sim, float dt, uint32_t steps, float learning_rate, float *losses,
                               float *relative_losses);
nbs_status nbs_surrogate_set_mode(nbs_sim *sim, int32_t mode, uint32_t check_interval, float max_density_error);
/* In-run training: with learning_rate > 0, every check step in surrogate mode
 * also takes one Adam step on its own frame, whose plain relaxation the check
 * computes anyway (0 turns it off; the normalisation must be set). */
nbs_status nbs_surrogate_set_online_training(nbs_sim *sim, float learning_rate);
/* Per-scene density errors of the last check step, errors[count] with count
 * = the scene count: sum over the scene of |rho(surrogate) - rho(plain)| /
 * sum of rho(plain) (FP64).  Zeros before the first check. */
nbs_status nbs_surrogate_get_scene_errors(nbs_sim *sim, double *errors, uint64_t count);
nbs_status nbs_surrogate_get_report(nbs_sim *sim, nbs_surrogate_report *report);
/* Runs one plain step and, without updating the weights, returns the frame
 * the network saw and its gradient: grad[NBS_SURROGATE_WEIGHT_COUNT], loss,
 * and per slot of the step's cell order (n entries each; any may be NULL):
 * out (network output o, 2 floats), x (predicted positions, 2), prev (2),
 * density (density and near density, 2), target (plain relaxation
 * displacement, 2), scene (int32). */
nbs_status nbs_surrogate_frame(nbs_sim *sim, float dt, float *grad, float *loss, float *out, float *x, float *prev,
                               float *density, float *target, int32_t *scene);

/* MFMA self-tests: c (32 x 32) = a (32 x K) b (K x 32), row-major FP32 inputs
 * rounded to BF16 and multiplied by one v_mfma_f32_32x32x8_bf16 (K = 8) or
 * one v_mfma_f32_32x32x16_bf16 (K = 16, the form the network's kernels use). */
nbs_status nbs_selftest_mfma(const float *a, const float *b, float *c);
nbs_status nbs_selftest_mfma16(const float *a, const float *b, float *c);

#ifdef __cplusplus
}
#endif

#endif /* NBS_SURROGATE_H */
