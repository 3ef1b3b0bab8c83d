/* velo_ghost.h -- learned ghost-track filter for consolidated VELO tracks.
 *
 * No counterpart in the reference (its VELO output keeps every track; the
 * dev_accepted_velo_tracks flags of ConsolidateVeloVisitor.cu:42-43 are all
 * set).  A 10 -> 32 -> 32 -> 1 perceptron (ReLU, ReLU, sigmoid) with FP16
 * weights and FP32 accumulation scores each track's probability of being a
 * ghost from ten features of its hits and fitted state; consolidation can
 * run it inline and clear dev_accepted_velo_tracks for tracks scored at or
 * above the model's threshold.  The plain output (hits, states, counts) is
 * never changed by the filter.
 *
 * Features (velo_track_features_launch, and inline in consolidation):
 *   0 number of hits
 *   1 first (lowest) module number, 2 last (highest) module number
 *   3 tx, 4 ty                      (fitted state)
 *   5 chi2 / ndof of the straight-line fit (state slot 10)
 *   6 mean, 7 max transverse distance of a hit from the fitted line (mm)
 *   8 module pairs between the first and last hit without a hit
 *   9 phi spread of the three seed hits (largest pairwise |dphi|, rad)
 * Parameter order (VELO_GHOST_PARAMS values): W1[32][10], b1[32],
 * W2[32][32], b2[32], w3[32], b3.
 */
#ifndef VELO_GHOST_H
#define VELO_GHOST_H

#include <stddef.h>
#include <stdint.h>

#include "velo_hlt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VELO_GHOST_FEATURES 10
#define VELO_GHOST_HIDDEN 32
#define VELO_GHOST_PARAMS 1441

typedef struct {
  float mean[VELO_GHOST_FEATURES];  /* input normalisation: (x - mean) * scale */
  float scale[VELO_GHOST_FEATURES];
  float threshold;                  /* a track is accepted when p(ghost) < threshold */
  uint32_t reserved;
  uint16_t weights[VELO_GHOST_PARAMS + 1]; /* IEEE binary16 bits; the last entry is padding */
} velo_ghost_model;

/* Features of every consolidated track (one workgroup per event, or
 * persistent).  dev_features: total_tracks x 10 floats, row-major. */
typedef struct {
  const int32_t* dev_atomics_velo;
  const uint32_t* dev_velo_track_hit_number;
  const char* dev_velo_track_hits;
  const char* dev_velo_states;
  float* dev_features;
  uint32_t n_events;
  velo_tuning tuning;
} velo_track_features_args;
int velo_track_features_launch(const velo_track_features_args* a, void* stream);

/* p(ghost) of n feature rows. */
typedef struct {
  const float* dev_features;
  const velo_ghost_model* dev_model;
  float* dev_score;
  uint32_t n;
  velo_tuning tuning;
} velo_ghost_infer_args;
int velo_ghost_infer_launch(const velo_ghost_infer_args* a, void* stream);

/* Gradient of the mean weighted binary cross-entropy over a minibatch
 *   L = -(1/B) sum_s w_s [y_s log p_s + (1 - y_s) log(1 - p_s)],
 *   w_s = pos_weight for ghosts (y = 1), 1 otherwise,
 * with respect to the model's parameters, evaluated with the FP16 weights
 * of dev_model.  Per-tile partial sums go to dev_partials (
 * velo_ghost_partials_size(batch_size) floats), then are summed in a fixed
 * order into dev_grad[0, VELO_GHOST_PARAMS) with the loss at
 * dev_grad[VELO_GHOST_PARAMS]: the result does not depend on the launch. */
typedef struct {
  const float* dev_features;
  const float* dev_labels;    /* 1 = ghost, 0 = not */
  const uint32_t* dev_batch;  /* sample indices of the minibatch */
  const velo_ghost_model* dev_model;
  float* dev_partials;
  float* dev_grad;
  uint32_t batch_size;
  float pos_weight;
  velo_tuning tuning;
} velo_ghost_grad_args;
size_t velo_ghost_partials_size(uint32_t batch_size);
int velo_ghost_grad_launch(const velo_ghost_grad_args* a, void* stream);

/* One Adam step on the FP32 master parameters, then the FP16 copy in
 * dev_model->weights is refreshed (round to nearest even). */
typedef struct {
  const float* dev_grad;
  float* dev_master;
  float* dev_m;
  float* dev_v;
  velo_ghost_model* dev_model;
  float lr, beta1, beta2, eps;
  uint32_t step; /* 1 for the first step (bias correction) */
} velo_ghost_adam_args;
int velo_ghost_adam_launch(const velo_ghost_adam_args* a, void* stream);

#ifdef __cplusplus
}
#endif

#endif
