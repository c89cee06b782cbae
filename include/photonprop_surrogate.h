/*
 * photonprop_surrogate.h - C ABI of the learned part of libphotonprop_hip.so:
 * an MLP that predicts, for a (cascade, module) pair, the expected number of
 * photons hitting the module and a 3-component gamma mixture of their arrival
 * time residuals. It has no counterpart in the reference; it stands beside the
 * Monte Carlo (include/photonprop_hip.h), is trained only on that Monte
 * Carlo's output, and the Monte Carlo stays the fallback and the judge.
 *
 * Network: n_in -> n_layers hidden layers of n_hidden (SiLU) -> n_out. Every
 * matrix product runs on gfx950 matrix cores (v_mfma_f32_32x32x16_bf16): bf16
 * weights and activations, fp32 accumulation, fp32 master weights and Adam
 * state. Parameters are one flat fp32 vector, layer by layer: W_l (out x in,
 * row-major) then b_l.
 *
 * Outputs (n_out = 10): y0 = log of the expected hits per unit exposure,
 * exposure = photons x (R / 2r)^2 (the fraction of isotropic, unabsorbed
 * light a sphere of radius R intercepts at distance r); y1..y3 mixture
 * logits; y4..y6 log shapes; y7..y9 log scales - 4 (scale in ns). Time
 * variable x = t - t_geo + PP_SUR_TIME_SHIFT, t_geo = (r - R) / c_ref.
 *
 * Training label of a sample: exposure n_mc (R/2r)^2 of the n_mc simulated
 * photons, absorption-weighted hit sum w, and the weighted histogram of x in
 * PP_SUR_BINS bins over [0, PP_SUR_TMAX) ns. Loss per sample:
 *   (lambda - w log lambda) + sum_bins h_j * (-log p(x_j)),
 *   lambda = exposure exp(y0), p = the mixture density, x_j = bin centres;
 * averaged over the batch.
 */
#ifndef PHOTONPROP_SURROGATE_H
#define PHOTONPROP_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_SUR_N_FEATURES 8
#define PP_SUR_N_OUT 10
#define PP_SUR_BINS 64
#define PP_SUR_TMAX 2000.0f
#define PP_SUR_TIME_SHIFT 20.0f
#define PP_SUR_LABEL_FLOATS (2 + PP_SUR_BINS) /* exposure, w, hist[64] */

typedef struct pp_mlp pp_mlp;

/* One cascade of a surrogate query: vertex, unit direction, energy (GeV) and
 * log of its photon yield (the photon count an emitter of this energy
 * carries). 32 bytes. */
typedef struct pp_cascade {
    float position[3];
    float direction[3];
    float energy;
    float log_yield;
} pp_cascade;

/* Domain of the surrogate; pairs outside it are flagged for the Monte Carlo. */
typedef struct pp_sur_domain {
    float min_dist, max_dist;     /* m, vertex to module centre           */
    float min_energy, max_energy; /* GeV                                  */
    float module_radius;          /* m                                    */
    float c_ref;                  /* m/ns, for t_geo                      */
    float reserved[2];
} pp_sur_domain;

typedef struct pp_adam {
    float lr, beta1, beta2, eps;
    float weight_decay;           /* decoupled (AdamW); 0 = plain Adam    */
    float reserved[3];
} pp_adam;

/* Create a network on `device` with parameters drawn from `seed`
 * (U(-1/sqrt(fan_in), 1/sqrt(fan_in)); output layer scaled by 0.1 with
 * biases y0 = -3, log shapes log 2, log scales -1, 0.5, 2). max_batch bounds
 * the rows of one training step or inference chunk. */
int pp_mlp_create(pp_mlp** out, int device, int n_in, int n_hidden, int n_layers, int n_out,
                  int max_batch, uint64_t seed);
int pp_mlp_destroy(pp_mlp* m);
int64_t pp_mlp_param_count(pp_mlp* m);
/* fp32 master parameters (host arrays of pp_mlp_param_count floats). Setting
 * them also refreshes the bf16 copy and clears the Adam state. */
int pp_mlp_get_params(pp_mlp* m, float* params);
int pp_mlp_set_params(pp_mlp* m, const float* params);

/* Forward pass of n rows of features (host, n x n_in) to outputs (host,
 * n x n_out); n <= max_batch. */
int pp_mlp_forward(pp_mlp* m, const float* x, int n, float* y);

/* Inference kernel choice (no counterpart in the reference: a tuning and
 * test switch). fused = 1 (the default) runs every layer of a 64-row tile in
 * one kernel with the activations kept in LDS, where the network has hidden
 * width 256 and input and output widths <= 32; 0 runs one GEMM kernel per
 * layer. Both give bitwise identical outputs. */
int pp_mlp_set_fused(pp_mlp* m, int fused);

/* Loss and gradient of one batch (host features n x n_in, labels n x
 * PP_SUR_LABEL_FLOATS): *loss and grads (pp_mlp_param_count floats, host)
 * of the mean loss. Does not change the parameters. */
int pp_mlp_loss_grad(pp_mlp* m, const float* x, const float* labels, int n, float* loss,
                     float* grads);

/* One Adam step with the gradient of the last pp_mlp_loss_grad. */
int pp_mlp_adam_step(pp_mlp* m, const pp_adam* opt);

/* Training on a device-resident dataset: n samples of features and labels
 * (host arrays, uploaded once). Each step draws `batch` samples with
 * replacement (Philox on the device, key = seed), runs forward, loss,
 * backward and Adam; the step is one hipGraph replayed n_steps times.
 * loss_curve (n_steps floats, host, may be NULL) receives the batch losses. */
int pp_mlp_train(pp_mlp* m, const float* x, const float* labels, int64_t n, int batch,
                 int n_steps, const pp_adam* opt, uint64_t seed, float* loss_curve);

/* Mean loss over n samples (forward only, chunks of max_batch). */
int pp_mlp_eval_loss(pp_mlp* m, const float* x, const float* labels, int64_t n, float* loss);

/* Features of (cascade, module) pairs, the network's inputs: n x n_in host
 * array for pairs (cascades[i], modules[i]) (host arrays of n each). */
int pp_sur_features(pp_mlp* m, const pp_cascade* cascades, const float* module_xyz, int64_t n,
                    float* features);

/* Batched surrogate query over every (event, module) pair: expected hits
 * (n_events x n_modules, host), the mixture parameters (n_events x
 * n_modules x 9: weights, shapes, scales in ns; may be NULL) and a fallback
 * mask (1 where the pair is outside `dom`, e.g. closer than min_dist: the
 * caller runs the Monte Carlo for those). All on the device, in chunks of
 * max_batch pairs; *kernel_ms (may be NULL) is the device time. */
int pp_sur_eval_pairs(pp_mlp* m, const pp_cascade* events, int n_events, const float* module_xyz,
                      int n_modules, const pp_sur_domain* dom, float* expected_hits,
                      float* mixture, uint8_t* fallback, float* kernel_ms);

/* Training labels of the samples of one propagation (the sources of the
 * last pp_propagate of `ctx`, one sample each), reduced on the device from the
 * context's hit records: labels (host, n_sources x (1 + PP_SUR_BINS)) gets
 * per source the absorption-weighted hit sum w and the weighted histogram of
 * x = t - tgeo[source] + PP_SUR_TIME_SHIFT over PP_SUR_BINS bins of
 * [0, PP_SUR_TMAX) ns. The weight of a hit is exp(-dist / (abs_scale *
 * abs_len(wl))), abs_len linear in the table (abs_wl, abs_len: n_abs knots,
 * clamped at its ends). Sums are exact to 2^-32 and independent of the hit
 * order. */
typedef struct pp_ctx pp_ctx;
int pp_sur_hit_labels(pp_ctx* ctx, const double* abs_wl, const double* abs_len, int n_abs,
                      double abs_scale, const double* tgeo, int n_sources, float* labels);

#ifdef __cplusplus
}
#endif

#endif /* PHOTONPROP_SURROGATE_H */
