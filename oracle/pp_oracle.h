/*
 * pp_oracle.h - CPU oracle of the photon-propagation hot path.
 * TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline).
 * Input PODs (pp_source, pp_medium, pp_target) and the random stream layout
 * constants are the library's boundary format, taken from its public header.
 */
#ifndef PP_ORACLE_H
#define PP_ORACLE_H

#include <stdint.h>

#include "../include/photonprop_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PPO_ABI_VERSION 1

typedef struct {
    float mono_wl;      /* used when table == NULL */
    const float* table; /* inverse CDF knots       */
    int n;
    int mapping;        /* enum pp_spectrum_mapping */
} ppo_spectrum;

typedef struct {
    float pos[3];
    float dir[3];
    float time;
    float wavelength;
    float sca_len;
    float c_grp;
} ppo_photon;

typedef struct {
    float pos[3];
    float dir[3];
    float init_dir[3];
    float wavelength;
    float time;
    float dist;
    int64_t module_ix; /* 1-based */
    int32_t n_steps;
    int32_t source_ix;
    uint64_t photon_id;
} ppo_hit; /* 72 bytes */

int ppo_abi_version(void);
int ppo_eval_fn(int fn, const pp_medium* m, const ppo_spectrum* sp, const float* in,
                int in_stride, float* out, int out_stride, int64_t n);
void ppo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float ppo_u01(uint32_t x);
float ppo_u01_open0(uint32_t x);
void ppo_sph_to_cart(float theta, float phi, float out[3]);
void ppo_isotropic_dir(float u1, float u2, float out[3]);
void ppo_rot_from_ez(const float a[3], const float b[3], float out[3]);
float ppo_hg_cos(float g, float u);
float ppo_es_cos(float b, float u);
float ppo_sl_cos(float g, float u);
float ppo_mixed_hges_cos(float g, float b, float hg_fraction, float u_choice, float u);
void ppo_update_direction(const float dir[3], float cos_sca_theta, float u_phi, float out[3]);
int ppo_intersect_sphere(const float c[3], float r, const float pos[3], const float dir[3],
                         float step, float* d_out);
int ppo_intersect_plane(int kind, const float c[3], float a, float b, const float pos[3],
                        const float dir[3], float step, float* d_out);
float ppo_step_length(float eta, float sca_len);
float ppo_kopelevich_sca_len(float wl, float vs, float vl, float sca_scale);
double ppo_quan_fry_n(double wl, double S, double T, double P);
double ppo_quan_fry_group_velocity(double wl, double S, double T, double P);
void ppo_medium_values(const pp_medium* m, float wl, float* sca_len, float* c_grp,
                       float* n_phase);
float ppo_phase_cos(const pp_medium* m, float u_choice, float u);
void ppo_cherenkov_track_dir(float u1, float u2, float out[3]);
float ppo_spectrum_sample(const ppo_spectrum* sp, float u);
float ppo_gamma_sample(float shape, float scale, uint64_t seed, uint64_t pid, uint32_t src_ix);
float ppo_beta_sample(float a, float b, uint64_t seed, uint64_t pid, uint32_t src_ix);
void ppo_init_photon(const pp_source* src, const pp_medium* m, const ppo_spectrum* sp,
                     const uint32_t key[2], uint32_t src_ix, uint64_t pid, ppo_photon* st);
int ppo_transport(ppo_photon* st, const pp_medium* m, const pp_target* targets, int n_targets,
                  const uint32_t key[2], uint32_t src_ix, uint64_t pid, int32_t n_steps,
                  ppo_hit* hit);
int64_t ppo_photon_count(int64_t photons, float photon_scaling);
int64_t ppo_propagate(const pp_source* sources, int n_sources, const pp_medium* m,
                      const ppo_spectrum* sp, const pp_target* targets, int n_targets,
                      uint64_t seed, int32_t n_steps, float photon_scaling, ppo_hit* hits,
                      int64_t capacity, int64_t* n_ph_sim, int n_threads);

/* Float32 port of the reference kernel (pp_ref_f32.c, the CPU baseline):
 * same arguments; hits unordered; *n_steps_done = photon steps executed;
 * returns -2 for a source, medium or target kind the port does not cover. */
int64_t ppf_propagate(const pp_source* sources, int n_sources, const pp_medium* m,
                      const ppo_spectrum* sp, const pp_target* targets, int n_targets,
                      uint64_t seed, int32_t n_steps, float photon_scaling, ppo_hit* hits,
                      int64_t capacity, int64_t* n_ph_sim, int64_t* n_steps_done,
                      int n_threads);

#ifdef __cplusplus
}
#endif

#endif
