/*
 * pp_oracle.c - CPU oracle of the photon-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain C, the per-photon
 * algorithm of PhotonPropagation.jl's CUDA.jl kernel so that the HIP library
 * can be checked against it. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the shipped library never does.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Each function follows the reference line it cites, in the reference's
 * operation order, in IEEE single precision without contraction (Julia does
 * not fuse a*b+c unless asked). Functions of third-party packages that are
 * not in the reference tree (CherenkovMediumBase 1.0.0, PhysicsTools 0.8.2)
 * are restated from their published algorithms; see the comment at each.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   philox          Random123 known-answer vectors + rocRAND host engine
 *   kopelevich      reference-held value 56.86782 m @ 488 nm (axicone.ipynb)
 *   intersect       closed-form hit point on the sphere, inside -> no hit,
 *                   brute-force first-hit order, no-scatter hit fraction
 *   plane           crossing point on the plane, inside/outside the shape by
 *                   construction, beams of known geometry through a disc
 *   isotropic       cos(theta) ~ U(-1,1), phi ~ U(0,2pi) (KS), unit norm
 *   hg / es / mixed moments and per-bin chi2 vs the analytic pdf
 *   rot_from_ez     R e_z = a, orthonormality, dot(R b, a) = b_z
 *   gamma / beta    KS vs scipy.stats at fixed parameters
 *   track_dir       KS of 1-cos(theta) vs the truncated Weibull CDF
 *   step length     KS vs Exp(sca_len)
 * Parity unpinned (no reference value exists for these, only invariants):
 *   quan_fry n / group velocity / Cherenkov angle (pressure term unknown),
 *   the exact conventions of rot_from_ez_fast, rand_gamma and
 *   sample_cherenkov_track_direction inside PhysicsTools / CherenkovMediumBase.
 */
#include "pp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11; Random123 constants)           */
/* ------------------------------------------------------------------ */
void ppo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float ppo_u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
float ppo_u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

/* One photon's draw block: counter (word, photon lo, photon hi, source). */
static void ppo_block(uint32_t word, uint64_t pid, uint32_t src, const uint32_t key[2],
                      uint32_t out[4]) {
    uint32_t ctr[4] = {word, (uint32_t)pid, (uint32_t)(pid >> 32), src};
    ppo_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------ */
/* Step functions (pure, explicit uniforms)                            */
/*                                                                     */
/* Each function takes the Float32 values the reference holds (inputs, */
/* uniforms), evaluates its mathematics in double precision and rounds */
/* its Float32 outputs once, so the oracle sits within 0.5 ulp of the  */
/* exact value of every step function.                                 */
/* ------------------------------------------------------------------ */
#define PPO_TWO_PI_F ((float)(2.0 * 3.14159265358979323846)) /* T(2*pi) */
#define PPO_TWO_PI_D (2.0 * 3.14159265358979323846)
#define PPO_C_VAC_M_NS 0.299792458 /* src/photon_prop_cuda.jl:37 */

/* uniform(min, max) = min + u*(max-min)   src/photon_prop_cuda.jl:48-51 */
static float ppo_uniform(float lo, float hi, float u) { return lo + u * (hi - lo); }

/* phi = uniform(0, T(2pi)) as a Float32 variable */
static float ppo_phi(float u) { return ppo_uniform(0.0f, PPO_TWO_PI_F, u); }

static void sph_to_cart_d(double ct, double st, double phi, float out[3]) {
    out[0] = (float)(cos(phi) * st);
    out[1] = (float)(sin(phi) * st);
    out[2] = (float)ct;
}

/* PhysicsTools.sph_to_cart(theta, phi) = (sin t cos p, sin t sin p, cos t) */
void ppo_sph_to_cart(float theta, float phi, float out[3]) {
    sph_to_cart_d(cos((double)theta), sin((double)theta), (double)phi, out);
}

/* initialize_direction_isotropic   src/photon_prop_cuda.jl:73-77:
 * theta = acos(uniform(-1, 1)), phi = uniform(0, 2pi) */
void ppo_isotropic_dir(float u1, float u2, float out[3]) {
    double theta = acos((double)ppo_uniform(-1.0f, 1.0f, u1));
    sph_to_cart_d(cos(theta), sin(theta), (double)ppo_phi(u2), out);
}

/* PhysicsTools.rot_from_ez_fast(a, b): rotate b by the rotation that takes
 * e_z to a. Restated as the minimal (Rodrigues) rotation about
 * k = e_z x a / |e_z x a| by theta = angle(e_z, a); a = -e_z turns by pi
 * about x. Convention unpinned (see file header). */
static void rot_from_ez_d(const double a[3], const double b[3], double out[3]) {
    double s = sqrt(a[0] * a[0] + a[1] * a[1]);
    double c = a[2];
    if (s == 0.0) {
        out[0] = b[0];
        out[1] = c > 0.0 ? b[1] : -b[1];
        out[2] = c > 0.0 ? b[2] : -b[2];
        return;
    }
    double kx = -a[1] / s, ky = a[0] / s; /* kz = 0 */
    double kxb0 = ky * b[2], kxb1 = -kx * b[2], kxb2 = kx * b[1] - ky * b[0];
    double kdb = kx * b[0] + ky * b[1];
    double omc = 1.0 - c;
    out[0] = b[0] * c + kxb0 * s + kx * kdb * omc;
    out[1] = b[1] * c + kxb1 * s + ky * kdb * omc;
    out[2] = b[2] * c + kxb2 * s;
}

void ppo_rot_from_ez(const float a[3], const float b[3], float out[3]) {
    double ad[3] = {a[0], a[1], a[2]}, bd[3] = {b[0], b[1], b[2]}, o[3];
    rot_from_ez_d(ad, bd, o);
    out[0] = (float)o[0]; out[1] = (float)o[1]; out[2] = (float)o[2];
}

static double clamp1d(double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); }

/* CherenkovMediumBase hg_scattering_func(g): inverse CDF of Henyey-Greenstein,
 * cos = (1 + g^2 - ((1 - g^2)/(1 + g(2u - 1)))^2) / (2g), clamped to [-1, 1]. */
static double hg_cos_d(double g, double u) {
    double t = (1.0 - g * g) / (1.0 + g * (2.0 * u - 1.0));
    return clamp1d((1.0 + g * g - t * t) / (2.0 * g));
}
float ppo_hg_cos(float g, float u) { return (float)hg_cos_d(g, u); }

/* Einstein-Smoluchowski phase function p(x) ~ 1 + b x^2 on [-1, 1]: solve
 * CDF(x) = u, i.e. (b/3)x^3 + x + (1+b/3) - u(2+2b/3) = 0, by Cardano's
 * formula for the depressed cubic x^3 + p x + q = 0 (one real root, p > 0),
 * clamped to [-1, 1]. */
static double es_cos_d(double b, double u) {
    double p = 3.0 / b;
    double q = (3.0 / b) * ((1.0 + b / 3.0) - u * (2.0 + 2.0 * b / 3.0));
    double disc = sqrt(q * q / 4.0 + p * p * p / 27.0);
    return clamp1d(cbrt(-q / 2.0 + disc) + cbrt(-q / 2.0 - disc));
}
float ppo_es_cos(float b, float u) { return (float)es_cos_d(b, u); }

/* Simplified Liu phase function of PPC (mixed_hg_sl_scattering_func_ppc of
 * CherenkovMediumBase, restated from PPC's published sampler):
 * p(cos) ~ (1 + cos)^a, a = 2g/(1-g), whose mean is g; inverse CDF
 * cos = 2 u^(1/(a+1)) - 1 = 2 u^((1-g)/(1+g)) - 1. */
static double sl_cos_d(double g, double u) {
    return clamp1d(2.0 * pow(u, (1.0 - g) / (1.0 + g)) - 1.0);
}
float ppo_sl_cos(float g, float u) { return (float)sl_cos_d(g, u); }

/* MixedHGES(g, b, hg_fraction): HG with probability hg_fraction, else ES.
 * Draw order: branch choice, then the cosine. */
float ppo_mixed_hges_cos(float g, float b, float hg_fraction, float u_choice, float u) {
    if (u_choice < hg_fraction) return ppo_hg_cos(g, u);
    return ppo_es_cos(b, u);
}

/* update_direction   src/photon_prop_cuda.jl:280-298 (cos theta given):
 * new direction (cos p sin t, sin p sin t, cos t) about e_z, rotated to dir */
void ppo_update_direction(const float dir[3], float cos_sca_theta, float u_phi, float out[3]) {
    double ct = cos_sca_theta;
    double st = sqrt(1.0 - ct * ct);
    double phi = ppo_phi(u_phi);
    double nd[3] = {cos(phi) * st, sin(phi) * st, ct};
    double a[3] = {dir[0], dir[1], dir[2]}, o[3];
    rot_from_ez_d(a, nd, o);
    out[0] = (float)o[0]; out[1] = (float)o[1]; out[2] = (float)o[2];
}

/* _intersect_nsphere + check_intersection(::Spherical)
 *   src/photon_prop_cuda.jl:309-351. Near root d = -a - sqrt(b) of the ray,
 *   hit iff b >= 0 and 0 < d < step. Returns 1 and *d on hit. */
int ppo_intersect_sphere(const float c[3], float r, const float pos[3], const float dir[3],
                         float step, float* d_out) {
    double target_rsq = (double)(r * r); /* radius^2 is a Float32 value */
    double dp0 = (double)pos[0] - c[0], dp1 = (double)pos[1] - c[1], dp2 = (double)pos[2] - c[2];
    double a = dir[0] * dp0 + dir[1] * dp1 + dir[2] * dp2;
    double pp = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
    double b = a * a - (pp - target_rsq);
    double d = b >= 0.0 ? (-a - sqrt(fabs(b))) : -1.0;
    if ((d > 0.0) && (d < (double)step)) {
        *d_out = (float)d;
        return 1;
    }
    return 0;
}

/* check_intersection(::Rectangular / ::Circular)   src/photon_prop_cuda.jl:353-404.
 *   No hit when dir_z == 0; d = (z_t - z) / dir_z, miss if d < 0 or d > step;
 *   hit iff the crossing point is strictly inside the rectangle
 *   (|x - x_t| < length_x/2, |y - y_t| < length_y/2) or the disc
 *   ((x - x_t)^2 + (y - y_t)^2 < r^2, r^2 a Float32 value). The crossing is
 *   evaluated exactly; d is rounded once. a = length_x or radius, b = length_y. */
int ppo_intersect_plane(int kind, const float c[3], float a, float b, const float pos[3],
                        const float dir[3], float step, float* d_out) {
    if (dir[2] == 0.0f) return 0;
    double d = ((double)c[2] - pos[2]) / dir[2];
    if (d < 0.0 || d > (double)step) return 0;
    double x = (double)pos[0] + (double)dir[0] * d - c[0];
    double y = (double)pos[1] + (double)dir[1] * d - c[1];
    int in;
    if (kind == PP_SHAPE_RECTANGLE)
        in = fabs(x) < (double)a / 2.0 && fabs(y) < (double)b / 2.0;
    else
        in = x * x + y * y < (double)(a * a);
    if (!in) return 0;
    *d_out = (float)d;
    return 1;
}

/* one target of any shape: check_intersection   src/photon_prop_cuda.jl:345-404 */
static int ppo_intersect_target(const pp_target* t, const float pos[3], const float dir[3],
                                float step, float* d_out) {
    if (t->shape_kind == PP_SHAPE_SPHERE)
        return ppo_intersect_sphere(t->position, t->radius, pos, dir, step, d_out);
    float a = t->shape_kind == PP_SHAPE_RECTANGLE ? t->length[0] : t->radius;
    return ppo_intersect_plane(t->shape_kind, t->position, a, t->length[1], pos, dir, step,
                               d_out);
}

/* step_size = -log(eta) * sca_len   src/photon_prop_cuda.jl:488-490 */
float ppo_step_length(float eta, float sca_len) { return (float)(-log((double)eta) * sca_len); }

/* Kopelevich scattering length (CherenkovMediumBase, KopelevichScatteringModel)
 * times sca_scale   src/water_properties.jl:95-100, constants :34-35.
 * b(l) = 0.0017 x^-4.3 + 1.34 vs x^-1.7 + 0.312 vl x^-0.3, x = l/550. */
float ppo_kopelevich_sca_len(float wl, float vs, float vl, float sca_scale) {
    double x = (double)wl / 550.0;
    double b = 0.0017 * pow(x, -4.3) + 1.34 * (double)vs * pow(x, -1.7) +
               0.312 * (double)vl * pow(x, -0.3);
    return (float)(1.0 / b * (double)sca_scale);
}

/* Quan & Fry (1995) phase refractive index of sea water (wl in nm, S permille,
 * T deg C) plus a linear pressure term 1.45e-5/atm; QuanFryDispersion in
 * CherenkovMediumBase (src/water_properties.jl:64, constants :37-40). */
static const double QF_N[10] = {1.31405,  1.779e-4, -1.05e-6,  1.6e-8,  -2.02e-6,
                                15.868,   0.01155,  -0.00423,  -4382.0, 1.1455e6};
#define QF_PRESSURE_COEF 1.45e-5

double ppo_quan_fry_n(double wl, double S, double T, double P) {
    const double* n = QF_N;
    return n[0] + (n[1] + n[2] * T + n[3] * T * T) * S + n[4] * T * T +
           (n[5] + n[6] * S + n[7] * T) / wl + n[8] / (wl * wl) + n[9] / (wl * wl * wl) +
           QF_PRESSURE_COEF * P;
}

/* group velocity c / (n - wl dn/dwl) in m/ns */
double ppo_quan_fry_group_velocity(double wl, double S, double T, double P) {
    const double* n = QF_N;
    double np = ppo_quan_fry_n(wl, S, T, P);
    double dndl = -(n[5] + n[6] * S + n[7] * T) / (wl * wl) - 2.0 * n[8] / (wl * wl * wl) -
                  3.0 * n[9] / (wl * wl * wl * wl);
    return PPO_C_VAC_M_NS / (np - wl * dndl);
}

/* Per-wavelength medium values: scattering length, group velocity,
 * phase refractive index (src/photon_prop_cuda.jl:480-483, :83). */
void ppo_medium_values(const pp_medium* m, float wl, float* sca_len, float* c_grp,
                       float* n_phase) {
    if (m->kind == PP_MEDIUM_SIMPLE) {
        *sca_len = m->sca_len;
        *c_grp = m->group_velocity;
        *n_phase = m->phase_refractive_index;
        return;
    }
    if (m->kind == PP_MEDIUM_ICE) {
        /* HomogenousIceProperties, src/ice_properties.jl:105-126, :156-168 */
        const double x = wl / 1000.0;
        const double np = 1.55749 - 1.57988 * x + 3.99993 * x * x - 4.68271 * x * x * x +
                          2.09354 * x * x * x * x;
        const double ng = np * (1 + 0.227106 - 0.954648 * x + 1.42568 * x * x - 0.711832 * x * x * x);
        const double eff = (double)m->ice_b_dust_400 * pow(wl / 400.0, -(double)m->ice_alpha_sca_dust);
        *sca_len = (float)(1.0 / (eff / (1.0 - (double)m->hg_g)) * m->sca_scale);
        *c_grp = (float)(PPO_C_VAC_M_NS / ng);
        *n_phase = (float)np;
        return;
    }
    *sca_len = ppo_kopelevich_sca_len(wl, m->vol_conc_small, m->vol_conc_large, m->sca_scale);
    *c_grp = (float)ppo_quan_fry_group_velocity(wl, m->salinity, m->temperature, m->pressure);
    *n_phase = (float)ppo_quan_fry_n(wl, m->salinity, m->temperature, m->pressure);
}

/* sample_scattering_function(medium): HG for SimpleMediumProperties
 * (src/simple_medium.jl:14), MixedHGES for Cascadia (src/water_properties.jl:57) */
float ppo_phase_cos(const pp_medium* m, float u_choice, float u) {
    if (m->phase_kind == PP_PHASE_HG) return ppo_hg_cos(m->hg_g, u);
    if (m->phase_kind == PP_PHASE_MIXED_HG_SL)
        return u_choice < m->hg_fraction ? ppo_hg_cos(m->hg_g, u) : ppo_sl_cos(m->hg_g, u);
    return ppo_mixed_hges_cos(m->hg_g, m->es_b, m->hg_fraction, u_choice, u);
}

/* sample_cherenkov_track_direction (CherenkovMediumBase): direction of a
 * cascade's emitting track about e_z with the clsim track-angle law
 * p(x) ~ x^(a-1) exp(-b x^a), x = 1 - cos, a = 0.39, b = 2.61, truncated at
 * x = 2 by inverse-CDF sampling with the Float32 normalisation
 * I = 1 - exp(-b 2^a). Draw order: cos theta, then phi. */
void ppo_cherenkov_track_dir(float u1, float u2, float out[3]) {
    const float a = 0.39f, b = 2.61f;
    const float I = 1.0f - expf(-b * powf(2.0f, a));
    double ct = 1.0 - pow(-log(1.0 - (double)u1 * I) / b, 1.0 / (double)a);
    if (ct < -1.0) ct = -1.0;
    /* cos_theta is a Float32 variable before sin_theta = sqrt(1 - cos^2) */
    ct = (double)(float)ct;
    sph_to_cart_d(ct, sqrt(1.0 - ct * ct), (double)(PPO_TWO_PI_F * u2), out);
}

/* ------------------------------------------------------------------ */
/* Samplers on a photon's sequential sampler sub-stream                */
/* ------------------------------------------------------------------ */
typedef struct {
    const uint32_t* key;
    uint64_t pid;
    uint32_t src;
} ppo_stream;

/* Standard normal from sampler sub-stream `sub`, block 0 (Box-Muller). */
static float ppo_normal(const ppo_stream* s, uint32_t sub) {
    uint32_t r[4];
    ppo_block(PP_CTR_SAMPLER | (sub << 24), s->pid, s->src, s->key, r);
    double u1 = ppo_u01_open0(r[0]), phi = PPO_TWO_PI_D * (double)ppo_u01(r[1]);
    return (float)(sqrt(-2.0 * log(u1)) * cos(phi));
}

/* PhysicsTools.rand_gamma(shape, scale): Marsaglia & Tsang (2000). Each
 * sampler block yields a Box-Muller pair (z0, z1) and two acceptance
 * uniforms; shape < 1 uses Gamma(shape+1) * U^(1/shape) with U from the
 * sub-stream's boost block. After PP_GAMMA_MAX_BLOCKS blocks without
 * acceptance (probability < 1e-100 for shape >= 0.1) returns d*scale. */
static float ppo_gamma_x(float shape, float scale, const ppo_stream* s, uint32_t sub,
                         float* x_acc);
float ppo_gamma(float shape, float scale, const ppo_stream* s, uint32_t sub) {
    return ppo_gamma_x(shape, scale, s, sub, NULL);
}

/* as ppo_gamma; *x_acc (may be NULL) receives the accepted normal variate
 * rounded to float (0 if none) */
static float ppo_gamma_x(float shape, float scale, const ppo_stream* s, uint32_t sub,
                         float* x_acc) {
    double boost = 1.0, k = shape;
    if (shape < 1.0f) {
        uint32_t r[4];
        ppo_block(PP_CTR_SAMPLER | (sub << 24) | PP_SAMPLER_BOOST_BLOCK, s->pid, s->src, s->key,
                  r);
        boost = pow((double)ppo_u01_open0(r[0]), 1.0 / k);
        k = k + 1.0;
    }
    double d = k - 1.0 / 3.0;
    double c = 1.0 / sqrt(9.0 * d);
    for (uint32_t i = 0; i < PP_GAMMA_MAX_BLOCKS; ++i) {
        uint32_t r[4];
        ppo_block(PP_CTR_SAMPLER | (sub << 24) | i, s->pid, s->src, s->key, r);
        double rad = sqrt(-2.0 * log((double)ppo_u01_open0(r[0])));
        double phi = PPO_TWO_PI_D * (double)ppo_u01(r[1]);  /* Box-Muller angle 2 pi u */
        double z[2] = {rad * cos(phi), rad * sin(phi)};
        double ua[2] = {ppo_u01_open0(r[2]), ppo_u01_open0(r[3])};
        for (int j = 0; j < 2; ++j) {
            double x = z[j];
            double v = 1.0 + c * x;
            if (v <= 0.0) continue;
            v = v * v * v;
            double u = ua[j];
            if (u < 1.0 - 0.0331 * (x * x) * (x * x) || log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) {
                if (x_acc) *x_acc = (float)x;
                return (float)(d * v * scale * boost);
            }
        }
    }
    if (x_acc) *x_acc = 0.0f;
    return (float)(d * scale * boost);
}

/* PhysicsTools.rand_beta(a, b) = X/(X+Y), X ~ Gamma(a), Y ~ Gamma(b) */
static float ppo_beta(float a, float b, const ppo_stream* s) {
    float x = ppo_gamma(a, 1.0f, s, PP_SUB_BETA_X);
    float y = ppo_gamma(b, 1.0f, s, PP_SUB_BETA_Y);
    return x / (x + y);
}

/* Inverse-CDF spectrum lookup (device texture of src/spectrum.jl:75-83). */
float ppo_spectrum_sample(const ppo_spectrum* sp, float u) {
    if (sp->table == NULL) return sp->mono_wl; /* Monochromatic: src/spectrum.jl:70 */
    int n = sp->n;
    float x = sp->mapping == PP_SPEC_TEXTURE ? u * (float)n - 0.5f : u * (float)(n - 1);
    if (x < 0.0f) x = 0.0f;
    if (x > (float)(n - 1)) x = (float)(n - 1);
    int i = (int)x;
    if (i > n - 2) i = n - 2;
    float f = x - (float)i;
    return sp->table[i] + f * (sp->table[i + 1] - sp->table[i]);
}

/* ------------------------------------------------------------------ */
/* Photon initialisation   src/photon_prop_cuda.jl:158-276             */
/* ------------------------------------------------------------------ */
static void ppo_add_scaled(const float p[3], float s, const float d[3], float out[3]) {
    out[0] = p[0] + s * d[0];
    out[1] = p[1] + s * d[1];
    out[2] = p[2] + s * d[2];
}

/* sample_cherenkov_direction (smear = 1), _no_smear (0), _custom_smear (2)
 *   src/photon_prop_cuda.jl:80-111, :134-154 */
static void ppo_cherenkov_direction(const pp_source* src, float n_phase, const ppo_stream* s,
                                    int smear, float out[3]) {
    uint32_t x[4];
    ppo_block(PP_CTR_INIT_EXTRA | 1u, s->pid, s->src, s->key, x);
    /* cherenkov_angle = acos(1/n) (beta = 1) */
    double cos_c = 1.0 / (double)n_phase;
    float ph[3], tmp[3];
    sph_to_cart_d(cos_c, sqrt(1.0 - cos_c * cos_c), (double)ppo_phi(ppo_u01(x[0])), ph);
    if (smear == 1) {
        float track[3];
        ppo_cherenkov_track_dir(ppo_u01(x[1]), ppo_u01(x[2]), track);
        ppo_rot_from_ez(track, ph, tmp);
        memcpy(ph, tmp, sizeof ph);
    } else if (smear == 2) {
        float tcos = ppo_beta(src->beta_a, src->beta_b, s) * 2.0f - 1.0f;
        float track[3];
        sph_to_cart_d(tcos, sqrt(1.0 - (double)tcos * tcos), (double)ppo_phi(ppo_u01(x[1])),
                      track);
        ppo_rot_from_ez(track, ph, tmp);
        memcpy(ph, tmp, sizeof ph);
    }
    ppo_rot_from_ez(src->direction, ph, out);
}

void ppo_init_photon(const pp_source* src, const pp_medium* m, const ppo_spectrum* sp,
                     const uint32_t key[2], uint32_t src_ix, uint64_t pid, ppo_photon* st) {
    ppo_stream s = {key, pid, src_ix};
    uint32_t r[4];
    ppo_block(PP_CTR_INIT, pid, src_ix, key, r);
    float wl = ppo_spectrum_sample(sp, ppo_u01(r[0]));
    float sca_len, c_grp, n_phase;
    ppo_medium_values(m, wl, &sca_len, &c_grp, &n_phase);
    st->wavelength = wl;
    st->sca_len = sca_len;
    st->c_grp = c_grp;
    memcpy(st->pos, src->position, sizeof st->pos);
    st->time = src->time;
    switch (src->kind) {
    case PP_SRC_POINTLIKE_ISOTROPIC: /* :158-163 */
        ppo_isotropic_dir(ppo_u01(r[1]), ppo_u01(r[2]), st->dir);
        break;
    case PP_SRC_POINTLIKE_TIME_RANGE: /* :165-171 */
        ppo_isotropic_dir(ppo_u01(r[1]), ppo_u01(r[2]), st->dir);
        st->time = ppo_uniform(src->time, src->time_end, ppo_u01(r[3]));
        break;
    case PP_SRC_AXICONE: { /* :173-184 */
        float phi = ppo_uniform(0.0f, PPO_TWO_PI_F, ppo_u01(r[1]));
        float theta = src->angle;
        if (src->beam_divergence > 0.0f) theta += ppo_normal(&s, PP_SUB_NORMAL) * src->beam_divergence;
        float v[3];
        ppo_sph_to_cart(theta, phi, v);
        ppo_rot_from_ez(src->direction, v, st->dir);
        break;
    }
    case PP_SRC_COLLIMATED_ISOTROPIC: { /* :186-195 */
        float phi = ppo_uniform(0.0f, PPO_TWO_PI_F, ppo_u01(r[1]));
        double ct = ppo_uniform(src->cut_angle_cos, 1.0f, ppo_u01(r[2]));
        sph_to_cart_d(ct, sqrt(1.0 - ct * ct), (double)phi, st->dir);
        break;
    }
    case PP_SRC_PENCIL: /* :197-208 */
        if (src->beam_divergence > 0.0f) {
            float theta = sqrtf(ppo_uniform(0.0f, src->beam_divergence * src->beam_divergence,
                                            ppo_u01(r[1])));
            float phi = ppo_uniform(0.0f, PPO_TWO_PI_F, ppo_u01(r[2]));
            float v[3];
            ppo_sph_to_cart(theta, phi, v);
            ppo_rot_from_ez(src->direction, v, st->dir);
        } else {
            memcpy(st->dir, src->direction, sizeof st->dir);
        }
        break;
    case PP_SRC_EXTENDED_CHERENKOV:              /* :211-222 */
    case PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR:     /* :225-236 */
    case PP_SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR: { /* :238-250 */
        float long_pos = ppo_gamma(src->long_a, 1.0f / src->long_b, &s, PP_SUB_LONG_GAMMA) *
                         src->long_lrad;
        ppo_add_scaled(src->position, long_pos, src->direction, st->pos);
        st->time = src->time + long_pos / (float)PPO_C_VAC_M_NS;
        int smear = src->kind == PP_SRC_EXTENDED_CHERENKOV ? 1
                    : src->kind == PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR ? 0 : 2;
        ppo_cherenkov_direction(src, n_phase, &s, smear, st->dir);
        break;
    }
    case PP_SRC_POINTLIKE_CHERENKOV: /* :254-263 */
        ppo_cherenkov_direction(src, n_phase, &s, 1, st->dir);
        break;
    case PP_SRC_CHERENKOV_TRACK: { /* :265-276 */
        float pos_along = ppo_uniform(0.0f, src->length, ppo_u01(r[1]));
        ppo_add_scaled(src->position, pos_along, src->direction, st->pos);
        st->time = src->time + pos_along / (float)PPO_C_VAC_M_NS;
        ppo_cherenkov_direction(src, n_phase, &s, 1, st->dir);
        break;
    }
    default:
        break;
    }
}

/* ------------------------------------------------------------------ */
/* Photon transport   src/photon_prop_cuda.jl:466-543                  */
/* ------------------------------------------------------------------ */
int ppo_transport(ppo_photon* st, const pp_medium* m, const pp_target* targets, int n_targets,
                  const uint32_t key[2], uint32_t src_ix, uint64_t pid, int32_t n_steps,
                  ppo_hit* hit) {
    float pos[3] = {st->pos[0], st->pos[1], st->pos[2]};
    float dir[3] = {st->dir[0], st->dir[1], st->dir[2]};
    float time = st->time, dist = 0.0f;
    for (int32_t nstep = 1; nstep <= n_steps; ++nstep) {
        uint32_t r[4];
        ppo_block((uint32_t)nstep, pid, src_ix, key, r);
        float step_size = ppo_step_length(ppo_u01_open0(r[0]), st->sca_len);
        int isec = 0, module_ix = 0;
        float dist_to_target = 0.0f;
        for (int ix = 0; ix < n_targets; ++ix) {
            float d;
            if (ppo_intersect_target(&targets[ix], pos, dir, step_size, &d)) {
                isec = 1;
                module_ix = ix + 1;
                dist_to_target = d;
                break;
            }
        }
        if (!isec) { /* :508-514 */
            float np[3];
            ppo_add_scaled(pos, step_size, dir, np);
            memcpy(pos, np, sizeof pos);
            dist += step_size;
            time += step_size / st->c_grp;
            float ct = ppo_phase_cos(m, ppo_u01(r[1]), ppo_u01(r[2]));
            float nd[3];
            ppo_update_direction(dir, ct, ppo_u01(r[3]), nd);
            memcpy(dir, nd, sizeof dir);
            continue;
        }
        /* :516-536 */
        float np[3];
        ppo_add_scaled(pos, dist_to_target, dir, np);
        dist += dist_to_target;
        time += dist_to_target / st->c_grp;
        memcpy(hit->pos, np, sizeof hit->pos);
        memcpy(hit->dir, dir, sizeof hit->dir);
        memcpy(hit->init_dir, st->dir, sizeof hit->init_dir);
        hit->wavelength = st->wavelength;
        hit->time = time;
        hit->dist = dist;
        hit->module_ix = module_ix;
        hit->n_steps = nstep;
        hit->photon_id = pid;
        hit->source_ix = (int32_t)src_ix;
        return 1;
    }
    return 0;
}

/* Photons of a source: floor(Float32(photons) * photon_scaling)
 *   src/photon_prop_cuda.jl:456-460 (Int64 * Float32 is a Float32 product). */
int64_t ppo_photon_count(int64_t photons, float photon_scaling) {
    float p = (float)photons * photon_scaling;
    return (int64_t)floorf(p);
}

static int ppo_hit_cmp(const void* a, const void* b) {
    const ppo_hit* x = (const ppo_hit*)a;
    const ppo_hit* y = (const ppo_hit*)b;
    if (x->source_ix != y->source_ix) return x->source_ix < y->source_ix ? -1 : 1;
    if (x->photon_id != y->photon_id) return x->photon_id < y->photon_id ? -1 : 1;
    return 0;
}

/* Whole propagation, hits ordered by (source, photon id). Returns the number
 * of hits, or -1 if more than `capacity`; *n_ph_sim counts photons. */
int64_t ppo_propagate(const pp_source* sources, int n_sources, const pp_medium* m,
                      const ppo_spectrum* sp, const pp_target* targets, int n_targets,
                      uint64_t seed, int32_t n_steps, float photon_scaling, ppo_hit* hits,
                      int64_t capacity, int64_t* n_ph_sim, int n_threads) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t total_hits = 0, total_ph = 0;
    int overflow = 0;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    for (int si = 0; si < n_sources; ++si) {
        int64_t nph = ppo_photon_count(sources[si].photons, photon_scaling);
        const pp_source* src = &sources[si];
#pragma omp parallel num_threads(n_threads) reduction(+ : total_ph)
        {
            ppo_hit local[256];
            int nloc = 0;
#pragma omp for schedule(dynamic, 4096)
            for (int64_t pid = 0; pid < nph; ++pid) {
                ppo_photon st;
                ppo_init_photon(src, m, sp, key, (uint32_t)si, (uint64_t)pid, &st);
                if (ppo_transport(&st, m, targets, n_targets, key, (uint32_t)si, (uint64_t)pid,
                                  n_steps, &local[nloc]))
                    ++nloc;
                total_ph += 1;
                if (nloc == 256) {
#pragma omp critical
                    {
                        if (total_hits + nloc <= capacity)
                            memcpy(hits + total_hits, local, nloc * sizeof(ppo_hit));
                        else
                            overflow = 1;
                        total_hits += nloc;
                    }
                    nloc = 0;
                }
            }
            if (nloc) {
#pragma omp critical
                {
                    if (total_hits + nloc <= capacity)
                        memcpy(hits + total_hits, local, nloc * sizeof(ppo_hit));
                    else
                        overflow = 1;
                    total_hits += nloc;
                }
            }
        }
    }
    *n_ph_sim = total_ph;
    if (overflow) return -1;
    qsort(hits, (size_t)total_hits, sizeof(ppo_hit), ppo_hit_cmp);
    return total_hits;
}

/* Draws of the sampler stream, exposed for distribution tests. */
float ppo_gamma_sample(float shape, float scale, uint64_t seed, uint64_t pid, uint32_t src_ix) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ppo_stream s = {key, pid, src_ix};
    return ppo_gamma(shape, scale, &s, PP_SUB_LONG_GAMMA);
}

float ppo_beta_sample(float a, float b, uint64_t seed, uint64_t pid, uint32_t src_ix) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ppo_stream s = {key, pid, src_ix};
    return ppo_beta(a, b, &s);
}

int ppo_abi_version(void) { return PPO_ABI_VERSION; }

/* Batch evaluation with the function ids and layouts of pp_eval_fn
 * (include/photonprop_hip.h), so that tests feed both the same rows. */
int ppo_eval_fn(int fn, const pp_medium* m, const ppo_spectrum* sp, const float* in,
                int in_stride, float* out, int out_stride, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const float* x = in + i * in_stride;
        float* y = out + i * out_stride;
        switch (fn) {
        case PP_FN_ISOTROPIC_DIR: ppo_isotropic_dir(x[0], x[1], y); break;
        case PP_FN_ROT_FROM_EZ: ppo_rot_from_ez(x, x + 3, y); break;
        case PP_FN_HG_COS: y[0] = ppo_hg_cos(x[0], x[1]); break;
        case PP_FN_ES_COS: y[0] = ppo_es_cos(x[0], x[1]); break;
        case PP_FN_UPDATE_DIR: ppo_update_direction(x, x[3], x[4], y); break;
        case PP_FN_SPHERE_ISEC: {
            float d = 0.0f;
            int h = ppo_intersect_sphere(x, x[3], x + 4, x + 7, x[10], &d);
            y[0] = h ? 1.0f : 0.0f;
            y[1] = h ? d : 0.0f;
            break;
        }
        case PP_FN_PLANE_ISEC: {
            float d = 0.0f;
            uint32_t kind;
            memcpy(&kind, x, 4);
            int h = ppo_intersect_plane((int)kind, x + 1, x[4], x[5], x + 6, x + 9, x[12], &d);
            y[0] = h ? 1.0f : 0.0f;
            y[1] = h ? d : 0.0f;
            break;
        }
        case PP_FN_STEP_LEN: y[0] = ppo_step_length(x[0], x[1]); break;
        case PP_FN_CASCADIA: {
            if (!m) return -1;
            ppo_medium_values(m, x[0], &y[0], &y[1], &y[2]);
            break;
        }
        case PP_FN_TRACK_DIR: ppo_cherenkov_track_dir(x[0], x[1], y); break;
        case PP_FN_PHILOX: {
            uint32_t c[4], k[2], r[4];
            memcpy(c, x, 16);
            memcpy(k, x + 4, 8);
            ppo_philox4x32_10(c, k, r);
            memcpy(y, r, 16);
            break;
        }
        case PP_FN_GAMMA: {
            uint32_t w[5];
            memcpy(w, x + 2, 20);
            uint32_t key[2] = {w[0], w[1]};
            ppo_stream st = {key, (uint64_t)w[2] | ((uint64_t)w[3] << 32), w[4]};
            y[0] = ppo_gamma_x(x[0], x[1], &st, PP_SUB_LONG_GAMMA, &y[1]);
            break;
        }
        case PP_FN_SPECTRUM:
            if (!sp) return -1;
            y[0] = ppo_spectrum_sample(sp, x[0]);
            break;
        case PP_FN_SL_COS: y[0] = ppo_sl_cos(x[0], x[1]); break;
        case PP_FN_MIXED_HGES: y[0] = ppo_mixed_hges_cos(x[0], x[1], x[2], x[3], x[4]); break;
        default: return -1;
        }
    }
    return 0;
}
