/*
 * pp_ref_f32.c - Float32 port of the reference propagation kernel for the CPU
 * baseline.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY. bench.py's cpu_baseline leg times
 * it and tests/ check it against the oracle (pp_oracle.c); the shipped
 * library never loads it.
 *
 * The oracle evaluates every step function in double precision and rounds
 * once, so that it is the exact value the GPU is measured against; that makes
 * it slower than the reference's own arithmetic. This file is the reference's
 * kernel as the reference computes it: every value a Float32, libm's float
 * functions (logf, sqrtf, sinf, cosf, acosf, powf, cbrtf), the every-target
 * scan in index order, n_steps steps, photons of a source spread over OpenMP
 * threads (schedule(dynamic, 4096)). It follows
 *   initialize_photon_state   src/photon_prop_cuda.jl:158-276
 *   update_direction          src/photon_prop_cuda.jl:280-298
 *   _intersect_nsphere        src/photon_prop_cuda.jl:309-351
 *   cuda_propagate_photons!   src/photon_prop_cuda.jl:429-549
 * with the oracle's random streams (per-photon Philox4x32-10, the same words),
 * so that it reproduces the oracle's photons up to Float32 rounding. Sources:
 * point isotropic, extended Cherenkov (cascade) and Cherenkov track; media:
 * Cascadia (Kopelevich, Quan-Fry, MixedHGES) and simple (HG); spherical
 * targets. Anything else returns -2.
 */
#include "pp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI_F 6.28318548202514648438f
#define C_VAC_F 0.299792458f

typedef struct {
    const uint32_t* key;
    uint64_t pid;
    uint32_t src;
} f_stream;

static void f_block(const f_stream* s, uint32_t word, uint32_t out[4]) {
    uint32_t ctr[4] = {word, (uint32_t)s->pid, (uint32_t)(s->pid >> 32), s->src};
    ppo_philox4x32_10(ctr, s->key, out);
}

static float f_u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
static float f_u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

static void f_sph_to_cart(float ct, float st, float phi, float o[3]) {
    o[0] = cosf(phi) * st;
    o[1] = sinf(phi) * st;
    o[2] = ct;
}

/* rot_from_ez_fast: the rotation about e_z x a taking e_z to a (the oracle's
 * convention), in Float32 */
static void f_rot_from_ez(const float a[3], const float b[3], float o[3]) {
    float s = sqrtf(a[0] * a[0] + a[1] * a[1]);
    float c = a[2];
    if (s == 0.0f) {
        o[0] = b[0];
        o[1] = c > 0.0f ? b[1] : -b[1];
        o[2] = c > 0.0f ? b[2] : -b[2];
        return;
    }
    float kx = -a[1] / s, ky = a[0] / s;
    float kdb = kx * b[0] + ky * b[1];
    float omc = 1.0f - c;
    float r0 = b[0] * c + ky * b[2] * s + kx * kdb * omc;
    float r1 = b[1] * c - kx * b[2] * s + ky * kdb * omc;
    float r2 = b[2] * c + (kx * b[1] - ky * b[0]) * s;
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
}

typedef struct {
    int simple, phase_kind;
    float g, es_b, hg_fraction;
    float sca_len, c_grp, n_phase;  /* SIMPLE */
    float vs, vl, sca_scale;        /* Kopelevich */
    float S, T, P;                  /* Quan-Fry */
} f_medium;

static float f_hg_cos(float g, float u) {
    float t = (1.0f - g * g) / (1.0f + g * (2.0f * u - 1.0f));
    float x = (1.0f + g * g - t * t) / (2.0f * g);
    return x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
}

static float f_es_cos(float b, float u) {
    float p = 3.0f / b;
    float q = (3.0f / b) * ((1.0f + b / 3.0f) - u * (2.0f + 2.0f * b / 3.0f));
    float disc = sqrtf(q * q / 4.0f + p * p * p / 27.0f);
    float x = cbrtf(-q / 2.0f + disc) + cbrtf(-q / 2.0f - disc);
    return x < -1.0f ? -1.0f : (x > 1.0f ? 1.0f : x);
}

static float f_phase_cos(const f_medium* m, float u_choice, float u) {
    if (m->phase_kind == PP_PHASE_HG) return f_hg_cos(m->g, u);
    return u_choice < m->hg_fraction ? f_hg_cos(m->g, u) : f_es_cos(m->es_b, u);
}

static void f_medium_values(const f_medium* m, float wl, float* sca_len, float* c_grp,
                            float* n_phase) {
    if (m->simple) {
        *sca_len = m->sca_len;
        *c_grp = m->c_grp;
        *n_phase = m->n_phase;
        return;
    }
    float x = wl / 550.0f;
    float b = 0.0017f * powf(x, -4.3f) + 1.34f * m->vs * powf(x, -1.7f) +
              0.312f * m->vl * powf(x, -0.3f);
    *sca_len = 1.0f / b * m->sca_scale;
    float S = m->S, T = m->T;
    float a1 = 15.868f + 0.01155f * S - 0.00423f * T;
    float n = 1.31405f + (1.779e-4f - 1.05e-6f * T + 1.6e-8f * T * T) * S - 2.02e-6f * T * T +
              a1 / wl - 4382.0f / (wl * wl) + 1.1455e6f / (wl * wl * wl) + 1.45e-5f * m->P;
    float dndl = -a1 / (wl * wl) + 2.0f * 4382.0f / (wl * wl * wl) -
                 3.0f * 1.1455e6f / (wl * wl * wl * wl);
    *n_phase = n;
    *c_grp = C_VAC_F / (n - wl * dndl);
}

/* Marsaglia-Tsang in Float32 with the oracle's draw layout */
static float f_gamma(float shape, float scale, const f_stream* s, uint32_t sub) {
    float boost = 1.0f, k = shape;
    if (shape < 1.0f) {
        uint32_t r[4];
        f_block(s, PP_CTR_SAMPLER | (sub << 24) | PP_SAMPLER_BOOST_BLOCK, r);
        boost = powf(f_u01_open0(r[0]), 1.0f / k);
        k += 1.0f;
    }
    float d = k - 1.0f / 3.0f;
    float c = 1.0f / sqrtf(9.0f * d);
    for (uint32_t i = 0; i < PP_GAMMA_MAX_BLOCKS; ++i) {
        uint32_t r[4];
        f_block(s, PP_CTR_SAMPLER | (sub << 24) | i, r);
        float rad = sqrtf(-2.0f * logf(f_u01_open0(r[0])));
        float phi = TWO_PI_F * f_u01(r[1]);
        float z[2] = {rad * cosf(phi), rad * sinf(phi)};
        float ua[2] = {f_u01_open0(r[2]), f_u01_open0(r[3])};
        for (int j = 0; j < 2; ++j) {
            float x = z[j];
            float v = 1.0f + c * x;
            if (v <= 0.0f) continue;
            v = v * v * v;
            float u = ua[j];
            if (u < 1.0f - 0.0331f * (x * x) * (x * x) ||
                logf(u) < 0.5f * x * x + d * (1.0f - v + logf(v)))
                return d * v * scale * boost;
        }
    }
    return d * scale * boost;
}

static void f_track_dir(float u1, float u2, float o[3]) {
    const float a = 0.39f, b = 2.61f;
    const float I = 1.0f - expf(-b * powf(2.0f, a));
    float ct = 1.0f - powf(-logf(1.0f - u1 * I) / b, 1.0f / a);
    if (ct < -1.0f) ct = -1.0f;
    f_sph_to_cart(ct, sqrtf(1.0f - ct * ct), TWO_PI_F * u2, o);
}

static float f_spectrum(const ppo_spectrum* sp, float u) {
    if (sp->table == NULL) return sp->mono_wl;
    int n = sp->n;
    float x = sp->mapping == PP_SPEC_TEXTURE ? u * (float)n - 0.5f : u * (float)(n - 1);
    if (x < 0.0f) x = 0.0f;
    if (x > (float)(n - 1)) x = (float)(n - 1);
    int i = (int)x;
    if (i > n - 2) i = n - 2;
    float f = x - (float)i;
    return sp->table[i] + f * (sp->table[i + 1] - sp->table[i]);
}

typedef struct {
    float pos[3], dir[3], time, wl, sca_len, c_grp;
} f_photon;

static void f_cherenkov_dir(const pp_source* src, float n_phase, const f_stream* s, int smear,
                            float o[3]) {
    uint32_t x[4];
    f_block(s, PP_CTR_INIT_EXTRA | 1u, x);
    float theta = acosf(1.0f / n_phase);
    float ph[3], t[3];
    f_sph_to_cart(cosf(theta), sinf(theta), TWO_PI_F * f_u01(x[0]), ph);
    if (smear) {
        float track[3];
        f_track_dir(f_u01(x[1]), f_u01(x[2]), track);
        f_rot_from_ez(track, ph, t);
        memcpy(ph, t, sizeof ph);
    }
    f_rot_from_ez(src->direction, ph, o);
}

static int f_init(const pp_source* src, const f_medium* m, const ppo_spectrum* sp,
                  const f_stream* s, f_photon* st) {
    uint32_t r[4];
    f_block(s, PP_CTR_INIT, r);
    float n_phase;
    st->wl = f_spectrum(sp, f_u01(r[0]));
    f_medium_values(m, st->wl, &st->sca_len, &st->c_grp, &n_phase);
    memcpy(st->pos, src->position, sizeof st->pos);
    st->time = src->time;
    switch (src->kind) {
    case PP_SRC_POINTLIKE_ISOTROPIC: {
        float ct = -1.0f + 2.0f * f_u01(r[1]);
        float theta = acosf(ct);
        f_sph_to_cart(cosf(theta), sinf(theta), TWO_PI_F * f_u01(r[2]), st->dir);
        return 0;
    }
    case PP_SRC_EXTENDED_CHERENKOV:
    case PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR: {
        float lp = f_gamma(src->long_a, 1.0f / src->long_b, s, PP_SUB_LONG_GAMMA) * src->long_lrad;
        for (int k = 0; k < 3; ++k) st->pos[k] = src->position[k] + lp * src->direction[k];
        st->time = src->time + lp / C_VAC_F;
        f_cherenkov_dir(src, n_phase, s, src->kind == PP_SRC_EXTENDED_CHERENKOV, st->dir);
        return 0;
    }
    case PP_SRC_CHERENKOV_TRACK: {
        float along = src->length * f_u01(r[1]);
        for (int k = 0; k < 3; ++k) st->pos[k] = src->position[k] + along * src->direction[k];
        st->time = src->time + along / C_VAC_F;
        f_cherenkov_dir(src, n_phase, s, 1, st->dir);
        return 0;
    }
    default:
        return -2;
    }
}

/* one photon: returns 1 with *hit filled on a hit */
static int f_transport(f_photon* st, const f_medium* m, const float* tgt, int n_targets,
                       const f_stream* s, int32_t n_steps, ppo_hit* hit) {
    float pos[3] = {st->pos[0], st->pos[1], st->pos[2]};
    float dir[3] = {st->dir[0], st->dir[1], st->dir[2]};
    float time = st->time, dist = 0.0f;
    for (int32_t nstep = 1; nstep <= n_steps; ++nstep) {
        uint32_t r[4];
        f_block(s, (uint32_t)nstep, r);
        float step = -logf(f_u01_open0(r[0])) * st->sca_len;
        int found = 0;
        float d_hit = 0.0f;
        for (int ix = 0; ix < n_targets; ++ix) {
            const float* t = tgt + 4 * ix;
            float dp0 = pos[0] - t[0], dp1 = pos[1] - t[1], dp2 = pos[2] - t[2];
            float a = dir[0] * dp0 + dir[1] * dp1 + dir[2] * dp2;
            float b = a * a - (dp0 * dp0 + dp1 * dp1 + dp2 * dp2 - t[3]);
            float d = b >= 0.0f ? (-a - sqrtf(fabsf(b))) : -1.0f;
            if (d > 0.0f && d < step) {
                found = ix + 1;
                d_hit = d;
                break;
            }
        }
        if (!found) {
            for (int k = 0; k < 3; ++k) pos[k] = pos[k] + step * dir[k];
            dist += step;
            time += step / st->c_grp;
            float ct = f_phase_cos(m, f_u01(r[1]), f_u01(r[2]));
            float sn = sqrtf(1.0f - ct * ct);
            float nd[3], o[3];
            f_sph_to_cart(ct, sn, TWO_PI_F * f_u01(r[3]), nd);
            f_rot_from_ez(dir, nd, o);
            memcpy(dir, o, sizeof dir);
            continue;
        }
        for (int k = 0; k < 3; ++k) hit->pos[k] = pos[k] + d_hit * dir[k];
        dist += d_hit;
        time += d_hit / st->c_grp;
        memcpy(hit->dir, dir, sizeof hit->dir);
        memcpy(hit->init_dir, st->dir, sizeof hit->init_dir);
        hit->wavelength = st->wl;
        hit->time = time;
        hit->dist = dist;
        hit->module_ix = found;
        hit->n_steps = nstep;
        hit->photon_id = s->pid;
        hit->source_ix = (int32_t)s->src;
        return 1;
    }
    return 0;
}

/* Whole propagation with the signature of ppo_propagate. Returns the number
 * of hits (unordered), -1 on overflow of `hits`, -2 for a source, medium or
 * target kind this port does not cover. *n_steps_done counts photon steps. */
int64_t ppf_propagate(const pp_source* sources, int n_sources, const pp_medium* pm,
                      const ppo_spectrum* sp, const pp_target* targets, int n_targets,
                      uint64_t seed, int32_t n_steps, float photon_scaling, ppo_hit* hits,
                      int64_t capacity, int64_t* n_ph_sim, int64_t* n_steps_done,
                      int n_threads) {
    f_medium m;
    memset(&m, 0, sizeof m);
    if (pm->kind != PP_MEDIUM_SIMPLE && pm->kind != PP_MEDIUM_CASCADIA) return -2;
    if (pm->phase_kind != PP_PHASE_HG && pm->phase_kind != PP_PHASE_MIXED_HG_ES) return -2;
    m.simple = pm->kind == PP_MEDIUM_SIMPLE;
    m.phase_kind = pm->phase_kind;
    m.g = pm->hg_g;
    m.es_b = pm->es_b;
    m.hg_fraction = pm->hg_fraction;
    m.sca_len = pm->sca_len;
    m.c_grp = pm->group_velocity;
    m.n_phase = pm->phase_refractive_index;
    m.vs = pm->vol_conc_small;
    m.vl = pm->vol_conc_large;
    m.sca_scale = pm->sca_scale;
    m.S = pm->salinity;
    m.T = pm->temperature;
    m.P = pm->pressure;
    float* tgt = (float*)malloc(sizeof(float) * 4 * (size_t)(n_targets > 0 ? n_targets : 1));
    if (!tgt) return -1;
    for (int i = 0; i < n_targets; ++i) {
        if (targets[i].shape_kind != PP_SHAPE_SPHERE) {
            free(tgt);
            return -2;
        }
        memcpy(tgt + 4 * i, targets[i].position, 3 * sizeof(float));
        tgt[4 * i + 3] = targets[i].radius * targets[i].radius;
    }
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t total_hits = 0, total_ph = 0, total_steps = 0;
    int overflow = 0, bad = 0;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    for (int si = 0; si < n_sources; ++si) {
        int64_t nph = ppo_photon_count(sources[si].photons, photon_scaling);
        const pp_source* src = &sources[si];
#pragma omp parallel num_threads(n_threads) reduction(+ : total_ph, total_steps)
        {
            ppo_hit local[256];
            int nloc = 0;
#pragma omp for schedule(dynamic, 4096)
            for (int64_t pid = 0; pid < nph; ++pid) {
                f_stream s = {key, (uint64_t)pid, (uint32_t)si};
                f_photon st;
                if (f_init(src, &m, sp, &s, &st)) {
                    bad = 1;
                    continue;
                }
                if (f_transport(&st, &m, tgt, n_targets, &s, n_steps, &local[nloc])) {
                    total_steps += local[nloc].n_steps;
                    ++nloc;
                } else {
                    total_steps += n_steps;
                }
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
    free(tgt);
    *n_ph_sim = total_ph;
    if (n_steps_done) *n_steps_done = total_steps;
    if (bad) return -2;
    if (overflow) return -1;
    return total_hits;
}
