/*
 * pp_pmt_oracle.c - CPU oracle of the POM's PMT-level detection,
 * check_pmt_hit(::POM{<:POMAcceptance}) (src/pom.jl:246-305), with the PMT
 * layout of make_pom_pmt_coordinates (src/pom.jl:108-141).
 *
 * TEST INFRASTRUCTURE ONLY (see pp_oracle.c). Computed in double from the
 * float inputs, with the photon's Philox draws of the library (word
 * PP_CTR_PMT), so the GPU's PMT choice can be compared hit by hit.
 *
 * Restated from third-party code absent here (parity unpinned beyond the
 * invariants the tests check): PhysicsTools.calc_rot_matrix(a, b) as the
 * rotation taking a onto b about a x b, and NeutrinoTelescopeBase's
 * get_pmt_positions as orientation x (sin t cos f, sin t sin f, cos t).
 */
#include <math.h>
#include <stdint.h>

#include "../include/photonprop_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void ppo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

void ppo_pom_pmt_dirs(const float orient[9], double dirs[48]) {
    const double d2r = M_PI / 180.0;
    double th[16], ph[16];
    const double lower2[4] = {M_PI / 2, 0.0, 3 * M_PI / 2, M_PI};
    const double lower[4] = {M_PI / 4, 7 * M_PI / 4, 5 * M_PI / 4, 3 * M_PI / 4};
    for (int k = 0; k < 4; ++k) {
        th[k] = (90.0 - 25.0) * d2r;     /* upper */
        ph[k] = k * M_PI / 2;
        th[4 + k] = (90.0 - 57.5) * d2r; /* upper 2 */
        ph[4 + k] = M_PI / 4 + k * M_PI / 2;
        th[8 + k] = (90.0 + 25.0) * d2r; /* lower 2 */
        ph[8 + k] = lower2[k];
        th[12 + k] = (90.0 + 57.5) * d2r; /* lower */
        ph[12 + k] = lower[k];
    }
    /* calc_rot_matrix(e_z, e_x): Rodrigues rotation about e_z x e_x = e_y by 90 deg */
    const double R[9] = {0, 0, 1, 0, 1, 0, -1, 0, 0};
    for (int j = 0; j < 16; ++j) {
        const double v[3] = {sin(th[j]) * cos(ph[j]), sin(th[j]) * sin(ph[j]), cos(th[j])};
        double w[3];
        for (int r = 0; r < 3; ++r) w[r] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
        for (int r = 0; r < 3; ++r)
            dirs[3 * j + r] = orient[3 * r] * w[0] + orient[3 * r + 1] * w[1] + orient[3 * r + 2] * w[2];
    }
}

static double interp0(const float* x, const float* y, int n, double v) {
    if (!(v >= x[0] && v <= x[n - 1])) return 0.0;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x[mid] <= v) lo = mid;
        else hi = mid;
    }
    const double t = (v - x[lo]) / ((double)x[hi] - x[lo]);
    return y[lo] + t * ((double)y[hi] - y[lo]);
}

/* PMT id (1..16, 0 = none) per hit; prob_sum (may be NULL) receives sum_j p_j. */
void ppo_pom_pmt_hits(const float orient[9], float sigma1, float sigma2, int n_wl,
                      const float* wl_grid, const float* acc1, const float* acc2, int64_t n,
                      const float* rel_pos3, const float* wl, const uint64_t* pid,
                      const int32_t* src, uint64_t seed, int32_t* out, double* prob_sum) {
    double dirs[48];
    ppo_pom_pmt_dirs(orient, dirs);
    const double sig[2] = {sigma1, sigma2};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t i = 0; i < n; ++i) {
        const double acc[2] = {interp0(wl_grid, acc1, n_wl, wl[i]), interp0(wl_grid, acc2, n_wl, wl[i])};
        double p[16], sum = 0.0;
        for (int j = 0; j < 16; ++j) {
            const int g = (j >> 2) & 1; /* groups 1,1,1,1,2,2,2,2,1,1,1,1,2,2,2,2 */
            double c = rel_pos3[3 * i] * dirs[3 * j] + rel_pos3[3 * i + 1] * dirs[3 * j + 1] +
                       rel_pos3[3 * i + 2] * dirs[3 * j + 2];
            c = c < -1 ? -1 : (c > 1 ? 1 : c);
            const double pt = acos(c), st = sin(pt);
            const double pdf = pt / (sig[g] * sig[g]) * exp(-pt * pt / (2 * sig[g] * sig[g]));
            p[j] = st != 0.0 ? pdf / (0.5 * st) * acc[g] / 8.0 : 0.0;
            sum += p[j];
        }
        if (prob_sum) prob_sum[i] = sum;
        const uint32_t ctr[4] = {PP_CTR_PMT, (uint32_t)pid[i], (uint32_t)(pid[i] >> 32), (uint32_t)src[i]};
        uint32_t r[4];
        ppo_philox4x32_10(ctr, key, r);
        const double u1 = (double)(r[0] >> 8) * 0x1p-24, u2 = (double)(r[1] >> 8) * 0x1p-24;
        int pmt = 0;
        if (u1 < sum) {
            const double u = u2 * sum;
            double cum = 0.0;
            for (int j = 0; j < 16; ++j) {
                cum += p[j];
                if (p[j] > 0.0) {
                    pmt = j + 1;
                    if (u < cum) break;
                }
            }
        }
        out[i] = pmt;
    }
}
