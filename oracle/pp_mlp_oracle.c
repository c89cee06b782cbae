/*
 * pp_mlp_oracle.c - CPU oracle of the surrogate MLP (include/photonprop_surrogate.h).
 *
 * TEST INFRASTRUCTURE ONLY, like pp_oracle.c: the reference values for the
 * GPU's forward pass, loss head, gradients and Adam step. Everything is computed in
 * double. With quant = 1 the values are rounded to bf16 (round to nearest
 * even) at exactly the points where the GPU kernels hold bf16: the input
 * features, the weights, each hidden activation, the output gradient and each
 * hidden pre-activation gradient; the GPU's remaining differences are fp32
 * accumulation and its exp/log. With quant = 0 it is the exact network, whose
 * gradient the tests check against finite differences.
 *
 * The surrogate has no counterpart in the reference; this file pins it by
 * mathematics only (finite differences, closed forms), see tests/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define N_OUT 10
#define BINS 64
#define TMAX 2000.0
#define LABEL_FLOATS (2 + BINS)

float ppo_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x; /* inf / nan */
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    u &= 0xFFFF0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

static double q(double v, int quant) { return quant ? (double)ppo_bf16((float)v) : v; }

/* digamma: recurrence to x >= 6, then the asymptotic series (Abramowitz &
 * Stegun 6.3.18), in double */
double ppo_digamma(double x) {
    double r = 0.0;
    while (x < 6.0) {
        r -= 1.0 / x;
        x += 1.0;
    }
    const double f = 1.0 / (x * x);
    return r + log(x) - 0.5 / x -
           f * (1.0 / 12 - f * (1.0 / 120 - f * (1.0 / 252 - f * (1.0 / 240 - f / 132))));
}

static double sigm(double z) { return 1.0 / (1.0 + exp(-z)); }

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Per-sample loss and d loss / d y (not divided by the batch size). */
void ppo_loss_head(const double* y, int n_out, const float* labels, int n, double* loss,
                   double* dy) {
    const double dx = TMAX / BINS;
    for (int b = 0; b < n; ++b) {
        const double* yb = y + (size_t)b * n_out;
        const float* lab = labels + (size_t)b * LABEL_FLOATS;
        const double n_mc = lab[0], w = lab[1];
        const double y0 = clampd(yb[0], -60.0, 10.0);
        const double lam = n_mc * exp(y0);
        double lb = lam - w * (log(n_mc) + y0);
        double g[N_OUT] = {0};
        g[0] = (yb[0] >= -60.0 && yb[0] <= 10.0) ? lam - w : 0.0;
        double amax = fmax(fmax(yb[1], yb[2]), yb[3]);
        double lse = amax + log(exp(yb[1] - amax) + exp(yb[2] - amax) + exp(yb[3] - amax));
        double la[3], k[3], lth[3], th[3], lgk[3], psi[3], gl[3] = {0, 0, 0}, gs[3] = {0, 0, 0},
                                                         gc[3] = {0, 0, 0};
        for (int j = 0; j < 3; ++j) {
            la[j] = yb[1 + j] - lse;
            k[j] = exp(clampd(yb[4 + j], -4.0, 6.0));
            lth[j] = clampd(yb[7 + j], -4.0, 6.0) + 4.0;
            th[j] = exp(lth[j]);
            lgk[j] = lgamma(k[j]);
            psi[j] = ppo_digamma(k[j]);
        }
        for (int i = 0; i < BINS; ++i) {
            const double h = lab[2 + i];
            if (h == 0.0) continue;
            const double x = (i + 0.5) * dx, lx = log(x);
            double l[3];
            for (int j = 0; j < 3; ++j)
                l[j] = la[j] + (k[j] - 1.0) * lx - x / th[j] - k[j] * lth[j] - lgk[j];
            const double lm = fmax(fmax(l[0], l[1]), l[2]);
            const double lp = lm + log(exp(l[0] - lm) + exp(l[1] - lm) + exp(l[2] - lm));
            lb -= h * lp;
            for (int j = 0; j < 3; ++j) {
                const double r = exp(l[j] - lp);
                gl[j] -= h * r;
                gs[j] -= h * r * k[j] * (lx - lth[j] - psi[j]);
                gc[j] -= h * r * (x / th[j] - k[j]);
            }
        }
        const double gsum = gl[0] + gl[1] + gl[2];
        for (int j = 0; j < 3; ++j) {
            g[1 + j] = gl[j] - exp(la[j]) * gsum;
            g[4 + j] = (yb[4 + j] < -4.0 || yb[4 + j] > 6.0) ? 0.0 : gs[j];
            g[7 + j] = (yb[7 + j] < -4.0 || yb[7 + j] > 6.0) ? 0.0 : gc[j];
        }
        if (loss) loss[b] = lb;
        if (dy)
            for (int o = 0; o < n_out; ++o) dy[(size_t)b * n_out + o] = o < N_OUT ? g[o] : 0.0;
    }
}

/* packed parameter offsets: W_l (dims[l+1] x dims[l]) then b_l */
static void offsets(const int* dims, int n_mat, int64_t* wo, int64_t* bo) {
    int64_t o = 0;
    for (int l = 0; l < n_mat; ++l) {
        wo[l] = o;
        o += (int64_t)dims[l + 1] * dims[l];
        bo[l] = o;
        o += dims[l + 1];
    }
}

/* Forward pass; keeps the activations (a[l], rows x dims[l]) and
 * pre-activations (z[l]) for the backward pass when they are not NULL. */
static void forward(const int* dims, int n_mat, const float* params, const float* x, int n,
                    int quant, double** a, double** z, double* y) {
    int64_t wo[32], bo[32];
    offsets(dims, n_mat, wo, bo);
    for (int64_t i = 0; i < (int64_t)n * dims[0]; ++i) a[0][i] = q(x[i], quant);
    for (int l = 0; l < n_mat; ++l) {
        const int fi = dims[l], fo = dims[l + 1];
        const float* W = params + wo[l];
        const float* bias = params + bo[l];
        for (int r = 0; r < n; ++r) {
            const double* in = a[l] + (size_t)r * fi;
            for (int o = 0; o < fo; ++o) {
                double s = 0.0;
                for (int i = 0; i < fi; ++i) s += q(W[(size_t)o * fi + i], quant) * in[i];
                s += bias[o];
                if (l + 1 < n_mat) {
                    z[l + 1][(size_t)r * fo + o] = s;
                    a[l + 1][(size_t)r * fo + o] = q((float)(s * sigm(s)), quant);
                } else {
                    y[(size_t)r * fo + o] = s;
                }
            }
        }
    }
}

static double** alloc_layers(const int* dims, int n_mat, int n) {
    double** p = calloc((size_t)n_mat + 1, sizeof(double*));
    for (int l = 0; l <= n_mat; ++l) p[l] = calloc((size_t)n * dims[l], sizeof(double));
    return p;
}

static void free_layers(double** p, int n_mat) {
    for (int l = 0; l <= n_mat; ++l) free(p[l]);
    free(p);
}

int ppo_mlp_forward(const int* dims, int n_mat, const float* params, const float* x, int n,
                    int quant, double* y) {
    if (n_mat < 1 || n_mat > 31 || n < 1) return -1;
    double** a = alloc_layers(dims, n_mat, n);
    double** z = alloc_layers(dims, n_mat, n);
    forward(dims, n_mat, params, x, n, quant, a, z, y);
    free_layers(a, n_mat);
    free_layers(z, n_mat);
    return 0;
}

/* Mean loss over the n samples and its gradient (packed layout). */
int ppo_mlp_loss_grad(const int* dims, int n_mat, const float* params, const float* x,
                      const float* labels, int n, int quant, double* loss, double* grads) {
    if (n_mat < 1 || n_mat > 31 || n < 1) return -1;
    int64_t wo[32], bo[32];
    offsets(dims, n_mat, wo, bo);
    const int n_out = dims[n_mat];
    double** a = alloc_layers(dims, n_mat, n);
    double** z = alloc_layers(dims, n_mat, n);
    double** d = alloc_layers(dims, n_mat, n); /* d[l]: gradient at layer l's output */
    double* y = calloc((size_t)n * n_out, sizeof(double));
    double* ls = calloc((size_t)n, sizeof(double));
    forward(dims, n_mat, params, x, n, quant, a, z, y);
    ppo_loss_head(y, n_out, labels, n, ls, d[n_mat]);
    double tot = 0.0;
    for (int b = 0; b < n; ++b) tot += ls[b];
    *loss = tot / n;
    for (int64_t i = 0; i < (int64_t)n * n_out; ++i) d[n_mat][i] = q(d[n_mat][i] / n, quant);
    for (int l = n_mat - 1; l >= 0; --l) {
        const int fi = dims[l], fo = dims[l + 1];
        const float* W = params + wo[l];
        double* gW = grads + wo[l];
        double* gb = grads + bo[l];
        for (int o = 0; o < fo; ++o) {
            double sb = 0.0;
            for (int r = 0; r < n; ++r) sb += d[l + 1][(size_t)r * fo + o];
            gb[o] = sb;
            for (int i = 0; i < fi; ++i) {
                double s = 0.0;
                for (int r = 0; r < n; ++r)
                    s += d[l + 1][(size_t)r * fo + o] * a[l][(size_t)r * fi + i];
                gW[(size_t)o * fi + i] = s;
            }
        }
        if (l == 0) break;
        for (int r = 0; r < n; ++r)
            for (int i = 0; i < fi; ++i) {
                double s = 0.0;
                for (int o = 0; o < fo; ++o)
                    s += d[l + 1][(size_t)r * fo + o] * q(W[(size_t)o * fi + i], quant);
                const double zz = z[l][(size_t)r * fi + i], sg = sigm(zz);
                d[l][(size_t)r * fi + i] = q((float)(s * (sg * (1.0 + zz * (1.0 - sg)))), quant);
            }
    }
    free(y);
    free(ls);
    free_layers(a, n_mat);
    free_layers(z, n_mat);
    free_layers(d, n_mat);
    return 0;
}

/* One Adam(W) step, in double, rounded to float (t = steps taken before). */
void ppo_adam(float* p, const float* g, float* m, float* v, int64_t n, int t, double lr,
              double b1, double b2, double eps, double wd) {
    const double c1 = 1.0 - pow(b1, t + 1), c2 = 1.0 - pow(b2, t + 1);
    for (int64_t i = 0; i < n; ++i) {
        const double mm = b1 * m[i] + (1.0 - b1) * g[i];
        const double vv = b2 * v[i] + (1.0 - b2) * (double)g[i] * g[i];
        m[i] = (float)mm;
        v[i] = (float)vv;
        p[i] = (float)(p[i] - lr * ((mm / c1) / (sqrt(vv / c2) + eps) + wd * p[i]));
    }
}

/* Network features of (cascade, module) pairs; cascade rows are
 * (x, y, z, dx, dy, dz, energy, log_yield). */
void ppo_sur_features(const float* cascades, const float* modules, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        const float* c = cascades + 8 * i;
        const double dx = modules[3 * i] - c[0], dy = modules[3 * i + 1] - c[1],
                     dz = modules[3 * i + 2] - c[2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double cs = (c[3] * dx + c[4] * dy + c[5] * dz) / fmax(r, 1e-6);
        cs = clampd(cs, -1.0, 1.0);
        const double sn = sqrt(fmax(0.0, (1.0 - cs) * (1.0 + cs)));
        const double lr = (log10(fmax(r, 1e-3)) - 1.2) / 0.7;
        float* f = out + 8 * i;
        f[0] = (float)lr;
        f[1] = (float)cs;
        f[2] = (float)sn;
        f[3] = (float)((log10(fmax((double)c[6], 1e-3)) - 4.0) / 1.2);
        f[4] = (float)(r * cs * 0.01);
        f[5] = (float)(r * sn * 0.01);
        f[6] = (float)(0.5 * log10(r * sn + 1.0));
        f[7] = (float)(cs * lr);
    }
}
