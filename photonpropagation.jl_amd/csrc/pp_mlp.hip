// pp_mlp.hip - the surrogate MLP of libphotonprop_hip.so (include/photonprop_surrogate.h).
//
// Every matrix product is one kernel, gemm_kernel<A_COL, B_KMAJ, EPI>:
// 64x64 output tiles per 256-thread workgroup, each wave one 32x32 tile of
// v_mfma_f32_32x32x16_bf16 (bf16 operands, fp32 accumulation), operand tiles
// staged through LDS as [row][k] with k contiguous so that a lane reads its 8
// k-values with one 16-byte LDS load. Operands in either orientation are
// transposed while being staged, so one kernel serves the forward product
// (H W^T), the data gradient (dZ W) and the weight gradient (dZ^T H). The
// epilogues fuse bias + SiLU (forward, writing the fp32 pre-activation and
// the bf16 activation), bias only (output layer), and the SiLU derivative
// (backward, writing bf16 dZ).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "photonprop_hip.h"
#include "photonprop_surrogate.h"
#include "pp_device.h"
#include "pp_hits.h"

// pp_device.h turns contraction off for the propagation step functions; the
// network's epilogues (bias, SiLU, Adam, loss) contract products and sums into
// fused multiply-adds, as hipcc does by default: their accuracy is set by
// the bf16 operands, not by the rounding of these sums.
#pragma clang fp contract(fast)

// error text shared with the Monte Carlo half of the library (pp_last_error)
void pp_set_error_message(const char* msg);

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 64, kBK = 32, kLdsK = kBK + 8;

enum Epi { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

struct GemmArgs {
    const bf16* A;  // A(m,k) = A_COL ? A[k*lda + m] : A[m*lda + k]
    int lda;
    const bf16* B;  // B(k,n) = B_KMAJ ? B[n*ldb + k] : B[k*ldb + n]
    int ldb;
    int M, N, K;
    float* C;  // fp32 output: EPI_F32, EPI_BIAS_F32, pre-activation of EPI_BIAS_SILU
    int ldc;
    bf16* H;  // bf16 output: SiLU(z) of EPI_BIAS_SILU, dz of EPI_DSILU
    int ldh;
    const float* bias;
    const float* Z;  // pre-activation read by EPI_DSILU
    int ldz;
    // split K: blockIdx.z covers K columns [z*k_chunk, (z+1)*k_chunk) and
    // writes its partial product (EPI_F32 only) to C + z*c_stride
    int k_chunk;
    int64_t c_stride;
};

// logistic function from the hardware exp2 and reciprocal (each about 1 ulp);
// the activations it feeds are rounded to bf16 (8-bit significand)
__device__ __forceinline__ float sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.44269504f));
}

// 8 consecutive bf16 at p[0..7] (zero where i >= valid) into dst[0..7]
__device__ __forceinline__ void load8(const bf16* p, int valid, bool vec, bf16* dst) {
    if (vec && valid >= 8) {
        *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(p);
    } else {
        for (int j = 0; j < 8; ++j) dst[j] = j < valid ? p[j] : (bf16)0.0f;
    }
}

// Stage a 64 x 32 tile of a [rows x K] operand into lds[row][k].
//   KCONTIG: element (r, k) at X[r*ld + k]; else at X[k*ld + r].
template <bool KCONTIG>
__device__ __forceinline__ void stage_tile(const bf16* X, int ld, int r0, int k0, int rows, int K,
                                           bf16* lds) {
    const int t = threadIdx.x;
    const bool vec = (ld & 7) == 0;
    if (KCONTIG) {
        const int r = t >> 2, kc = (t & 3) * 8;
        alignas(16) bf16 v[8];
        const int gr = r0 + r, gk = k0 + kc;
        const int left = K - gk;
        const int valid = (gr < rows) ? (left < 0 ? 0 : (left > 8 ? 8 : left)) : 0;
        load8(X + (size_t)gr * ld + gk, valid, vec, v);
        *reinterpret_cast<uint4*>(&lds[r * kLdsK + kc]) = *reinterpret_cast<const uint4*>(v);
    } else {
        const int kk = t >> 3, rc = (t & 7) * 8;
        alignas(16) bf16 v[8];
        const int gk = k0 + kk, gr = r0 + rc;
        const int left = rows - gr;
        const int valid = (gk < K) ? (left < 0 ? 0 : (left > 8 ? 8 : left)) : 0;
        load8(X + (size_t)gk * ld + gr, valid, vec, v);
        for (int j = 0; j < 8; ++j) lds[(rc + j) * kLdsK + kk] = v[j];
    }
}

template <bool A_COL, bool B_KMAJ, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) bf16 As[kBM * kLdsK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[kBN * kLdsK];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    // acc[r] holds row (r/4)*8 + (lane/32)*4 + r%4, column lane%32 of the wave's
    // tile; a bias epilogue starts the sum from the column's bias (the first
    // product adds to it, as in mlp_fused_kernel)
    const int n = n0 + wn + (lane & 31);
    const float bias0 =
        (EPI == EPI_BIAS_F32 || EPI == EPI_BIAS_SILU) && n < g.N ? g.bias[n] : 0.0f;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = bias0;
    const int k_lo = blockIdx.z * g.k_chunk;
    const int k_hi = k_lo + g.k_chunk < g.K ? k_lo + g.k_chunk : g.K;
    float* const C = g.C ? g.C + blockIdx.z * g.c_stride : nullptr;
    for (int k0 = k_lo; k0 < k_hi; k0 += kBK) {
        stage_tile<!A_COL>(g.A, g.lda, m0, k0, g.M, g.K, As);
        stage_tile<B_KMAJ>(g.B, g.ldb, n0, k0, g.N, g.K, Bs);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wm + (lane & 31)) * kLdsK + kq]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wn + (lane & 31)) * kLdsK + kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    if (n >= g.N) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (m >= g.M) continue;
        const float v = acc[r];
        if (EPI == EPI_F32) {
            C[(size_t)m * g.ldc + n] = v;
        } else if (EPI == EPI_BIAS_F32) {
            C[(size_t)m * g.ldc + n] = v;
        } else if (EPI == EPI_BIAS_SILU) {
            const float z = v;
            if (C) C[(size_t)m * g.ldc + n] = z;
            g.H[(size_t)m * g.ldh + n] = (bf16)(z * sigmoid(z));
        } else {
            const float z = g.Z[(size_t)m * g.ldz + n];
            const float s = sigmoid(z);
            g.H[(size_t)m * g.ldh + n] = (bf16)(v * (s * (1.0f + z * (1.0f - s))));
        }
    }
}

// Column sums of X (rows x cols, bf16) in row slices: blockIdx.y sums rows
// [y*row_chunk, (y+1)*row_chunk) into part[y*cols + c] (bias gradients,
// finished by reduce_partials_kernel in a fixed order).
__global__ void __launch_bounds__(256) colsum_kernel(const bf16* X, int ld, int rows, int cols,
                                                     int row_chunk, float* part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), r0 = threadIdx.x >> 6;
    const int lo = blockIdx.y * row_chunk;
    const int hi = lo + row_chunk < rows ? lo + row_chunk : rows;
    float s = 0.0f;
    if (c < cols)
        for (int r = lo + r0; r < hi; r += 4) s += (float)X[(size_t)r * ld + c];
    red[r0][threadIdx.x & 63] = s;
    __syncthreads();
    if (r0 == 0 && c < cols)
        part[(size_t)blockIdx.y * cols + c] =
            (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// out[i] = sum over s < S of part[s*n + i], in order s = 0..S-1
__global__ void __launch_bounds__(256) reduce_partials_kernel(const float* part, int S, int64_t n,
                                                              float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += part[k * n + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// loss head
// ---------------------------------------------------------------------------
// digamma(x), x > 0: recurrence up to x >= 6, then the asymptotic series
// psi(x) = ln x - 1/(2x) - sum B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18)
__device__ __forceinline__ float digamma_f(float x) {
    float r = 0.0f;
    while (x < 6.0f) {
        r -= 1.0f / x;
        x += 1.0f;
    }
    const float f = 1.0f / (x * x);
    return r + logf(x) - 0.5f / x -
           f * (1.0f / 12 - f * (1.0f / 120 - f * (1.0f / 252 - f * (1.0f / 240 - f / 132))));
}

constexpr float kShapeLo = -4.0f, kShapeHi = 6.0f;  // log shape range
constexpr float kScaleLo = -4.0f, kScaleHi = 6.0f;  // log scale - 4 range
constexpr float kY0Lo = -60.0f, kY0Hi = 10.0f;

struct LossArgs {
    const float* Y;  // rows x n_out
    int n_out;
    const float* L;  // rows x PP_SUR_LABEL_FLOATS
    int rows;
    float inv_rows;
    float* dY;   // rows x n_out (fp32), may be null
    bf16* dYb;   // rows x n_out (bf16), may be null
    float* loss; // accumulated sum of per-sample losses x inv_rows
    const uint32_t* step;  // if set, loss[*step] is the target
};

// Per-sample loss and its gradient w.r.t. the outputs (see the header).
__global__ void __launch_bounds__(256) loss_head_kernel(LossArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float lb = 0.0f;
    if (b < a.rows) {
        const float* y = a.Y + (size_t)b * a.n_out;
        const float* lab = a.L + (size_t)b * PP_SUR_LABEL_FLOATS;
        const float n_mc = lab[0], w = lab[1];
        const float y0 = fminf(fmaxf(y[0], kY0Lo), kY0Hi);
        const float lam = n_mc * __expf(y0);
        lb = lam - w * (logf(n_mc) + y0);
        float g0 = (y[0] >= kY0Lo && y[0] <= kY0Hi) ? lam - w : 0.0f;
        // mixture
        float la[3], k[3], lth[3], th[3], lgk[3], psi[3], ga[3] = {0, 0, 0}, gs[3] = {0, 0, 0},
                                                           gc[3] = {0, 0, 0};
        const float amax = fmaxf(fmaxf(y[1], y[2]), y[3]);
        const float lse = amax + logf(__expf(y[1] - amax) + __expf(y[2] - amax) + __expf(y[3] - amax));
        for (int j = 0; j < 3; ++j) {
            la[j] = y[1 + j] - lse;
            k[j] = __expf(fminf(fmaxf(y[4 + j], kShapeLo), kShapeHi));
            lth[j] = fminf(fmaxf(y[7 + j], kScaleLo), kScaleHi) + 4.0f;
            th[j] = __expf(lth[j]);
            lgk[j] = lgammaf(k[j]);
            psi[j] = digamma_f(k[j]);
        }
        const float dx = PP_SUR_TMAX / PP_SUR_BINS;
        float gl_sum = 0.0f, gl[3] = {0, 0, 0};
        for (int i = 0; i < PP_SUR_BINS; ++i) {
            const float h = lab[2 + i];
            if (h == 0.0f) continue;
            const float x = (i + 0.5f) * dx, lx = logf(x);
            float l[3];
            for (int j = 0; j < 3; ++j)
                l[j] = la[j] + (k[j] - 1.0f) * lx - x / th[j] - k[j] * lth[j] - lgk[j];
            const float lm = fmaxf(fmaxf(l[0], l[1]), l[2]);
            const float lp = lm + logf(__expf(l[0] - lm) + __expf(l[1] - lm) + __expf(l[2] - lm));
            lb -= h * lp;
            for (int j = 0; j < 3; ++j) {
                const float r = __expf(l[j] - lp);
                gl[j] -= h * r;
                gs[j] -= h * r * k[j] * (lx - lth[j] - psi[j]);
                gc[j] -= h * r * (x / th[j] - k[j]);
            }
        }
        gl_sum = gl[0] + gl[1] + gl[2];
        for (int j = 0; j < 3; ++j) {
            ga[j] = gl[j] - __expf(la[j]) * gl_sum;
            if (y[4 + j] < kShapeLo || y[4 + j] > kShapeHi) gs[j] = 0.0f;
            if (y[7 + j] < kScaleLo || y[7 + j] > kScaleHi) gc[j] = 0.0f;
        }
        const float gout[PP_SUR_N_OUT] = {g0, ga[0], ga[1], ga[2], gs[0], gs[1], gs[2],
                                          gc[0], gc[1], gc[2]};
        for (int o = 0; o < a.n_out; ++o) {
            const float gv = o < PP_SUR_N_OUT ? gout[o] * a.inv_rows : 0.0f;
            if (a.dY) a.dY[(size_t)b * a.n_out + o] = gv;
            if (a.dYb) a.dYb[(size_t)b * a.n_out + o] = (bf16)gv;
        }
    }
    // block sum of the losses, one atomic per block
    __shared__ float red[256];
    red[threadIdx.x] = lb * a.inv_rows;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(a.step ? a.loss + *a.step : a.loss, red[0]);
}

// ---------------------------------------------------------------------------
// optimiser, batches, features, inference head
// ---------------------------------------------------------------------------
struct AdamArgs {
    float* p;
    const float* g;
    float* m;
    float* v;
    bf16* pb;
    int64_t n;
    pp_adam o;
    const uint32_t* step;  // number of steps taken before this one
};

__global__ void __launch_bounds__(256) adam_kernel(AdamArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float t = (float)(*a.step + 1);
    const float g = a.g[i];
    const float m = a.o.beta1 * a.m[i] + (1.0f - a.o.beta1) * g;
    const float v = a.o.beta2 * a.v[i] + (1.0f - a.o.beta2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float mh = m / (1.0f - powf(a.o.beta1, t));
    const float vh = v / (1.0f - powf(a.o.beta2, t));
    float p = a.p[i];
    p -= a.o.lr * (mh / (sqrtf(vh) + a.o.eps) + a.o.weight_decay * p);
    a.p[i] = p;
    a.pb[i] = (bf16)p;
}

__global__ void bump_kernel(uint32_t* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// Weights of one layer in the order the fused inference kernel loads them:
// for 32-column tile T and 16-deep K fragment i, the 64 lanes' 8 bf16 each,
// contiguous: chunk (T*nf + i)*64 + lane holds W[32T + lane%32][16i +
// (lane/32)*8 + j], j = 0..7, zero beyond K and N.
__global__ void pack_fused_kernel(const bf16* W, int K, int N, int nf, int64_t chunks,
                                  bf16* out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= chunks) return;
    const int lane = (int)(c & 63);
    const int64_t ti = c >> 6;
    const int i = (int)(ti % nf), T = (int)(ti / nf);
    const int n = 32 * T + (lane & 31), k0 = 16 * i + (lane >> 5) * 8;
    for (int j = 0; j < 8; ++j)
        out[c * 8 + j] = (n < N && k0 + j < K) ? W[(size_t)n * K + k0 + j] : (bf16)0.0f;
}

__global__ void to_bf16_kernel(const float* x, bf16* y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

// Training batch: `rows` samples drawn with replacement, row b of step s from
// Philox block (b, s, 0x5EEDBA7C, 0) under the training key.
struct GatherArgs {
    const float* X;  // n x n_in
    const float* L;  // n x PP_SUR_LABEL_FLOATS
    int64_t n;
    int n_in;
    int rows;
    uint32_t k0, k1;
    const uint32_t* step;
    bf16* Xb;  // rows x n_in
    float* Lb; // rows x PP_SUR_LABEL_FLOATS
};

__global__ void __launch_bounds__(256) gather_kernel(GatherArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.rows) return;
    const uint4 r = pp::philox(make_uint4((uint32_t)b, *a.step, 0x5EEDBA7Cu, 0u), a.k0, a.k1);
    const uint64_t u = ((uint64_t)r.x << 32) | r.y;
    const int64_t idx = (int64_t)__umul64hi(u, (uint64_t)a.n);
    for (int j = 0; j < a.n_in; ++j) a.Xb[(size_t)b * a.n_in + j] = (bf16)a.X[idx * a.n_in + j];
    for (int j = 0; j < PP_SUR_LABEL_FLOATS; ++j)
        a.Lb[(size_t)b * PP_SUR_LABEL_FLOATS + j] = a.L[idx * PP_SUR_LABEL_FLOATS + j];
}

// Features of the pair (cascade c, module at m): r = |m - p|, cos = dir . (m - p)/r,
// z = r cos, rho = r sin, lE = log10 E:
//   (log10 r - 1.2)/0.7, cos, sin, (lE - 4)/1.2, z/100, rho/100,
//   log10(rho + 1)/2, cos (log10 r - 1.2)/0.7
__device__ __forceinline__ void pair_features(const pp_cascade& c, float mx, float my, float mz,
                                              float* f, float& r_out) {
    const float dx = mx - c.position[0], dy = my - c.position[1], dz = mz - c.position[2];
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    const float ir = 1.0f / fmaxf(r, 1e-6f);
    const float cs = fminf(fmaxf((c.direction[0] * dx + c.direction[1] * dy + c.direction[2] * dz) * ir,
                                 -1.0f), 1.0f);
    const float sn = sqrtf(fmaxf(0.0f, (1.0f - cs) * (1.0f + cs)));
    const float lr = (log10f(fmaxf(r, 1e-3f)) - 1.2f) / 0.7f;
    f[0] = lr;
    f[1] = cs;
    f[2] = sn;
    f[3] = (log10f(fmaxf(c.energy, 1e-3f)) - 4.0f) / 1.2f;
    f[4] = r * cs * 0.01f;
    f[5] = r * sn * 0.01f;
    f[6] = 0.5f * log10f(r * sn + 1.0f);
    f[7] = cs * lr;
    r_out = r;
}

struct PairArgs {
    const pp_cascade* ev;  // per event (eval_pairs) or per pair (features)
    const float* mod;      // module xyz: per module (eval_pairs) or per pair (features)
    int n_mod;             // eval_pairs: modules per event; 0 = one cascade and module per row
    int64_t row0, rows;
    int n_in;
    float* xf;   // rows x n_in fp32 (may be null)
    bf16* xb;    // rows x n_in bf16 (may be null)
    float* logy; // rows: log yield (may be null)
    uint8_t* fb; // rows: fallback flags (may be null)
    pp_sur_domain dom;
};

// Row `row` of a pair list: its features, the log of yield x geometric
// exposure (R/2r)^2 (expected hits = that x exp(y0)) and the fallback flag
// (the pair outside the network's domain).
__device__ __forceinline__ void pair_row(const PairArgs& a, int64_t row, float* f, float& logy,
                                         bool& fb) {
    const int64_t e = a.n_mod ? row / a.n_mod : row;
    const int64_t j = a.n_mod ? row % a.n_mod : row;
    const pp_cascade c = a.ev[e];
    float r;
    pair_features(c, a.mod[3 * j], a.mod[3 * j + 1], a.mod[3 * j + 2], f, r);
    logy = c.log_yield + 2.0f * logf(a.dom.module_radius / (2.0f * fmaxf(r, a.dom.module_radius)));
    fb = !(r >= a.dom.min_dist && r <= a.dom.max_dist && c.energy >= a.dom.min_energy &&
           c.energy <= a.dom.max_energy);
}

__global__ void __launch_bounds__(256) pair_features_kernel(PairArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.rows) return;
    float f[PP_SUR_N_FEATURES], logy;
    bool fb;
    pair_row(a, a.row0 + i, f, logy, fb);
    for (int k = 0; k < a.n_in; ++k) {
        const float v = k < PP_SUR_N_FEATURES ? f[k] : 0.0f;
        if (a.xf) a.xf[i * a.n_in + k] = v;
        if (a.xb) a.xb[i * a.n_in + k] = (bf16)v;
    }
    if (a.logy) a.logy[i] = logy;
    if (a.fb) a.fb[i] = fb;
}

// Inference head of one row: expected hits = exp(y0 + log yield x exposure);
// mixture (weights, shapes, scales in ns) into mix[0..8] when mix is set.
__device__ __forceinline__ void sur_head_row(const float* y, float logy, float* hits, float* mix) {
    *hits = __expf(fminf(fmaxf(y[0], kY0Lo), kY0Hi) + logy);
    if (!mix) return;
    const float amax = fmaxf(fmaxf(y[1], y[2]), y[3]);
    float e[3], s = 0.0f;
    for (int j = 0; j < 3; ++j) s += (e[j] = __expf(y[1 + j] - amax));
    for (int j = 0; j < 3; ++j) {
        mix[j] = e[j] / s;
        mix[3 + j] = __expf(fminf(fmaxf(y[4 + j], kShapeLo), kShapeHi));
        mix[6 + j] = __expf(fminf(fmaxf(y[7 + j], kScaleLo), kScaleHi) + 4.0f);
    }
}

__global__ void __launch_bounds__(256) sur_head_kernel(const float* Y, int n_out, int64_t rows,
                                                       const float* logy, float* hits,
                                                       float* mix) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    sur_head_row(Y + i * n_out, logy[i], hits + i, mix ? mix + i * 9 : nullptr);
}

// ---------------------------------------------------------------------------
// Fused inference: the whole network for a tile of kFRows rows per
// workgroup, the activations resident in LDS between layers (no HBM round
// trip). The products are taken transposed, C^T = W . act^T: the weights are
// MFMA operand A, held in registers and loaded straight from L2 in the packed
// order of pack_fused_kernel (A row = output column n = lane%32 of the tile,
// k = 16i + (lane/32)*8 + j; each wave load is 1 KB contiguous), the
// activations operand B, read from LDS (B column = row m = lane%32). Wave w
// owns output columns [64w, 64w+64) (two 32-column tiles) of all kFRows rows
// (four 32-row tiles): each weight fragment feeds four MFMAs and each
// activation fragment two, and a lane's four consecutive accumulator entries
// are four consecutive columns of one row, written with one 8-byte LDS store.
// One activation buffer serves as a layer's input and output (a barrier
// between the last read and the first write). Each output element is the
// same 16-term MFMA dot products, K ascending with zero padding to a
// multiple of 32, and the same bias/SiLU/bf16 epilogue as gemm_kernel.
// Shapes: hidden width kFH, input and output widths <= 32, <= kFMaxLayers
// weight matrices.
// ---------------------------------------------------------------------------
constexpr int kFH = 256;            // hidden width
constexpr int kFRows = 128;         // rows per workgroup (four 32-row MFMA tiles)
constexpr int kFThreads = 256;      // four waves
constexpr int kFActLd = kFH + 8;    // LDS row stride of the activations (bf16)
constexpr int kFFrag = kFH / 16;    // 16-deep K fragments of a hidden layer
constexpr int kFMaxLayers = 8;
static_assert(kFThreads == kFH, "one thread per bias entry of a layer");

struct FusedArgs {
    const bf16* X;   // rows x n_in (bf16); unused for pairs
    int64_t rows;
    int n_in, n_out, n_mat;
    const bf16* W[kFMaxLayers];    // layer l packed by pack_fused_kernel
    const float* b[kFMaxLayers];
    float* Y;        // rows x n_out (fp32); unused for pairs
    // pairs: the rows are the (event, module) pairs of pr (ev, mod, n_mod,
    // dom), whose features are computed in the kernel; pr.fb receives the
    // fallback flags, and the inference head writes hits and, when set, mix
    PairArgs pr;
    float* hits;
    float* mix;
};

enum FusedLayer { kFInput = 0, kFHidden = 1, kFOutput = 2 };

// One layer for wave w: products over K in batches of eight weight
// fragments, barrier (every wave's reads of the layer input done), epilogue
// into `act` (bf16 SiLU activations, row stride kFActLd) or, for the output
// layer, fp32 outputs into `act` viewed as float[kFRows][32], barrier. In the
// output layer (N <= 32, one column tile) wave w takes row tile w.
template <int MODE>
__device__ __forceinline__ void fused_layer(const FusedArgs& f, int l, bf16* act,
                                            const float* bias_l, int lane, int w) {
    constexpr int NT = MODE == kFOutput ? 1 : 2;          // 32-column tiles per wave
    constexpr int MT = MODE == kFOutput ? 1 : 4;          // 32-row tiles per wave
    constexpr int NF = MODE == kFInput ? 2 : kFFrag;      // K fragments
    constexpr int NB = NF < 8 ? NF : 8;                   // fragments per load batch
    const int kl = (lane >> 5) * 8;
    const int t0 = MODE == kFOutput ? 0 : 2 * w;          // first column tile
    const int m0 = MODE == kFOutput ? 32 * w : 0;         // first row
    const bf16x8* wsrc = reinterpret_cast<const bf16x8*>(f.W[l]) + (size_t)t0 * NF * 64 + lane;
    const bf16* bsrc = act + (m0 + (lane & 31)) * kFActLd + kl;
    // the sums start from the biases: acc[nt][mt][4g + j] belongs to column
    // 32(t0 + nt) + 8g + (lane/32)*4 + j of any row tile, so one block of
    // biases per column tile is the first product's accumulator input
    f32x16 bb[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 b4 = *reinterpret_cast<const float4*>(bias_l + 32 * (t0 + nt) + 8 * g +
                                                               (lane >> 5) * 4);
            bb[nt][4 * g] = b4.x;
            bb[nt][4 * g + 1] = b4.y;
            bb[nt][4 * g + 2] = b4.z;
            bb[nt][4 * g + 3] = b4.w;
        }
    f32x16 acc[NT][MT];
    // one batch of NB weight fragments from fragment i0; FIRST: the batch at
    // fragment 0, whose first products take the bias blocks
    auto batch = [&](int i0, auto first) {
        constexpr bool FIRST = decltype(first)::value;
        bf16x8 wa[NT][NB];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int i = 0; i < NB; ++i) wa[nt][i] = wsrc[(nt * NF + i0 + i) * 64];
        // the batch's loads all in flight before the first product waits on one
        __builtin_amdgcn_sched_barrier(0);
        // issue priority to the waves feeding the matrix cores, so that the
        // other workgroup's epilogue fills the vector slots they leave
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            bf16x8 b[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                b[mt] = *reinterpret_cast<const bf16x8*>(bsrc + 32 * mt * kFActLd + 16 * (i0 + i));
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[nt][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        wa[nt][i], b[mt], (FIRST && i == 0) ? bb[nt] : acc[nt][mt], 0, 0, 0);
        }
        __builtin_amdgcn_s_setprio(0);
    };
    batch(0, std::true_type{});
#pragma unroll 1
    for (int i0 = NB; i0 < NF; i0 += NB) batch(i0, std::false_type{});
    __syncthreads();
    // acc[nt][mt][4g + j]: column 32(t0 + nt) + 8g + (lane/32)*4 + j, row m0 + 32mt + lane%32
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int n0 = 32 * (t0 + nt) + 8 * g + (lane >> 5) * 4;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int m = m0 + 32 * mt + (lane & 31);
                float z[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) z[j] = acc[nt][mt][4 * g + j];
                if (MODE == kFOutput) {
                    *reinterpret_cast<float4*>(reinterpret_cast<float*>(act) + m * 32 + n0) =
                        make_float4(z[0], z[1], z[2], z[3]);
                } else {
                    bf16 h[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) h[j] = (bf16)(z[j] * sigmoid(z[j]));
                    *reinterpret_cast<uint2*>(act + m * kFActLd + n0) =
                        *reinterpret_cast<const uint2*>(h);
                }
            }
        }
    }
    __syncthreads();
}

// two workgroups per CU: LDS 76.3 KB each, <= 256 VGPRs per lane.
// PAIRS: the input tile is computed from (event, module) pairs (pair_row, as
// pair_features_kernel) and the outputs go through the inference head
// (sur_head_row, as sur_head_kernel), so that a pair list is evaluated in one
// launch with the same values as the three-kernel sequence.
template <bool PAIRS>
__global__ void __launch_bounds__(kFThreads, 2) mlp_fused_kernel(FusedArgs f) {
    __shared__ __attribute__((aligned(16))) bf16 act[kFRows * kFActLd];
    __shared__ __attribute__((aligned(16))) float bias[kFMaxLayers][kFH];
    __shared__ float logy[PAIRS ? kFRows : 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t row0 = (int64_t)blockIdx.x * kFRows;
    // every layer's biases, zero beyond the layer's width
    for (int l = 0; l < f.n_mat; ++l)
        bias[l][t] = t < (l + 1 == f.n_mat ? f.n_out : kFH) ? f.b[l][t] : 0.0f;
    // input tile, zero beyond n_in (up to 32 columns) and beyond the last row
    if (PAIRS) {
        if (t < kFRows) {
            bf16* a = act + t * kFActLd;
            const int64_t row = row0 + t;
            if (row < f.rows) {
                float feat[PP_SUR_N_FEATURES];
                bool fb;
                pair_row(f.pr, row, feat, logy[t], fb);
                f.pr.fb[row] = fb;
                for (int k = 0; k < 32; ++k)
                    a[k] = k < f.n_in ? (bf16)(k < PP_SUR_N_FEATURES ? feat[k] : 0.0f) : (bf16)0.0f;
            } else {
                for (int k = 0; k < 32; ++k) a[k] = (bf16)0.0f;
            }
        }
    } else {
        for (int i = t; i < kFRows * 32; i += kFThreads) {
            const int r = i >> 5, k = i & 31;
            const int64_t gr = row0 + r;
            act[r * kFActLd + k] = (gr < f.rows && k < f.n_in) ? f.X[gr * f.n_in + k]
                                                               : (bf16)0.0f;
        }
    }
    __syncthreads();
    fused_layer<kFInput>(f, 0, act, bias[0], lane, w);
    for (int l = 1; l + 1 < f.n_mat; ++l) fused_layer<kFHidden>(f, l, act, bias[l], lane, w);
    fused_layer<kFOutput>(f, f.n_mat - 1, act, bias[f.n_mat - 1], lane, w);
    const float* zs = reinterpret_cast<const float*>(act);
    const int valid_rows = f.rows - row0 < kFRows ? (int)(f.rows - row0) : kFRows;
    if (PAIRS) {
        if (t < valid_rows) {
            const int64_t row = row0 + t;
            sur_head_row(zs + t * 32, logy[t], f.hits + row, f.mix ? f.mix + row * 9 : nullptr);
        }
    } else {
        for (int i = t; i < valid_rows * f.n_out; i += kFThreads) {
            const int r = i / f.n_out, c = i - r * f.n_out;
            f.Y[row0 * f.n_out + i] = zs[r * 32 + c];
        }
    }
}

int fail(int code, const std::string& msg) {
    pp_set_error_message(msg.c_str());
    return code;
}

#define PP_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(PP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

struct pp_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<int> dims;  // n_in, hidden..., n_out
    int n_mat = 0, max_batch = 0;
    // parameters: internal layout with every block at a 64-element boundary
    std::vector<int64_t> w_off, b_off;
    int64_t n_int = 0, n_params = 0;
    bool fused = true;  // inference through mlp_fused_kernel where the shape allows
    // the fused kernel's copy of the bf16 weights (pack_fused_kernel order),
    // refreshed by the first inference after the weights change
    bf16* wpk = nullptr;
    std::vector<int64_t> pk_off;
    bool pk_ok = false;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16* pb = nullptr;
    // activations (max_batch rows): H[l] bf16 with dims[l] columns (H[0] = the
    // input), Z[l] fp32 pre-activations of the hidden layers, Y fp32 outputs
    std::vector<bf16*> H;
    std::vector<float*> Z;
    float* Y = nullptr;
    float* dY = nullptr;
    bf16* dYb = nullptr;
    bf16 *dA = nullptr, *dB = nullptr;  // backward ping-pong
    float* Lb = nullptr;                // batch labels
    float* part = nullptr;              // split-K / column-sum partials
    int64_t part_cap = 0;               // floats
    float* xs = nullptr;                // fp32 staging of host features
    float* loss = nullptr;              // loss accumulator(s)
    int64_t loss_cap = 0;
    uint32_t* step = nullptr;
    // training set
    float *dsX = nullptr, *dsL = nullptr;
    int64_t ds_cap = 0;
    // pair evaluation
    float* logy = nullptr;
    float* hits = nullptr;
    float* mix = nullptr;
    uint8_t* fb = nullptr;
    int64_t out_cap = 0;
    pp_cascade* d_ev = nullptr;
    float* d_mod = nullptr;
    int64_t ev_cap = 0, mod_cap = 0;
};

namespace {

template <bool A_COL, bool B_KMAJ, int EPI>
int gemm(pp_mlp* M, const GemmArgs& a) {
    if (a.M <= 0 || a.N <= 0) return PP_OK;
    GemmArgs b = a;
    if (b.k_chunk <= 0) b.k_chunk = b.K;
    const dim3 grid(blocks_for(b.N, kBN), blocks_for(b.M, kBM), blocks_for(b.K, b.k_chunk));
    hipLaunchKernelGGL((gemm_kernel<A_COL, B_KMAJ, EPI>), grid, dim3(256), 0, M->stream, b);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// The fused inference kernel takes hidden layers of width kFH and input and
// output widths of at most 32.
bool fused_ok(const pp_mlp* M) {
    if (M->n_mat > kFMaxLayers || M->dims.front() > 32 || M->dims.back() > 32) return false;
    for (size_t i = 1; i + 1 < M->dims.size(); ++i)
        if (M->dims[i] != kFH) return false;
    return true;
}

// the fused kernel's packed copy of the bf16 weights, refreshed after a change
int ensure_packed(pp_mlp* M) {
    if (M->pk_ok) return PP_OK;
    for (int l = 0; l < M->n_mat; ++l) {
        const int K = M->dims[l], N = M->dims[l + 1];
        const int nf = (K + 31) / 32 * 2;
        const int64_t chunks = (M->pk_off[l + 1] - M->pk_off[l]) / 8;
        hipLaunchKernelGGL(pack_fused_kernel, dim3(blocks_for(chunks, 256)), dim3(256), 0,
                           M->stream, M->pb + M->w_off[l], K, N, nf, chunks, M->wpk + M->pk_off[l]);
        PP_HIP(hipGetLastError());
    }
    M->pk_ok = true;
    return PP_OK;
}

// The fused kernel over `rows` rows: the bf16 input in H[0] to the fp32
// outputs in Y, or, with pairs set, the pair list's features to the head's
// outputs (hits, mix, pairs->fb).
int forward_fused(pp_mlp* M, int64_t rows, const PairArgs* pairs = nullptr, float* hits = nullptr,
                  float* mix = nullptr) {
    if (int rc = ensure_packed(M)) return rc;
    if (rows <= 0) return PP_OK;
    FusedArgs f{};
    f.X = M->H[0];
    f.rows = rows;
    f.n_in = M->dims.front();
    f.n_out = M->dims.back();
    f.n_mat = M->n_mat;
    for (int l = 0; l < M->n_mat; ++l) {
        f.W[l] = M->wpk + M->pk_off[l];
        f.b[l] = M->p + M->b_off[l];
    }
    f.Y = M->Y;
    const dim3 grid(blocks_for(rows, kFRows));
    if (pairs) {
        f.pr = *pairs;
        f.hits = hits;
        f.mix = mix;
        hipLaunchKernelGGL(mlp_fused_kernel<true>, grid, dim3(kFThreads), 0, M->stream, f);
    } else {
        hipLaunchKernelGGL(mlp_fused_kernel<false>, grid, dim3(kFThreads), 0, M->stream, f);
    }
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// forward of `rows` rows whose bf16 input is in H[0]; keep_z stores the fp32
// pre-activations that the backward pass needs (and uses the layer-by-layer
// kernels); inference runs the fused kernel where the shape allows
int forward(pp_mlp* M, int rows, bool keep_z) {
    if (!keep_z && M->fused && M->wpk) return forward_fused(M, rows);
    for (int l = 0; l < M->n_mat; ++l) {
        GemmArgs a{};
        a.A = M->H[l];
        a.lda = M->dims[l];
        a.B = M->pb + M->w_off[l];
        a.ldb = M->dims[l];
        a.M = rows;
        a.N = M->dims[l + 1];
        a.K = M->dims[l];
        a.bias = M->p + M->b_off[l];
        if (l + 1 < M->n_mat) {
            a.C = keep_z ? M->Z[l + 1] : nullptr;
            a.ldc = a.N;
            a.H = M->H[l + 1];
            a.ldh = a.N;
            if (int rc = gemm<false, true, EPI_BIAS_SILU>(M, a)) return rc;
        } else {
            a.C = M->Y;
            a.ldc = a.N;
            if (int rc = gemm<false, true, EPI_BIAS_F32>(M, a)) return rc;
        }
    }
    return PP_OK;
}

int loss_head(pp_mlp* M, int rows, bool grad, bool per_step) {
    LossArgs a{};
    a.Y = M->Y;
    a.n_out = M->dims.back();
    a.L = M->Lb;
    a.rows = rows;
    a.inv_rows = 1.0f / (float)rows;
    a.dY = grad ? M->dY : nullptr;
    a.dYb = grad ? M->dYb : nullptr;
    a.loss = M->loss;
    a.step = per_step ? M->step : nullptr;
    hipLaunchKernelGGL(loss_head_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, M->stream, a);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// K slices of a weight-gradient product of `outputs` elements over `rows`
// batch rows: about 256 workgroups, at least 256 rows per slice, and within
// the partial buffer.
int split_count(const pp_mlp* M, int rows, int64_t outputs) {
    const int64_t tiles = (outputs + (int64_t)kBM * kBN - 1) / ((int64_t)kBM * kBN);
    int s = (int)std::max<int64_t>(1, 256 / std::max<int64_t>(tiles, 1));
    s = std::min(s, std::max(1, rows / 256));
    s = (int)std::min<int64_t>(s, std::max<int64_t>(1, M->part_cap / std::max<int64_t>(outputs, 1)));
    return std::max(1, s);
}

// gradients of the parameters into M->g (internal layout) from dYb
int backward(pp_mlp* M, int rows) {
    const bf16* dcur = M->dYb;
    bf16* bufs[2] = {M->dA, M->dB};
    int nb = 0;
    for (int l = M->n_mat - 1; l >= 0; --l) {
        const int n_o = M->dims[l + 1], n_i = M->dims[l];
        GemmArgs w{};
        w.A = dcur;  // dZ^T: (m = out unit, k = row) at dcur[row * n_o + m]
        w.lda = n_o;
        w.B = M->H[l];  // (k = row, n = in unit) at H[row * n_i + n]
        w.ldb = n_i;
        w.M = n_o;
        w.N = n_i;
        w.K = rows;
        w.ldc = n_i;
        // split the batch (K) so that the weight gradient fills the device;
        // partial products are summed in a fixed order (deterministic)
        const int s_w = split_count(M, rows, (int64_t)n_o * n_i);
        w.k_chunk = (rows + s_w - 1) / s_w;
        w.k_chunk = (w.k_chunk + kBK - 1) / kBK * kBK;
        const int s_used = (rows + w.k_chunk - 1) / w.k_chunk;
        w.C = s_used > 1 ? M->part : M->g + M->w_off[l];
        w.c_stride = (int64_t)n_o * n_i;
        if (int rc = gemm<true, false, EPI_F32>(M, w)) return rc;
        if (s_used > 1) {
            hipLaunchKernelGGL(reduce_partials_kernel, dim3(blocks_for(w.c_stride, 256)), dim3(256),
                               0, M->stream, M->part, s_used, w.c_stride, M->g + M->w_off[l]);
            PP_HIP(hipGetLastError());
        }
        const int rc_chunk = 64;
        const int s_b = (rows + rc_chunk - 1) / rc_chunk;
        hipLaunchKernelGGL(colsum_kernel, dim3(blocks_for(n_o, 64), s_b), dim3(256), 0, M->stream,
                           dcur, n_o, rows, n_o, rc_chunk, M->part);
        PP_HIP(hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(blocks_for(n_o, 256)), dim3(256), 0,
                           M->stream, M->part, s_b, (int64_t)n_o, M->g + M->b_off[l]);
        PP_HIP(hipGetLastError());
        if (l == 0) break;
        GemmArgs d{};
        d.A = dcur;  // (m = row, k = out unit)
        d.lda = n_o;
        d.B = M->pb + M->w_off[l];  // (k = out unit, n = in unit) at W[k * n_i + n]
        d.ldb = n_i;
        d.M = rows;
        d.N = n_i;
        d.K = n_o;
        d.Z = M->Z[l];
        d.ldz = n_i;
        d.H = bufs[nb];
        d.ldh = n_i;
        if (int rc = gemm<false, false, EPI_DSILU>(M, d)) return rc;
        dcur = bufs[nb];
        nb ^= 1;
    }
    return PP_OK;
}

int adam(pp_mlp* M, const pp_adam& o) {
    M->pk_ok = false;
    AdamArgs a{M->p, M->g, M->m, M->v, M->pb, M->n_int, o, M->step};
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(M->n_int, 256)), dim3(256), 0, M->stream, a);
    PP_HIP(hipGetLastError());
    hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(64), 0, M->stream, M->step);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// packed (API) <-> internal parameter layouts
void pack_params(const pp_mlp* M, const std::vector<float>& in, float* out) {
    int64_t o = 0;
    for (int l = 0; l < M->n_mat; ++l) {
        const int64_t nw = (int64_t)M->dims[l + 1] * M->dims[l], nb = M->dims[l + 1];
        std::memcpy(out + o, in.data() + M->w_off[l], nw * 4);
        o += nw;
        std::memcpy(out + o, in.data() + M->b_off[l], nb * 4);
        o += nb;
    }
}

void unpack_params(const pp_mlp* M, const float* in, std::vector<float>& out) {
    out.assign((size_t)M->n_int, 0.0f);
    int64_t o = 0;
    for (int l = 0; l < M->n_mat; ++l) {
        const int64_t nw = (int64_t)M->dims[l + 1] * M->dims[l], nb = M->dims[l + 1];
        std::memcpy(out.data() + M->w_off[l], in + o, nw * 4);
        o += nw;
        std::memcpy(out.data() + M->b_off[l], in + o, nb * 4);
        o += nb;
    }
}

int upload_params(pp_mlp* M, const std::vector<float>& internal) {
    M->pk_ok = false;
    PP_HIP(hipMemcpyAsync(M->p, internal.data(), M->n_int * 4, hipMemcpyHostToDevice, M->stream));
    PP_HIP(hipMemsetAsync(M->m, 0, M->n_int * 4, M->stream));
    PP_HIP(hipMemsetAsync(M->v, 0, M->n_int * 4, M->stream));
    PP_HIP(hipMemsetAsync(M->step, 0, 4, M->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(M->n_int, 256)), dim3(256), 0, M->stream,
                       M->p, M->pb, M->n_int);
    PP_HIP(hipGetLastError());
    PP_HIP(hipStreamSynchronize(M->stream));
    return PP_OK;
}

int set_dev(pp_mlp* M) {
    PP_HIP(hipSetDevice(M->device));
    (void)hipGetLastError();
    return PP_OK;
}

// stage host features (n x n_in fp32) as the bf16 input H[0]
int stage_x(pp_mlp* M, const float* x, int n) {
    float* tmp = M->xs;
    const int64_t cnt = (int64_t)n * M->dims[0];
    PP_HIP(hipMemcpyAsync(tmp, x, cnt * 4, hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(cnt, 256)), dim3(256), 0, M->stream, tmp,
                       M->H[0], cnt);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

int ensure_loss(pp_mlp* M, int64_t n) {
    if (n <= M->loss_cap) return PP_OK;
    if (M->loss) PP_HIP(hipFree(M->loss));
    M->loss = nullptr;
    PP_HIP(hipMalloc(&M->loss, n * 4));
    M->loss_cap = n;
    return PP_OK;
}

}  // namespace

extern "C" {

int pp_mlp_create(pp_mlp** out, int device, int n_in, int n_hidden, int n_layers, int n_out,
                  int max_batch, uint64_t seed) {
    if (!out) return fail(PP_ERR_INVALID, "pp_mlp_create: out is NULL");
    *out = nullptr;
    if (n_in < 1 || n_in > 4096 || n_hidden < 1 || n_hidden > 4096 || n_layers < 1 ||
        n_layers > 16 || n_out < 1 || n_out > 4096)
        return fail(PP_ERR_INVALID, "pp_mlp_create: bad layer sizes");
    if (n_out < PP_SUR_N_OUT)
        return fail(PP_ERR_INVALID, "pp_mlp_create: the loss head needs n_out >= 10");
    if (max_batch < 1 || max_batch > (1 << 24))
        return fail(PP_ERR_INVALID, "pp_mlp_create: max_batch must be in [1, 2^24]");
    if (max_batch < n_in) return fail(PP_ERR_INVALID, "pp_mlp_create: max_batch < n_in");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(PP_ERR_INVALID, "pp_mlp_create: no such HIP device");
    auto* M = new pp_mlp();
    auto bail = [&](int rc) {
        pp_mlp_destroy(M);
        return rc;
    };
    M->device = device;
    M->max_batch = max_batch;
    M->dims.push_back(n_in);
    for (int l = 0; l < n_layers; ++l) M->dims.push_back(n_hidden);
    M->dims.push_back(n_out);
    M->n_mat = n_layers + 1;
    int64_t o = 0;
    for (int l = 0; l < M->n_mat; ++l) {
        M->w_off.push_back(o);
        o += ((int64_t)M->dims[l + 1] * M->dims[l] + 63) / 64 * 64;
        M->b_off.push_back(o);
        o += ((int64_t)M->dims[l + 1] + 63) / 64 * 64;
        M->n_params += (int64_t)M->dims[l + 1] * M->dims[l] + M->dims[l + 1];
    }
    M->n_int = o;
    if (hipSetDevice(device) != hipSuccess) return bail(fail(PP_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&M->ev0) != hipSuccess || hipEventCreate(&M->ev1) != hipSuccess)
        return bail(fail(PP_ERR_HIP, "pp_mlp_create: stream/event creation failed"));
    auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 256)); };
    const size_t rows = (size_t)max_batch;
    bool ok = alloc((void**)&M->p, M->n_int * 4) == hipSuccess &&
              alloc((void**)&M->g, M->n_int * 4) == hipSuccess &&
              alloc((void**)&M->m, M->n_int * 4) == hipSuccess &&
              alloc((void**)&M->v, M->n_int * 4) == hipSuccess &&
              alloc((void**)&M->pb, M->n_int * 2) == hipSuccess &&
              alloc((void**)&M->Y, rows * n_out * 4) == hipSuccess &&
              alloc((void**)&M->dY, rows * n_out * 4) == hipSuccess &&
              alloc((void**)&M->dYb, rows * n_out * 2) == hipSuccess &&
              alloc((void**)&M->dA, rows * n_hidden * 2) == hipSuccess &&
              alloc((void**)&M->dB, rows * n_hidden * 2) == hipSuccess &&
              alloc((void**)&M->Lb, rows * PP_SUR_LABEL_FLOATS * 4) == hipSuccess &&
              alloc((void**)&M->xs, rows * n_in * 4) == hipSuccess &&
              alloc((void**)&M->step, 4) == hipSuccess;
    M->H.assign(M->n_mat, nullptr);
    M->Z.assign(M->n_mat, nullptr);
    for (int l = 0; ok && l < M->n_mat; ++l) {
        ok = alloc((void**)&M->H[l], rows * M->dims[l] * 2) == hipSuccess;
        if (ok && l > 0) ok = alloc((void**)&M->Z[l], rows * M->dims[l] * 4) == hipSuccess;
    }
    // packed weights of the fused inference kernel: 32-column tiles x 16-deep
    // K fragments (K padded to a multiple of 32) x 512 bf16
    M->pk_off.assign(1, 0);
    for (int l = 0; l < M->n_mat; ++l)
        M->pk_off.push_back(M->pk_off.back() + (int64_t)((M->dims[l + 1] + 31) / 32) *
                                                   ((M->dims[l] + 31) / 32 * 2) * 512);
    if (ok && fused_ok(M)) ok = alloc((void**)&M->wpk, M->pk_off.back() * 2) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        return bail(fail(PP_ERR_NOMEM, "pp_mlp_create: device allocation failed"));
    }
    {
        int64_t wmax = 0, cmax = 0;
        for (int l = 0; l < M->n_mat; ++l) {
            wmax = std::max<int64_t>(wmax, (int64_t)M->dims[l + 1] * M->dims[l]);
            cmax = std::max<int64_t>(cmax, M->dims[l + 1]);
        }
        M->part_cap = std::max<int64_t>(16 * wmax, ((int64_t)max_batch / 64 + 1) * cmax);
        if (hipMalloc(&M->part, M->part_cap * 4) != hipSuccess) {
            (void)hipGetLastError();
            return bail(fail(PP_ERR_NOMEM, "pp_mlp_create: device allocation failed"));
        }
    }
    if (int rc = ensure_loss(M, 1)) return bail(rc);
    // initial parameters
    std::mt19937_64 rng(seed);
    auto unif = [&]() { return (double)(rng() >> 11) * 0x1p-53; };
    std::vector<float> init((size_t)M->n_int, 0.0f);
    for (int l = 0; l < M->n_mat; ++l) {
        const int fi = M->dims[l], fo = M->dims[l + 1];
        const double bound = 1.0 / std::sqrt((double)fi) * (l + 1 == M->n_mat ? 0.1 : 1.0);
        for (int64_t i = 0; i < (int64_t)fo * fi; ++i)
            init[(size_t)(M->w_off[l] + i)] = (float)((2.0 * unif() - 1.0) * bound);
        for (int i = 0; i < fo; ++i)
            init[(size_t)(M->b_off[l] + i)] = (float)((2.0 * unif() - 1.0) * bound);
    }
    float* bl = init.data() + M->b_off[M->n_mat - 1];
    const float head_bias[PP_SUR_N_OUT] = {-3.0f, 0.0f, 0.0f, 0.0f, 0.6931472f, 0.6931472f,
                                           0.6931472f, -1.0f, 0.5f, 2.0f};
    for (int i = 0; i < PP_SUR_N_OUT; ++i) bl[i] = head_bias[i];
    if (int rc = upload_params(M, init)) return bail(rc);
    *out = M;
    return PP_OK;
}

int pp_mlp_destroy(pp_mlp* M) {
    if (!M) return PP_OK;
    (void)hipSetDevice(M->device);
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    void* ptrs[] = {M->p, M->g, M->m, M->v, M->pb, M->Y, M->dY, M->dYb, M->dA, M->dB, M->Lb, M->xs,
                    M->part, M->wpk,
                    M->loss, M->step, M->dsX, M->dsL, M->logy, M->hits, M->mix, M->fb,
                    M->d_ev, M->d_mod};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto* p : M->H)
        if (p) (void)hipFree(p);
    for (auto* p : M->Z)
        if (p) (void)hipFree(p);
    if (M->ev0) (void)hipEventDestroy(M->ev0);
    if (M->ev1) (void)hipEventDestroy(M->ev1);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
    return PP_OK;
}

int64_t pp_mlp_param_count(pp_mlp* M) { return M ? M->n_params : -1; }

int pp_mlp_set_fused(pp_mlp* M, int fused) {
    if (!M) return fail(PP_ERR_INVALID, "pp_mlp_set_fused: NULL model");
    M->fused = fused != 0;
    return PP_OK;
}

int pp_mlp_get_params(pp_mlp* M, float* params) {
    if (!M || !params) return fail(PP_ERR_INVALID, "pp_mlp_get_params: NULL argument");
    if (int rc = set_dev(M)) return rc;
    std::vector<float> internal((size_t)M->n_int);
    PP_HIP(hipMemcpyAsync(internal.data(), M->p, M->n_int * 4, hipMemcpyDeviceToHost, M->stream));
    PP_HIP(hipStreamSynchronize(M->stream));
    pack_params(M, internal, params);
    return PP_OK;
}

int pp_mlp_set_params(pp_mlp* M, const float* params) {
    if (!M || !params) return fail(PP_ERR_INVALID, "pp_mlp_set_params: NULL argument");
    for (int64_t i = 0; i < M->n_params; ++i)
        if (!std::isfinite(params[i])) return fail(PP_ERR_INVALID, "pp_mlp_set_params: non-finite");
    if (int rc = set_dev(M)) return rc;
    std::vector<float> internal;
    unpack_params(M, params, internal);
    return upload_params(M, internal);
}

int pp_mlp_forward(pp_mlp* M, const float* x, int n, float* y) {
    if (!M || (n > 0 && (!x || !y))) return fail(PP_ERR_INVALID, "pp_mlp_forward: NULL argument");
    if (n < 0 || n > M->max_batch) return fail(PP_ERR_INVALID, "pp_mlp_forward: n > max_batch");
    if (n == 0) return PP_OK;
    if (int rc = set_dev(M)) return rc;
    if (int rc = stage_x(M, x, n)) return rc;
    if (int rc = forward(M, n, false)) return rc;
    PP_HIP(hipMemcpyAsync(y, M->Y, (size_t)n * M->dims.back() * 4, hipMemcpyDeviceToHost,
                          M->stream));
    PP_HIP(hipStreamSynchronize(M->stream));
    return PP_OK;
}

int pp_mlp_loss_grad(pp_mlp* M, const float* x, const float* labels, int n, float* loss,
                     float* grads) {
    if (!M || !x || !labels) return fail(PP_ERR_INVALID, "pp_mlp_loss_grad: NULL argument");
    if (n < 1 || n > M->max_batch) return fail(PP_ERR_INVALID, "pp_mlp_loss_grad: n out of range");
    if (int rc = set_dev(M)) return rc;
    if (int rc = stage_x(M, x, n)) return rc;
    PP_HIP(hipMemcpyAsync(M->Lb, labels, (size_t)n * PP_SUR_LABEL_FLOATS * 4,
                          hipMemcpyHostToDevice, M->stream));
    PP_HIP(hipMemsetAsync(M->loss, 0, 4, M->stream));
    if (int rc = forward(M, n, true)) return rc;
    if (int rc = loss_head(M, n, true, false)) return rc;
    if (int rc = backward(M, n)) return rc;
    float l = 0.0f;
    PP_HIP(hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, M->stream));
    std::vector<float> gi((size_t)M->n_int);
    if (grads)
        PP_HIP(hipMemcpyAsync(gi.data(), M->g, M->n_int * 4, hipMemcpyDeviceToHost, M->stream));
    PP_HIP(hipStreamSynchronize(M->stream));
    if (loss) *loss = l;
    if (grads) pack_params(M, gi, grads);
    return PP_OK;
}

int pp_mlp_adam_step(pp_mlp* M, const pp_adam* opt) {
    if (!M || !opt) return fail(PP_ERR_INVALID, "pp_mlp_adam_step: NULL argument");
    if (int rc = set_dev(M)) return rc;
    if (int rc = adam(M, *opt)) return rc;
    PP_HIP(hipStreamSynchronize(M->stream));
    return PP_OK;
}

int pp_mlp_train(pp_mlp* M, const float* x, const float* labels, int64_t n, int batch,
                 int n_steps, const pp_adam* opt, uint64_t seed, float* loss_curve) {
    if (!M || !x || !labels || !opt) return fail(PP_ERR_INVALID, "pp_mlp_train: NULL argument");
    if (n < 1 || batch < 1 || batch > M->max_batch || n_steps < 0)
        return fail(PP_ERR_INVALID, "pp_mlp_train: bad sizes");
    if (int rc = set_dev(M)) return rc;
    M->pk_ok = false;  // the replayed steps update the weights
    const int n_in = M->dims[0];
    if (n > M->ds_cap) {
        if (M->dsX) PP_HIP(hipFree(M->dsX));
        if (M->dsL) PP_HIP(hipFree(M->dsL));
        M->dsX = M->dsL = nullptr;
        M->ds_cap = 0;
        PP_HIP(hipMalloc(&M->dsX, (size_t)n * n_in * 4));
        PP_HIP(hipMalloc(&M->dsL, (size_t)n * PP_SUR_LABEL_FLOATS * 4));
        M->ds_cap = n;
    }
    PP_HIP(hipMemcpyAsync(M->dsX, x, (size_t)n * n_in * 4, hipMemcpyHostToDevice, M->stream));
    PP_HIP(hipMemcpyAsync(M->dsL, labels, (size_t)n * PP_SUR_LABEL_FLOATS * 4,
                          hipMemcpyHostToDevice, M->stream));
    // the step counter indexes the loss curve: count this call's steps from 0,
    // and keep the Adam step count in a second word
    uint32_t t0 = 0;
    PP_HIP(hipMemcpyAsync(&t0, M->step, 4, hipMemcpyDeviceToHost, M->stream));
    PP_HIP(hipStreamSynchronize(M->stream));
    if (int rc = ensure_loss(M, (int64_t)t0 + n_steps + 1)) return rc;
    PP_HIP(hipMemsetAsync(M->loss, 0, ((size_t)t0 + n_steps + 1) * 4, M->stream));
    // one training step, captured once and replayed
    GatherArgs ga{M->dsX, M->dsL, n, n_in, batch, (uint32_t)seed, (uint32_t)(seed >> 32),
                  M->step, M->H[0], M->Lb};
    auto enqueue_step = [&]() -> int {
        hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(batch, 256)), dim3(256), 0, M->stream, ga);
        PP_HIP(hipGetLastError());
        if (int rc = forward(M, batch, true)) return rc;
        if (int rc = loss_head(M, batch, true, true)) return rc;
        if (int rc = backward(M, batch)) return rc;
        return adam(M, *opt);
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (n_steps > 0) {
        PP_HIP(hipStreamBeginCapture(M->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_step();
        hipGraph_t g2 = nullptr;
        const hipError_t e = hipStreamEndCapture(M->stream, &g2);
        if (rc != PP_OK) {
            if (g2) (void)hipGraphDestroy(g2);
            return rc;
        }
        PP_HIP(e);
        graph = g2;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(PP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    int rc = PP_OK;
    for (int s = 0; s < n_steps; ++s) {
        const hipError_t e = hipGraphLaunch(exec, M->stream);
        if (e != hipSuccess) {
            rc = fail(PP_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
            break;
        }
    }
    std::vector<float> curve((size_t)t0 + n_steps + 1);
    if (rc == PP_OK && loss_curve && n_steps > 0) {
        const hipError_t e = hipMemcpyAsync(curve.data(), M->loss, curve.size() * 4,
                                            hipMemcpyDeviceToHost, M->stream);
        if (e != hipSuccess) rc = fail(PP_ERR_HIP, "pp_mlp_train: loss copy");
    }
    const hipError_t es = hipStreamSynchronize(M->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != PP_OK) return rc;
    PP_HIP(es);
    if (loss_curve)
        for (int s = 0; s < n_steps; ++s) loss_curve[s] = curve[(size_t)t0 + s];
    return PP_OK;
}

int pp_mlp_eval_loss(pp_mlp* M, const float* x, const float* labels, int64_t n, float* loss) {
    if (!M || !x || !labels || !loss) return fail(PP_ERR_INVALID, "pp_mlp_eval_loss: NULL argument");
    if (n < 1) return fail(PP_ERR_INVALID, "pp_mlp_eval_loss: n < 1");
    if (int rc = set_dev(M)) return rc;
    double total = 0.0;
    for (int64_t r0 = 0; r0 < n; r0 += M->max_batch) {
        const int rows = (int)std::min<int64_t>(M->max_batch, n - r0);
        if (int rc = stage_x(M, x + r0 * M->dims[0], rows)) return rc;
        PP_HIP(hipMemcpyAsync(M->Lb, labels + r0 * PP_SUR_LABEL_FLOATS,
                              (size_t)rows * PP_SUR_LABEL_FLOATS * 4, hipMemcpyHostToDevice,
                              M->stream));
        PP_HIP(hipMemsetAsync(M->loss, 0, 4, M->stream));
        if (int rc = forward(M, rows, false)) return rc;
        if (int rc = loss_head(M, rows, false, false)) return rc;
        float l = 0.0f;
        PP_HIP(hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, M->stream));
        PP_HIP(hipStreamSynchronize(M->stream));
        total += (double)l * rows;
    }
    *loss = (float)(total / (double)n);
    return PP_OK;
}

int pp_sur_features(pp_mlp* M, const pp_cascade* cascades, const float* module_xyz, int64_t n,
                    float* features) {
    if (!M || (n > 0 && (!cascades || !module_xyz || !features)))
        return fail(PP_ERR_INVALID, "pp_sur_features: NULL argument");
    if (n <= 0) return PP_OK;
    if (int rc = set_dev(M)) return rc;
    const int n_in = M->dims[0];
    pp_cascade* dc = nullptr;
    float *dm = nullptr, *df = nullptr;
    auto cleanup = [&]() {
        if (dc) (void)hipFree(dc);
        if (dm) (void)hipFree(dm);
        if (df) (void)hipFree(df);
    };
    if (hipMalloc(&dc, n * sizeof(pp_cascade)) != hipSuccess ||
        hipMalloc(&dm, n * 12) != hipSuccess || hipMalloc(&df, n * n_in * 4) != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        return fail(PP_ERR_NOMEM, "pp_sur_features: device allocation failed");
    }
    PairArgs a{};
    a.ev = dc;
    a.mod = dm;
    a.n_mod = 0;
    a.row0 = 0;
    a.rows = n;
    a.n_in = n_in;
    a.xf = df;
    int rc = PP_OK;
    if (hipMemcpyAsync(dc, cascades, n * sizeof(pp_cascade), hipMemcpyHostToDevice, M->stream) !=
            hipSuccess ||
        hipMemcpyAsync(dm, module_xyz, n * 12, hipMemcpyHostToDevice, M->stream) != hipSuccess)
        rc = fail(PP_ERR_HIP, "pp_sur_features: upload failed");
    if (rc == PP_OK) {
        hipLaunchKernelGGL(pair_features_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, M->stream, a);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(features, df, n * n_in * 4, hipMemcpyDeviceToHost, M->stream) !=
                hipSuccess ||
            hipStreamSynchronize(M->stream) != hipSuccess)
            rc = fail(PP_ERR_HIP, "pp_sur_features: kernel or copy failed");
    }
    cleanup();
    return rc;
}

int pp_sur_eval_pairs(pp_mlp* M, const pp_cascade* events, int n_events, const float* module_xyz,
                      int n_modules, const pp_sur_domain* dom, float* expected_hits,
                      float* mixture, uint8_t* fallback, float* kernel_ms) {
    if (!M || !events || !module_xyz || !dom || !expected_hits)
        return fail(PP_ERR_INVALID, "pp_sur_eval_pairs: NULL argument");
    if (n_events < 0 || n_modules < 0) return fail(PP_ERR_INVALID, "pp_sur_eval_pairs: bad sizes");
    const int64_t n = (int64_t)n_events * n_modules;
    if (n == 0) return PP_OK;
    if (int rc = set_dev(M)) return rc;
    // device outputs for all pairs, inputs per event / module
    if (n > M->out_cap) {
        for (void** p : {(void**)&M->hits, (void**)&M->mix, (void**)&M->fb, (void**)&M->logy})
            if (*p) {
                PP_HIP(hipFree(*p));
                *p = nullptr;
            }
        M->out_cap = 0;
        PP_HIP(hipMalloc(&M->hits, n * 4));
        PP_HIP(hipMalloc(&M->mix, n * 9 * 4));
        PP_HIP(hipMalloc(&M->fb, n));
        PP_HIP(hipMalloc(&M->logy, (size_t)M->max_batch * 4));
        M->out_cap = n;
    }
    if (n_events > M->ev_cap) {
        if (M->d_ev) PP_HIP(hipFree(M->d_ev));
        M->d_ev = nullptr;
        PP_HIP(hipMalloc(&M->d_ev, (size_t)n_events * sizeof(pp_cascade)));
        M->ev_cap = n_events;
    }
    if (n_modules > M->mod_cap) {
        if (M->d_mod) PP_HIP(hipFree(M->d_mod));
        M->d_mod = nullptr;
        PP_HIP(hipMalloc(&M->d_mod, (size_t)n_modules * 12));
        M->mod_cap = n_modules;
    }
    PP_HIP(hipMemcpyAsync(M->d_ev, events, (size_t)n_events * sizeof(pp_cascade),
                          hipMemcpyHostToDevice, M->stream));
    PP_HIP(hipMemcpyAsync(M->d_mod, module_xyz, (size_t)n_modules * 12, hipMemcpyHostToDevice,
                          M->stream));
    PP_HIP(hipEventRecord(M->ev0, M->stream));
    if (M->fused && M->wpk) {
        // one launch over every pair: features, network and head in one kernel
        PairArgs a{};
        a.ev = M->d_ev;
        a.mod = M->d_mod;
        a.n_mod = n_modules;
        a.rows = n;
        a.n_in = M->dims[0];
        a.fb = M->fb;
        a.dom = *dom;
        if (int rc = forward_fused(M, n, &a, M->hits, mixture ? M->mix : nullptr)) return rc;
    }
    for (int64_t r0 = 0; !(M->fused && M->wpk) && r0 < n; r0 += M->max_batch) {
        const int rows = (int)std::min<int64_t>(M->max_batch, n - r0);
        PairArgs a{};
        a.ev = M->d_ev;
        a.mod = M->d_mod;
        a.n_mod = n_modules;
        a.row0 = r0;
        a.rows = rows;
        a.n_in = M->dims[0];
        a.xb = M->H[0];
        a.logy = M->logy;
        a.fb = M->fb + r0;
        a.dom = *dom;
        hipLaunchKernelGGL(pair_features_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0,
                           M->stream, a);
        PP_HIP(hipGetLastError());
        if (int rc = forward(M, rows, false)) return rc;
        hipLaunchKernelGGL(sur_head_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, M->stream,
                           M->Y, M->dims.back(), (int64_t)rows, M->logy, M->hits + r0,
                           mixture ? M->mix + r0 * 9 : nullptr);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(M->ev1, M->stream));
    PP_HIP(hipMemcpyAsync(expected_hits, M->hits, n * 4, hipMemcpyDeviceToHost, M->stream));
    if (mixture)
        PP_HIP(hipMemcpyAsync(mixture, M->mix, n * 9 * 4, hipMemcpyDeviceToHost, M->stream));
    if (fallback) PP_HIP(hipMemcpyAsync(fallback, M->fb, n, hipMemcpyDeviceToHost, M->stream));
    PP_HIP(hipStreamSynchronize(M->stream));
    if (kernel_ms) PP_HIP(hipEventElapsedTime(kernel_ms, M->ev0, M->ev1));
    return PP_OK;
}

}  // extern "C"

namespace {

// Training labels from the hit records of a propagation whose sources are
// the samples: per hit the absorption weight w = exp(-dist / abs_len(wl))
// (abs_len linear in the table, clamped at its ends, times abs_scale, in
// FP64 as the host computes it) and the residual bin of
// x = t - t_geo(source) + PP_SUR_TIME_SHIFT. Weights are summed as 2^-32
// fixed point in 64-bit integer atomics, so that the sums do not depend on
// the order in which hits arrive. acc: (1 + PP_SUR_BINS) per source.
constexpr int kLabelMaxKnots = 1024;
constexpr double kFix = 4294967296.0;  // 2^32

__global__ void __launch_bounds__(256) hit_label_kernel(const pp::HitRecord* __restrict__ rec,
                                                        int64_t n, const double* tab, int n_abs,
                                                        double abs_scale, const double* tgeo,
                                                        int n_src, unsigned long long* acc) {
    __shared__ double t[2 * kLabelMaxKnots];
    for (int i = threadIdx.x; i < 2 * n_abs; i += blockDim.x) t[i] = tab[i];
    __syncthreads();
    const double* xw = t;
    const double* yl = t + n_abs;
    const double bin_w = (double)PP_SUR_TMAX / PP_SUR_BINS;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float4* r = reinterpret_cast<const float4*>(rec + i);
        const float4 c = r[2], d = r[3];
        const int s = (int)__float_as_uint(d.w);
        if (s < 0 || s >= n_src) continue;
        const double wl = c.y, dist = c.w, time = c.z;
        // np.interp: clamped ends, slope * (x - x_j) + y_j inside
        double len;
        if (wl <= xw[0]) {
            len = yl[0];
        } else if (wl >= xw[n_abs - 1]) {
            len = yl[n_abs - 1];
        } else {
            int lo = 0, hi = n_abs - 1;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (xw[mid] <= wl) lo = mid;
                else hi = mid;
            }
            const double slope = (yl[hi] - yl[lo]) / (xw[hi] - xw[lo]);
            len = slope * (wl - xw[lo]) + yl[lo];
        }
        const double w = exp(-dist / (len * abs_scale));
        const unsigned long long wf = (unsigned long long)llrint(w * kFix);
        unsigned long long* row = acc + (int64_t)s * (1 + PP_SUR_BINS);
        atomicAdd(row, wf);
        const double x = time - tgeo[s] + (double)PP_SUR_TIME_SHIFT;
        const double b = floor(x / bin_w);
        if (b >= 0.0 && b < (double)PP_SUR_BINS) atomicAdd(row + 1 + (int)b, wf);
    }
}

__global__ void label_finish_kernel(const unsigned long long* acc, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)((double)acc[i] / kFix);
}

int label_fail(int code, const std::string& msg) {
    pp_set_error_message(msg.c_str());
    return code;
}

}  // namespace

extern "C" {

int pp_sur_hit_labels(pp_ctx* ctx, const double* abs_wl, const double* abs_len, int n_abs,
                      double abs_scale, const double* tgeo, int n_sources, float* labels) {
    if (!ctx) return label_fail(PP_ERR_INVALID, "pp_sur_hit_labels: NULL ctx");
    if (n_abs < 2 || n_abs > kLabelMaxKnots || !abs_wl || !abs_len)
        return label_fail(PP_ERR_INVALID, "pp_sur_hit_labels: absorption table of 2..1024 knots");
    for (int i = 1; i < n_abs; ++i)
        if (!(abs_wl[i] > abs_wl[i - 1]))
            return label_fail(PP_ERR_INVALID, "pp_sur_hit_labels: wavelengths must increase");
    if (n_sources < 1 || !tgeo || !labels || !(abs_scale > 0.0))
        return label_fail(PP_ERR_INVALID, "pp_sur_hit_labels: bad sources or output");
    const pp::HitRecord* rec = nullptr;
    int64_t n = 0;
    hipStream_t st = nullptr;
    int device = 0;
    if (int rc = pp::ctx_hit_records(ctx, &rec, &n, &st, &device)) return rc;
    const size_t per = 1 + PP_SUR_BINS;
    const size_t acc_b = (size_t)n_sources * per * 8;
    const size_t off_tab = (acc_b + 255) & ~(size_t)255;
    const size_t off_tg = off_tab + (((size_t)2 * n_abs * 8 + 255) & ~(size_t)255);
    const size_t off_out = off_tg + (((size_t)n_sources * 8 + 255) & ~(size_t)255);
    const size_t bytes = off_out + (size_t)n_sources * per * 4;
    char* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return label_fail(PP_ERR_NOMEM, "pp_sur_hit_labels: device allocation failed");
    }
    std::vector<double> tab(2 * (size_t)n_abs);
    std::copy(abs_wl, abs_wl + n_abs, tab.begin());
    std::copy(abs_len, abs_len + n_abs, tab.begin() + n_abs);
    int rc = PP_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == PP_OK)
            rc = label_fail(PP_ERR_HIP, std::string("pp_sur_hit_labels: ") + what + ": " +
                                            hipGetErrorString(e));
    };
    chk(hipMemsetAsync(d, 0, acc_b, st), "clear");
    chk(hipMemcpyAsync(d + off_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st), "table");
    chk(hipMemcpyAsync(d + off_tg, tgeo, (size_t)n_sources * 8, hipMemcpyHostToDevice, st), "tgeo");
    if (rc == PP_OK && n > 0) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
        hipLaunchKernelGGL(hit_label_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rec, n,
                           (const double*)(d + off_tab), n_abs, abs_scale,
                           (const double*)(d + off_tg), n_sources, (unsigned long long*)d);
        chk(hipGetLastError(), "label kernel");
    }
    if (rc == PP_OK) {
        const int64_t m = (int64_t)n_sources * per;
        hipLaunchKernelGGL(label_finish_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                           (const unsigned long long*)d, m, (float*)(d + off_out));
        chk(hipGetLastError(), "label finish");
        chk(hipMemcpyAsync(labels, d + off_out, (size_t)m * 4, hipMemcpyDeviceToHost, st), "copy");
    }
    chk(hipStreamSynchronize(st), "sync");
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
