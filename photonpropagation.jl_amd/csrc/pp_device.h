// pp_device.h - device-side step functions of the gfx950 photon propagator.
//
// Each function is the device form of one step of the per-photon algorithm of
// PhotonPropagation.jl's kernel (src/photon_prop_cuda.jl), written as a pure
// function of explicit uniforms so that it can be evaluated stand-alone
// (pp_eval_fn) and inside the propagation kernel from the same inlined code.
// Forms are chosen for the FP32 VALU: reciprocal multiplies instead of
// divisions where the value is reused, sqrt((1-x)(1+x)) instead of
// sin(acos(x)), a one-division rotation, a one-cube-root Cardano solution.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/photonprop_hip.h"

// No contraction from here on in a translation unit that includes this
// header: each fused multiply-add is written as fmaf and every other product
// and sum is rounded on its own. A function then gives the same bits wherever
// it is inlined (the propagation kernels inline the step functions into
// several schedules whose results must agree bit for bit). The setting is
// taken when a function is parsed, so a file that switches contraction back
// on for its own code after the include leaves these functions as they are.
#pragma clang fp contract(off)

namespace pp {

constexpr float kTwoPi = 6.28318548202514648438f;  // float(2*pi), as T(2*pi) in Julia
constexpr float kCVacMNs = 0.299792458f;          // src/photon_prop_cuda.jl:37

// ---------------------------------------------------------------------------
// Philox4x32-10. Round: (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0));
// the key is bumped by the Weyl constants between rounds. Each 32x32->64
// product is one v_mad_u64_u32; the key schedule stays in SGPRs.
// ---------------------------------------------------------------------------
// 3-input XOR in one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// R rounds on c with the round keys (k0, k1) of the first of them
template <int R>
__device__ __forceinline__ uint4 philox_rounds(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * (uint64_t)c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * (uint64_t)c.z;
        c = make_uint4(xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1,
                       xor3((uint32_t)(p0 >> 32), c.w, k1), (uint32_t)p0);
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c;
}

__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
    return philox_rounds<10>(c, k0, k1);
}

// [0,1) and (0,1] with 24 random bits (exact integer->float conversion)
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
__device__ __forceinline__ float u01_open0(uint32_t x) {
    return (float)((x >> 8) + 1u) * 0x1p-24f;
}

// Identity of one photon's random stream. block(word) is Philox4x32-10 of
// the counter (word, pid_lo, pid_hi, src) under the key (k0, k1), evaluated
// as StepStream (below) does: the word-independent half of rounds 1 and 2 is
// the same for every block of a photon, so where the blocks are inlined into
// one function it is computed once, and the word's own products are scalar.
struct Stream {
    uint32_t k0, k1;    // seed
    uint32_t pid_lo, pid_hi;
    uint32_t src;
    __device__ __forceinline__ uint4 block(uint32_t word) const;
};

// The step draws of one photon: Stream::block(word), bit for bit, with the
// word-independent half of rounds 1 and 2 done once per photon. Round 1
// multiplies c.z (the photon id's high word) and round 2 the x word that
// product yields, neither of which depends on the word; the word enters
// through c.x, and while a wave steps together it is wave-uniform, so its
// products and the round keys stay scalar. Per photon: y1 = lo(M1 pid_hi),
// and h2, w2 the high and low halves of M0 x1, x1 = hi(M1 pid_hi)^pid_lo^k0.
struct StepStream {
    uint32_t k0, k1, src;
    uint32_t y1, h2, w2;
    __device__ __forceinline__ explicit StepStream(const Stream& s)
        : k0(s.k0), k1(s.k1), src(s.src) {
        const uint64_t p1 = (uint64_t)kPhiloxM1 * (uint64_t)s.pid_hi;
        const uint32_t x1 = xor3((uint32_t)(p1 >> 32), s.pid_lo, s.k0);
        const uint64_t q = (uint64_t)kPhiloxM0 * (uint64_t)x1;
        y1 = (uint32_t)p1;
        h2 = (uint32_t)(q >> 32);
        w2 = (uint32_t)q;
    }
    __device__ __forceinline__ uint4 block(uint32_t word) const {
        // round 1: (x1, y1, z1, w1) with z1, w1 from the word
        const uint64_t p0 = (uint64_t)kPhiloxM0 * (uint64_t)word;
        const uint32_t z1 = (uint32_t)(p0 >> 32) ^ src ^ k1;
        const uint32_t w1 = (uint32_t)p0;
        // round 2, keys + (W0, W1): M1 z1 from the word, M0 x1 per photon
        const uint64_t p1 = (uint64_t)kPhiloxM1 * (uint64_t)z1;
        const uint32_t ka = (uint32_t)(p1 >> 32) ^ (k0 + kPhiloxW0);
        const uint32_t kb = w1 ^ (k1 + kPhiloxW1);
        const uint32_t x2 = y1 ^ ka, y2 = (uint32_t)p1, z2 = h2 ^ kb;
        // round 3, keys + 2 (W0, W1): its y word (y2) is still scalar
        const uint64_t q0 = (uint64_t)kPhiloxM0 * (uint64_t)x2;
        const uint64_t q1 = (uint64_t)kPhiloxM1 * (uint64_t)z2;
        const uint4 c = make_uint4((uint32_t)(q1 >> 32) ^ (y2 ^ (k0 + 2u * kPhiloxW0)), (uint32_t)q1,
                                   xor3((uint32_t)(q0 >> 32), w2, k1 + 2u * kPhiloxW1), (uint32_t)q0);
        return philox_rounds<7>(c, k0 + 3u * kPhiloxW0, k1 + 3u * kPhiloxW1);
    }
};

__device__ __forceinline__ uint4 Stream::block(uint32_t word) const {
    return StepStream(*this).block(word);
}

struct f3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ f3 make3(float x, float y, float z) { return f3{x, y, z}; }

// ---------------------------------------------------------------------------
// Arithmetic helpers for the FP32 VALU
// ---------------------------------------------------------------------------

// 1/x: hardware reciprocal (1 ulp) refined by one Newton step (~0.5 ulp),
// 3 instructions instead of the IEEE division sequence.
__device__ __forceinline__ float rcp_nr(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}

// 1/x in FP64 for a normal x: the hardware estimate refined by two Newton
// steps (each doubles the correct bits; ~1 ulp of FP64), 5 instructions
// instead of the IEEE division sequence.
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = ::fma(r, ::fma(-x, r, 1.0), r);
    return ::fma(r, ::fma(-x, r, 1.0), r);
}

// Square root by the hardware instruction (1 ulp). The correctly rounded
// sequence the compiler emits for sqrtf adds two Newton corrections and a
// denormal rescale (~14 instructions); every argument here is zero or a
// normal number, and the per-step results only need a few ulp.
__host__ __device__ __forceinline__ float sqrt_hw(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// Cube root of x > 0: hardware log2/exp2 estimate, one Newton step.
__device__ __forceinline__ float cbrt_pos(float x) {
    const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
    // y' = y + (x / y^2 - y) / 3
    return fmaf(fmaf(x, rcp_nr(y * y), -y), 1.0f / 3.0f, y);
}

// sin(2 pi u), cos(2 pi u) for u in [0, 1): the hardware v_sin_f32/v_cos_f32,
// whose argument is in revolutions (no reduction needed). Measured on gfx950
// over all 2^24 u = k 2^-24: absolute error <= 1.25e-7 (1.05 ulp of 1), the
// scale at which direction components are compared.
__device__ __forceinline__ void sincos_2pi(float u, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(u);
    *c = __builtin_amdgcn_cosf(u);
}

// ---------------------------------------------------------------------------
// Directions
// ---------------------------------------------------------------------------

// (sin t cos p, sin t sin p, cos t) from cos t, sin t and p = 2*pi*u
__device__ __forceinline__ f3 dir_from_cos_phi(float ct, float st, float u_phi) {
    float sp, cp;
    sincos_2pi(u_phi, &sp, &cp);
    return make3(cp * st, sp * st, ct);
}

// initialize_direction_isotropic: cos t = U(-1,1), phi = U(0,2pi)
//   src/photon_prop_cuda.jl:73-77
__device__ __forceinline__ f3 isotropic_dir(float u1, float u2) {
    const float ct = -1.0f + u1 * 2.0f;
    const float st = sqrt_hw((1.0f - ct) * (1.0f + ct));
    return dir_from_cos_phi(ct, st, u2);
}

// rot_from_ez_fast(a, b): the minimal rotation taking e_z to a, applied to b.
// With v = e_z x a = (-ay, ax, 0), c = az:  R b = c b + v x b + h v (v.b),
// h = 1/(1+c) for c >= 0 and (1-c)/|v|^2 for c < 0 (same value for unit a,
// no cancellation on either side). a = -e_z exactly turns by pi about x.
__device__ __forceinline__ f3 rot_from_ez(f3 a, f3 b) {
    const float c = a.z;
    const float s2 = fmaf(a.x, a.x, a.y * a.y);
    const bool up = c >= 0.0f;
    const float h = (up ? 1.0f : (1.0f - c)) * rcp_nr(up ? (1.0f + c) : s2);
    const float hv = h * fmaf(a.x, b.y, -(a.y * b.x));  // h (v.b)
    f3 o;
    o.x = fmaf(-a.y, hv, fmaf(c, b.x, a.x * b.z));
    o.y = fmaf(a.x, hv, fmaf(c, b.y, a.y * b.z));
    o.z = fmaf(c, b.z, -fmaf(a.x, b.x, a.y * b.y));
    if (!up && s2 == 0.0f) o = make3(b.x, -b.y, -b.z);
    return o;
}

// ---------------------------------------------------------------------------
// Phase functions (CherenkovMediumBase samplers; inverse CDFs)
// ---------------------------------------------------------------------------
struct Phase {
    int kind;          // pp_phase_kind
    float hg_fraction;
    // HG: cos = A - C t^2, t = B / (D0 + D1 u)
    float hg_A, hg_B, hg_C, hg_D0, hg_D1;
    // ES: q = Q (1 - 2u); x = sgn (W - P3/W), W = cbrt(|q|/2 + sqrt(q^2/4 + P3^3))
    float es_Q, es_P3, es_P3cube;
    // simplified Liu: p(cos) ~ (1 + cos)^(2g/(1-g)), cos = 2 u^((1-g)/(1+g)) - 1
    float sl_exp;
    // the HG/ES mixture on the random words (hg_es_cos_words): hg_D1 2^-24,
    // and the HG choice u < hg_fraction as (word >> 8) < hg_words
    float hg_D1w;
    uint32_t hg_words;
};

__host__ __device__ inline Phase make_phase(int kind, float g, float b, float hg_fraction) {
    Phase p;
    p.kind = kind;
    p.hg_fraction = hg_fraction;
    // 1 +- g^2 with one rounding: 1 - g^2 cancels for g near 1
    p.hg_A = fmaf(g, g, 1.0f) / (2.0f * g);
    p.hg_B = fmaf(-g, g, 1.0f);
    p.hg_C = 1.0f / (2.0f * g);
    p.hg_D0 = 1.0f - g;
    p.hg_D1 = 2.0f * g;
    p.es_Q = (3.0f + b) / b;
    p.es_P3 = 1.0f / b;
    p.es_P3cube = p.es_P3 * p.es_P3 * p.es_P3;
    p.sl_exp = (1.0f - g) / (1.0f + g);
    p.hg_D1w = p.hg_D1 * 0x1p-24f;
    // u01(w) = n 2^-24 < f  <=>  n < f 2^24  <=>  n < ceil(f 2^24) (n an integer;
    // f 2^24 is exact)
    const float f24 = hg_fraction * 0x1p24f;
    p.hg_words = !(f24 > 0.0f) ? 0u : f24 >= 0x1p24f ? (1u << 24) : (uint32_t)ceilf(f24);
    return p;
}

// x limited to [-1, 1] by one v_med3_f32 (fminf/fmaxf add an instruction that
// canonicalises x first; x is never a NaN here)
__device__ __forceinline__ float clamp1(float x) { return __builtin_amdgcn_fmed3f(x, -1.0f, 1.0f); }

__device__ __forceinline__ float hg_cos(const Phase& p, float u) {
    const float t = p.hg_B * rcp_nr(fmaf(p.hg_D1, u, p.hg_D0));
    return clamp1(fmaf(-p.hg_C, t * t, p.hg_A));
}

__device__ __forceinline__ float es_cos(const Phase& p, float u) {
    const float q = p.es_Q * fmaf(-2.0f, u, 1.0f);
    const float hq = 0.5f * fabsf(q);
    const float w = cbrt_pos(hq + sqrt_hw(fmaf(hq, hq, p.es_P3cube)));
    const float x = fmaf(-p.es_P3, rcp_nr(w), w);
    return clamp1(q > 0.0f ? -x : x);
}

// simplified Liu (PPC's ice phase function): inverse CDF 2 u^e - 1, e = (1-g)/(1+g)
__device__ __forceinline__ float sl_cos(const Phase& p, float u) {
    return clamp1(fmaf(2.0f, exp2f(p.sl_exp * log2f(u)), -1.0f));
}

// The HG/ES mixture without a branch: nearly every wave holds photons of both
// kinds, so both samplers run anyway; they share the one Newton reciprocal,
// of HG's denominator or of ES's root w (the same values as hg_cos and es_cos).
__device__ __forceinline__ float hg_es_cos(const Phase& p, float u_choice, float u) {
    const bool hg = u_choice < p.hg_fraction;
    const float q = p.es_Q * fmaf(-2.0f, u, 1.0f);
    const float hq = 0.5f * fabsf(q);
    const float w = cbrt_pos(hq + sqrt_hw(fmaf(hq, hq, p.es_P3cube)));
    const float r = rcp_nr(hg ? fmaf(p.hg_D1, u, p.hg_D0) : w);
    const float t = p.hg_B * r;
    const float x = fmaf(-p.es_P3, r, w);
    return clamp1(hg ? fmaf(-p.hg_C, t * t, p.hg_A) : (q > 0.0f ? -x : x));
}

// hg_es_cos(p, u01(w_choice), u01(w_u)), bit for bit, from the two random
// words: the choice as an integer compare, and the 2^-24 of u folded into the
// constants it multiplies (-2 u = -2^-23 n and D1 u = (D1 2^-24) n exactly, so
// each FMA rounds the same value)
__device__ __forceinline__ float hg_es_cos_words(const Phase& p, uint32_t w_choice, uint32_t w_u) {
    const bool hg = (w_choice >> 8) < p.hg_words;
    const float n = (float)(w_u >> 8);
    const float q = p.es_Q * fmaf(-0x1p-23f, n, 1.0f);
    const float hq = 0.5f * fabsf(q);
    const float w = cbrt_pos(hq + sqrt_hw(fmaf(hq, hq, p.es_P3cube)));
    const float r = rcp_nr(hg ? fmaf(p.hg_D1w, n, p.hg_D0) : w);
    const float t = p.hg_B * r;
    const float x = fmaf(-p.es_P3, r, w);
    return clamp1(hg ? fmaf(-p.hg_C, t * t, p.hg_A) : (q > 0.0f ? -x : x));
}

__device__ __forceinline__ float phase_cos(const Phase& p, float u_choice, float u) {
    if (p.kind == PP_PHASE_HG) return hg_cos(p, u);
    if (p.kind == PP_PHASE_MIXED_HG_SL) return u_choice < p.hg_fraction ? hg_cos(p, u) : sl_cos(p, u);
    return hg_es_cos(p, u_choice, u);
}

// phase_cos(p, u01(w_choice), u01(w_u)), the form the propagation kernels call
__device__ __forceinline__ float phase_cos_words(const Phase& p, uint32_t w_choice, uint32_t w_u) {
    if (p.kind == PP_PHASE_HG || p.kind == PP_PHASE_MIXED_HG_SL)
        return phase_cos(p, u01(w_choice), u01(w_u));
    return hg_es_cos_words(p, w_choice, w_u);
}

// update_direction   src/photon_prop_cuda.jl:280-298
__device__ __forceinline__ f3 update_direction(f3 dir, float ct, float u_phi) {
    const float st = sqrt_hw(fmaf(-ct, ct, 1.0f));
    return rot_from_ez(dir, dir_from_cos_phi(ct, st, u_phi));
}

// ---------------------------------------------------------------------------
// Step length   src/photon_prop_cuda.jl:488-490
// step = -log(eta) sca_len = -log2(eta) (ln2 sca_len): the hardware log2
// (eta >= 2^-24 is a normal number) times a per-photon constant, instead of
// the library logf's denormal rescale and extended-precision ln2 product.
// ---------------------------------------------------------------------------
constexpr float kLn2 = 0.693147180559945309f;
__device__ __forceinline__ float step_length(float eta, float ln2_sca_len) {
    return -__builtin_amdgcn_logf(eta) * ln2_sca_len;
}

// ---------------------------------------------------------------------------
// Ray-sphere test: _intersect_nsphere   src/photon_prop_cuda.jl:309-351
// Near root d = -a - sqrt(b), a = dir.dp, b = a^2 - (|dp|^2 - r^2); hit iff
// b >= 0 and 0 < d < step. b is evaluated as
//     b = r^2 - |dp - a dir|^2 + a^2 (|dir|^2 - 1),
// the same value without the cancellation of a^2 against |dp|^2 for a far
// sphere; the last term keeps it equal to the reference's discriminant for a
// direction that has drifted from unit length. dir_excess() gives
// |dir|^2 - 1 once per step (FP64 FMAs: the squares of floats are exact).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ float dir_excess(f3 dir) {
    const double x = dir.x, y = dir.y, z = dir.z;
    return (float)::fma(x, x, ::fma(y, y, ::fma(z, z, -1.0)));
}

// Every product and sum is spelled out (fmaf, or a plain product that cannot
// be fused), so that the result does not depend on where the compiler inlines
// it: the grid and the brute-force search give identical distances.
__host__ __device__ __forceinline__ bool sphere_hit(float4 t, f3 pos, f3 dir, float excess,
                                                    float step, float& d) {
    const float dx = pos.x - t.x, dy = pos.y - t.y, dz = pos.z - t.z;
    const float a = fmaf(dir.z, dz, fmaf(dir.y, dy, dir.x * dx));
    const float qx = fmaf(-a, dir.x, dx), qy = fmaf(-a, dir.y, dy), qz = fmaf(-a, dir.z, dz);
    const float q2 = fmaf(qz, qz, fmaf(qy, qy, qx * qx));
    const float b = fmaf(a * excess, a, t.w - q2);  // t.w = r^2
    d = (b >= 0.0f) ? (-a - sqrt_hw(b)) : -1.0f;
    return (d > 0.0f) & (d < step);
}
// Planar targets: check_intersection(::Rectangular / ::Circular)
//   src/photon_prop_cuda.jl:353-404. The plane z = t.z is crossed at
// d = (t.z - z) / dir_z (correctly rounded division, as the reference's
// Float32 `/`), a hit iff dir_z != 0, 0 <= d <= step and the crossing point
// lies strictly inside the shape: |x - t.x| < length_x/2 and
// |y - t.y| < length_y/2 for a rectangle (e = (_, lx/2, ly/2, _)),
// (x - t.x)^2 + (y - t.y)^2 < r^2 for a disc (t.w = r^2).
__host__ __device__ __forceinline__ bool plane_hit(uint32_t kind, float4 t, float4 e, f3 pos,
                                                   f3 dir, float step, float& d) {
    d = -1.0f;
    if (dir.z == 0.0f) return false;
    d = (t.z - pos.z) / dir.z;
    if (d < 0.0f || d > step) return false;
    const float x = fmaf(dir.x, d, pos.x) - t.x, y = fmaf(dir.y, d, pos.y) - t.y;
    return kind == PP_SHAPE_RECTANGLE ? (fabsf(x) < e.y && fabsf(y) < e.z)
                                      : (x * x + y * y < t.w);
}

// ---------------------------------------------------------------------------
// Medium: per-wavelength values   src/photon_prop_cuda.jl:480-483
// ---------------------------------------------------------------------------
struct Medium {
    int kind;  // pp_medium_kind
    Phase phase;
    // Values derived from launch constants are stored here, computed on the
    // host: the kernels would otherwise hoist each conversion or product into
    // a register that spills, and reload it from scratch for every photon.
    // SIMPLE (inv_c_grp = 1 / c_grp)
    float sca_len, c_grp, n_phase, inv_c_grp;
    // CASCADIA: Kopelevich b = k1 x^-4.3 + k2 x^-1.7 + k3 x^-0.3, x = wl/550
    // (FP64 copies of the Float32 coefficients)
    double kop1, kop2, kop3, sca_scale;
    // Quan-Fry: n = a0 + a1/wl + a2/wl^2 + a3/wl^3 ; n_g = a0 + 2a1/wl + 3a2/wl^2 + 4a3/wl^3
    // (qg1..qg3 = 2 a1, 3 a2, 4 a3 in Float32)
    float qf0, qf1, qf2, qf3, qg1, qg2, qg3;
    // ICE: sca_len = ice_pref (wl/400)^ice_alpha, ice_pref = sca_scale (1-g) / b_dust_400
    float ice_pref, ice_alpha;
};

struct MediumValues {
    float sca_len, inv_c_grp, cos_c, sin_c;
};

__device__ __forceinline__ MediumValues medium_values(const Medium& m, float wl) {
    MediumValues v;
    float n;
    if (m.kind == PP_MEDIUM_SIMPLE) {
        v.sca_len = m.sca_len;
        v.inv_c_grp = m.inv_c_grp;
        n = m.n_phase;
    } else if (m.kind == PP_MEDIUM_ICE) {
        // src/ice_properties.jl:105-126 (indices), :164-168 (scattering)
        v.sca_len = m.ice_pref * exp2f(m.ice_alpha * log2f(wl * (1.0f / 400.0f)));
        // the index polynomials cancel to ~1.3: evaluated in double (once per
        // photon) so that n and the group velocity keep float accuracy
        const double x = (double)wl * 1e-3;
        const double nd = 1.55749 + x * (-1.57988 + x * (3.99993 + x * (-4.68271 + x * 2.09354)));
        const double ng = nd * (1.227106 + x * (-0.954648 + x * (1.42568 + x * -0.711832)));
        n = (float)nd;
        v.inv_c_grp = (float)(ng * (1.0 / (double)kCVacMNs));
    } else {
        // b = k1 x^-4.3 + k2 x^-1.7 + k3 x^-0.3, x = wl/550, with one
        // transcendental t = x^-0.3 = q^0.3 (q = 1/x; the small exponent keeps
        // the hardware log2's error out of the result) and x^-4.3 = q^4 t,
        // x^-1.7 = q^2 / t in FP64 (a handful of FP64 operations once per
        // photon; q in [0.6, 2] and b are normal numbers)
        const double q = 550.0 * rcp64((double)wl);
        const double q2 = q * q;
        const float tf = __builtin_amdgcn_exp2f(0.3f * __builtin_amdgcn_logf((float)q));
        const double t = (double)tf;
        const double b = fma(m.kop1, q2 * q2 * t, fma(m.kop2, q2 * rcp64(t), m.kop3 * t));
        v.sca_len = (float)(m.sca_scale * rcp64(b));
        const float il = rcp_nr(wl);
        // Horner with FMAs
        n = fmaf(il, fmaf(il, fmaf(il, m.qf3, m.qf2), m.qf1), m.qf0);
        const float ng = fmaf(il, fmaf(il, fmaf(il, m.qg3, m.qg2), m.qg1), m.qf0);
        v.inv_c_grp = ng * (1.0f / kCVacMNs);
    }
    v.cos_c = rcp_nr(n);  // cherenkov_angle: acos(1/n), beta = 1
    v.sin_c = sqrt_hw((1.0f - v.cos_c) * (1.0f + v.cos_c));
    return v;
}

// sample_cherenkov_track_direction: 1 - cos t follows x^(a-1) exp(-b x^a),
// a = 0.39, b = 2.61, truncated at x = 2 (inverse CDF); phi uniform.
__device__ __forceinline__ f3 cherenkov_track_dir(float u1, float u2) {
    constexpr float kInvA = 1.0f / 0.39f;
    constexpr float kInvB = 1.0f / 2.61f;
    constexpr float kI = 0.9672917f;  // 1 - exp(-2.61 * 2^0.39)
    const float t = -logf(1.0f - u1 * kI) * kInvB;
    // t^(1/a) as 2^(log2(t) / a) with the hardware log2/exp2: where the power
    // matters (t > 2^-9) the exponent carries < 1.5e-6 absolute error
    const float ct = fmaxf(1.0f - __builtin_amdgcn_exp2f(kInvA * __builtin_amdgcn_logf(t)), -1.0f);
    const float st = sqrt_hw((1.0f - ct) * (1.0f + ct));
    return dir_from_cos_phi(ct, st, u2);
}

// ---------------------------------------------------------------------------
// Samplers on the photon's sampler sub-streams
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sampler_word(uint32_t sub, uint32_t i) {
    return PP_CTR_SAMPLER | (sub << 24) | i;
}

__device__ __forceinline__ float normal_sample(const Stream& s, uint32_t sub) {
    const uint4 r = s.block(sampler_word(sub, 0));
    float sp, cp;
    sincos_2pi(u01(r.y), &sp, &cp);
    return sqrt_hw(-2.0f * logf(u01_open0(r.x))) * cp;
}

// Marsaglia-Tsang gamma(shape, scale). Each sampler block: a Box-Muller pair
// (angle 2 pi u) and two acceptance uniforms (two attempts).
// shape < 1 boosts Gamma(shape+1) by U^(1/shape), U from the sub-stream's
// boost block. The accepted value d (1 + c x)^3 scale boost amplifies the
// relative error of 1 + c x three times, so that factor is formed with one
// FMA from the float variate and the cube and the product are taken in FP64
// (a handful of FP64 operations once per photon) and rounded once: a draw
// whose decisions agree with an exact evaluation is then within ~2 ulp of it.
// x_acc (may be NULL) receives the accepted normal variate (0 if none is).
__device__ inline float gamma_sample(float shape, float scale, const Stream& s, uint32_t sub,
                                     float* x_acc = nullptr) {
    float boost = 1.0f;
    if (shape < 1.0f) {
        const uint4 r = s.block(sampler_word(sub, PP_SAMPLER_BOOST_BLOCK));
        boost = powf(u01_open0(r.x), 1.0f / shape);
        shape += 1.0f;
    }
    const double dd = (double)shape - 1.0 / 3.0;
    const float d = (float)dd;
    const float c = 1.0f / sqrtf(9.0f * d);
    for (uint32_t i = 0; i < PP_GAMMA_MAX_BLOCKS; ++i) {
        const uint4 r = s.block(sampler_word(sub, i));
        // the Box-Muller radius and angle from the library logarithm, a
        // correctly rounded square root and sincospi (each within ~1 ulp):
        // the cube below amplifies the variate's error
        const float rad = sqrtf(-2.0f * logf(u01_open0(r.x)));
        float sp, cp;
        sincospif(2.0f * u01(r.y), &sp, &cp);
        const float z[2] = {rad * cp, rad * sp};
        const float ua[2] = {u01_open0(r.z), u01_open0(r.w)};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float x = z[k];
            const double v1 = fma((double)c, (double)x, 1.0);
            if (v1 <= 0.0) continue;
            const double v3 = v1 * v1 * v1;
            const float v = (float)v3;
            const float u = ua[k];
            const float x2 = x * x;
            if (u < 1.0f - 0.0331f * (x2 * x2) || logf(u) < 0.5f * x2 + d * (1.0f - v + logf(v))) {
                if (x_acc) *x_acc = x;
                return (float)(dd * v3 * (double)scale * (double)boost);
            }
        }
    }
    if (x_acc) *x_acc = 0.0f;
    return (float)(dd * (double)scale * (double)boost);
}

// ---------------------------------------------------------------------------
// Spectrum: inverse-CDF knots, linear interpolation
// ---------------------------------------------------------------------------
// n_f = float(n), last_f = float(n - 1): passed in, computed once on the host
// (a conversion of a launch constant would otherwise be hoisted into a
// register that spills)
__device__ __forceinline__ float spectrum_sample(const float* tab, int n, float n_f, float last_f,
                                                 int mapping, float mono, float u) {
    if (n == 0) return mono;
    // separately rounded products and sums (this header turns contraction
    // off): the knot position and a + f (b - a) as the reference evaluates them
    float x = mapping == PP_SPEC_TEXTURE ? u * n_f - 0.5f : u * last_f;
    x = fminf(fmaxf(x, 0.0f), last_f);
    const int i = min((int)x, n - 2);
    const float f = x - (float)i;
    const float a = tab[i], b = tab[i + 1];
    return a + f * (b - a);
}

}  // namespace pp
