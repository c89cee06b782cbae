// pp_search.h - target search of one photon step.
//
// Semantics of src/photon_prop_cuda.jl:494-506: among the targets whose near
// root d satisfies 0 < d < step, report the one with the smallest index
// ("first target in index order", not the nearest).
//
// BruteSearch tests every target (wave-uniform loop over LDS, broadcast
// reads) - the reference's own algorithm. MixedSearch is the same loop over
// lists that hold planar targets (rectangles, discs); for those the hit
// window is 0 <= d <= step (src/photon_prop_cuda.jl:353-404).
// GridSearch walks the step segment through a uniform grid of box cells
// (3-D DDA, Amanatides & Woo), cubic or taller along z. Each cell record holds the bounding box of
// the targets overlapping the cell; a slab test of the segment against that
// box decides whether the cell's targets are tested. Every target whose hit
// point lies on the segment overlaps a visited cell (cells are binned with a
// margin far above float rounding of the walk) and lies inside that cell's
// box (grown by a further margin, so the slab test can only err towards
// testing), and the smallest hit index over all tested candidates is kept:
// the result equals BruteSearch's.
#pragma once

#include "pp_device.h"

namespace pp {

struct GridDesc {
    float ox, oy, oz;  // lower corner
    float h, inv_h;    // cell edge along x and y
    float hz, inv_hz;  // cell edge along z (strings are vertical: cells may be columns)
    int nx, ny, nz;
    int n_cells, n_items;
    float zpad;        // largest target radius plus the member boxes' margin
    float ex, ey, ez;  // upper corner, ox + nx h etc. (rounded once, on the host)
};

// The step segment clipped to the grid box: reciprocal direction components,
// plane offsets and the parameter interval [t_in, t_out] inside the box
// (empty if t_in > t_out). Every operation is spelled out (explicit FMAs,
// products that cannot fuse), so two evaluations with the same inputs give
// the same bits wherever the compiler inlines them.
struct Clip {
    float ix, iy, iz, px, py, pz, t_in, t_out;
};

// Work done by one search (instrumented builds only).
struct Work {
    uint32_t tests = 0, cells = 0, walks = 0;
};

struct BruteSearch {
    const float4* tgt;
    int n;
    __host__ __device__ __forceinline__ int operator()(f3 pos, f3 dir, float step, float& dt) const {
        Work w;
        return find<false>(pos, dir, step, dt, w);
    }
    template <bool COUNT>
    __host__ __device__ __forceinline__ int find(f3 pos, f3 dir, float step, float& dt,
                                                 Work& w) const {
        if (COUNT) w.tests += (uint32_t)n;
        const float ex = dir_excess(dir);
        int found = -1;
        for (int t = 0; t < n; ++t) {
            float d;
            if (sphere_hit(tgt[t], pos, dir, ex, step, d) && found < 0) {
                found = t;
                dt = d;
            }
        }
        return found;
    }
};

// Reciprocal for the grid walk, clamped to +-1e20 (1/+-0 = +-inf becomes
// +-1e20): the hardware estimate on the device (the walk only needs its
// boundary parameters to be consistent, not correctly rounded); a plain
// division when the same code is compiled for the host.
__host__ __device__ __forceinline__ float rcp_clamped(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(__builtin_amdgcn_rcpf(x), -1e20f, 1e20f);
#else
    return fminf(fmaxf(1.0f / x, -1e20f), 1e20f);
#endif
}

// a * b for 0 <= a, b and a * b < 2^24 (cell indices: a grid has at most
// 2^21 cells, its tables' 64 MB budget): the full-rate 24-bit multiply
__host__ __device__ __forceinline__ int mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// x, opaque to the optimiser: a value computed from it stays where it is
// written instead of being hoisted out of a rarely taken branch
__host__ __device__ __forceinline__ float pinned(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(x));
#endif
    return x;
}

// v limited to [0, hi], hi >= 0 (one v_med3_i32 on the device)
__host__ __device__ __forceinline__ int clampi(int v, int hi) {
    return min(max(v, 0), hi);
}

__host__ __device__ __forceinline__ uint32_t float_bits(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(x);
#else
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
#endif
}

// Target lists with planar shapes: every target in index order, as
// BruteSearch, with the shape test chosen per target (the kind is the same
// for all lanes, so the branch is wave-uniform). Two records per target:
// (x, y, z, r^2) and (kind bits, length_x/2, length_y/2, 0).
struct MixedSearch {
    const float4* tgt;
    int n;
    __host__ __device__ __forceinline__ int operator()(f3 pos, f3 dir, float step, float& dt) const {
        Work w;
        return find<false>(pos, dir, step, dt, w);
    }
    template <bool COUNT>
    __host__ __device__ __forceinline__ int find(f3 pos, f3 dir, float step, float& dt,
                                                 Work& w) const {
        if (COUNT) w.tests += (uint32_t)n;
        const float ex = dir_excess(dir);
        int found = -1;
        for (int t = 0; t < n; ++t) {
            const float4 a = tgt[2 * t], e = tgt[2 * t + 1];
            const uint32_t kind = float_bits(e.x);
            float d;
            const bool h = kind == PP_SHAPE_SPHERE ? sphere_hit(a, pos, dir, ex, step, d)
                                                   : plane_hit(kind, a, e, pos, dir, step, d);
            if (h && found < 0) {
                found = t;
                dt = d;
            }
        }
        return found;
    }
};

struct GridSearch {
    const float4* tgt;
    const float4* cells;    // 2 per cell: (lo xyz, start bits), (hi xyz, count bits)
    const uint16_t* items;  // target indices per cell
    GridDesc g;

    __host__ __device__ __forceinline__ int operator()(f3 pos, f3 dir, float step, float& dt) const {
        Work w;
        return find<false>(pos, dir, step, dt, w);
    }
    template <bool COUNT>
    __host__ __device__ __forceinline__ int find(f3 pos, f3 dir, float step, float& dt,
                                                 Work& w) const {
        // one layer of cells (columns over the whole height): a 2-D walk
        return g.nz == 1 ? walk<COUNT, true>(pos, dir, step, dt, w)
                         : walk<COUNT, false>(pos, dir, step, dt, w);
    }
    __host__ __device__ __forceinline__ Clip clip(f3 pos, f3 dir, float step) const {
        Clip c;
        // reciprocal direction components, finite (|1/d| <= 1e20, the sign
        // of a zero component kept); the walk's step signs come from these
        // same values so that they agree with ix..iz
        c.ix = rcp_clamped(dir.x);
        c.iy = rcp_clamped(dir.y);
        c.iz = rcp_clamped(dir.z);
        // plane parameters as one FMA per plane: (b - p) / d = b (1/d) - p (1/d);
        // the rounding of p (1/d) moves a plane by about |p| 2^-24 along its
        // axis, far inside the binning margin and the boxes' 1 cm margin
        c.px = pos.x * c.ix;
        c.py = pos.y * c.iy;
        c.pz = pos.z * c.iz;
        // clip the segment [0, step] to the grid box
        const float tx0 = fmaf(g.ox, c.ix, -c.px), tx1 = fmaf(g.ex, c.ix, -c.px);
        const float ty0 = fmaf(g.oy, c.iy, -c.py), ty1 = fmaf(g.ey, c.iy, -c.py);
        const float tz0 = fmaf(g.oz, c.iz, -c.pz), tz1 = fmaf(g.ez, c.iz, -c.pz);
        c.t_in = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        c.t_out = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), step));
        return c;
    }
    // Whether the bounding box of the segment a-b meets the grid box: a
    // superset of the segments that clip() keeps (the grid box holds every
    // target's box with a margin of at least 2 mm, far beyond the rounding of
    // the end point b), by compares alone, combined without branches
    __host__ __device__ __forceinline__ bool touches(f3 a, f3 b) const {
        const bool x = ((a.x >= g.ox) | (b.x >= g.ox)) & ((a.x <= g.ex) | (b.x <= g.ex));
        const bool y = ((a.y >= g.oy) | (b.y >= g.oy)) & ((a.y <= g.ey) | (b.y <= g.ey));
        const bool z = ((a.z >= g.oz) | (b.z >= g.oz)) & ((a.z <= g.ez) | (b.z <= g.ez));
        return x & y & z;
    }
    // touches() of every lane as a wave mask: each compare is a ballot of its
    // own (one v_cmp writing a scalar mask) and the masks are combined by
    // scalar instructions; as one lane condition the compiler merged the
    // pairs of compares into minima and maxima, which need a canonicalising
    // instruction per operand, and copied the result out of and back into a
    // mask
    __device__ __forceinline__ uint64_t touch_mask(f3 a, f3 b) const {
        const uint64_t x = (__ballot(a.x >= g.ox) | __ballot(b.x >= g.ox)) &
                           (__ballot(a.x <= g.ex) | __ballot(b.x <= g.ex));
        const uint64_t y = (__ballot(a.y >= g.oy) | __ballot(b.y >= g.oy)) &
                           (__ballot(a.y <= g.ey) | __ballot(b.y <= g.ey));
        const uint64_t z = (__ballot(a.z >= g.oz) | __ballot(b.z >= g.oz)) &
                           (__ballot(a.z <= g.ez) | __ballot(b.z <= g.ez));
        return x & y & z;
    }
    template <bool COUNT, bool COLS>
    __host__ __device__ __forceinline__ int walk(f3 pos, f3 dir, float step, float& dt,
                                                 Work& w) const {
        return walk_from<COUNT, COLS>(clip(pos, dir, step), pos, dir, step, dt, w);
    }
    // the walk over the clipped segment; cl = clip(pos, dir, step)
    template <bool COUNT, bool COLS>
    __host__ __device__ __forceinline__ int walk_from(const Clip& cl, f3 pos, f3 dir, float step,
                                                      float& dt, Work& w) const {
        const float ix_ = cl.ix, iy_ = cl.iy, iz_ = cl.iz, px_ = cl.px, py_ = cl.py, pz_ = cl.pz;
        const float t_in = cl.t_in, t_out = cl.t_out;
        int found = -1;
        if (!(t_in <= t_out)) return found;
        // entry cell
        const float px = pos.x + dir.x * t_in, py = pos.y + dir.y * t_in, pz = pos.z + dir.z * t_in;
        const int cx = clampi((int)floorf((px - g.ox) * g.inv_h), g.nx - 1);
        const int cy = clampi((int)floorf((py - g.oy) * g.inv_h), g.ny - 1);
        const int cz = COLS ? 0 : clampi((int)floorf((pz - g.oz) * g.inv_hz), g.nz - 1);
        const int sx = ix_ > 0.0f ? 1 : -1, sy = iy_ > 0.0f ? 1 : -1, sz = iz_ > 0.0f ? 1 : -1;
        // ray parameter of the next cell boundary on each axis (in one layer
        // the walk never crosses a z face: the clip already ends it there)
        float tmx = fmaf(g.ox + (float)(cx + (sx > 0)) * g.h, ix_, -px_);
        float tmy = fmaf(g.oy + (float)(cy + (sy > 0)) * g.h, iy_, -py_);
        float tmz = COLS ? INFINITY : fmaf(g.oz + (float)(cz + (sz > 0)) * g.hz, iz_, -pz_);
        const float tdx = g.h * fabsf(ix_), tdy = g.h * fabsf(iy_), tdz = g.hz * fabsf(iz_);
        // cell index and its increments: the walk advances with selects, not
        // branches. It ends at the first boundary at or beyond t_out, which is
        // at most the grid exit; where rounding lets it cross the grid's face
        // first, the next index (wrapped into another row, or clamped into the
        // table) names a cell off the segment, whose targets are tested
        // exactly and cannot change the smallest hit index.
        // the cell as the byte offset of its record pair (32 bytes a cell);
        // its steps are selects between uniform constants, not products
        const int row = 32 * g.nx, layer = 32 * g.nx * g.ny;
        const int dcx = sx > 0 ? 32 : -32, dcy = sy > 0 ? row : -row;
        const int dcz = COLS ? 0 : (sz > 0 ? layer : -layer);
        const int c_max = 32 * (g.n_cells - 1);
        const char* const rec = reinterpret_cast<const char*>(cells);
        int c = 32 * (mul24(mul24(cz, g.ny) + cy, g.nx) + cx);
        const int max_visits = g.nx + g.ny + g.nz;
        for (int v = 0; v < max_visits; ++v) {
            // this cell's record; the walk's next step is computed while it loads
            const float4 lo = *reinterpret_cast<const float4*>(rec + c);
            const float4 hi = *reinterpret_cast<const float4*>(rec + c + 16);
            // axis of the nearest boundary (ties: x, then y), by compares and
            // selects rather than fminf, whose NaN rules cost a canonicalising
            // instruction per operand here
            const bool ax = COLS ? tmx <= tmy : tmx <= tmy && tmx <= tmz;
            const bool ay = COLS ? !ax : !ax && tmy <= tmz;
            const bool az = !ax && !ay;
            const float tn = ax ? tmx : (ay ? tmy : tmz);
            const bool last = tn >= t_out;
            const int cn = min(max(c + (ax ? dcx : (COLS || ay ? dcy : dcz)), 0), c_max);
            if (COUNT) w.cells += 1;
            // slab test of the segment [0, step] against the box of this
            // cell's targets (an empty cell holds a box at +inf, which fails
            // it), then the targets themselves. In one layer of columns the
            // boxes' z slab is left out: the clip already holds the segment
            // inside the grid's z range, a column's box spans its string,
            // and a wider [b_in, b_out] only admits candidates that the z
            // filter and the exact sphere test below then settle.
            const float ax0 = fmaf(lo.x, ix_, -px_), ax1 = fmaf(hi.x, ix_, -px_);
            const float ay0 = fmaf(lo.y, iy_, -py_), ay1 = fmaf(hi.y, iy_, -py_);
            // the end of the segment as a compare of its own (a minimum with
            // step would canonicalise step in every iteration)
            float b_in, b_out;
            if (COLS) {
                b_in = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), 0.0f);
                b_out = fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1));
            } else {
                const float az0 = fmaf(lo.z, iz_, -pz_), az1 = fmaf(hi.z, iz_, -pz_);
                b_in = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), 0.0f));
                b_out = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fmaxf(az0, az1));
            }
            if (b_in <= b_out && b_in <= step) {
                b_out = fminf(b_out, step);
                const uint32_t b = float_bits(lo.w), cnt = float_bits(hi.w);
                // rarely needed: computed here, not hoisted in front of the walk
                const float exc = dir_excess(make3(pinned(dir.x), pinned(dir.y), pinned(dir.z)));
                // a target is hit inside the box, so within its radius of the
                // segment's z range there: targets farther in z are skipped
                // (columns hold a whole string; a ray crossing it meets one or
                // two of its modules)
                const float za = fmaf(dir.z, b_in, pos.z), zb = fmaf(dir.z, b_out, pos.z);
                const float zlo = fminf(za, zb) - g.zpad, zhi = fmaxf(za, zb) + g.zpad;
                for (uint32_t k = b; k < b + cnt; ++k) {
                    const int t = items[k];
                    const float4 c = tgt[t];
                    if (c.z < zlo || c.z > zhi) continue;
                    if (COUNT) w.tests += 1;
                    float d;
                    if (sphere_hit(c, pos, dir, exc, step, d) && (found < 0 || t < found)) {
                        found = t;
                        dt = d;
                    }
                }
            }
            if (last) break;
            c = cn;
            tmx += ax ? tdx : 0.0f;
            tmy += ay ? tdy : 0.0f;
            if (!COLS) tmz += az ? tdz : 0.0f;
        }
        return found;
    }
};

}  // namespace pp
