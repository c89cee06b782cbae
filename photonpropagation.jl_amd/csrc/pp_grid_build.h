// pp_grid_build.h - host construction of the uniform target grid searched by
// GridSearch (pp_search.h): cell edge choice, binning, per-cell index lists
// and per-cell bounding boxes of the members.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pp_search.h"

namespace pp {

// LDS layout of the propagation kernel: targets float4[n_t] | cell records
// float4[2 n_cells] | cell items u16[n_items, padded to even] | spectrum
// float[n_spec]
__host__ __device__ inline size_t lds_bytes(int n_t, int n_cells, int n_items, int n_spec,
                                            bool grid) {
    size_t b = (size_t)n_t * 16;
    if (grid) b += (size_t)n_cells * 32 + (size_t)((n_items + 1) & ~1) * 2;
    return b + (size_t)n_spec * 4;
}

struct GridTables {
    GridDesc g{};
    std::vector<float4> cells;    // 2 per cell: (lo xyz, start), (hi xyz, count)
    std::vector<uint16_t> items;  // at least one entry
};

inline float u32_as_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Bin every target (centre xyz, radius) into the cells its bounding box plus
// a margin m = 1e-3 max(h, hz) + 1 mm overlaps (the margin is far above the
// float rounding of the DDA walk). Each cell lists its targets in ascending
// index order and keeps the union of their boxes, grown by a further
// 1 cm + 1e-4 max(h, hz), so that a float slab test of a segment against it can
// only err towards testing; an empty cell keeps a box at +inf. cell_req > 0
// sets the xy edge; <= 0 picks, among 1-2.5x the edge of one cubic cell per
// target, the edge of least modelled walk cost (below). cell_z_req < 0 makes
// one layer over the whole height (columns: strings are vertical), > 0 sets
// the z edge, 0 makes the cells cubic - or, with cell_req <= 0 as well, picks
// cubic cells or columns by the same model. A set edge grows by 1.25x until the tables
// fit `budget` bytes of LDS together with the targets and n_spec spectrum
// knots. Returns false if nothing fits.
inline bool build_grid_tables(const float* xyz, const float* radius, int n, float cell_req,
                              float cell_z_req, int n_spec, size_t budget, GridTables& out) {
    if (n <= 0 || n > 65535) return false;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], xyz[3 * i + a] - radius[i]);
            hi[a] = std::max(hi[a], xyz[3 * i + a] + radius[i]);
        }
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = std::max(hi[a] - lo[a], 1e-3f);
    // one grid of xy edge h; z edge h (zmode 0, cubic), hz0 (zmode > 0) or
    // the whole height (zmode < 0, columns); false if its tables exceed the
    // budget
    auto try_build = [&](float h, int zmode, float hz0, GridTables& out) -> bool {
        // column cells: one layer whose edge covers the extent with margins
        float hz = zmode == 0 ? h : hz0;
        float m = 1e-3f * std::max(h, hz) + 1e-3f;
        if (zmode < 0) {
            hz = ext[2] + 4.0f * (1e-3f * std::max(h, ext[2]) + 1e-3f) + 1e-2f;
            m = 1e-3f * std::max(h, hz) + 1e-3f;
        }
        const float edge[3] = {h, h, hz};
        float o[3];
        int dim[3];
        for (int a = 0; a < 3; ++a) {
            // whole cells over the extent and margins, centred on the targets
            // (targets on a regular pitch equal to the edge sit mid-cell)
            dim[a] = std::max(1, (int)std::ceil((ext[a] + 4.0f * m) / edge[a]));
            o[a] = lo[a] - 2.0f * m - 0.5f * ((float)dim[a] * edge[a] - (ext[a] + 4.0f * m));
        }
        const int64_t n_cells = (int64_t)dim[0] * dim[1] * dim[2];
        if (n_cells * 32 > (int64_t)budget) return false;
        auto range = [&](int i, int a, int& c0, int& c1) {
            const float p = xyz[3 * i + a], r = radius[i];
            c0 = std::max(0, (int)std::floor((p - r - m - o[a]) / edge[a]));
            c1 = std::min(dim[a] - 1, (int)std::floor((p + r + m - o[a]) / edge[a]));
        };
        std::vector<uint32_t> count((size_t)n_cells, 0);
        int64_t n_items = 0;
        for (int i = 0; i < n; ++i) {
            int x0, x1, y0, y1, z0, z1;
            range(i, 0, x0, x1);
            range(i, 1, y0, y1);
            range(i, 2, z0, z1);
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x) {
                        ++count[(size_t)((z * dim[1] + y) * dim[0] + x)];
                        ++n_items;
                    }
        }
        if (lds_bytes(n, (int)n_cells, (int)n_items, n_spec, true) > budget) return false;
        std::vector<uint32_t> start((size_t)n_cells + 1, 0);
        for (int64_t k = 0; k < n_cells; ++k)
            start[(size_t)k + 1] = start[(size_t)k] + count[(size_t)k];
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        out.items.assign((size_t)std::max<int64_t>(n_items, 1), 0);
        std::vector<float> box_lo(3 * (size_t)n_cells, INFINITY), box_hi(3 * (size_t)n_cells, -INFINITY);
        const float mb = m + 1e-2f + 1e-4f * std::max(h, hz);
        for (int i = 0; i < n; ++i) {
            int x0, x1, y0, y1, z0, z1;
            range(i, 0, x0, x1);
            range(i, 1, y0, y1);
            range(i, 2, z0, z1);
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x) {
                        const size_t c = (size_t)((z * dim[1] + y) * dim[0] + x);
                        out.items[fill[c]++] = (uint16_t)i;
                        for (int a = 0; a < 3; ++a) {
                            box_lo[3 * c + a] = std::min(box_lo[3 * c + a], xyz[3 * i + a] - radius[i] - mb);
                            box_hi[3 * c + a] = std::max(box_hi[3 * c + a], xyz[3 * i + a] + radius[i] + mb);
                        }
                    }
        }
        out.cells.assign(2 * (size_t)n_cells, float4{});
        for (int64_t k = 0; k < n_cells; ++k) {
            const size_t c = (size_t)k;
            if (count[c] == 0)  // a box at +inf: every slab test against it fails
                for (int a = 0; a < 3; ++a) box_lo[3 * c + a] = box_hi[3 * c + a] = INFINITY;
            out.cells[2 * c] = make_float4(box_lo[3 * c], box_lo[3 * c + 1], box_lo[3 * c + 2],
                                           u32_as_float(start[c]));
            out.cells[2 * c + 1] = make_float4(box_hi[3 * c], box_hi[3 * c + 1], box_hi[3 * c + 2],
                                               u32_as_float(count[c]));
        }
        GridDesc& g = out.g;
        g.ox = o[0];
        g.oy = o[1];
        g.oz = o[2];
        g.h = h;
        g.inv_h = 1.0f / h;
        g.hz = hz;
        g.inv_hz = 1.0f / hz;
        g.nx = dim[0];
        g.ny = dim[1];
        g.nz = dim[2];
        g.n_cells = (int)n_cells;
        g.n_items = (int)n_items;
        float r_max = 0.0f;
        for (int i = 0; i < n; ++i) r_max = std::max(r_max, radius[i]);
        g.zpad = r_max + mb;
        g.ex = g.ox + (float)g.nx * g.h;
        g.ey = g.oy + (float)g.ny * g.h;
        g.ez = g.oz + (float)g.nz * g.hz;
        return true;
    };
    const float hz_req = cell_z_req > 0.0f ? std::max(cell_z_req, 1e-3f) : 0.0f;
    const int zmode_req = cell_z_req > 0.0f ? 1 : cell_z_req < 0.0f ? -1 : 0;
    if (cell_req > 0.0f) {
        float h = std::max(cell_req, 1e-3f), hz0 = hz_req;
        for (int attempt = 0; attempt < 200; ++attempt, h *= 1.25f, hz0 *= 1.25f)
            if (try_build(h, zmode_req, hz0, out)) return true;
        return false;
    }
    // Automatic edge (and, when cell_z_req is 0 too, shape): the candidate of
    // least modelled cost per unit of photon path. Inside the grid box, a
    // segment of unit length with a uniform position and direction crosses
    // (1/h + 1/h + 1/hz) / 2 cell faces on average and meets a box of surface
    // S with probability S / (4 V) (Cauchy's formula, V the grid volume). A
    // visit costs ~42 instructions (~37 in the 2-D walk of one layer); a box
    // that is met costs 8 per member (the z-range filter) plus 40 per member
    // left to test, a fraction (2 zpad + w) / height of them for a box of
    // width w. Bigger cells are visited less, but once a cell holds targets far
    // apart (two strings) its member box spans the gap and is met far more
    // often. Each walk also costs ~75 instructions of set-up (clip, entry cell,
    // batch bookkeeping), and a step walks when its end points' box meets the
    // grid box: for steps of length L ~ 40 m (scattering lengths of deep
    // water and ice) and isotropic directions, when it starts in the grid box
    // grown by L/2 along each axis. Per unit of path in a region around the
    // detector: cost = 75 prod(X_a + L/2) / L + V (visits and boxes per unit
    // length), X_a the grid's sides - so a grid whose cells overshoot the
    // targets' extent pays for its larger box. The candidates: 1-2.5x the edge
    // of one cubic cell per target, and the edges of 1.5-2.5x it that divide
    // the larger xy extent (with margins) into a whole number of cells exactly
    // (smaller exact edges, one column per line of a dense lattice, measured
    // slower than the model puts them: the member loop of a met box is bound
    // by its LDS latency, which the model leaves out).
    const float base = std::cbrt(ext[0] * ext[1] * ext[2] / (float)n);
    std::vector<float> cand = {1.0f * base, 1.25f * base, 1.5f * base, 1.75f * base,
                               2.0f * base, 2.25f * base, 2.5f * base};
    const float ext_xy = std::max(ext[0], ext[1]);
    for (int k = 1; k <= 4096; ++k) {
        // fixed point of h = (ext_xy + 4 m) / k with the margin m of try_build
        // for cubic cells (hz = h) and for columns (hz the whole height), a
        // little above it so that ceil gives k
        float hk[2];
        for (int col = 0; col < 2; ++col) {
            float h = ext_xy / (float)k;
            for (int it = 0; it < 3; ++it) {
                const float hz = col ? ext[2] + 4.0f * (1e-3f * std::max(h, ext[2]) + 1e-3f) + 1e-2f : h;
                h = (ext_xy + 4.0f * (1e-3f * std::max(h, hz) + 1e-3f)) / (float)k;
            }
            hk[col] = h * (1.0f + 4e-6f);
        }
        if (std::max(hk[0], hk[1]) < 1.5f * base) break;
        for (float h : hk)
            if (h >= 1.5f * base && h <= 2.5f * base) cand.push_back(h);
    }
    constexpr double kWalkSetup = 75.0, kStepLen = 40.0;
    bool found = false;
    double best = INFINITY;
    const int zmodes[2] = {zmode_req, -1};
    const int n_modes = cell_z_req == 0.0f ? 2 : 1;  // fully automatic: cubic or columns
    for (int mi = 0; mi < n_modes; ++mi) {
        for (float h0 : cand) {
            const float h = std::max(h0, 1e-3f);
            GridTables t;
            if (!try_build(h, zmodes[mi], hz_req, t)) continue;
            const GridDesc& g = t.g;
            const double sx = (double)g.nx * g.h, sy = (double)g.ny * g.h, sz = (double)g.nz * g.hz;
            const double vol = sx * sy * sz;
            double met = 0.0;
            for (int c = 0; c < g.n_cells; ++c) {
                const float4 a = t.cells[2 * (size_t)c], b = t.cells[2 * (size_t)c + 1];
                uint32_t cnt;
                std::memcpy(&cnt, &b.w, 4);
                if (cnt == 0) continue;
                const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                const double surf = 2.0 * (dx * dy + dy * dz + dx * dz);
                const double frac = std::min(1.0, (2.0 * g.zpad + std::min(dx, dy)) / dz);
                met += surf * cnt * (8.0 + 40.0 * frac);
            }
            const double per_visit = g.nz == 1 ? 37.0 : 42.0;
            const double per_length = per_visit * 0.5 * (2.0 / g.h + (g.nz == 1 ? 0.0 : 1.0 / g.hz)) +
                                      met / (4.0 * vol);
            const double half = 0.5 * kStepLen;
            const double cost = kWalkSetup * (sx + half) * (sy + half) * (sz + half) / kStepLen +
                                per_length * vol;
            if (cost < best) {
                best = cost;
                out = std::move(t);
                found = true;
            }
        }
    }
    if (found) return true;
    float h = 2.5f * base, hz0 = hz_req;
    for (int attempt = 0; attempt < 200; ++attempt, h *= 1.25f, hz0 *= 1.25f)
        if (try_build(h, zmode_req, hz0, out)) return true;
    return false;
}

}  // namespace pp
