// Host-side check of the target grid: builds the tables with
// pp_grid_build.h and runs GridSearch and BruteSearch (pp_search.h, the code
// the propagation kernel inlines) on the CPU over random segments, counting
// disagreements. Built by tests/test_grid_host.py with hipcc for the host,
// under AddressSanitizer, so any out-of-bounds table access aborts the run.
//   usage: grid_check <scene> <seed> <n_rays> <cell> [cell_z]
//          (cell <= 0: auto; cell_z 0: cubic, < 0: columns, > 0: z edge)
//   prints: rays hits mismatches nx ny nz n_items far_fat
//   (far_fat: brute-force hits of a direction longer than unit whose hit point
//   lies outside the sphere by more than the grid's margin; not mismatches)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../photonpropagation.jl_amd/csrc/pp_grid_build.h"

using namespace pp;

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    const int scene = atoi(argv[1]);
    const unsigned seed = (unsigned)atoi(argv[2]);
    const long n_rays = atol(argv[3]);
    const float cell = (float)atof(argv[4]);
    const float cell_z = argc > 5 ? (float)atof(argv[5]) : 0.0f;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    std::normal_distribution<float> N(0.0f, 1.0f);
    std::vector<float> xyz, rad;
    if (scene == 0) {  // 16 strings x 20 modules, r = 0.3
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 20; ++k) {
                    xyz.insert(xyz.end(), {-75.0f + 50.0f * i, -75.0f + 50.0f * j, -95.0f + 10.0f * k});
                    rad.push_back(0.3f);
                }
    } else if (scene == 1) {  // dense random, overlapping spheres of mixed size
        for (int i = 0; i < 600; ++i) {
            xyz.insert(xyz.end(), {20.0f * (U(rng) - 0.5f), 20.0f * (U(rng) - 0.5f), 20.0f * (U(rng) - 0.5f)});
            rad.push_back(0.05f + 1.5f * U(rng));
        }
    } else {  // 1000 modules, 10 x 10 strings
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j)
                for (int k = 0; k < 10; ++k) {
                    xyz.insert(xyz.end(), {-225.0f + 50.0f * i, -225.0f + 50.0f * j, -45.0f + 10.0f * k});
                    rad.push_back(0.3f);
                }
    }
    const int n = (int)rad.size();
    std::vector<float4> tgt(n);
    for (int i = 0; i < n; ++i)
        tgt[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], rad[i] * rad[i]);
    GridTables t;
    if (!build_grid_tables(xyz.data(), rad.data(), n, cell, cell_z, 600, 64 * 1024, t)) {
        printf("build failed\n");
        return 3;
    }
    // exact-size copies so that the sanitizer sees every overrun
    std::vector<float4> cells(t.cells);
    std::vector<uint16_t> items(t.items.begin(), t.items.begin() + std::max(t.g.n_items, 1));
    GridSearch grid{tgt.data(), cells.data(), items.data(), t.g};
    BruteSearch brute{tgt.data(), n};
    long hits = 0, mism = 0, n_fat = 0;
    const float span = scene == 1 ? 30.0f : 300.0f;
    for (long r = 0; r < n_rays; ++r) {
        f3 pos, dir;
        if (r % 2 == 0 && n > 0) {  // aim near a random target, every 13th from afar
            const int k = (int)(U(rng) * n) % n;
            const float d = r % 13 == 4 ? 2000.0f + 3000.0f * U(rng) : 1.0f + 60.0f * U(rng);
            f3 a = make3(N(rng), N(rng), N(rng));
            const float an = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
            dir = make3(a.x / an, a.y / an, a.z / an);
            pos = make3(xyz[3 * k] - d * dir.x + 0.5f * N(rng), xyz[3 * k + 1] - d * dir.y + 0.5f * N(rng),
                        xyz[3 * k + 2] - d * dir.z + 0.5f * N(rng));
        } else {
            pos = make3(span * (U(rng) - 0.5f), span * (U(rng) - 0.5f), span * (U(rng) - 0.5f));
            f3 a = make3(N(rng), N(rng), N(rng));
            if (r % 7 == 1) a.x = 0.0f;  // axis-parallel segments
            if (r % 11 == 1) a.y = 0.0f, a.z = 0.0f;
            const float an = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z) + 1e-30f;
            dir = make3(a.x / an, a.y / an, a.z / an);
        }
        float step = -logf(1.0f - U(rng)) * 40.0f;
        if (r % 13 == 4) step += 5000.0f;  // long enough to reach the far target
        float d1 = 0.0f, d2 = 0.0f;
        const int f1 = brute(pos, dir, step, d1);
        const int f2 = grid(pos, dir, step, d2);
        if (f1 >= 0) ++hits;
        if (f1 != f2 || (f1 >= 0 && d1 != d2)) {
            // The reference's discriminant keeps a^2 (|dir|^2 - 1): for a
            // direction longer than unit its hit point lies sqrt(r^2 + e d^2)
            // from the centre, outside the sphere. A brute-force hit whose
            // point lies outside the sphere by more than the binning margin is
            // such a hit (rays from afar), which the grid need not find.
            bool fat = false;
            if (f1 >= 0 && (f2 < 0 || f2 > f1)) {
                const double hx = pos.x + (double)d1 * dir.x - xyz[3 * f1];
                const double hy = pos.y + (double)d1 * dir.y - xyz[3 * f1 + 1];
                const double hz = pos.z + (double)d1 * dir.z - xyz[3 * f1 + 2];
                const double m = 1e-3 * std::max(t.g.h, t.g.hz) + 1e-3;
                fat = std::sqrt(hx * hx + hy * hy + hz * hz) > rad[f1] + m;
            }
            if (fat) {
                ++n_fat;
            } else {
                if (mism < 5)
                    printf("mismatch ray %ld: brute %d (%g) grid %d (%g)\n", r, f1, d1, f2, d2);
                ++mism;
            }
        }
    }
    printf("%ld %ld %ld %d %d %d %d %ld\n", n_rays, hits, mism, t.g.nx, t.g.ny, t.g.nz, t.g.n_items,
           n_fat);
    return 0;
}
