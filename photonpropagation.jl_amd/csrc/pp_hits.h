// pp_hits.h - the device hit record shared by the propagation kernels
// (pp_lib.hip) and the PMT detection kernels (pp_pmt.hip), and the internal
// interface between the two translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/photonprop_hip.h"

namespace pp {

// One hit record: 64 bytes, 64-byte aligned, so that a hit is written as two
// full 32-byte sectors (four 16-byte stores by one lane).
//   f[0..2] position, f[3..5] direction, f[6..8] initial direction,
//   f[9] wavelength, f[10] time, f[11] distance travelled,
//   id = photon id (bits 0..39) | module index, 1-based (bits 40..55),
//   n_steps, source index.
struct alignas(64) HitRecord {
    float f[12];
    uint64_t id;
    int32_t n_steps;
    int32_t source_ix;
};
static_assert(sizeof(HitRecord) == 64, "hit record is 64 bytes");
constexpr uint64_t kPidBits = 40;
constexpr uint64_t kMaxPhotonsPerSource = 1ull << kPidBits;

// PMT detection set-up of one call: the 16 PMT directions of the module
// orientation, 1/sigma^2 of both PMT groups, the acceptance table length and
// the Philox key of the draws.
struct PmtSetup {
    float dir[16][3];
    float inv_s2[2];
    int32_t n_wl;
    uint32_t k0, k1;
};

// Validates the model and its acceptance table and fills `s` and the table
// wl | acc1 | acc2 (3 * n_wl floats) for the kernels. Host code.
int pmt_setup(const pp_pom_model* m, const float* wl_grid, const float* acc1, const float* acc2,
              uint64_t seed, PmtSetup* s, float* tab);

// The detection kernel over n context hit records (module centre and radius
// from `targets`, one (x, y, z, radius) per target, indexed by the record's
// 1-based module index), writing out[i] and adding the detected count to
// *count (which the caller clears). `tab` is the device copy of the table.
int pmt_launch_records(const PmtSetup& s, const float* tab, const HitRecord* rec, int64_t n,
                       const float4* targets, int32_t* out, unsigned long long* count,
                       hipStream_t stream, int n_cu);

// The hit records of a context's last propagation (device pointer, in
// buffer order), their count, the context's stream and device. Host code,
// for the surrogate's label reduction.
int ctx_hit_records(pp_ctx* ctx, const HitRecord** rec, int64_t* n, hipStream_t* stream,
                    int* device);

}  // namespace pp
