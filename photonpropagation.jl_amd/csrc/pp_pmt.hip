// pp_pmt.hip - PMT-level detection of the POM (include/photonprop_hip.h,
// pp_pom_pmt_hits): check_pmt_hit(::POM{<:POMAcceptance}), src/pom.jl:246-305,
// one lane per photon hit. The 16 PMT directions and both group widths are
// kernel arguments (SGPRs); the acceptance tables sit in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "photonprop_hip.h"
#include "pp_device.h"
#include "pp_hits.h"

void pp_set_error_message(const char* msg);

namespace {

constexpr int kPmts = 16;
constexpr int kMaxWl = 1024;
// PMT group of PMT j (src/pom.jl:172-174): PMTs 0-3 and 8-11 (25 deg from the
// equator) are group 1, 4-7 and 12-15 (57.5 deg) group 2: index (j >> 2) & 1

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

// Inputs of the detection kernel, one of two layouts:
//   SoA      position relative to the module centre in radii, wavelength,
//            photon id and source index per hit (pp_pom_pmt_hits)
//   records  the context's hit records, module centre and radius from the
//            target table (pp_pom_pmt_hits_ctx)
struct PmtSoA {
    const float* rel_pos;
    const float* wl;
    const uint64_t* pid;
    const int32_t* src;
    __device__ __forceinline__ void load(int64_t i, float& px, float& py, float& pz, float& w,
                                         uint64_t& id, uint32_t& sx) const {
        px = rel_pos[3 * i];
        py = rel_pos[3 * i + 1];
        pz = rel_pos[3 * i + 2];
        w = wl[i];
        id = pid[i];
        sx = (uint32_t)src[i];
    }
};

struct PmtRecords {
    const pp::HitRecord* rec;
    const float4* targets;  // x, y, z, radius
    __device__ __forceinline__ void load(int64_t i, float& px, float& py, float& pz, float& w,
                                         uint64_t& id, uint32_t& sx) const {
        const float4* r = reinterpret_cast<const float4*>(rec + i);
        const float4 a = r[0], c = r[2], d = r[3];
        const uint64_t raw = (uint64_t)__float_as_uint(d.x) | ((uint64_t)__float_as_uint(d.y) << 32);
        const float4 t = targets[((raw >> pp::kPidBits) & 0xFFFF) - 1];
        // (position - centre) / radius, as the host computes it for a hit table
        px = (a.x - t.x) / t.w;
        py = (a.y - t.y) / t.w;
        pz = (a.z - t.z) / t.w;
        w = c.y;
        id = raw & (pp::kMaxPhotonsPerSource - 1);
        sx = __float_as_uint(d.w);
    }
};

template <typename In>
struct PmtArgs {
    pp::PmtSetup s;
    const float* tab;  // wl | acc1 | acc2, n_wl each
    In in;
    int64_t n;
    int32_t* out;
    unsigned long long* n_det;
};

// linear interpolation in a sorted table, 0 outside (extrapolation_bc = 0)
__device__ __forceinline__ float interp0(const float* x, const float* y, int n, float v) {
    if (!(v >= x[0] && v <= x[n - 1])) return 0.0f;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] <= v) lo = mid;
        else hi = mid;
    }
    const float t = (v - x[lo]) / (x[hi] - x[lo]);
    return y[lo] + t * (y[hi] - y[lo]);
}

// Grid-stride over the hits (a few workgroups per CU), so that the detected
// count takes one atomic per workgroup: a wave ballot per pass, the waves'
// sums in LDS at the end.
template <typename In>
__global__ void __launch_bounds__(256) pom_pmt_kernel(PmtArgs<In> a) {
    __shared__ float tab[3 * kMaxWl];
    __shared__ unsigned long long wave_det[4];
    const int n_wl = a.s.n_wl;
    for (int i = threadIdx.x; i < 3 * n_wl; i += blockDim.x) tab[i] = a.tab[i];
    __syncthreads();
    unsigned long long n_det = 0;  // wave-uniform
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.n; base += stride) {
        const int64_t i = base + threadIdx.x;
        bool det = false;
        if (i < a.n) {
            float px, py, pz, wl;
            uint64_t id;
            uint32_t sx;
            a.in.load(i, px, py, pz, wl, id, sx);
            const float acc[2] = {interp0(tab, tab + n_wl, n_wl, wl),
                                  interp0(tab, tab + 2 * n_wl, n_wl, wl)};
            float p[kPmts], sum = 0.0f;
#pragma unroll
            for (int j = 0; j < kPmts; ++j) {
                const int g = (j >> 2) & 1;
                const float c = fminf(fmaxf(px * a.s.dir[j][0] + py * a.s.dir[j][1] +
                                                pz * a.s.dir[j][2],
                                            -1.0f), 1.0f);
                const float pt = acosf(c);
                // sin(acos c) = sqrt((1 - c)(1 + c)), exact factors (no sinf)
                const float st = __builtin_amdgcn_sqrtf((1.0f - c) * (1.0f + c));
                // Rayleigh pdf over the density 0.5 sin(pt) of a uniform flux
                const float pdf = pt * a.s.inv_s2[g] * __expf(-0.5f * pt * pt * a.s.inv_s2[g]);
                p[j] = st != 0.0f ? pdf * (2.0f * __builtin_amdgcn_rcpf(st)) * acc[g] * 0.125f : 0.0f;
                sum += p[j];
            }
            const uint4 r = pp::philox(make_uint4(PP_CTR_PMT, (uint32_t)id, (uint32_t)(id >> 32), sx),
                                       a.s.k0, a.s.k1);
            int pmt = 0;
            if (pp::u01(r.x) < sum) {
                // categorical draw: first j whose cumulative weight exceeds u * sum
                const float u = pp::u01(r.y) * sum;
                float cum = 0.0f;
                for (int j = 0; j < kPmts; ++j) {
                    cum += p[j];
                    if (p[j] > 0.0f) {
                        pmt = j + 1;  // the last PMT with weight if rounding leaves u >= cum
                        if (u < cum) break;
                    }
                }
                det = true;
            }
            a.out[i] = pmt;
        }
        n_det += (unsigned long long)__popcll(__ballot(det));
    }
    if (__lane_id() == 0) wave_det[threadIdx.x >> 6] = n_det;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += wave_det[w];
        if (s) atomicAdd(a.n_det, s);
    }
}

template <typename In>
int launch(const PmtArgs<In>& a, hipStream_t stream, int n_cu) {
    const int64_t blocks = std::min<int64_t>((a.n + 255) / 256, (int64_t)std::max(n_cu, 1) * 8);
    hipLaunchKernelGGL(pom_pmt_kernel<In>, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256),
                       0, stream, a);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// PMT (theta, phi) of the POM, src/pom.jl:108-128
void pom_angles(double th[kPmts], double ph[kPmts]) {
    const double d2r = M_PI / 180.0;
    for (int k = 0; k < 4; ++k) {
        th[k] = (90.0 - 25.0) * d2r;
        ph[k] = k * M_PI / 2;
        th[4 + k] = (90.0 - 57.5) * d2r;
        ph[4 + k] = M_PI / 4 + k * M_PI / 2;
        th[8 + k] = (90.0 + 25.0) * d2r;
        th[12 + k] = (90.0 + 57.5) * d2r;
    }
    const double lower2[4] = {M_PI / 2, 0.0, 3 * M_PI / 2, M_PI};
    const double lower[4] = {M_PI / 4, 7 * M_PI / 4, 5 * M_PI / 4, 3 * M_PI / 4};
    for (int k = 0; k < 4; ++k) {
        ph[8 + k] = lower2[k];
        ph[12 + k] = lower[k];
    }
}

int validate(const pp_pom_model* m) {
    if (!m) return fail(PP_ERR_INVALID, "pom: NULL model");
    if (!(m->sigma_grp1 > 0.0f && m->sigma_grp2 > 0.0f) || !std::isfinite(m->sigma_grp1) ||
        !std::isfinite(m->sigma_grp2))
        return fail(PP_ERR_INVALID, "pom: sigmas must be finite and > 0");
    for (float v : m->orientation)
        if (!std::isfinite(v)) return fail(PP_ERR_INVALID, "pom: non-finite orientation");
    if (m->n_wl < 2 || m->n_wl > kMaxWl)
        return fail(PP_ERR_INVALID, "pom: n_wl must be in [2, 1024]");
    return PP_OK;
}

}  // namespace

namespace pp {

int pmt_setup(const pp_pom_model* m, const float* wl_grid, const float* acc1, const float* acc2,
              uint64_t seed, PmtSetup* s, float* tab) {
    if (int rc = validate(m)) return rc;
    if (!wl_grid || !acc1 || !acc2) return fail(PP_ERR_INVALID, "pom: NULL acceptance table");
    for (int i = 1; i < m->n_wl; ++i)
        if (!(wl_grid[i] > wl_grid[i - 1]))
            return fail(PP_ERR_INVALID, "pom: wavelength grid must increase");
    float dirs[kPmts * 3];
    if (int rc = pp_pom_pmt_directions(m, dirs)) return rc;
    for (int j = 0; j < kPmts; ++j)
        for (int r = 0; r < 3; ++r) s->dir[j][r] = dirs[3 * j + r];
    s->inv_s2[0] = 1.0f / (m->sigma_grp1 * m->sigma_grp1);
    s->inv_s2[1] = 1.0f / (m->sigma_grp2 * m->sigma_grp2);
    s->n_wl = m->n_wl;
    s->k0 = (uint32_t)seed;
    s->k1 = (uint32_t)(seed >> 32);
    const size_t nw = (size_t)m->n_wl;
    for (size_t k = 0; k < nw; ++k) {
        tab[k] = wl_grid[k];
        tab[nw + k] = acc1[k];
        tab[2 * nw + k] = acc2[k];
    }
    return PP_OK;
}

int pmt_launch_records(const PmtSetup& s, const float* tab, const HitRecord* rec, int64_t n,
                       const float4* targets, int32_t* out, unsigned long long* count,
                       hipStream_t stream, int n_cu) {
    PmtArgs<PmtRecords> a{};
    a.s = s;
    a.tab = tab;
    a.in.rec = rec;
    a.in.targets = targets;
    a.n = n;
    a.out = out;
    a.n_det = count;
    return launch(a, stream, n_cu);
}

}  // namespace pp

namespace {

// Scratch of pp_pom_pmt_hits_noalloc: the acceptance table (3 x kMaxWl
// floats) and the detected-hit counter, in a file-scope device array that
// each GPU the library runs on holds until the process ends. Host code uses
// the per-device address from hipGetSymbolAddress, never the array's name.
constexpr size_t kScratchTableBytes = 3 * kMaxWl * sizeof(float);
constexpr size_t kScratchBytes = kScratchTableBytes + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// The scratch array's address on `device` (current), looked up once per
// device on its first call.
int scratch_address(int device, char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PP_ERR_HIP, std::string("pp_pom_pmt_hits_noalloc: hipGetSymbolAddress: ") +
                                        hipGetErrorString(e));
        g_scratch[device] = (char*)p;
    }
    *out = g_scratch[device];
    return PP_OK;
}

// Clears the counter with a kernel (no memset node where a caller captures
// the stream into a graph).
__global__ void clear_count_kernel(unsigned long long* c) {
    if (threadIdx.x == 0) *c = 0ull;
}

}  // namespace

extern "C" {

int pp_pom_pmt_directions(const pp_pom_model* m, float* out) {
    if (int rc = validate(m)) return rc;
    if (!out) return fail(PP_ERR_INVALID, "pp_pom_pmt_directions: NULL output");
    double th[kPmts], ph[kPmts];
    pom_angles(th, ph);
    for (int j = 0; j < kPmts; ++j) {
        const double v[3] = {std::sin(th[j]) * std::cos(ph[j]), std::sin(th[j]) * std::sin(ph[j]),
                             std::cos(th[j])};
        // R takes e_z to e_x (rotation by +90 deg about e_y): (x, y, z) -> (z, y, -x)
        const double w[3] = {v[2], v[1], -v[0]};
        for (int r = 0; r < 3; ++r)
            out[3 * j + r] = (float)(m->orientation[3 * r] * w[0] + m->orientation[3 * r + 1] * w[1] +
                                     m->orientation[3 * r + 2] * w[2]);
    }
    return PP_OK;
}

int pp_pom_pmt_hits(int device, const pp_pom_model* m, const float* wl_grid, const float* acc1,
                    const float* acc2, int64_t n, const float* rel_pos3, const float* wavelength,
                    const uint64_t* photon_id, const int32_t* source_ix, uint64_t seed,
                    int32_t* pmt_id_out, int64_t* n_detected, float* kernel_ms) {
    if (int rc = validate(m)) return rc;
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(PP_ERR_INVALID, "pp_pom_pmt_hits: n out of range");
    if (n_detected) *n_detected = 0;
    if (kernel_ms) *kernel_ms = 0.0f;
    if (n == 0) return PP_OK;
    if (!rel_pos3 || !wavelength || !photon_id || !source_ix || !pmt_id_out)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits: NULL array");
    PmtArgs<PmtSoA> a{};
    std::vector<float> tab(3 * (size_t)m->n_wl);
    if (int rc = pp::pmt_setup(m, wl_grid, acc1, acc2, seed, &a.s, tab.data())) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits: no such HIP device");
    PP_HIP(hipSetDevice(device));
    (void)hipGetLastError();
    a.n = n;
    const size_t nw = (size_t)m->n_wl;
    // one allocation: table | rel_pos | wl | pid | src | out | counter
    const size_t off_pos = (3 * nw * 4 + 255) & ~(size_t)255;
    const size_t off_wl = off_pos + (((size_t)n * 12 + 255) & ~(size_t)255);
    const size_t off_pid = off_wl + (((size_t)n * 4 + 255) & ~(size_t)255);
    const size_t off_src = off_pid + (((size_t)n * 8 + 255) & ~(size_t)255);
    const size_t off_out = off_src + (((size_t)n * 4 + 255) & ~(size_t)255);
    const size_t off_cnt = off_out + (((size_t)n * 4 + 255) & ~(size_t)255);
    const size_t bytes = off_cnt + 8;
    char* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PP_ERR_NOMEM, "pp_pom_pmt_hits: device allocation failed");
    }
    int rc = PP_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == PP_OK) rc = fail(PP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(hipMemcpy(d, tab.data(), 3 * nw * 4, hipMemcpyHostToDevice), "upload table");
    chk(hipMemcpy(d + off_pos, rel_pos3, (size_t)n * 12, hipMemcpyHostToDevice), "upload positions");
    chk(hipMemcpy(d + off_wl, wavelength, (size_t)n * 4, hipMemcpyHostToDevice), "upload wavelengths");
    chk(hipMemcpy(d + off_pid, photon_id, (size_t)n * 8, hipMemcpyHostToDevice), "upload ids");
    chk(hipMemcpy(d + off_src, source_ix, (size_t)n * 4, hipMemcpyHostToDevice), "upload sources");
    chk(hipMemset(d + off_cnt, 0, 8), "clear counter");
    if (rc == PP_OK) {
        a.tab = (const float*)d;
        a.in.rel_pos = (const float*)(d + off_pos);
        a.in.wl = (const float*)(d + off_wl);
        a.in.pid = (const uint64_t*)(d + off_pid);
        a.in.src = (const int32_t*)(d + off_src);
        a.out = (int32_t*)(d + off_out);
        a.n_det = (unsigned long long*)(d + off_cnt);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (kernel_ms) {
            chk(hipEventCreate(&e0), "event");
            chk(hipEventCreate(&e1), "event");
            if (rc == PP_OK) chk(hipEventRecord(e0, 0), "event record");
        }
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
            cus <= 0)
            cus = 256;
        if (rc == PP_OK) rc = launch(a, 0, cus);
        if (kernel_ms && rc == PP_OK) {
            chk(hipEventRecord(e1, 0), "event record");
            chk(hipEventSynchronize(e1), "event sync");
            if (rc == PP_OK) chk(hipEventElapsedTime(kernel_ms, e0, e1), "event time");
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        chk(hipMemcpy(pmt_id_out, d + off_out, (size_t)n * 4, hipMemcpyDeviceToHost), "copy ids");
        unsigned long long cnt = 0;
        chk(hipMemcpy(&cnt, d + off_cnt, 8, hipMemcpyDeviceToHost), "copy count");
        if (n_detected) *n_detected = (int64_t)cnt;
    }
    (void)hipFree(d);
    return rc;
}

int pp_pom_pmt_hits_noalloc(int device, const pp_pom_model* m, const float* wl_grid,
                             const float* acc1, const float* acc2, int64_t n,
                             const float* d_rel_pos3, const float* d_wavelength,
                             const uint64_t* d_photon_id, const int32_t* d_source_ix,
                             uint64_t seed, int32_t* d_pmt_id_out, int64_t* n_detected,
                             void* stream) {
    if (n_detected) *n_detected = 0;
    if (int rc = validate(m)) return rc;  // n_wl <= kMaxWl: the table fits the scratch
    if (n < 0 || n > ((int64_t)1 << 31))
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_noalloc: n out of range");
    if (n == 0) return PP_OK;
    if (!d_rel_pos3 || !d_wavelength || !d_photon_id || !d_source_ix || !d_pmt_id_out)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_noalloc: NULL array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev ||
        device >= kMaxDevices)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_noalloc: no such HIP device");
    PP_HIP(hipSetDevice(device));
    (void)hipGetLastError();
    char* sc = nullptr;
    if (int rc = scratch_address(device, &sc)) return rc;
    PmtArgs<PmtSoA> a{};
    float tab[3 * kMaxWl];
    if (int rc = pp::pmt_setup(m, wl_grid, acc1, acc2, seed, &a.s, tab)) return rc;
    const hipStream_t st = (hipStream_t)stream;
    auto* d_cnt = (unsigned long long*)(sc + kScratchTableBytes);
    // from pageable host memory: HIP stages the bytes before the call returns,
    // so the stack copy may go out of scope while the transfer is queued
    PP_HIP(hipMemcpyAsync(sc, tab, 3 * (size_t)m->n_wl * sizeof(float), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(clear_count_kernel, dim3(1), dim3(64), 0, st, d_cnt);
    PP_HIP(hipGetLastError());
    a.tab = (const float*)sc;
    a.in.rel_pos = d_rel_pos3;
    a.in.wl = d_wavelength;
    a.in.pid = d_photon_id;
    a.in.src = d_source_ix;
    a.n = n;
    a.out = d_pmt_id_out;
    a.n_det = d_cnt;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (int rc = launch(a, st, cus)) return rc;
    if (n_detected) {
        unsigned long long cnt = 0;
        PP_HIP(hipMemcpyAsync(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, st));
        PP_HIP(hipStreamSynchronize(st));
        *n_detected = (int64_t)cnt;
    }
    return PP_OK;
}

}  // extern "C"
