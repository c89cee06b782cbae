// pp_bench.hip - issue-rate microbenchmarks of the vector ALU, for the
// roofline of the propagation kernel (an addition with no reference
// counterpart). Each lane runs 8 independent dependency chains, so that
// latency is hidden and the SIMD issues every cycle it can; the chains'
// results are stored, so nothing is optimised away.
//   PP_MB_FMA_F32   v_fma_f32 (2 FLOP each; the build issues no packed FP32)
//   PP_MB_TRANS_F32 v_rcp_f32 (transcendental unit)
//   PP_MB_MAD_U64   v_mad_u64_u32 (the Philox product)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/photonprop_hip.h"

// the library's per-thread error message (pp_last_error)
void pp_set_error_message(const char* msg);

namespace {

constexpr int kIters = 2048;
constexpr int kChains = 8;

__global__ void __launch_bounds__(256) mb_fma(float* out, float a, float b) {
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = (float)(threadIdx.x + c) * 1e-3f;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = fmaf(x[c], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) mb_trans(float* out, float a) {
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = 1.0f + (float)(threadIdx.x + c) * a;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = __builtin_amdgcn_rcpf(x[c]);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) mb_mad64(float* out, uint32_t m) {
    uint32_t x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = threadIdx.x * 2654435761u + (uint32_t)c;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            const uint64_t p = (uint64_t)m * (uint64_t)x[c];
            x[c] = (uint32_t)(p >> 32) ^ (uint32_t)p;
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s ^= x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = __uint_as_float(s);
}

}  // namespace

extern "C" int pp_microbench(int device, int which, double* rate) {
    if (!rate || which < 0 || which > 2) {
        pp_set_error_message("pp_microbench: bad arguments");
        return PP_ERR_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) {
        pp_set_error_message("pp_microbench: hipSetDevice failed");
        return PP_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        pp_set_error_message("pp_microbench: hipGetDeviceProperties failed");
        return PP_ERR_HIP;
    }
    // 8 waves of 64 per SIMD, 4 SIMDs per CU
    const int block = 256;
    const int blocks = prop.multiProcessorCount * 8;
    const size_t n = (size_t)blocks * block;
    float* out = nullptr;
    if (hipMalloc(&out, n * sizeof(float)) != hipSuccess) {
        pp_set_error_message("pp_microbench: device allocation failed");
        return PP_ERR_HIP;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    int rc = PP_OK;
    for (int rep = 0; rep < 4 && rc == PP_OK; ++rep) {
        (void)hipEventRecord(e0, 0);
        if (which == PP_MB_FMA_F32)
            hipLaunchKernelGGL(mb_fma, dim3(blocks), dim3(block), 0, 0, out, 0.999f, 1e-3f);
        else if (which == PP_MB_TRANS_F32)
            hipLaunchKernelGGL(mb_trans, dim3(blocks), dim3(block), 0, 0, out, 1e-6f);
        else
            hipLaunchKernelGGL(mb_mad64, dim3(blocks), dim3(block), 0, 0, out, 0xD2511F53u);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) rc = PP_ERR_HIP;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;  // first launch warms up
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc != PP_OK) {
        pp_set_error_message("pp_microbench: kernel failed");
        return rc;
    }
    const double ops = (double)n * kIters * kChains * (which == PP_MB_FMA_F32 ? 2.0 : 1.0);
    *rate = ops / (best * 1e-3);
    return PP_OK;
}
