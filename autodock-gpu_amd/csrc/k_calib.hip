// k_calib.hip -- the device's VALU issue rate, measured: the peak that the
// local-search kernels' VALU instruction rate is reported against.  No
// counterpart in the reference (measurement only).
#include <hip/hip_runtime.h>

#include "adgpu_internal.h"

namespace adgpu {

namespace {

constexpr int kCalibChains = 8;     // independent FMA chains per lane
constexpr int kCalibUnroll = 4;     // FMAs per chain per loop trip

// Each lane runs kCalibChains independent chains of v_fma_f32 (full-rate FP32
// VALU instructions), kCalibUnroll per chain per trip: iters * 32 VALU
// instructions per wavefront, plus the loop's scalar bookkeeping.
__global__ __launch_bounds__(256) void k_valu_rate(float* out, int iters, float b, float c) {
    float a[kCalibChains];
#pragma unroll
    for (int k = 0; k < kCalibChains; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kCalibUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < kCalibChains; ++k) a[k] = __builtin_fmaf(a[k], b, c);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kCalibChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

hipError_t measure_valu_rate(hipStream_t stream, float* scratch, int blocks, int iters, double* wave_instr_per_s) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {   // warm-up launch, then the timed one
        hipLaunchKernelGGL(k_valu_rate, dim3(blocks), dim3(256), 0, stream, scratch, iters / 8 + 1, 0.999f, 1e-3f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_valu_rate, dim3(blocks), dim3(256), 0, stream, scratch, iters, 0.999f, 1e-3f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) {
        const double waves = (double)blocks * (256 / 64);
        const double instr = waves * (double)iters * (kCalibChains * kCalibUnroll);
        *wave_instr_per_s = instr / (ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return e;
}

}  // namespace adgpu
