// k_eval.hip -- batched pose evaluation, the per-pose parity entry point
// (gpu_calc_energy calcenergy.cl:132-554, gpu_calc_energrad
// calcMergedEneGra.cl:46-936, gpu_calc_gradient calcgradient.cl:109-1003).
#include <mutex>
#include <type_traits>

#include "adgpu_device.h"

namespace adgpu {

// One team per pose.  Teams past the end evaluate the last pose and write
// nothing, so every wavefront runs the evaluation with all lanes.
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_eval_poses(Launch a) {
    KERNEL_PROLOGUE(TW, NC, PL)
    const bool live = team_idx < a.n;
    if (!TM::any_in_wave(live)) return;
    const int id = live ? team_idx : a.n - 1;
    float gene[GPL], grad[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) gene[j] = a.in_geno[(size_t)id * kGenes + tl + TW * j];
    if (a.mode == ADGPU_EVAL_GRADIENT) {
        float_gradient<TW, NC, MF>(X, gene, grad);
    } else if (a.mode == ADGPU_EVAL_ENERGRAD) {
        const float e = evaluate<TW, NC, MF, true>(X, gene, grad);
        if (live && tl == 0 && a.out_e) a.out_e[id] = e;
    } else {
        const float e = evaluate<TW, NC, MF, false>(X, gene, grad);
        if (live && tl == 0 && a.out_e) a.out_e[id] = e;
        return;
    }
    if (live && a.out_g) {
#pragma unroll
        for (int j = 0; j < GPL; ++j) a.out_g[(size_t)id * kGenes + tl + TW * j] = grad[j];
    }
}

// Staging area of adgpu_eval_poses_noalloc: genotypes [max][64], gradients
// [max][64], energies [max].  Each device holds its own copy for the life of
// the process; the host reaches it only through the address that
// hipGetSymbolAddress returns on that device.
static __device__ __align__(256) float scratch[(size_t)ADGPU_NOALLOC_MAX_POSES * (2 * kGenes + 1)];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device, float** out) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the caller has made `device` current
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Host buffers in and out through the device's staging area: no allocation.
hipError_t eval_poses_noalloc(const LaunchArgs& la, float* scratch_dev, int mode, const float* genotypes, int n,
                              float* energies, float* grads) {
    if (n <= 0) return hipSuccess;
    if (n > ADGPU_NOALLOC_MAX_POSES || !scratch_dev) return hipErrorInvalidValue;
    const size_t cap = (size_t)ADGPU_NOALLOC_MAX_POSES;
    float* d_geno = scratch_dev;
    float* d_grad = scratch_dev + cap * kGenes;
    float* d_ene = scratch_dev + 2 * cap * kGenes;
    hipError_t e = hipMemcpyAsync(d_geno, genotypes, sizeof(float) * kGenes * (size_t)n, hipMemcpyHostToDevice,
                                  la.stream);
    if (e == hipSuccess) e = launch_eval_poses(la, mode, d_geno, n, d_ene, d_grad);
    if (e == hipSuccess && energies && mode != ADGPU_EVAL_GRADIENT)
        e = hipMemcpyAsync(energies, d_ene, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, la.stream);
    if (e == hipSuccess && grads && mode != ADGPU_EVAL_ENERGY)
        e = hipMemcpyAsync(grads, d_grad, sizeof(float) * kGenes * (size_t)n, hipMemcpyDeviceToHost, la.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(la.stream);
    return e;
}

size_t lds_bytes(const LigandDev& Lh, int team_hint, int* team_width) {
    Shape s = choose_shape(Lh, team_hint);
    if (team_width) *team_width = s.tw;
    return s.lds;
}

hipError_t launch_eval_poses(const LaunchArgs& la, int mode, const float* genotypes, int n, float* energies,
                             float* grads) {
    if (n <= 0) return hipSuccess;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.in_geno = genotypes; a.out_e = energies; a.out_g = grads; a.n = n;
    a.mode = mode;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint);
    const int per_group = 256 / sh.tw;
    const dim3 grid((n + per_group - 1) / per_group);
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_eval_poses, sh, la.grid.format, {
        hipLaunchKernelGGL(kfn, grid, dim3(256), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
