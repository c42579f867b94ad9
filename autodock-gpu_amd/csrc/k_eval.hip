// k_eval.hip -- batched pose evaluation, the per-pose parity entry point
// (gpu_calc_energy calcenergy.cl:132-554, gpu_calc_energrad
// calcMergedEneGra.cl:46-936, gpu_calc_gradient calcgradient.cl:109-1003).
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
