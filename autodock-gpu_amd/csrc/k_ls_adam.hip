// k_ls_adam.hip -- ADAM local search (cuda/kernel_adam.cu:53-416, without
// AD_RHO_CRITERION: a fixed number of iterations), as a persistent
// work-queue kernel.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_launch.h"

namespace adgpu {

// Each team takes LS slots (run, entity) from the queue until it is empty.
// Per slot: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// x -= (m / (1 - b1^(t+1))) / (sqrt(v / (1 - b2^(t+1))) + eps)
// (kernel_adam.cu:280-298); the best genotype so far is kept; max_iters
// evaluations (kernel_adam.cu:360).
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_adam(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    const float b1 = S.adam_beta1, b2 = S.adam_beta2, eps = S.adam_epsilon;
    float geno[GPL], best[GPL], mt[GPL], vt[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) { geno[j] = 0.f; best[j] = 0.f; mt[j] = 0.f; vt[j] = 0.f; }
    float best_e = INFINITY;
    int iters = 0;
    size_t gi = 0;
    bool done = false, need = true;
    uint32_t steps = 0;              // wavefront evaluation steps (tallied once per wavefront)
    for (;;) {
        if (need) {   // uniform within the team
            int t = 0;
            if (tl == 0) t = atomicAdd(a.B.ls_counter, 1);
            t = TM::bcast_i(t, 0);
            need = false;
            if (t >= ntot) {
                done = true;
#pragma unroll
                for (int j = 0; j < GPL; ++j) geno[j] = 0.f;
            } else {
                const int run = t / nls, le = t - run * nls;
                const int ent = ls_entity<TW>(S, le, a.B.prng + (size_t)t * 64);
                gi = (size_t)run * pop + ent;
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const int g = tl + TW * j;
                    geno[j] = g < NG ? a.B.pop_next[gi * kGenes + g] : 0.f;
                    best[j] = geno[j]; mt[j] = 0.f; vt[j] = 0.f;
                }
                best_e = INFINITY; iters = 0;
            }
        }
        if (!TM::any_in_wave(!done)) {
            if (lane_id() == 0 && steps != 0u) atomicAdd(&a.B.ls_counter[8], (int)steps);
            break;
        }
        ++steps;
        float g[GPL];
        const float e = evaluate<TW, NC, MF, true>(X, geno, g);
        if (!done) {
            const float b1p = 1.0f - powf(b1, 1.0f + (float)iters);
            const float b2p = 1.0f - powf(b2, 1.0f + (float)iters);
            const bool better = e < best_e;
#pragma unroll
            for (int j = 0; j < GPL; ++j) {
                if (j > 0 && NG <= TW) break;   // genes beyond the first lane set are all zero
                if (better) best[j] = geno[j];
                mt[j] = b1 * mt[j] + (1.0f - b1) * g[j];
                vt[j] = b2 * vt[j] + (1.0f - b2) * g[j] * g[j];
                const float mp = mt[j] / b1p, vp = vt[j] / b2p;
                geno[j] -= mp / (sqrtf(vp) + eps);
            }
            iters++;
            if (better) best_e = e;
            if (!(iters < S.max_iters)) {
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const int gg = tl + TW * j;
                    float b = best[j];
                    if (gg >= 3) map_angle(b);
                    if (gg < NG) a.B.pop_next[gi * kGenes + gg] = b;
                }
                if (tl == 0) { a.B.evals_new[gi] += iters; a.B.ene_next[gi] = best_e; }
                need = true;
            }
        }
    }
}

hipError_t launch_ls_adam(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.num_ls;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint, ls_group_threads());
    const int per_group = sh.threads / sh.tw;
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_ls_adam, sh, la.grid.format, {
        const int want = (n + per_group - 1) / per_group;
        const int cap = persistent_cap(reinterpret_cast<const void*>(kfn), sh.lds, sh.threads, la.share);
        const int groups = cap > 0 && cap < want ? cap : want;
        hipLaunchKernelGGL(kfn, dim3(groups), dim3(sh.threads), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
