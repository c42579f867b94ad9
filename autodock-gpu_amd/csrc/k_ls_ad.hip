// k_ls_ad.hip -- ADADELTA local search (kernel_ad.cl:60-549, with the
// ADADELTA_AUTOSTOP varrho schedule), as a persistent work-queue kernel.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_launch.h"

namespace adgpu {

// Each team takes LS slots (run, entity) from the queue until it is empty.
// Per slot: gradient steps delta = -g sqrt((E[dx^2]+eps)/(E[g^2]+eps)) with
// rho = 0.8, eps = 1e-2 (kernel_ad.cl:40-41, 443-452); the best genotype so far is
// kept; varrho doubles after 4 successes and halves after 4 failures; the
// search ends at max_iters or varrho < 0.01 (kernel_ad.cl:476-515).
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_ad(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    const float RHO = 0.8f, EPS = 1e-2f;
    float geno[GPL], best[GPL], sqg[GPL], sqd[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) { geno[j] = 0.f; best[j] = 0.f; sqg[j] = 0.f; sqd[j] = 0.f; }
    float best_e = INFINITY, varrho = 1.0f;
    int cons_succ = 0, cons_fail = 0, iters = 0;
    size_t gi = 0;
    bool done = false, need = true;
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
                    best[j] = geno[j]; sqg[j] = 0.f; sqd[j] = 0.f;
                }
                best_e = INFINITY; varrho = 1.0f; cons_succ = 0; cons_fail = 0; iters = 0;
            }
        }
        if (!TM::any_in_wave(!done)) break;
        float g[GPL];
        const float e = evaluate<TW, NC, MF, true>(X, geno, g);
        if (!done) {
            const bool better = e < best_e;
#pragma unroll
            for (int j = 0; j < GPL; ++j) {
                if (j > 0 && NG <= TW) break;   // genes beyond the first lane set are all zero
                if (better) best[j] = geno[j];
                sqg[j] *= RHO;
                sqg[j] += (1.0f - RHO) * g[j] * g[j];
                const float delta = -1.0f * g[j] * sqrtf((sqd[j] + EPS) / (sqg[j] + EPS));
                sqd[j] *= RHO;
                sqd[j] += (1.0f - RHO) * delta * delta;
                geno[j] += delta;
            }
            if (better) { best_e = e; cons_succ++; cons_fail = 0; }
            else { cons_succ = 0; cons_fail++; }
            iters++;
            if (cons_succ >= 4) { varrho *= 2.0f; cons_succ = 0; }
            else if (cons_fail >= 4) { varrho *= 0.5f; cons_fail = 0; }
            if (!(iters < S.max_iters && varrho > 0.01f)) {
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

hipError_t launch_ls_ad(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.num_ls;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint, ls_group_threads());
    const int per_group = sh.threads / sh.tw;
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_ls_ad, sh, la.grid.format, {
        const int want = (n + per_group - 1) / per_group;
        const int cap = persistent_cap(reinterpret_cast<const void*>(kfn), sh.lds, sh.threads, la.share);
        const int groups = cap > 0 && cap < want ? cap : want;
        hipLaunchKernelGGL(kfn, dim3(groups), dim3(sh.threads), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
