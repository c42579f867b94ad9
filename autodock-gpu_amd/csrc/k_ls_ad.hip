// k_ls_ad.hip -- ADADELTA local search (kernel_ad.cl:60-549, with the
// ADADELTA_AUTOSTOP varrho schedule), as a persistent work-queue kernel.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_items.h"
#include "adgpu_launch.h"

namespace adgpu {

// Each team takes LS slots (run, entity) from the queue until it is empty
// and runs one ADADELTA search per slot (adgpu_items.h).
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_ad(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    AdItem<TW> it;
#pragma unroll
    for (int j = 0; j < GPL; ++j) { it.geno[j] = 0.f; it.best[j] = 0.f; it.sqg[j] = 0.f; it.sqd[j] = 0.f; }
    it.best_e = INFINITY; it.varrho = 1.0f; it.cons_succ = 0; it.cons_fail = 0; it.iters = 0;
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
                for (int j = 0; j < GPL; ++j) it.geno[j] = 0.f;
            } else {
                const int run = t / nls, le = t - run * nls;
                const int ent = ls_entity<TW>(S, le, a.B.prng + (size_t)t * 64);
                gi = (size_t)run * pop + ent;
                ad_begin<TW>(it, a.B.pop_next + gi * kGenes, NG);
            }
        }
        if (!TM::any_in_wave(!done)) break;
        float g[GPL];
        const float e = evaluate<TW, NC, MF, true>(X, it.geno, g);
        if (!done && ad_step<TW>(it, e, g, NG, S.max_iters)) {
            ad_store<TW>(it, NG, a.B.pop_next + gi * kGenes, a.B.ene_next + gi, a.B.evals_new + gi);
            need = true;
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
