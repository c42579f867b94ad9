// k_ls_sw.hip -- Solis-Wets local search with SWAT3 step scaling
// (perform_LS, kernel3.cl:30-436), as a persistent work-queue kernel.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_launch.h"

namespace adgpu {

// Each team takes LS slots (run, entity) from the queue until it is empty.
// One loop trip is one evaluation; a slot's iteration is either one trip
// (the "+ deviation" candidate improved) or two (then "- deviation" is tried):
//   dev_g = rho * U(-1,1) * [U < 1/sqrt(genes)] * scale_g     (kernel3.cl:155-193)
//   scale = base_dmov*sqrt3 (translation), base_dang*sqrt3/sqrt(atoms)
//           (orientation), base_dang*sqrt3/sqrt(genes) (torsions)
//   success: bias = 0.6 bias +- 0.4 dev; failure of both: bias *= 0.5
//   rho *= 2 after cons_limit successes, *= 0.5 after cons_limit failures;
//   stop at max_iters or rho <= rho_lower_bound (kernel3.cl:395-414).
// Gene g draws from its own LCG stream (work-item g of the reference).
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_sw(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    const int gene_start = (a.L->true_ligand_atoms == 0) * 6;
    const float gene_scale = a.gene_scale;
    const float sdmov = S.base_dmov_mul_sqrt3, sdang_l = S.base_dang_mul_sqrt3 * a.lig_scale,
                sdang_g = S.base_dang_mul_sqrt3 * gene_scale;
    uint32_t state[GPL];
    float off[GPL], cand[GPL], bias[GPL], dev[GPL];
    bool active[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) {
        const int g = tl + TW * j;
        active[j] = g >= gene_start && g < NG;
        state[j] = 0u; off[j] = 0.f; cand[j] = 0.f; bias[j] = 0.f; dev[j] = 0.f;
    }
    float off_e = 0.f, rho = 1.0f;
    int cons_succ = 0, cons_fail = 0, iters = 0, evals = 0, phase = 0, t = 0;
    size_t gi = 0;
    bool done = false, need = true;
    uint32_t steps = 0;              // wavefront evaluation steps (tallied once per wavefront)
    for (;;) {
        while (need) {   // uniform within the team
            int tt = 0;
            if (tl == 0) tt = atomicAdd(a.B.ls_counter, 1);
            t = TM::bcast_i(tt, 0);
            if (t >= ntot) {
                done = true; need = false;
#pragma unroll
                for (int j = 0; j < GPL; ++j) cand[j] = 0.f;
                break;
            }
            const int run = t / nls, le = t - run * nls;
            uint32_t* st = a.B.prng + (size_t)t * 64;
            const int ent = ls_entity<TW>(S, le, st);
            gi = (size_t)run * pop + ent;
#pragma unroll
            for (int j = 0; j < GPL; ++j) {
                const int g = tl + TW * j;
                state[j] = st[g];
                off[j] = a.B.pop_next[gi * kGenes + g];
                cand[j] = off[j]; bias[j] = 0.f; dev[j] = 0.f;
            }
            off_e = a.B.ene_next[gi];
            rho = 1.0f; cons_succ = 0; cons_fail = 0; iters = 0; evals = 0; phase = 0;
            if (iters < S.max_iters && rho > S.rho_lower_bound) {
                need = false;
            } else {   // nothing to do: the slot keeps its genotype
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const int g = tl + TW * j;
                    float o = off[j];
                    if (g >= 3) map_angle(o);
                    if (g < NG) a.B.pop_next[gi * kGenes + g] = o;
                }
                if (tl == 0) a.B.ene_next[gi] = off_e;
            }
        }
        if (!TM::any_in_wave(!done)) {
            if (lane_id() == 0 && steps != 0u) atomicAdd(&a.B.ls_counter[8], (int)steps);
            break;
        }
        ++steps;
        if (!done && phase == 0) {
#pragma unroll
            for (int j = 0; j < GPL; ++j) {
                if (j > 0 && NG <= TW) break;   // genes beyond the first lane set are inactive
                if (!active[j]) continue;
                const int g = tl + TW * j;
                const float r1 = randf(state[j]);
                const float r2 = randf(state[j]);
                float d = rho * (2.0f * r1 - 1.0f) * (float)(r2 < gene_scale);
                d *= (g < 3) ? sdmov : ((g < 6) ? sdang_l : sdang_g);
                dev[j] = d;
                cand[j] = off[j] + d + bias[j];
            }
        }
        float grad[GPL];
        const float ce = evaluate<TW, NC, MF, false>(X, cand, grad);
        if (!done) {
            evals++;
            bool iter_end = false;
            if (phase == 0) {
                if (ce < off_e) {
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        if (active[j]) { off[j] = cand[j]; bias[j] = 0.6f * bias[j] + 0.4f * dev[j]; }
                    off_e = ce; cons_succ++; cons_fail = 0;
                    iter_end = true;
                } else {
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        if (active[j]) cand[j] = off[j] - dev[j] - bias[j];
                    phase = 1;
                }
            } else {
                if (ce < off_e) {
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        if (active[j]) { off[j] = cand[j]; bias[j] = 0.6f * bias[j] - 0.4f * dev[j]; }
                    off_e = ce; cons_succ++; cons_fail = 0;
                } else {
#pragma unroll
                    for (int j = 0; j < GPL; ++j)
                        if (active[j]) bias[j] = 0.5f * bias[j];
                    cons_succ = 0; cons_fail++;
                }
                phase = 0;
                iter_end = true;
            }
            if (iter_end) {
                iters++;
                if (cons_succ >= S.cons_limit) { rho *= 2.0f; cons_succ = 0; }
                else if (cons_fail >= S.cons_limit) { rho *= 0.5f; cons_fail = 0; }
                if (!(iters < S.max_iters && rho > S.rho_lower_bound)) {
                    uint32_t* st = a.B.prng + (size_t)t * 64;
#pragma unroll
                    for (int j = 0; j < GPL; ++j) {
                        const int g = tl + TW * j;
                        st[g] = state[j];
                        float o = off[j];
                        if (g >= 3) map_angle(o);
                        if (g < NG) a.B.pop_next[gi * kGenes + g] = o;
                    }
                    if (tl == 0) { a.B.evals_new[gi] += evals; a.B.ene_next[gi] = off_e; }
                    need = true;
                }
            }
        }
    }
}

hipError_t launch_ls_sw(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.num_ls;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    a.lig_scale = 1.0f / sqrtf((float)la.lig_host->num_atoms);
    a.gene_scale = 1.0f / sqrtf((float)la.lig_host->num_genes);
    const Shape sh = choose_shape(*la.lig_host, la.team_hint, ls_group_threads());
    const int per_group = sh.threads / sh.tw;
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_ls_sw, sh, la.grid.format, {
        const int want = (n + per_group - 1) / per_group;
        const int cap = persistent_cap(reinterpret_cast<const void*>(kfn), sh.lds, sh.threads, la.share);
        const int groups = cap > 0 && cap < want ? cap : want;
        hipLaunchKernelGGL(kfn, dim3(groups), dim3(sh.threads), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
