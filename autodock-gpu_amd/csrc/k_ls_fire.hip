// k_ls_fire.hip -- FIRE local search (gradient_minFire, kernel_fire.cl:75-895).
// A first pass chooses every slot's entity and computes its float gradient
// (gpu_calc_gradient, kernel_fire.cl:318); the persistent search kernel then
// runs the FIRE iterations from the work queue.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_launch.h"

namespace adgpu {

// One team per LS slot: entity choice (the slot's LCG draw) and the float
// gradient of its genotype, kept in ls_ent / ls_grad for k_ls_fire.
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_fire_init(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    const bool live = team_idx < ntot;
    if (!TM::any_in_wave(live)) return;
    const int t = live ? team_idx : ntot - 1;
    const int run = t / nls, le = t - run * nls;
    int ent = 0;
    if (live) ent = ls_entity<TW>(S, le, a.B.prng + (size_t)t * 64);
    const size_t gi = (size_t)run * pop + ent;
    float geno[GPL], grad[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) {
        const int g = tl + TW * j;
        geno[j] = g < NG ? a.B.pop_next[gi * kGenes + g] : 0.f;
    }
    float_gradient<TW, NC, MF>(X, geno, grad);
    if (!live) return;
#pragma unroll
    for (int j = 0; j < GPL; ++j) a.B.ls_grad[(size_t)t * kGenes + tl + TW * j] = grad[j];
    if (tl == 0) a.B.ls_ent[t] = ent;
}

// Per slot (kernel_fire.cl:318-880): v = -g a0/|g| (a0 = 0.2), dt = 10/3;
// each step x' = x + dt v, energy and merged gradient at x'; P = -g'.v;
// P < 0: v = -g a0/|g|, dt *= 0.8, alpha = a0; otherwise after more than 5
// successes dt = min(1.2 dt, 10), alpha *= 0.99; then
// v = (1-alpha) v - g' alpha |v|/|g'|; best so far kept; stop at max_iters or
// dt <= 1e-6.
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_fire(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    float geno[GPL], best[GPL], grad[GPL], vel[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) { geno[j] = 0.f; best[j] = 0.f; grad[j] = 0.f; vel[j] = 0.f; }
    float energy = 0.f, best_e = 0.f, alpha = 0.2f, dt = 10.0f / 3.0f;
    int count_success = 0, iters = 0;
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
                for (int j = 0; j < GPL; ++j) { geno[j] = 0.f; vel[j] = 0.f; grad[j] = 0.f; }
            } else {
                const int run = t / nls;
                const int ent = a.B.ls_ent[t];
                gi = (size_t)run * pop + ent;
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const int g = tl + TW * j;
                    geno[j] = g < NG ? a.B.pop_next[gi * kGenes + g] : 0.f;
                    grad[j] = a.B.ls_grad[(size_t)t * kGenes + g];
                    best[j] = geno[j];
                }
                energy = a.B.ene_next[gi];
                best_e = energy;
                alpha = 0.2f; dt = 10.0f / 3.0f; count_success = 0; iters = 0;
            }
            float gg = 0.f;
#pragma unroll
            for (int j = 0; j < GPL; ++j) gg += grad[j] * grad[j];
            const float inv_gn = 0.2f * (1.0f / sqrtf(TM::sum(gg)));
#pragma unroll
            for (int j = 0; j < GPL; ++j) vel[j] = done ? 0.f : -grad[j] * inv_gn;
        }
        if (!TM::any_in_wave(!done)) {
            if (lane_id() == 0 && steps != 0u) atomicAdd(&a.B.ls_counter[8], (int)steps);
            break;
        }
        ++steps;
        float cand[GPL], cg[GPL];
#pragma unroll
        for (int j = 0; j < GPL; ++j) cand[j] = geno[j] + dt * vel[j];
        const float ce = evaluate<TW, NC, MF, true>(X, cand, cg);
        float pw = 0.f, og = 0.f, ng = 0.f;
#pragma unroll
        for (int j = 0; j < GPL; ++j) { pw += -cg[j] * vel[j]; og += grad[j] * grad[j]; ng += cg[j] * cg[j]; }
        const float power = TM::sum(pw);
        const float old_gn2 = TM::sum(og);
        const float new_gn2 = TM::sum(ng);
        if (!done) {
            if (power < 0.0f) {
                const float inv_gn = 0.2f * (1.0f / sqrtf(old_gn2));
#pragma unroll
                for (int j = 0; j < GPL; ++j) vel[j] = -grad[j] * inv_gn;
                count_success = 0; alpha = 0.2f; dt = dt * 0.8f;
            } else {
                count_success++;
                if (count_success > 5) { dt = fminf(dt * 1.2f, 10.0f); alpha = alpha * 0.99f; }
            }
            energy = ce;
#pragma unroll
            for (int j = 0; j < GPL; ++j) { geno[j] = cand[j]; grad[j] = cg[j]; }
        }
        float vv = 0.f;
#pragma unroll
        for (int j = 0; j < GPL; ++j) vv += vel[j] * vel[j];
        const float vn = sqrtf(TM::sum(vv));
        if (!done) {
            const float ign = 1.0f / sqrtf(new_gn2);
            const float vdg = alpha * vn * ign;
#pragma unroll
            for (int j = 0; j < GPL; ++j) vel[j] = (1 - alpha) * vel[j] - grad[j] * vdg;
            iters++;
            if (energy < best_e) {
                best_e = energy;
#pragma unroll
                for (int j = 0; j < GPL; ++j) best[j] = geno[j];
            }
            if (!(iters < S.max_iters && dt > 1e-6f)) {
#pragma unroll
                for (int j = 0; j < GPL; ++j) {
                    const int g = tl + TW * j;
                    float b = best[j];
                    if (g >= 3) map_angle(b);
                    if (g < NG) a.B.pop_next[gi * kGenes + g] = b;
                }
                if (tl == 0) { a.B.evals_new[gi] += iters; a.B.ene_next[gi] = best_e; }
                need = true;
            }
        }
    }
}

hipError_t launch_ls_fire(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.num_ls;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint);
    const int per_group = 256 / sh.tw;
    const int want = (n + per_group - 1) / per_group;
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_fire_init, sh, la.grid.format, {
        hipLaunchKernelGGL(kfn, dim3(want), dim3(256), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    if (err != hipSuccess) return err;
    // the persistent loop, in workgroups of ls_group_threads()
    const Shape shl = choose_shape(*la.lig_host, la.team_hint, ls_group_threads());
    const int want_l = (n + shl.threads / shl.tw - 1) / (shl.threads / shl.tw);
    ADGPU_DISPATCH(k_ls_fire, shl, la.grid.format, {
        const int cap = persistent_cap(reinterpret_cast<const void*>(kfn), shl.lds, shl.threads, la.share);
        const int groups = cap > 0 && cap < want_l ? cap : want_l;
        hipLaunchKernelGGL(kfn, dim3(groups), dim3(shl.threads), shl.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
