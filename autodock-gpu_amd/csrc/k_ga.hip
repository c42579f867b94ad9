// k_ga.hip -- K1 (initial energies), K4 (GA generation), K2 (eval counts)
// and the local-search dispatch.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_launch.h"

namespace adgpu {

// K1: energies of the initial population (kernel1.cl:28-145).
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_initpop(Launch a) {
    KERNEL_PROLOGUE(TW, NC, PL)
    const int ne = a.S.num_runs * a.S.pop_size;
    const bool live = team_idx < ne;
    if (!TM::any_in_wave(live)) return;
    const int id = live ? team_idx : ne - 1;
    float gene[GPL], grad[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) gene[j] = a.B.pop_cur[(size_t)id * kGenes + tl + TW * j];
    const float e = evaluate<TW, NC, MF, false>(X, gene, grad);
    if (live && tl == 0) { a.B.ene_cur[id] = e; a.B.evals_new[id] = 1; }
}

// Best entity of a run as gpu_perform_elitist_selection finds it with 64
// work-items (auxiliary_genetic.cl:89-162): work-item v keeps the first strict
// minimum of entities v, v+64, ...; work-item 0 then scans work-items 1..63 in
// order and takes strictly smaller energies.  Emulated on a team: each lane
// holds work-items l and l+32 (TW = 32) or l (TW = 64), a tree over the
// work-items keeps the lowest index among equal energies and skips NaN, and
// work-item 0's own result wins unless the tree found something strictly
// smaller (a NaN there is kept, as the serial scan keeps it).
template <int TW>
DEV void elite_of_run(const float* ecur, int pop, float& best_e, int& best_id) {
    const int tl = Team<TW>::lane();
    float ve = 0.f;          // running candidate of this lane over its work-items
    int vw = 64, vi = 0;     // work-item (64 = none) and entity index
#pragma unroll
    for (int j = 0; j < Team<TW>::kGPL; ++j) {
        const int w = tl + TW * j;
        if (w >= pop) continue;
        float be = ecur[w];
        int bi = w;
        for (int k = w + 64; k < pop; k += 64) {
            const float ek = ecur[k];
            if (ek < be) { be = ek; bi = k; }
        }
        const bool ok = !(be != be);
        if (ok && (vw == 64 || be < ve)) { ve = be; vw = w; vi = bi; }
    }
    // own result of work-item 0 (every lane computes it)
    float e0 = ecur[0];
    int i0 = 0;
    for (int k = 64; k < pop; k += 64) {
        const float ek = ecur[k];
        if (ek < e0) { e0 = ek; i0 = k; }
    }
#pragma unroll
    for (int o = TW / 2; o > 0; o >>= 1) {
        const float oe = __shfl_xor(ve, o);
        const int ow = __shfl_xor(vw, o), oi = __shfl_xor(vi, o);
        const bool take = (ow < 64) && ((vw >= 64) || (oe < ve) || (!(ve < oe) && ow < vw));
        if (take) { ve = oe; vw = ow; vi = oi; }
    }
    if (!(e0 == e0) || vw >= 64 || !(ve < e0)) { best_e = e0; best_id = i0; }
    else { best_e = ve; best_id = vi; }
}

// K4: elitism, tournament selection, crossover, mutation and offspring
// energies (kernel4.cl:32-336).  The elitist team (entity 0) evaluates its
// copy as well, keeping every wavefront's lanes together, but keeps the
// copied energy and its LCG streams unchanged as the reference does.
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_gen(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const int pop = a.S.pop_size, NG = a.S.num_genes;
    const int ne = a.S.num_runs * pop;
    const bool live = team_idx < ne;
    if (!TM::any_in_wave(live)) return;
    const int idx = live ? team_idx : ne - 1;
    const int run = idx / pop, ent = idx - run * pop;
    const float* cur = a.B.pop_cur + (size_t)run * pop * kGenes;
    const float* ecur = a.B.ene_cur + (size_t)run * pop;
    uint32_t* st = a.B.prng + (size_t)idx * 64;
    float child[GPL];
    uint32_t state[GPL];
    float elite_e = 0.f;
    const bool elite = ent == 0;
    if (elite) {
        int bi;
        elite_of_run<TW>(ecur, pop, elite_e, bi);
#pragma unroll
        for (int j = 0; j < GPL; ++j) child[j] = cur[(size_t)bi * kGenes + tl + TW * j];
    } else {
#pragma unroll
        for (int j = 0; j < GPL; ++j) state[j] = st[tl + TW * j];
        float rn = 0.f;
        if (tl < 10) rn = randf(state[0]);
        // parent candidates (team lanes 0..3), binary tournaments
        const int cand = (int)((float)pop * rn);
        const float ce = (tl < 4) ? ecur[cand] : 0.f;
        const int c0 = TM::bcast_i(cand, 0), c1 = TM::bcast_i(cand, 1), c2 = TM::bcast_i(cand, 2),
                  c3 = TM::bcast_i(cand, 3);
        const float e0 = TM::bcast(ce, 0), e1 = TM::bcast(ce, 1), e2 = TM::bcast(ce, 2), e3 = TM::bcast(ce, 3);
        const float r4 = TM::bcast(rn, 4), r5 = TM::bcast(rn, 5), r6 = TM::bcast(rn, 6), r7 = TM::bcast(rn, 7),
                    r8 = TM::bcast(rn, 8);
        const float tr = a.S.tournament_rate;
        const int p0 = (e0 < e1) ? ((r4 < tr) ? c0 : c1) : ((r4 < tr) ? c1 : c0);
        const int p1 = (e2 < e3) ? ((r5 < tr) ? c2 : c3) : ((r5 < tr) ? c3 : c2);
        const bool cross = r6 < a.S.crossover_rate;
        int cp0 = (int)((float)(NG - 1) * r7), cp1 = (int)((float)(NG - 1) * r8);
        if (cp1 < cp0) { int t = cp1; cp1 = cp0; cp0 = t; }
#pragma unroll
        for (int j = 0; j < GPL; ++j) {
            const int g = tl + TW * j;
            const float v0 = cur[(size_t)p0 * kGenes + g], v1 = cur[(size_t)p1 * kGenes + g];
            float c = v0;
            if (cross) {
                if (cp0 != cp1) c = ((g <= cp0) || (g > cp1)) ? v0 : v1;
                else c = (g <= cp0) ? v0 : v1;
            }
            if (g < NG) {
                // separate rounding of every operation, as the reference source
                // writes it (no fused multiply-add): the new genes are bit-exact
#pragma clang fp contract(off)
                if (randf(state[j]) < a.S.mutation_rate) {
                    if (g < 3) c += a.S.abs_max_dmov * (2 * randf(state[j]) - 1);
                    else { c += a.S.abs_max_dang * (2 * randf(state[j]) - 1); map_angle(c); }
                }
            }
            child[j] = c;
        }
    }
    float in[GPL], grad[GPL];
#pragma unroll
    for (int j = 0; j < GPL; ++j) in[j] = (tl + TW * j) < NG ? child[j] : 0.f;
    const float e = evaluate<TW, NC, MF, false>(X, in, grad);
    if (!live) return;
#pragma unroll
    for (int j = 0; j < GPL; ++j)
        if (tl + TW * j < NG) a.B.pop_next[(size_t)idx * kGenes + tl + TW * j] = child[j];
    if (elite) {
        if (tl == 0) { a.B.ene_next[idx] = elite_e; a.B.evals_new[idx] = 0; }
    } else {
#pragma unroll
        for (int j = 0; j < GPL; ++j) st[tl + TW * j] = state[j];
        if (tl == 0) { a.B.ene_next[idx] = e; a.B.evals_new[idx] = 1; }
    }
}

// K2: per-run eval totals (kernel2.cl:26-62); one wavefront per run.
__global__ __launch_bounds__(64) void k_sum_evals(SearchDev S, JobBuffers B) {
    if (JOB_STOPPED(B)) return;
    const int run = blockIdx.x;
    const int lane = lane_id();
    int s = 0;
    for (int e = lane; e < S.pop_size; e += 64) s += B.evals_new[run * S.pop_size + e];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) B.evals_runs[run] += s;
}

// Clears the local-search work queue (a kernel rather than a memset, so the
// launch sequence can be captured into a graph).
__global__ __launch_bounds__(64) void k_reset_queue(int32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

hipError_t launch_initpop(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.pop_size;
    if (n <= 0) return hipSuccess;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint);
    const int per_group = 256 / sh.tw;
    const dim3 grid((n + per_group - 1) / per_group);
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_initpop, sh, la.grid.format, {
        hipLaunchKernelGGL(kfn, grid, dim3(256), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

hipError_t launch_gen(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.pop_size;
    if (n <= 0) return hipSuccess;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint);
    const int per_group = 256 / sh.tw;
    const dim3 grid((n + per_group - 1) / per_group);
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_gen, sh, la.grid.format, {
        hipLaunchKernelGGL(kfn, grid, dim3(256), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

hipError_t launch_sum_evals(hipStream_t s, const SearchDev& sp, const JobBuffers& b) {
    hipLaunchKernelGGL(k_sum_evals, dim3(sp.num_runs), dim3(64), 0, s, sp, b);
    return hipGetLastError();
}

hipError_t launch_ls(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b, int method) {
    if (sp.num_runs * sp.num_ls <= 0 || method == ADGPU_LS_NONE) return hipSuccess;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, la.stream, b.ls_counter);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    switch (method) {
        case ADGPU_LS_SW: return launch_ls_sw(la, sp, b);
        case ADGPU_LS_AD: return launch_ls_ad(la, sp, b);
        case ADGPU_LS_ADAM: return launch_ls_adam(la, sp, b);
        case ADGPU_LS_FIRE: return launch_ls_fire(la, sp, b);
        default: return hipSuccess;
    }
}

}  // namespace adgpu
