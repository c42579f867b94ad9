// k_ga.hip -- K1 (initial energies), K4 (GA generation), K2 (eval counts)
// and the local-search dispatch.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_items.h"
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
    uint32_t* st = a.B.prng + (size_t)idx * 64;
    K4Item<TW> it;
    k4_prepare<TW>(a.S, a.B.pop_cur + (size_t)run * pop * kGenes, a.B.ene_cur + (size_t)run * pop, st, ent, it);
    float in[GPL], grad[GPL];
    k4_genes<TW>(it, NG, in);
    const float e = evaluate<TW, NC, MF, false>(X, in, grad);
    if (!live) return;
    k4_store<TW>(it, NG, e, a.B.pop_next + (size_t)idx * kGenes, a.B.ene_next + idx, a.B.evals_new + idx, st);
}

// K2: per-run eval totals (kernel2.cl:26-62); one wavefront per run.
__global__ __launch_bounds__(64) void k_sum_evals(SearchDev S, JobBuffers B) {
    const int run = blockIdx.x;
    const int lane = lane_id();
    if (!JOB_STOPPED(B)) {
        int s = 0;
        for (int e = lane; e < S.pop_size; e += 64) s += B.evals_new[run * S.pop_size + e];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) B.evals_runs[run] += s;
    }
    if (lane == 0 && B.evals_copy != nullptr) B.evals_copy[run] = B.evals_runs[run];
}

// Clears the local-search work queue and, for ADADELTA hand-offs, its
// counters, the n entries of the hand-off ring and the 2n regrouping seats
// that follow it (a kernel rather than a memset, so the launch sequence can be
// captured into a graph).  The fault flag counter[3] and the job's tallies
// (counter[4], [7], [8]) stay: they are cleared when a search starts.
__global__ __launch_bounds__(256) void k_reset_queue(int32_t* counter, int32_t* ring, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 || i == 5 || i == 6) counter[i] = 0;
    if (ring != nullptr && i < 3 * n) ring[i] = -1;
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
    const int n = sp.num_runs * sp.num_ls;
    int32_t* ring = method == ADGPU_LS_AD && sp.ls_yield > 0 ? b.ls_ring : nullptr;
    hipLaunchKernelGGL(k_reset_queue, dim3(ring != nullptr ? (3 * n + 255) / 256 : 1), dim3(256), 0, la.stream,
                       b.ls_counter, ring, n);
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
