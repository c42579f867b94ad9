// k_autostop.hip -- the AutoStop test on the device.
//
// The reference tests the population every as_frequency generations on the
// host: AutoStop::check_if_satisfactory (host/inc/autostop.hpp:181-231) on the
// downloaded energies, leaving the generation loop when the test passes and
// total_evals > min_as_evals (host/src/performdocking.cpp.Cuda:687-696).  Here
// one workgroup runs the same test between two generations on the resident
// energies and raises the job's stop flag, so the host never has to drain the
// device to decide.
//
// The result must equal the reference host's bit for bit: its sums are
// float sums in population order, so the accumulation stays sequential in
// that order (one lane for the total, one lane per energy bin), while the bin
// of every energy -- the part with real work, a search over pop_size bin
// edges compared in double -- is found in parallel by binary search over the
// monotone edge sequence.  Division and square root go through double,
// which rounds to the correctly rounded float result.
#include "adgpu_internal.h"

namespace adgpu {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 2048;   // energies staged in LDS per pass

__device__ __forceinline__ float div_rn(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ float sqrt_rn(float a) { return (float)__builtin_sqrt((double)a); }

// average(), autostop.hpp:62-67
__device__ float avg3(const float* s) {
    if (s[2] < 1.0f) return 0.0f;
    return div_rn(s[0], s[2]);
}

// stddev(), autostop.hpp:69-76
__device__ float std3(const float* s) {
#pragma clang fp contract(off)
    if (s[2] < 1.0f) return 0.0f;
    const float a = s[1] * s[2];
    const float b = s[0] * s[0];
    const float sq = a - b;
    if (((double)__builtin_fabsf(sq) <= 0.000001) || (sq < 0.0f)) return 0.0f;
    return div_rn(sqrt_rn(sq), s[2]);
}

// Index of the first bin m in [0, ntop) with e < lo + m * delta
// (autostop.hpp:88-95), ntop if none, -1 if e is not below the threshold.
// The edges lo + (double)((float)m * delta) do not decrease with m when
// delta >= 0, so the first one above e is found by bisection.
__device__ int bin_of(float e, float threshold, double lo, float delta, int ntop) {
#pragma clang fp contract(off)
    if (!(e < threshold)) return -1;
    const double ed = (double)e;
    if (!(delta >= 0.0f)) {
        for (int m = 0; m < ntop; ++m)
            if (ed < lo + (double)((float)m * delta)) return m;
        return ntop;
    }
    int a = 0, b = ntop;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (ed < lo + (double)((float)m * delta)) b = m;
        else a = m + 1;
    }
    return a;
}

// tabulate_energies(), autostop.hpp:78-100, adding to acc (average_sd2_N in
// LDS).  Lane 0 keeps the sums over all samples and the lowest energy; lane
// k > 0 owns bins k-1, k-1+255, ...; each adds its samples in population
// order, as the reference's single loop does.
__device__ void tabulate(const float* ene, int n, int ntop, float threshold, float thres_stddev, float delta,
                         float* E, int* B, float* acc, float* best_out) {
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    const double lo = (double)threshold - 2.0 * (double)thres_stddev;
    float s0 = acc[0], s1 = acc[1], s2 = acc[2];
    float best = 16777216.0f;
    for (int c0 = 0; c0 < n; c0 += kChunk) {
        const int len = n - c0 < kChunk ? n - c0 : kChunk;
        for (int i = t; i < len; i += kThreads) {
            const float e = ene[c0 + i];
            E[i] = e;
            B[i] = bin_of(e, threshold, lo, delta, ntop);
        }
        __syncthreads();
        if (t == 0) {
            for (int i = 0; i < len; ++i) {
                const float e = E[i];
                if (e < best) best = e;
                if (B[i] >= 0) {
                    s0 += e;
                    const float sq = e * e;
                    s1 += sq;
                    s2 += 1.0f;
                }
            }
        } else {
            for (int m = t - 1; m < ntop; m += kThreads - 1) {
                float* a = acc + 3 * (m + 1);
                float a0 = a[0], a1 = a[1], a2 = a[2];
                for (int i = 0; i < len; ++i) {
                    if (B[i] == m) {
                        const float e = E[i];
                        a0 += e;
                        const float sq = e * e;
                        a1 += sq;
                        a2 += 1.0f;
                    }
                }
                a[0] = a0; a[1] = a1; a[2] = a2;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        acc[0] = s0; acc[1] = s1; acc[2] = s2;
        *best_out = best;
    }
    __syncthreads();
}

// set_stats(), autostop.hpp:102-126 (one lane).
__device__ void set_stats(AutoStopDev& S, float* acc, int ntop, uint32_t ncream, float stopstd) {
#pragma clang fp contract(off)
    S.curr_avg = avg3(acc);
    S.curr_std = std3(acc);
    S.best_n = (uint32_t)acc[2];
    acc[0] = 0.0f; acc[1] = 0.0f; acc[2] = 0.0f;
    uint32_t taken = 0;
    for (int m = 0; m < ntop; ++m) {
        const float* b = acc + 3 * (m + 1);
        if (!(b[2] >= 1.0f && taken < ncream)) continue;
        if (acc[2] < 4.0f || (double)__builtin_fabsf(avg3(acc) - avg3(b)) < 2.0 * (double)stopstd) {
            acc[0] += b[0];
            acc[1] += b[1];
            acc[2] += b[2];
            taken++;
        }
    }
    if (taken > 0) {
        S.curr_avg = avg3(acc);
        S.curr_std = std3(acc);
        S.best_n = (uint32_t)acc[2];
    }
}

__device__ void fill_record(adgpu_autostop_record& r, const AutoStopDev& S, const float* sums, int64_t generation,
                            int64_t total, float threshold, int stop) {
    r.generation = generation;
    r.total_evals = total;
    r.threshold = threshold;
    r.average = S.curr_avg;
    r.stddev = S.curr_std;
    r.samples = (int32_t)S.best_n;
    r.best_energy = S.overall_best;
    r.autostopped = S.autostopped;
    r.stop = stop;
    r.roll_average = avg3(sums);
    r.roll_stddev = std3(sums);
    r.roll_samples = (int32_t)(uint32_t)sums[2];
}

}  // namespace

// mode 0: one test (check_if_satisfactory + the min_as_evals rule); a job
// whose stop flag is already raised is left unchanged.  mode 1: the final
// statistics of output_final_stddev (autostop.hpp:233-244) into *out.
__global__ __launch_bounds__(kThreads) void k_autostop(AutoStopParams P, const float* ene,
                                                       const int32_t* evals_runs, AutoStopDev* st, float* gbins,
                                                       adgpu_autostop_record* log, int32_t* stop, int64_t generation,
                                                       int mode, adgpu_autostop_record* out) {
#pragma clang fp contract(off)
    __shared__ float E[kChunk];
    __shared__ int B[kChunk];
    __shared__ AutoStopDev S;
    __shared__ int64_t part[kThreads];
    __shared__ float best_sh;
    extern __shared__ float acc[];   // average_sd2_N: 3 * (pop_size + 1)
    if (mode == 0 && *stop) return;
    const int t = threadIdx.x;
    const int ntop = P.pop_size, n = P.num_runs * P.pop_size, nb = 3 * (ntop + 1);
    int64_t s = 0;
    for (int r = t; r < P.num_runs; r += kThreads) s += evals_runs[r];
    part[t] = s;
    if (t == 0) S = *st;
    for (int i = t; i < nb; i += kThreads) acc[i] = gbins[i];
    __syncthreads();
    int64_t total = 0;
    for (int i = 0; i < kThreads; ++i) total += part[i];

    if (mode == 0) {
        const int reps = generation == 0 ? 9 : 1;
        for (int c = 0; c < reps; ++c) {
            if (!((double)__builtin_fabsf(S.curr_avg) > 0.00001)) break;   // S is uniform here
            for (int i = t; i < nb; i += kThreads) acc[i] = 0.0f;
            __syncthreads();
            if (t == 0) S.threshold_used = S.threshold;
            tabulate(ene, n, ntop, S.threshold, S.thres_stddev, S.delta_energy, E, B, acc, &best_sh);
            if (t == 0) {
                S.overall_best = best_sh;
                if (S.first_time) {
                    S.curr_avg = avg3(acc);
                    S.curr_std = std3(acc);
                    S.best_n = (uint32_t)acc[2];
                    S.thres_stddev = S.curr_std;
                    S.threshold = S.curr_avg + S.thres_stddev;
                    S.first_time = 0;
                } else {
                    set_stats(S, acc, ntop, (uint32_t)P.ncream, P.stopstd);
                    S.thres_stddev = (S.curr_std < 0.5f * P.stopstd) ? P.stopstd : S.curr_std;
                    const float num = (float)(uint32_t)P.ncream * S.thres_stddev;
                    S.threshold = S.curr_avg + div_rn(num, (float)S.best_n);
                }
                S.delta_energy = (float)(2.0 * (double)S.thres_stddev / (double)(uint32_t)(ntop - 1));
            }
            __syncthreads();
        }
        if (t == 0) {
            // rolling window of the last four tests, autostop.hpp:205-224
            float* r = S.rolling + 4 * S.roll_count;
            const float bn = (float)S.best_n;
            r[0] = S.curr_avg * bn;
            const float v2 = S.curr_std * S.curr_std;
            const float a2 = S.curr_avg * S.curr_avg;
            r[1] = (v2 + a2) * bn;
            r[2] = bn;
            r[3] = S.overall_best;
            S.roll_count = (S.roll_count + 1u) % 4u;
            const float* q = S.rolling;
            acc[0] = q[0] + q[4] + q[8] + q[12];
            acc[1] = q[1] + q[5] + q[9] + q[13];
            acc[2] = q[2] + q[6] + q[10] + q[14];
            float best_avg = q[3] + q[7] + q[11] + q[15];
            best_avg = div_rn(best_avg, 4.0f);
            const float b3 = q[3] * q[3], b7 = q[7] * q[7], b11 = q[11] * q[11], b15 = q[15] * q[15];
            float best_var = b3 + b7 + b11 + b15;
            best_var = div_rn(best_var, 4.0f);
            const float ba2 = best_avg * best_avg;
            best_var -= ba2;
            if ((std3(acc) < P.stopstd) && (generation >= (int64_t)(4u * (uint32_t)P.as_frequency)) &&
                (best_var <= 0.00002f))
                S.autostopped = 1;
            const int halt = S.autostopped && total > P.min_as_evals;
            adgpu_autostop_record rec;
            fill_record(rec, S, acc, generation, total, S.threshold_used, halt);
            log[S.num_tests % P.log_cap] = rec;
            S.num_tests++;
            if (halt) *stop = 1;
        }
    } else {
        if (!S.autostopped) {
            tabulate(ene, n, ntop, S.threshold, S.thres_stddev, S.delta_energy, E, B, acc, &best_sh);
            if (t == 0) {
                S.overall_best = best_sh;
                set_stats(S, acc, ntop, (uint32_t)P.ncream, P.stopstd);
            }
        }
        if (t == 0) {
            adgpu_autostop_record rec;
            fill_record(rec, S, acc, generation, total, S.threshold, 0);
            *out = rec;
        }
    }
    __syncthreads();
    for (int i = t; i < nb; i += kThreads) gbins[i] = acc[i];
    if (t == 0) *st = S;
}

hipError_t launch_autostop(hipStream_t s, const AutoStopParams& p, const float* ene, const int32_t* evals_runs,
                           AutoStopDev* st, float* bins, adgpu_autostop_record* log, int32_t* stop,
                           int64_t generation, int mode, adgpu_autostop_record* out) {
    const size_t lds = sizeof(float) * 3 * ((size_t)p.pop_size + 1);
    hipLaunchKernelGGL(k_autostop, dim3(1), dim3(kThreads), lds, s, p, ene, evals_runs, st, bins, log, stop,
                       generation, mode, out);
    return hipGetLastError();
}

}  // namespace adgpu
