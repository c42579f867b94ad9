/*
 * autostop_oracle.c -- CPU restatement of AutoDock-GPU's AutoStop test.
 *
 * TEST INFRASTRUCTURE ONLY (see adock_oracle.h): the checker for the device
 * AutoStop kernel (autodock-gpu_amd/csrc/k_autostop.hip).
 *
 * Follows the AutoStop class of host/inc/autostop.hpp:36-244 as the reference
 * host build compiles it (g++ -O3 -std=c++17, x86-64 SSE, no FMA contraction):
 * running sums in float in population order, the bin edges compared in double
 * where the reference's expression promotes to double ((threshold - 2.0 *
 * thres_stddev) + m * delta_energy), and the float overloads of fabs/sqrt
 * that <math.h> brings into the global namespace in C++.
 *
 * Pinned by tests/test_oracle_autostop.py against the reference's own class
 * compiled in oracle/_ref (every state member, bit for bit, after every test,
 * on synthetic energy sequences that stop, that never stop, and on degenerate
 * populations).
 */
#include "adock_oracle.h"

#include <math.h>
#include <string.h>

/* average(), autostop.hpp:62-67 */
static float as_average(const float* s) {
    if (s[2] < 1.0f) return 0.0f;
    return s[0] / s[2];
}

/* stddev(), autostop.hpp:69-76 */
static float as_stddev(const float* s) {
    if (s[2] < 1.0f) return 0.0f;
    float sq = s[1] * s[2] - s[0] * s[0];
    if ((fabsf(sq) <= 0.000001) || (sq < 0.0)) return 0.0f;
    return sqrtf(sq) / s[2];
}

void orc_autostop_init(orc_autostop* a, int pop_size, int num_runs, float stopstd, int as_frequency) {
    memset(a, 0, sizeof(*a));
    a->num_runs = num_runs;
    a->ntop = (uint32_t)pop_size;
    a->ncream = (uint32_t)pop_size / 10u;
    a->stopstd = stopstd;
    a->as_frequency = (uint32_t)as_frequency;
    /* constructor, autostop.hpp:129-146 */
    a->first_time = 1;
    a->autostopped = 0;
    a->threshold = (float)(1 << 24);
    a->threshold_used = a->threshold;
    a->thres_stddev = a->threshold;
    a->curr_avg = (float)(-(1 << 24));
    a->curr_std = a->thres_stddev;
    a->roll_count = 0;
    a->best_n = 1;
    a->delta_energy = (float)(2.0 * (double)a->thres_stddev / (double)a->ntop);
    a->overall_best = (float)(1 << 24);
}

/* tabulate_energies(), autostop.hpp:78-100: adds to bins[0..2] every sample
 * below the threshold and to bins[3(m+1)..] those of the first bin m whose
 * upper edge lies above the energy; overall_best = lowest energy.  It does not
 * clear the sums: the test clears them first, the final statistics
 * (output_final_stddev) add to what the last test left. */
static void as_tabulate(orc_autostop* a, const float* e, float* bins) {
    const size_t n = (size_t)a->num_runs * a->ntop;
    a->overall_best = (float)(1 << 24);
    const double lo = (double)a->threshold - 2.0 * (double)a->thres_stddev;
    for (size_t i = 0; i < n; ++i) {
        const float v = e[i];
        if (v < a->overall_best) a->overall_best = v;
        if (!(v < a->threshold)) continue;
        bins[0] += v;
        bins[1] += v * v;
        bins[2] = (float)((double)bins[2] + 1.0);
        for (uint32_t m = 0; m < a->ntop; ++m) {
            const float step = (float)m * a->delta_energy;
            if ((double)v < lo + (double)step) {
                float* b = bins + 3 * ((size_t)m + 1);
                b[0] += v;
                b[1] += v * v;
                b[2] = (float)((double)b[2] + 1.0);
                break;
            }
        }
    }
}

/* set_stats(), autostop.hpp:102-126 */
static void as_set_stats(orc_autostop* a, float* bins) {
    a->curr_avg = as_average(bins);
    a->curr_std = as_stddev(bins);
    a->best_n = (uint32_t)bins[2];
    bins[0] = bins[1] = bins[2] = 0.0f;
    uint32_t taken = 0;
    for (uint32_t m = 0; m < a->ntop; ++m) {
        const float* b = bins + 3 * ((size_t)m + 1);
        if (!(b[2] >= 1.0 && taken < a->ncream)) continue;
        if (bins[2] < 4.0 || (double)fabsf(as_average(bins) - as_average(b)) < 2.0 * (double)a->stopstd) {
            bins[0] += b[0];
            bins[1] += b[1];
            bins[2] += b[2];
            taken++;
        }
    }
    if (taken > 0) {
        a->curr_avg = as_average(bins);
        a->curr_std = as_stddev(bins);
        a->best_n = (uint32_t)bins[2];
    }
}

/* check_if_satisfactory(), autostop.hpp:181-231, plus the caller's
 * "total_evals > min_as_evals" condition (performdocking.cpp.Cuda:692-696).
 * bins: the class's average_sd2_N, 3*(pop_size+1) floats kept by the caller
 * between calls (zero before the first).  Returns 1 when the run stops. */
int orc_autostop_check(orc_autostop* a, int generation, const float* energies, int64_t total_evals,
                       int64_t min_as_evals, float* bins, adgpu_autostop_record* rec) {
    const int reps = 1 + 8 * (generation == 0);
    for (int c = 0; c < reps && fabsf(a->curr_avg) > 0.00001; ++c) {
        a->threshold_used = a->threshold;
        memset(bins, 0, sizeof(float) * 3 * ((size_t)a->ntop + 1));
        as_tabulate(a, energies, bins);
        if (a->first_time) {
            a->curr_avg = as_average(bins);
            a->curr_std = as_stddev(bins);
            a->best_n = (uint32_t)bins[2];
            a->thres_stddev = a->curr_std;
            a->threshold = a->curr_avg + a->thres_stddev;
            a->delta_energy = (float)(2.0 * (double)a->thres_stddev / (double)(a->ntop - 1u));
            a->first_time = 0;
        } else {
            as_set_stats(a, bins);
            a->thres_stddev = (a->curr_std < 0.5f * a->stopstd) ? a->stopstd : a->curr_std;
            a->threshold = a->curr_avg + (float)a->ncream * a->thres_stddev / (float)a->best_n;
            a->delta_energy = (float)(2.0 * (double)a->thres_stddev / (double)(a->ntop - 1u));
        }
    }
    /* rolling window of the last four tests, autostop.hpp:205-224 */
    float* r = a->rolling + 4 * a->roll_count;
    r[0] = a->curr_avg * (float)a->best_n;
    r[1] = (a->curr_std * a->curr_std + a->curr_avg * a->curr_avg) * (float)a->best_n;
    r[2] = (float)a->best_n;
    r[3] = a->overall_best;
    a->roll_count = (a->roll_count + 1u) % 4u;
    const float* q = a->rolling;
    a->sums[0] = q[0] + q[4] + q[8] + q[12];
    a->sums[1] = q[1] + q[5] + q[9] + q[13];
    a->sums[2] = q[2] + q[6] + q[10] + q[14];
    memcpy(bins, a->sums, 3 * sizeof(float));
    float best_avg = q[3] + q[7] + q[11] + q[15];
    best_avg /= 4;
    float best_var = q[3] * q[3] + q[7] * q[7] + q[11] * q[11] + q[15] * q[15];
    best_var /= 4;
    best_var -= best_avg * best_avg;
    if ((as_stddev(a->sums) < a->stopstd) && (generation >= (int)(4u * a->as_frequency)) && (best_var <= 0.00002f))
        a->autostopped = 1;
    const int stop = a->autostopped && total_evals > min_as_evals;
    if (rec) {
        memset(rec, 0, sizeof(*rec));
        rec->generation = generation;
        rec->total_evals = total_evals;
        rec->threshold = a->threshold_used;
        rec->average = a->curr_avg;
        rec->stddev = a->curr_std;
        rec->samples = (int32_t)a->best_n;
        rec->best_energy = a->overall_best;
        rec->autostopped = a->autostopped;
        rec->stop = stop;
        rec->roll_average = as_average(a->sums);
        rec->roll_stddev = as_stddev(a->sums);
        rec->roll_samples = (int32_t)(uint32_t)a->sums[2];
    }
    return stop;
}

/* The statistics output_final_stddev() prints (autostop.hpp:233-244): the
 * rolling window when the run was autostopped, otherwise one more tabulation
 * of the final energies. */
void orc_autostop_final(orc_autostop* a, int generation, const float* energies, int64_t total_evals, float* bins,
                        adgpu_autostop_record* rec) {
    memset(rec, 0, sizeof(*rec));
    if (!a->autostopped) {
        as_tabulate(a, energies, bins);
        as_set_stats(a, bins);
        memcpy(a->sums, bins, 3 * sizeof(float));
    }
    rec->generation = generation;
    rec->total_evals = total_evals;
    rec->threshold = a->threshold;
    rec->average = a->curr_avg;
    rec->stddev = a->curr_std;
    rec->samples = (int32_t)a->best_n;
    rec->best_energy = a->overall_best;
    rec->autostopped = a->autostopped;
    rec->stop = 0;
    rec->roll_average = as_average(a->sums);
    rec->roll_stddev = as_stddev(a->sums);
    rec->roll_samples = (int32_t)(uint32_t)a->sums[2];
}
