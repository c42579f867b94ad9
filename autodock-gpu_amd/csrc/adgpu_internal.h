// adgpu_internal.h -- device data layout shared by the HIP kernels and the
// host side of the library.  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/adgpu.h"

namespace adgpu {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kGroupThreads = 256;   // threads per workgroup (its teams share the staged ligand tables)
constexpr int kGenes = ADGPU_GENOTYPE_STRIDE;

// One intramolecular pair, precomputed on the host from the reference's
// per-type tables (the lookups of calcenergy.cl:397-536 folded per pair), in
// two 16-byte halves stored as separate arrays so that consecutive lanes read
// consecutive 16-byte records (no LDS bank conflicts).
struct __attribute__((aligned(16))) PairA {
    uint32_t ij;      // atom1 | atom2 << 16
    uint32_t flags;   // m | n << 8 | (CG-G0 bond) << 16 | (fast 12-6 / 12-10 form) << 17
    float ac;         // VWpars_AC of the type pair
    float bd;         // VWpars_BD
};
struct __attribute__((aligned(16))) PairB {
    float reqm;       // reqm_AB (optimum distance)
    float cq;         // coeff_elec * q1 * q2 (same operation order as the reference)
    float dsol;       // (S1 + qasp|q1|) V2 + (S2 + qasp|q2|) V1
    float g;          // CG-G0 bond constant G: 50 for a CG-G0 pair, else 0
};

// Everything the kernels need about one ligand, in one device allocation.
// Offsets are in bytes from the start of the block.
struct LigandDev {
    int32_t num_atoms, true_ligand_atoms, num_rotbonds, num_genes;
    int32_t num_pairs, nchunk, any_general, pad1;   // any_general: some pair is not 12-6 / 12-10
    float coeff_elec, coeff_desolv, qasp, half_smooth;
    float elec_min_distance, inv_elec_min, pad3, pad4;
    uint32_t off_atom_ref;    // float4 [N]  x, y, z (grid units), charge
    uint32_t off_atom_info;   // int4   [N]  map index, type, ignore_inter, 0
    uint32_t off_atom_mask;   // uint2  [N]  bit t = moves with rotatable bond t
    uint32_t off_tor_unit;    // float4 [T]  unit vector
    uint32_t off_tor_move;    // float4 [T]  moving vector
    uint32_t off_tor_atoms;   // int2   [T]  bond atoms
    uint32_t off_pair_a;      // PairA[P]
    uint32_t off_pair_b;      // PairB[P]
    uint32_t off_type_ds;     // float2 [num_atypes] (dspars_S, dspars_V) per ligand type (FP64 re-scoring)
    uint32_t total_bytes, pad6, pad7;
};

struct GridDev {
    int32_t gx, gy, gz, nmaps;        // nmaps = type maps + 2
    uint32_t g1, g2, g3, format;      // g1 = gx, g2 = gx*gy, g3 = gx*gy*gz
    float spacing, inv_spacing, pad0, pad1;
    const void* data;                 // [nmaps][g3] points of 4 values (float4 or half4/bf16x4)
};

// Search parameters scaled as the device sees them
// (host/src/performdocking.cpp.Cuda:427-463).
struct SearchDev {
    int32_t num_runs, pop_size, num_genes, num_ls;
    float tournament_rate, crossover_rate, mutation_rate;   // in [0,1]
    float abs_max_dmov, abs_max_dang, lsearch_rate;        // lsearch_rate in percent
    float rho_lower_bound, base_dmov_mul_sqrt3, base_dang_mul_sqrt3;
    int32_t cons_limit, max_iters, ls_method;
    float adam_beta1, adam_beta2, adam_epsilon;
    int32_t ls_yield;   // ADADELTA: iterations after which a search may hand off (0: never)
    int32_t ls_regroup; // ADADELTA with hand-off: iterations a team runs a search before the search may
                        // move to a seat in another wavefront once the queue is drained (0: never)
};

// State of an ADADELTA search handed off between teams: genes, best genes,
// E[g^2], E[dx^2] (kGenes floats each), then best energy, varrho,
// consecutive successes, consecutive failures, iterations and the
// population index (integers stored as their bit patterns).
// Lanes per team for a ligand of n atoms and a width hint (16, 32, 64, or
// another value for the default): 16 only up to 32 atoms, 32 only up to 128.
inline int team_lanes_for(int n_atoms, int hint) {
    int tw = hint == 64 ? 64 : (hint == 32 ? 32 : (n_atoms <= 128 ? 32 : 64));
    if (hint == 16 && n_atoms <= 32) tw = 16;
    if (tw == 32 && n_atoms > 128) tw = 64;
    return tw;
}

constexpr int kAdStateFloats = 4 * kGenes + 8;

struct JobBuffers {
    float* pop_cur;
    float* pop_next;
    float* ene_cur;
    float* ene_next;
    int32_t* evals_new;
    int32_t* evals_runs;
    uint32_t* prng;
    int32_t* ls_counter;   // work queue of the persistent local-search kernels: [0] next unstarted
                           // search; ADADELTA hand-off: [2] next ring ticket, [3] fault flag
                           // (sticky over a run), [5] next seat ticket, [6] next seat claim;
                           // per job: [4] hand-offs, [7] regroups, [8] wavefront evaluation steps
    int32_t* ls_ent;       // FIRE: entity chosen per LS slot [runs*num_ls]
    float* ls_grad;        // FIRE: initial gradient per LS slot [runs*num_ls*64]
    float* ls_state;       // ADADELTA hand-off: saved search per LS slot [runs*num_ls*kAdStateFloats]
    int32_t* ls_ring;      // ADADELTA hand-off: per claim of the queue, the search handed off there
                           // [runs*num_ls], then the regrouping seats [2*runs*num_ls]
    const int32_t* stop;   // set by a passed AutoStop test: generation kernels return at once
    int32_t* evals_copy;   // when set, K2 also writes the new per-run totals here (a generation
                           // inside a captured graph block: its ring slot of the device copy)
};

// The AutoStop class's members (host/inc/autostop.hpp:36-58) in device
// memory; bins = average_sd2_N, 3*(pop_size+1) floats, follows the struct.
struct AutoStopDev {
    float threshold, threshold_used, thres_stddev, curr_avg, curr_std;
    float delta_energy, overall_best;
    float rolling[16];
    uint32_t roll_count, best_n;
    int32_t first_time, autostopped, num_tests;
};

struct AutoStopParams {
    int32_t num_runs, pop_size, ncream, as_frequency;
    float stopstd;
    int32_t log_cap;           // records kept (the latest log_cap tests)
    int64_t min_as_evals;
};

// What every launcher needs about the ligand and the maps.
struct LaunchArgs {
    hipStream_t stream;
    const LigandDev* lig_dev;    // device copy of the ligand block
    const LigandDev* lig_host;   // host copy of its header
    GridDev grid;
    int team_hint;               // 0 = automatic, 32 or 64 = team width
    int share;                   // concurrent launches sharing the device (run groups); persistent
                                 // local-search grids take 1/share of the resident capacity
};

// Launchers (k_*.hip).
// FP64 energies of n poses (k_rescore.hip): total, inter and intra per pose.
hipError_t launch_rescore_f64(const LaunchArgs& la, const float* genotypes, int n, double* out3);
hipError_t launch_eval_poses(const LaunchArgs& la, int mode, const float* genotypes, int n, float* energies,
                             float* grads);
hipError_t launch_initpop(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
hipError_t launch_gen(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
hipError_t launch_sum_evals(hipStream_t s, const SearchDev& sp, const JobBuffers& b);
// One local-search launch (resets the work queue first; FIRE also runs its
// initial-gradient pass).
hipError_t launch_ls(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b, int method);
hipError_t launch_ls_sw(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
hipError_t launch_ls_ad(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
hipError_t launch_ls_adam(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
hipError_t launch_ls_fire(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b);
// One AutoStop test (mode 0) or the final statistics (mode 1) on the
// energies ene[num_runs*pop_size]; k_autostop.hip.
hipError_t launch_autostop(hipStream_t s, const AutoStopParams& p, const float* ene, const int32_t* evals_runs,
                           AutoStopDev* st, float* bins, adgpu_autostop_record* log, int32_t* stop,
                           int64_t generation, int mode, adgpu_autostop_record* out);
// Address of the staging area of adgpu_eval_poses_noalloc on `device` (which
// must be current), looked up once per device (k_eval.hip).
hipError_t noalloc_scratch_setup(int device, float** out);
// adgpu_eval_poses through that staging area; n <= ADGPU_NOALLOC_MAX_POSES.
hipError_t eval_poses_noalloc(const LaunchArgs& la, float* scratch_dev, int mode, const float* genotypes, int n,
                              float* energies, float* grads);
// Wavefront VALU instructions per second of a kernel of independent FP32
// FMA chains over `blocks` workgroups of 256 threads (k_calib.hip); scratch
// holds blocks * 256 floats.
hipError_t measure_valu_rate(hipStream_t stream, float* scratch, int blocks, int iters, double* wave_instr_per_s);
// LDS bytes and team width the launchers use for a ligand.
size_t lds_bytes(const LigandDev& lig_host, int team_hint, int* team_width);

}  // namespace adgpu
