/*
 * adgpu.h -- C ABI of the MI355X (gfx950) docking hot path.
 *
 * This is the drop-in boundary between AutoDock-GPU's host driver and the
 * HIP kernels of this library.  Every entry point takes plain pointers and
 * sizes, returns an adgpu_status code instead of calling exit(), and never
 * exposes HIP or torch types.  Each declaration names the reference
 * interface it replaces (file:line in ccsb-scripps/AutoDock-GPU v1.5).
 *
 * Units and layouts follow the reference device contract:
 *   - genotypes are [x, y, z (grid units), phi, theta, alpha (deg),
 *     torsion_1 .. torsion_T (deg)] padded to ADGPU_GENOTYPE_STRIDE floats
 *     (GENOTYPE_LENGTH_IN_GLOBMEM, common/defines.h:94);
 *   - populations are [run][individual][64], energies [run][individual]
 *     (host/src/performdocking.cpp.Cuda:293-299);
 *   - receptor maps use the packed-4 layout of host/src/processgrid.cpp:234-249:
 *     each grid point of each map stores v(x,y,z), v(x,y+1,z), v(x,y,z+1),
 *     v(x,y+1,z+1); maps are ordered [type maps..., electrostatic, desolvation].
 */
#ifndef ADGPU_H_
#define ADGPU_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limits (common/defines.h:83-95). */
#define ADGPU_MAX_ATOMS          256
#define ADGPU_MAX_ROTBONDS       57
#define ADGPU_MAX_ATYPES         32
#define ADGPU_MAX_PAIRS          (ADGPU_MAX_ATOMS * ADGPU_MAX_ATOMS)
#define ADGPU_MAX_POPSIZE        2048
#define ADGPU_MAX_RUNS           8192
#define ADGPU_MAX_GRIDPOINTS     256
#define ADGPU_GENOTYPE_STRIDE    64
/* Number of PRNG streams per individual.  The reference keeps one LCG state
 * per work-item (device/auxiliary_genetic.cl:29-51), i.e. NUM_OF_THREADS_PER_BLOCK
 * states per block; this library always uses the 64-work-item layout. */
#define ADGPU_NWI                64

typedef enum adgpu_status {
    ADGPU_OK            = 0,
    ADGPU_ERR_ARG       = 1,   /* null pointer, bad id, inconsistent sizes     */
    ADGPU_ERR_LIMIT     = 2,   /* a size exceeds one of the limits above        */
    ADGPU_ERR_HIP       = 3,   /* a HIP runtime call failed                     */
    ADGPU_ERR_NOMEM     = 4,   /* device memory could not be allocated          */
    ADGPU_ERR_STATE     = 5,   /* call not valid in the current object state    */
    ADGPU_ERR_TIMEOUT   = 6    /* a wait ended before the request finished      */
} adgpu_status;

typedef enum adgpu_map_format {
    ADGPU_MAP_F32  = 0,
    ADGPU_MAP_F16  = 1,
    ADGPU_MAP_BF16 = 2
} adgpu_map_format;

typedef enum adgpu_ls_method {
    ADGPU_LS_NONE  = 0,        /* plain GA (lsearch_rate == 0)                  */
    ADGPU_LS_SW    = 1,        /* Solis-Wets, device/kernel3.cl                 */
    ADGPU_LS_AD    = 2,        /* ADADELTA, device/kernel_ad.cl                 */
    ADGPU_LS_ADAM  = 3,        /* ADAM, cuda/kernel_adam.cu                     */
    ADGPU_LS_FIRE  = 4         /* FIRE, device/kernel_fire.cl                   */
} adgpu_ls_method;

typedef enum adgpu_eval_mode {
    ADGPU_EVAL_ENERGY   = 0,   /* gpu_calc_energy   (device/calcenergy.cl:132)        */
    ADGPU_EVAL_ENERGRAD = 1,   /* gpu_calc_energrad (device/calcMergedEneGra.cl:46)   */
    ADGPU_EVAL_GRADIENT = 2    /* gpu_calc_gradient (device/calcgradient.cl:109)      */
} adgpu_eval_mode;

/* Receptor grid.  Mirrors the fields of Gridinfo the kernels read
 * (host/inc/processgrid.h:43-57; dockpars gridsize/spacing at
 * host/src/performdocking.cpp.Cuda:420-426). */
typedef struct adgpu_grid {
    int32_t      size_x, size_y, size_z;   /* grid points per axis (<= 256)      */
    int32_t      num_map_atypes;           /* type maps; e and d follow them     */
    float        spacing;                  /* Angstrom per grid unit             */
    const float* packed;                   /* 4*size_x*size_y*size_z*(num_map_atypes+2) */
} adgpu_grid;

/* Ligand (+ flexible residue) device constants.  Mirrors the content of the
 * kernelconstant_* blocks built by prepare_const_fields_for_gpu
 * (host/src/calcenergy.cpp:31-345, host/inc/calcenergy.h:87-136) without the
 * MAX_* padding.  Coordinates are already centred and in grid units, as
 * gen_initpop_and_reflig leaves them (host/src/getparameters.cpp:2624-2630). */
typedef struct adgpu_ligand {
    int32_t         num_atoms;          /* ligand + flexres atoms (<= 256)        */
    int32_t         true_ligand_atoms;  /* atoms 0..true_ligand_atoms-1 are ligand */
    int32_t         num_rotbonds;       /* genes = num_rotbonds + 6 (<= 57)       */
    int32_t         num_atypes;         /* ligand atom types (<= 32)              */
    int32_t         num_pairs;          /* intramolecular contributors            */
    const float*    charges;            /* [num_atoms]                            */
    const int32_t*  atom_types;         /* [num_atoms] ligand type index          */
    const int32_t*  atom_map;           /* [num_atoms] index of the type map      */
    const uint8_t*  ignore_inter;       /* [num_atoms] 1 = no receptor energy     */
    const float*    ref_coords;         /* [3*num_atoms]                          */
    const float*    rotbond_moving;     /* [3*num_rotbonds] moving vectors        */
    const float*    rotbond_unit;       /* [3*num_rotbonds] unit vectors          */
    const int32_t*  rotbond_atoms;      /* [2*num_rotbonds] the bond's two atoms  */
    const uint8_t*  atom_rotbonds;      /* [num_atoms*num_rotbonds] 1 = atom moves */
    const int32_t*  pairs;              /* [2*num_pairs] (i<j)                    */
    const uint32_t* atype_reqm;         /* [num_atypes] base type index           */
    const uint16_t* vw_exp;             /* [num_atypes^2] (m<<8)|n                */
    const float*    reqm_ab;            /* [num_atypes^2]                         */
    const float*    vw_ac;              /* [num_atypes^2] A (vdW) or C (H-bond)   */
    const float*    vw_bd;              /* [num_atypes^2] B (vdW) or D (H-bond)   */
    const float*    ds_s;               /* [num_atypes] solvation parameter       */
    const float*    ds_v;               /* [num_atypes] atomic volume             */
} adgpu_ligand;

/* Scoring-function scalars (GpuDockparameters, cuda/GpuData.h:74-110;
 * set at host/src/performdocking.cpp.Cuda:427-441). */
typedef struct adgpu_energy_params {
    float coeff_elec;          /* scaled_AD4_coeff_elec                      */
    float coeff_desolv;        /* AD4_coeff_desolv                           */
    float qasp;                /* 0.01097 by default                         */
    float smooth;              /* 0.5 A by default                           */
    float elec_min_distance;   /* 0.01 A by default                          */
} adgpu_energy_params;

/* Search parameters (Dockpars defaults at host/inc/getparameters.h:66-145;
 * device scaling at host/src/performdocking.cpp.Cuda:427-463).  Rates are
 * given in percent exactly as Dockpars holds them. */
typedef struct adgpu_search_params {
    int32_t  num_runs;
    int32_t  pop_size;
    int64_t  max_evals;            /* num_of_energy_evals, per run on average     */
    int64_t  max_generations;      /* num_of_generations                           */
    float    tournament_rate;      /* percent                                      */
    float    crossover_rate;       /* percent                                      */
    float    mutation_rate;        /* percent                                      */
    float    abs_max_dmov;         /* grid units                                   */
    float    abs_max_dang;         /* degrees                                      */
    float    lsearch_rate;         /* percent                                      */
    int32_t  ls_method;            /* adgpu_ls_method                              */
    int32_t  initial_sw_generations;
    float    rho_lower_bound;
    float    base_dmov_mul_sqrt3;  /* grid units                                   */
    float    base_dang_mul_sqrt3;  /* degrees                                      */
    int32_t  cons_limit;
    int32_t  max_num_of_iters;
    float    adam_beta1, adam_beta2, adam_epsilon;
} adgpu_search_params;

typedef struct adgpu_context* adgpu_ctx;
typedef struct adgpu_job*     adgpu_job_t;

/* Number of usable GPUs (replaces the device enumeration of get_gpu_pool,
 * host/src/performdocking.cpp.Cuda:96-115). */
int  adgpu_device_count(int* out_count);

/* --- context ---------------------------------------------------------------
 * Replaces setup_gpu_for_docking / finish_gpu_from_docking
 * (host/inc/performdocking.h.Cuda:65-75, host/src/performdocking.cpp.Cuda:117-222). */
int  adgpu_open(int device, adgpu_ctx* out);
int  adgpu_close(adgpu_ctx ctx);
/* Message of the calling thread's last failed call (the reference prints it
 * and exits, cuda/GpuData.h:40-72). */
const char* adgpu_last_error(adgpu_ctx ctx);
/* Version string; also usable without a GPU. */
const char* adgpu_version(void);
/* Lanes that evaluate one individual: 0 = chosen per ligand (the team table
 * below if set; else for a job, with the local searches in flight per
 * generation counted as the job's own times a server's slots: up to 32 atoms
 * 64 for Solis-Wets below 7000 and ADADELTA below 10000 searches in flight,
 * else 32; 33-64 atoms 64 below 20000, else 32; 64 above 64 atoms; for
 * per-pose calls 32 up to 128 atoms), 16 (four individuals per
 * wavefront, ligands of at most 32 atoms; wider teams otherwise), 32 (two
 * individuals per wavefront) or 64.  Plays the role of the reference's
 * build-time NUM_OF_THREADS_PER_BLOCK (common/defines.h:29-49) as a run-time
 * choice; results agree within the FP32 tolerance either way. */
int  adgpu_set_team_width(adgpu_ctx ctx, int width);
/* Team width per ligand size, used when no width is set: a ligand of N atoms
 * takes widths[i] of the first entry with N <= max_atoms[i] (ascending; 16,
 * 32 or 64, 16 only up to 32 atoms); larger ligands keep the built-in rule.
 * n = 0 clears the table.  Like the width itself it stands for the
 * reference's build-time NUM_OF_THREADS_PER_BLOCK (common/defines.h:29-49);
 * adgpu.autotune fills it from measured local-search launches.  Set it before
 * creating the jobs that should use it. */
int  adgpu_set_team_table(adgpu_ctx ctx, int n, const int32_t* max_atoms, const int32_t* widths);
/* ADADELTA hand-off: a search that has run `iterations` iterations while
 * unstarted searches of the same launch remain is saved and resumed later by
 * whichever team is free, so long searches drawn late do not end the launch
 * alone.  Every search computes exactly what it computes without hand-off
 * (bit-identical results); 0 turns it off.  Default 32 (environment variable
 * ADGPU_LS_YIELD at adgpu_open); applies to jobs created afterwards.  The
 * reference runs each ADADELTA search to its end in a block of its own
 * (cuda/kernel_ad.cu:60-120: one block per run and LS entity); this has no
 * counterpart there and changes only when each search runs. */
int  adgpu_set_ls_handoff(adgpu_ctx ctx, int iterations);
/* ADADELTA regrouping, with the hand-off on and teams narrower than a
 * wavefront: once a launch's searches have all started, a team left without a
 * search beside a team that still runs one takes a seat, and a search that is
 * the last work of its wavefront moves (once) to a free seat, trying every
 * `iterations` iterations, so wavefronts that would run one team's search
 * beside idle lanes merge and the freed lanes run other launches' work.
 * Bit-identical results, like the hand-off; 0 turns it off; -1 (the default,
 * also environment variable ADGPU_LS_REGROUP at adgpu_open) regroups every 8
 * iterations in jobs of four or more run groups and in servers of four or
 * more slots, whose launches overlap, every 32 with two or three run groups,
 * and not with one.  Applies to jobs created afterwards.  No counterpart in the
 * reference (one search per block, cuda/kernel_ad.cu:60-120). */
int  adgpu_set_ls_regroup(adgpu_ctx ctx, int iterations);

/* --- resident inputs ---------------------------------------------------------
 * Replace the per-job uploads in docking_with_gpu
 * (host/src/performdocking.cpp.Cuda:363-410).  Maps and ligands stay resident
 * until released, so several jobs can share one map set. */
int  adgpu_maps_upload(adgpu_ctx ctx, const adgpu_grid* grid, int format, int* out_map_id);
int  adgpu_maps_release(adgpu_ctx ctx, int map_id);
int  adgpu_ligand_upload(adgpu_ctx ctx, const adgpu_ligand* lig,
                         const adgpu_energy_params* ep, int* out_ligand_id);
int  adgpu_ligand_release(adgpu_ctx ctx, int ligand_id);

/* --- pose evaluation (the per-pose parity entry point) ------------------------
 * Evaluates n host genotypes (stride ADGPU_GENOTYPE_STRIDE).  mode selects the
 * reference device function: energy only (device/calcenergy.cl:132-554),
 * merged energy + gradient (device/calcMergedEneGra.cl:46-936) or the float
 * gradient used for FIRE's first step (device/calcgradient.cl:109-1003).
 * energies[n] is written for ENERGY and ENERGRAD; grads[n*64] for ENERGRAD
 * and GRADIENT (either may be NULL when not wanted). */
int  adgpu_eval_poses(adgpu_ctx ctx, int map_id, int ligand_id, int mode,
                      const float* genotypes, int n, float* energies, float* grads);

/* As adgpu_eval_poses, but allocates nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * genotypes, energies and gradients pass through a fixed per-device staging
 * area of ADGPU_NOALLOC_MAX_POSES poses (4.2 MB, held by each device the
 * library runs on until the process ends); n above that returns
 * ADGPU_ERR_LIMIT.  Calls on one device share that area: calls that overlap
 * on one device (from different contexts on it) race.  Calls on one context
 * are serialized by the context. */
#define ADGPU_NOALLOC_MAX_POSES  8192
int  adgpu_eval_poses_noalloc(adgpu_ctx ctx, int map_id, int ligand_id, int mode,
                              const float* genotypes, int n, float* energies, float* grads);

/* FP64 energies of n host genotypes on the GPU: total, inter and intra per
 * pose (any output may be NULL).  Same terms as gpu_calc_energy
 * (device/calcenergy.cl:132-554) evaluated in double; the final re-scoring
 * step that the reference does on the CPU in FP32 (process_result,
 * host/src/processresult.cpp:430-466) for each run's best pose. */
int  adgpu_rescore_f64(adgpu_ctx ctx, int map_id, int ligand_id, const float* genotypes, int n,
                       double* total, double* inter, double* intra);

/* --- docking jobs --------------------------------------------------------------
 * A job owns the device state docking_with_gpu allocates per ligand
 * (host/src/performdocking.cpp.Cuda:325-342): two population/energy buffers,
 * per-entity eval counters, per-run eval totals and the LCG states.
 * init_population: num_runs*pop_size*64 floats (gen_initpop_and_reflig,
 * host/src/getparameters.cpp:2543-2622).  prng_seeds: num_runs*pop_size*64
 * uint32 (host/src/performdocking.cpp.Cuda:310-317). */
int  adgpu_job_create(adgpu_ctx ctx, int map_id, int ligand_id,
                      const adgpu_search_params* sp,
                      const float* init_population, const uint32_t* prng_seeds,
                      adgpu_job_t* out);
int  adgpu_job_destroy(adgpu_job_t job);

/* K1 + K2: energies of the current population and the first eval counts
 * (gpu_calc_initpop, device/kernel1.cl:28-145; gpu_sum_evals, device/kernel2.cl:26-62). */
int  adgpu_job_initpop(adgpu_job_t job);
/* K4 alone: elitism, selection, crossover, mutation and offspring energies from
 * the current into the next buffers (gpu_gen_and_eval_newpops, device/kernel4.cl:32-336). */
int  adgpu_job_kernel_gen(adgpu_job_t job);
/* The local-search launch alone, on the next buffers, with the job's method
 * (perform_LS kernel3.cl:30-436, gradient_minAD kernel_ad.cl:60-549,
 * gpu_gradient_minAdam kernel_adam.cu:53-416, gradient_minFire kernel_fire.cl:75-895). */
int  adgpu_job_kernel_ls(adgpu_job_t job, int generation);
/* K2 alone, then the pointer flip (host/src/performdocking.cpp.Cuda:792-824). */
int  adgpu_job_kernel_sum_and_swap(adgpu_job_t job);
/* One whole generation: K4, LS, K2, flip (host/src/performdocking.cpp.Cuda:685-831). */
int  adgpu_job_generation(adgpu_job_t job);
/* Runs generations until check_progress reaches 100 % (host/src/performdocking.cpp.Cuda:
 * 685, 908-942).  Returns the number of generations in *out_generations. */
int  adgpu_job_run(adgpu_job_t job, int64_t* out_generations, int64_t* out_total_evals);
/* As adgpu_job_run, but runs at most max_new_generations generations in this
 * call, so a caller can stop between generations (the AutoStop test every
 * as_frequency generations, performdocking.cpp.Cuda:687-696).  The loop has
 * finished when fewer generations than asked for were run. */
int  adgpu_job_run_generations(adgpu_job_t job, int64_t max_new_generations,
                               int64_t* out_generations, int64_t* out_total_evals);
/* Host copies of the current population [runs*pop*64], energies [runs*pop],
 * per-run eval totals [runs] and LCG states [runs*pop*64]; any may be NULL.
 * Mirrors the final downloads at host/src/performdocking.cpp.Cuda:856-859. */
int  adgpu_job_download(adgpu_job_t job, float* population, float* energies,
                        int32_t* evals_of_runs, uint32_t* prng_states);
/* Overwrites the current population, energies and LCG states (any may be NULL). */
int  adgpu_job_upload(adgpu_job_t job, const float* population, const float* energies,
                      const uint32_t* prng_states);
/* Best individual of each run of the current population (lowest FP32
 * energy, first index on ties) and its FP64 re-scored energies; each array
 * has num_runs entries (genotypes: num_runs*64 floats); any may be NULL. */
int  adgpu_job_best_f64(adgpu_job_t job, float* best_genotypes, float* best_energies, int32_t* best_index,
                        double* total, double* inter, double* intra);
/* Average device time in milliseconds of the last launches of one kernel
 * class, measured with HIP events on the job's stream (0 = gen, 1 = ls). */
int  adgpu_job_kernel_time(adgpu_job_t job, int kernel_class, double* avg_ms, int64_t* launches);
/* Device time during which at least one launch of the class ran (the union
 * of the launch intervals of all run groups since the job's first run call),
 * and the number of launches.  Launches inside graph blocks
 * (adgpu_set_graph_block) are not timed and not counted here.  No counterpart in the reference: measurement
 * for callers that time the kernels of overlapping run groups. */
int  adgpu_job_kernel_busy(adgpu_job_t job, int kernel_class, double* busy_ms, int64_t* launches);

/* Searches the ADADELTA hand-off passed from one team to another since the
 * job started (adgpu_set_ls_handoff).  No counterpart in the reference, where
 * every search runs to its end in its own block (cuda/kernel_ad.cu:60-120):
 * measurement, for callers and tests that check the hand-off took place. */
int  adgpu_job_ls_handoffs(adgpu_job_t job, int64_t* count);
/* Local-search tallies since the job started: stats[0] hand-offs (as
 * adgpu_job_ls_handoffs), stats[1] regroups (searches moved to a seat,
 * adgpu_set_ls_regroup), stats[2] wavefront evaluation steps of the local-search
 * launches (each counts one evaluation for every team of the wavefront, idle
 * teams included, modulo 2^32 per run group), stats[3] the job's team width
 * TW (lanes per team).  With the job's local-search evaluations E,
 * 1 - E / (stats[2] * 64 / TW) is the share of team evaluation steps that
 * idle teams spent.  No counterpart
 * in the reference: measurement. */
int  adgpu_job_ls_stats(adgpu_job_t job, int64_t stats[4]);

/* Wavefront VALU instructions per second the device issues on independent
 * full-rate FP32 FMA chains (v_fma_f32, 32 wavefronts per compute unit): the
 * peak that a kernel's measured VALU instruction rate is reported against.
 * No counterpart in the reference: measurement. */
int  adgpu_measure_valu_rate(adgpu_ctx ctx, double* wave_instr_per_s);

/* Generations per captured graph (0, 2, 4 or 8; default 0, or
 * ADGPU_GRAPH_BLOCK) for jobs started afterwards.  The reference launches
 * every kernel of every generation from its host loop
 * (host/src/performdocking.cpp.Cuda:685-831); with a block size B each run
 * group captures B consecutive generations (K4, local search, K2 and the
 * eval-total copy, performdocking.cpp.Cuda:700-831) into a hipGraph once and
 * replays it whenever the stop rule has proven all B generations run, so the
 * host issues one launch per B generations.  Results are bit-identical to
 * launching the generations one by one. */
int  adgpu_set_graph_block(adgpu_ctx ctx, int generations);

/* ---- AutoStop on the device -------------------------------------------------
 * The reference tests its population every as_frequency generations on the
 * host (AutoStop::check_if_satisfactory, host/inc/autostop.hpp:181-231, called
 * at host/src/performdocking.cpp.Cuda:687-696) and leaves the generation loop
 * when the test passes and total_evals > min_as_evals.  Here the test runs as
 * a kernel between generations, so adgpu_job_run keeps the device busy
 * across tests: generations enqueued behind a passed test do nothing. */

/* One test, as the reference prints it (autostop.hpp:201) and as
 * output_final_stddev reports the end (autostop.hpp:233-244). */
typedef struct adgpu_autostop_record {
    int64_t generation;     /* generation_cnt of the test                        */
    int64_t total_evals;    /* sum of evals_of_runs at the test                  */
    float   threshold;      /* threshold the test used (threshold_used)          */
    float   average;        /* curr_avg: average energy of the best samples      */
    float   stddev;         /* curr_std                                          */
    int32_t samples;        /* bestN                                             */
    float   best_energy;    /* overall_best_energy                               */
    int32_t autostopped;    /* check_if_satisfactory's result (sticky)           */
    int32_t stop;           /* autostopped && total_evals > min_as_evals         */
    float   roll_average;   /* average, stddev and samples of the last four tests */
    float   roll_stddev;
    int32_t roll_samples;
} adgpu_autostop_record;

/* Enables (or disables, enable = 0) AutoStop for the job's next
 * adgpu_job_run / adgpu_job_run_generations calls and resets its state (the
 * AutoStop constructor, autostop.hpp:129-146).  stopstd in kcal/mol,
 * as_frequency in generations, min_as_evals summed over all runs
 * (performdocking.cpp.Cuda:465-511). */
int  adgpu_job_set_autostop(adgpu_job_t job, int enable, float stopstd, int as_frequency, int64_t min_as_evals);
/* One test now, on the current population, at the job's generation count;
 * once a test has stopped the job, the job runs no more generations. */
int  adgpu_job_autostop_check(adgpu_job_t job, adgpu_autostop_record* out);
/* Records of the tests run so far: copies records [first, first + max) and
 * returns the number of tests in *out_count. */
int  adgpu_job_autostop_log(adgpu_job_t job, int first, int max, adgpu_autostop_record* out, int* out_count);
/* The end-of-docking statistics (output_final_stddev, autostop.hpp:233-244). */
int  adgpu_job_autostop_final(adgpu_job_t job, adgpu_autostop_record* out);

/* ---- serving loop -----------------------------------------------------------
 * Docking requests (a ligand against one of the context's resident map sets)
 * arrive one by one and run concurrently.  The reference docks the ligands
 * of a file list one after another per device (main.cpp:135-187,312-327,
 * 401-434; the OVERLAP pipeline only overlaps a ligand's setup with the
 * previous docking).  Here a fixed number of slots run requests at the same
 * time, each on its own streams, so that the generations of different
 * ligands fill the device together.  A slot keeps its buffers for the next
 * request (no allocation in steady state), and the memory budget bounds what
 * all slots hold, idle ones included: a request is admitted, in arrival
 * order, into the idle slot that needs the least growth to run it, when every
 * slot's buffers after that growth fit the budget; idle slots free their
 * buffers when only that makes room.  A request's result does not depend on
 * what else runs: every search is seeded by its own request. */
typedef struct adgpu_server* adgpu_server_t;

typedef struct adgpu_server_config {
    int32_t slots;            /* requests on the device at once (1..256)             */
    int32_t run_groups;       /* streams per request (0 = 1; at most one per run)     */
    int64_t memory_budget;    /* device bytes the slots' buffers may hold (0 = 90 % of
                                 the free device memory when the server starts)       */
} adgpu_server_config;

#define ADGPU_REQ_KEEP_POPULATION  1   /* keep the final populations and energies   */
#define ADGPU_REQ_RESCORE_F64      2   /* FP64 re-scoring of each run's best pose   */

/* One docking: everything adgpu_ligand_upload + adgpu_job_create take, plus
 * AutoStop (adgpu_job_set_autostop) and the result options.  The server
 * copies every array before adgpu_server_submit returns. */
typedef struct adgpu_request {
    int32_t                     map_id;
    const adgpu_ligand*         ligand;
    const adgpu_energy_params*  energy;
    const adgpu_search_params*  search;
    const float*                init_population;   /* num_runs*pop_size*64            */
    const uint32_t*             prng_seeds;        /* num_runs*pop_size*64            */
    int32_t                     autostop;          /* 1 = AutoStop on                 */
    int32_t                     as_frequency;
    float                       stopstd;
    int64_t                     min_as_evals;
    int32_t                     flags;             /* ADGPU_REQ_* bits                */
} adgpu_request;

typedef struct adgpu_request_result {
    int32_t status;           /* adgpu_status of the docking                         */
    int32_t slot;             /* slot that ran it                                    */
    int32_t num_runs, pop_size;
    int64_t generations;      /* generation_cnt                                      */
    int64_t total_evals;      /* sum of evals_of_runs                                */
    int64_t device_bytes;     /* device bytes the request needs in an empty slot      */
    double  t_submit_ms;      /* server clock: milliseconds since adgpu_server_create */
    double  t_admit_ms;
    double  t_done_ms;
    char    message[160];     /* error text when status != 0                         */
} adgpu_request_result;

typedef struct adgpu_server_stats {
    int64_t submitted, completed, failed;
    int64_t queued, active, peak_active;
    int64_t bytes_in_use, peak_bytes, memory_budget;   /* bytes the slots' buffers hold */
    int64_t slot_reallocations;   /* slot buffer growths after their first allocation */
    double  now_ms;
} adgpu_server_stats;

int  adgpu_server_create(adgpu_ctx ctx, const adgpu_server_config* cfg, adgpu_server_t* out);
/* Queues a request and returns its ticket.  Fails at once (nothing queued)
 * for an invalid ligand or search, an unknown map set, or a request larger
 * than the whole memory budget (ADGPU_ERR_LIMIT). */
int  adgpu_server_submit(adgpu_server_t srv, const adgpu_request* req, int64_t* ticket);
/* Oldest finished request not yet returned here; waits up to timeout_ms
 * (< 0: no limit) and returns ADGPU_ERR_TIMEOUT when none finished. */
int  adgpu_server_next_done(adgpu_server_t srv, int timeout_ms, int64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for the request, copies its result
 * and forgets the ticket (one thread at a time per ticket; a second one gets
 * ADGPU_ERR_STATE).  best_genotypes: num_runs*64 floats, best_energies:
 * num_runs, best_f64: 3*num_runs doubles (total, inter, intra; needs
 * ADGPU_REQ_RESCORE_F64), population: num_runs*pop_size*64 and energies:
 * num_runs*pop_size (need ADGPU_REQ_KEEP_POPULATION).  Any pointer may be
 * null. */
int  adgpu_server_result(adgpu_server_t srv, int64_t ticket, int timeout_ms, adgpu_request_result* out,
                         float* best_genotypes, float* best_energies, double* best_f64,
                         float* population, float* energies);
int  adgpu_server_stats_get(adgpu_server_t srv, adgpu_server_stats* out);
/* Finishes the running requests, drops the queued ones and frees the slots. */
int  adgpu_server_destroy(adgpu_server_t srv);

#ifdef __cplusplus
}
#endif
#endif /* ADGPU_H_ */
