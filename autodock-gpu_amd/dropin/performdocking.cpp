// performdocking.cpp -- the MI355X backend of AutoDock-GPU's docking driver.
//
// Implements the boundary the reference's main.cpp calls
// (host/inc/performdocking.h.Cuda:65-96): get_gpu_pool, setup_gpu_for_docking,
// finish_gpu_from_docking, docking_with_gpu and check_progress.  The host
// preparation uses the reference's own host functions (gen_initpop_and_reflig,
// LocalRNG, prepare_const_fields_for_gpu, and AutoStop's table header); the
// generations and the AutoStop tests run on the GPU through the C ABI of
// include/adgpu.h (libadgpu_hip.so).
// Built with the reference's host sources by the Makefile beside this file.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "performdocking.h"
#include "autostop.hpp"

namespace {

double elapsed_seconds(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

// Runtime errors end the program, as RTERROR does (cuda/GpuData.h:40-72).
void check(int rc, const char* what, adgpu_ctx ctx) {
    if (rc == ADGPU_OK) return;
    printf("Error: %s failed (status %d): %s\n", what, rc, ctx ? adgpu_last_error(ctx) : "");
    exit(-1);
}

int ls_method_code(const char* name) {
    if (strcmp(name, "sw") == 0) return ADGPU_LS_SW;
    if (strcmp(name, "ad") == 0) return ADGPU_LS_AD;
    if (strcmp(name, "adam") == 0) return ADGPU_LS_ADAM;
    if (strcmp(name, "fire") == 0) return ADGPU_LS_FIRE;
    return -1;
}

const char* ls_method_label(int m) {
    switch (m) {
        case ADGPU_LS_SW: return "Solis-Wets (sw)";
        case ADGPU_LS_AD: return "ADADELTA (ad)";
        case ADGPU_LS_ADAM: return "ADAM (adam)";
        case ADGPU_LS_FIRE: return "FIRE (fire)";
        default: return "GA";
    }
}

// Map storage precision: ADGPU_MAP_FORMAT = f32 (default), f16 or bf16.
int map_format() {
    const char* f = getenv("ADGPU_MAP_FORMAT");
    if (!f) return ADGPU_MAP_F32;
    if (strcmp(f, "f16") == 0) return ADGPU_MAP_F16;
    if (strcmp(f, "bf16") == 0) return ADGPU_MAP_BF16;
    return ADGPU_MAP_F32;
}

// The ligand tables prepare_const_fields_for_gpu builds, as adgpu_ligand.
struct LigandTables {
    std::vector<int32_t> atom_types, atom_map, rotbond_atoms, pairs;
    std::vector<uint8_t> ignore_inter, atom_rotbonds;
    adgpu_ligand L{};
};

void to_adgpu_ligand(const Liganddata& lig, const kernelconstant_interintra& ki, const kernelconstant_intracontrib& kc,
                     const kernelconstant_intra& kin, const kernelconstant_conform& kcf, const kernelconstant_grads& kg,
                     LigandTables& t) {
    const int N = lig.num_of_atoms, T = lig.num_of_rotbonds, P = lig.num_of_intraE_contributors;
    t.atom_types.assign(ki.atom_types_const, ki.atom_types_const + N);
    t.atom_map.assign(ki.atom_types_map_const, ki.atom_types_map_const + N);
    t.ignore_inter.resize(N);
    for (int a = 0; a < N; ++a) t.ignore_inter[a] = ki.ignore_inter_const[a] ? 1 : 0;
    t.rotbond_atoms.assign(kg.rotbonds, kg.rotbonds + 2 * T);
    t.atom_rotbonds.assign((size_t)N * T, 0);
    for (int r = 0; r < T; ++r)
        for (int k = 0; k < kg.num_rotating_atoms_per_rotbond[r]; ++k)
            t.atom_rotbonds[(size_t)kg.rotbonds_atoms[MAX_NUM_OF_ATOMS * r + k] * T + r] = 1;
    t.pairs.assign(kc.intraE_contributors_const, kc.intraE_contributors_const + 2 * P);
    adgpu_ligand& L = t.L;
    L.num_atoms = N;
    L.true_ligand_atoms = lig.true_ligand_atoms;
    L.num_rotbonds = T;
    L.num_atypes = lig.num_of_atypes;
    L.num_pairs = P;
    L.charges = ki.atom_charges_const;
    L.atom_types = t.atom_types.data();
    L.atom_map = t.atom_map.data();
    L.ignore_inter = t.ignore_inter.data();
    L.ref_coords = kcf.ref_coords_const;
    L.rotbond_moving = kcf.rotbonds_moving_vectors_const;
    L.rotbond_unit = kcf.rotbonds_unit_vectors_const;
    L.rotbond_atoms = t.rotbond_atoms.data();
    L.atom_rotbonds = t.atom_rotbonds.data();
    L.pairs = t.pairs.data();
    L.atype_reqm = kin.atom_types_reqm_const;
    L.vw_exp = kin.VWpars_exp_const;
    L.reqm_ab = kin.reqm_AB_const;
    L.vw_ac = kin.VWpars_AC_const;
    L.vw_bd = kin.VWpars_BD_const;
    L.ds_s = kin.dspars_S_const;
    L.ds_v = kin.dspars_V_const;
}

}  // namespace

std::vector<int> get_gpu_pool() {
    // eight hardware queues per process (HIP's default is 4) before the HIP
    // runtime starts, so the library runs each job in four run groups
    setenv("GPU_MAX_HW_QUEUES", "8", 0);
    int n = 0;
    if (adgpu_device_count(&n) != ADGPU_OK || n <= 0) {
        printf("No HIP devices found, exiting.\n");
        exit(-1);
    }
    std::vector<int> pool(n);
    std::iota(pool.begin(), pool.end(), 0);
    return pool;
}

void setup_gpu_for_docking(GpuData& cData, GpuTempData& tData) {
    if (cData.devnum < -1) return;   // device already set up
    auto const t0 = std::chrono::steady_clock::now();
    int n = 0;
    if (adgpu_device_count(&n) != ADGPU_OK || n <= 0) {
        printf("No HIP devices found, exiting.\n");
        exit(-1);
    }
    if (cData.devnum >= n) {
        printf("Error: Requested device %i does not exist (only %i devices available).\n", cData.devnum + 1, n);
        exit(-1);
    }
    const int dev = cData.devnum < 0 ? 0 : cData.devnum;
    cData.ctx = nullptr;
    check(adgpu_open(dev, &cData.ctx), "adgpu_open", nullptr);
    const char* name = adgpu_version();
    tData.device_name = (char*)malloc(strlen(name) + 32);
    strcpy(tData.device_name, name);
    if (n > 1) snprintf(tData.device_name + strlen(name), 32, " (#%d / %d)", dev + 1, n);
    printf("HIP device:                               %s\n", tData.device_name);
    cData.devid = dev;
    cData.devnum = -2;
    cData.map_id = 0;
    cData.map_key[0] = '\0';
    tData.pMem_fgrids = nullptr;
    auto const t1 = std::chrono::steady_clock::now();
    printf("\nHIP Setup time %fs\n", elapsed_seconds(t0, t1));
}

void finish_gpu_from_docking(GpuData& cData, GpuTempData& tData) {
    if (cData.devnum > -2) return;   // device not set up
    if (cData.ctx) adgpu_close(cData.ctx);
    cData.ctx = nullptr;
    cData.map_id = 0;
    free(tData.device_name);
    tData.device_name = nullptr;
}

// check_progress, host/src/performdocking.cpp.Cuda:908-942: the larger of the
// average evals per run and the generation count, in percent of their limits.
double check_progress(int* evals_of_runs, int generation_cnt, int max_num_of_evals, int max_num_of_gens,
                      int num_of_runs, unsigned long& total_evals) {
    total_evals = 0;
    for (int i = 0; i < num_of_runs; i++) total_evals += evals_of_runs[i];
    const double evals_progress = (double)total_evals / ((double)num_of_runs) / max_num_of_evals * 100.0;
    const double gens_progress = ((double)generation_cnt) / ((double)max_num_of_gens) * 100.0;
    return evals_progress > gens_progress ? evals_progress : gens_progress;
}

// docking_with_gpu, host/src/performdocking.cpp.Cuda:224-905.
int docking_with_gpu(const Gridinfo* mygrid, Dockpars* mypars, const Liganddata* myligand_init,
                     const Liganddata* myxrayligand, Profile& profile, const int* argc, char** argv,
                     SimulationState& sim_state, GpuData& cData, GpuTempData& tData, std::string* output) {
    (void)myxrayligand; (void)argc; (void)argv; (void)tData;
    char* outbuf = nullptr;   // used by para_printf when output != NULL
    if (output != NULL) outbuf = (char*)malloc(256 * sizeof(char));
    auto const t1 = std::chrono::steady_clock::now();
    adgpu_ctx ctx = cData.ctx;

    const int nrun = (int)mypars->num_of_runs, pop = (int)mypars->pop_size;
    const size_t ne = (size_t)nrun * pop;
    const size_t size_populations = ne * GENOTYPE_LENGTH_IN_GLOBMEM * sizeof(float);
    sim_state.cpu_populations.assign(size_populations, 0.0f);
    sim_state.cpu_energies.assign(ne * sizeof(float), 0.0f);
    sim_state.cpu_evals_of_runs.assign((size_t)nrun * sizeof(int), 0);

    // initial population and reference ligand (moved to the origin, scaled),
    // and the per-thread LCG seeds of 64 work-items per individual
    Liganddata myligand_reference = *myligand_init;
    gen_initpop_and_reflig(mypars, sim_state.cpu_populations.data(), &myligand_reference, mygrid);
    std::vector<uint32_t> seeds(ne * ADGPU_NWI);
    LocalRNG r(mypars->seed);
    for (size_t i = 0; i < seeds.size(); i++) seeds[i] = r.random_uint();

    // ligand tables as the reference prepares them for its kernels
    std::unique_ptr<kernelconstant_interintra> ki(new kernelconstant_interintra);
    std::unique_ptr<kernelconstant_intracontrib> kc(new kernelconstant_intracontrib);
    std::unique_ptr<kernelconstant_intra> kin(new kernelconstant_intra);
    std::unique_ptr<kernelconstant_rotlist> krl(new kernelconstant_rotlist);
    std::unique_ptr<kernelconstant_conform> kcf(new kernelconstant_conform);
    std::unique_ptr<kernelconstant_grads> kg(new kernelconstant_grads);
    if (prepare_const_fields_for_gpu(&myligand_reference, mypars, ki.get(), kc.get(), kin.get(), krl.get(), kcf.get(),
                                     kg.get()) == 1) {
        if (outbuf) free(outbuf);
        return 1;
    }
    LigandTables lt;
    to_adgpu_ligand(myligand_reference, *ki, *kc, *kin, *kcf, *kg, lt);
    adgpu_energy_params ep{};
    ep.coeff_elec = (float)mypars->coeffs.scaled_AD4_coeff_elec;
    ep.coeff_desolv = (float)mypars->coeffs.AD4_coeff_desolv;
    ep.qasp = mypars->qasp;
    ep.smooth = mypars->smooth;
    ep.elec_min_distance = mypars->elec_min_distance;
    int lig_id = 0;
    check(adgpu_ligand_upload(ctx, &lt.L, &ep, &lig_id), "adgpu_ligand_upload", ctx);

    // receptor maps stay resident while consecutive jobs use the same grid
    char key[sizeof(cData.map_key)];
    snprintf(key, sizeof(key), "%s|%s|%d %d %d|%.6f|%d|%zu", mygrid->receptor_name.c_str(),
             mygrid->map_base_name.c_str(), mygrid->size_xyz[0], mygrid->size_xyz[1], mygrid->size_xyz[2],
             mygrid->spacing, mygrid->num_of_map_atypes, mygrid->grids.size());
    if (cData.map_id == 0 || strcmp(key, cData.map_key) != 0) {
        if (cData.map_id) adgpu_maps_release(ctx, cData.map_id);
        adgpu_grid g{};
        g.size_x = mygrid->size_xyz[0];
        g.size_y = mygrid->size_xyz[1];
        g.size_z = mygrid->size_xyz[2];
        g.num_map_atypes = mygrid->num_of_map_atypes;
        g.spacing = (float)mygrid->spacing;
        g.packed = mygrid->grids.data();
        check(adgpu_maps_upload(ctx, &g, map_format(), &cData.map_id), "adgpu_maps_upload", ctx);
        strcpy(cData.map_key, key);
    }

    // search parameters (the device-side scaling happens in the library)
    const int method = mypars->lsearch_rate != 0.0f ? ls_method_code(mypars->ls_method) : ADGPU_LS_NONE;
    if (method < 0) {
        para_printf("\nError: LS method %s is not supported.\n", mypars->ls_method);
        exit(-1);
    }

    // heuristics: evaluation budget from the torsion count,
    // performdocking.cpp.Cuda:465-511
    unsigned long min_as_evals = 0;
    if (mypars->use_heuristics) {
        unsigned long heur_evals;
        const unsigned long nev = mypars->num_of_energy_evals;
        const int T = myligand_init->num_of_rotbonds;
        if (strcmp(mypars->ls_method, "sw") == 0) {
            heur_evals = (unsigned long)ceil(1000 * pow(2.0, 1.3 * T + 4.0));
        } else if (strcmp(mypars->ls_method, "ad") == 0) {
            heur_evals = (unsigned long)ceil(64000 * pow(2.0, (0.5 - 0.2 * T / (20.0f + T)) * T));
        } else {
            para_printf("\nError: LS method \"%s\" is not supported by heuristics.\n       Please choose Solis-Wets (sw), "
                        "Adadelta (ad),\n       or switch off the heuristics.\n", mypars->ls_method);
            exit(-1);
        }
        if (heur_evals < 500000) heur_evals = 500000;
        heur_evals *= 50.0f / mypars->num_of_runs;
        mypars->num_of_energy_evals = (unsigned long)ceil(heur_evals * (float)mypars->heuristics_max /
                                                          (mypars->heuristics_max + heur_evals));
        para_printf("    Using heuristics: (capped) number of evaluations set to %lu\n", mypars->num_of_energy_evals);
        if (mypars->nev_provided && (mypars->num_of_energy_evals > nev)) {
            para_printf("    Overriding heuristics, setting number of evaluations to --nev = %lu instead.\n", nev);
            mypars->num_of_energy_evals = nev;
            profile.capped = true;
        }
        const float cap_fraction = (float)mypars->num_of_energy_evals / heur_evals;
        const float a = 27.0 / 26.0;   // 10 % at a cap fraction of 50 %
        const float min_frac = a / (1 + cap_fraction * cap_fraction * (a / (a - 1.0f) - 1.0f)) + 1.0f - a;
        min_as_evals = (unsigned long)ceil(mypars->num_of_energy_evals * min_frac) * mypars->num_of_runs;
        if (cap_fraction < 0.5f) {
            para_printf("    Warning: The set number of evals is %.2f%% of the uncapped heuristics estimate of %lu "
                        "evals.\n", cap_fraction * 100.0f, heur_evals);
            if (mypars->autostop)
                para_printf("             AutoStop will not stop before %.2f%% (%lu) of the set number of "
                            "evaluations.\n", min_frac * 100.0f, min_as_evals / mypars->num_of_runs);
        }
    }
    para_printf("    Local-search chosen method is: %s\n", ls_method_label(method));
    if ((mypars->initial_sw_generations > 0) && (strcmp(mypars->ls_method, "sw") != 0))
        para_printf("    Using Solis-Wets (sw) for the first %d generations.\n", mypars->initial_sw_generations);
    profile.adadelta = (strcmp(mypars->ls_method, "ad") == 0);
    profile.n_evals = mypars->num_of_energy_evals;
    profile.num_atoms = myligand_reference.num_of_atoms;
    profile.num_rotbonds = myligand_init->num_of_rotbonds;

    adgpu_search_params sp{};
    sp.num_runs = nrun;
    sp.pop_size = pop;
    sp.max_evals = (int64_t)mypars->num_of_energy_evals;
    sp.max_generations = (int64_t)mypars->num_of_generations;
    sp.tournament_rate = mypars->tournament_rate;
    sp.crossover_rate = mypars->crossover_rate;
    sp.mutation_rate = mypars->mutation_rate;
    sp.abs_max_dmov = mypars->abs_max_dmov;
    sp.abs_max_dang = mypars->abs_max_dang;
    sp.lsearch_rate = mypars->lsearch_rate;
    sp.ls_method = method;
    sp.initial_sw_generations = mypars->initial_sw_generations;
    sp.rho_lower_bound = mypars->rho_lower_bound;
    sp.base_dmov_mul_sqrt3 = mypars->base_dmov_mul_sqrt3;
    sp.base_dang_mul_sqrt3 = mypars->base_dang_mul_sqrt3;
    sp.cons_limit = (int32_t)mypars->cons_limit;
    sp.max_num_of_iters = (int32_t)mypars->max_num_of_iters;
    sp.adam_beta1 = mypars->adam_beta1;
    sp.adam_beta2 = mypars->adam_beta2;
    sp.adam_epsilon = mypars->adam_epsilon;

    clock_t clock_start_docking = clock();
    adgpu_job_t job = nullptr;
    check(adgpu_job_create(ctx, cData.map_id, lig_id, &sp, sim_state.cpu_populations.data(), seeds.data(), &job),
          "adgpu_job_create", ctx);
    check(adgpu_job_initpop(job), "adgpu_job_initpop", ctx);
    auto const t2 = std::chrono::steady_clock::now();
    para_printf("\nRest of Setup time %fs\n", elapsed_seconds(t1, t2));

    // AutoStop runs on the device (adgpu_job_set_autostop); ADGPU_HOST_AUTOSTOP=1
    // tests on the host with the reference's AutoStop class instead, as the
    // reference does (performdocking.cpp.Cuda:687-696).
    const bool host_autostop = mypars->autostop && getenv("ADGPU_HOST_AUTOSTOP") &&
                               strcmp(getenv("ADGPU_HOST_AUTOSTOP"), "1") == 0;
    AutoStop autostop(mypars->pop_size, mypars->num_of_runs, mypars->stopstd, mypars->as_frequency, output);
    if (mypars->autostop) {
        autostop.print_intro(mypars->num_of_generations, mypars->num_of_energy_evals);
        if (!host_autostop)
            check(adgpu_job_set_autostop(job, 1, mypars->stopstd, (int)mypars->as_frequency, (int64_t)min_as_evals),
                  "adgpu_job_set_autostop", ctx);
    } else {
        para_printf("\nExecuting docking runs:\n");
        para_printf("        20%%        40%%       60%%       80%%       100%%\n");
        para_printf("---------+---------+---------+---------+---------+\n");
    }

    int64_t gens = 0, total = 0;
    int curr_progress_cnt = 0;
    bool autostopped = false;
    std::vector<int32_t> evals(nrun);
    if (mypars->autostop && !host_autostop) {
        // the tests run between generations on the device; the table lines
        // (autostop.hpp:201) are printed after every chunk of generations
        const int64_t chunk = 20 * (int64_t)mypars->as_frequency;
        int printed = 0;
        for (;;) {
            const int64_t before = gens;
            check(adgpu_job_run_generations(job, chunk, &gens, &total), "adgpu_job_run_generations", ctx);
            int n = 0;
            check(adgpu_job_autostop_log(job, 0, 0, nullptr, &n), "adgpu_job_autostop_log", ctx);
            std::vector<adgpu_autostop_record> recs(n > printed ? n - printed : 0);
            if (!recs.empty())
                check(adgpu_job_autostop_log(job, printed, (int)recs.size(), recs.data(), &n), "adgpu_job_autostop_log",
                      ctx);
            for (const auto& r : recs) {
                para_printf("%11u | %12lu |%8.2f kcal/mol |%8.2f +/-%8.2f kcal/mol |%8i |%8.2f kcal/mol\n",
                            (unsigned)r.generation, (unsigned long)r.total_evals / mypars->num_of_runs, r.threshold,
                            r.average, r.stddev, r.samples, r.best_energy);
                autostopped = autostopped || r.autostopped;
            }
            printed += (int)recs.size();
            fflush(stdout);
            if (gens - before < chunk) break;
        }
    } else {
        // generation loop: without AutoStop one call runs it to the end; with
        // the host AutoStop the energies are tested every as_frequency
        // generations before the generation runs (performdocking.cpp.Cuda:687-696)
        for (;;) {
            check(adgpu_job_download(job, nullptr, nullptr, evals.data(), nullptr), "adgpu_job_download", ctx);
            unsigned long total_evals = 0;
            const double progress = check_progress(evals.data(), (int)gens, (int)mypars->num_of_energy_evals,
                                                   (int)mypars->num_of_generations, nrun, total_evals);
            total = (int64_t)total_evals;
            if (progress >= 100.0) break;
            int64_t step = INT64_MAX;
            if (mypars->autostop) {
                const int asf = (int)mypars->as_frequency;
                if (gens % asf == 0) {
                    check(adgpu_job_download(job, nullptr, sim_state.cpu_energies.data(), nullptr, nullptr),
                          "adgpu_job_download", ctx);
                    if (autostop.check_if_satisfactory((int)gens, sim_state.cpu_energies.data(), total_evals))
                        if (total_evals > min_as_evals) break;
                }
                step = asf - gens % asf;
            } else {
                int new_progress_cnt = (int)(progress / 2.0 + 0.5);
                if (new_progress_cnt > 50) new_progress_cnt = 50;
                while (curr_progress_cnt < new_progress_cnt) { curr_progress_cnt++; para_printf("*"); }
                fflush(stdout);
                // progress bar updates about every 2 %
                const double per_gen = gens > 0 ? progress / (double)gens : 0.0;
                if (per_gen > 0.0) step = std::max<int64_t>(1, (int64_t)(2.0 / per_gen));
            }
            int64_t g_after = 0, t_after = 0;
            check(adgpu_job_run_generations(job, step, &g_after, &t_after), "adgpu_job_run_generations", ctx);
            gens = g_after;
        }
        autostopped = autostop.did_stop();
    }
    profile.nev_at_stop = total / mypars->num_of_runs;
    profile.autostopped = autostopped;
    clock_t clock_stop_docking = clock();
    if (mypars->autostop == 0) {
        while (curr_progress_cnt < 50) { curr_progress_cnt++; para_printf("*"); }
        para_printf("\n");
    }
    auto const t3 = std::chrono::steady_clock::now();
    check(adgpu_job_download(job, sim_state.cpu_populations.data(), sim_state.cpu_energies.data(),
                             sim_state.cpu_evals_of_runs.data(), nullptr), "adgpu_job_download", ctx);
    if (host_autostop) {
        autostop.output_final_stddev((int)gens, sim_state.cpu_energies.data(), total);
    } else if (mypars->autostop) {
        // output_final_stddev (autostop.hpp:233-244) from the device statistics
        adgpu_autostop_record f{};
        check(adgpu_job_autostop_final(job, &f), "adgpu_job_autostop_final", ctx);
        const char* rule = "------------+--------------+------------------+------------------------------+---------+-------------------\n";
        if (f.autostopped) {
            para_printf("%s", rule);
            para_printf("\n%43s evaluation after reaching\n%40.2f +/-%8.2f kcal/mol combined.\n%31i samples, best inter + intra %8.2f kcal/mol.\n",
                        "Finished", f.roll_average, f.roll_stddev, f.roll_samples, f.best_energy);
        } else {
            para_printf("%11u | %12lu |%8.2f kcal/mol |%8.2f +/-%8.2f kcal/mol |%8i |%8.2f kcal/mol\n", (unsigned)gens,
                        (unsigned long)total / mypars->num_of_runs, f.threshold, f.average, f.stddev, f.samples,
                        f.best_energy);
            para_printf("%s", rule);
            para_printf("\n%43s evaluation after reaching\n%30lu evaluations. Best inter + intra %8.2f kcal/mol.\n",
                        "Finished", (unsigned long)total / mypars->num_of_runs, f.best_energy);
        }
        fflush(stdout);
    }
    para_printf("\nDocking time %fs\n", elapsed_seconds(t2, t3));

    // Each run's best pose (the individual process_result puts first, the
    // lowest FP32 energy of the run) re-scored in FP64 on the GPU.  The dlg's
    // energies still come from the reference's CPU re-scoring of every
    // individual (processresult.cpp:430-526), whose intramolecular energy
    // uses distance tables; these are the same terms evaluated analytically.
    // For flexible residues the split between the columns differs from the
    // dlg's (residue-ligand pairs count as intra here); the totals agree.
    {
        std::vector<float> best_geno((size_t)nrun * ADGPU_GENOTYPE_STRIDE), best_e(nrun);
        std::vector<int32_t> best_i(nrun);
        std::vector<double> e_tot(nrun), e_inter(nrun), e_intra(nrun);
        check(adgpu_job_best_f64(job, best_geno.data(), best_e.data(), best_i.data(), e_tot.data(), e_inter.data(),
                                 e_intra.data()), "adgpu_job_best_f64", ctx);
        para_printf("\nFP64 re-scoring on the GPU of each run's best pose (kcal/mol):\n");
        para_printf("    Run | Entity |  FP32 energy |   FP64 total |   FP64 inter |   FP64 intra\n");
        for (int r = 0; r < nrun; ++r)
            para_printf("FP64 %3d | %6d | %12.4f | %12.4f | %12.4f | %12.4f\n", r + 1, best_i[r] + 1,
                        (double)best_e[r], e_tot[r], e_inter[r], e_intra[r]);
        fflush(stdout);
    }

    sim_state.myligand_reference = myligand_reference;
    sim_state.generation_cnt = (int)gens;
    sim_state.sec_per_run = ((float)(clock_stop_docking - clock_start_docking) / (float)CLOCKS_PER_SEC) /
                            mypars->num_of_runs;
    sim_state.total_evals = (unsigned long)total;

    adgpu_job_destroy(job);
    adgpu_ligand_release(ctx, lig_id);
    auto const t4 = std::chrono::steady_clock::now();
    para_printf("\nShutdown time %fs\n", elapsed_seconds(t3, t4));
    if (outbuf) free(outbuf);
    return 0;
}
