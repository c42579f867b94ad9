// ref_driver.cpp -- C entry points over the reference's own host C++ code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference sources where they lie (host/src/{calcenergy,getparameters,
// miscellaneous,processgrid,processligand,processresult,setup}.cpp)
// into oracle/_ref/libadref.so.  It gives tests the reference's own:
//   - ligand preparation (pdbqt -> device constant tables,
//     prepare_const_fields_for_gpu, host/src/calcenergy.cpp:31-345),
//   - packed map layout (get_gridvalues, host/src/processgrid.cpp:186-253),
//   - initial population and PRNG seeds (docking_with_gpu,
//     host/src/performdocking.cpp.Cuda:306-317),
//   - CPU scoring of a genotype exactly as make_resfiles does it
//     (host/src/processresult.cpp:430-466),
//   - the AutoStop test (host/inc/autostop.hpp).
#include <cstring>
#include <cstdio>
#include <vector>
#include <string>

#include "calcenergy.h"
#include "processgrid.h"
#include "processligand.h"
#include "getparameters.h"
#include "filelist.hpp"
#include "setup.hpp"

namespace {
struct RefHandle {
    Dockpars pars;
    Gridinfo grid;
    Liganddata lig_init, lig_xray, lig_ref;
    std::vector<float> init_pop;
    std::vector<uint32_t> seeds;
    kernelconstant_interintra ki;
    kernelconstant_intracontrib kc;
    kernelconstant_intra kin;
    kernelconstant_rotlist krl;
    kernelconstant_conform kcf;
    kernelconstant_grads kg;
    IntraTables* tables = nullptr;
};
}  // namespace

extern "C" {

// argv as for the autodock_gpu CLI, e.g. {"adgpu", "--ffile", f, "--lfile", l,
// "--nrun", "4", "--psize", "8", "--seed", "11,23,0", ...}.
void* ref_open(int argc, char** argv, int nwi) {
    RefHandle* H = new RefHandle();
    FileList filelist;
    if (initial_commandpars(&argc, argv, &H->pars, &H->grid, filelist) != 0) { delete H; return nullptr; }
    if (setup(&H->grid, &H->pars, H->lig_init, H->lig_xray, filelist, 0, argc, argv) != 0) { delete H; return nullptr; }
    H->lig_ref = H->lig_init;
    size_t ne = (size_t)H->pars.num_of_runs * H->pars.pop_size;
    H->init_pop.assign(ne * GENOTYPE_LENGTH_IN_GLOBMEM, 0.0f);
    gen_initpop_and_reflig(&H->pars, H->init_pop.data(), &H->lig_ref, &H->grid);
    LocalRNG r(H->pars.seed);
    H->seeds.resize(ne * (size_t)nwi);
    for (size_t i = 0; i < H->seeds.size(); ++i) H->seeds[i] = r.random_uint();
    if (prepare_const_fields_for_gpu(&H->lig_ref, &H->pars, &H->ki, &H->kc, &H->kin, &H->krl, &H->kcf, &H->kg) != 0) {
        delete H; return nullptr;
    }
    H->tables = new IntraTables(&H->lig_ref, H->pars.coeffs.scaled_AD4_coeff_elec, H->pars.coeffs.AD4_coeff_desolv,
                                H->pars.qasp, H->pars.nr_mod_atype_pairs, H->pars.mod_atype_pairs);
    return H;
}

void ref_close(void* h) {
    RefHandle* H = (RefHandle*)h;
    if (!H) return;
    delete H->tables;
    delete H;
}

// ints: num_atoms, true_ligand_atoms, num_atypes, num_rotbonds, num_pairs,
//       num_map_atypes, gx, gy, gz, num_rotcyc, num_runs, pop_size, num_lsentities_x, nwi_unused
// dbls: spacing, coeff_elec, coeff_desolv, qasp, smooth, elec_min_distance,
//       size_angstr x/y/z, origo x/y/z, abs_max_dmov, abs_max_dang,
//       base_dmov_mul_sqrt3, base_dang_mul_sqrt3, lsearch_rate, rho_lower_bound
void ref_info(void* h, int* iv, double* dv) {
    RefHandle* H = (RefHandle*)h;
    iv[0] = H->lig_ref.num_of_atoms;
    iv[1] = H->lig_ref.true_ligand_atoms;
    iv[2] = H->lig_ref.num_of_atypes;
    iv[3] = H->lig_ref.num_of_rotbonds;
    iv[4] = H->lig_ref.num_of_intraE_contributors;
    iv[5] = H->grid.num_of_map_atypes;
    iv[6] = H->grid.size_xyz[0]; iv[7] = H->grid.size_xyz[1]; iv[8] = H->grid.size_xyz[2];
    iv[9] = H->lig_ref.num_of_rotcyc;
    iv[10] = (int)H->pars.num_of_runs; iv[11] = (int)H->pars.pop_size;
    iv[12] = 0; iv[13] = 0;
    dv[0] = H->grid.spacing;
    dv[1] = H->pars.coeffs.scaled_AD4_coeff_elec;
    dv[2] = H->pars.coeffs.AD4_coeff_desolv;
    dv[3] = H->pars.qasp;
    dv[4] = H->pars.smooth;
    dv[5] = H->pars.elec_min_distance;
    for (int k = 0; k < 3; ++k) { dv[6 + k] = H->grid.size_xyz_angstr[k]; dv[9 + k] = H->grid.origo_real_xyz[k]; }
    dv[12] = H->pars.abs_max_dmov; dv[13] = H->pars.abs_max_dang;
    dv[14] = H->pars.base_dmov_mul_sqrt3; dv[15] = H->pars.base_dang_mul_sqrt3;
    dv[16] = H->pars.lsearch_rate; dv[17] = H->pars.rho_lower_bound;
}

// Pointers to the reference's own arrays (valid until ref_close).
const void* ref_ptr(void* h, const char* name) {
    RefHandle* H = (RefHandle*)h;
    std::string n(name);
    if (n == "charges") return H->ki.atom_charges_const;
    if (n == "atom_types") return H->ki.atom_types_const;
    if (n == "atom_map") return H->ki.atom_types_map_const;
    if (n == "ignore_inter") return H->ki.ignore_inter_const;
    if (n == "pairs") return H->kc.intraE_contributors_const;
    if (n == "atype_reqm") return H->kin.atom_types_reqm_const;
    if (n == "vw_exp") return H->kin.VWpars_exp_const;
    if (n == "reqm_ab") return H->kin.reqm_AB_const;
    if (n == "vw_ac") return H->kin.VWpars_AC_const;
    if (n == "vw_bd") return H->kin.VWpars_BD_const;
    if (n == "ds_s") return H->kin.dspars_S_const;
    if (n == "ds_v") return H->kin.dspars_V_const;
    if (n == "rotlist") return H->krl.rotlist_const;
    if (n == "ref_coords") return H->kcf.ref_coords_const;
    if (n == "rotbond_moving") return H->kcf.rotbonds_moving_vectors_const;
    if (n == "rotbond_unit") return H->kcf.rotbonds_unit_vectors_const;
    if (n == "rotbonds") return H->kg.rotbonds;
    if (n == "rotbonds_atoms") return H->kg.rotbonds_atoms;
    if (n == "num_rotating_atoms") return H->kg.num_rotating_atoms_per_rotbond;
    if (n == "grids") return H->grid.grids.data();
    if (n == "init_pop") return H->init_pop.data();
    if (n == "seeds") return H->seeds.data();
    return nullptr;
}

// CPU re-scoring of one genotype as make_resfiles does it
// (host/src/processresult.cpp:430-466).  Writes inter and intra energies,
// and, if xyz != NULL, the atom coordinates in grid units after
// change_conform_f, and per-atom vdW/elec inter energies if peratom != NULL
// (2*num_atoms floats).  Inter includes the grid energy of flexible-residue
// atoms.
void ref_score(void* h, const float* genotype, double* inter, double* intra, double* xyz, float* peratom) {
    RefHandle* H = (RefHandle*)h;
    Liganddata lig = H->lig_ref;
    change_conform_f(&lig, &H->grid, genotype, 0);
    if (xyz)
        for (int a = 0; a < lig.num_of_atoms; ++a)
            for (int k = 0; k < 3; ++k) xyz[3 * a + k] = lig.atom_idxyzq[a][k + 1];
    float intraflexE = 0.0f, interflexE = 0.0f, elecE = 0.0f;
    float ie;
    if (peratom)
        ie = calc_interE_f(&H->grid, &lig, 0.0005, 0, intraflexE, &elecE, peratom, peratom + lig.num_of_atoms);
    else
        ie = calc_interE_f(&H->grid, &lig, 0.0005, 0, intraflexE);
    scale_ligand(&lig, H->grid.spacing);
    float ia = calc_intraE_f(&lig, 8, H->pars.smooth, 0, H->pars.elec_min_distance, H->tables, 0, interflexE);
    // calc_interE_f books the grid energy of flexible-residue atoms in
    // intraflexE; the GPU energy (calcenergy.cl:224-315) counts every atom as
    // intermolecular, so both parts are returned together.
    *inter = (double)ie + (double)intraflexE;
    *intra = ia;
}

}  // extern "C"

extern "C" {
// Per-atom-pair matrices of the prepared reference ligand (num_atoms^2 bytes):
// "bonds" (get_bonds) or "intra" (get_intraE_contributors); "rigid" writes
// num_atoms ints (atom_rigid_structures).
int ref_matrix(void* h, const char* name, void* out) {
    RefHandle* H = (RefHandle*)h;
    std::string n(name);
    int N = H->lig_ref.num_of_atoms;
    if (n == "bonds" || n == "intra") {
        unsigned char* o = (unsigned char*)out;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                o[i * N + j] = (unsigned char)(n == "bonds" ? H->lig_ref.bonds[i][j] : H->lig_ref.intraE_contributors[i][j]);
        return 0;
    }
    if (n == "rigid") {
        int* o = (int*)out;
        for (int i = 0; i < N; ++i) o[i] = H->lig_ref.atom_rigid_structures[i];
        return 0;
    }
    return 1;
}
}

// The reference's AutoStop class (host/inc/autostop.hpp), with its private
// members opened (declared as a struct) so tests can compare the whole state
// after every test.  Its own includes are included first.
#include <cmath>
#include <stdio.h>
#include <stdlib.h>
#include "miscellaneous.h"
#define class struct
#include "autostop.hpp"
#undef class

extern "C" {

void* ref_autostop_new(int pop_size, int num_runs, float stopstd, int as_frequency) {
    return new AutoStop(pop_size, num_runs, stopstd, as_frequency, nullptr);
}

void ref_autostop_free(void* h) { delete (AutoStop*)h; }

// State after a call: 33 words, floats as floats and counters as uint32:
// threshold, threshold_used, thres_stddev, curr_avg, curr_std, delta_energy,
// overall_best_energy, rolling[16], average_sd2_N[0..2], roll_count, bestN,
// first_time, autostopped.
static void ref_autostop_state(const AutoStop* a, uint32_t* out) {
    const float f[7] = {a->threshold, a->threshold_used, a->thres_stddev, a->curr_avg, a->curr_std,
                        a->delta_energy, a->overall_best_energy};
    std::memcpy(out, f, sizeof(f));
    std::memcpy(out + 7, a->rolling.data(), 16 * sizeof(float));
    std::memcpy(out + 23, a->average_sd2_N.data(), 3 * sizeof(float));
    out[26] = a->roll_count;
    out[27] = a->bestN;
    out[28] = a->first_time;
    out[29] = a->autostopped;
    out[30] = a->Ntop;
    out[31] = a->Ncream;
    out[32] = a->as_frequency;
}

int ref_autostop_check(void* h, int generation, const float* energies, unsigned long total_evals, uint32_t* state) {
    AutoStop* a = (AutoStop*)h;
    const bool r = a->check_if_satisfactory(generation, energies, total_evals);
    if (state) ref_autostop_state(a, state);
    return r ? 1 : 0;
}

void ref_autostop_final(void* h, int generation, const float* energies, unsigned long total_evals, uint32_t* state) {
    AutoStop* a = (AutoStop*)h;
    a->output_final_stddev(generation, energies, total_evals);
    if (state) ref_autostop_state(a, state);
}

}  // extern "C"
