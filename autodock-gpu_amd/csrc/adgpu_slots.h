// adgpu_slots.h -- interface between the job layer (adgpu_host.cpp) and the
// serving loop (adgpu_serve.cpp).  Not part of the public C ABI.
//
// A slot is a docking job whose device buffers (populations, energies, PRNG
// states, the ligand block, AutoStop state) outlive the request it runs: the
// next request reuses them, so a serving loop in steady state allocates and
// frees no device memory (hipFree synchronizes the whole device).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "adgpu_internal.h"

namespace adgpu {

// A ligand validated and packed on the host (the device block of LigandDev),
// ready to be copied into a slot.
struct PreparedLigand {
    std::vector<char> blob;
    LigandDev hd{};
    int max_map = 0;        // largest atom_map index of an atom with receptor energy
};

int prepare_ligand(adgpu_context* c, const adgpu_ligand* L, const adgpu_energy_params* ep, PreparedLigand& out);
// Checks a request against the context: the map set exists and has the
// ligand's type maps, and the search parameters are valid.
int check_request(adgpu_context* c, int map_id, const PreparedLigand& lig, const adgpu_search_params* sp);
// Adds delta to the uses of a map set (a queued or running request holds
// one, so the map set cannot be released under it).
int map_hold(adgpu_context* c, int map_id, int delta);
// Device bytes one request occupies in an empty slot (population buffers,
// PRNG states, LS state, ligand block, AutoStop state when as_frequency > 0,
// re-scoring scratch when rescore).
size_t request_device_bytes(const adgpu_context* c, const PreparedLigand& lig, const adgpu_search_params* sp,
                            int as_frequency, bool rescore);
// Device bytes a slot's size-dependent buffers hold now (they outlive the
// request, for the next one), and what they will hold once the slot has
// bound the request (buffers only grow).
size_t slot_held_bytes(const adgpu_job* j);
size_t slot_bytes_after(const adgpu_job* j, const PreparedLigand& lig, const adgpu_search_params* sp,
                        int as_frequency, bool rescore);
// Frees an idle slot's size-dependent buffers (after which it holds 0 bytes).
void slot_shrink(adgpu_job* j);

// sharers: requests that may run on the device at the same time (the
// server's slots); the job's team width counts their searches too.
int slot_open(adgpu_context* c, int run_groups, int sharers, adgpu_job** out);
// Binds the slot to a request: holds a use of the map set, grows buffers
// that are too small, uploads the ligand block, the initial population and
// the PRNG seeds, and resets the generation state.
int slot_bind(adgpu_job* j, int map_id, const PreparedLigand& lig, const adgpu_search_params* sp,
              const float* init_pop, const uint32_t* seeds);
// Releases the map-set use taken by slot_bind.
void slot_unbind(adgpu_job* j);
// Per-run best individual of the current populations (the first minimum of
// each run's energies), optionally its FP64 re-scoring (total, inter, intra
// per run) and the whole populations.  Any output may be null.
int slot_results(adgpu_job* j, float* best_geno, float* best_e, double* best_f64, float* pop, float* ene);
int64_t slot_reallocations(const adgpu_job* j);
int context_device(const adgpu_context* c);
int set_error(int code, const std::string& msg);
const char* thread_error();

}  // namespace adgpu
