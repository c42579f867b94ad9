// adgpu_host.cpp -- host side of the C ABI declared in include/adgpu.h.
//
// Owns the HIP context objects (one stream per context, default priority),
// the resident map sets and ligands, and the docking jobs.  The generation
// loop of adgpu_job_run follows docking_with_gpu
// (host/src/performdocking.cpp.Cuda:685-831): K4, the local search, K2 and the
// population pointer flip, stopped by check_progress (:908-942).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "adgpu_internal.h"
#include "adgpu_slots.h"

using namespace adgpu;

namespace {

struct MapSet {
    GridDev dev{};
    void* data = nullptr;
    size_t bytes = 0;
    int jobs = 0;         // jobs created on these maps and not yet destroyed
};

struct LigandSet {
    LigandDev host{};
    void* dev = nullptr;
    int max_map = 0;      // largest atom_map index (must be < the grid's type maps)
    int jobs = 0;         // jobs created for this ligand and not yet destroyed
};

}  // namespace

struct adgpu_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, MapSet> maps;
    std::map<int, LigandSet> ligands;
    int next_id = 1;
    std::mutex mu;
    std::atomic<int> team_hint{0};   // 0 = automatic team width, else 16, 32 or 64 (ADGPU_TEAM_WIDTH)
    std::atomic<int> ls_yield{32};   // ADADELTA hand-off point in iterations, 0 = off (ADGPU_LS_YIELD)
    std::atomic<int> ls_regroup{-1}; // ADADELTA regrouping once the queue drains: iterations, 0 = off,
                                     // -1 = kRegroupAuto with 4 or more run groups (ADGPU_LS_REGROUP)
    std::atomic<int> graph_block{0}; // generations per captured graph, 0 = launched one by one (ADGPU_GRAPH_BLOCK)
    std::mutex table_mu;
    std::vector<std::pair<int, int>> team_table;   // (max atoms, width), ascending; adgpu_set_team_table
    // scratch for adgpu_eval_poses
    float* d_geno = nullptr;
    float* d_ene = nullptr;
    float* d_grad = nullptr;
    int scratch_n = 0;
    double* d_f64 = nullptr;   // FP64 re-scoring results [3n]
    int f64_n = 0;
    float* noalloc = nullptr;  // this device's staging area of adgpu_eval_poses_noalloc
};

// Runs of a job are split into groups, each with its own stream, so that the
// generations of different groups overlap on the device (a group's local
// search tail is filled with other groups' work).
constexpr int kRing = 16;   // generations a group may have in flight

struct RunGroup {
    int r0 = 0, r1 = 0;                 // runs [r0, r1)
    hipStream_t stream = nullptr;
    int parity = 0;                     // current population in buffer A (0) or B (1)
    int parity0 = 0;                    // parity when the current adgpu_job_run started
    int64_t done = 0, enq = 0;          // generations completed / enqueued in this adgpu_job_run
    std::vector<int64_t> totals;        // evals of the group's runs after each completed generation
    bool ls_used[kRing] = {};           // the slot's LS launch was timed by ev_ls
    bool gen_timed[kRing] = {};         // the slot's K4 launch was timed by ev_gen
    hipEvent_t ev_done[kRing] = {}, ev_gen[2][kRing] = {}, ev_ls[2][kRing] = {};
    // graph-captured blocks of generations, by the ring slot they start at:
    // the block's launches, events and eval-total copies are those of its
    // slots, so a block replays unchanged whenever its first slot comes round
    // again with the same parity and local search
    hipGraphExec_t blk[kRing] = {};
    int blk_parity[kRing] = {}, blk_method[kRing] = {}, blk_len[kRing] = {};
};

constexpr int kMaxGroups = 8;
constexpr int64_t kGroupSearches = 1500;   // local searches per generation per default run group

// Flags of the events the host waits on: blocking (the thread sleeps until
// the device signals) when ADGPU_HOST_WAIT=block, else the runtime's wait.
static std::atomic<unsigned> g_wait_event_flags{hipEventDisableTiming};

struct adgpu_job {
    adgpu_context* ctx = nullptr;
    int map_id = 0, ligand_id = 0;
    bool counted = false;               // holds a use of its maps and ligand
    // capacities of the device buffers; a serving slot reuses them for the
    // next request and grows them only when a request needs more
    size_t cap_ne = 0, cap_ls = 0;
    int cap_runs = 0;
    size_t as_bins_cap = 0;
    int as_log_alloc = 0;
    std::atomic<int64_t> reallocs{0};   // buffer growths after the first allocation (read by server stats)
    // serving slot: the ligand block belongs to the job (not a context id)
    bool own_lig = false;
    int want_groups = 1;                // run groups of a serving slot (fewer when a request has fewer runs)
    int sharers = 1;                    // requests that may run at once with this job (a server's slots)
    int team_hint = 0;                  // team width of the job's kernels, fixed when it starts
    int team_lanes = 0;                 // the lanes per team that width gives the job's ligand
    int graph_block = 0;                // generations per captured graph, fixed when it starts (0: none)
    LigandSet own{};
    size_t own_cap = 0;
    float* d_best = nullptr;            // per-run best genotypes for FP64 re-scoring
    double* d_f64 = nullptr;
    int best_cap = 0;
    SearchDev sd{};
    adgpu_search_params sp{};
    float* pop[2] = {nullptr, nullptr};
    float* ene[2] = {nullptr, nullptr};
    int32_t* evals_new = nullptr;
    int32_t* evals_runs = nullptr;
    uint32_t* prng = nullptr;
    int32_t* counters = nullptr;        // one LS work-queue counter per group (16 ints apart)
    int32_t* ls_ent = nullptr;
    float* ls_grad = nullptr;
    float* ls_state = nullptr;          // ADADELTA hand-off: saved searches, ring (cap_yield LS slots)
    int32_t* ls_ring = nullptr;
    size_t cap_yield = 0;
    int32_t* h_ring = nullptr;          // pinned [kRing][num_runs] eval totals per generation
    int32_t* d_ring = nullptr;          // device [kRing][num_runs]: the same, written by K2 in captured blocks
    std::vector<RunGroup> groups;
    int64_t generation = 0;
    double kernel_ms[2] = {0.0, 0.0};
    int64_t kernel_n[2] = {0, 0};
    // launch intervals (ms after ev_origin) per kernel class, for the time the
    // device spent in that class when launches of several groups overlap
    hipEvent_t ev_origin = nullptr;
    bool origin_set = false;
    std::vector<std::pair<double, double>> spans[2];
    // AutoStop on the device (k_autostop.hip)
    int32_t* d_stop = nullptr;          // raised by a passed test; every generation kernel then returns
    int32_t* h_stop = nullptr;          // pinned copy, written after each test
    bool stopped = false;               // a passed test has ended the job
    bool as_enabled = false;
    AutoStopParams as_p{};
    AutoStopDev* d_as = nullptr;
    float* d_as_bins = nullptr;
    adgpu_autostop_record* d_as_log = nullptr;
    adgpu_autostop_record* d_as_out = nullptr;
    hipEvent_t ev_test = nullptr;
};

namespace {

// The context's team width for a ligand of num_atoms atoms: its fixed width,
// else its per-size table, else 0 (the kernels' built-in rule).
int context_team_width(adgpu_context* c, int num_atoms) {
    int w = c->team_hint.load(std::memory_order_relaxed);
    if (w == 0) {
        std::lock_guard<std::mutex> lk(c->table_mu);
        for (const auto& e : c->team_table)
            if (num_atoms <= e.first) { w = e.second; break; }
    }
    return w;
}

// team_hint < 0: the context's width for this ligand; else that width (a
// job's, fixed when it starts).
LaunchArgs launch_args(adgpu_context* c, const LigandSet& ls, const MapSet& ms, hipStream_t stream = nullptr,
                       int share = 1, int team_hint = -1) {
    LaunchArgs la{};
    la.stream = stream ? stream : c->stream;
    la.share = share;
    la.lig_dev = reinterpret_cast<const LigandDev*>(ls.dev);
    la.lig_host = &ls.host;
    la.grid = ms.dev;
    la.team_hint = team_hint >= 0 ? team_hint : context_team_width(c, ls.host.num_atoms);
    return la;
}

// Local searches in flight per generation below which 64-lane teams are
// faster for ligands of 33-128 atoms.  With one team per wavefront a
// launch's end does not leave half-idle wavefronts, which pays while the
// searches are few enough for the launch tails to matter; with more, 32-lane
// teams (two per wavefront) do the same work in fewer instructions.
constexpr int64_t kWideTeamSearches = 20000;

// A job's team width, fixed when it starts: the context's width or table,
// else (up to 32 atoms, see below) 32 lanes, and 64 above 64; between, 64 lanes when the
// searches in flight -- the job's own times the requests that may share the
// device with it (a server's slots) -- are below kWideTeamSearches, else 32.
// Above 64 atoms a 32-lane team carries three or four atom chunks per lane;
// measured with 16 requests of nrun 20 served at once (48000 searches), M
// evals/s at 72 / 90 / 110 atoms, 32 vs 64 lanes: ADADELTA 52 / 33 / 19 vs
// 86 / 57 / 35, FIRE 135 / 99 / 71 vs 135 / 100 / 74.
// Measured, one job alone: lognormal 8-90-atom mix, FIRE nrun 20 (3000
// searches) 0.101 s per ligand (32) vs 0.081 s (64); ligand + 2 flexible
// residues (54 atoms), ADADELTA nrun 100 (15000 searches) 198 vs 226 M
// evals/s; ADAM nrun 64 (9600 searches) at 50 / 60 atoms 240 / 164 vs 260 /
// 176 M evals/s.  Shared: 4 slots x ADAM nrun 50 (30000 searches), 10-60
// atoms, 16.0 vs 14.5 k ligands/h; 32 slots x FIRE nrun 20 (96000 searches),
// 85.4 vs 82.8 k ligands/h (64 from 33 atoms).
//
// Up to 32 atoms the local search decides: Solis-Wets and ADADELTA searches
// end after uneven iteration counts, so the two searches of a 32-lane
// wavefront seldom end together, and while the searches are few the
// half-idle wavefronts are a large share of each launch; 64-lane teams have
// none, at more instructions per evaluation.  Measured on the 30-atom ligand
// alone (2.5 M evaluations per run), M evals/s, 32 vs 64 lanes: Solis-Wets
// nrun 10 / 20 / 30 / 40 / 45 / 50: 232 / 448 / 631 / 781 / 865 / 941 vs
// 290 / 539 / 748 / 846 / 887 / 924; ADADELTA nrun 20 / 30 / 40 / 50 / 60 /
// 80 / 100: 211 / 288 / 356 / 431 / 502 / 646 / 740 vs 227 / 324 / 452 / 518
// / 533 / 538 / 571.  FIRE and ADAM run fixed iteration counts and keep 32
// lanes (nrun 20: FIRE 436 vs 423, ADAM 414 vs 380).
constexpr int64_t kWideTeamSwSearches = 7000;
constexpr int64_t kWideTeamAdSearches = 10000;

int job_team_width(adgpu_context* c, int num_atoms, int64_t searches, int sharers, int ls_method) {
    const int w = context_team_width(c, num_atoms);
    if (w) return w;
    const int64_t in_flight = searches * std::max(1, sharers);
    if (num_atoms <= 32) {
        const int64_t few = ls_method == ADGPU_LS_SW ? kWideTeamSwSearches
                          : ls_method == ADGPU_LS_AD ? kWideTeamAdSearches : 0;
        return in_flight < few ? 64 : 32;
    }
    if (num_atoms > 64) return 64;
    return in_flight < kWideTeamSearches ? 64 : 32;
}

// Message of the calling thread's last failed call (jobs of one context may
// be driven from several threads, e.g. the serving loop's workers).
thread_local std::string t_err;

int fail(adgpu_context* c, int code, const std::string& msg) {
    (void)c;
    t_err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), ADGPU_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// FP16 storage saturates at the largest finite half (maps clip repulsive
// cores at 1e5, above the FP16 range).
uint16_t f32_to_f16(float f) {
    if (f > 65504.0f) f = 65504.0f;
    if (f < -65504.0f) f = -65504.0f;
    _Float16 h = (_Float16)f;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

uint16_t f32_to_bf16(float f) {   // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

// Builds the single device block of one ligand (see LigandDev).
int build_ligand(adgpu_context* c, const adgpu_ligand* L, const adgpu_energy_params* ep,
                 std::vector<char>& blob, LigandDev& hd) {
    const int N = L->num_atoms, T = L->num_rotbonds, NT = L->num_atypes, P = L->num_pairs;
    if (N <= 0 || N > ADGPU_MAX_ATOMS) return fail(c, ADGPU_ERR_LIMIT, "num_atoms out of range (1..256)");
    if (T < 0 || T > ADGPU_MAX_ROTBONDS) return fail(c, ADGPU_ERR_LIMIT, "num_rotbonds out of range (0..57)");
    if (NT <= 0 || NT > ADGPU_MAX_ATYPES) return fail(c, ADGPU_ERR_LIMIT, "num_atypes out of range (1..32)");
    if (P < 0 || P > N * (N - 1) / 2) return fail(c, ADGPU_ERR_LIMIT, "num_pairs out of range");
    if (L->true_ligand_atoms < 0 || L->true_ligand_atoms > N) return fail(c, ADGPU_ERR_ARG, "true_ligand_atoms out of range");
    if (!(ep->elec_min_distance > 0.0f)) return fail(c, ADGPU_ERR_ARG, "elec_min_distance must be positive");
    if (!L->charges || !L->atom_types || !L->atom_map || !L->ref_coords || (P && !L->pairs) ||
        !L->atype_reqm || !L->vw_exp || !L->reqm_ab || !L->vw_ac || !L->vw_bd || !L->ds_s || !L->ds_v ||
        (T && (!L->rotbond_moving || !L->rotbond_unit || !L->rotbond_atoms || !L->atom_rotbonds)))
        return fail(c, ADGPU_ERR_ARG, "null ligand array");
    std::memset(&hd, 0, sizeof(hd));
    hd.num_atoms = N;
    hd.true_ligand_atoms = L->true_ligand_atoms;
    hd.num_rotbonds = T;
    hd.num_genes = T + 6;
    hd.num_pairs = P;
    hd.nchunk = (N + 31) / 32;
    hd.coeff_elec = ep->coeff_elec;
    hd.coeff_desolv = ep->coeff_desolv;
    hd.qasp = ep->qasp;
    hd.half_smooth = 0.5f * ep->smooth;
    hd.elec_min_distance = ep->elec_min_distance;
    hd.inv_elec_min = 1.0f / ep->elec_min_distance;
    size_t off = (sizeof(LigandDev) + 255) & ~size_t(255);
    auto take = [&](size_t bytes) { size_t o = off; off = (off + bytes + 255) & ~size_t(255); return (uint32_t)o; };
    hd.off_atom_ref = take(sizeof(float4) * N);
    hd.off_atom_info = take(sizeof(int4) * N);
    hd.off_atom_mask = take(sizeof(uint2) * N);
    hd.off_tor_unit = take(sizeof(float4) * (T ? T : 1));
    hd.off_tor_move = take(sizeof(float4) * (T ? T : 1));
    hd.off_tor_atoms = take(sizeof(int2) * (T ? T : 1));
    hd.off_pair_a = take(sizeof(PairA) * (P ? P : 1));
    hd.off_pair_b = take(sizeof(PairB) * (P ? P : 1));
    hd.off_type_ds = take(sizeof(float2) * NT);
    hd.total_bytes = (uint32_t)off;
    blob.assign(off, 0);
    float4* ar = reinterpret_cast<float4*>(blob.data() + hd.off_atom_ref);
    int4* ai = reinterpret_cast<int4*>(blob.data() + hd.off_atom_info);
    uint2* am = reinterpret_cast<uint2*>(blob.data() + hd.off_atom_mask);
    for (int a = 0; a < N; ++a) {
        if (L->atom_types[a] < 0 || L->atom_types[a] >= NT) return fail(c, ADGPU_ERR_ARG, "atom type out of range");
        ar[a] = make_float4(L->ref_coords[3 * a], L->ref_coords[3 * a + 1], L->ref_coords[3 * a + 2], L->charges[a]);
        ai[a] = make_int4(L->atom_map[a], L->atom_types[a], L->ignore_inter ? (int)L->ignore_inter[a] : 0, 0);
        uint64_t m = 0;
        for (int t = 0; t < T; ++t)
            if (L->atom_rotbonds[(size_t)a * T + t]) m |= (uint64_t)1 << t;
        am[a] = make_uint2((uint32_t)m, (uint32_t)(m >> 32));
    }
    float4* tu = reinterpret_cast<float4*>(blob.data() + hd.off_tor_unit);
    float4* tm = reinterpret_cast<float4*>(blob.data() + hd.off_tor_move);
    int2* ta = reinterpret_cast<int2*>(blob.data() + hd.off_tor_atoms);
    for (int t = 0; t < T; ++t) {
        tu[t] = make_float4(L->rotbond_unit[3 * t], L->rotbond_unit[3 * t + 1], L->rotbond_unit[3 * t + 2], 0.f);
        tm[t] = make_float4(L->rotbond_moving[3 * t], L->rotbond_moving[3 * t + 1], L->rotbond_moving[3 * t + 2], 0.f);
        int a0 = L->rotbond_atoms[2 * t], a1 = L->rotbond_atoms[2 * t + 1];
        if (a0 < 0 || a0 >= N || a1 < 0 || a1 >= N) return fail(c, ADGPU_ERR_ARG, "rotbond atom out of range");
        ta[t] = make_int2(a0, a1);
    }
    // Pair order on the device: by index distance j - i, then i.  Consecutive
    // lanes then handle pairs with distinct first and distinct second atoms, so
    // the per-atom LDS force accumulations of one instruction rarely collide
    // (the reference's pair list runs i-major, where up to N-1 lanes would hit
    // the same atom).  Only the order of the FP32 energy sum changes.
    std::vector<int> order(P);
    for (int p = 0; p < P; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        int dx = std::abs(L->pairs[2 * x + 1] - L->pairs[2 * x]), dy = std::abs(L->pairs[2 * y + 1] - L->pairs[2 * y]);
        if (dx != dy) return dx < dy;
        return std::min(L->pairs[2 * x], L->pairs[2 * x + 1]) < std::min(L->pairs[2 * y], L->pairs[2 * y + 1]);
    });
    float2* tds = reinterpret_cast<float2*>(blob.data() + hd.off_type_ds);
    for (int t = 0; t < NT; ++t) tds[t] = make_float2(L->ds_s[t], L->ds_v[t]);
    PairA* pra = reinterpret_cast<PairA*>(blob.data() + hd.off_pair_a);
    PairB* prb = reinterpret_cast<PairB*>(blob.data() + hd.off_pair_b);
    for (int q = 0; q < P; ++q) {
        const int p = order[q];
        PairA& ra = pra[q];
        PairB& rb = prb[q];
        int i = L->pairs[2 * p], j = L->pairs[2 * p + 1];
        if (i < 0 || i >= N || j < 0 || j >= N) return fail(c, ADGPU_ERR_ARG, "pair atom out of range");
        int t1 = L->atom_types[i], t2 = L->atom_types[j];
        int idx = t1 * NT + t2;
        uint32_t m = (L->vw_exp[idx] >> 8) & 0xffu, n = L->vw_exp[idx] & 0xffu;
        uint32_t r1 = L->atype_reqm[t1], r2 = L->atype_reqm[t2];
        bool cgg0 = (r1 == 24 && r2 == 25) || (r1 == 25 && r2 == 24);   // ATYPE_CG_IDX / ATYPE_G0_IDX
        bool fast = (m == 12) && (n == 6 || n == 10);
        if (!fast) hd.any_general = 1;
        ra.ij = (uint32_t)i | ((uint32_t)j << 16);
        ra.flags = m | (n << 8) | ((uint32_t)cgg0 << 16) | ((uint32_t)fast << 17);
        ra.ac = L->vw_ac[idx];
        ra.bd = L->vw_bd[idx];
        rb.reqm = L->reqm_ab[idx];
        float q1 = L->charges[i], q2 = L->charges[j];
        rb.cq = ep->coeff_elec * q1 * q2;                       // calcenergy.cl:527 order
        rb.dsol = (L->ds_s[t1] + ep->qasp * std::fabs(q1)) * L->ds_v[t2] +
                  (L->ds_s[t2] + ep->qasp * std::fabs(q2)) * L->ds_v[t1];   // calcenergy.cl:488-491
        rb.g = cgg0 ? 50.0f : 0.0f;                             // calcenergy.cl:452-458, G = 50
    }
    std::memcpy(blob.data(), &hd, sizeof(hd));
    return ADGPU_OK;
}

int check_search(adgpu_context* c, const adgpu_search_params* sp, int num_genes) {
    if (sp->num_runs <= 0 || sp->num_runs > ADGPU_MAX_RUNS) return fail(c, ADGPU_ERR_LIMIT, "num_runs out of range");
    if (sp->pop_size <= 1 || sp->pop_size > ADGPU_MAX_POPSIZE) return fail(c, ADGPU_ERR_LIMIT, "pop_size out of range");
    if (sp->ls_method < ADGPU_LS_NONE || sp->ls_method > ADGPU_LS_FIRE) return fail(c, ADGPU_ERR_ARG, "bad ls_method");
    if (sp->max_num_of_iters < 0) return fail(c, ADGPU_ERR_ARG, "max_num_of_iters < 0");
    (void)num_genes;
    return ADGPU_OK;
}

SearchDev to_dev(const adgpu_search_params* sp, int num_genes) {
    SearchDev s{};
    s.num_runs = sp->num_runs;
    s.pop_size = sp->pop_size;
    s.num_genes = num_genes;
    s.num_ls = sp->lsearch_rate != 0.0f ? (int)(unsigned)(sp->lsearch_rate / 100.0 * sp->pop_size + 0.5) : 0;
    s.tournament_rate = sp->tournament_rate / 100.0f;
    s.crossover_rate = sp->crossover_rate / 100.0f;
    s.mutation_rate = sp->mutation_rate / 100.f;
    s.abs_max_dmov = sp->abs_max_dmov;
    s.abs_max_dang = sp->abs_max_dang;
    s.lsearch_rate = sp->lsearch_rate;
    s.rho_lower_bound = sp->rho_lower_bound;
    s.base_dmov_mul_sqrt3 = sp->base_dmov_mul_sqrt3;
    s.base_dang_mul_sqrt3 = sp->base_dang_mul_sqrt3;
    s.cons_limit = sp->cons_limit;
    s.max_iters = sp->max_num_of_iters;
    s.ls_method = sp->ls_method;
    s.adam_beta1 = sp->adam_beta1;
    s.adam_beta2 = sp->adam_beta2;
    s.adam_epsilon = sp->adam_epsilon;
    return s;
}

}  // namespace

extern "C" {

const char* adgpu_version(void) { return "adgpu-mi355x 0.1 (gfx950)"; }

const char* adgpu_last_error(adgpu_ctx ctx) { return ctx ? t_err.c_str() : "null context"; }

int adgpu_set_team_width(adgpu_ctx c, int width) {
    if (!c) return ADGPU_ERR_ARG;
    if (width != 0 && width != 16 && width != 32 && width != 64)
        return fail(c, ADGPU_ERR_ARG, "team width must be 0, 16, 32 or 64");
    std::lock_guard<std::mutex> lk(c->mu);
    c->team_hint = width;
    return ADGPU_OK;
}

int adgpu_set_ls_handoff(adgpu_ctx c, int iterations) {
    if (!c) return ADGPU_ERR_ARG;
    if (iterations < 0) return fail(c, ADGPU_ERR_ARG, "hand-off iterations must be >= 0");
    c->ls_yield = iterations;
    return ADGPU_OK;
}

int adgpu_set_ls_regroup(adgpu_ctx c, int iterations) {
    if (!c) return ADGPU_ERR_ARG;
    if (iterations < -1) return fail(c, ADGPU_ERR_ARG, "regroup iterations must be >= -1");
    c->ls_regroup = iterations;
    return ADGPU_OK;
}

int adgpu_set_team_table(adgpu_ctx c, int n, const int32_t* max_atoms, const int32_t* widths) {
    if (!c || n < 0 || n > 16 || (n > 0 && (!max_atoms || !widths)))
        return fail(c, ADGPU_ERR_ARG, "team table: n must be 0..16 with both arrays");
    std::vector<std::pair<int, int>> t;
    for (int i = 0; i < n; ++i) {
        const int a = max_atoms[i], w = widths[i];
        if (a < 1 || a > ADGPU_MAX_ATOMS || (i > 0 && a <= max_atoms[i - 1]))
            return fail(c, ADGPU_ERR_ARG, "team table: max_atoms must ascend within 1..256");
        if (w != 16 && w != 32 && w != 64) return fail(c, ADGPU_ERR_ARG, "team table: widths are 16, 32 or 64");
        if (w == 16 && a > 32) return fail(c, ADGPU_ERR_ARG, "team table: width 16 only up to 32 atoms");
        t.emplace_back(a, w);
    }
    std::lock_guard<std::mutex> lk(c->table_mu);
    c->team_table = std::move(t);
    return ADGPU_OK;
}

int adgpu_device_count(int* out_count) {
    if (!out_count) return ADGPU_ERR_ARG;
    *out_count = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return ADGPU_ERR_HIP;
    *out_count = n;
    return ADGPU_OK;
}

int adgpu_open(int device, adgpu_ctx* out) {
    if (!out) return ADGPU_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ADGPU_ERR_HIP;
    if (device < 0) device = 0;
    if (device >= n) return ADGPU_ERR_ARG;
    std::unique_ptr<adgpu_context> c(new adgpu_context());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return ADGPU_ERR_HIP;
    // How host threads wait for the device (ADGPU_HOST_WAIT = spin, yield or
    // block): many threads waiting at once (the serving loop's workers) share
    // the host cores.  Applies to the process; the runtime's default otherwise.
    if (const char* w = std::getenv("ADGPU_HOST_WAIT")) {
        const std::string v(w);
        const unsigned f = v == "spin" ? hipDeviceScheduleSpin : v == "yield" ? hipDeviceScheduleYield
                         : v == "block" ? hipDeviceScheduleBlockingSync : hipDeviceScheduleAuto;
        if (hipSetDeviceFlags(f) != hipSuccess) {
            (void)hipGetLastError();
            std::fprintf(stderr, "adgpu: ADGPU_HOST_WAIT=%s not applied (device already initialized)\n", w);
        }
        g_wait_event_flags = hipEventDisableTiming | (v == "block" ? hipEventBlockingSync : 0u);
    }
    if (const char* y = std::getenv("ADGPU_LS_YIELD")) c->ls_yield = std::max(0, std::atoi(y));
    if (const char* y = std::getenv("ADGPU_LS_REGROUP")) c->ls_regroup = std::max(-1, std::atoi(y));
    if (const char* b = std::getenv("ADGPU_GRAPH_BLOCK")) {
        const int v = std::atoi(b);
        c->graph_block = (v == 2 || v == 4 || v == 8) ? v : 0;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return ADGPU_ERR_HIP;
    if (noalloc_scratch_setup(device, &c->noalloc) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        return ADGPU_ERR_HIP;
    }
    if (const char* tw = std::getenv("ADGPU_TEAM_WIDTH")) {
        int v = std::atoi(tw);
        c->team_hint = (v == 16 || v == 32 || v == 64) ? v : 0;
    }
    *out = c.release();
    return ADGPU_OK;
}

int adgpu_close(adgpu_ctx c) {
    if (!c) return ADGPU_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->maps) (void)hipFree(kv.second.data);
    for (auto& kv : c->ligands) (void)hipFree(kv.second.dev);
    (void)hipFree(c->d_geno);
    (void)hipFree(c->d_ene);
    (void)hipFree(c->d_grad);
    (void)hipFree(c->d_f64);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return ADGPU_OK;
}

int adgpu_maps_upload(adgpu_ctx c, const adgpu_grid* g, int format, int* out_id) {
    if (!c || !g || !out_id || !g->packed) return fail(c, ADGPU_ERR_ARG, "null argument");
    if (g->size_x < 2 || g->size_y < 2 || g->size_z < 2 || g->size_x > ADGPU_MAX_GRIDPOINTS ||
        g->size_y > ADGPU_MAX_GRIDPOINTS || g->size_z > ADGPU_MAX_GRIDPOINTS)
        return fail(c, ADGPU_ERR_LIMIT, "grid size out of range (2..256 points per axis)");
    if (g->num_map_atypes <= 0 || g->num_map_atypes > ADGPU_MAX_ATYPES)
        return fail(c, ADGPU_ERR_LIMIT, "num_map_atypes out of range");
    if (!(g->spacing > 0.f) || g->spacing > 1.0f) return fail(c, ADGPU_ERR_ARG, "spacing must be in (0, 1]");
    if (format != ADGPU_MAP_F32 && format != ADGPU_MAP_F16 && format != ADGPU_MAP_BF16)
        return fail(c, ADGPU_ERR_ARG, "unknown map format");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    MapSet ms;
    const size_t g3 = (size_t)g->size_x * g->size_y * g->size_z;
    const size_t nmaps = (size_t)g->num_map_atypes + 2;
    const size_t nval = 4 * g3 * nmaps;
    std::vector<uint16_t> half;
    const void* src = g->packed;
    size_t bytes = nval * sizeof(float);
    if (format != ADGPU_MAP_F32) {
        half.resize(nval);
        for (size_t i = 0; i < nval; ++i)
            half[i] = format == ADGPU_MAP_F16 ? f32_to_f16(g->packed[i]) : f32_to_bf16(g->packed[i]);
        src = half.data();
        bytes = nval * sizeof(uint16_t);
    }
    if (hipMalloc(&ms.data, bytes) != hipSuccess) return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(maps) failed");
    hipError_t e = hipMemcpy(ms.data, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(ms.data); return fail(c, ADGPU_ERR_HIP, hipGetErrorString(e)); }
    ms.bytes = bytes;
    GridDev& d = ms.dev;
    d.gx = g->size_x; d.gy = g->size_y; d.gz = g->size_z;
    d.nmaps = (int)nmaps;
    d.g1 = (uint32_t)g->size_x;
    d.g2 = (uint32_t)g->size_x * g->size_y;
    d.g3 = (uint32_t)g3;
    d.format = (uint32_t)format;
    d.spacing = g->spacing;
    d.inv_spacing = 1.0f / g->spacing;
    d.data = ms.data;
    int id = c->next_id++;
    c->maps[id] = ms;
    *out_id = id;
    return ADGPU_OK;
}

int adgpu_maps_release(adgpu_ctx c, int id) {
    if (!c) return ADGPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->maps.find(id);
    if (it == c->maps.end()) return fail(c, ADGPU_ERR_ARG, "unknown map id");
    if (it->second.jobs > 0) return fail(c, ADGPU_ERR_STATE, "maps still used by a job (destroy the job first)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(it->second.data);
    c->maps.erase(it);
    return ADGPU_OK;
}

int adgpu_ligand_upload(adgpu_ctx c, const adgpu_ligand* L, const adgpu_energy_params* ep, int* out_id) {
    if (!c || !L || !ep || !out_id) return fail(c, ADGPU_ERR_ARG, "null argument");
    PreparedLigand pl;
    int rc = prepare_ligand(c, L, ep, pl);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    LigandSet ls;
    ls.host = pl.hd;
    ls.max_map = pl.max_map;
    if (hipMalloc(&ls.dev, pl.blob.size()) != hipSuccess) return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(ligand) failed");
    hipError_t e = hipMemcpy(ls.dev, pl.blob.data(), pl.blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(ls.dev); return fail(c, ADGPU_ERR_HIP, hipGetErrorString(e)); }
    int id = c->next_id++;
    c->ligands[id] = ls;
    *out_id = id;
    return ADGPU_OK;
}

int adgpu_ligand_release(adgpu_ctx c, int id) {
    if (!c) return ADGPU_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->ligands.find(id);
    if (it == c->ligands.end()) return fail(c, ADGPU_ERR_ARG, "unknown ligand id");
    if (it->second.jobs > 0) return fail(c, ADGPU_ERR_STATE, "ligand still used by a job (destroy the job first)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(it->second.dev);
    c->ligands.erase(it);
    return ADGPU_OK;
}

int adgpu_eval_poses(adgpu_ctx c, int map_id, int lig_id, int mode, const float* genotypes, int n,
                     float* energies, float* grads) {
    if (!c || !genotypes || n < 0) return fail(c, ADGPU_ERR_ARG, "null argument");
    if (mode < ADGPU_EVAL_ENERGY || mode > ADGPU_EVAL_GRADIENT) return fail(c, ADGPU_ERR_ARG, "bad mode");
    if (n == 0) return ADGPU_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    auto li = c->ligands.find(lig_id);
    if (mi == c->maps.end() || li == c->ligands.end()) return fail(c, ADGPU_ERR_ARG, "unknown map or ligand id");
    if (li->second.max_map >= mi->second.dev.nmaps - 2)
        return fail(c, ADGPU_ERR_ARG, "ligand atom_map index exceeds the grid's type maps");
    HIPCHK(c, hipSetDevice(c->device));
    if (n > c->scratch_n) {
        (void)hipFree(c->d_geno); (void)hipFree(c->d_ene); (void)hipFree(c->d_grad);
        c->d_geno = c->d_ene = c->d_grad = nullptr;
        c->scratch_n = 0;
        if (hipMalloc(&c->d_geno, sizeof(float) * kGenes * (size_t)n) != hipSuccess ||
            hipMalloc(&c->d_ene, sizeof(float) * (size_t)n) != hipSuccess ||
            hipMalloc(&c->d_grad, sizeof(float) * kGenes * (size_t)n) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(eval scratch) failed");
        c->scratch_n = n;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_geno, genotypes, sizeof(float) * kGenes * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_eval_poses(launch_args(c, li->second, mi->second), mode, c->d_geno, n, c->d_ene, c->d_grad));
    if (energies && mode != ADGPU_EVAL_GRADIENT)
        HIPCHK(c, hipMemcpyAsync(energies, c->d_ene, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (grads && mode != ADGPU_EVAL_ENERGY)
        HIPCHK(c, hipMemcpyAsync(grads, c->d_grad, sizeof(float) * kGenes * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADGPU_OK;
}

int adgpu_eval_poses_noalloc(adgpu_ctx c, int map_id, int lig_id, int mode, const float* genotypes, int n,
                             float* energies, float* grads) {
    if (!c || !genotypes || n < 0) return fail(c, ADGPU_ERR_ARG, "null argument");
    if (mode < ADGPU_EVAL_ENERGY || mode > ADGPU_EVAL_GRADIENT) return fail(c, ADGPU_ERR_ARG, "bad mode");
    if (n > ADGPU_NOALLOC_MAX_POSES)
        return fail(c, ADGPU_ERR_LIMIT, "adgpu_eval_poses_noalloc: n = " + std::to_string(n) +
                                            " exceeds ADGPU_NOALLOC_MAX_POSES = " +
                                            std::to_string(ADGPU_NOALLOC_MAX_POSES));
    if (n == 0) return ADGPU_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    auto li = c->ligands.find(lig_id);
    if (mi == c->maps.end() || li == c->ligands.end()) return fail(c, ADGPU_ERR_ARG, "unknown map or ligand id");
    if (li->second.max_map >= mi->second.dev.nmaps - 2)
        return fail(c, ADGPU_ERR_ARG, "ligand atom_map index exceeds the grid's type maps");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, eval_poses_noalloc(launch_args(c, li->second, mi->second), c->noalloc, mode, genotypes, n, energies,
                                 grads));
    return ADGPU_OK;
}

int adgpu_rescore_f64(adgpu_ctx c, int map_id, int lig_id, const float* genotypes, int n, double* total,
                      double* inter, double* intra) {
    if (!c || !genotypes || n < 0) return fail(c, ADGPU_ERR_ARG, "null argument");
    if (n == 0) return ADGPU_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    auto li = c->ligands.find(lig_id);
    if (mi == c->maps.end() || li == c->ligands.end()) return fail(c, ADGPU_ERR_ARG, "unknown map or ligand id");
    if (li->second.max_map >= mi->second.dev.nmaps - 2)
        return fail(c, ADGPU_ERR_ARG, "ligand atom_map index exceeds the grid's type maps");
    HIPCHK(c, hipSetDevice(c->device));
    if (n > c->scratch_n) {
        (void)hipFree(c->d_geno); (void)hipFree(c->d_ene); (void)hipFree(c->d_grad);
        c->d_geno = c->d_ene = c->d_grad = nullptr;
        c->scratch_n = 0;
        if (hipMalloc(&c->d_geno, sizeof(float) * kGenes * (size_t)n) != hipSuccess ||
            hipMalloc(&c->d_ene, sizeof(float) * (size_t)n) != hipSuccess ||
            hipMalloc(&c->d_grad, sizeof(float) * kGenes * (size_t)n) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(eval scratch) failed");
        c->scratch_n = n;
    }
    if (n > c->f64_n) {
        (void)hipFree(c->d_f64);
        c->d_f64 = nullptr;
        c->f64_n = 0;
        if (hipMalloc(&c->d_f64, sizeof(double) * 3 * (size_t)n) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(FP64 scratch) failed");
        c->f64_n = n;
    }
    std::vector<double> h(3 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(c->d_geno, genotypes, sizeof(float) * kGenes * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_rescore_f64(launch_args(c, li->second, mi->second), c->d_geno, n, c->d_f64));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_f64, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
        if (total) total[i] = h[3 * (size_t)i];
        if (inter) inter[i] = h[3 * (size_t)i + 1];
        if (intra) intra[i] = h[3 * (size_t)i + 2];
    }
    return ADGPU_OK;
}

// Run groups: streams and their per-generation events.
static bool make_group(RunGroup& g) {
    bool ok = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) == hipSuccess;
    for (int r = 0; r < kRing && ok; ++r)
        ok = hipEventCreateWithFlags(&g.ev_done[r], g_wait_event_flags.load()) == hipSuccess &&
             hipEventCreate(&g.ev_gen[0][r]) == hipSuccess && hipEventCreate(&g.ev_gen[1][r]) == hipSuccess &&
             hipEventCreate(&g.ev_ls[0][r]) == hipSuccess && hipEventCreate(&g.ev_ls[1][r]) == hipSuccess;
    return ok;
}

// Drops a group's captured generation blocks (their kernels hold the job's
// buffers and the group's slice, which a new search may change).
static void drop_blocks(RunGroup& g) {
    for (int r = 0; r < kRing; ++r)
        if (g.blk[r]) { (void)hipGraphExecDestroy(g.blk[r]); g.blk[r] = nullptr; }
}

static void destroy_group(RunGroup& g) {
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    drop_blocks(g);   // before the stream its blocks were launched on
    if (g.stream) (void)hipStreamDestroy(g.stream);
    for (int r = 0; r < kRing; ++r) {
        if (g.ev_done[r]) (void)hipEventDestroy(g.ev_done[r]);
        for (int h = 0; h < 2; ++h) {
            if (g.ev_gen[h][r]) (void)hipEventDestroy(g.ev_gen[h][r]);
            if (g.ev_ls[h][r]) (void)hipEventDestroy(g.ev_ls[h][r]);
        }
    }
    g = RunGroup{};
    (void)hipGetLastError();   // nothing of a torn-down group may surface at the next launch
}

// Sets the number of run groups (at most one per run) and splits the runs
// evenly over them.
static int job_set_groups(adgpu_job* j, int ng, int nrun) {
    ng = std::max(1, std::min(ng, std::min(nrun, kMaxGroups)));
    while ((int)j->groups.size() > ng) {
        destroy_group(j->groups.back());
        j->groups.pop_back();
    }
    while ((int)j->groups.size() < ng) {
        j->groups.emplace_back();
        if (!make_group(j->groups.back())) return fail(j->ctx, ADGPU_ERR_NOMEM, "stream or event creation failed");
    }
    for (int k = 0; k < ng; ++k) {
        j->groups[k].r0 = (int)((int64_t)nrun * k / ng);
        j->groups[k].r1 = (int)((int64_t)nrun * (k + 1) / ng);
    }
    return ADGPU_OK;
}

// Allocations of a job that do not depend on the request's size.
static int job_open(adgpu_job* j) {
    bool ok = hipMalloc(&j->counters, 16 * kMaxGroups * sizeof(int32_t)) == hipSuccess &&
              hipMalloc(&j->d_stop, 16 * sizeof(int32_t)) == hipSuccess &&
              hipHostMalloc(&j->h_stop, 16 * sizeof(int32_t)) == hipSuccess &&
              hipEventCreateWithFlags(&j->ev_test, g_wait_event_flags.load()) == hipSuccess &&
              hipEventCreate(&j->ev_origin) == hipSuccess;
    return ok ? ADGPU_OK : fail(j->ctx, ADGPU_ERR_NOMEM, "job allocation failed");
}

// Grows the population, energy, PRNG and FIRE buffers to hold nrun runs of
// pop individuals with nls local-search slots per run.  Buffers that already
// fit are kept; growing frees the old ones (hipFree synchronizes the device).
static int job_reserve(adgpu_job* j, int nrun, int pop, int nls, bool fire, bool yield) {
    const size_t ne = (size_t)nrun * pop;
    if (ne > j->cap_ne) {
        if (j->cap_ne) j->reallocs++;
        (void)hipFree(j->pop[0]); (void)hipFree(j->pop[1]);
        (void)hipFree(j->ene[0]); (void)hipFree(j->ene[1]);
        (void)hipFree(j->evals_new); (void)hipFree(j->prng);
        j->pop[0] = j->pop[1] = j->ene[0] = j->ene[1] = nullptr;
        j->evals_new = nullptr;
        j->prng = nullptr;
        j->cap_ne = 0;
        bool ok = hipMalloc(&j->pop[0], ne * kGenes * sizeof(float)) == hipSuccess &&
                  hipMalloc(&j->pop[1], ne * kGenes * sizeof(float)) == hipSuccess &&
                  hipMalloc(&j->ene[0], ne * sizeof(float)) == hipSuccess &&
                  hipMalloc(&j->ene[1], ne * sizeof(float)) == hipSuccess &&
                  hipMalloc(&j->evals_new, ne * sizeof(int32_t)) == hipSuccess &&
                  hipMalloc(&j->prng, ne * ADGPU_NWI * sizeof(uint32_t)) == hipSuccess;
        if (!ok) return fail(j->ctx, ADGPU_ERR_NOMEM, "job allocation failed");
        j->cap_ne = ne;
    }
    if (nrun > j->cap_runs) {
        if (j->cap_runs) j->reallocs++;
        (void)hipFree(j->evals_runs);
        (void)hipFree(j->d_ring);
        if (j->h_ring) (void)hipHostFree(j->h_ring);
        j->evals_runs = nullptr;
        j->d_ring = nullptr;
        j->h_ring = nullptr;
        j->cap_runs = 0;
        bool ok = hipMalloc(&j->evals_runs, nrun * sizeof(int32_t)) == hipSuccess &&
                  hipMalloc(&j->d_ring, (size_t)kRing * nrun * sizeof(int32_t)) == hipSuccess &&
                  hipHostMalloc(&j->h_ring, (size_t)kRing * nrun * sizeof(int32_t)) == hipSuccess;
        if (!ok) return fail(j->ctx, ADGPU_ERR_NOMEM, "job allocation failed");
        j->cap_runs = nrun;
    }
    const size_t nls_all = (size_t)nrun * (nls > 0 ? nls : 1);
    if (fire && nls_all > j->cap_ls) {
        if (j->cap_ls) j->reallocs++;
        (void)hipFree(j->ls_ent); (void)hipFree(j->ls_grad);
        j->ls_ent = nullptr;
        j->ls_grad = nullptr;
        j->cap_ls = 0;
        bool ok = hipMalloc(&j->ls_ent, nls_all * sizeof(int32_t)) == hipSuccess &&
                  hipMalloc(&j->ls_grad, nls_all * kGenes * sizeof(float)) == hipSuccess;
        if (!ok) return fail(j->ctx, ADGPU_ERR_NOMEM, "job allocation failed");
        j->cap_ls = nls_all;
    }
    if (yield && nls_all > j->cap_yield) {
        if (j->cap_yield) j->reallocs++;
        (void)hipFree(j->ls_state); (void)hipFree(j->ls_ring);
        j->ls_state = nullptr;
        j->ls_ring = nullptr;
        j->cap_yield = 0;
        bool ok = hipMalloc(&j->ls_state, nls_all * kAdStateFloats * sizeof(float)) == hipSuccess &&
                  hipMalloc(&j->ls_ring, 3 * nls_all * sizeof(int32_t)) == hipSuccess;
        if (!ok) return fail(j->ctx, ADGPU_ERR_NOMEM, "job allocation failed");
        j->cap_yield = nls_all;
    }
    return ADGPU_OK;
}

// Regrouping interval of the automatic setting, used when a job runs in four
// or more run groups or in a server of four or more slots: the launches
// overlap, so the lanes that regrouping frees in one launch's last searches
// run another launch's work.  Measured on the 30-atom ligand, ADADELTA nrun
// 100 x pop 150 (M evals/s, interval 8 / 32 / off): 4 groups 736 / 726 / 711;
// 2 groups 469 / 685 / 676; 1 group 247 / 524 / 575.  Served, 32 slots x 128
// requests of nrun 20 (k ligands/h, 4 repetitions): 112-115 / 109-114 /
// 102-108.  With fewer groups and no other requests the last searches of a
// launch set its length, and each hand-over attempt (a few device-scope
// memory round trips on one counter that every lone search uses) lengthens
// them.
constexpr int kRegroupAuto = 8;
// ... and with two or three run groups (2 groups, M evals/s, interval 32 /
// off, three dockings each: 680 / 686 / 690 against 670 / 678 / 681).
constexpr int kRegroupAutoFew = 32;

// True when the job's ADADELTA searches hand off between teams.
static bool yield_wanted(const adgpu_context* c, const adgpu_search_params* sp) {
    return sp->ls_method == ADGPU_LS_AD && c->ls_yield.load() > 0;
}

// Starts a search: uploads the initial population and PRNG seeds, clears
// energies, eval counters, work queues and the stop flag, and resets the
// generation state (the job's groups and buffers are already sized).
static int job_start(adgpu_job* j, const adgpu_search_params* sp, const LigandDev& lig, const float* init_pop,
                     const uint32_t* seeds) {
    adgpu_context* c = j->ctx;
    j->sp = *sp;
    j->sd = to_dev(sp, lig.num_genes);
    if (yield_wanted(c, sp) && j->cap_yield >= (size_t)sp->num_runs * std::max(1, j->sd.num_ls)) {
        j->sd.ls_yield = c->ls_yield.load();
        const int rg = c->ls_regroup.load();
        j->sd.ls_regroup = rg >= 0 ? rg
                         : j->groups.size() >= 4 || j->sharers >= 4 ? kRegroupAuto
                         : j->groups.size() >= 2 ? kRegroupAutoFew : 0;
    }
    j->team_hint = job_team_width(c, lig.num_atoms, (int64_t)sp->num_runs * std::max(1, j->sd.num_ls), j->sharers,
                                  sp->ls_method);
    j->team_lanes = team_lanes_for(lig.num_atoms, j->team_hint);
    const int nrun = sp->num_runs;
    const size_t ne = (size_t)nrun * sp->pop_size;
    j->graph_block = c->graph_block.load();
    for (auto& g : j->groups) {
        drop_blocks(g);
        g.parity = g.parity0 = 0;
        g.done = g.enq = 0;
        g.totals.clear();
    }
    j->generation = 0;
    j->stopped = false;
    j->as_enabled = false;
    j->origin_set = false;
    j->spans[0].clear();
    j->spans[1].clear();
    j->kernel_ms[0] = j->kernel_ms[1] = 0.0;
    j->kernel_n[0] = j->kernel_n[1] = 0;
    hipStream_t s = j->groups[0].stream;
    HIPCHK(c, hipMemcpyAsync(j->pop[0], init_pop, ne * kGenes * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(j->pop[1], init_pop, ne * kGenes * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(j->prng, seeds, ne * ADGPU_NWI * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(j->evals_runs, 0, nrun * sizeof(int32_t), s));
    HIPCHK(c, hipMemsetAsync(j->evals_new, 0, ne * sizeof(int32_t), s));
    HIPCHK(c, hipMemsetAsync(j->ene[0], 0, ne * sizeof(float), s));
    HIPCHK(c, hipMemsetAsync(j->ene[1], 0, ne * sizeof(float), s));
    HIPCHK(c, hipMemsetAsync(j->counters, 0, 16 * kMaxGroups * sizeof(int32_t), s));
    HIPCHK(c, hipMemsetAsync(j->d_stop, 0, 16 * sizeof(int32_t), s));
    HIPCHK(c, hipStreamSynchronize(s));
    j->h_stop[0] = 0;
    return ADGPU_OK;
}

int adgpu_job_create(adgpu_ctx c, int map_id, int lig_id, const adgpu_search_params* sp, const float* init_pop,
                     const uint32_t* seeds, adgpu_job_t* out) {
    if (!c || !sp || !init_pop || !seeds || !out) return fail(c, ADGPU_ERR_ARG, "null argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    auto li = c->ligands.find(lig_id);
    if (mi == c->maps.end() || li == c->ligands.end()) return fail(c, ADGPU_ERR_ARG, "unknown map or ligand id");
    if (li->second.max_map >= mi->second.dev.nmaps - 2)
        return fail(c, ADGPU_ERR_ARG, "ligand atom_map index exceeds the grid's type maps");
    int rc = check_search(c, sp, li->second.host.num_genes);
    if (rc) return rc;
    const int num_genes = li->second.host.num_genes;
    if (to_dev(sp, num_genes).num_ls > sp->pop_size) return fail(c, ADGPU_ERR_ARG, "lsearch_rate above 100 %");
    HIPCHK(c, hipSetDevice(c->device));
    std::unique_ptr<adgpu_job> j(new adgpu_job());
    j->ctx = c;
    j->map_id = map_id;
    j->ligand_id = lig_id;
    const int nrun = sp->num_runs;
    // run groups: ADGPU_RUN_GROUPS streams.  Groups whose streams share a
    // hardware queue serialize, so the default follows the HIP runtime's
    // hardware queues per process (GPU_MAX_HW_QUEUES, 4 unless set): 2 groups
    // (with the context's stream, 3 of 4 queues), or 4 groups when at least
    // 6 queues are available (the faster setting on MI355X).  A group's
    // local-search launch should still fill a good part of the device, so
    // small jobs get fewer groups: one per kGroupSearches searches of a
    // generation (measured: nrun 20 x 150 searches fastest in 2 groups,
    // nrun 50 x 150 in 4)
    int ng = 2;
    if (const char* q = std::getenv("GPU_MAX_HW_QUEUES"))
        if (std::atoi(q) >= 6) ng = 4;
    const int64_t searches = (int64_t)nrun * std::max(1, to_dev(sp, num_genes).num_ls);
    ng = (int)std::max<int64_t>(1, std::min<int64_t>(ng, searches / kGroupSearches));
    if (const char* e = std::getenv("ADGPU_RUN_GROUPS")) ng = std::atoi(e);
    rc = job_open(j.get());
    if (!rc) rc = job_set_groups(j.get(), ng, nrun);
    if (!rc)
        rc = job_reserve(j.get(), nrun, sp->pop_size, to_dev(sp, num_genes).num_ls, sp->ls_method == ADGPU_LS_FIRE,
                         yield_wanted(c, sp));
    if (!rc) rc = job_start(j.get(), sp, li->second.host, init_pop, seeds);
    if (rc) {
        const std::string msg = t_err;
        adgpu_job_destroy(j.release());
        return fail(c, rc, msg);
    }
    mi->second.jobs++;
    li->second.jobs++;
    j->counted = true;
    *out = j.release();
    return ADGPU_OK;
}

int adgpu_job_destroy(adgpu_job_t j) {
    if (!j) return ADGPU_ERR_ARG;
    if (j->ctx) (void)hipSetDevice(j->ctx->device);
    if (j->ctx && j->counted) {
        std::lock_guard<std::mutex> lk(j->ctx->mu);
        auto mi = j->ctx->maps.find(j->map_id);
        auto li = j->ctx->ligands.find(j->ligand_id);
        if (mi != j->ctx->maps.end()) mi->second.jobs--;
        if (!j->own_lig && li != j->ctx->ligands.end()) li->second.jobs--;
    }
    for (auto& g : j->groups) destroy_group(g);
    (void)hipFree(j->pop[0]); (void)hipFree(j->pop[1]);
    (void)hipFree(j->ene[0]); (void)hipFree(j->ene[1]);
    (void)hipFree(j->evals_new); (void)hipFree(j->evals_runs); (void)hipFree(j->prng);
    (void)hipFree(j->d_ring);
    (void)hipFree(j->counters); (void)hipFree(j->ls_ent); (void)hipFree(j->ls_grad);
    (void)hipFree(j->ls_state); (void)hipFree(j->ls_ring);
    (void)hipFree(j->d_stop); (void)hipFree(j->d_as); (void)hipFree(j->d_as_bins);
    (void)hipFree(j->d_as_log); (void)hipFree(j->d_as_out);
    (void)hipFree(j->own.dev); (void)hipFree(j->d_best); (void)hipFree(j->d_f64);
    if (j->ev_test) (void)hipEventDestroy(j->ev_test);
    if (j->ev_origin) (void)hipEventDestroy(j->ev_origin);
    if (j->h_stop) (void)hipHostFree(j->h_stop);
    if (j->h_ring) (void)hipHostFree(j->h_ring);
    delete j;
    return ADGPU_OK;
}

// The job's maps and ligand cannot be released while the job exists (their
// job counts), so the pointers stay valid for the job's lifetime.
static int job_lookup(adgpu_job_t j, const LigandSet** ls, const MapSet** ms) {
    adgpu_context* c = j->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(j->map_id);
    if (mi == c->maps.end()) return fail(c, ADGPU_ERR_STATE, "map set released");
    *ms = &mi->second;
    if (j->own_lig) {   // a serving slot's own ligand block
        if (!j->counted) return fail(c, ADGPU_ERR_STATE, "serving slot not bound to a request");
        *ls = &j->own;
        return ADGPU_OK;
    }
    auto li = c->ligands.find(j->ligand_id);
    if (li == c->ligands.end()) return fail(c, ADGPU_ERR_STATE, "ligand released");
    *ls = &li->second;
    return ADGPU_OK;
}

// Device buffers of runs [r0, r1) with the given current-population parity.
static JobBuffers slice(const adgpu_job* j, int r0, int r1, int parity, int counter) {
    const size_t pop = (size_t)j->sp.pop_size, off = (size_t)r0 * pop;
    JobBuffers b{};
    b.pop_cur = j->pop[parity] + off * kGenes;
    b.pop_next = j->pop[parity ^ 1] + off * kGenes;
    b.ene_cur = j->ene[parity] + off;
    b.ene_next = j->ene[parity ^ 1] + off;
    b.evals_new = j->evals_new + off;
    b.evals_runs = j->evals_runs + r0;
    b.prng = j->prng + off * ADGPU_NWI;
    b.ls_counter = j->counters + 16 * counter;
    const size_t nls = (size_t)(j->sd.num_ls > 0 ? j->sd.num_ls : 1);
    b.ls_ent = j->ls_ent ? j->ls_ent + (size_t)r0 * nls : nullptr;
    b.ls_grad = j->ls_grad ? j->ls_grad + (size_t)r0 * nls * kGenes : nullptr;
    b.ls_state = j->ls_state ? j->ls_state + (size_t)r0 * nls * kAdStateFloats : nullptr;
    b.ls_ring = j->ls_ring ? j->ls_ring + 3 * (size_t)r0 * nls : nullptr;   // ring, then seats
    b.stop = j->d_stop;
    (void)r1;
    return b;
}

static SearchDev slice_sd(const adgpu_job* j, int r0, int r1) {
    SearchDev s = j->sd;
    s.num_runs = r1 - r0;
    return s;
}

static int ls_method_for(adgpu_job_t j, int64_t generation) {
    // performdocking.cpp.Cuda:727-731: AD falls back to SW for the first
    // initial_sw_generations generations
    int m = j->sp.ls_method;
    if (j->sd.num_ls == 0 || j->sp.lsearch_rate == 0.0f) return ADGPU_LS_NONE;
    if (m == ADGPU_LS_AD && generation < j->sp.initial_sw_generations) m = ADGPU_LS_SW;
    return m;
}

// Whole-job steps on the first group's stream (all groups share one parity
// outside adgpu_job_run).
static int whole_parity(adgpu_job_t j) { return j->groups[0].parity; }
static void flip_all(adgpu_job_t j) { for (auto& g : j->groups) g.parity ^= 1; }

int adgpu_job_initpop(adgpu_job_t j) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    const LigandSet* ls; const MapSet* ms;
    int rc = job_lookup(j, &ls, &ms);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = j->groups[0].stream;
    const JobBuffers b = slice(j, 0, j->sp.num_runs, whole_parity(j), 0);
    HIPCHK(c, launch_initpop(launch_args(c, *ls, *ms, s, 1, j->team_hint), j->sd, b));
    HIPCHK(c, launch_sum_evals(s, j->sd, b));
    HIPCHK(c, hipStreamSynchronize(s));
    return ADGPU_OK;
}

int adgpu_job_kernel_gen(adgpu_job_t j) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    const LigandSet* ls; const MapSet* ms;
    int rc = job_lookup(j, &ls, &ms);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = j->groups[0].stream;
    HIPCHK(c, launch_gen(launch_args(c, *ls, *ms, s, 1, j->team_hint), j->sd,
                         slice(j, 0, j->sp.num_runs, whole_parity(j), 0)));
    HIPCHK(c, hipStreamSynchronize(s));
    return ADGPU_OK;
}

// The ADADELTA hand-off's fault flags (one per group): a team that polled a
// reserved ring entry past its bound stopped and left that search
// unfinished.  Read once the job's kernels have completed.
static int check_handoff(adgpu_job_t j) {
    if (j->sd.ls_yield <= 0) return ADGPU_OK;
    int32_t q[16 * kMaxGroups];
    HIPCHK(j->ctx, hipMemcpy(q, j->counters, sizeof(q), hipMemcpyDeviceToHost));
    for (int k = 0; k < kMaxGroups; ++k) {
        if (q[16 * k + 3] == 2)
            return fail(j->ctx, ADGPU_ERR_HIP, "ADADELTA hand-off: a saved search record was invalid when resumed");
        if (q[16 * k + 3]) return fail(j->ctx, ADGPU_ERR_HIP, "ADADELTA hand-off between teams did not complete");
    }
    return ADGPU_OK;
}

int adgpu_set_graph_block(adgpu_ctx c, int generations) {
    if (!c) return ADGPU_ERR_ARG;
    if (generations != 0 && generations != 2 && generations != 4 && generations != 8)
        return fail(c, ADGPU_ERR_ARG, "graph block must be 0, 2, 4 or 8 generations");
    c->graph_block = generations;
    return ADGPU_OK;
}

int adgpu_measure_valu_rate(adgpu_ctx c, double* wave_instr_per_s) {
    if (!c || !wave_instr_per_s) return fail(c, ADGPU_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipSetDevice(c->device));
    int cus = 0;
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const int blocks = std::max(1, cus) * 8;   // 32 wavefronts per compute unit, 8 per SIMD
    float* buf = nullptr;
    if (hipMalloc(&buf, (size_t)blocks * 256 * sizeof(float)) != hipSuccess)
        return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(calibration) failed");
    const hipError_t e = measure_valu_rate(c->stream, buf, blocks, 20000, wave_instr_per_s);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(c, ADGPU_ERR_HIP, std::string("VALU rate: ") + hipGetErrorString(e));
    return ADGPU_OK;
}

int adgpu_job_ls_stats(adgpu_job_t j, int64_t* stats) {
    if (!j || !stats) return ADGPU_ERR_ARG;
    HIPCHK(j->ctx, hipSetDevice(j->ctx->device));
    for (auto& g : j->groups) HIPCHK(j->ctx, hipStreamSynchronize(g.stream));
    int32_t q[16 * kMaxGroups];
    HIPCHK(j->ctx, hipMemcpy(q, j->counters, sizeof(q), hipMemcpyDeviceToHost));
    int64_t h = 0, r = 0, w = 0;
    for (int k = 0; k < kMaxGroups; ++k) {
        h += q[16 * k + 4];
        r += q[16 * k + 7];
        w += (uint32_t)q[16 * k + 8];
    }
    stats[0] = h; stats[1] = r; stats[2] = w; stats[3] = j->team_lanes;
    return ADGPU_OK;
}

int adgpu_job_ls_handoffs(adgpu_job_t j, int64_t* count) {
    if (!j || !count) return ADGPU_ERR_ARG;
    HIPCHK(j->ctx, hipSetDevice(j->ctx->device));
    for (auto& g : j->groups) HIPCHK(j->ctx, hipStreamSynchronize(g.stream));
    int32_t q[16 * kMaxGroups];
    HIPCHK(j->ctx, hipMemcpy(q, j->counters, sizeof(q), hipMemcpyDeviceToHost));
    int64_t n = 0;
    for (int k = 0; k < kMaxGroups; ++k) n += q[16 * k + 4];
    *count = n;
    return ADGPU_OK;
}

int adgpu_job_kernel_ls(adgpu_job_t j, int generation) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    const LigandSet* ls; const MapSet* ms;
    int rc = job_lookup(j, &ls, &ms);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int m = ls_method_for(j, generation);
    if (m == ADGPU_LS_NONE) return ADGPU_OK;
    hipStream_t s = j->groups[0].stream;
    HIPCHK(c, launch_ls(launch_args(c, *ls, *ms, s, 1, j->team_hint), j->sd,
                        slice(j, 0, j->sp.num_runs, whole_parity(j), 0), m));
    HIPCHK(c, hipStreamSynchronize(s));
    return check_handoff(j);
}

int adgpu_job_kernel_sum_and_swap(adgpu_job_t j) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = j->groups[0].stream;
    HIPCHK(c, launch_sum_evals(s, j->sd, slice(j, 0, j->sp.num_runs, whole_parity(j), 0)));
    HIPCHK(c, hipStreamSynchronize(s));
    flip_all(j);
    j->generation++;
    return ADGPU_OK;
}

int adgpu_job_generation(adgpu_job_t j) {
    if (!j) return ADGPU_ERR_ARG;
    int rc = adgpu_job_kernel_gen(j);
    if (!rc) rc = adgpu_job_kernel_ls(j, (int)j->generation);
    if (!rc) rc = adgpu_job_kernel_sum_and_swap(j);
    return rc;
}

// Enqueues one generation of group k (K4, local search, K2, eval totals to
// the pinned ring) on its stream; performdocking.cpp.Cuda:700-831.
// captured: the generation goes into a graph block (enqueue_block), which
// holds only the kernels; K2 then leaves the totals in the device ring, and
// the block's launch copies them and records the completion events on the
// stream after the graph.  Launches inside blocks are not timed.
static int enqueue_generation(adgpu_job_t j, int k, const LigandSet* ls, const MapSet* ms, bool captured = false) {
    adgpu_context* c = j->ctx;
    RunGroup& g = j->groups[k];
    const int slot = (int)(g.enq % kRing);
    const LaunchArgs la = launch_args(c, *ls, *ms, g.stream, (int)j->groups.size(), j->team_hint);
    const SearchDev sd = slice_sd(j, g.r0, g.r1);
    JobBuffers b = slice(j, g.r0, g.r1, g.parity, k);
    const int m = ls_method_for(j, j->generation + g.enq);
    if (captured) b.evals_copy = j->d_ring + (size_t)slot * j->sp.num_runs + g.r0;
    if (!captured) HIPCHK(c, hipEventRecord(g.ev_gen[0][slot], g.stream));
    HIPCHK(c, launch_gen(la, sd, b));
    if (!captured) HIPCHK(c, hipEventRecord(g.ev_gen[1][slot], g.stream));
    g.ls_used[slot] = m != ADGPU_LS_NONE && !captured;   // timed launches only
    g.gen_timed[slot] = !captured;
    if (m != ADGPU_LS_NONE) {
        if (!captured) HIPCHK(c, hipEventRecord(g.ev_ls[0][slot], g.stream));
        HIPCHK(c, launch_ls(la, sd, b, m));
        if (!captured) HIPCHK(c, hipEventRecord(g.ev_ls[1][slot], g.stream));
    }
    HIPCHK(c, launch_sum_evals(g.stream, sd, b));
    if (!captured) {
        HIPCHK(c, hipMemcpyAsync(j->h_ring + (size_t)slot * j->sp.num_runs + g.r0, j->evals_runs + g.r0,
                                 (size_t)(g.r1 - g.r0) * sizeof(int32_t), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(c, hipEventRecord(g.ev_done[slot], g.stream));
    }
    g.parity ^= 1;
    g.enq++;
    return ADGPU_OK;
}

// Enqueues the next B generations of group k (B divides kRing; the first is
// at a ring slot that is a multiple of B) as one graph launch.  The block is
// captured from enqueue_generation the first time its first slot comes round
// with this parity and local search, and replayed afterwards: its kernels are
// those of its slots (K2 leaves each generation's totals in the slot's row of
// the device ring), so a replay runs exactly the kernels of the B calls of
// enqueue_generation.  After the graph, on the same stream, one copy brings
// the B rows of totals to the host and the B completion events are recorded:
// the host learns of each generation in stream order, as without the graph.
// Generations a group must have in flight before it captures a new block.
constexpr int kCaptureDepth = 2;

// True when group k's ring slot for its next generation holds a captured
// block of B generations that matches them (parity and local search).
static bool block_cached(adgpu_job_t j, int k, int B) {
    const RunGroup& g = j->groups[k];
    const int slot0 = (int)(g.enq % kRing);
    return g.blk[slot0] && g.blk_parity[slot0] == g.parity &&
           g.blk_method[slot0] == ls_method_for(j, j->generation + g.enq) && g.blk_len[slot0] == B;
}

static int enqueue_block(adgpu_job_t j, int k, const LigandSet* ls, const MapSet* ms, int B) {
    adgpu_context* c = j->ctx;
    RunGroup& g = j->groups[k];
    const int slot0 = (int)(g.enq % kRing);
    const int m = ls_method_for(j, j->generation + g.enq);
    if (g.blk[slot0] && (g.blk_parity[slot0] != g.parity || g.blk_method[slot0] != m || g.blk_len[slot0] != B)) {
        (void)hipGraphExecDestroy(g.blk[slot0]);
        g.blk[slot0] = nullptr;
    }
    if (!g.blk[slot0]) {
        const int parity = g.parity;
        const int64_t enq = g.enq;
        HIPCHK(c, hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
        int rc = ADGPU_OK;
        for (int i = 0; i < B && rc == ADGPU_OK; ++i) rc = enqueue_generation(j, k, ls, ms, true);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(g.stream, &graph);
        g.parity = parity;   // the launch below runs the captured generations
        g.enq = enq;
        if (rc != ADGPU_OK || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc != ADGPU_OK ? rc : fail(c, ADGPU_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                                   hipGetErrorString(ec));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(c, ADGPU_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        g.blk[slot0] = exec;
        g.blk_parity[slot0] = parity;
        g.blk_method[slot0] = m;
        g.blk_len[slot0] = B;
    }
    HIPCHK(c, hipGraphLaunch(g.blk[slot0], g.stream));
    const size_t nrun = (size_t)j->sp.num_runs, w = (size_t)(g.r1 - g.r0) * sizeof(int32_t);
    HIPCHK(c, hipMemcpy2DAsync(j->h_ring + (size_t)slot0 * nrun + g.r0, nrun * sizeof(int32_t),
                               j->d_ring + (size_t)slot0 * nrun + g.r0, nrun * sizeof(int32_t), w, (size_t)B,
                               hipMemcpyDeviceToHost, g.stream));
    for (int i = 0; i < B; ++i) {
        const int slot = (slot0 + i) % kRing;
        HIPCHK(c, hipEventRecord(g.ev_done[slot], g.stream));
        g.ls_used[slot] = false;   // not timed
        g.gen_timed[slot] = false;
        g.parity ^= 1;
        g.enq++;
    }
    return ADGPU_OK;
}

// Milliseconds between two recorded events; false when they cannot be timed
// (the failure is not left behind for the next launch's error check).
static bool elapsed_ms(hipEvent_t a, hipEvent_t b, float* ms) {
    if (hipEventElapsedTime(ms, a, b) == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
}

// Waits for the oldest in-flight generation of group k and records it.
static int complete_generation(adgpu_job_t j, int k) {
    adgpu_context* c = j->ctx;
    RunGroup& g = j->groups[k];
    const int slot = (int)(g.done % kRing);
    HIPCHK(c, hipEventSynchronize(g.ev_done[slot]));
    int64_t t = 0;
    for (int r = g.r0; r < g.r1; ++r) t += j->h_ring[(size_t)slot * j->sp.num_runs + r];
    g.totals.push_back(t);
    float ms = 0.f, a = 0.f, b = 0.f;
    if (g.gen_timed[slot] && elapsed_ms(g.ev_gen[0][slot], g.ev_gen[1][slot], &ms)) {
        j->kernel_ms[0] += ms; j->kernel_n[0]++;
        if (elapsed_ms(j->ev_origin, g.ev_gen[0][slot], &a) && elapsed_ms(j->ev_origin, g.ev_gen[1][slot], &b))
            j->spans[0].emplace_back(a, b);
    }
    if (g.ls_used[slot] && elapsed_ms(g.ev_ls[0][slot], g.ev_ls[1][slot], &ms)) {
        j->kernel_ms[1] += ms; j->kernel_n[1]++;
        if (elapsed_ms(j->ev_origin, g.ev_ls[0][slot], &a) && elapsed_ms(j->ev_origin, g.ev_ls[1][slot], &b))
            j->spans[1].emplace_back(a, b);
    }
    g.done++;
    return ADGPU_OK;
}

// check_progress (host/src/performdocking.cpp.Cuda:908-942)
static double progress_of(const adgpu_job_t j, int64_t total, int64_t generation) {
    double ep = (double)total / (double)j->sp.num_runs / (double)j->sp.max_evals * 100.0;
    double gp = (double)generation / (double)j->sp.max_generations * 100.0;
    return ep > gp ? ep : gp;
}

// The generation loop of docking_with_gpu (performdocking.cpp.Cuda:685-831),
// stopped by check_progress (:908-942) exactly as the reference: generation
// c+1 runs iff progress after c generations is below 100 %.  Groups of runs
// advance on their own streams; a group enqueues generation c+1 only when an
// upper bound of the total evals after generation c (groups that have not
// reported c yet are charged the most evals a generation can take) already
// proves that c is not the last generation.  No group ever runs past the
// final generation, so results equal the one-stream loop.
int adgpu_job_run(adgpu_job_t j, int64_t* out_gens, int64_t* out_total) {
    return adgpu_job_run_generations(j, INT64_MAX, out_gens, out_total);
}

// Enqueues the AutoStop test after cg generations of this call on the first
// group's stream, behind every group's generation cg-1; the groups' later
// generations wait for it (performdocking.cpp.Cuda:687-696).
static int enqueue_test(adgpu_job_t j, int64_t cg) {
    adgpu_context* c = j->ctx;
    hipStream_t s0 = j->groups[0].stream;
    if (cg > 0)
        for (size_t k = 1; k < j->groups.size(); ++k)
            HIPCHK(c, hipStreamWaitEvent(s0, j->groups[k].ev_done[(cg - 1) % kRing], 0));
    const float* ene = j->ene[j->groups[0].parity0 ^ (int)(cg & 1)];
    HIPCHK(c, launch_autostop(s0, j->as_p, ene, j->evals_runs, j->d_as, j->d_as_bins, j->d_as_log, j->d_stop,
                              j->generation + cg, 0, nullptr));
    HIPCHK(c, hipMemcpyAsync(j->h_stop, j->d_stop, sizeof(int32_t), hipMemcpyDeviceToHost, s0));
    HIPCHK(c, hipEventRecord(j->ev_test, s0));
    for (size_t k = 1; k < j->groups.size(); ++k) HIPCHK(c, hipStreamWaitEvent(j->groups[k].stream, j->ev_test, 0));
    return ADGPU_OK;
}

// Generation count of the test that raised the stop flag (the last test run).
static int stopped_generation(adgpu_job_t j, int64_t* gen) {
    adgpu_context* c = j->ctx;
    AutoStopDev st{};
    HIPCHK(c, hipStreamSynchronize(j->groups[0].stream));
    HIPCHK(c, hipMemcpy(&st, j->d_as, sizeof(st), hipMemcpyDeviceToHost));
    if (st.num_tests < 1) return fail(c, ADGPU_ERR_STATE, "stop flag without an AutoStop test");
    adgpu_autostop_record r{};
    HIPCHK(c, hipMemcpy(&r, j->d_as_log + (st.num_tests - 1) % j->as_p.log_cap, sizeof(r), hipMemcpyDeviceToHost));
    *gen = r.generation;
    return ADGPU_OK;
}

// Most evals one run can add in one generation: K4 offspring plus every LS
// slot at its iteration cap: two evaluations per Solis-Wets iteration, one per
// ADADELTA, ADAM or FIRE iteration (ADADELTA may run Solis-Wets first).
static int64_t max_evals_per_generation(const adgpu_job* j) {
    const bool sw = j->sp.ls_method == ADGPU_LS_SW ||
                    (j->sp.ls_method == ADGPU_LS_AD && j->sp.initial_sw_generations > 0);
    return (int64_t)(j->sp.pop_size - 1) +
           (int64_t)j->sd.num_ls * (sw ? 2 : 1) * std::max<int64_t>(1, j->sp.max_num_of_iters);
}

int adgpu_job_run_generations(adgpu_job_t j, int64_t max_new, int64_t* out_gens, int64_t* out_total) {
    if (!j || max_new < 0) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    const LigandSet* ls; const MapSet* ms;
    int rc = job_lookup(j, &ls, &ms);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int nrun = j->sp.num_runs;
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    HIPCHK(c, hipMemcpy(j->h_ring, j->evals_runs, nrun * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (auto& g : j->groups) {
        int64_t t = 0;
        for (int r = g.r0; r < g.r1; ++r) t += j->h_ring[r];
        g.totals.assign(1, t);
        g.done = g.enq = 0;
        g.parity0 = g.parity;
    }
    if (j->stopped) max_new = 0;   // a passed AutoStop test ended the job
    if (!j->origin_set) {          // time origin of the launch intervals
        HIPCHK(c, hipEventRecord(j->ev_origin, j->groups[0].stream));
        HIPCHK(c, hipEventSynchronize(j->ev_origin));
        j->origin_set = true;
    }
    // most evals one run can add in one generation: K4 offspring plus every LS
    // slot at its iteration cap (two evaluations per Solis-Wets iteration)
    const int64_t emax = max_evals_per_generation(j);
    auto bound = [&](int64_t cg) {   // upper bound of the total after cg generations
        int64_t t = 0;
        for (auto& g : j->groups) {
            if (cg <= g.done) t += g.totals[cg];
            else t += g.totals[g.done] + (cg - g.done) * (int64_t)(g.r1 - g.r0) * emax;
        }
        return t;
    };
    auto surely_continues = [&](int64_t cg) {
        return cg < max_new && progress_of(j, bound(cg), j->generation + cg) < 100.0;
    };
    // AutoStop: the test after cg generations (generation count j->generation
    // + cg, a multiple of as_frequency) precedes generation cg + 1 of every
    // group; next_test is the first test point not enqueued yet
    const int64_t asf = j->as_enabled ? j->as_p.as_frequency : 0;
    int64_t next_test = asf ? (asf - j->generation % asf) % asf : INT64_MAX;
    bool test_pending = false, stop_known = false;
    int64_t stop_cg = -1;
    auto poll_test = [&]() -> int {
        if (!test_pending) return ADGPU_OK;
        const hipError_t q = hipEventQuery(j->ev_test);
        if (q == hipErrorNotReady) return ADGPU_OK;
        HIPCHK(c, q);
        test_pending = false;
        if (j->h_stop[0] && !stop_known) {
            int64_t gen = 0;
            const int r = stopped_generation(j, &gen);
            if (r) return r;
            stop_known = true;
            stop_cg = gen - j->generation;
        }
        return ADGPU_OK;
    };
    for (;;) {
        for (bool moved = true; moved && !stop_known;) {
            moved = false;
            bool all_at_test = asf != 0;
            for (auto& g : j->groups) all_at_test = all_at_test && g.enq >= next_test;
            if (all_at_test && surely_continues(next_test)) {
                if ((rc = enqueue_test(j, next_test))) return rc;
                test_pending = true;
                next_test += asf;
                moved = true;
            }
            for (int k = 0; k < (int)j->groups.size(); ++k) {
                RunGroup& g = j->groups[k];
                while (g.enq - g.done < kRing && g.enq < next_test && surely_continues(g.enq)) {
                    // a whole captured block when every generation of it surely runs
                    // with one local search
                    const int B = j->graph_block;
                    bool block = B > 0 && g.enq % B == 0 && g.enq - g.done + B <= kRing && g.enq + B <= next_test &&
                                 surely_continues(g.enq + B - 1);
                    for (int i = 1; block && i < B; ++i)
                        block = ls_method_for(j, j->generation + g.enq + i) == ls_method_for(j, j->generation + g.enq);
                    // capturing and instantiating a block takes the host a few milliseconds:
                    // only while the group has generations in flight to hide them (a
                    // block already captured for this ring slot launches at once)
                    if (block && !block_cached(j, k, B) && g.enq - g.done < kCaptureDepth) block = false;
                    rc = block ? enqueue_block(j, k, ls, ms, B) : enqueue_generation(j, k, ls, ms);
                    if (rc) return rc;
                    moved = true;
                }
            }
        }
        if ((rc = poll_test())) return rc;
        int wait_k = -1;
        for (int k = 0; k < (int)j->groups.size(); ++k) {
            const RunGroup& g = j->groups[k];
            if (g.enq > g.done && (wait_k < 0 || g.done < j->groups[wait_k].done)) wait_k = k;
        }
        if (wait_k < 0) {
            if (!test_pending) break;   // nothing in flight and nothing may be enqueued
            HIPCHK(c, hipEventSynchronize(j->ev_test));
            continue;
        }
        if ((rc = complete_generation(j, wait_k))) return rc;
    }
    int64_t gens = j->groups[0].done;
    if (stop_known) {
        // generations enqueued behind the passed test did nothing
        gens = stop_cg;
        for (auto& g : j->groups) {
            if (g.done < gens) return fail(c, ADGPU_ERR_STATE, "run group behind the AutoStop test");
            g.parity = g.parity0 ^ (int)(gens & 1);
        }
        j->stopped = true;
    }
    if ((rc = check_handoff(j))) return rc;
    int64_t total = 0;
    for (auto& g : j->groups) {
        if (!stop_known && g.done != gens) return fail(c, ADGPU_ERR_STATE, "run groups stopped at different generations");
        total += g.totals[gens];
    }
    j->generation += gens;
    if (out_gens) *out_gens = j->generation;
    if (out_total) *out_total = total;
    return ADGPU_OK;
}

int adgpu_job_set_autostop(adgpu_job_t j, int enable, float stopstd, int as_frequency, int64_t min_as_evals) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    if (enable && (as_frequency < 1 || !(stopstd > 0.0f) || min_as_evals < 0))
        return fail(c, ADGPU_ERR_ARG, "AutoStop needs as_frequency >= 1, stopstd > 0 and min_as_evals >= 0");
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    const int pop = j->sp.pop_size;
    AutoStopParams p{};
    p.num_runs = j->sp.num_runs;
    p.pop_size = pop;
    p.ncream = pop / 10;
    p.as_frequency = as_frequency;
    p.stopstd = stopstd;
    p.min_as_evals = min_as_evals;
    // one record per test of the longest allowed search, at most 16384 kept
    const int64_t tests = as_frequency > 0 ? j->sp.max_generations / as_frequency + 2 : 2;
    p.log_cap = (int32_t)std::max<int64_t>(16, std::min<int64_t>(tests, 16384));
    const size_t nb = 3 * ((size_t)pop + 1);
    if (!j->d_as) {
        if (hipMalloc(&j->d_as, sizeof(AutoStopDev)) != hipSuccess ||
            hipMalloc(&j->d_as_out, sizeof(adgpu_autostop_record)) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "AutoStop allocation failed");
    }
    if (nb > j->as_bins_cap) {   // kept when large enough (a serving slot's next request)
        (void)hipFree(j->d_as_bins);
        j->d_as_bins = nullptr;
        j->as_bins_cap = 0;
        if (hipMalloc(&j->d_as_bins, nb * sizeof(float)) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "AutoStop allocation failed");
        j->as_bins_cap = nb;
    }
    if (p.log_cap > j->as_log_alloc) {
        (void)hipFree(j->d_as_log);
        j->d_as_log = nullptr;
        j->as_log_alloc = 0;
        if (hipMalloc(&j->d_as_log, (size_t)p.log_cap * sizeof(adgpu_autostop_record)) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "AutoStop allocation failed");
        j->as_log_alloc = p.log_cap;
    }
    // the AutoStop constructor, autostop.hpp:129-146
    AutoStopDev st{};
    st.first_time = 1;
    st.autostopped = 0;
    st.threshold = (float)(1 << 24);
    st.threshold_used = st.threshold;
    st.thres_stddev = st.threshold;
    st.curr_avg = (float)(-(1 << 24));
    st.curr_std = st.thres_stddev;
    st.roll_count = 0;
    st.best_n = 1;
    st.delta_energy = (float)(2.0 * (double)st.thres_stddev / (double)(uint32_t)pop);
    st.overall_best = (float)(1 << 24);
    st.num_tests = 0;
    HIPCHK(c, hipMemcpy(j->d_as, &st, sizeof(st), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(j->d_as_bins, 0, nb * sizeof(float)));
    HIPCHK(c, hipMemset(j->d_as_log, 0, (size_t)p.log_cap * sizeof(adgpu_autostop_record)));
    HIPCHK(c, hipMemset(j->d_stop, 0, 16 * sizeof(int32_t)));
    j->h_stop[0] = 0;
    j->stopped = false;
    j->as_p = p;
    j->as_enabled = enable != 0;
    return ADGPU_OK;
}

int adgpu_job_autostop_check(adgpu_job_t j, adgpu_autostop_record* out) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    if (!j->d_as) return fail(c, ADGPU_ERR_STATE, "AutoStop not set up (adgpu_job_set_autostop)");
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    hipStream_t s0 = j->groups[0].stream;
    HIPCHK(c, launch_autostop(s0, j->as_p, j->ene[whole_parity(j)], j->evals_runs, j->d_as, j->d_as_bins,
                              j->d_as_log, j->d_stop, j->generation, 0, nullptr));
    HIPCHK(c, hipStreamSynchronize(s0));
    AutoStopDev st{};
    HIPCHK(c, hipMemcpy(&st, j->d_as, sizeof(st), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(j->h_stop, j->d_stop, sizeof(int32_t), hipMemcpyDeviceToHost));
    j->stopped = j->h_stop[0] != 0;
    if (out) {
        if (st.num_tests < 1) return fail(c, ADGPU_ERR_STATE, "no AutoStop test recorded");
        HIPCHK(c, hipMemcpy(out, j->d_as_log + (st.num_tests - 1) % j->as_p.log_cap, sizeof(*out),
                            hipMemcpyDeviceToHost));
    }
    return ADGPU_OK;
}

int adgpu_job_autostop_log(adgpu_job_t j, int first, int max, adgpu_autostop_record* out, int* out_count) {
    if (!j || first < 0 || max < 0 || (max > 0 && !out)) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    if (!j->d_as) {
        if (out_count) *out_count = 0;
        return max > 0 ? fail(c, ADGPU_ERR_STATE, "AutoStop not set up (adgpu_job_set_autostop)") : ADGPU_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    AutoStopDev st{};
    HIPCHK(c, hipMemcpy(&st, j->d_as, sizeof(st), hipMemcpyDeviceToHost));
    const int n = st.num_tests, cap = j->as_p.log_cap;
    if (out_count) *out_count = n;
    const int last = std::min(n, first + max);
    if (first < last && first < n - cap)
        return fail(c, ADGPU_ERR_ARG, "AutoStop records before the last log_cap tests are no longer kept");
    for (int i = first; i < last; ++i)
        HIPCHK(c, hipMemcpy(out + (i - first), j->d_as_log + i % cap, sizeof(*out), hipMemcpyDeviceToHost));
    return ADGPU_OK;
}

int adgpu_job_autostop_final(adgpu_job_t j, adgpu_autostop_record* out) {
    if (!j || !out) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    if (!j->d_as) return fail(c, ADGPU_ERR_STATE, "AutoStop not set up (adgpu_job_set_autostop)");
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    hipStream_t s0 = j->groups[0].stream;
    HIPCHK(c, launch_autostop(s0, j->as_p, j->ene[whole_parity(j)], j->evals_runs, j->d_as, j->d_as_bins,
                              j->d_as_log, j->d_stop, j->generation, 1, j->d_as_out));
    HIPCHK(c, hipStreamSynchronize(s0));
    HIPCHK(c, hipMemcpy(out, j->d_as_out, sizeof(*out), hipMemcpyDeviceToHost));
    return ADGPU_OK;
}

int adgpu_job_download(adgpu_job_t j, float* pop, float* ene, int32_t* evals_runs, uint32_t* prng) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    const size_t pz = (size_t)j->sp.pop_size;
    for (auto& g : j->groups) {   // each group's current buffer
        const size_t off = (size_t)g.r0 * pz, n = (size_t)(g.r1 - g.r0) * pz;
        if (pop) HIPCHK(c, hipMemcpy(pop + off * kGenes, j->pop[g.parity] + off * kGenes, n * kGenes * sizeof(float),
                                     hipMemcpyDeviceToHost));
        if (ene) HIPCHK(c, hipMemcpy(ene + off, j->ene[g.parity] + off, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    const size_t ne = (size_t)j->sp.num_runs * pz;
    if (evals_runs) HIPCHK(c, hipMemcpy(evals_runs, j->evals_runs, j->sp.num_runs * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (prng) HIPCHK(c, hipMemcpy(prng, j->prng, ne * ADGPU_NWI * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return ADGPU_OK;
}

int adgpu_job_upload(adgpu_job_t j, const float* pop, const float* ene, const uint32_t* prng) {
    if (!j) return ADGPU_ERR_ARG;
    adgpu_context* c = j->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : j->groups) HIPCHK(c, hipStreamSynchronize(g.stream));
    const size_t pz = (size_t)j->sp.pop_size;
    for (auto& g : j->groups) {
        const size_t off = (size_t)g.r0 * pz, n = (size_t)(g.r1 - g.r0) * pz;
        if (pop) HIPCHK(c, hipMemcpy(j->pop[g.parity] + off * kGenes, pop + off * kGenes, n * kGenes * sizeof(float),
                                     hipMemcpyHostToDevice));
        if (ene) HIPCHK(c, hipMemcpy(j->ene[g.parity] + off, ene + off, n * sizeof(float), hipMemcpyHostToDevice));
    }
    const size_t ne = (size_t)j->sp.num_runs * pz;
    if (prng) HIPCHK(c, hipMemcpy(j->prng, prng, ne * ADGPU_NWI * sizeof(uint32_t), hipMemcpyHostToDevice));
    return ADGPU_OK;
}

int adgpu_job_best_f64(adgpu_job_t j, float* best_geno, float* best_e, int32_t* best_idx, double* total,
                       double* inter, double* intra) {
    if (!j) return ADGPU_ERR_ARG;
    const int nrun = j->sp.num_runs, pz = j->sp.pop_size;
    std::vector<float> pop((size_t)nrun * pz * kGenes), ene((size_t)nrun * pz);
    int rc = adgpu_job_download(j, pop.data(), ene.data(), nullptr, nullptr);
    if (rc) return rc;
    std::vector<float> geno((size_t)nrun * kGenes, 0.0f);
    for (int r = 0; r < nrun; ++r) {
        int bi = 0;
        for (int k = 1; k < pz; ++k)
            if (ene[(size_t)r * pz + k] < ene[(size_t)r * pz + bi]) bi = k;
        std::memcpy(&geno[(size_t)r * kGenes], &pop[((size_t)r * pz + bi) * kGenes], kGenes * sizeof(float));
        if (best_idx) best_idx[r] = bi;
        if (best_e) best_e[r] = ene[(size_t)r * pz + bi];
    }
    if (best_geno) std::memcpy(best_geno, geno.data(), geno.size() * sizeof(float));
    if (!total && !inter && !intra) return ADGPU_OK;
    return adgpu_rescore_f64(j->ctx, j->map_id, j->ligand_id, geno.data(), nrun, total, inter, intra);
}

int adgpu_job_kernel_busy(adgpu_job_t j, int cls, double* busy_ms, int64_t* launches) {
    if (!j || cls < 0 || cls > 1) return ADGPU_ERR_ARG;
    std::vector<std::pair<double, double>> v = j->spans[cls];
    std::sort(v.begin(), v.end());
    double busy = 0.0, s0 = 0.0, e0 = -1.0;
    for (const auto& iv : v) {
        if (iv.first > e0) {
            if (e0 > s0) busy += e0 - s0;
            s0 = iv.first;
            e0 = iv.second;
        } else if (iv.second > e0) {
            e0 = iv.second;
        }
    }
    if (e0 > s0) busy += e0 - s0;
    if (busy_ms) *busy_ms = busy;
    if (launches) *launches = (int64_t)v.size();
    return ADGPU_OK;
}

int adgpu_job_kernel_time(adgpu_job_t j, int cls, double* avg_ms, int64_t* launches) {
    if (!j || cls < 0 || cls > 1) return ADGPU_ERR_ARG;
    if (avg_ms) *avg_ms = j->kernel_n[cls] ? j->kernel_ms[cls] / (double)j->kernel_n[cls] : 0.0;
    if (launches) *launches = j->kernel_n[cls];
    return ADGPU_OK;
}

}  // extern "C"

// ---- serving slots (adgpu_slots.h) ------------------------------------------

namespace adgpu {

int set_error(int code, const std::string& msg) { return fail(nullptr, code, msg); }
const char* thread_error() { return t_err.c_str(); }
int context_device(const adgpu_context* c) { return c->device; }
int64_t slot_reallocations(const adgpu_job* j) { return j->reallocs; }

int prepare_ligand(adgpu_context* c, const adgpu_ligand* L, const adgpu_energy_params* ep, PreparedLigand& out) {
    if (!L || !ep) return fail(c, ADGPU_ERR_ARG, "null argument");
    int rc = build_ligand(c, L, ep, out.blob, out.hd);
    if (rc) return rc;
    out.max_map = 0;
    for (int a = 0; a < L->num_atoms; ++a) {
        if (L->ignore_inter && L->ignore_inter[a]) continue;
        if (L->atom_map[a] < 0) return fail(c, ADGPU_ERR_ARG, "negative atom_map index");
        out.max_map = std::max(out.max_map, (int)L->atom_map[a]);
    }
    return ADGPU_OK;
}

int check_request(adgpu_context* c, int map_id, const PreparedLigand& lig, const adgpu_search_params* sp) {
    if (!sp) return fail(c, ADGPU_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    if (mi == c->maps.end()) return fail(c, ADGPU_ERR_ARG, "unknown map id");
    if (lig.max_map >= mi->second.dev.nmaps - 2)
        return fail(c, ADGPU_ERR_ARG, "ligand atom_map index exceeds the grid's type maps");
    int rc = check_search(c, sp, lig.hd.num_genes);
    if (rc) return rc;
    if (to_dev(sp, lig.hd.num_genes).num_ls > sp->pop_size) return fail(c, ADGPU_ERR_ARG, "lsearch_rate above 100 %");
    return ADGPU_OK;
}

int map_hold(adgpu_context* c, int map_id, int delta) {
    std::lock_guard<std::mutex> lk(c->mu);
    auto mi = c->maps.find(map_id);
    if (mi == c->maps.end()) return fail(c, ADGPU_ERR_ARG, "unknown map id");
    mi->second.jobs += delta;
    return ADGPU_OK;
}

// Capacities of a slot's size-dependent device buffers: what it holds now,
// or what it holds after binding a request (each buffer only grows, so the
// component-wise maximum of both).  The sizes follow job_reserve, slot_bind,
// adgpu_job_set_autostop and slot_results.
struct SlotCaps {
    size_t ne = 0, ls = 0, yield = 0, own = 0, as_bins = 0;
    int runs = 0, as_log = 0, best = 0;
    bool as = false;   // AutoStop state and final record
};

static size_t caps_bytes(const SlotCaps& k) {
    size_t b = k.ne * (2 * kGenes * sizeof(float) + 2 * sizeof(float) + sizeof(int32_t) +
                       ADGPU_NWI * sizeof(uint32_t));
    b += (size_t)k.runs * sizeof(int32_t) * (1 + kRing);
    b += k.ls * (sizeof(int32_t) + kGenes * sizeof(float));
    b += k.yield * (kAdStateFloats * sizeof(float) + 3 * sizeof(int32_t));
    b += k.own;
    if (k.as) b += sizeof(AutoStopDev) + sizeof(adgpu_autostop_record);
    b += k.as_bins * sizeof(float) + (size_t)k.as_log * sizeof(adgpu_autostop_record);
    b += (size_t)k.best * (kGenes * sizeof(float) + 3 * sizeof(double));
    return b;
}

static SlotCaps caps_of(const adgpu_job* j) {
    SlotCaps k;
    k.ne = j->cap_ne; k.ls = j->cap_ls; k.yield = j->cap_yield; k.own = j->own_cap;
    k.as_bins = j->as_bins_cap; k.runs = j->cap_runs; k.as_log = j->as_log_alloc; k.best = j->best_cap;
    k.as = j->d_as != nullptr;
    return k;
}

static SlotCaps caps_needed(const adgpu_context* c, const PreparedLigand& lig, const adgpu_search_params* sp,
                            int as_frequency, bool rescore) {
    SlotCaps k;
    const int nrun = sp->num_runs;
    const size_t nls_all = (size_t)nrun * std::max(1, to_dev(sp, lig.hd.num_genes).num_ls);
    k.ne = (size_t)nrun * sp->pop_size;
    k.runs = nrun;
    k.ls = sp->ls_method == ADGPU_LS_FIRE ? nls_all : 0;
    k.yield = yield_wanted(c, sp) ? nls_all : 0;
    k.own = (lig.blob.size() + 65535) & ~size_t(65535);
    if (as_frequency > 0) {
        k.as = true;
        k.as_bins = 3 * ((size_t)sp->pop_size + 1);
        const int64_t tests = sp->max_generations / as_frequency + 2;
        k.as_log = (int)std::max<int64_t>(16, std::min<int64_t>(tests, 16384));
    }
    k.best = rescore ? nrun : 0;
    return k;
}

static SlotCaps caps_max(const SlotCaps& a, const SlotCaps& b) {
    SlotCaps k;
    k.ne = std::max(a.ne, b.ne); k.ls = std::max(a.ls, b.ls); k.yield = std::max(a.yield, b.yield);
    k.own = std::max(a.own, b.own); k.as_bins = std::max(a.as_bins, b.as_bins);
    k.runs = std::max(a.runs, b.runs); k.as_log = std::max(a.as_log, b.as_log); k.best = std::max(a.best, b.best);
    k.as = a.as || b.as;
    return k;
}

size_t request_device_bytes(const adgpu_context* c, const PreparedLigand& lig, const adgpu_search_params* sp,
                            int as_frequency, bool rescore) {
    return caps_bytes(caps_needed(c, lig, sp, as_frequency, rescore));
}

size_t slot_held_bytes(const adgpu_job* j) { return caps_bytes(caps_of(j)); }

size_t slot_bytes_after(const adgpu_job* j, const PreparedLigand& lig, const adgpu_search_params* sp,
                        int as_frequency, bool rescore) {
    return caps_bytes(caps_max(caps_of(j), caps_needed(j->ctx, lig, sp, as_frequency, rescore)));
}

void slot_shrink(adgpu_job* j) {
    (void)hipSetDevice(j->ctx->device);
    for (auto& g : j->groups) (void)hipStreamSynchronize(g.stream);
    (void)hipFree(j->pop[0]); (void)hipFree(j->pop[1]);
    (void)hipFree(j->ene[0]); (void)hipFree(j->ene[1]);
    (void)hipFree(j->evals_new); (void)hipFree(j->prng); (void)hipFree(j->evals_runs); (void)hipFree(j->d_ring);
    (void)hipFree(j->ls_ent); (void)hipFree(j->ls_grad);
    (void)hipFree(j->ls_state); (void)hipFree(j->ls_ring);
    (void)hipFree(j->own.dev);
    (void)hipFree(j->d_as); (void)hipFree(j->d_as_out); (void)hipFree(j->d_as_bins); (void)hipFree(j->d_as_log);
    (void)hipFree(j->d_best); (void)hipFree(j->d_f64);
    if (j->h_ring) (void)hipHostFree(j->h_ring);
    j->pop[0] = j->pop[1] = j->ene[0] = j->ene[1] = nullptr;
    j->evals_new = nullptr; j->prng = nullptr; j->evals_runs = nullptr; j->h_ring = nullptr; j->d_ring = nullptr;
    j->ls_ent = nullptr; j->ls_grad = nullptr; j->ls_state = nullptr; j->ls_ring = nullptr;
    j->own.dev = nullptr;
    j->d_as = nullptr; j->d_as_out = nullptr; j->d_as_bins = nullptr; j->d_as_log = nullptr;
    j->d_best = nullptr; j->d_f64 = nullptr;
    j->cap_ne = j->cap_ls = j->cap_yield = j->own_cap = j->as_bins_cap = 0;
    j->cap_runs = j->as_log_alloc = j->best_cap = 0;
    j->as_enabled = false;
}

int slot_open(adgpu_context* c, int run_groups, int sharers, adgpu_job** out) {
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    std::unique_ptr<adgpu_job> j(new adgpu_job());
    j->ctx = c;
    j->own_lig = true;
    j->sharers = std::max(1, sharers);
    j->want_groups = std::max(1, std::min(run_groups, kMaxGroups));
    int rc = job_open(j.get());
    if (!rc) rc = job_set_groups(j.get(), run_groups, run_groups);
    if (rc) {
        const std::string msg = t_err;
        adgpu_job_destroy(j.release());
        return fail(c, rc, msg);
    }
    *out = j.release();
    return ADGPU_OK;
}

int slot_bind(adgpu_job* j, int map_id, const PreparedLigand& lig, const adgpu_search_params* sp,
              const float* init_pop, const uint32_t* seeds) {
    adgpu_context* c = j->ctx;
    if (!j->own_lig || j->counted) return fail(c, ADGPU_ERR_STATE, "serving slot already bound");
    HIPCHK(c, hipSetDevice(c->device));
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto mi = c->maps.find(map_id);
        if (mi == c->maps.end()) return fail(c, ADGPU_ERR_ARG, "unknown map id");
        mi->second.jobs++;
        j->map_id = map_id;
        j->counted = true;
    }
    const int nrun = sp->num_runs;
    const SearchDev sd = to_dev(sp, lig.hd.num_genes);
    int rc = job_set_groups(j, j->want_groups, nrun);
    if (!rc) rc = job_reserve(j, nrun, sp->pop_size, sd.num_ls, sp->ls_method == ADGPU_LS_FIRE, yield_wanted(c, sp));
    if (!rc && lig.blob.size() > j->own_cap) {
        if (j->own_cap) j->reallocs++;
        (void)hipFree(j->own.dev);
        j->own.dev = nullptr;
        j->own_cap = 0;
        const size_t cap = (lig.blob.size() + 65535) & ~size_t(65535);
        if (hipMalloc(&j->own.dev, cap) != hipSuccess) rc = fail(c, ADGPU_ERR_NOMEM, "hipMalloc(ligand) failed");
        else j->own_cap = cap;
    }
    if (!rc) {
        j->own.host = lig.hd;
        j->own.max_map = lig.max_map;
        const hipError_t e = hipMemcpyAsync(j->own.dev, lig.blob.data(), lig.blob.size(), hipMemcpyHostToDevice,
                                            j->groups[0].stream);
        if (e != hipSuccess) rc = fail(c, ADGPU_ERR_HIP, std::string("hipMemcpyAsync(ligand): ") + hipGetErrorString(e));
    }
    if (!rc) rc = job_start(j, sp, lig.hd, init_pop, seeds);   // waits for the copies
    if (rc) {
        const std::string msg = t_err;
        slot_unbind(j);
        return fail(c, rc, msg);
    }
    return ADGPU_OK;
}

void slot_unbind(adgpu_job* j) {
    if (!j->counted) return;
    std::lock_guard<std::mutex> lk(j->ctx->mu);
    auto mi = j->ctx->maps.find(j->map_id);
    if (mi != j->ctx->maps.end()) mi->second.jobs--;
    j->counted = false;
}

int slot_results(adgpu_job* j, float* best_geno, float* best_e, double* best_f64, float* pop, float* ene) {
    adgpu_context* c = j->ctx;
    const int nrun = j->sp.num_runs, pz = j->sp.pop_size;
    const size_t ne = (size_t)nrun * pz;
    std::vector<float> P(ne * kGenes), E(ne);
    int rc = adgpu_job_download(j, P.data(), E.data(), nullptr, nullptr);
    if (rc) return rc;
    std::vector<float> geno((size_t)nrun * kGenes);
    for (int r = 0; r < nrun; ++r) {
        int bi = 0;
        for (int k = 1; k < pz; ++k)
            if (E[(size_t)r * pz + k] < E[(size_t)r * pz + bi]) bi = k;
        std::memcpy(&geno[(size_t)r * kGenes], &P[((size_t)r * pz + bi) * kGenes], kGenes * sizeof(float));
        if (best_e) best_e[r] = E[(size_t)r * pz + bi];
    }
    if (best_geno) std::memcpy(best_geno, geno.data(), geno.size() * sizeof(float));
    if (pop) std::memcpy(pop, P.data(), P.size() * sizeof(float));
    if (ene) std::memcpy(ene, E.data(), E.size() * sizeof(float));
    if (!best_f64) return ADGPU_OK;
    const LigandSet* ls; const MapSet* ms;
    if ((rc = job_lookup(j, &ls, &ms))) return rc;
    if (nrun > j->best_cap) {
        (void)hipFree(j->d_best); (void)hipFree(j->d_f64);
        j->d_best = nullptr;
        j->d_f64 = nullptr;
        j->best_cap = 0;
        if (hipMalloc(&j->d_best, (size_t)nrun * kGenes * sizeof(float)) != hipSuccess ||
            hipMalloc(&j->d_f64, (size_t)nrun * 3 * sizeof(double)) != hipSuccess)
            return fail(c, ADGPU_ERR_NOMEM, "hipMalloc(re-scoring scratch) failed");
        j->best_cap = nrun;
    }
    hipStream_t s = j->groups[0].stream;
    HIPCHK(c, hipMemcpyAsync(j->d_best, geno.data(), geno.size() * sizeof(float), hipMemcpyHostToDevice, s));
    HIPCHK(c, launch_rescore_f64(launch_args(c, *ls, *ms, s, 1, j->team_hint), j->d_best, nrun, j->d_f64));
    HIPCHK(c, hipMemcpyAsync(best_f64, j->d_f64, (size_t)nrun * 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return ADGPU_OK;
}

}  // namespace adgpu
