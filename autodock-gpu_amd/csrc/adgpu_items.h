// adgpu_items.h -- the per-item work of the GA generation (one offspring)
// and of the ADADELTA local search (one search), as device functions of one
// team (k_ga.hip, k_ls_ad.hip).
#pragma once

#include "adgpu_device.h"

namespace adgpu {

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

// One offspring of K4 (kernel4.cl:32-336) for a team: entity `ent` of a run
// whose current genotypes and energies are cur / ecur; st = the entity's 64
// LCG streams.  Entity 0 is the elite copy (its energy and streams are kept).
template <int TW>
struct K4Item {
    float child[64 / TW];
    uint32_t state[64 / TW];
    float elite_e;
    bool elite;
};

template <int TW>
DEV void k4_prepare(const SearchDev& S, const float* cur, const float* ecur, const uint32_t* st, int ent,
                    K4Item<TW>& it) {
    using TM = Team<TW>;
    constexpr int GPL = TM::kGPL;
    const int tl = TM::lane();
    const int pop = S.pop_size, NG = S.num_genes;
    it.elite_e = 0.f;
    it.elite = ent == 0;
    if (it.elite) {
        int bi;
        elite_of_run<TW>(ecur, pop, it.elite_e, bi);
#pragma unroll
        for (int j = 0; j < GPL; ++j) { it.child[j] = cur[(size_t)bi * kGenes + tl + TW * j]; it.state[j] = 0u; }
        return;
    }
#pragma unroll
    for (int j = 0; j < GPL; ++j) it.state[j] = st[tl + TW * j];
    float rn = 0.f;
    if (tl < 10) rn = randf(it.state[0]);
    // parent candidates (team lanes 0..3), binary tournaments
    const int cand = (int)((float)pop * rn);
    const float ce = (tl < 4) ? ecur[cand] : 0.f;
    const int c0 = TM::bcast_i(cand, 0), c1 = TM::bcast_i(cand, 1), c2 = TM::bcast_i(cand, 2),
              c3 = TM::bcast_i(cand, 3);
    const float e0 = TM::bcast(ce, 0), e1 = TM::bcast(ce, 1), e2 = TM::bcast(ce, 2), e3 = TM::bcast(ce, 3);
    const float r4 = TM::bcast(rn, 4), r5 = TM::bcast(rn, 5), r6 = TM::bcast(rn, 6), r7 = TM::bcast(rn, 7),
                r8 = TM::bcast(rn, 8);
    const float tr = S.tournament_rate;
    const int p0 = (e0 < e1) ? ((r4 < tr) ? c0 : c1) : ((r4 < tr) ? c1 : c0);
    const int p1 = (e2 < e3) ? ((r5 < tr) ? c2 : c3) : ((r5 < tr) ? c3 : c2);
    const bool cross = r6 < S.crossover_rate;
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
            if (randf(it.state[j]) < S.mutation_rate) {
                if (g < 3) c += S.abs_max_dmov * (2 * randf(it.state[j]) - 1);
                else { c += S.abs_max_dang * (2 * randf(it.state[j]) - 1); map_angle(c); }
            }
        }
        it.child[j] = c;
    }
}

// Genes to evaluate (zero beyond the genotype).
template <int TW>
DEV void k4_genes(const K4Item<TW>& it, int NG, float (&in)[64 / TW]) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) in[j] = (tl + TW * j) < NG ? it.child[j] : 0.f;
}

// Writes the offspring (genes, energy e, eval count, LCG streams);
// next = the entity's genotype in the next population.
template <int TW>
DEV void k4_store(const K4Item<TW>& it, int NG, float e, float* next, float* ene, int32_t* evals, uint32_t* st) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j)
        if (tl + TW * j < NG) next[tl + TW * j] = it.child[j];
    if (it.elite) {
        if (tl == 0) { *ene = it.elite_e; *evals = 0; }
    } else {
#pragma unroll
        for (int j = 0; j < 64 / TW; ++j) st[tl + TW * j] = it.state[j];
        if (tl == 0) { *ene = e; *evals = 1; }
    }
}

// One ADADELTA search (kernel_ad.cl:60-549, ADADELTA_AUTOSTOP schedule):
// gradient steps delta = -g sqrt((E[dx^2]+eps)/(E[g^2]+eps)), rho = 0.8,
// eps = 1e-2 (kernel_ad.cl:40-41, 443-452); the best genotype so far is kept;
// varrho doubles after 4 successes and halves after 4 failures; the search
// ends at max_iters or varrho < 0.01 (kernel_ad.cl:476-515).
template <int TW>
struct AdItem {
    float geno[64 / TW], best[64 / TW], sqg[64 / TW], sqd[64 / TW];
    float best_e, varrho;
    int cons_succ, cons_fail, iters;
};

template <int TW>
DEV void ad_begin(AdItem<TW>& s, const float* genes, int NG) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) {
        const int g = tl + TW * j;
        s.geno[j] = g < NG ? genes[g] : 0.f;
        s.best[j] = s.geno[j]; s.sqg[j] = 0.f; s.sqd[j] = 0.f;
    }
    s.best_e = INFINITY; s.varrho = 1.0f; s.cons_succ = 0; s.cons_fail = 0; s.iters = 0;
}

// One step after the evaluation (energy e, gradient g) of s.geno; true when
// the search has ended.
template <int TW>
DEV bool ad_step(AdItem<TW>& s, float e, const float (&g)[64 / TW], int NG, int max_iters) {
    const float RHO = 0.8f, EPS = 1e-2f;
    const bool better = e < s.best_e;
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) {
        if (j > 0 && NG <= TW) break;   // genes beyond the first lane set are all zero
        if (better) s.best[j] = s.geno[j];
        s.sqg[j] *= RHO;
        s.sqg[j] += (1.0f - RHO) * g[j] * g[j];
        // native square root of a native quotient, as kernel_ad.cl:447 computes it
        const float delta = -1.0f * g[j] * __builtin_amdgcn_sqrtf((s.sqd[j] + EPS) * __builtin_amdgcn_rcpf(s.sqg[j] + EPS));
        s.sqd[j] *= RHO;
        s.sqd[j] += (1.0f - RHO) * delta * delta;
        s.geno[j] += delta;
    }
    if (better) { s.best_e = e; s.cons_succ++; s.cons_fail = 0; }
    else { s.cons_succ = 0; s.cons_fail++; }
    s.iters++;
    if (s.cons_succ >= 4) { s.varrho *= 2.0f; s.cons_succ = 0; }
    else if (s.cons_fail >= 4) { s.varrho *= 0.5f; s.cons_fail = 0; }
    return !(s.iters < max_iters && s.varrho > 0.01f);
}

// Writes the search's best genotype (angles mapped), energy and eval count.
template <int TW>
DEV void ad_store(const AdItem<TW>& s, int NG, float* genes, float* ene, int32_t* evals) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) {
        const int gg = tl + TW * j;
        float b = s.best[j];
        if (gg >= 3) map_angle(b);
        if (gg < NG) genes[gg] = b;
    }
    if (tl == 0) { *evals += s.iters; *ene = s.best_e; }
}

// Saves a search in progress and its population index to rec
// (kAdStateFloats floats); ad_load restores both exactly, so a search
// continued by another team computes what its first team would have.  The
// record moves between compute units of different XCDs within one launch:
// it is written and read with device-scope accesses (sc1: coherent across
// the XCDs' L2 caches without writing back or invalidating them).
DEV void st_dev(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV float ld_dev(const float* p) {
    return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int TW>
DEV void ad_save(const AdItem<TW>& s, int gi, float* rec) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) {
        const int g = tl + TW * j;
        st_dev(rec + g, s.geno[j]);
        st_dev(rec + kGenes + g, s.best[j]);
        st_dev(rec + 2 * kGenes + g, s.sqg[j]);
        st_dev(rec + 3 * kGenes + g, s.sqd[j]);
    }
    if (tl == 0) {
        float* r = rec + 4 * kGenes;
        st_dev(r + 0, s.best_e); st_dev(r + 1, s.varrho);
        st_dev(r + 2, __int_as_float(s.cons_succ)); st_dev(r + 3, __int_as_float(s.cons_fail));
        st_dev(r + 4, __int_as_float(s.iters)); st_dev(r + 5, __int_as_float(gi));
    }
}

template <int TW>
DEV int ad_load(AdItem<TW>& s, const float* rec) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int j = 0; j < 64 / TW; ++j) {
        const int g = tl + TW * j;
        s.geno[j] = ld_dev(rec + g);
        s.best[j] = ld_dev(rec + kGenes + g);
        s.sqg[j] = ld_dev(rec + 2 * kGenes + g);
        s.sqd[j] = ld_dev(rec + 3 * kGenes + g);
    }
    const float* r = rec + 4 * kGenes;
    s.best_e = ld_dev(r + 0); s.varrho = ld_dev(r + 1);
    s.cons_succ = __float_as_int(ld_dev(r + 2)); s.cons_fail = __float_as_int(ld_dev(r + 3));
    s.iters = __float_as_int(ld_dev(r + 4));
    return __float_as_int(ld_dev(r + 5));
}

}  // namespace adgpu
