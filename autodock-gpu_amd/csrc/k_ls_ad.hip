// k_ls_ad.hip -- ADADELTA local search (kernel_ad.cl:60-549, with the
// ADADELTA_AUTOSTOP varrho schedule), as a persistent work-queue kernel.
#include <type_traits>

#include "adgpu_device.h"
#include "adgpu_items.h"
#include "adgpu_launch.h"

namespace adgpu {

// Polls a team makes for one ring entry before it reports a fault and stops
// (an entry is written right after its claim, and claims all precede the
// first ticket, so this bound is never reached by a correct run).
constexpr int kHandoffPolls = 1 << 22;

// s_waitcnt vmcnt(0): every vector memory access of the wavefront, the
// saved record's device-scope stores among them, has completed.  The hand-off
// is ordered this way rather than by release/acquire fences, which write back
// and invalidate the whole L2 of an XCD (and with it the grid maps every
// evaluation reads).
constexpr int kWaitStores = 0x0F70;

// Ring entries: one per LS slot claimed from the queue, written by the team
// that claimed it, right after the claim.
constexpr int kRingUnset = -1;   // not written yet (cleared before the launch)
constexpr int kRingNone = -2;    // claimed without a hand-off

// Team lane 0: takes tickets over the ring in claim order.  Returns the LS
// slot of a handed-off search to resume, -1 when the team should look again in
// its next iteration (its ticket's entry is not written yet), -2 when the
// tickets are used up, -3 on a fault.  `ticket` keeps lane 0's ticket across
// calls (-1: none).
DEV int take_handoff(int32_t* Q, const int32_t* ring, int ntot, int& ticket, int& polls) {
    for (int k = 0; k < 4; ++k) {   // a few entries without a hand-off per call
        if (ticket < 0) {
            ticket = atomicAdd(&Q[2], 1);
            polls = 0;
        }
        if (ticket >= ntot) return -2;
        const int v = __hip_atomic_load(const_cast<int32_t*>(&ring[ticket]), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        if (v >= 0 && v < ntot) {
            ticket = -1;
            return v;
        }
        if (v == kRingUnset) {
            if (++polls > kHandoffPolls) {
                __hip_atomic_store(&Q[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return -3;
            }
            return -1;
        }
        ticket = -1;   // kRingNone: no hand-off at this claim
    }
    return -1;
}

// Seat entries (regrouping): kRingUnset until a search is handed to the seat
// (its LS slot) or the seat is given up.
constexpr int kSeatCancelled = -3;

// Team lane 0, a team with a search that no other team of its wavefront can
// share: claims the oldest seat ticket no search has been handed to yet
// (Q[6] < Q[5]) if it is below `cap`.  Returns the ticket, or -1 when none is
// free or another team claimed it first.
DEV int claim_seat(int32_t* Q, int cap) {
    int t = __hip_atomic_load(&Q[6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int h = __hip_atomic_load(&Q[5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t >= h || t >= cap) return -1;
    // one attempt: many lone searches try at once, and retries would queue on
    // the one counter
    if (__hip_atomic_compare_exchange_strong(&Q[6], &t, t + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
        return t;
    return -1;
}

// Team lane 0: moves seat entry *e from kRingUnset to v.  Returns the entry's
// value afterwards (v on success; the search handed to it, or kSeatCancelled,
// when another team got there first).
DEV int seat_swap(int32_t* e, int v) {
    int u = kRingUnset;
    __hip_atomic_compare_exchange_strong(e, &u, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return u == kRingUnset ? v : u;
}

// Each team takes LS slots (run, entity) from the queue and runs one
// ADADELTA search per slot (adgpu_items.h).
//
// Hand-off (S.ls_yield > 0 and a ring): searches last from one to max_iters
// iterations, so teams that drew long searches late would finish the launch
// alone.  A search that reaches ls_yield iterations while unstarted slots
// remain is saved to ls_state, and its team starts the unstarted slot it
// claimed.  The ring has one entry per claim: the claiming team writes the
// handed-off search there, or kRingNone.  Once no unstarted slot is left,
// teams take tickets over the ring in claim order (one atomic add each, no
// compare-and-swap for many teams to contend on) and resume the searches they
// find.  Each search is handed off at most once, and a resumed search
// computes exactly what it would have computed without the hand-off.  Every
// claim precedes the first ticket, so a ticket waits only for an entry its
// claimer is about to write, and waits by looking once per iteration, so its
// wavefront partner keeps running.
//
// Regrouping (S.ls_regroup > 0, teams narrower than the wavefront): once the
// queue and the ring are drained, a team whose wavefront still runs another
// team's search, but has none of its own, takes a seat (a ticket Q[5] and a
// seat entry) and looks at it once per iteration; a team whose search is the
// only work left in its wavefront tries, whenever the search's iteration count
// is a multiple of ls_regroup, to save it and hand it to the oldest free seat
// (a ticket Q[6] < Q[5]) if that seat is older than every seat its own
// wavefront holds: two half-idle wavefronts become one busy and one finished,
// and the lanes the idle teams spent on evaluations nobody used go to other
// launches.  A search moves at most once.  A seated
// team whose wavefront has no other search left gives its seat up; giving up
// and handing over race on the seat entry (one compare-and-swap each), so a
// search handed to a seat is taken exactly once, and a hand-over that loses
// keeps the search where it is.  Like the ring hand-off it computes exactly
// what the search would have computed on its first team.
template <int TW, int NC, int MF, bool PL>
__global__ ADGPU_GROUP_KERNEL void k_ls_ad(Launch a) {
    if (JOB_STOPPED(a.B)) return;
    KERNEL_PROLOGUE(TW, NC, PL)
    const SearchDev& S = a.S;
    const int nls = S.num_ls, pop = S.pop_size, NG = S.num_genes;
    const int ntot = S.num_runs * nls;
    int32_t* const Q = a.B.ls_counter;
    int32_t* const ring = a.B.ls_ring;
    const int yq = ring != nullptr && a.B.ls_state != nullptr ? S.ls_yield : 0;
    const int rk = TW < 64 && yq > 0 ? S.ls_regroup : 0;   // regrouping: seats follow the ring
    int32_t* const seats = rk > 0 ? ring + ntot : nullptr;
    const int seat_cap = 2 * ntot;
    // While a team has no search, its AdItem is dead: cons_succ and cons_fail
    // then hold lane 0's ring ticket (-1: none) and its poll count, and once
    // the ring is drained cons_fail holds the team's seat while it is seated.
    AdItem<TW> it;
#pragma unroll
    for (int j = 0; j < GPL; ++j) { it.geno[j] = 0.f; it.best[j] = 0.f; it.sqg[j] = 0.f; it.sqd[j] = 0.f; }
    it.best_e = INFINITY; it.varrho = 1.0f; it.cons_succ = -1; it.cons_fail = 0; it.iters = 0;
    int gi = 0;                      // population index of the search (runs * pop_size < 2^31)
    int t = 0;                       // LS slot of the search in progress
    bool active = false, done = false, need = true;
    bool fresh = true;               // unstarted slots may remain
    bool may_yield = false;          // the search in progress may still be handed off
    bool seated = false;             // holds seat it.cons_fail, waiting for a search
    bool moved = false;              // the search in progress came from a seat (it moves once)
    int peek = kRingUnset;           // lane 0 of a seated team: its seat entry, read last iteration
    uint32_t steps = 0;              // wavefront evaluation steps (tallied once per wavefront)
    for (;;) {
        if (need) {   // uniform within the team
            int c = -1;
            if (fresh) {
                if (tl == 0) {
                    c = atomicAdd(&Q[0], 1);
                    if (yq > 0 && c < ntot) __hip_atomic_store(&ring[c], kRingNone, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
                }
                c = TM::bcast_i(c, 0);
                if (c >= ntot) { fresh = false; c = -1; }
            }
            if (c < 0 && yq > 0) {
                int v = -2;
                if (tl == 0) v = take_handoff(Q, ring, ntot, it.cons_succ, it.cons_fail);
                v = TM::bcast_i(v, 0);
                if (v >= 0) {
                    t = v;   // the record's loads depend on v: they cannot pass the ring load
                    gi = ad_load<TW>(it, a.B.ls_state + (size_t)t * kAdStateFloats);
                    // a record is saved at exactly yq iterations, for an entity of its
                    // slot's run: anything else is a stale or torn record, reported
                    // through the fault flag instead of being resumed
                    const int run = t / nls;
                    const bool valid = it.iters == yq && gi >= run * pop && gi < (run + 1) * pop;
                    if (valid) {
                        may_yield = false;
                        active = true;
                        need = false;
                        moved = false;
                    } else {
                        if (tl == 0) __hip_atomic_store(&Q[3], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        it.cons_succ = -1;
                        it.cons_fail = 0;
                        active = false;
                        need = true;
                    }
                } else {
                    active = false;
                    done = v != -1;
                    need = !done;
                }
            } else if (c >= 0) {
                t = c;
                const int run = t / nls, le = t - run * nls;
                gi = run * pop + ls_entity<TW>(S, le, a.B.prng + (size_t)t * 64);
                ad_begin<TW>(it, a.B.pop_next + (size_t)gi * kGenes, NG);
                may_yield = yq > 0;
                active = true;
                need = false;
                moved = false;
            } else {
                active = false;
                done = true;
                need = false;
            }
            if (!active) {
#pragma unroll
                for (int j = 0; j < GPL; ++j) it.geno[j] = 0.f;
            }
        }
        if (rk > 0) {   // regrouping (uniform over the wavefront from here to its end)
            // another team of the wavefront has a search (teams are uniform: TW lanes each)
            const bool others = __popcll(__ballot(active)) > (active ? TW : 0);
            if (seated) {
                // the entry as read before the last evaluation (a written entry does
                // not change again); giving the seat up re-reads it by compare-and-swap
                int v = peek;
                if (tl == 0 && v == kRingUnset && !others) v = seat_swap(&seats[it.cons_fail], kSeatCancelled);
                v = TM::bcast_i(v, 0);
                if (v >= 0) {   // a search was handed to this seat
                    seated = false;
                    t = v;
                    gi = ad_load<TW>(it, a.B.ls_state + (size_t)t * kAdStateFloats);
                    const int run = t / nls;
                    const bool valid = it.iters >= 1 && it.iters < S.max_iters &&
                                       gi >= run * pop && gi < (run + 1) * pop;
                    if (valid) {
                        may_yield = false;
                        active = true;
                        moved = true;
                    } else {
                        if (tl == 0) __hip_atomic_store(&Q[3], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        it.cons_succ = -1;
                        it.cons_fail = 0;
                    }
                } else if (v == kSeatCancelled) {
                    seated = false;
                    it.cons_fail = 0;
                }
            } else if (done && !active && others) {
                int h = 0;
                if (tl == 0) h = atomicAdd(&Q[5], 1);
                h = TM::bcast_i(h, 0);
                if (h < seat_cap) { seated = true; it.cons_fail = h; peek = kRingUnset; }
            }
        }
        if (!TM::any_in_wave(!done || active || seated)) {
            if (lane_id() == 0 && steps != 0u) atomicAdd(&Q[8], (int)steps);
            break;
        }
        if (!TM::any_in_wave(active)) {   // every team of the wavefront waits for a ring entry or a seat
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        if (rk > 0 && seated && tl == 0)   // read now, used next iteration: the evaluation hides the latency
            peek = __hip_atomic_load(&seats[it.cons_fail], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        float g[GPL];
        ++steps;
        const float e = evaluate<TW, NC, MF, true>(X, it.geno, g);
        // no other team of the wavefront has a search or may still get one from
        // the queue or the ring; a search moves only to a seat older than every
        // seat its wavefront holds, so two half-idle wavefronts merge toward the
        // older seat instead of trading searches
        const bool alone = rk > 0 && __popcll(__ballot(active || !done)) == ((active || !done) ? TW : 0);
        int lim = seat_cap;
        if (rk > 0) {
            const int ms = seated ? it.cons_fail : seat_cap;
#pragma unroll
            for (int k = 0; k < 64 / TW; ++k) lim = min(lim, rdl_i(ms, k * TW));
        }
        if (!active) continue;
        if (ad_step<TW>(it, e, g, NG, S.max_iters)) {
            ad_store<TW>(it, NG, a.B.pop_next + (size_t)gi * kGenes, a.B.ene_next + gi, a.B.evals_new + gi);
            it.cons_succ = -1;
            it.cons_fail = 0;
            active = false;
            need = true;
        } else if (rk > 0 && alone && !moved && it.iters % rk == 0) {
            // (alone: the wavefront's other teams drew no work from the queue or
            // the ring, so both are drained)
            // the wavefront's last search, every rk iterations: hand it to a seat
            // in a busier wavefront (a search moved there runs at least one
            // iteration before it may move again)
            int tk = -1;
            if (tl == 0) {
                tk = claim_seat(Q, lim);
                // a seat given up already: no record to save (the ticket is spent)
                if (tk >= 0 && __hip_atomic_load(&seats[tk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kRingUnset)
                    tk = -1;
            }
            tk = TM::bcast_i(tk, 0);
            if (tk >= 0) {
                ad_save<TW>(it, gi, a.B.ls_state + (size_t)t * kAdStateFloats);
                __builtin_amdgcn_s_waitcnt(kWaitStores);   // the record is written before the seat entry
                int r = 0;
                if (tl == 0) {
                    r = seat_swap(&seats[tk], t);
                    if (r == t) atomicAdd(&Q[7], 1);   // regroups of the job (not cleared per launch)
                }
                r = TM::bcast_i(r, 0);
                if (r == t) {
                    // drained: the next acquisition finds the queue and the ring used up
                    it.cons_succ = -1;
                    it.cons_fail = 0;
                    active = false;
                    need = true;
                }
            }
        } else if (may_yield && it.iters == yq) {
            may_yield = false;
            int c = -1;
            if (fresh) {
                if (tl == 0) c = atomicAdd(&Q[0], 1);
                c = TM::bcast_i(c, 0);
                if (c >= ntot) { fresh = false; c = -1; }
            }
            if (c >= 0) {
                // hand the search off at claim c (one entry per claim, so at most once per slot)
                ad_save<TW>(it, gi, a.B.ls_state + (size_t)t * kAdStateFloats);
                __builtin_amdgcn_s_waitcnt(kWaitStores);   // the record is written before the ring entry
                if (tl == 0) {
                    __hip_atomic_store(&ring[c], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicAdd(&Q[4], 1);   // hand-offs of the job (not cleared per launch)
                }
                t = c;
                const int run = t / nls, le = t - run * nls;
                gi = run * pop + ls_entity<TW>(S, le, a.B.prng + (size_t)t * 64);
                ad_begin<TW>(it, a.B.pop_next + (size_t)gi * kGenes, NG);
                may_yield = true;
            }
        }
    }
}

hipError_t launch_ls_ad(const LaunchArgs& la, const SearchDev& sp, const JobBuffers& b) {
    const int n = sp.num_runs * sp.num_ls;
    Launch a{};
    a.L = la.lig_dev; a.G = la.grid; a.S = sp; a.B = b;
    const Shape sh = choose_shape(*la.lig_host, la.team_hint, ls_group_threads());
    const int per_group = sh.threads / sh.tw;
    hipError_t err = hipSuccess;
    ADGPU_DISPATCH(k_ls_ad, sh, la.grid.format, {
        const int want = (n + per_group - 1) / per_group;
        const int cap = persistent_cap(reinterpret_cast<const void*>(kfn), sh.lds, sh.threads, la.share);
        const int groups = cap > 0 && cap < want ? cap : want;
        hipLaunchKernelGGL(kfn, dim3(groups), dim3(sh.threads), sh.lds, la.stream, a);
        err = hipGetLastError();
    });
    return err;
}

}  // namespace adgpu
