// adgpu_serve.cpp -- the serving loop of include/adgpu.h (adgpu_server_*).
//
// Requests queue in arrival order.  Each slot has one worker thread and one
// reusable job (adgpu_slots.h) whose device buffers outlive its request.  The
// memory budget bounds what the slots hold, idle ones included: the head of
// the queue goes to the idle slot that needs the least growth to bind it,
// when the bytes of every slot after that growth fit the budget; when no idle
// slot fits but freeing the buffers of idle slots would make room, those
// slots free them first.  A worker runs the whole docking (initial
// population, generation loop with AutoStop, per-run best and optional FP64
// re-scoring) on its own streams and files the result.  The reference's counterpart is the per-ligand loop
// of main.cpp (:312-327,401-434), which docks one ligand per device at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "adgpu_slots.h"

using namespace adgpu;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0;
    int map_id = 0;
    PreparedLigand lig;
    adgpu_search_params sp{};
    std::vector<float> init;
    std::vector<uint32_t> seeds;
    int autostop = 0, as_frequency = 0;
    float stopstd = 0.f;
    int64_t min_as_evals = 0;
    int flags = 0;
    size_t bytes = 0;
    bool done = false;
    bool claimed = false;   // a thread is in adgpu_server_result for this ticket
    adgpu_request_result res{};
    std::vector<float> best_geno, best_e, pop, ene;
    std::vector<double> best_f64;
};

}  // namespace

struct adgpu_server {
    adgpu_context* ctx = nullptr;
    int device = 0;
    size_t budget = 0;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;   // queue or memory changed, or stopping
    std::condition_variable cv_done;   // a request finished
    std::deque<Request*> queue;        // submitted, not admitted (arrival order)
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> finished;      // finished and not yet returned by next_done
    int64_t next_ticket = 1, submitted = 0, completed = 0, failed = 0;
    size_t in_use = 0, peak_bytes = 0;   // device bytes the slots hold (sum of held)
    int active = 0, peak_active = 0;
    bool stopping = false;
    std::vector<adgpu_job*> slots;
    std::vector<size_t> held;            // per slot: bytes of its buffers (after growth, once assigned)
    std::vector<Request*> assigned;      // per slot: request handed to it, not yet taken by its worker
    std::vector<char> busy;              // per slot: assigned or running a request
    std::vector<char> shrink;            // per slot: idle, asked to free its buffers
    std::vector<std::thread> workers;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

void run_request(adgpu_server* s, adgpu_job* j, Request* r) {
    adgpu_request_result& res = r->res;
    const int nrun = r->sp.num_runs;
    int rc = slot_bind(j, r->map_id, r->lig, &r->sp, r->init.data(), r->seeds.data());
    std::vector<float>().swap(r->init);       // uploaded; free the host copies
    std::vector<uint32_t>().swap(r->seeds);
    if (!rc) {
        if (r->autostop) rc = adgpu_job_set_autostop(j, 1, r->stopstd, r->as_frequency, r->min_as_evals);
        if (!rc) rc = adgpu_job_initpop(j);
        int64_t gens = 0, total = 0;
        if (!rc) rc = adgpu_job_run(j, &gens, &total);
        res.generations = gens;
        res.total_evals = total;
        if (!rc) {
            r->best_geno.resize((size_t)nrun * ADGPU_GENOTYPE_STRIDE);
            r->best_e.resize(nrun);
            if (r->flags & ADGPU_REQ_RESCORE_F64) r->best_f64.resize((size_t)nrun * 3);
            if (r->flags & ADGPU_REQ_KEEP_POPULATION) {
                r->pop.resize((size_t)nrun * r->sp.pop_size * ADGPU_GENOTYPE_STRIDE);
                r->ene.resize((size_t)nrun * r->sp.pop_size);
            }
            rc = slot_results(j, r->best_geno.data(), r->best_e.data(),
                              r->best_f64.empty() ? nullptr : r->best_f64.data(),
                              r->pop.empty() ? nullptr : r->pop.data(), r->ene.empty() ? nullptr : r->ene.data());
        }
        slot_unbind(j);
    }
    res.status = rc;
    if (rc) std::snprintf(res.message, sizeof(res.message), "%s", thread_error());
    (void)s;
}

bool rescore_of(const Request* r) { return (r->flags & ADGPU_REQ_RESCORE_F64) != 0; }

// Admission, with s->mu held: hands queued requests, in arrival order, to idle
// slots while their bytes fit, or asks idle slots to free their buffers when
// only that makes room for the head.  Returns true when a worker has work.
bool schedule(adgpu_server* s) {
    bool work = false;
    while (!s->queue.empty()) {
        Request* r = s->queue.front();
        const int as_freq = r->autostop ? r->as_frequency : 0;
        int best = -1;
        size_t best_after = 0, best_growth = SIZE_MAX;
        for (size_t k = 0; k < s->slots.size(); ++k) {
            if (s->busy[k] || s->shrink[k]) continue;
            const size_t after = slot_bytes_after(s->slots[k], r->lig, &r->sp, as_freq, rescore_of(r));
            if (s->in_use - s->held[k] + after > s->budget) continue;
            if (after - s->held[k] < best_growth) {
                best = (int)k;
                best_after = after;
                best_growth = after - s->held[k];
            }
        }
        if (best >= 0) {
            s->queue.pop_front();
            s->in_use += best_after - s->held[best];
            s->held[best] = best_after;
            s->peak_bytes = std::max(s->peak_bytes, s->in_use);
            s->busy[best] = 1;
            s->assigned[best] = r;
            s->active++;
            s->peak_active = std::max(s->peak_active, s->active);
            r->res.t_admit_ms = s->now_ms();
            r->res.slot = best;
            work = true;
            continue;
        }
        // No idle slot fits.  Freeing idle slots' buffers makes room for the
        // head (in an emptied slot it needs r->bytes) if the bytes held by
        // running slots leave that much; free the largest first.
        size_t running = 0, releasing = 0;
        std::vector<int> idle;
        for (size_t k = 0; k < s->slots.size(); ++k) {
            if (s->busy[k]) running += s->held[k];
            else if (s->shrink[k]) releasing += s->held[k];
            else if (s->held[k] > 0) idle.push_back((int)k);
        }
        if (running + r->bytes <= s->budget && !idle.empty()) {
            std::sort(idle.begin(), idle.end(), [&](int a, int b) { return s->held[a] > s->held[b]; });
            size_t freed = releasing;
            for (int k : idle) {
                if (s->in_use - freed + r->bytes <= s->budget) break;
                s->shrink[k] = 1;
                freed += s->held[k];
                work = true;
            }
        }
        break;   // the head waits for running requests or for buffers being freed
    }
    return work;
}

void worker(adgpu_server* s, int slot) {
    (void)hipSetDevice(s->device);
    adgpu_job* j = s->slots[slot];
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stopping || s->assigned[slot] || s->shrink[slot]; });
            if (s->stopping) return;
            r = s->assigned[slot];
            s->assigned[slot] = nullptr;
        }
        if (!r) {   // asked to free this idle slot's buffers
            slot_shrink(j);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->in_use -= s->held[slot];
                s->held[slot] = slot_held_bytes(j);
                s->in_use += s->held[slot];
                s->shrink[slot] = 0;
                schedule(s);
            }
            s->cv_work.notify_all();
            continue;
        }
        try {
            run_request(s, j, r);
        } catch (const std::exception& ex) {   // host memory for the results, for instance
            slot_unbind(j);
            r->res.status = ADGPU_ERR_NOMEM;
            std::snprintf(r->res.message, sizeof(r->res.message), "%s", ex.what());
        }
        (void)map_hold(s->ctx, r->map_id, -1);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            // what the slot keeps for its next request (a failed growth may
            // have left less than was charged)
            s->in_use -= s->held[slot];
            s->held[slot] = slot_held_bytes(j);
            s->in_use += s->held[slot];
            s->busy[slot] = 0;
            s->active--;
            r->res.t_done_ms = s->now_ms();
            r->done = true;
            s->completed++;
            if (r->res.status) s->failed++;
            s->finished.push_back(r->ticket);
            schedule(s);
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
}

template <class Pred>
bool wait_for(std::condition_variable& cv, std::unique_lock<std::mutex>& lk, int timeout_ms, Pred pred) {
    if (timeout_ms < 0) {
        cv.wait(lk, pred);
        return true;
    }
    return cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
}

}  // namespace

extern "C" {

int adgpu_server_create(adgpu_ctx ctx, const adgpu_server_config* cfg, adgpu_server_t* out) {
    if (!ctx || !cfg || !out) return set_error(ADGPU_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 256) return set_error(ADGPU_ERR_ARG, "slots must be in 1..256");
    if (cfg->run_groups < 0 || cfg->memory_budget < 0) return set_error(ADGPU_ERR_ARG, "negative run_groups or budget");
    std::unique_ptr<adgpu_server> s(new adgpu_server());
    s->ctx = ctx;
    s->device = context_device(ctx);
    if (hipSetDevice(s->device) != hipSuccess) return set_error(ADGPU_ERR_HIP, "hipSetDevice failed");
    if (cfg->memory_budget > 0) {
        s->budget = (size_t)cfg->memory_budget;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return set_error(ADGPU_ERR_HIP, "hipMemGetInfo failed");
        s->budget = free_b / 10 * 9;
    }
    s->t0 = Clock::now();
    for (int k = 0; k < cfg->slots; ++k) {
        adgpu_job* j = nullptr;
        int rc = slot_open(ctx, cfg->run_groups > 0 ? cfg->run_groups : 1, cfg->slots, &j);
        if (rc) {
            const std::string msg = thread_error();
            for (adgpu_job* o : s->slots) adgpu_job_destroy(o);
            return set_error(rc, msg);
        }
        s->slots.push_back(j);
    }
    s->held.assign(s->slots.size(), 0);
    s->assigned.assign(s->slots.size(), nullptr);
    s->busy.assign(s->slots.size(), 0);
    s->shrink.assign(s->slots.size(), 0);
    for (int k = 0; k < cfg->slots; ++k) s->workers.emplace_back(worker, s.get(), k);
    *out = s.release();
    return ADGPU_OK;
}

int adgpu_server_submit(adgpu_server_t s, const adgpu_request* q, int64_t* ticket) {
    if (!s || !q || !ticket || !q->ligand || !q->energy || !q->search || !q->init_population || !q->prng_seeds)
        return set_error(ADGPU_ERR_ARG, "null argument");
    if (q->autostop && (q->as_frequency < 1 || !(q->stopstd > 0.0f) || q->min_as_evals < 0))
        return set_error(ADGPU_ERR_ARG, "AutoStop needs as_frequency >= 1, stopstd > 0 and min_as_evals >= 0");
    std::unique_ptr<Request> r(new Request());
    int rc = prepare_ligand(s->ctx, q->ligand, q->energy, r->lig);
    if (!rc) rc = check_request(s->ctx, q->map_id, r->lig, q->search);
    if (rc) return rc;
    r->map_id = q->map_id;
    r->sp = *q->search;
    const size_t ne = (size_t)r->sp.num_runs * r->sp.pop_size;
    r->init.assign(q->init_population, q->init_population + ne * ADGPU_GENOTYPE_STRIDE);
    r->seeds.assign(q->prng_seeds, q->prng_seeds + ne * ADGPU_NWI);
    r->autostop = q->autostop != 0;
    r->as_frequency = q->as_frequency;
    r->stopstd = q->stopstd;
    r->min_as_evals = q->min_as_evals;
    r->flags = q->flags;
    r->bytes = request_device_bytes(s->ctx, r->lig, &r->sp, r->autostop ? r->as_frequency : 0, rescore_of(r.get()));
    r->res.num_runs = r->sp.num_runs;
    r->res.pop_size = r->sp.pop_size;
    r->res.device_bytes = (int64_t)r->bytes;
    r->res.slot = -1;
    if (r->bytes > s->budget) return set_error(ADGPU_ERR_LIMIT, "request larger than the server's memory budget");
    if ((rc = map_hold(s->ctx, r->map_id, +1))) return rc;   // released when the request ends
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) {
            (void)map_hold(s->ctx, r->map_id, -1);
            return set_error(ADGPU_ERR_STATE, "server is shutting down");
        }
        r->ticket = s->next_ticket++;
        r->res.t_submit_ms = s->now_ms();
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->reqs[r->ticket] = std::move(r);
        s->submitted++;
        schedule(s);
    }
    s->cv_work.notify_all();
    return ADGPU_OK;
}

int adgpu_server_next_done(adgpu_server_t s, int timeout_ms, int64_t* ticket) {
    if (!s || !ticket) return set_error(ADGPU_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!wait_for(s->cv_done, lk, timeout_ms, [&] { return !s->finished.empty(); }))
        return set_error(ADGPU_ERR_TIMEOUT, "no request finished within the timeout");
    *ticket = s->finished.front();
    s->finished.pop_front();
    return ADGPU_OK;
}

int adgpu_server_result(adgpu_server_t s, int64_t ticket, int timeout_ms, adgpu_request_result* out,
                        float* best_genotypes, float* best_energies, double* best_f64, float* population,
                        float* energies) {
    if (!s || !out) return set_error(ADGPU_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(ADGPU_ERR_ARG, "unknown ticket");
    Request* r = it->second.get();
    if (r->claimed) return set_error(ADGPU_ERR_STATE, "another thread is collecting this ticket");
    r->claimed = true;
    if (!wait_for(s->cv_done, lk, timeout_ms, [&] { return r->done; })) {
        r->claimed = false;
        return set_error(ADGPU_ERR_TIMEOUT, "request not finished within the timeout");
    }
    r->claimed = false;
    if (best_f64 && r->res.status == ADGPU_OK && r->best_f64.empty())   // the ticket stays collectable
        return set_error(ADGPU_ERR_STATE, "request ran without ADGPU_REQ_RESCORE_F64");
    if ((population || energies) && r->res.status == ADGPU_OK && r->pop.empty())
        return set_error(ADGPU_ERR_STATE, "request ran without ADGPU_REQ_KEEP_POPULATION");
    *out = r->res;
    auto copy = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    copy(best_genotypes, r->best_geno);
    copy(best_energies, r->best_e);
    copy(best_f64, r->best_f64);
    copy(population, r->pop);
    copy(energies, r->ene);
    auto f = std::find(s->finished.begin(), s->finished.end(), ticket);
    if (f != s->finished.end()) s->finished.erase(f);
    s->reqs.erase(it);
    return ADGPU_OK;
}

int adgpu_server_stats_get(adgpu_server_t s, adgpu_server_stats* out) {
    if (!s || !out) return set_error(ADGPU_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->queued = (int64_t)s->queue.size();
    out->active = s->active;
    out->peak_active = s->peak_active;
    out->bytes_in_use = (int64_t)s->in_use;
    out->peak_bytes = (int64_t)s->peak_bytes;
    out->memory_budget = (int64_t)s->budget;
    int64_t re = 0;
    for (const adgpu_job* j : s->slots) re += slot_reallocations(j);
    out->slot_reallocations = re;
    out->now_ms = s->now_ms();
    return ADGPU_OK;
}

int adgpu_server_destroy(adgpu_server_t s) {
    if (!s) return set_error(ADGPU_ERR_ARG, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        for (Request* r : s->queue) (void)map_hold(s->ctx, r->map_id, -1);   // dropped unrun
        s->queue.clear();
        for (Request*& r : s->assigned)
            if (r) { (void)map_hold(s->ctx, r->map_id, -1); r = nullptr; }    // admitted, not started
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (adgpu_job* j : s->slots) adgpu_job_destroy(j);
    delete s;
    return ADGPU_OK;
}

}  // extern "C"
