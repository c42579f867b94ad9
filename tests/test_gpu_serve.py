"""Serving loop (adgpu_server_*) through the C ABI.

* Each request's result is the same alone and in the mix.  Uneven ligands
  (10-90 atoms, 0-25 torsions) against three resident map sets (FP32 and
  FP16), all four local searches, one request with AutoStop, are served one
  at a time (nothing else on the device), then all at once by the same
  server configuration, whose slots run them concurrently.  Per-run best
  genotypes and energies, generation counts and eval totals must be
  bit-identical: every search is seeded by its own request, and nothing of
  one request reaches another's kernels.  With the context's team width
  fixed, served results also equal the plain job API's.
* Admission by memory: a budget that holds only a few requests keeps the
  bytes in use under it at every moment, and the results are again the same.
* The per-run best energies equal the oracle's FP32 energy of the returned
  genotypes (the bar of helpers.check_energies), and the FP64 re-scoring of
  each run's best pose equals the FP64 oracle to 1e-10 of the sum of |terms|.
* Refused at submit: a request larger than the whole budget, an unknown map
  set.  A map set cannot be released while a queued or running request
  uses it.
"""
import numpy as np
import pytest

import adgpu
import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params, search_params

pytestmark = pytest.mark.gpu

NRUN, POP = 4, 40


def _ligand(n_atoms, n_tors, seed):
    for t in range(n_tors, -1, -1):
        try:
            return synth.make_ligand(n_atoms, t, seed=seed)
        except RuntimeError:
            continue
    raise RuntimeError("no ligand")


def _rounded(packed, fmt):
    if fmt == adgpu.MAP_F16:
        return np.clip(packed, -65504.0, 65504.0).astype(np.float16).astype(np.float32)
    return packed


@pytest.fixture(scope="module")
def mix(gpu_ctx):
    """Ten requests over three resident map sets."""
    rng = np.random.default_rng(4242)
    atoms = [10, 90] + [int(a) for a in rng.integers(11, 90, 8)]
    ligs = [_ligand(a, int(rng.integers(0, 26)), 500 + k) for k, a in enumerate(atoms)]
    types = sorted({t for lig in ligs for t in lig.type_names})
    sets = []
    for npts, seed, fmt in ((40, 51, adgpu.MAP_F32), (48, 52, adgpu.MAP_F16), (56, 53, adgpu.MAP_F32)):
        maps = synth.make_maps(types, npts=npts, seed=seed)
        packed = maps.packed()
        mid = gpu_ctx.maps_upload(packed, maps.size, maps.num_map_atypes, maps.spacing, fmt)
        sets.append((maps, mid, CGrid(_rounded(packed, fmt), maps.size, maps.num_map_atypes, maps.spacing)))
    ep = energy_params()
    methods = ["ad", "sw", "adam", "fire"]
    reqs = []
    for k, lig in enumerate(ligs):
        maps, mid, cg = sets[k % 3]
        T = synth.ligand_tables(lig, maps.spacing, maps.type_names)
        sp = search_params(num_runs=NRUN, pop_size=POP, max_evals=12000, ls=methods[k % 4], spacing=maps.spacing)
        init = O.initial_population((k, 17), NRUN, POP, maps.size_angstr, maps.spacing)
        seeds = O.cmwc_uints((k, 17), NRUN * POP * 64)
        autostop = (0.15, 5, 0) if k == 3 else None
        reqs.append(dict(k=k, mid=mid, T=T, sp=sp, init=init, seeds=seeds, autostop=autostop, cg=cg, ep=ep,
                         types=types))
    yield reqs
    for _, mid, _ in sets:
        gpu_ctx.maps_release(mid)


def _alone(ctx, q):
    """One request through the plain job API (its own ligand upload)."""
    lid = ctx.ligand_upload(q["T"], q["ep"])
    job = ctx.job(q["mid"], lid, q["sp"], q["init"], q["seeds"])
    if q["autostop"] is not None:
        job.set_autostop(True, *q["autostop"])
    job.initpop()
    gens, total = job.run()
    b = job.best_f64()
    job.destroy()
    ctx.ligand_release(lid)
    return dict(generations=gens, total_evals=total, best_genotypes=b["genotypes"], best_energies=b["energy_f32"],
                best_f64=np.stack([b["total"], b["inter"], b["intra"]], axis=1))


def _serve(ctx, reqs, slots, budget=0, one_at_a_time=False):
    """Results of the requests through one server; one_at_a_time: each
    request finishes before the next is submitted (nothing else on the
    device)."""
    with adgpu.Server(ctx, slots=slots, memory_budget=budget) as srv:
        def sub(q):
            return srv.submit(q["mid"], q["T"], q["sp"], q["init"], q["seeds"], ep=q["ep"], autostop=q["autostop"],
                              rescore_f64=True)
        if one_at_a_time:
            out = [srv.result(sub(q)) for q in reqs]
        else:
            tickets = [sub(q) for q in reqs]
            print("submitted", len(tickets), "slots", slots, "budget", budget, flush=True)
            out = [srv.result(t) for t in tickets]
        stats = srv.stats()
    return out, stats


def _same(a, b, k):
    assert a["generations"] == b["generations"], (k, a["generations"], b["generations"])
    assert a["total_evals"] == b["total_evals"], (k, a["total_evals"], b["total_evals"])
    assert np.array_equal(a["best_energies"], b["best_energies"]), k
    assert np.array_equal(a["best_genotypes"], b["best_genotypes"]), k
    assert np.array_equal(a["best_f64"], b["best_f64"]), k


def test_each_result_is_the_same_alone_and_in_the_mix(gpu_ctx, mix):
    alone, _ = _serve(gpu_ctx, mix, slots=5, one_at_a_time=True)
    served, stats = _serve(gpu_ctx, mix, slots=5)
    print("stats", stats, flush=True)
    assert stats["completed"] == len(mix) and stats["failed"] == 0
    assert stats["peak_active"] >= 2          # the requests overlapped on the device
    for q, a, s in zip(mix, alone, served):
        assert s["status"] == 0, s["message"]
        assert s["t_submit_ms"] <= s["t_admit_ms"] <= s["t_done_ms"]
        _same(a, s, q["k"])
    # the AutoStop request either used its whole budget or stopped at a test generation
    assert served[3]["total_evals"] >= NRUN * 12000 or served[3]["generations"] % 5 == 0


def test_served_results_equal_the_job_api(gpu_ctx, mix):
    """With the team width fixed for the context, a served request and the
    same docking through the job API (its own ligand upload, job, run, FP64
    re-scoring) are bit-identical.  Without a fixed width a server that
    shares the device between requests may pick 32-lane teams where a job on
    its own picks 64 (adgpu_set_team_width)."""
    gpu_ctx.set_team_width(32)
    try:
        alone = []
        for q in mix:
            print("job api", q["k"], "atoms", q["T"].num_atoms, "genes", q["T"].num_rotbonds + 6, flush=True)
            alone.append(_alone(gpu_ctx, q))
        served, _ = _serve(gpu_ctx, mix, slots=4)
    finally:
        gpu_ctx.set_team_width(0)
    for q, a, s in zip(mix, alone, served):
        assert s["status"] == 0, s["message"]
        _same(a, s, q["k"])


def test_admission_by_memory_keeps_bytes_under_the_budget(gpu_ctx, mix):
    ref, _ = _serve(gpu_ctx, mix, slots=3)
    sizes = sorted(r["device_bytes"] for r in ref)
    budget = sizes[-1] + sizes[-2]            # any two requests fit, not every three
    served, stats = _serve(gpu_ctx, mix, slots=8, budget=budget)
    print("stats", stats, "sizes", sizes, flush=True)
    assert stats["peak_bytes"] <= budget
    fit = 0
    while fit < len(sizes) and sum(sizes[:fit + 1]) <= budget:
        fit += 1
    assert 2 <= stats["peak_active"] <= fit
    for q, a, s in zip(mix, ref, served):
        assert s["status"] == 0, s["message"]
        _same(a, s, q["k"])


def test_served_results_match_the_oracle(gpu_ctx, mix):
    served, _ = _serve(gpu_ctx, mix, slots=4)
    for q, s in zip(mix, served):
        orc = O.Oracle(q["cg"], CLigand(q["T"]), q["ep"])
        helpers.check_energies(orc, s["best_genotypes"], s["best_energies"], edge_share=1.0)
        for r in range(NRUN):
            e64, ei, ea, sc = orc.energy64(s["best_genotypes"][r])
            tol = 1e-10 * max(1.0, sc)
            assert abs(s["best_f64"][r, 0] - e64) <= tol, (q["k"], r, s["best_f64"][r], e64)
            assert abs(s["best_f64"][r, 1] - ei) <= tol and abs(s["best_f64"][r, 2] - ea) <= tol


def test_refusals(gpu_ctx, mix):
    q = mix[0]
    with adgpu.Server(gpu_ctx, slots=2, memory_budget=4096) as srv:
        with pytest.raises(adgpu.AdgpuError, match="budget"):
            srv.submit(q["mid"], q["T"], q["sp"], q["init"], q["seeds"], ep=q["ep"])
    with adgpu.Server(gpu_ctx, slots=2) as srv:
        with pytest.raises(adgpu.AdgpuError, match="map id"):
            srv.submit(987654, q["T"], q["sp"], q["init"], q["seeds"], ep=q["ep"])
        t = srv.submit(q["mid"], q["T"], q["sp"], q["init"], q["seeds"], ep=q["ep"])
        r = srv.result(t)
        assert r["status"] == 0
        assert srv.next_done(0) is None       # result() already took it
        assert srv.stats()["completed"] == 1


def test_map_set_is_held_by_queued_and_running_requests(gpu_ctx, mix):
    q = mix[1]
    maps = synth.make_maps(q["types"], npts=40, seed=77)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    with adgpu.Server(gpu_ctx, slots=1) as srv:
        t = [srv.submit(mid, q["T"], q["sp"], q["init"], q["seeds"], ep=q["ep"]) for _ in range(2)]
        with pytest.raises(adgpu.AdgpuError, match="still used"):
            gpu_ctx.maps_release(mid)          # one request running, one queued
        assert all(srv.result(x)["status"] == 0 for x in t)
    gpu_ctx.maps_release(mid)                  # both finished: the map set can go


def test_one_collector_per_ticket(gpu_ctx, mix):
    """While one thread waits for a ticket, another thread's call for the
    same ticket is refused (ADGPU_ERR_STATE); the first still gets the result."""
    import threading
    import time
    q = mix[1]
    sp = search_params(num_runs=NRUN, pop_size=POP, max_evals=3000000, ls="ad", spacing=0.375)
    with adgpu.Server(gpu_ctx, slots=1) as srv:
        t = srv.submit(q["mid"], q["T"], sp, q["init"], q["seeds"], ep=q["ep"])
        got = {}
        th = threading.Thread(target=lambda: got.setdefault("r", srv.result(t)))
        th.start()
        time.sleep(0.05)
        if th.is_alive():                      # the first collector is waiting
            with pytest.raises(adgpu.AdgpuError, match="another thread"):
                srv.result(t, timeout_ms=0)
        th.join(timeout=120)
        assert not th.is_alive() and got["r"]["status"] == 0


def _device_free_bytes():
    """hipMemGetInfo's free bytes of the current device (the HIP runtime the
    library already loaded)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_admission_bounds_the_device_memory_slots_hold(gpu_ctx, mix):
    """Slots keep their buffers for the next request, so the budget must bound
    what idle slots hold too.  Six large requests (64 runs x 150, ADADELTA
    with its hand-off state: ~18 MB each) alternate with small ones on eight
    slots under a budget of 2.5 large requests.  Device memory measured with
    hipMemGetInfo, while the requests run and after they finish, stays within
    the budget (plus 10 % and 8 MB for the allocator's rounding); without the
    bound, every slot that once ran a large request would keep its 18 MB.
    Results are again the same as alone, with the team width fixed: left to
    the default, a server of one slot and one of eight may pick different
    widths for the same request (the searches in flight differ)."""
    q = mix[0]
    big_sp = search_params(num_runs=64, pop_size=150, max_evals=20000, ls="ad", spacing=0.375)
    g = q["cg"].s
    size_a = tuple((n - 1) * 0.375 for n in (g.size_x, g.size_y, g.size_z))
    big_init = O.initial_population((9, 1), 64, 150, size_a, 0.375)
    big_seeds = O.cmwc_uints((9, 1), 64 * 150 * 64)
    reqs = []
    for k in range(12):
        if k % 2 == 0:
            reqs.append(dict(q, sp=big_sp, init=big_init, seeds=big_seeds, autostop=None, k="big%d" % k))
        else:
            reqs.append(dict(mix[k % len(mix)], k="small%d" % k))
    gpu_ctx.set_team_width(32)
    try:
        ref, served, stats, peak, held, budget, big = _bounded_run(gpu_ctx, reqs)
    finally:
        gpu_ctx.set_team_width(0)
    print("budget", budget, "big", big, "peak", peak[0], "held after", held, "stats", stats, flush=True)
    slack = 0.10 * budget + (8 << 20)
    assert stats["peak_bytes"] <= budget and stats["bytes_in_use"] <= budget
    assert peak[0] <= budget + slack, (peak[0], budget)
    assert held <= budget + slack, (held, budget)
    assert 6 * big > budget + slack        # what slots would keep without the bound
    for r, a, s in zip(reqs, ref, served):
        assert s["status"] == 0, s["message"]
        _same(a, s, r["k"])


def _bounded_run(gpu_ctx, reqs):
    """The requests one at a time through one slot, then all at once through
    eight slots under a budget of 2.5 large requests, sampling the device
    memory in use."""
    import threading
    import time
    ref, _ = _serve(gpu_ctx, reqs, slots=1, one_at_a_time=True)      # also warms up the runtime
    big = max(r["device_bytes"] for r in ref)
    budget = int(2.5 * big)
    with adgpu.Server(gpu_ctx, slots=8, memory_budget=budget) as srv:
        time.sleep(0.05)
        base = _device_free_bytes()
        peak = [0]
        stop = threading.Event()

        def poll():
            while not stop.is_set():
                peak[0] = max(peak[0], base - _device_free_bytes())
                time.sleep(0.001)
        th = threading.Thread(target=poll)
        th.start()
        tickets = [srv.submit(r["mid"], r["T"], r["sp"], r["init"], r["seeds"], ep=r["ep"], autostop=r["autostop"],
                              rescore_f64=True) for r in reqs]
        served = [srv.result(t) for t in tickets]
        stop.set()
        th.join()
        held = base - _device_free_bytes()
        stats = srv.stats()
    return ref, served, stats, peak, held, budget, big
