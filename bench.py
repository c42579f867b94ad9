#!/usr/bin/env python3
"""Docking throughput benchmark (one MI355X).

Workload (the LGA configuration of the headline metric, on one GPU): a seeded
30-atom ligand (types C, A, N, OA, HD; charges uniform in [-0.5, 0.5]; 6
rotatable torsions in a random tree) on seeded 60^3 FP32 maps at 0.375 A (200
Gaussian wells, repulsive cores clipped at 1e5); LGA with ADADELTA local
search, nrun = 100, pop 150, 2.5 M evaluations per run, no heuristics, no
AutoStop.  One step = one complete docking of the ligand (all 100 runs to the
evaluation budget), inputs resident in HBM.

Prints one JSON line: evals/s (value), ligands/hour, p50/p99 per-ligand
latency, the roofline of the dominant kernel (the local-search launch) and a
CPU baseline: the CPU-runnable configuration (Solis-Wets, nrun 20, pop 150) run
by the oracle's LGA on the host cores with a cut evaluation budget, timed to
completion.  "serving" adds the latency half of the metric under load: a
short run of the serving workload below (96 ligands: saturation throughput,
p50/p99 latency of Poisson arrivals at 70 % of it, and the plain one-at-a-time
FP32 form on the first 8 ligands); --no-serving leaves it out.

--workload poses evaluates 1e5 seeded random poses of the same ligand (energy
and gradient) and compares each FP32 energy with the FP64 re-scoring of the
pose on the GPU, by pose class.  --workload flexres docks the ligand plus two
flexible residues (tables built by the reference's host code, kept in
tests/golden/ref_flex.npz) on 126^3 BF16 maps with ADADELTA, nrun 100, and
reports the same per-pose comparison.  --workload screen measures a batched screen (seeded ligands of 10-60 atoms
and 0-20 torsions, one 80^3 map set, ADAM, nrun 50) through the same server
with FP32 and FP16 maps.  --workload serve measures the serving loop instead: seeded ligands of uneven
size (atoms lognormal with median 35 and at most 90, torsions 0-25) docked
against three resident FP16 map sets (60^3, 80^3, 100^3 at 0.375 A), FIRE
local search, 20 runs x 1 M evaluations per ligand, pop 150.  Three phases:
all ligands submitted at once (throughput at saturation), Poisson arrivals at
a fixed rate below it (p50/p99 latency from submission to result), and the
plain form -- FP32 maps, one ligand at a time through the job API -- on the
first ligands of the same list, with the median best energies of both forms.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "autodock-gpu_amd"))


HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, 8 TB/s

# Memory traffic (L2 <-> memory) of the local-search kernel per local-search
# evaluation, from two rocprofv3 --pmc runs of this workload (FETCH_SIZE, then
# WRITE_SIZE) scaled by the same counters read on a 1 GiB device-to-device
# copy kernel (FETCH_SIZE reports 0.5 GiB for the 1 GiB read: factor 2.0;
# WRITE_SIZE reports 1 GiB: factor 1.0).  Summary and commands:
# profiles/k_ls_ad_c1_traffic.md.  Keyed by (maps, nrun, pop, ls).
TRAFFIC_BYTES_PER_LS_EVAL = {("f32", 100, 150, "ad"): 494.3}

# Wavefront instructions of the local-search kernel per local-search
# evaluation (SQ_INSTS_VALU and SQ_INSTS_LDS summed over its dispatches, over
# the LS evaluations of those dispatches), from a rocprofv3 --pmc pass over
# this workload in its default four run groups (with regrouping):
# profiles/k_ls_ad_c1_counters.md.  With the measured evaluation rate they give
# the kernel's VALU issue rate, reported against the device's VALU issue rate
# on independent FP32 FMA chains measured live (adgpu_measure_valu_rate).
# Keyed by (maps, nrun, pop, ls).
LS_INSTR_PER_EVAL = {("f32", 100, 150, "ad"): (892.9, 73.2)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nrun", type=int, default=100)
    ap.add_argument("--pop", type=int, default=150)
    ap.add_argument("--evals", type=int, default=2500000)
    ap.add_argument("--ls", default="ad")
    ap.add_argument("--maps", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="lga: leave out the short serving-loop measurement (a child process running "
                         "--workload serve on --serving-ligands ligands) attached to the line")
    ap.add_argument("--serving-ligands", type=int, default=96)
    ap.add_argument("--cpu-evals", type=int, default=1000000,
                    help="evaluations per run of the CPU baseline's docking (of the 2.5 M)")
    ap.add_argument("--team-width", type=int, default=0, choices=[0, 16, 32, 64])
    ap.add_argument("--graph-block", type=int, default=0, choices=[0, 2, 4, 8],
                    help="generations per captured hipGraph (adgpu_set_graph_block; 0 = launched one by one)")
    # hardware queues of this process, set before the HIP runtime starts (HIP's
    # default is 4): with 8 the library runs the job's runs in four groups on
    # four streams, each on its own queue
    ap.add_argument("--hw-queues", type=int, default=8, choices=range(1, 33), metavar="1..32")
    ap.add_argument("--workload", default="lga", choices=["lga", "serve", "screen", "poses", "flexres"])
    ap.add_argument("--poses", type=int, default=100000)
    ap.add_argument("--screen-ligands", type=int, default=64)
    ap.add_argument("--autotune", action="store_true",
                    help="serve/screen: measure the team width per ligand size first (adgpu.autotune)")
    ap.add_argument("--team-table", default="",
                    help="serve/screen: fixed team widths per ligand size, e.g. 32:32,128:64 (max atoms:width)")
    ap.add_argument("--serve-ligands", type=int, default=300)
    ap.add_argument("--serve-distinct", type=int, default=0,
                    help="serve: distinct ligands to generate (0: one per request); requests beyond them "
                         "reuse the ligands in turn with their own seeds and initial populations")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-evals", type=int, default=1000000)
    ap.add_argument("--serve-nrun", type=int, default=20)
    ap.add_argument("--serve-qps", type=float, default=0.0,
                    help="arrival rate of the latency phase (0: 70 %% of the measured saturation rate)")
    ap.add_argument("--serve-plain", type=int, default=12,
                    help="ligands docked one at a time with FP32 maps for the plain-form baseline")
    ap.add_argument("--serve-phases", default="saturation,poisson,f32,plain",
                    help="comma-separated subset of saturation, poisson, f32 (the same requests on FP32 "
                         "maps), plain")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one 1 GiB device-to-device copy kernel (the reference point for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only measure the device's VALU issue rate on independent FP32 FMA chains "
                         "(adgpu_measure_valu_rate) and exit")
    return ap.parse_args()


def valu_calibration():
    """The VALU issue-rate calibration kernel alone (under rocprofv3 its
    counters give the counter readings at the peak rate)."""
    import adgpu
    ctx = adgpu.Context(0)
    r = [ctx.measure_valu_rate() for _ in range(3)]
    ctx.close()
    print(json.dumps({"valu_wave_instr_per_s": r}), flush=True)


def copy_calibration():
    """One plain copy kernel of known size: 2^28 floats (1 GiB) read and written.
    Under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE give the counters' scale."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)


def build_inputs():
    from adgpu import synth
    from adgpu._abi import energy_params
    lig = synth.make_ligand(30, 6, seed=7)
    maps = synth.make_maps(lig.type_names, npts=60, spacing=0.375, seed=11)
    tables = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    return lig, maps, tables, energy_params()


def initial_state(seed, nrun, pop, maps):
    """Initial population and LCG seeds, generated on the host from numpy's
    seeded generator: translations uniform in the box, orientations from
    uniform random quaternions converted to (phi, theta, alpha) as
    gen_initpop_and_reflig does (getparameters.cpp:2594-2619), torsions
    uniform in [0, 360), and one LCG state per work-item.  The reference
    draws the same quantities from its CMWC4096 generator seeded with the
    job seed; the docking does not depend on which generator made them."""
    rng = np.random.default_rng(seed)
    n = nrun * pop
    init = np.zeros((n, 64), np.float32)
    size = np.array(maps.size_angstr) / maps.spacing
    init[:, 0:3] = rng.random((n, 3)) * size
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    qw = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    qx = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    qy = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    qz = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    s = np.sqrt(np.maximum(1 - qw * qw, 1e-12))
    init[:, 3] = np.arctan2(qy / s, qx / s) / 0.0174533
    init[:, 4] = np.arccos(np.clip(qz / s, -1, 1)) / 0.0174533
    init[:, 5] = 2 * np.arccos(np.clip(qw, -1, 1)) / 0.0174533
    init[:, 6:63] = rng.random((n, 57)) * 360.0
    seeds = rng.integers(0, 2 ** 32, size=n * 64, dtype=np.uint64).astype(np.uint32)
    return init, seeds


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(maps, tables, ep, evals_per_run):
    """The CPU-runnable configuration of the metric on the host cores: the
    same ligand and maps, Solis-Wets, nrun 20, pop 150, no heuristics or
    AutoStop, run by the oracle's LGA (the FP32 restatement of the reference
    kernels: kind "port"; the reference's OpenCL CPU build needs an OpenCL CPU
    device these machines do not have).  The evaluation budget is cut to
    evals_per_run of the 2.5 M and the docking is timed to completion; each
    generation's (run, individual) work is split over the host threads (at
    most 16, the GPU's share of the host), as the OpenCL CPU runtime schedules
    work-groups.  A single-thread rate comes from one run on one thread."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    from adgpu._abi import CGrid, CLigand, search_params
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    cg = CGrid(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    cl = CLigand(tables)
    nrun, pop, full = 20, 150, 2500000
    sp = search_params(num_runs=nrun, pop_size=pop, max_evals=evals_per_run, ls="sw", spacing=maps.spacing)
    init, seeds = initial_state(99, nrun, pop, maps)
    job = O.OracleJob(cg, cl, ep, sp, init, seeds)
    t0 = time.time()
    gens, total = job.run(nthreads=threads)
    dt = time.time() - t0
    # one thread: one run with a tenth of the per-run budget
    sp1 = search_params(num_runs=1, pop_size=pop, max_evals=max(50000, evals_per_run // 10), ls="sw",
                        spacing=maps.spacing)
    j1 = O.OracleJob(cg, cl, ep, sp1, init[:pop], seeds[:pop * 64])
    t1 = time.time()
    _, total1 = j1.run(nthreads=1)
    dt1 = time.time() - t1
    full_s = dt * full / evals_per_run
    return dict(value=total / dt, unit="evals/s", cores=threads, kind="port",
                sample="Solis-Wets LGA, nrun %d, pop %d, the same ligand and maps; %d of %d evals per run, timed "
                       "to completion: %d generations, %d evals in %.1f s on %d threads"
                       % (nrun, pop, evals_per_run, full, gens, total, dt, threads),
                cpu_model=_cpu_model(), nproc=os.cpu_count(),
                single_thread_evals_per_s=total1 / dt1,
                ligands_per_hour=3600.0 / full_s,
                ligands_per_hour_note="full 2.5 M-eval docking time extrapolated linearly in evaluations "
                                      "from the timed sample (%.1f s per ligand)" % full_s)


def _pose_errors(ctx, mid, lid, poses):
    """FP32 energies and gradients of the poses (the kernels' arithmetic,
    adgpu_eval_poses) against the FP64 energies of the same poses
    (adgpu_rescore_f64), on the GPU.  Classes: out of the grid (some atom
    outside the maps: +2^24 each), clash (|E| > 1000 kcal/mol), normal.
    Error = |E32 - E64| / max(1, |E64|)."""
    import adgpu
    t0 = time.perf_counter()
    e32, _ = ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    dt = time.perf_counter() - t0
    tot, _, _ = ctx.rescore_f64(mid, lid, poses)
    rel = np.abs(e32.astype(np.float64) - tot) / np.maximum(1.0, np.abs(tot))
    out_grid = tot >= 0.5 * 2.0 ** 24
    clash = ~out_grid & (np.abs(tot) > 1000.0)
    normal = ~out_grid & ~clash
    res = {"poses": len(poses), "energrad_poses_per_s": len(poses) / dt,
           "energrad_note": "host buffers in and out (PCIe included)"}
    for name, m in (("out_of_grid", out_grid), ("clash", clash), ("normal", normal)):
        r = rel[m]
        res[name] = {"fraction": float(np.mean(m)),
                     "rel_err_median": float(np.median(r)) if r.size else None,
                     "rel_err_p99": float(np.percentile(r, 99)) if r.size else None,
                     "rel_err_max": float(np.max(r)) if r.size else None}
    return res


def poses(a):
    """Per-pose measurement on the LGA workload's ligand and maps."""
    import adgpu
    from adgpu import synth
    lig, maps, tables, ep = build_inputs()
    ctx = adgpu.Context(0)
    mid = ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing, adgpu.MAP_F32)
    lid = ctx.ligand_upload(tables, ep)
    ps = synth.random_poses(a.poses, tables.num_genes, maps.size, seed=2024)
    ctx.eval_poses(mid, lid, ps[:1000], adgpu.EVAL_ENERGRAD)   # warm-up
    out = {"metric": "per-pose FP32 energy+gradient against FP64 re-scoring", "unit": "poses/s",
           "config": {"workload": "%d seeded random poses, 30 atoms / 6 torsions, 60^3 FP32 maps" % a.poses},
           "data": "synthetic (seeded ligand, maps and poses)"}
    out.update(_pose_errors(ctx, mid, lid, ps))
    out["value"] = out["energrad_poses_per_s"]
    ctx.close()
    print(json.dumps(out), flush=True)


def _flex_tables():
    """Ligand (30 atoms / 6 torsions) plus two flexible residues (12 atoms /
    4 torsions each), tables as the reference's host code builds them, from
    tests/golden/ref_flex.npz; returns (tables, map types, energy params)."""
    from types import SimpleNamespace
    from adgpu._abi import energy_params
    f = np.load(os.path.join(HERE, "tests", "golden", "ref_flex.npz"))
    keys = ("charges", "atom_types", "atom_map", "ignore_inter", "ref_coords", "rotbond_moving", "rotbond_unit",
            "rotbond_atoms", "atom_rotbonds", "pairs", "atype_reqm", "vw_exp", "reqm_ab", "vw_ac", "vw_bd", "ds_s",
            "ds_v")
    t = SimpleNamespace(**{k: f[k] for k in keys})
    t.num_atoms, t.true_ligand_atoms = int(f["num_atoms"]), int(f["true_ligand_atoms"])
    t.num_rotbonds, t.num_atypes = int(f["num_rotbonds"]), int(f["num_atypes"])
    t.num_genes = t.num_rotbonds + 6
    ep = energy_params(coeff_elec=float(f["coeff_elec"]), coeff_desolv=float(f["coeff_desolv"]), qasp=float(f["qasp"]),
                       smooth=float(f["smooth"]), elec_min_distance=float(f["elec_min"]))
    return t, [str(x) for x in f["map_types"]], ep


def flexres(a):
    """Flexible-receptor docking: ligand + two flexible residues on 126^3
    BF16 maps, ADADELTA, nrun 100; throughput and per-pose FP32 vs FP64."""
    import adgpu
    from adgpu import synth
    tables, types, ep = _flex_tables()
    maps = synth.make_maps(types, npts=126, spacing=0.375, seed=11)
    ctx = adgpu.Context(0)
    mid = ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing, adgpu.MAP_BF16)
    lid = ctx.ligand_upload(tables, ep)
    sp = adgpu.search_params(num_runs=a.nrun, pop_size=a.pop, max_evals=a.evals, ls="ad", spacing=maps.spacing)
    lat, evals = [], []
    for step in range(a.warmup + a.steps):
        init, seeds = initial_state(3000 + step, a.nrun, a.pop, maps)
        job = ctx.job(mid, lid, sp, init, seeds)
        t0 = time.perf_counter()
        job.initpop()
        gens, total = job.run()
        dt = time.perf_counter() - t0
        job.destroy()
        print("step %d: %.3f s, %d generations, %d evals" % (step, dt, gens, total), file=sys.stderr, flush=True)
        if step >= a.warmup:
            lat.append(dt)
            evals.append(total)
    out = {"metric": "LGA energy evals/sec, ligand + 2 flexible residues", "value": float(np.sum(evals) / np.sum(lat)),
           "unit": "evals/s", "ligands_per_hour": 3600.0 / float(np.mean(lat)),
           "config": {"workload": "ligand 30 atoms / 6 torsions + 2 flexible residues (12 atoms / 4 torsions), "
                                  "126^3 BF16 maps at 0.375 A, ADADELTA, nrun %d, pop %d, %d evals per run"
                                  % (a.nrun, a.pop, a.evals)},
           "data": "synthetic (seeded maps; reference-built ligand and residue tables)"}
    ps = synth.random_poses(10000, tables.num_genes, maps.size, seed=77)
    out["per_pose"] = _pose_errors(ctx, mid, lid, ps)
    ctx.close()
    print(json.dumps(out), flush=True)


def _lognormal_ligand(k):
    """Ligand k of the serving workload: atoms lognormal (median 35, 8..90),
    torsions uniform in 0..25 (fewer when the seeded tree has fewer rotatable
    bonds)."""
    rng = np.random.default_rng(90000 + k)
    n_atoms = int(np.clip(np.round(rng.lognormal(np.log(35.0), 0.35)), 8, 90))
    return _ligand_with(n_atoms, int(rng.integers(0, 26)), 90000 + k)


def _uniform_ligand(k):
    """Ligand k of the screening workload: atoms uniform in 10..60, torsions
    uniform in 0..20."""
    rng = np.random.default_rng(60000 + k)
    return _ligand_with(int(rng.integers(10, 61)), int(rng.integers(0, 21)), 60000 + k)


def _ligand_with(n_atoms, n_tors, seed):
    from adgpu import synth
    for t in range(n_tors, -1, -1):
        try:
            return synth.make_ligand(n_atoms, t, seed=seed)
        except RuntimeError:
            continue
    raise RuntimeError("no ligand for seed %d" % seed)


def _prepare(make, n, map_points, nrun, pop):
    """n seeded ligands (made in a process pool, before any HIP call), map
    sets of the given sizes over their atom types, and each ligand's tables,
    initial population and LCG seeds; ligand k goes to map set k mod len."""
    import multiprocessing as mp
    from adgpu import synth
    t_prep = time.perf_counter()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        ligs = []
        for lig in pool.imap(make, range(n), chunksize=4):
            ligs.append(lig)
            if len(ligs) % 250 == 0:   # progress of long preparations
                print("  %d ligands made, %.1f s" % (len(ligs), time.perf_counter() - t_prep), file=sys.stderr,
                      flush=True)
    types = sorted({t for lig in ligs for t in lig.type_names})
    map_sets = [synth.make_maps(types, npts=p, spacing=0.375, seed=70 + p) for p in map_points]
    reqs = []
    for k, lig in enumerate(ligs):
        ms = map_sets[k % len(map_sets)]
        tables = synth.ligand_tables(lig, ms.spacing, ms.type_names)
        init, seeds = initial_state(5000 + k, nrun, pop, ms)
        reqs.append((k % len(map_sets), tables, init, seeds))
    print("prepared %d ligands in %.1f s" % (len(reqs), time.perf_counter() - t_prep), file=sys.stderr, flush=True)
    return map_sets, reqs


def _run_served(ctx, slots, reqs, mids, sp, ep, qps=0.0, seed=0):
    """Submits the requests to a server (all at once, or Poisson arrivals at
    qps per second) and collects every result; returns (wall s, results,
    server statistics with the host CPU seconds used)."""
    import adgpu
    rng = np.random.default_rng(seed)
    with adgpu.Server(ctx, slots=slots) as srv:
        cpu0 = os.times()
        t0 = time.perf_counter()
        tickets = []
        next_t = 0.0
        for m, tables, init, seeds in reqs:
            if qps > 0:
                next_t += rng.exponential(1.0 / qps)
                delay = t0 + next_t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            tickets.append(srv.submit(mids[m], tables, sp, init, seeds, ep=ep))
        res = []
        for k, t in enumerate(tickets):
            res.append(srv.result(t))
            if (k + 1) % 1000 == 0:   # progress of long runs (10,000-ligand serving)
                print("  %d results, %.1f s" % (k + 1, time.perf_counter() - t0), file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        cpu1 = os.times()
        st = srv.stats()
        st["host_cpu_s"] = (cpu1.user - cpu0.user) + (cpu1.system - cpu0.system)
    bad = [r["message"] for r in res if r["status"] != 0]
    if bad:
        raise RuntimeError("served requests failed: %s" % bad[:3])
    return wall, res, st


def _saturation(ctx, slots, reqs, mids, sp, ep, label):
    wall, res, st = _run_served(ctx, slots, reqs, mids, sp, ep)
    evals = float(sum(r["total_evals"] for r in res))
    lph = 3600.0 * len(res) / wall
    print("%s: %d ligands in %.2f s, %.0f ligands/h, %.1f M evals/s, peak active %d, host CPU %.1f s" %
          (label, len(res), wall, lph, evals / wall / 1e6, st["peak_active"], st["host_cpu_s"]),
          file=sys.stderr, flush=True)
    return res, {"wall_s": wall, "ligands_per_hour": lph, "evals_per_s": evals / wall,
                 "host_cpu_s": st["host_cpu_s"], "peak_active": st["peak_active"],
                 "slot_reallocations": st["slot_reallocations"]}


def _plain(ctx, reqs, mids32, sp, ep, out, served_value):
    """The plain form: FP32 maps, one ligand at a time through the job API,
    with the default run groups and team widths, and with one stream (the
    reference's loop structure); per-ligand FP64 best energies."""
    ctx.set_team_table([])
    for label, groups in (("plain_f32_one_at_a_time", None), ("plain_f32_one_stream", "1")):
        if groups is None:
            os.environ.pop("ADGPU_RUN_GROUPS", None)
        else:
            os.environ["ADGPU_RUN_GROUPS"] = groups
        plain_t, plain_best = [], []
        for m, tables, init, seeds in reqs:
            lid = ctx.ligand_upload(tables, ep)
            t0 = time.perf_counter()
            job = ctx.job(mids32[m], lid, sp, init, seeds)
            job.initpop()
            job.run()
            b = job.best_f64()
            plain_t.append(time.perf_counter() - t0)
            plain_best.append(float(np.min(b["total"])))
            job.destroy()
            ctx.ligand_release(lid)
        lph = 3600.0 / float(np.mean(plain_t))
        print("%s: %.3f s per ligand, %.0f ligands/h" % (label, np.mean(plain_t), lph), file=sys.stderr, flush=True)
        out[label] = {"ligands": len(plain_t), "s_per_ligand": float(np.mean(plain_t)), "ligands_per_hour": lph,
                      "latency_s": {"p50": float(np.percentile(plain_t, 50)),
                                    "p99": float(np.percentile(plain_t, 99))},
                      "median_best_energy_f64": float(np.median(plain_best))}
        if served_value:
            out["speedup_vs_" + label.replace("plain_f32_", "")] = served_value / lph
    os.environ.pop("ADGPU_RUN_GROUPS", None)


def _best_on_f32(ctx, reqs, res, mids32, ep):
    """Per ligand the lowest FP64 energy of its runs' best poses re-scored on
    the FP32 maps."""
    best = []
    for (m, tables, _, _), r in zip(reqs, res):
        lid = ctx.ligand_upload(tables, ep)
        tot, _, _ = ctx.rescore_f64(mids32[m], lid, r["best_genotypes"])
        best.append(float(np.min(tot)))
        ctx.ligand_release(lid)
    return best


def _autotune(ctx, a, ls):
    """--team-table, or with --autotune the team-width table measured for
    this local search with the device shared by many requests
    (adgpu.autotune, mode "served"), installed in the context for the served
    phases; else None (built-in rule)."""
    if a.team_table:
        table = [tuple(int(x) for x in e.split(":")) for e in a.team_table.split(",")]
        ctx.set_team_table(table)
        return {"table": table, "seconds": 0.0, "evals_per_s": None}
    if not a.autotune:
        return None
    from adgpu import autotune
    t0 = time.perf_counter()
    table, rates = autotune.tune(ctx, ls=ls, mode="served", nrun=20, evals=100000, reps=1)
    dt = time.perf_counter() - t0
    print("autotune (%s, %.1f s): %s" % (ls, dt, table), file=sys.stderr, flush=True)
    return {"table": table, "seconds": dt,
            "evals_per_s": {str(n): {str(w): v for w, v in r.items()} for n, r in rates.items()}}


def serve(a):
    """The serving-loop measurement (module docstring)."""
    from adgpu._abi import energy_params
    distinct = a.serve_distinct if 0 < a.serve_distinct < a.serve_ligands else a.serve_ligands
    map_sets, reqs = _prepare(_lognormal_ligand, distinct, (60, 80, 100), a.serve_nrun, a.pop)
    for k in range(distinct, a.serve_ligands):   # the same ligands again, own seeds and initial populations
        m, tables, _, _ = reqs[k % distinct]
        init, seeds = initial_state(5000 + k, a.serve_nrun, a.pop, map_sets[m])
        reqs.append((m, tables, init, seeds))
    ep = energy_params()
    # the workers wait for the device in the kernel, not by spinning on host cores
    os.environ.setdefault("ADGPU_HOST_WAIT", "block")
    import adgpu
    ctx = adgpu.Context(0)
    if a.graph_block:
        ctx.set_graph_block(a.graph_block)
    sp = adgpu.search_params(num_runs=a.serve_nrun, pop_size=a.pop, max_evals=a.serve_evals, ls="fire",
                             spacing=0.375)
    mids16 = [ctx.maps_upload(m.packed(), m.size, m.num_map_atypes, m.spacing, adgpu.MAP_F16) for m in map_sets]
    phases = set(a.serve_phases.split(","))
    tuned = _autotune(ctx, a, "fire")
    out = {
        "metric": "ligands/hour of the serving loop (FP16 maps, concurrent slots); p50/p99 ligand latency",
        "value": None,
        "unit": "ligands/h",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 (FP16 map storage)",
        "data": "synthetic (seeded ligands and maps generated in-process)",
        "config": {"workload": "serving loop: %d ligands%s (atoms lognormal median 35, max 90; torsions 0-25) on "
                               "three resident FP16 map sets (60^3, 80^3, 100^3), FIRE, nrun %d x %d evals, pop %d"
                               % (len(reqs), (" (%d distinct)" % distinct) if distinct < len(reqs) else "",
                                  a.serve_nrun, a.serve_evals, a.pop),
                   "slots": a.serve_slots, "hw_queues": a.hw_queues,
                   "host_wait": os.environ.get("ADGPU_HOST_WAIT")},
        "team_table": tuned,
    }
    _run_served(ctx, a.serve_slots, reqs[:a.serve_slots], mids16, sp, ep)   # warm-up
    res = []
    if "saturation" in phases or "poisson" in phases:
        res, out["saturation"] = _saturation(ctx, a.serve_slots, reqs, mids16, sp, ep, "saturation")
        out["value"] = out["saturation"]["ligands_per_hour"]
    if "poisson" in phases:
        qps = a.serve_qps if a.serve_qps > 0 else 0.7 * out["saturation"]["ligands_per_hour"] / 3600.0
        wall_q, res_q, st_q = _run_served(ctx, a.serve_slots, reqs, mids16, sp, ep, qps, 3)
        lat = np.array([(r["t_done_ms"] - r["t_submit_ms"]) / 1000.0 for r in res_q])
        queue = np.array([(r["t_admit_ms"] - r["t_submit_ms"]) / 1000.0 for r in res_q])
        print("poisson %.2f/s: p50 %.2f s, p99 %.2f s" % (qps, np.percentile(lat, 50), np.percentile(lat, 99)),
              file=sys.stderr, flush=True)
        out["poisson"] = {"qps": qps, "wall_s": wall_q, "ligands_per_hour": 3600.0 * len(res_q) / wall_q,
                          "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                          "queue_s": {"p50": float(np.percentile(queue, 50)), "p99": float(np.percentile(queue, 99))},
                          "peak_active": st_q["peak_active"]}
    mids32 = [ctx.maps_upload(m.packed(), m.size, m.num_map_atypes, m.spacing, adgpu.MAP_F32) for m in map_sets]
    if "f32" in phases and res:
        # the same requests on FP32 maps: what FP16 storage changes in the results
        res32, out["saturation_f32"] = _saturation(ctx, a.serve_slots, reqs, mids32, sp, ep, "saturation FP32")
        b16 = np.array(_best_on_f32(ctx, reqs, res, mids32, ep))
        b32 = np.array(_best_on_f32(ctx, reqs, res32, mids32, ep))
        d = b16 - b32
        out["best_energy_f64"] = {"median_f16": float(np.median(b16)), "median_f32": float(np.median(b32)),
                                  "paired_diff_mean": float(np.mean(d)),
                                  "paired_diff_sem": float(np.std(d, ddof=1) / np.sqrt(len(d))),
                                  "ligands": len(d),
                                  "note": "per ligand the lowest FP64 energy over its runs' best poses, "
                                          "re-scored on the FP32 maps"}
    if "plain" in phases and a.serve_plain > 0:
        sub = reqs[:a.serve_plain]
        _plain(ctx, sub, mids32, sp, ep, out, out["value"])
        if res:
            served_best = _best_on_f32(ctx, sub, res, mids32, ep)
            out["median_best_energy_f64"] = {
                "served_f16": float(np.median(served_best)),
                "plain_f32": out["plain_f32_one_at_a_time"]["median_best_energy_f64"],
                "ligands": len(served_best),
                "note": "per ligand the lowest FP64 energy over its runs' best poses, re-scored on the FP32 maps"}
    ctx.close()
    print(json.dumps(out), flush=True)


def screen(a):
    """Batched screen against one receptor: seeded ligands (atoms uniform in
    10..60, torsions 0..20) on one 80^3 map set, ADAM, nrun 50, pop 150,
    `--evals` per run, all submitted at once to the server, with FP32 and
    with FP16 maps; the plain form (FP32, one ligand at a time) on the first
    ligands, and the paired FP16 - FP32 difference of each ligand's best
    energy (FP64, re-scored on the FP32 maps)."""
    from adgpu._abi import energy_params
    map_sets, reqs = _prepare(_uniform_ligand, a.screen_ligands, (80,), 50, a.pop)
    ep = energy_params()
    os.environ.setdefault("ADGPU_HOST_WAIT", "block")
    import adgpu
    ctx = adgpu.Context(0)
    if a.graph_block:
        ctx.set_graph_block(a.graph_block)
    sp = adgpu.search_params(num_runs=50, pop_size=a.pop, max_evals=a.evals, ls="adam", spacing=0.375)
    mids32 = [ctx.maps_upload(m.packed(), m.size, m.num_map_atypes, m.spacing, adgpu.MAP_F32) for m in map_sets]
    mids16 = [ctx.maps_upload(m.packed(), m.size, m.num_map_atypes, m.spacing, adgpu.MAP_F16) for m in map_sets]
    tuned = _autotune(ctx, a, "adam")
    out = {
        "team_table": tuned,
        "metric": "ligands/hour of a batched screen against one receptor",
        "value": None,
        "unit": "ligands/h",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 (FP32 and FP16 map storage)",
        "data": "synthetic (seeded ligands and maps generated in-process)",
        "config": {"workload": "batched screen: %d ligands (atoms 10-60, torsions 0-20) on one 80^3 map set, "
                               "ADAM, nrun 50 x %d evals, pop %d" % (len(reqs), a.evals, a.pop),
                   "slots": a.serve_slots, "hw_queues": a.hw_queues,
                   "host_wait": os.environ.get("ADGPU_HOST_WAIT")},
    }
    _run_served(ctx, a.serve_slots, reqs[:a.serve_slots], mids16, sp, ep)   # warm-up
    res32, out["served_f32"] = _saturation(ctx, a.serve_slots, reqs, mids32, sp, ep, "served FP32")
    res16, out["served_f16"] = _saturation(ctx, a.serve_slots, reqs, mids16, sp, ep, "served FP16")
    out["value"] = out["served_f16"]["ligands_per_hour"]
    b32 = _best_on_f32(ctx, reqs, res32, mids32, ep)
    b16 = _best_on_f32(ctx, reqs, res16, mids32, ep)
    d = np.array(b16) - np.array(b32)
    out["best_energy_f64"] = {"median_f32": float(np.median(b32)), "median_f16": float(np.median(b16)),
                              "paired_diff_mean": float(np.mean(d)),
                              "paired_diff_sem": float(np.std(d, ddof=1) / np.sqrt(len(d))) if len(d) > 1 else None,
                              "ligands": len(d)}
    if a.serve_plain > 0:
        _plain(ctx, reqs[:a.serve_plain], mids32, sp, ep, out, out["value"])
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.valu_calibration:
        valu_calibration()
        return
    os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    if a.gpus != 1:
        print("bench.py measures one GPU: the docking path shards by ligand or by run with no "
              "exchange between GPUs during the search, so it runs one process per GPU; "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "serve":
        serve(a)
        return
    if a.workload == "screen":
        screen(a)
        return
    if a.workload == "poses":
        poses(a)
        return
    if a.workload == "flexres":
        flexres(a)
        return
    import adgpu
    lig, maps, tables, ep = build_inputs()
    fmt = {"f32": adgpu.MAP_F32, "f16": adgpu.MAP_F16, "bf16": adgpu.MAP_BF16}[a.maps]
    ctx = adgpu.Context(0)
    if a.team_width:
        ctx.set_team_width(a.team_width)
    if a.graph_block:
        ctx.set_graph_block(a.graph_block)
    mid = ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing, fmt)
    lid = ctx.ligand_upload(tables, ep)
    sp = adgpu.search_params(num_runs=a.nrun, pop_size=a.pop, max_evals=a.evals, ls=a.ls, spacing=maps.spacing)
    n_inter = int(tables.num_atoms - int(np.sum(tables.ignore_inter)))
    bytes_per_eval = n_inter * 3 * 8 * {"f32": 4, "f16": 2, "bf16": 2}[a.maps]

    lat, evals_list, gens_list, ls_ms, ls_n, gen_ms, ls_busy = [], [], [], [], [], [], []
    ls_evals_list, ls_stats_list = [], []
    for step in range(a.warmup + a.steps):
        init, seeds = initial_state(1000 + step, a.nrun, a.pop, maps)
        job = ctx.job(mid, lid, sp, init, seeds)      # allocations outside the timed region
        t0 = time.perf_counter()
        job.initpop()
        gens, total = job.run()
        dt = time.perf_counter() - t0
        ms_ls, n_ls = job.kernel_time(1)
        busy_ls, _ = job.kernel_busy(1)
        ms_gen, n_gen = job.kernel_time(0)
        lss = job.ls_stats()
        st = job.download()
        job.destroy()
        if step >= a.warmup:
            lat.append(dt)
            evals_list.append(total)
            gens_list.append(gens)
            ls_ms.append(ms_ls)
            ls_n.append(n_ls)
            ls_busy.append(busy_ls)
            gen_ms.append(ms_gen)
            # LS evaluations: total minus K1 (nrun*pop) and K4 (nrun*(pop-1) per generation)
            ls_evals_list.append(total - a.nrun * a.pop - gens * a.nrun * (a.pop - 1))
            ls_stats_list.append(lss)
        print("step %d: %.3f s, %d generations, %d evals, best %.3f" %
              (step, dt, gens, total, float(st["ene"].min())), file=sys.stderr, flush=True)

    t = float(np.sum(lat))
    evals = float(np.sum(evals_list))
    evals_per_s = evals / t
    lat_arr = np.array(lat)
    # roofline of the dominant kernel: algorithmic map-gather bytes per LS launch
    # divided by the LS launch's duration, from HIP events on the launches'
    # streams.  The run groups' launches overlap, so the duration per launch is
    # the device time during which any LS launch ran (the union of the launch
    # intervals) divided by the number of launches; each launch's own
    # event-to-event time, which includes co-running launches, is reported as
    # ls_avg_ms.
    tot_ls_ms = float(np.sum(np.array(ls_ms) * np.array(ls_n)))
    ls_evals = float(np.sum(ls_evals_list))
    n_launch = max(1, int(np.sum(ls_n)))
    avg_ls_ms = tot_ls_ms / n_launch
    eff_ls_ms = float(np.sum(ls_busy)) / n_launch
    bytes_per_launch = ls_evals * bytes_per_eval / n_launch
    # launches inside captured graph blocks are not timed (--graph-block): the
    # per-launch figure then needs a run without blocks
    achieved = bytes_per_launch / (eff_ls_ms * 1e-3) / 1e9 if a.graph_block == 0 and eff_ls_ms > 0 else None
    tkey = (a.maps, a.nrun, a.pop, a.ls)
    traffic = (TRAFFIC_BYTES_PER_LS_EVAL[tkey] * ls_evals / max(1, int(np.sum(ls_n)))
               if tkey in TRAFFIC_BYTES_PER_LS_EVAL else None)
    # the run groups' launches overlap on the device, so per-launch durations
    # include co-running work; the whole-job rate over the generation loop is
    # reported beside it
    aggregate = ls_evals * bytes_per_eval / t / 1e9
    out = {
        "metric": "LGA energy evals/sec + ligands/hour at 2.5M evals; p50/p99 ligand latency",
        "value": evals_per_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000.0 * t / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ligand and maps generated in-process)",
        "config": {"workload": "LGA %s, 1 ligand of 30 atoms / 6 torsions, 60^3 maps at 0.375 A (%s), "
                               "nrun %d, pop %d, %d evals per run"
                               % ({"ad": "ADADELTA", "sw": "Solis-Wets", "adam": "ADAM", "fire": "FIRE"}.get(a.ls, a.ls),
                                  a.maps.upper(), a.nrun, a.pop, a.evals),
                   "nrun": a.nrun, "pop_size": a.pop, "evals_per_run": a.evals, "ls": a.ls, "maps": a.maps,
                   "hw_queues": a.hw_queues, "graph_block": a.graph_block},
        "ligands_per_hour": 3600.0 / float(np.mean(lat_arr)),
        "latency_s": {"p50": float(np.percentile(lat_arr, 50)), "p99": float(np.percentile(lat_arr, 99))},
        "generations_per_ligand": float(np.mean(gens_list)),
        "kernels": {"ls_avg_ms": avg_ls_ms, "ls_busy_ms_per_launch": eff_ls_ms, "gen_avg_ms": float(np.mean(gen_ms)),
                    "ls_launches": int(np.sum(ls_n)), "ls_busy_share_of_wall": float(np.sum(ls_busy)) / 1000.0 / t,
                    # ADADELTA: searches handed off and regrouped per docking; every local
                    # search: the share of team evaluation steps idle teams spent beside
                    # busy ones
                    "ls_handoffs": float(np.mean([x["handoffs"] for x in ls_stats_list])),
                    "ls_regroups": float(np.mean([x["regroups"] for x in ls_stats_list])),
                    "idle_team_share": (1.0 - ls_evals / sum(x["wave_steps"] * 64.0 / x["team_width"]
                                                             for x in ls_stats_list)
                                        if sum(x["wave_steps"] for x in ls_stats_list) else None),
                    "team_width": int(ls_stats_list[0]["team_width"]) if ls_stats_list else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "duration": "LS busy time (union of launch intervals) / launches",
                     "bytes_per_eval": bytes_per_eval, "aggregate_achieved": aggregate,
                     "aggregate_frac": aggregate / HBM_PEAK_GBS},
    }
    if tkey in LS_INSTR_PER_EVAL and achieved:
        # VALU issue rate of the LS kernel while it runs (LS busy time) against
        # the calibrated rate of full-rate FP32 VALU instructions
        valu_per_eval, lds_per_eval = LS_INSTR_PER_EVAL[tkey]
        peak = ctx.measure_valu_rate()
        rate = valu_per_eval * ls_evals / (float(np.sum(ls_busy)) * 1e-3)
        out["roofline"]["valu"] = {"achieved": rate, "peak": peak, "unit": "wave-instr/s", "frac": rate / peak,
                                   "valu_instr_per_eval": valu_per_eval, "lds_instr_per_eval": lds_per_eval,
                                   "per_cu_per_clock_at_2p4ghz": rate / (256 * 2.4e9),
                                   "source": "instructions per evaluation from dispatch counters, four run "
                                             "groups with regrouping (dispatches serialized); peak = FP32 FMA "
                                             "chains measured live by adgpu_measure_valu_rate; frac above 1: the "
                                             "overlapping launches issue this mix faster than the FMA loop (its "
                                             "v_mov and DPP moves take a fraction of an FMA's issue slot)"}
        out["roofline"]["note"] = "the launch is VALU-issue-bound (valu); HBM traffic is 0.17 of the " \
                                  "algorithmic bytes (the map set stays in L2 and the Infinity Cache)"
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(maps, tables, ep, a.cpu_evals)
    ctx.close()
    if not a.no_serving:
        out["serving"] = serving_summary(a)
    print(json.dumps(out), flush=True)


def serving_summary(a):
    """The per-ligand latency part of the metric under load: the serving loop
    (--workload serve) on a.serving_ligands uneven ligands, run in a child
    process (its host threads wait for the device by blocking, which must be
    set before the HIP runtime starts): saturation throughput, p50/p99 latency
    of Poisson arrivals at 70 % of it, and the plain one-at-a-time FP32 form on
    the first ligands."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "serve", "--serve-ligands",
           str(a.serving_ligands), "--serve-phases", "saturation,poisson,plain", "--serve-plain", "8",
           "--hw-queues", str(a.hw_queues)]
    t0 = time.perf_counter()
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError("serving measurement failed: %s" % p.stderr[-2000:])
    d = json.loads(p.stdout.strip().splitlines()[-1])
    return {"workload": d["config"]["workload"], "slots": d["config"]["slots"],
            "ligands_per_hour": d["value"], "evals_per_s": d["saturation"]["evals_per_s"],
            "poisson": d.get("poisson"),
            "plain_f32_ligands_per_hour": d.get("plain_f32_one_at_a_time", {}).get("ligands_per_hour"),
            "speedup_vs_one_at_a_time": d.get("speedup_vs_one_at_a_time"),
            "median_best_energy_f64": d.get("median_best_energy_f64"),
            "seconds": time.perf_counter() - t0}


if __name__ == "__main__":
    main()
