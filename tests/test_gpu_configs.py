"""Parity on the other workload shapes of the benchmark set, through the C ABI:

* a ligand set (10-60 atoms, 0-20 torsions) against one resident 80^3 map
  set, FP32 and FP16 storage, ADAM local search;
* ligands of lognormal size (median 35, at most 90 atoms, 0-25 torsions)
  against three resident FP16 map sets of different sizes (60^3, 80^3,
  100^3), FIRE local search.

Bars as in test_gpu_parity.py: the energy bar of helpers.check_energies
(1e-4 of the sum of |terms|, plus the coordinate-rounding sensitivity), on
random poses and on the final populations of short LGA jobs (each downloaded
energy is the oracle's FP32 energy of the downloaded genotype)."""
import numpy as np
import pytest

import adgpu
import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params, search_params

pytestmark = pytest.mark.gpu


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


def _dock_and_check(ctx, mid, lid, orc, maps, tables, ls, seed):
    nrun, pop = 4, 30
    sp = search_params(num_runs=nrun, pop_size=pop, max_evals=15000, ls=ls, spacing=maps.spacing)
    init = O.initial_population((seed, 7), nrun, pop, maps.size_angstr, maps.spacing)
    job = ctx.job(mid, lid, sp, init, O.cmwc_uints((seed, 7), nrun * pop * 64))
    job.initpop()
    gens, total = job.run()
    st = job.download()
    job.destroy()
    assert gens > 0 and total >= nrun * 15000
    # a converged population may hold many copies of a pose on the grid edge
    helpers.check_energies(orc, st["pop"][:, :64], st["ene"], edge_share=1.0)


def test_ligand_set_on_one_80cube_map_set(gpu_ctx):
    rng = np.random.default_rng(2024)
    specs = [(int(rng.integers(10, 61)), int(rng.integers(0, 21)), 100 + k) for k in range(10)]
    ligs = [_ligand(a, t, s) for a, t, s in specs]
    types = sorted({t for lig in ligs for t in lig.type_names})
    maps = synth.make_maps(types, npts=80, seed=31)
    packed = maps.packed()
    ep = energy_params()
    for fmt in (adgpu.MAP_F32, adgpu.MAP_F16):
        mid = gpu_ctx.maps_upload(packed, maps.size, maps.num_map_atypes, maps.spacing, fmt)
        cg = CGrid(_rounded(packed, fmt), maps.size, maps.num_map_atypes, maps.spacing)
        for k, lig in enumerate(ligs):
            T = synth.ligand_tables(lig, maps.spacing, maps.type_names)
            lid = gpu_ctx.ligand_upload(T, ep)
            cl = CLigand(T)
            orc = O.Oracle(cg, cl, ep)
            case = helpers.SimpleNamespace(ng=T.num_genes, size=maps.size)
            poses = helpers.poses_for(case, 200, seed=900 + k)
            print("ligand", k, "atoms", len(lig.types), "fmt", fmt, flush=True)
            e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
            helpers.check_energies(orc, poses, e)
            if fmt == adgpu.MAP_F32 and k % 3 == 0:
                _dock_and_check(gpu_ctx, mid, lid, orc, maps, T, "adam", seed=k + 1)
            gpu_ctx.ligand_release(lid)
        gpu_ctx.maps_release(mid)


def test_lognormal_ligands_on_three_resident_fp16_map_sets(gpu_ctx):
    rng = np.random.default_rng(77)
    sizes = np.clip(np.round(rng.lognormal(np.log(35.0), 0.35, 9)), 8, 90).astype(int)
    ligs = [_ligand(int(a), int(rng.integers(0, 26)), 300 + k) for k, a in enumerate(sizes)]
    types = sorted({t for lig in ligs for t in lig.type_names})
    sets = []
    for npts, seed in ((60, 41), (80, 42), (100, 43)):
        maps = synth.make_maps(types, npts=npts, seed=seed)
        packed = maps.packed()
        mid = gpu_ctx.maps_upload(packed, maps.size, maps.num_map_atypes, maps.spacing, adgpu.MAP_F16)
        sets.append((maps, mid, CGrid(_rounded(packed, adgpu.MAP_F16), maps.size, maps.num_map_atypes,
                                      maps.spacing)))
    ep = energy_params()
    for k, lig in enumerate(ligs):
        maps, mid, cg = sets[k % 3]
        T = synth.ligand_tables(lig, maps.spacing, maps.type_names)
        lid = gpu_ctx.ligand_upload(T, ep)
        orc = O.Oracle(cg, CLigand(T), ep)
        case = helpers.SimpleNamespace(ng=T.num_genes, size=maps.size)
        poses = helpers.poses_for(case, 150, seed=1200 + k)
        print("ligand", k, "atoms", int(sizes[k]), "maps", maps.size, flush=True)
        e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
        helpers.check_energies(orc, poses, e)
        if k % 2 == 0:
            _dock_and_check(gpu_ctx, mid, lid, orc, maps, T, "fire", seed=k + 3)
        gpu_ctx.ligand_release(lid)
    for _, mid, _ in sets:
        gpu_ctx.maps_release(mid)


def test_screen_fp16_maps_give_the_fp32_median_best_energy(gpu_ctx):
    """The batched screen at its own scale: 16 seeded ligands (atoms 10-60,
    torsions 0-20) against one 80^3 map set, ADAM, nrun 50, pop 150, 2.5 M
    evaluations per run, all submitted at once to a server, once with FP32
    and once with FP16 map storage.  Each run's best pose is re-scored in FP64
    on the FP32 maps; a ligand's best energy is the lowest over its runs.  The
    median over the ligands of the best energy with FP16 maps equals the one
    with FP32 maps within 0.1 kcal/mol (FP16 storage at equal median best
    energy), and so does the mean paired difference."""
    import time
    ctx = gpu_ctx
    rng = np.random.default_rng(4711)
    specs = [(int(rng.integers(10, 61)), int(rng.integers(0, 21)), 700 + k) for k in range(16)]
    ligs = [_ligand(a, t, s) for a, t, s in specs]
    types = sorted({t for lig in ligs for t in lig.type_names})
    maps = synth.make_maps(types, npts=80, seed=33)
    packed = maps.packed()
    ep = energy_params()
    mids = {fmt: ctx.maps_upload(packed, maps.size, maps.num_map_atypes, maps.spacing, fmt)
            for fmt in (adgpu.MAP_F32, adgpu.MAP_F16)}
    nrun, pop = 50, 150
    tables = [synth.ligand_tables(lig, maps.spacing, maps.type_names) for lig in ligs]
    inits = [O.initial_population((k, 5), nrun, pop, maps.size_angstr, maps.spacing) for k in range(len(ligs))]
    seeds = [O.cmwc_uints((k, 5), nrun * pop * 64) for k in range(len(ligs))]
    best = {}
    for fmt, mid in mids.items():
        t0 = time.time()
        with adgpu.Server(ctx, slots=16) as srv:
            tickets = []
            for k, T in enumerate(tables):
                sp = search_params(num_runs=nrun, pop_size=pop, max_evals=2500000, ls="adam", spacing=maps.spacing)
                tickets.append(srv.submit(mid, T, sp, inits[k], seeds[k], ep=ep))
            res = [srv.result(t) for t in tickets]
        assert all(r["status"] == 0 for r in res)
        b = []
        for k, (T, r) in enumerate(zip(tables, res)):
            lid = ctx.ligand_upload(T, ep)
            tot, _, _ = ctx.rescore_f64(mids[adgpu.MAP_F32], lid, r["best_genotypes"])
            ctx.ligand_release(lid)
            b.append(float(np.min(tot)))
        best[fmt] = np.array(b)
        print("maps", fmt, "%.1f s" % (time.time() - t0), "best", np.round(best[fmt], 3), flush=True)
    for mid in mids.values():
        ctx.maps_release(mid)
    d = best[adgpu.MAP_F16] - best[adgpu.MAP_F32]
    med16, med32 = np.median(best[adgpu.MAP_F16]), np.median(best[adgpu.MAP_F32])
    print("median best FP16 %.3f FP32 %.3f; paired difference mean %.4f sd %.4f" % (med16, med32, d.mean(), d.std()))
    assert abs(med16 - med32) <= 0.1
    assert abs(d.mean()) <= 0.1


def test_flexres_on_126cube_bf16_maps(gpu_ctx):
    """The flexible-receptor configuration at its own map size: the ligand and
    two flexible residues whose tables the reference's host code built
    (tests/golden/ref_flex.npz) on 126^3 maps stored in BF16 (FP32
    interpolation and accumulation).  Against the oracle on the same rounded
    maps: the energy bar and the merged-gradient bar of test_gpu_parity.py, and
    the FP64 re-scoring within 1e-10 of the sum of |terms|.  The FP32 error
    against the FP64 oracle is reported per pose class."""
    f = helpers.load_ref_fixture("flex")
    T = helpers.tables_from_fixture(f)
    types = [str(t) for t in f["map_types"]]
    maps = synth.make_maps(types, npts=126, seed=int(f["map_seed"]))
    packed = maps.packed()
    u = packed.view(np.uint32).astype(np.uint64)
    rounded = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    del u
    ep = energy_params(coeff_elec=float(f["coeff_elec"]), coeff_desolv=float(f["coeff_desolv"]),
                       qasp=float(f["qasp"]), smooth=float(f["smooth"]), elec_min_distance=float(f["elec_min"]))
    cg = CGrid(rounded, maps.size, maps.num_map_atypes, maps.spacing)
    orc = O.Oracle(cg, CLigand(T), ep)
    mid = gpu_ctx.maps_upload(packed, maps.size, maps.num_map_atypes, maps.spacing, adgpu.MAP_BF16)
    lid = gpu_ctx.ligand_upload(T, ep)
    ng = T.num_genes
    # poses around the residues (the fixture's grid centre, 30 grid units)
    # and uniform over the 126^3 box
    near = helpers.poses_for(helpers.SimpleNamespace(ng=ng, size=(60, 60, 60)), 1500, seed=606, frac_central=1.0,
                             spread=14.0)
    far = synth.random_poses(300, ng, maps.size, seed=607)
    poses = np.concatenate([f["init_pop"], near, far]).astype(np.float32)
    e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
    helpers.check_energies(orc, poses, e)
    e2, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    helpers.check_energies(orc, poses, e2, e_ref=e)
    case = helpers.SimpleNamespace(spacing=maps.spacing)
    qv = np.array([helpers.gene_quantum(case, k) for k in range(ng)])
    well = [0, 1, 2] + list(range(6, ng))
    _, g32 = orc.energrad_batch(poses)
    d = np.abs(g[:, well] - g32[:, well])
    viol = d > 1e-4 * np.abs(g32[:, well]) + 4 * qv[well]
    assert viol.sum() <= 0.002 * viol.size, np.argwhere(viol)[:10]
    assert np.all(g[:, ng:] == 0.0)
    tot, inter, intra = gpu_ctx.rescore_f64(mid, lid, poses)
    e64, i64, s64 = orc.energy_batch(poses, f64=True)
    a64 = e64 - i64
    tol = 1e-10 * np.maximum(1.0, s64)
    assert np.all(np.abs(tot - e64) <= tol) and np.all(np.abs(inter - i64) <= tol)
    assert np.all(np.abs(intra - a64) <= tol + 1e-9 * np.abs(e64))
    rel = np.abs(e.astype(np.float64) - e64) / np.maximum(1.0, np.abs(e64))
    out = i64 >= 0.99 * 2.0 ** 24
    clash = ~out & (np.abs(e64) > 1000.0)
    normal = ~out & ~clash
    for name, m in (("normal", normal), ("clash", clash), ("out of grid", out)):
        if m.any():
            print("%s: %.1f %% of poses, |E32-E64|/max(1,|E64|) median %.1e p99 %.1e max %.1e"
                  % (name, 100 * m.mean(), np.median(rel[m]), np.percentile(rel[m], 99), rel[m].max()))
    assert normal.sum() >= 300 and np.percentile(rel[normal], 99) <= 1e-4   # 370 of 1864 poses
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)
