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
