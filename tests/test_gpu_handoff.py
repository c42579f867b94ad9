"""ADADELTA hand-off between teams (adgpu_set_ls_handoff) through the C ABI.

A search still running after the hand-off point, while unstarted searches of
the launch remain, is saved to device memory and resumed later by whichever
team is free.  Every search computes what it computes without hand-off, so a
docking is bit-identical with hand-off off, after one iteration (every search
that can be handed off is), or at the default 32: generation count, eval
total, final population, energies, per-run eval counters and LCG streams.
Ligand sizes and team widths cover 16-, 32- and 64-lane teams and one, two
and four atom chunks per lane, where the kernel variants (and their register
budgets) differ.  The reference runs each search to its end in its own block
(cuda/kernel_ad.cu:60-120); its results do not depend on the order in which
searches run, and neither do these.

Regrouping (adgpu_set_ls_regroup) moves, once a launch's searches have all
started, a search that is the last work of its wavefront to a team left idle
beside a busy one; it goes through the same saved records, so the dockings
are bit-identical with regrouping off, trying every iteration, or every 8,
and the tests check that searches did move.
"""
import numpy as np
import pytest

import adgpu
from adgpu import synth
from adgpu._abi import energy_params, search_params

pytestmark = pytest.mark.gpu

# 200 runs of 150 searches a generation (lsearch_rate 100 %) in two run
# groups: 15000 searches per launch, more than the resident teams of every
# case here (at most 8192 16-lane teams for half the device), so unstarted
# searches remain when the first searches reach the hand-off point.
NRUN, POP, EVALS = 200, 150, 300000


def _ligand(n_atoms, n_tors, seed):
    for t in range(n_tors, -1, -1):
        try:
            return synth.make_ligand(n_atoms, t, seed=seed)
        except RuntimeError:
            continue
    raise RuntimeError("no ligand")


def _initial(nrun, pop, maps, seed):
    rng = np.random.default_rng(seed)
    n = nrun * pop
    init = np.zeros((n, 64), np.float32)
    init[:, 0:3] = rng.random((n, 3)) * (np.array(maps.size, np.float64) - 1.0)
    init[:, 3:6] = rng.random((n, 3)) * np.array([360.0, 180.0, 360.0])
    init[:, 6:] = rng.random((n, 58)) * 360.0
    seeds = rng.integers(0, 2 ** 32, size=n * 64, dtype=np.uint64).astype(np.uint32)
    return init, seeds


def _dock(ctx, mid, lid, sp, init, seeds, handoff, regroup=8):
    ctx.set_ls_handoff(handoff)
    ctx.set_ls_regroup(regroup)
    job = ctx.job(mid, lid, sp, init, seeds)
    job.initpop()
    gens, total = job.run()
    st = job.download()
    n = job.ls_handoffs()
    stats = job.ls_stats()
    job.destroy()
    assert stats["handoffs"] == n
    return gens, total, st, n, stats


@pytest.mark.parametrize("atoms,tors,width", [(30, 6, 0), (20, 4, 16), (20, 4, 64), (70, 12, 32), (110, 20, 0),
                                              (110, 20, 64)])
def test_handoff_is_bit_identical(gpu_ctx, atoms, tors, width):
    lig = _ligand(atoms, tors, 100 + atoms)
    maps = synth.make_maps(lig.type_names, npts=48, spacing=0.375, seed=atoms)
    tables = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(tables, energy_params())
    sp = search_params(num_runs=NRUN, pop_size=POP, max_evals=EVALS, ls="ad", spacing=maps.spacing)
    init, seeds = _initial(NRUN, POP, maps, atoms)
    gpu_ctx.set_team_width(width)
    try:
        ref = _dock(gpu_ctx, mid, lid, sp, init, seeds, 0)
        assert ref[3] == 0 and ref[4]["regroups"] == 0
        steps = {}
        for h, rg in ((1, 8), (32, 0), (32, 8), (32, 1)):
            got = _dock(gpu_ctx, mid, lid, sp, init, seeds, h, rg)
            st = got[4]
            steps[(h, rg)] = st["wave_steps"]
            print("atoms", atoms, "width", width, "hand-off", h, "regroup", rg, "searches handed off", got[3],
                  "regrouped", st["regroups"], "wavefront steps", st["wave_steps"], flush=True)
            # the hand-off took place (the bit-identity below is not vacuous)
            assert got[3] > 0, (h, got[3])
            # searches regrouped, except with one team per wavefront (nothing to merge)
            # default widths of these jobs: 32 lanes at 30 atoms, 64 above 64 atoms
            tw = width if width else (64 if atoms > 64 else 32)
            assert st["team_width"] == tw, st
            if rg > 0 and tw < 64:
                assert st["regroups"] > 0, st
            else:
                assert st["regroups"] == 0, st
            assert got[0] == ref[0] and got[1] == ref[1], (h, rg, got[:2], ref[:2])
            for k in ref[2]:
                assert np.array_equal(got[2][k], ref[2][k]), (h, rg, k)
        tw = width if width else (64 if atoms > 64 else 32)
        if tw < 64:
            # the same evaluations in fewer wavefront steps: fewer idle teams beside busy ones
            assert steps[(32, 8)] < steps[(32, 0)], steps
    finally:
        gpu_ctx.set_ls_handoff(32)
        gpu_ctx.set_ls_regroup(-1)
        gpu_ctx.set_team_width(0)
        gpu_ctx.ligand_release(lid)
        gpu_ctx.maps_release(mid)


@pytest.mark.parametrize("width", [16, 32])
def test_regroup_with_few_searches(gpu_ctx, width):
    """Few searches per launch (8 runs x 3 of pop 50, lsearch_rate 6 %) against
    thousands of teams: most teams find nothing from the start and the seat
    tickets (twice the searches) run out; every regrouping interval still
    gives the same docking as none."""
    lig = _ligand(24, 5, 321)
    maps = synth.make_maps(lig.type_names, npts=48, spacing=0.375, seed=24)
    tables = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(tables, energy_params())
    sp = search_params(num_runs=8, pop_size=50, max_evals=60000, ls="ad", spacing=maps.spacing, lsearch_rate=6.0)
    init, seeds = _initial(8, 50, maps, 24)
    gpu_ctx.set_team_width(width)
    try:
        ref = _dock(gpu_ctx, mid, lid, sp, init, seeds, 32, 0)
        for rg in (1, 8):
            got = _dock(gpu_ctx, mid, lid, sp, init, seeds, 32, rg)
            print("few searches: width", width, "regroup", rg, got[4], flush=True)
            assert got[0] == ref[0] and got[1] == ref[1], (rg, got[:2], ref[:2])
            for k in ref[2]:
                assert np.array_equal(got[2][k], ref[2][k]), (rg, k)
    finally:
        gpu_ctx.set_ls_regroup(-1)
        gpu_ctx.set_team_width(0)
        gpu_ctx.ligand_release(lid)
        gpu_ctx.maps_release(mid)


@pytest.mark.parametrize("ls,nrun,want", [("sw", 20, 64), ("sw", 50, 32), ("ad", 20, 64), ("ad", 60, 64),
                                          ("ad", 100, 32), ("fire", 20, 32), ("adam", 20, 32)])
def test_default_team_width_for_small_ligands(gpu_ctx, ls, nrun, want):
    """Up to 32 atoms the default width follows the local search: Solis-Wets
    below 7000 and ADADELTA below 10000 searches in flight per generation run
    in 64-lane teams (their searches end unevenly, and few searches leave
    many half-idle wavefronts), FIRE and ADAM, and more searches, in 32-lane
    teams.  adgpu_job_ls_stats reports the width."""
    lig = _ligand(30, 6, 130)
    maps = synth.make_maps(lig.type_names, npts=40, spacing=0.375, seed=30)
    tables = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(tables, energy_params())
    try:
        sp = search_params(num_runs=nrun, pop_size=POP, max_evals=20000, ls=ls, spacing=maps.spacing)
        init, seeds = _initial(nrun, POP, maps, 30)
        job = gpu_ctx.job(mid, lid, sp, init, seeds)
        try:
            assert job.ls_stats()["team_width"] == want
        finally:
            job.destroy()
        gpu_ctx.set_team_width(32)   # a fixed width wins
        job = gpu_ctx.job(mid, lid, sp, init, seeds)
        try:
            assert job.ls_stats()["team_width"] == 32
        finally:
            job.destroy()
    finally:
        gpu_ctx.set_team_width(0)
        gpu_ctx.ligand_release(lid)
        gpu_ctx.maps_release(mid)


def test_handoff_argument(gpu_ctx):
    with pytest.raises(adgpu.AdgpuError):
        gpu_ctx.set_ls_handoff(-1)
    with pytest.raises(adgpu.AdgpuError):
        gpu_ctx.set_ls_regroup(-2)
