"""Per-size team widths (adgpu_set_team_table) and the autotuner that fills
them, through the C ABI.

A ligand whose size class maps to width w docks exactly as with width w set
for the whole context (bit-identical populations, energies and eval counts),
so the table changes only the launch shape, never the arithmetic of a
width; the default rule (32 lanes up to 32 atoms, 64 above) is the same as
setting that width.  The autotuner returns a valid table from measured
local-search jobs and installs it."""
import numpy as np
import pytest

import adgpu
import oracle as O
from adgpu import autotune, synth
from adgpu._abi import energy_params, search_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lig40(gpu_ctx):
    lig = synth.make_ligand(40, 8, seed=41)
    maps = synth.make_maps(lig.type_names, npts=48, seed=42)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(synth.ligand_tables(lig, maps.spacing, maps.type_names), energy_params())
    yield mid, lid, maps
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)


def _dock(ctx, mid, lid, maps, ls):
    sp = search_params(num_runs=4, pop_size=40, max_evals=15000, ls=ls, spacing=maps.spacing)
    init = O.initial_population((8, 8), 4, 40, maps.size_angstr, maps.spacing)
    job = ctx.job(mid, lid, sp, init, O.cmwc_uints((8, 8), 4 * 40 * 64))
    job.initpop()
    gens, total = job.run()
    st = job.download()
    job.destroy()
    return gens, total, st


@pytest.mark.parametrize("ls", ["ad", "fire"])
def test_table_entry_equals_the_width_it_names(gpu_ctx, lig40, ls):
    mid, lid, maps = lig40
    runs = {}
    try:
        for key, width, table in (("w32", 32, []), ("t32", 0, [(24, 64), (48, 32)]), ("w64", 64, []),
                                  ("t64", 0, [(48, 64)]), ("default", 0, [])):
            gpu_ctx.set_team_width(width)
            gpu_ctx.set_team_table(table)
            print("team", key, flush=True)
            runs[key] = _dock(gpu_ctx, mid, lid, maps, ls)
    finally:
        gpu_ctx.set_team_width(0)
        gpu_ctx.set_team_table([])
    for a, b in (("w32", "t32"), ("w64", "t64"), ("w64", "default")):
        ga, ta, sa = runs[a]
        gb, tb, sb = runs[b]
        assert (ga, ta) == (gb, tb), (a, b)
        for k in ("pop", "ene", "evals_runs", "prng"):
            assert np.array_equal(sa[k], sb[k]), (a, b, k)


def test_table_validation(gpu_ctx):
    for bad in ([(40, 32), (30, 64)], [(40, 16)], [(20, 48)], [(0, 32)], [(300, 64)], [(k + 1, 32) for k in range(17)]):
        with pytest.raises(adgpu.AdgpuError, match="team table"):
            gpu_ctx.set_team_table(bad)
    gpu_ctx.set_team_table([(32, 16), (64, 64)])
    gpu_ctx.set_team_table([])


@pytest.mark.parametrize("mode", ["single", "served"])
def test_autotune_installs_a_measured_table(gpu_ctx, mode):
    try:
        table, rates = autotune.tune(gpu_ctx, sizes=(24, 40), nrun=8, pop=40, evals=20000, reps=1, mode=mode,
                                     requests=6, slots=3)
    finally:
        gpu_ctx.set_team_table([])
    print("table", table, "rates", rates, flush=True)
    assert set(rates) == {24, 40}
    assert set(rates[24]) == {16, 32, 64} and set(rates[40]) == {32, 64}
    assert all(v > 0 for r in rates.values() for v in r.values())
    assert table[-1][0] == 40 and all(w in (16, 32, 64) for _, w in table)
    assert [m for m, _ in table] == sorted(m for m, _ in table)
