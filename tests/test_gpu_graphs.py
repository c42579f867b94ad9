"""Graph-captured generations (adgpu_set_graph_block) through the C ABI.

With a block size B, each run group captures B consecutive generations (K4,
the local search, K2 and the eval-total copy) into a hipGraph the first time
the block's ring slot comes round, and replays it whenever the stop rule has
proven that all B generations run; the last generations before the stop, and
blocks that would cross an AutoStop test or a change of local search, are
launched one by one.  A docking is bit-identical either way: generation
count, eval total, final populations, energies, per-run eval counters and
LCG states, for every local search, with two run groups, with AutoStop, and
with ADADELTA falling back to Solis-Wets for its first generations.  Each job
replays its captured blocks many times, with the host's own eager launches
and synchronizations between the replays; two jobs in a row capture afresh.
The reference launches every kernel of every generation from its host loop
(host/src/performdocking.cpp.Cuda:685-831).
"""
import numpy as np
import pytest

import adgpu
from adgpu import synth
from adgpu._abi import energy_params, search_params

pytestmark = pytest.mark.gpu

NRUN, POP = 40, 100      # 40 x 100 local searches a generation: two run groups


@pytest.fixture(scope="module")
def c1(gpu_ctx):
    lig = synth.make_ligand(30, 6, seed=7)
    maps = synth.make_maps(lig.type_names, npts=60, spacing=0.375, seed=11)
    tables = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    mid = gpu_ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(tables, energy_params())
    rng = np.random.default_rng(5)
    n = NRUN * POP
    init = np.zeros((n, 64), np.float32)
    init[:, 0:3] = rng.random((n, 3)) * (np.array(maps.size, np.float64) - 1.0)
    init[:, 3:6] = rng.random((n, 3)) * np.array([360.0, 180.0, 360.0])
    init[:, 6:] = rng.random((n, 58)) * 360.0
    seeds = rng.integers(0, 2 ** 32, size=n * 64, dtype=np.uint64).astype(np.uint32)
    yield gpu_ctx, mid, lid, maps, init, seeds
    gpu_ctx.set_graph_block(0)
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)


def _dock(ctx, mid, lid, sp, init, seeds, block, autostop=None):
    ctx.set_graph_block(block)
    job = ctx.job(mid, lid, sp, init, seeds)
    if autostop is not None:
        job.set_autostop(True, *autostop)
    job.initpop()
    gens, total = job.run()
    st = job.download()
    busy, launches = job.kernel_busy(1)
    job.destroy()
    return gens, total, st, busy, launches


def _same(a, b, what):
    assert a[0] == b[0] and a[1] == b[1], (what, a[:2], b[:2])
    for k in a[2]:
        assert np.array_equal(a[2][k], b[2][k]), (what, k)


# Searches capped at 50 iterations: the stop rule can then prove several
# generations ahead (a generation adds at most 99 + 100 x 50 evaluations per
# run, twice that for Solis-Wets), so most generations run inside blocks.
@pytest.mark.parametrize("ls,evals", [("ad", 300000), ("sw", 600000), ("adam", 300000), ("fire", 300000)])
def test_graph_blocks_are_bit_identical(c1, ls, evals):
    ctx, mid, lid, maps, init, seeds = c1
    sp = search_params(num_runs=NRUN, pop_size=POP, max_evals=evals, ls=ls, spacing=maps.spacing, max_iters=50)
    ref = _dock(ctx, mid, lid, sp, init, seeds, 0)
    assert ref[0] >= 16, ref[0]                 # enough generations for several block replays
    for block in (8, 4, 2):
        got = _dock(ctx, mid, lid, sp, init, seeds, block)
        print(ls, "block", block, "generations", got[0], "LS launches", got[4], "LS busy ms %.2f" % got[3], flush=True)
        _same(ref, got, (ls, block))
        # only the local-search launches outside captured blocks are timed: fewer
        # timed launches show that blocks ran
        assert got[4] < ref[4], (got[4], ref[4])
    again = _dock(ctx, mid, lid, sp, init, seeds, 8)   # a second job captures its own blocks
    _same(ref, again, (ls, "again"))


def test_graph_blocks_with_autostop_and_initial_solis_wets(c1):
    ctx, mid, lid, maps, init, seeds = c1
    sp = search_params(num_runs=NRUN, pop_size=POP, max_evals=600000, ls="ad", spacing=maps.spacing,
                       initial_sw_generations=5, max_iters=50)
    ref = _dock(ctx, mid, lid, sp, init, seeds, 0, autostop=(0.5, 5, 0))
    got = _dock(ctx, mid, lid, sp, init, seeds, 4, autostop=(0.5, 5, 0))
    print("autostop: generations", ref[0], got[0], "timed LS launches", ref[4], got[4], flush=True)
    _same(ref, got, "autostop")
    sp2 = search_params(num_runs=NRUN, pop_size=POP, max_evals=300000, ls="ad", spacing=maps.spacing,
                        initial_sw_generations=5, max_iters=50)
    a, b = _dock(ctx, mid, lid, sp2, init, seeds, 0), _dock(ctx, mid, lid, sp2, init, seeds, 4)
    _same(a, b, "initial SW")
    assert b[4] < a[4]


def test_graph_block_argument(gpu_ctx):
    for bad in (-1, 3, 16):
        with pytest.raises(adgpu.AdgpuError):
            gpu_ctx.set_graph_block(bad)
    gpu_ctx.set_graph_block(0)
