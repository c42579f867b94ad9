"""GPU GA generation, local searches and LGA jobs through the C ABI against
the oracle's restatement of the reference control flow (same LCG streams)."""
import numpy as np
import pytest

import adgpu
import helpers
import oracle as O
from adgpu._abi import CGrid, CLigand, energy_params, search_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1g(gpu_ctx):
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    case.mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    case.lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    return case


def _pair_jobs(ctx, case, sp, seed=(5, 9)):
    init = O.initial_population(seed, sp.num_runs, sp.pop_size, case.maps.size_angstr, case.spacing)
    seeds = O.cmwc_uints(seed, sp.num_runs * sp.pop_size * 64)
    job = ctx.job(case.mid, case.lid, sp, init, seeds)
    oj = O.OracleJob(case.cg, case.cl, case.ep, sp, init, seeds)
    return job, oj


@pytest.mark.parametrize("width", [0, 16])
def test_initpop_and_ga_generation_exact(gpu_ctx, c1g, width):
    """K1 energies, then K4: selection, crossover and mutation consume the
    same LCG draws, so genes and PRNG states are bit-identical to the oracle;
    offspring energies meet the energy bar (default and 16-lane teams)."""
    gpu_ctx.set_team_width(width)
    try:
        _ga_generations(gpu_ctx, c1g)
    finally:
        gpu_ctx.set_team_width(0)


def _ga_generations(gpu_ctx, c1g):
    sp = search_params(num_runs=6, pop_size=64, ls="sw", spacing=c1g.spacing)
    job, oj = _pair_jobs(gpu_ctx, c1g, sp)
    job.initpop()
    oj.initpop()
    a, b = job.download(), oj.state()
    assert np.array_equal(a["evals_runs"], b["evals_runs"])
    helpers.check_energies(c1g.orc, b["pop"], a["ene"], e_ref=b["ene"])
    job.upload(ene=b["ene"])
    for gen in range(3):
        job.kernel_gen()
        oj.gen()
        job.sum_and_swap()
        oj.sum_and_swap()
        a, b = job.download(), oj.state()
        ng = c1g.ng
        assert np.array_equal(a["pop"][:, :ng], b["pop"][:, :ng]), gen
        assert np.array_equal(a["prng"], b["prng"])
        assert np.array_equal(a["evals_runs"], b["evals_runs"])
        helpers.check_energies(c1g.orc, b["pop"], a["ene"], e_ref=b["ene"])
        # keep the two populations identical (energies may differ in the last bits)
        job.upload(ene=b["ene"])


@pytest.mark.parametrize("ls,expected", [("sw", (56, 57)), ("ad", (29, 30))])
def test_local_search_schedule_on_flat_energy(gpu_ctx, ls, expected):
    """Flat maps: SW halves rho every 4 failures and stops after 28 iterations
    (56 evaluations); ADADELTA improves once on +inf, then stops after 28 more
    (kernel3.cl:405-409, kernel_ad.cl:506-515)."""
    n = 16
    vals = np.full((3, n, n, n), -1.0, np.float32)
    from adgpu import synth
    mid = gpu_ctx.maps_upload(synth.pack_maps(vals), (n, n, n), 1, 0.375)
    lid = gpu_ctx.ligand_upload(helpers.single_atom_tables(0, 0.0), energy_params())
    sp = search_params(num_runs=3, pop_size=2, ls=ls)
    init = np.zeros((6, 64), np.float32)
    init[:, 0:3] = 5.0
    job = gpu_ctx.job(mid, lid, sp, init, O.cmwc_uints((1,), 6 * 64))
    job.initpop()
    job.kernel_gen()
    job.kernel_ls(0)
    job.sum_and_swap()
    ev = job.download()["evals_runs"]
    # per run: K1 2 evals, K4 1 (elite copied), LS on both entities
    assert list(ev) == [2 + 1 + expected[0] + expected[0]] * 3


@pytest.mark.parametrize("width", [0, 16])
@pytest.mark.parametrize("ls", ["sw", "ad", "adam", "fire"])
def test_short_local_search_parity(gpu_ctx, c1g, ls, width):
    """A few local-search iterations from identical populations: the GPU and
    the oracle take the same discrete decisions for nearly all entities and
    their continuous state stays within the gradient bar (default teams and
    16-lane teams)."""
    sp = search_params(num_runs=4, pop_size=32, ls=ls, max_iters=3, spacing=c1g.spacing)
    gpu_ctx.set_team_width(width)
    try:
        _short_ls(gpu_ctx, c1g, sp)
    finally:
        gpu_ctx.set_team_width(0)


def _short_ls(gpu_ctx, c1g, sp):
    job, oj = _pair_jobs(gpu_ctx, c1g, sp, seed=(7, 3))
    job.initpop()
    oj.initpop()
    job.kernel_gen()
    oj.gen()
    # start the LS from the oracle's exact next population and energies
    st = oj.state()
    job.sum_and_swap()
    job.upload(pop=st["pop_next"], ene=st["ene_next"])
    job.sum_and_swap()   # swap back: uploaded buffers become "next"
    job.kernel_ls(0)
    oj.ls(0)
    job.sum_and_swap()
    oj.sum_and_swap()
    a, b = job.download(), oj.state()
    ng = c1g.ng
    d = np.zeros((a["pop"].shape[0], ng))
    d[:, :3] = np.abs(a["pop"][:, :3] - b["pop"][:, :3])
    d[:, 3:] = helpers.angle_diff(a["pop"][:, 3:ng], b["pop"][:, 3:ng])
    close = np.all(d <= np.maximum(1e-2, 1e-3 * np.abs(b["pop"][:, :ng])), axis=1)
    assert close.mean() >= 0.9, (sp.ls_method, close.mean())
    de = np.abs(a["ene"] - b["ene"]) / np.maximum(1.0, np.abs(b["ene"]))
    assert np.mean(de[close] <= 1e-3) >= 0.95


def test_job_determinism_and_eval_accounting(gpu_ctx, c1g):
    """Two identical jobs give bit-identical results; the stopping rule of
    check_progress (performdocking.cpp.Cuda:908-942) is applied to the
    per-run eval totals."""
    sp = search_params(num_runs=8, pop_size=50, max_evals=30000, ls="ad", spacing=c1g.spacing)
    out = []
    for _ in range(2):
        init = O.initial_population((21, 2), 8, 50, c1g.maps.size_angstr, c1g.spacing)
        job = gpu_ctx.job(c1g.mid, c1g.lid, sp, init, O.cmwc_uints((21, 2), 8 * 50 * 64))
        job.initpop()
        gens, total = job.run()
        out.append((gens, total, job.download()))
        job.destroy()
    (g1, t1, s1), (g2, t2, s2) = out
    assert (g1, t1) == (g2, t2)
    assert np.array_equal(s1["pop"], s2["pop"]) and np.array_equal(s1["ene"], s2["ene"])
    assert t1 == int(s1["evals_runs"].sum())
    assert t1 / 8 >= 30000
    # one generation less would not have reached the budget: each generation adds
    # at most pop * (1 + max_iters) evals per run
    assert t1 / 8 < 30000 + 50 * 301


def test_lga_statistical_parity_with_oracle(gpu_ctx, c1g):
    """Whole LGA runs (ADADELTA) on GPU and oracle from the same seeds: the
    trajectories diverge after the first energy ties, so the comparison is on
    the distribution of the best energy of each run."""
    nrun = 24
    sp = search_params(num_runs=nrun, pop_size=40, max_evals=40000, ls="ad", spacing=c1g.spacing)
    job, oj = _pair_jobs(gpu_ctx, c1g, sp, seed=(77,))
    job.initpop()
    job.run()
    oj.run(nthreads=8)
    a, b = job.download(), oj.state()
    best_gpu = a["ene"].reshape(nrun, -1).min(axis=1)
    best_orc = b["ene"].reshape(nrun, -1).min(axis=1)
    assert np.all(np.isfinite(best_gpu))
    # medians of 24 runs; spread of the per-run best is a few kcal/mol
    assert abs(np.median(best_gpu) - np.median(best_orc)) < 1.0, (np.median(best_gpu), np.median(best_orc))


_SHAPE_CHILD = r"""
import sys, numpy as np
sys.path[:0] = [{paths}]
import adgpu, helpers, oracle as O
from adgpu._abi import search_params
case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
ctx = adgpu.Context(0)
mid = ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
lid = ctx.ligand_upload(case.tables, case.ep)
sp = search_params(num_runs=6, pop_size=40, max_evals=40000, ls="{ls}", spacing=case.spacing)
init = O.initial_population((3, 1), 6, 40, case.maps.size_angstr, case.spacing)
job = ctx.job(mid, lid, sp, init, O.cmwc_uints((3, 1), 6 * 40 * 64))
job.initpop()
job.run()
st = job.download()
np.savez(sys.argv[1], pop=st["pop"], ene=st["ene"], ev=st["evals_runs"])
"""


@pytest.mark.parametrize("ls", ["ad", "sw", "fire"])
def test_results_do_not_depend_on_launch_shape(gpu_ctx, c1g, ls, tmp_path):
    """Three run groups, two groups with 64-thread workgroups, one group, and
    three groups each on the whole device give bit-identical jobs: an
    item's result depends only on its own LCG streams."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    repo = os.path.dirname(here)
    paths = ", ".join(repr(p) for p in (os.path.join(repo, "autodock-gpu_amd"), os.path.join(repo, "oracle"), here))
    code = _SHAPE_CHILD.format(paths=paths, ls=ls)
    outs = []
    for k, env in enumerate(({"ADGPU_RUN_GROUPS": "3"}, {"ADGPU_LS_BLOCK": "64", "ADGPU_RUN_GROUPS": "2"},
                             {"ADGPU_RUN_GROUPS": "1", "ADGPU_LS_BLOCK": "128"},
                             {"ADGPU_LS_SHARE": "0", "ADGPU_RUN_GROUPS": "3"})):
        f = os.path.join(str(tmp_path), "r%d.npz" % k)
        print("launch shape", env, flush=True)
        p = subprocess.run([sys.executable, "-c", code, f], env=dict(os.environ, **env), capture_output=True,
                           text=True, timeout=100)
        assert p.returncode == 0, (env, p.stderr[-2000:])
        outs.append(np.load(f))
    for o in outs[1:]:
        for k in ("pop", "ene", "ev"):
            assert np.array_equal(o[k], outs[0][k]), k


def test_maps_and_ligand_outlive_their_jobs(gpu_ctx):
    """A map set or ligand that a job uses cannot be released until the job
    is destroyed (the job keeps device pointers to both)."""
    case = helpers.synthetic_case(12, 2, 3, 5, npts=40)
    mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    sp = search_params(num_runs=2, pop_size=10, max_evals=2000, ls="sw", spacing=case.spacing)
    init = O.initial_population((1,), 2, 10, case.maps.size_angstr, case.spacing)
    job = gpu_ctx.job(mid, lid, sp, init, O.cmwc_uints((1,), 2 * 10 * 64))
    for release, ident in ((gpu_ctx.maps_release, mid), (gpu_ctx.ligand_release, lid)):
        with pytest.raises(adgpu.AdgpuError):
            release(ident)
    job.initpop()
    job.run()
    job.destroy()
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)
