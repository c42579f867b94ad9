"""AutoStop on the device (k_autostop.hip) through the C ABI.

Bars: every field of every test record equals the oracle's AutoStop
(oracle/autostop_oracle.c, itself pinned bit for bit to the reference's class)
bit for bit; a job run with the device test stops at the same generation,
with the same population, as the same job driven by the oracle's test on
downloaded energies, which is how the reference's host loop decides
(performdocking.cpp.Cuda:687-696)."""
import numpy as np
import pytest

import adgpu
import helpers
import oracle as O
from adgpu._abi import search_params
from test_oracle_autostop import converging_sequence

pytestmark = pytest.mark.gpu

FIELDS_F = ("threshold", "average", "stddev", "best_energy", "roll_average", "roll_stddev")
FIELDS_I = ("generation", "total_evals", "samples", "autostopped", "stop", "roll_samples")


def _same_record(a, b, where):
    for k in FIELDS_F:
        assert np.float32(a[k]).view(np.uint32) == np.float32(b[k]).view(np.uint32), (where, k, a[k], b[k])
    for k in FIELDS_I:
        assert a[k] == b[k], (where, k, a[k], b[k])


@pytest.fixture(scope="module")
def c1g(gpu_ctx):
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    case.mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    case.lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    return case


def _job(ctx, case, nrun, pop, **kw):
    sp = search_params(num_runs=nrun, pop_size=pop, spacing=case.spacing, **kw)
    init = O.initial_population((5, 9), nrun, pop, case.maps.size_angstr, case.spacing)
    return ctx.job(case.mid, case.lid, sp, init, O.cmwc_uints((5, 9), nrun * pop * 64))


@pytest.mark.parametrize("nrun,pop,seed,kind", [(10, 150, 1, "converging"), (20, 60, 2, "converging"),
                                                (8, 100, 4, "improving"), (4, 50, 0, "flat"),
                                                (6, 8, 3, "small")])
def test_device_test_records_match_oracle(gpu_ctx, c1g, nrun, pop, seed, kind):
    """Energies set by upload, generation count advanced by K2 + flip only
    (each adds pop evals per run): every record and the final statistics
    equal the oracle's."""
    if kind == "converging":
        seq = converging_sequence(nrun, pop, 30, seed)
    elif kind == "improving":
        seq = converging_sequence(nrun, pop, 30, seed, tau=1e9, noise=1.0, drift=0.5)
    elif kind == "flat":
        seq = [np.full(nrun * pop, -5.5, np.float32)] * 10
    else:
        rng = np.random.default_rng(seed)
        seq = [rng.normal(-4, 0.01, nrun * pop).astype(np.float32) for _ in range(12)]
    job = _job(gpu_ctx, c1g, nrun, pop, ls="sw")
    job.initpop()                      # every individual counted once: pop evals per run
    asf = 5
    job.set_autostop(True, stopstd=0.15, as_frequency=asf, min_as_evals=10 ** 12)   # tests never end the job
    orc = O.OracleAutoStop(pop, nrun, 0.15, asf)
    for c, e in enumerate(seq):
        if c:
            for _ in range(asf):
                job.sum_and_swap()
        job.upload(ene=e)
        rec = job.autostop_check()
        total = nrun * pop * (1 + asf * c)
        _, ref = orc.check(c * asf, e, total, 10 ** 12)
        _same_record(rec, ref, (kind, c))
    log = job.autostop_log()
    assert len(log) == len(seq) and [r["generation"] for r in log] == [c * asf for c in range(len(seq))]
    fin = job.autostop_final()
    _same_record(fin, orc.final(asf * (len(seq) - 1), seq[-1], nrun * pop * (1 + asf * (len(seq) - 1))), "final")
    job.destroy()


@pytest.mark.parametrize("ls,stopstd,min_frac", [("ad", 1.0, 0.0), ("sw", 0.5, 0.0), ("ad", 1.0, 0.3)])
def test_run_stops_where_the_host_loop_stops(gpu_ctx, c1g, ls, stopstd, min_frac, monkeypatch):
    """The device test inside adgpu_job_run against the reference's loop
    shape run on the host: every as_frequency generations download the
    energies, test them with the oracle, stop when the test passes and the
    evals exceed min_as_evals.  Four run groups, so the test waits for every
    group's generation and holds back their next ones."""
    monkeypatch.setenv("ADGPU_RUN_GROUPS", "4")
    nrun, pop, asf, nev = 8, 60, 5, 400_000
    min_as = int(np.ceil(nev * min_frac)) * nrun
    a = _job(gpu_ctx, c1g, nrun, pop, ls=ls, max_evals=nev)
    a.initpop()
    a.set_autostop(True, stopstd=stopstd, as_frequency=asf, min_as_evals=min_as)
    gens_a, total_a = a.run()
    sa = a.download()
    log = a.autostop_log()

    b = _job(gpu_ctx, c1g, nrun, pop, ls=ls, max_evals=nev)
    b.initpop()
    orc = O.OracleAutoStop(pop, nrun, stopstd, asf)
    gens = 0
    recs = []
    while True:
        st = b.download()
        total = int(st["evals_runs"].sum())
        if max(total / nrun / nev, gens / b.sp.max_generations) * 100.0 >= 100.0:     # check_progress
            break
        if gens % asf == 0:
            stop, rec = orc.check(gens, st["ene"], total, min_as)
            recs.append(rec)
            if stop:
                break
        gens, _ = b.run(max_new_generations=asf - gens % asf)
    sb = b.download()
    assert gens_a == gens and total_a == int(sb["evals_runs"].sum()), (gens_a, gens)
    assert np.array_equal(sa["pop"], sb["pop"]) and np.array_equal(sa["ene"], sb["ene"])
    assert len(log) == len(recs)
    for i, (x, y) in enumerate(zip(log, recs)):
        _same_record(x, y, i)
    stopped = bool(recs and recs[-1]["stop"])
    if stopped:
        # a stopped job runs no more generations
        g2, t2 = a.run()
        assert (g2, t2) == (gens_a, total_a)
        assert np.array_equal(a.download()["pop"], sa["pop"])
    a.destroy()
    b.destroy()
