"""The oracle's AutoStop test against the reference's own AutoStop class
(host/inc/autostop.hpp, compiled in oracle/_ref): every state member, bit for
bit, after every test, on energy sequences of converging and non-converging
searches and on degenerate populations."""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (reference host build) not built")


def converging_sequence(nrun, pop, checks, seed, tau=6.0, noise=0.0, wander=0.0, drift=0.0):
    """Population energies at successive tests of a search whose runs converge:
    each run's best approaches its own minimum, the population collapses onto
    it (elitism keeps copies of the best), a few clashing poses stay high."""
    rng = np.random.default_rng(seed)
    final = rng.uniform(-9.0, -6.0, nrun).astype(np.float32)
    final[: nrun // 2] = final[0]        # half of the runs find the same minimum
    out = []
    for c in range(checks):
        scale = np.float32(np.exp(-c / tau))
        best = final + 3.0 * scale - drift * c + rng.normal(0.0, wander, nrun).astype(np.float32)
        spread = rng.exponential(2.0 * scale + noise, size=(nrun, pop)).astype(np.float32)
        e = best[:, None] + spread
        e[:, : pop // 3] = best[:, None]   # copies of the best individual
        clash = rng.random((nrun, pop)) < 0.05
        e[clash] = rng.uniform(1e3, 1e5, clash.sum())
        out.append(e.astype(np.float32).reshape(-1))
    return out


def _bits(x):
    return np.asarray(x, np.float32).view(np.uint32)


def _assert_same_state(a, b, where):
    for k in O.AUTOSTOP_FLOATS + ("rolling", "sums"):
        assert np.array_equal(_bits(a[k]), _bits(b[k])), (where, k, a[k], b[k])
    for k in O.AUTOSTOP_INTS:
        assert a[k] == b[k], (where, k, a[k], b[k])


def _run_pair(seq, nrun, pop, stopstd=0.15, asf=5, evals_per_check=50000):
    ref = O.RefAutoStop(pop, nrun, stopstd, asf)
    orc = O.OracleAutoStop(pop, nrun, stopstd, asf)
    stops = []
    for c, e in enumerate(seq):
        gen = c * asf
        total = (c + 1) * evals_per_check
        r = ref.check(gen, e, total)
        s, rec = orc.check(gen, e, total, 0)
        assert r == s == bool(rec["autostopped"]), (gen, r, s)
        _assert_same_state(orc.state(), ref.state(), gen)
        stops.append(r)
    return ref, orc, stops


@needs_ref
@pytest.mark.parametrize("nrun,pop,seed", [(10, 150, 1), (20, 60, 2), (3, 200, 3)])
def test_converging_search_stops_like_reference(nrun, pop, seed):
    seq = converging_sequence(nrun, pop, 40, seed)
    ref, orc, stops = _run_pair(seq, nrun, pop)
    assert any(stops), "sequence chosen to converge"
    first = stops.index(True)
    assert first * 5 >= 20 and all(stops[first:])   # not before 4 * as_frequency; sticky
    e = seq[-1]
    ref.final(len(seq) * 5, e, 123)
    rec = orc.final(len(seq) * 5, e, 123)
    _assert_same_state(orc.state(), ref.state(), "final")
    assert rec["autostopped"] == 1


@needs_ref
@pytest.mark.parametrize("nrun,pop,seed", [(8, 100, 4), (5, 150, 5)])
def test_improving_search_never_stops_like_reference(nrun, pop, seed):
    """Every run keeps improving by 0.5 kcal/mol per test: the average of the
    best samples moves, so the test never passes."""
    seq = converging_sequence(nrun, pop, 30, seed, tau=1e9, noise=1.0, drift=0.5)
    ref, orc, stops = _run_pair(seq, nrun, pop)
    assert not any(stops)
    ref.final(150, seq[-1], 999)
    rec = orc.final(150, seq[-1], 999)
    _assert_same_state(orc.state(), ref.state(), "final")
    assert rec["autostopped"] == 0 and rec["samples"] == orc.state()["best_n"]


@needs_ref
def test_degenerate_populations_like_reference():
    rng = np.random.default_rng(9)
    cases = [
        # every energy equal: zero spread, delta_energy 0
        (4, 50, [np.full(200, -5.5, np.float32)] * 12),
        # small population: Ncream = 0, threshold = curr_avg + 0 / bestN
        (6, 8, [rng.normal(-4, 0.01, 48).astype(np.float32) for _ in range(12)]),
        # nothing below the initial 2^24 threshold: the average stays 0 and the
        # test loop never runs again
        (2, 30, [np.full(60, 2.0 ** 25, np.float32)] * 6),
        # infinities and NaNs among the energies
        (3, 40, [np.where(rng.random(120) < 0.1, np.float32(np.inf),
                          np.where(rng.random(120) < 0.05, np.float32(np.nan),
                                   rng.normal(-6, 0.05, 120).astype(np.float32))) for _ in range(10)]),
        # the threshold falls below a worse population: no samples, the average
        # is 0 and the statistics freeze, so the test passes (as the reference's)
        (8, 100, converging_sequence(8, 100, 30, 4, tau=1e9, noise=3.0, wander=1.0)),
        # ties on bin edges: energies on a coarse lattice
        (5, 64, [(np.round(rng.normal(-7, 0.3, 320) * 4) / 4).astype(np.float32) for _ in range(10)]),
    ]
    for nrun, pop, seq in cases:
        _run_pair(seq, nrun, pop)


@needs_ref
def test_min_as_evals_defers_the_stop():
    """The caller stops only once total_evals > min_as_evals
    (performdocking.cpp.Cuda:692-696); the test itself stays passed."""
    nrun, pop = 10, 150
    seq = converging_sequence(nrun, pop, 40, 1)
    _, _, plain = _run_pair(seq, nrun, pop)
    first = plain.index(True)
    min_as_evals = (first + 3) * 50000        # passes 2 tests before the budget allows the stop
    orc = O.OracleAutoStop(pop, nrun)
    ref = O.RefAutoStop(pop, nrun)
    stopped = None
    for c, e in enumerate(seq):
        total = (c + 1) * 50000
        r = ref.check(c * 5, e, total)
        s, rec = orc.check(c * 5, e, total, min_as_evals=min_as_evals)
        assert s == (r and total > min_as_evals) and rec["stop"] == s
        if s and stopped is None:
            stopped = c
    assert stopped == first + 3
