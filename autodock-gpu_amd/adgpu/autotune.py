"""Team-width autotuner: measures the local search on the device for
ligands of several sizes at every team width and installs the fastest width
per size class (adgpu_set_team_table).

The reference fixes its work-group size at build time
(NUM_OF_THREADS_PER_BLOCK, common/defines.h:29-49; Makefile.Cuda); here the
width is a run-time choice per ligand, and this module picks it from
measurements: for each size, a seeded synthetic ligand (torsions ~ atoms / 4)
docks at each width, and the width with the most evaluations per second
wins.  The best width depends on how full the device is: a ligand docked
alone leaves much of the device idle, and 64-lane teams (twice the
wavefronts of 32-lane teams) fill it better, while with many ligands served
at once the device is full and 32-lane teams do the same work with fewer
instructions up to ~48 atoms.  So measure() times either one job at a time
(mode "single") or a server full of requests (mode "served").

The table changes only speed: every width meets the same FP32 parity bar,
but results are not bit-identical across widths (the team sums add in a
different order), so a serving run keeps one table for its whole life.
"""
from __future__ import annotations

import time

import numpy as np

from . import synth
from ._abi import MAP_F16, energy_params, search_params

DEFAULT_SIZES = (24, 32, 40, 48, 64, 80, 96, 128)


def _ligand(n_atoms, seed):
    for t in range(min(20, n_atoms // 4), -1, -1):
        try:
            return synth.make_ligand(n_atoms, t, seed=seed)
        except RuntimeError:
            continue
    raise RuntimeError("no synthetic ligand of %d atoms" % n_atoms)


def _initial(nrun, pop, maps, seed):
    rng = np.random.default_rng(seed)
    n = nrun * pop
    init = np.zeros((n, 64), np.float32)
    init[:, 0:3] = rng.random((n, 3)) * (np.array(maps.size, np.float64) - 1.0)
    init[:, 3:6] = rng.random((n, 3)) * np.array([360.0, 180.0, 360.0])
    init[:, 6:] = rng.random((n, 58)) * 360.0
    seeds = rng.integers(0, 2 ** 32, size=n * 64, dtype=np.uint64).astype(np.uint32)
    return init, seeds


def _served_rate(ctx, mid, lig_tables, ep, sp, maps, requests, slots, seed):
    """Evaluations per second of `requests` copies of one ligand (own seeds)
    served by `slots` concurrent slots."""
    from . import Server
    subs = [_initial(sp.num_runs, sp.pop_size, maps, seed + k) for k in range(requests)]
    with Server(ctx, slots=slots) as srv:
        t0 = time.perf_counter()
        tickets = [srv.submit(mid, lig_tables, sp, init, seeds, ep=ep) for init, seeds in subs]
        total = sum(srv.result(t)["total_evals"] for t in tickets)
        return total / (time.perf_counter() - t0)


def measure(ctx, sizes=DEFAULT_SIZES, ls="fire", nrun=64, pop=150, evals=200000, fmt=MAP_F16, reps=2,
            seed=4000, mode="single", requests=32, slots=16):
    """Evaluations per second per (size, width).  mode "single": one bulk
    job of nrun runs at a time (the device runs one ligand); mode "served":
    `requests` jobs of nrun runs through a server with `slots` slots (the
    device is shared by many ligands).  Returns {size: {width: evals/s}} and
    leaves the context's width at 0."""
    ligs = {n: _ligand(n, seed + n) for n in sizes}
    types = sorted({t for lig in ligs.values() for t in lig.type_names})
    maps = synth.make_maps(types, npts=60, spacing=0.375, seed=seed)
    mid = ctx.maps_upload(maps.packed(), maps.size, maps.num_map_atypes, maps.spacing, fmt)
    ep = energy_params()
    sp = search_params(num_runs=nrun, pop_size=pop, max_evals=evals, ls=ls, spacing=maps.spacing)
    init, seeds = _initial(nrun, pop, maps, seed)
    out = {}
    try:
        for n in sizes:
            tables = synth.ligand_tables(ligs[n], maps.spacing, maps.type_names)
            lid = ctx.ligand_upload(tables, ep)
            rates = {}
            for w in ((16, 32, 64) if n <= 32 else (32, 64)):
                ctx.set_team_width(w)
                best = 0.0
                for r in range(reps):
                    if mode == "served":
                        best = max(best, _served_rate(ctx, mid, tables, ep, sp, maps, requests, slots,
                                                      seed + 100 * r))
                        continue
                    job = ctx.job(mid, lid, sp, init, seeds)
                    t0 = time.perf_counter()
                    job.initpop()
                    _, total = job.run()
                    dt = time.perf_counter() - t0
                    job.destroy()
                    best = max(best, total / dt)
                rates[w] = best
            ctx.set_team_width(0)
            ctx.ligand_release(lid)
            out[n] = rates
    finally:
        ctx.set_team_width(0)
        ctx.maps_release(mid)
    return out


def table_from(rates):
    """[(max_atoms, width)] from measure(): each measured size bounds a class
    of the ligands above the previous size; neighbouring classes with the
    same width merge."""
    table = []
    for n in sorted(rates):
        w = max(rates[n], key=rates[n].get)
        if table and table[-1][1] == w:
            table[-1] = (n, w)
        else:
            table.append((n, w))
    return table


def tune(ctx, **kw):
    """Measures, installs the table in the context and returns (table, rates)."""
    rates = measure(ctx, **kw)
    table = table_from(rates)
    ctx.set_team_table(table)
    return table, rates
