"""adgpu -- MI355X (gfx950) docking hot path of AutoDock-GPU behind a C ABI.

Python mirror of the host interface (include/adgpu.h).  The compute path is
the HIP library libadgpu_hip.so; this module only marshals arrays.  There is
no CPU fallback: every call raises if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (EVAL_ENERGRAD, EVAL_ENERGY, EVAL_GRADIENT, GENOTYPE_STRIDE, LS_AD, LS_ADAM, LS_FIRE, LS_NONE,
                   LS_SW, MAP_BF16, MAP_F16, MAP_F32, NWI, CGrid, CLigand, energy_params, search_params)

__all__ = ["Context", "Job", "Server", "AdgpuError", "energy_params", "search_params", "MAP_F32", "MAP_F16", "MAP_BF16",
           "EVAL_ENERGY", "EVAL_ENERGRAD", "EVAL_GRADIENT", "LS_NONE", "LS_SW", "LS_AD", "LS_ADAM", "LS_FIRE",
           "GENOTYPE_STRIDE", "NWI"]


class AdgpuError(RuntimeError):
    pass


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class Context:
    """A device, its stream, and the resident map sets and ligands
    (replaces setup_gpu_for_docking / finish_gpu_from_docking,
    host/src/performdocking.cpp.Cuda:117-222)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        rc = self.lib.adgpu_open(device, C.byref(h))
        if rc != 0:
            raise AdgpuError("adgpu_open(%d) failed with status %d (no usable GPU?)" % (device, rc))
        self.h = h
        self._keep = {}

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.adgpu_last_error(self.h)
            raise AdgpuError("%s failed (status %d): %s" % (what, rc, msg.decode() if msg else ""))

    def rescore_f64(self, map_id, lig_id, genotypes):
        """FP64 total, inter and intra energies of host genotypes on the GPU."""
        g = _f32(genotypes).reshape(-1, GENOTYPE_STRIDE)
        n = g.shape[0]
        tot, inter, intra = np.zeros(n), np.zeros(n), np.zeros(n)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.adgpu_rescore_f64(self.h, map_id, lig_id, g.ctypes.data_as(_fp), n,
                                               tot.ctypes.data_as(dp), inter.ctypes.data_as(dp),
                                               intra.ctypes.data_as(dp)), "adgpu_rescore_f64")
        return tot, inter, intra

    def set_team_width(self, width: int):
        """0 = per ligand, 16 = four individuals per wavefront (ligands of at
        most 32 atoms), 32 = two, 64 = one."""
        self._check(self.lib.adgpu_set_team_width(self.h, int(width)), "adgpu_set_team_width")

    def set_ls_handoff(self, iterations: int):
        """ADADELTA searches still running after this many iterations pass to
        idle teams while unstarted searches remain (0: never).  Results are
        bit-identical either way; applies to jobs created afterwards."""
        self._check(self.lib.adgpu_set_ls_handoff(self.h, int(iterations)), "adgpu_set_ls_handoff")

    def set_ls_regroup(self, iterations: int):
        """Once a launch's ADADELTA searches have all started, a search that is
        the last work of its wavefront moves to a team left idle beside a busy
        one, trying every this many iterations (0: never; -1: every 8 in jobs
        of four or more run groups, the default).  Bit-identical results;
        applies to jobs created afterwards."""
        self._check(self.lib.adgpu_set_ls_regroup(self.h, int(iterations)), "adgpu_set_ls_regroup")

    def set_team_table(self, table):
        """Per-size team widths [(max_atoms, width), ...] ascending; [] clears."""
        a = np.array([int(m) for m, _ in table] or [0], np.int32)
        w = np.array([int(x) for _, x in table] or [0], np.int32)
        self._check(self.lib.adgpu_set_team_table(self.h, len(table), a.ctypes.data_as(_ip), w.ctypes.data_as(_ip)),
                    "adgpu_set_team_table")

    def close(self):
        if getattr(self, "h", None):
            self.lib.adgpu_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def maps_upload(self, packed, size, num_map_atypes, spacing, fmt=MAP_F32) -> int:
        g = CGrid(packed, size, num_map_atypes, spacing)
        mid = C.c_int()
        self._check(self.lib.adgpu_maps_upload(self.h, C.byref(g.s), int(fmt), C.byref(mid)), "adgpu_maps_upload")
        return mid.value

    def maps_release(self, map_id: int):
        self._check(self.lib.adgpu_maps_release(self.h, map_id), "adgpu_maps_release")

    def ligand_upload(self, tables, ep=None) -> int:
        cl = CLigand(tables)
        ep = ep or energy_params()
        lid = C.c_int()
        self._check(self.lib.adgpu_ligand_upload(self.h, C.byref(cl.s), C.byref(ep), C.byref(lid)),
                    "adgpu_ligand_upload")
        return lid.value

    def ligand_release(self, ligand_id: int):
        self._check(self.lib.adgpu_ligand_release(self.h, ligand_id), "adgpu_ligand_release")

    def eval_poses(self, map_id, ligand_id, genotypes, mode=EVAL_ENERGY):
        g = _f32(genotypes)
        if g.ndim != 2 or g.shape[1] != GENOTYPE_STRIDE:
            raise ValueError("genotypes must be (n, 64)")
        n = g.shape[0]
        e = np.zeros(n, np.float32)
        gr = np.zeros((n, GENOTYPE_STRIDE), np.float32)
        self._check(self.lib.adgpu_eval_poses(self.h, map_id, ligand_id, int(mode), g.ctypes.data_as(_fp), n,
                                              e.ctypes.data_as(_fp), gr.ctypes.data_as(_fp)), "adgpu_eval_poses")
        return e, gr

    def set_graph_block(self, generations: int):
        """Generations per captured hipGraph for jobs started afterwards (0, 2, 4 or 8; 0 = one by one)."""
        self._check(self.lib.adgpu_set_graph_block(self.h, int(generations)), "adgpu_set_graph_block")

    def measure_valu_rate(self):
        """Wavefront VALU instructions per second on independent FP32 FMA chains."""
        r = C.c_double()
        self._check(self.lib.adgpu_measure_valu_rate(self.h, C.byref(r)), "adgpu_measure_valu_rate")
        return r.value

    def job(self, map_id, ligand_id, sp, init_pop, seeds) -> "Job":
        return Job(self, map_id, ligand_id, sp, init_pop, seeds)


class Job:
    """Device state of one docking (docking_with_gpu's per-ligand buffers)."""

    def __init__(self, ctx: Context, map_id, ligand_id, sp, init_pop, seeds):
        self.ctx, self.lib, self.sp = ctx, ctx.lib, sp
        self.ne = sp.num_runs * sp.pop_size
        ip = _f32(init_pop).reshape(-1)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1)
        if ip.size != self.ne * GENOTYPE_STRIDE or sd.size != self.ne * NWI:
            raise ValueError("init_pop / seeds sizes do not match num_runs * pop_size")
        h = C.c_void_p()
        ctx._check(self.lib.adgpu_job_create(ctx.h, map_id, ligand_id, C.byref(sp), ip.ctypes.data_as(_fp),
                                             sd.ctypes.data_as(_up), C.byref(h)), "adgpu_job_create")
        self.h = h

    def destroy(self):
        if getattr(self, "h", None):
            self.lib.adgpu_job_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx._check(rc, what)

    def initpop(self):
        self._c(self.lib.adgpu_job_initpop(self.h), "adgpu_job_initpop")

    def kernel_gen(self):
        self._c(self.lib.adgpu_job_kernel_gen(self.h), "adgpu_job_kernel_gen")

    def kernel_ls(self, generation=0):
        self._c(self.lib.adgpu_job_kernel_ls(self.h, generation), "adgpu_job_kernel_ls")

    def sum_and_swap(self):
        self._c(self.lib.adgpu_job_kernel_sum_and_swap(self.h), "adgpu_job_kernel_sum_and_swap")

    def generation(self):
        self._c(self.lib.adgpu_job_generation(self.h), "adgpu_job_generation")

    def run(self, max_new_generations=None):
        """Generations until check_progress reaches 100 % (or at most
        max_new_generations more); returns (generations so far, total evals)."""
        g, t = C.c_int64(), C.c_int64()
        if max_new_generations is None:
            self._c(self.lib.adgpu_job_run(self.h, C.byref(g), C.byref(t)), "adgpu_job_run")
        else:
            self._c(self.lib.adgpu_job_run_generations(self.h, int(max_new_generations), C.byref(g), C.byref(t)),
                    "adgpu_job_run_generations")
        return g.value, t.value

    def best_f64(self):
        """Each run's best individual and its FP64 re-scored energies."""
        nr = self.sp.num_runs
        geno = np.zeros((nr, GENOTYPE_STRIDE), np.float32)
        e32 = np.zeros(nr, np.float32)
        idx = np.zeros(nr, np.int32)
        tot, inter, intra = np.zeros(nr), np.zeros(nr), np.zeros(nr)
        dp = C.POINTER(C.c_double)
        self._c(self.lib.adgpu_job_best_f64(self.h, geno.ctypes.data_as(_fp), e32.ctypes.data_as(_fp),
                                            idx.ctypes.data_as(_ip), tot.ctypes.data_as(dp),
                                            inter.ctypes.data_as(dp), intra.ctypes.data_as(dp)),
                "adgpu_job_best_f64")
        return dict(genotypes=geno, energy_f32=e32, index=idx, total=tot, inter=inter, intra=intra)

    def download(self):
        pop = np.zeros((self.ne, GENOTYPE_STRIDE), np.float32)
        ene = np.zeros(self.ne, np.float32)
        ev = np.zeros(self.sp.num_runs, np.int32)
        pr = np.zeros(self.ne * NWI, np.uint32)
        self._c(self.lib.adgpu_job_download(self.h, pop.ctypes.data_as(_fp), ene.ctypes.data_as(_fp),
                                            ev.ctypes.data_as(_ip), pr.ctypes.data_as(_up)), "adgpu_job_download")
        return dict(pop=pop, ene=ene, evals_runs=ev, prng=pr)

    def upload(self, pop=None, ene=None, prng=None):
        p = _f32(pop).reshape(-1) if pop is not None else None
        e = _f32(ene).reshape(-1) if ene is not None else None
        r = np.ascontiguousarray(prng, dtype=np.uint32).reshape(-1) if prng is not None else None
        self._c(self.lib.adgpu_job_upload(self.h, p.ctypes.data_as(_fp) if p is not None else None,
                                          e.ctypes.data_as(_fp) if e is not None else None,
                                          r.ctypes.data_as(_up) if r is not None else None), "adgpu_job_upload")

    def kernel_time(self, cls):
        ms, n = C.c_double(), C.c_int64()
        self._c(self.lib.adgpu_job_kernel_time(self.h, cls, C.byref(ms), C.byref(n)), "adgpu_job_kernel_time")
        return ms.value, n.value

    def kernel_busy(self, cls):
        """Union of the launch intervals of one kernel class (ms) and the launches."""
        ms, n = C.c_double(), C.c_int64()
        self._c(self.lib.adgpu_job_kernel_busy(self.h, cls, C.byref(ms), C.byref(n)), "adgpu_job_kernel_busy")
        return ms.value, n.value

    def ls_handoffs(self):
        """ADADELTA searches handed off between teams since the job started."""
        n = C.c_int64()
        self._c(self.lib.adgpu_job_ls_handoffs(self.h, C.byref(n)), "adgpu_job_ls_handoffs")
        return n.value

    def ls_stats(self):
        """{"handoffs", "regroups", "wave_steps"} since the job started, and
        the job's "team_width" (adgpu_job_ls_stats)."""
        a = (C.c_int64 * 4)()
        self._c(self.lib.adgpu_job_ls_stats(self.h, a), "adgpu_job_ls_stats")
        return {"handoffs": a[0], "regroups": a[1], "wave_steps": a[2], "team_width": a[3]}

    def set_autostop(self, enable=True, stopstd=0.15, as_frequency=5, min_as_evals=0):
        """AutoStop tested on the device every as_frequency generations."""
        self._c(self.lib.adgpu_job_set_autostop(self.h, int(bool(enable)), float(stopstd), int(as_frequency),
                                                int(min_as_evals)), "adgpu_job_set_autostop")

    def autostop_check(self):
        r = _abi.AutostopRecord()
        self._c(self.lib.adgpu_job_autostop_check(self.h, C.byref(r)), "adgpu_job_autostop_check")
        return r.as_dict()

    def autostop_log(self):
        n = C.c_int()
        self._c(self.lib.adgpu_job_autostop_log(self.h, 0, 0, None, C.byref(n)), "adgpu_job_autostop_log")
        recs = (_abi.AutostopRecord * max(1, n.value))()
        self._c(self.lib.adgpu_job_autostop_log(self.h, 0, n.value, recs, C.byref(n)), "adgpu_job_autostop_log")
        return [recs[i].as_dict() for i in range(n.value)]

    def autostop_final(self):
        r = _abi.AutostopRecord()
        self._c(self.lib.adgpu_job_autostop_final(self.h, C.byref(r)), "adgpu_job_autostop_final")
        return r.as_dict()


class Server:
    """Serving loop (adgpu_server_*): docking requests run concurrently in
    `slots` slots, admitted in arrival order when a slot is free and the
    device memory every slot's buffers hold, idle slots included, fits
    `memory_budget` bytes (0 = 90 % of free memory)."""

    def __init__(self, ctx: Context, slots: int = 8, run_groups: int = 0, memory_budget: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        cfg = _abi.ServerConfig(int(slots), int(run_groups), int(memory_budget))
        h = C.c_void_p()
        ctx._check(self.lib.adgpu_server_create(ctx.h, C.byref(cfg), C.byref(h)), "adgpu_server_create")
        self.h = h
        self._shapes = {}

    def submit(self, map_id, tables, sp, init_pop, seeds, ep=None, autostop=None, keep_population=False,
               rescore_f64=False) -> int:
        """Queues one docking; autostop = (stopstd, as_frequency, min_as_evals)
        or None.  Returns the ticket."""
        cl = CLigand(tables)
        ep = ep or energy_params()
        ne = sp.num_runs * sp.pop_size
        ip = _f32(init_pop).reshape(-1)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1)
        if ip.size != ne * GENOTYPE_STRIDE or sd.size != ne * NWI:
            raise ValueError("init_pop / seeds sizes do not match num_runs * pop_size")
        q = _abi.Request()
        q.map_id = int(map_id)
        q.ligand = C.pointer(cl.s)
        q.energy = C.pointer(ep)
        q.search = C.pointer(sp)
        q.init_population = ip.ctypes.data_as(_fp)
        q.prng_seeds = sd.ctypes.data_as(_up)
        if autostop is not None:
            q.autostop = 1
            q.stopstd, q.as_frequency, q.min_as_evals = float(autostop[0]), int(autostop[1]), int(autostop[2])
        q.flags = (_abi.REQ_KEEP_POPULATION if keep_population else 0) | (_abi.REQ_RESCORE_F64 if rescore_f64 else 0)
        t = C.c_int64()
        self.ctx._check(self.lib.adgpu_server_submit(self.h, C.byref(q), C.byref(t)), "adgpu_server_submit")
        self._shapes[t.value] = (sp.num_runs, sp.pop_size, bool(keep_population), bool(rescore_f64))
        return t.value

    def next_done(self, timeout_ms: int = -1):
        """Ticket of the oldest finished request not returned yet, or None
        after timeout_ms."""
        t = C.c_int64()
        rc = self.lib.adgpu_server_next_done(self.h, int(timeout_ms), C.byref(t))
        if rc == _abi.ERR_TIMEOUT:
            return None
        self.ctx._check(rc, "adgpu_server_next_done")
        return t.value

    def result(self, ticket: int, timeout_ms: int = -1):
        """Waits for the request and returns its result (and forgets it):
        a dict with status, timing (ms on the server clock), generations,
        total_evals, best_genotypes, best_energies, and best_f64 /
        population / energies when requested at submit."""
        nr, pz, keep, f64 = self._shapes[ticket]
        r = _abi.RequestResult()
        geno = np.zeros((nr, GENOTYPE_STRIDE), np.float32)
        e = np.zeros(nr, np.float32)
        b64 = np.zeros((nr, 3)) if f64 else None
        pop = np.zeros((nr * pz, GENOTYPE_STRIDE), np.float32) if keep else None
        ene = np.zeros(nr * pz, np.float32) if keep else None
        dp = C.POINTER(C.c_double)
        rc = self.lib.adgpu_server_result(self.h, int(ticket), int(timeout_ms), C.byref(r), geno.ctypes.data_as(_fp),
                                          e.ctypes.data_as(_fp), b64.ctypes.data_as(dp) if f64 else None,
                                          pop.ctypes.data_as(_fp) if keep else None,
                                          ene.ctypes.data_as(_fp) if keep else None)
        if rc == _abi.ERR_TIMEOUT:
            return None
        self.ctx._check(rc, "adgpu_server_result")
        del self._shapes[ticket]
        d = r.as_dict()
        d.update(best_genotypes=geno, best_energies=e)
        if f64:
            d["best_f64"] = b64
        if keep:
            d.update(population=pop, energies=ene)
        return d

    def stats(self):
        s = _abi.ServerStats()
        self.ctx._check(self.lib.adgpu_server_stats_get(self.h, C.byref(s)), "adgpu_server_stats_get")
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.adgpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
