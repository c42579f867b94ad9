"""ctypes mirror of include/adgpu.h and the loader of the HIP library.

The library is built in-tree (autodock-gpu_amd/adgpu/libadgpu_hip.so) by
adgpu.build.  Loading fails loudly when it is missing: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libadgpu_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

GENOTYPE_STRIDE = 64
NWI = 64

MAP_F32, MAP_F16, MAP_BF16 = 0, 1, 2
LS_NONE, LS_SW, LS_AD, LS_ADAM, LS_FIRE = 0, 1, 2, 3, 4
EVAL_ENERGY, EVAL_ENERGRAD, EVAL_GRADIENT = 0, 1, 2
LS_NAMES = {"none": LS_NONE, "ga": LS_NONE, "sw": LS_SW, "ad": LS_AD, "adam": LS_ADAM, "fire": LS_FIRE}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_u16 = C.POINTER(C.c_uint16)


class Grid(C.Structure):
    _fields_ = [("size_x", C.c_int32), ("size_y", C.c_int32), ("size_z", C.c_int32),
                ("num_map_atypes", C.c_int32), ("spacing", C.c_float), ("packed", _fp)]


class Ligand(C.Structure):
    _fields_ = [("num_atoms", C.c_int32), ("true_ligand_atoms", C.c_int32), ("num_rotbonds", C.c_int32),
                ("num_atypes", C.c_int32), ("num_pairs", C.c_int32),
                ("charges", _fp), ("atom_types", _ip), ("atom_map", _ip), ("ignore_inter", _u8),
                ("ref_coords", _fp), ("rotbond_moving", _fp), ("rotbond_unit", _fp),
                ("rotbond_atoms", _ip), ("atom_rotbonds", _u8), ("pairs", _ip),
                ("atype_reqm", _up), ("vw_exp", _u16), ("reqm_ab", _fp), ("vw_ac", _fp), ("vw_bd", _fp),
                ("ds_s", _fp), ("ds_v", _fp)]


class EnergyParams(C.Structure):
    _fields_ = [("coeff_elec", C.c_float), ("coeff_desolv", C.c_float), ("qasp", C.c_float),
                ("smooth", C.c_float), ("elec_min_distance", C.c_float)]


class SearchParams(C.Structure):
    _fields_ = [("num_runs", C.c_int32), ("pop_size", C.c_int32), ("max_evals", C.c_int64),
                ("max_generations", C.c_int64), ("tournament_rate", C.c_float), ("crossover_rate", C.c_float),
                ("mutation_rate", C.c_float), ("abs_max_dmov", C.c_float), ("abs_max_dang", C.c_float),
                ("lsearch_rate", C.c_float), ("ls_method", C.c_int32), ("initial_sw_generations", C.c_int32),
                ("rho_lower_bound", C.c_float), ("base_dmov_mul_sqrt3", C.c_float),
                ("base_dang_mul_sqrt3", C.c_float), ("cons_limit", C.c_int32), ("max_num_of_iters", C.c_int32),
                ("adam_beta1", C.c_float), ("adam_beta2", C.c_float), ("adam_epsilon", C.c_float)]


class AutostopRecord(C.Structure):
    _fields_ = [("generation", C.c_int64), ("total_evals", C.c_int64), ("threshold", C.c_float),
                ("average", C.c_float), ("stddev", C.c_float), ("samples", C.c_int32), ("best_energy", C.c_float),
                ("autostopped", C.c_int32), ("stop", C.c_int32), ("roll_average", C.c_float),
                ("roll_stddev", C.c_float), ("roll_samples", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("run_groups", C.c_int32), ("memory_budget", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("map_id", C.c_int32), ("ligand", C.POINTER(Ligand)), ("energy", C.POINTER(EnergyParams)),
                ("search", C.POINTER(SearchParams)), ("init_population", _fp), ("prng_seeds", _up),
                ("autostop", C.c_int32), ("as_frequency", C.c_int32), ("stopstd", C.c_float),
                ("min_as_evals", C.c_int64), ("flags", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("slot", C.c_int32), ("num_runs", C.c_int32), ("pop_size", C.c_int32),
                ("generations", C.c_int64), ("total_evals", C.c_int64), ("device_bytes", C.c_int64),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double),
                ("message", C.c_char * 160)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["message"] = self.message.decode(errors="replace")
        return d


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("active", C.c_int64), ("peak_active", C.c_int64),
                ("bytes_in_use", C.c_int64), ("peak_bytes", C.c_int64), ("memory_budget", C.c_int64),
                ("slot_reallocations", C.c_int64), ("now_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_KEEP_POPULATION = 1
REQ_RESCORE_F64 = 2
ERR_TIMEOUT = 6


def _ptr(a, ctype):
    return a.ctypes.data_as(ctype)


class CGrid:
    """Keeps the numpy buffer alive behind a Grid struct."""

    def __init__(self, packed: np.ndarray, size, num_map_atypes: int, spacing: float):
        self.packed = np.ascontiguousarray(packed, dtype=np.float32)
        nx, ny, nz = size
        need = 4 * nx * ny * nz * (num_map_atypes + 2)
        if self.packed.size != need:
            raise ValueError("packed maps hold %d floats, expected %d" % (self.packed.size, need))
        self.s = Grid(nx, ny, nz, num_map_atypes, spacing, _ptr(self.packed, _fp))


class CLigand:
    """Keeps the numpy arrays of a LigandTables alive behind a Ligand struct."""

    FIELDS = [("charges", np.float32, _fp), ("atom_types", np.int32, _ip), ("atom_map", np.int32, _ip),
              ("ignore_inter", np.uint8, _u8), ("ref_coords", np.float32, _fp),
              ("rotbond_moving", np.float32, _fp), ("rotbond_unit", np.float32, _fp),
              ("rotbond_atoms", np.int32, _ip), ("atom_rotbonds", np.uint8, _u8), ("pairs", np.int32, _ip),
              ("atype_reqm", np.uint32, _up), ("vw_exp", np.uint16, _u16), ("reqm_ab", np.float32, _fp),
              ("vw_ac", np.float32, _fp), ("vw_bd", np.float32, _fp), ("ds_s", np.float32, _fp),
              ("ds_v", np.float32, _fp)]

    def __init__(self, t):
        self.arrays = {}
        kw = {}
        for name, dt, ct in self.FIELDS:
            a = np.ascontiguousarray(getattr(t, name), dtype=dt).reshape(-1)
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self.arrays[name] = a
            kw[name] = _ptr(a, ct)
        self.s = Ligand(num_atoms=t.num_atoms, true_ligand_atoms=t.true_ligand_atoms,
                        num_rotbonds=t.num_rotbonds, num_atypes=t.num_atypes,
                        num_pairs=int(np.asarray(t.pairs).size // 2), **kw)


def energy_params(coeff_elec=None, coeff_desolv=None, qasp=0.01097, smooth=0.5, elec_min_distance=0.01):
    from .synth import AD4_COEFF_DESOLV, AD4_COEFF_ELEC
    return EnergyParams(AD4_COEFF_ELEC if coeff_elec is None else coeff_elec,
                        AD4_COEFF_DESOLV if coeff_desolv is None else coeff_desolv,
                        qasp, smooth, elec_min_distance)


def search_params(num_runs=20, pop_size=150, max_evals=2500000, max_generations=42000, spacing=0.375,
                  ls="ad", lsearch_rate=100.0, max_iters=300, **kw):
    """Dockpars defaults (getparameters.h:66-145) with the spacing-dependent
    ones of getparameters.cpp:1399-1400."""
    import math
    sp = SearchParams()
    sp.num_runs = num_runs
    sp.pop_size = pop_size
    sp.max_evals = max_evals
    sp.max_generations = max_generations
    sp.tournament_rate = 60.0
    sp.crossover_rate = 80.0
    sp.mutation_rate = 2.0
    sp.abs_max_dmov = 6.0 / spacing
    sp.abs_max_dang = 90.0
    sp.lsearch_rate = lsearch_rate
    sp.ls_method = LS_NAMES[ls] if isinstance(ls, str) else int(ls)
    sp.initial_sw_generations = 0
    sp.rho_lower_bound = 0.01
    sp.base_dmov_mul_sqrt3 = 2.0 / spacing * math.sqrt(3.0)
    sp.base_dang_mul_sqrt3 = 75.0 * math.sqrt(3.0)
    sp.cons_limit = 4
    sp.max_num_of_iters = max_iters
    sp.adam_beta1 = 0.9
    sp.adam_beta2 = 0.999
    sp.adam_epsilon = 1.0e-8
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


_LIB = None


def load_library(path: str | None = None):
    """Loads the HIP library (ADGPU_LIBRARY names another build of it);
    raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or os.environ.get("ADGPU_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("HIP library %s is missing: run __graft_entry__.build() (or python -m adgpu.build)" % p)
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.adgpu_version.restype = C.c_char_p
    lib.adgpu_last_error.restype = C.c_char_p
    lib.adgpu_last_error.argtypes = [vp]
    lib.adgpu_open.argtypes = [C.c_int, C.POINTER(vp)]
    lib.adgpu_close.argtypes = [vp]
    lib.adgpu_set_team_width.argtypes = [vp, C.c_int]
    lib.adgpu_set_team_table.argtypes = [vp, C.c_int, _ip, _ip]
    lib.adgpu_set_ls_handoff.argtypes = [vp, C.c_int]
    lib.adgpu_maps_upload.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(C.c_int)]
    lib.adgpu_maps_release.argtypes = [vp, C.c_int]
    lib.adgpu_ligand_upload.argtypes = [vp, C.POINTER(Ligand), C.POINTER(EnergyParams), C.POINTER(C.c_int)]
    lib.adgpu_ligand_release.argtypes = [vp, C.c_int]
    lib.adgpu_eval_poses.argtypes = [vp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp, _fp]
    lib.adgpu_job_ls_handoffs.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.adgpu_job_ls_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.adgpu_set_ls_regroup.argtypes = [vp, C.c_int]
    lib.adgpu_measure_valu_rate.argtypes = [vp, C.POINTER(C.c_double)]
    lib.adgpu_set_graph_block.argtypes = [vp, C.c_int]
    lib.adgpu_eval_poses_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, _fp, C.c_int, _fp, _fp]
    lib.adgpu_job_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(SearchParams), _fp, _up, C.POINTER(vp)]
    lib.adgpu_job_destroy.argtypes = [vp]
    for fn in ("adgpu_job_initpop", "adgpu_job_kernel_gen", "adgpu_job_kernel_sum_and_swap",
               "adgpu_job_generation"):
        getattr(lib, fn).argtypes = [vp]
    lib.adgpu_job_kernel_ls.argtypes = [vp, C.c_int]
    lib.adgpu_job_run.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.adgpu_job_run_generations.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.adgpu_device_count.argtypes = [C.POINTER(C.c_int)]
    _dp = C.POINTER(C.c_double)
    lib.adgpu_rescore_f64.argtypes = [vp, C.c_int, C.c_int, _fp, C.c_int, _dp, _dp, _dp]
    lib.adgpu_job_best_f64.argtypes = [vp, _fp, _fp, _ip, _dp, _dp, _dp]
    lib.adgpu_job_download.argtypes = [vp, _fp, _fp, _ip, _up]
    lib.adgpu_job_upload.argtypes = [vp, _fp, _fp, _up]
    lib.adgpu_job_kernel_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.adgpu_job_kernel_busy.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    rp = C.POINTER(AutostopRecord)
    lib.adgpu_job_set_autostop.argtypes = [vp, C.c_int, C.c_float, C.c_int, C.c_int64]
    lib.adgpu_job_autostop_check.argtypes = [vp, rp]
    lib.adgpu_job_autostop_log.argtypes = [vp, C.c_int, C.c_int, rp, C.POINTER(C.c_int)]
    lib.adgpu_job_autostop_final.argtypes = [vp, rp]
    lib.adgpu_server_create.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(vp)]
    lib.adgpu_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
    lib.adgpu_server_next_done.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    lib.adgpu_server_result.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(RequestResult), _fp, _fp, _dp, _fp, _fp]
    lib.adgpu_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
    lib.adgpu_server_destroy.argtypes = [vp]
    if path is None:
        _LIB = lib
    return lib


EXPORTED = ["adgpu_open", "adgpu_close", "adgpu_last_error", "adgpu_version", "adgpu_set_team_width",
            "adgpu_set_ls_handoff",
            "adgpu_set_team_table", "adgpu_device_count", "adgpu_job_run_generations", "adgpu_rescore_f64", "adgpu_job_best_f64", "adgpu_maps_upload",
            "adgpu_maps_release", "adgpu_ligand_upload", "adgpu_ligand_release", "adgpu_eval_poses", "adgpu_eval_poses_noalloc", "adgpu_job_ls_handoffs", "adgpu_job_ls_stats", "adgpu_set_ls_regroup", "adgpu_measure_valu_rate", "adgpu_set_graph_block",
            "adgpu_job_create", "adgpu_job_destroy", "adgpu_job_initpop", "adgpu_job_kernel_gen",
            "adgpu_job_kernel_ls", "adgpu_job_kernel_sum_and_swap", "adgpu_job_generation", "adgpu_job_run",
            "adgpu_job_download", "adgpu_job_upload", "adgpu_job_kernel_time", "adgpu_job_kernel_busy", "adgpu_job_set_autostop",
            "adgpu_job_autostop_check", "adgpu_job_autostop_log", "adgpu_job_autostop_final",
            "adgpu_server_create", "adgpu_server_submit", "adgpu_server_next_done", "adgpu_server_result",
            "adgpu_server_stats_get", "adgpu_server_destroy"]
