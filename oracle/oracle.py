"""ctypes access to the oracle (TEST INFRASTRUCTURE ONLY).

liboracle.so  -- FP32/FP64 C restatement of the reference device path.
_ref/libadref.so -- the reference's own host code (ligand prep, packed maps,
                    CPU scoring); present only where the reference tree was
                    available at build time.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "autodock-gpu_amd"))
from adgpu._abi import AutostopRecord, EnergyParams, Grid, Ligand, SearchParams  # noqa: E402

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint32)
_ip = C.POINTER(C.c_int32)


class OrcJob(C.Structure):
    _fields_ = [("grid", C.POINTER(Grid)), ("lig", C.POINTER(Ligand)), ("ep", C.POINTER(EnergyParams)),
                ("sp", SearchParams), ("pop_cur", _fp), ("pop_next", _fp), ("ene_cur", _fp), ("ene_next", _fp),
                ("evals_new", _ip), ("evals_runs", _ip), ("prng", _up), ("generation", C.c_int64)]


class OrcAutostop(C.Structure):
    _fields_ = [("num_runs", C.c_int32), ("ntop", C.c_uint32), ("ncream", C.c_uint32), ("as_frequency", C.c_uint32),
                ("stopstd", C.c_float), ("first_time", C.c_int32), ("autostopped", C.c_int32),
                ("threshold", C.c_float), ("threshold_used", C.c_float), ("thres_stddev", C.c_float),
                ("curr_avg", C.c_float), ("curr_std", C.c_float), ("roll_count", C.c_uint32),
                ("best_n", C.c_uint32), ("rolling", C.c_float * 16), ("delta_energy", C.c_float),
                ("overall_best", C.c_float), ("sums", C.c_float * 3)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        # ADGPU_ORACLE_LIB names another build of the oracle (the mutation
        # tests load deliberately broken copies through it)
        p = os.environ.get("ADGPU_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(p)
        gp, lp, ep = C.POINTER(Grid), C.POINTER(Ligand), C.POINTER(EnergyParams)
        L.orc_energy_f32.restype = C.c_float
        L.orc_energy_f32.argtypes = [gp, lp, ep, _fp, _fp, _fp, _fp]
        L.orc_energy_f64.restype = C.c_double
        L.orc_energy_f64.argtypes = [gp, lp, ep, _fp, _dp, _dp, _dp]
        L.orc_energrad_f32.restype = C.c_float
        L.orc_energrad_f32.argtypes = [gp, lp, ep, _fp, C.c_int, _fp]
        L.orc_energrad_f64.restype = C.c_double
        L.orc_energrad_f64.argtypes = [gp, lp, ep, _fp, C.c_int, _dp]
        L.orc_atom_grad_f64.restype = C.c_double
        L.orc_atom_grad_f64.argtypes = [gp, lp, ep, _fp, _dp, _dp]
        L.orc_gradient_f32.argtypes = [gp, lp, ep, _fp, _fp]
        L.orc_gradient_f64.argtypes = [gp, lp, ep, _fp, _dp]
        for t, pt in (("f32", _fp), ("f64", _dp)):
            getattr(L, "orc_energy_batch_" + t).argtypes = [gp, lp, ep, _fp, C.c_int64, pt, pt, pt, C.c_int]
            getattr(L, "orc_energrad_batch_" + t).argtypes = [gp, lp, ep, _fp, C.c_int64, C.c_int, pt, pt, C.c_int]
            getattr(L, "orc_gradient_batch_" + t).argtypes = [gp, lp, ep, _fp, C.c_int64, pt, C.c_int]
            getattr(L, "orc_sensitivity_batch_" + t).argtypes = [gp, lp, ep, _fp, C.c_int64, pt, C.c_int]
        L.orc_fast_acos_f64.restype = C.c_double
        L.orc_fast_acos_f64.argtypes = [C.c_double]
        L.orc_coords_f32.argtypes = [lp, _fp, _fp]
        L.orc_coords_f64.argtypes = [lp, _fp, _dp]
        L.orc_rand.restype = C.c_uint32
        L.orc_rand.argtypes = [_up]
        L.orc_randf.restype = C.c_float
        L.orc_randf.argtypes = [_up]
        jp = C.POINTER(OrcJob)
        L.orc_job_init.argtypes = [jp, gp, lp, ep, C.POINTER(SearchParams), _fp, _up]
        L.orc_job_free.argtypes = [jp]
        for fn in ("orc_initpop", "orc_gen", "orc_sum_evals"):
            getattr(L, fn).argtypes = [jp, C.c_int, C.c_int]
        L.orc_ls.argtypes = [jp, C.c_int, C.c_int, C.c_int64]
        L.orc_swap.argtypes = [jp]
        L.orc_lga_run.restype = C.c_int64
        L.orc_lga_run.argtypes = [jp, C.c_int, C.POINTER(C.c_int64)]
        L.orc_lga_generations.restype = C.c_int64
        L.orc_lga_generations.argtypes = [jp, C.c_int, C.c_int, C.c_int64, C.c_int]
        L.orc_cmwc_uints.argtypes = [_up, _up, C.c_int64]
        L.orc_initial_population.argtypes = [_up, C.c_int, C.c_int, _fp, C.c_float, _fp]
        ap, rp = C.POINTER(OrcAutostop), C.POINTER(AutostopRecord)
        L.orc_autostop_init.argtypes = [ap, C.c_int, C.c_int, C.c_float, C.c_int]
        L.orc_autostop_check.restype = C.c_int
        L.orc_autostop_check.argtypes = [ap, C.c_int, _fp, C.c_int64, C.c_int64, _fp, rp]
        L.orc_autostop_final.argtypes = [ap, C.c_int, _fp, C.c_int64, _fp, rp]
        _LIB = L
    return _LIB


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Oracle:
    """Oracle bound to one grid + ligand + energy parameter set."""

    def __init__(self, cgrid, cligand, ep):
        self.cg, self.cl, self.ep = cgrid, cligand, ep
        self.ng = cligand.s.num_rotbonds + 6
        self.n = cligand.s.num_atoms

    def _args(self):
        return C.byref(self.cg.s), C.byref(self.cl.s), C.byref(self.ep)

    def energy32(self, geno):
        g = _f(geno)
        i, a, s = C.c_float(), C.c_float(), C.c_float()
        e = lib().orc_energy_f32(*self._args(), g.ctypes.data_as(_fp), C.byref(i), C.byref(a), C.byref(s))
        return float(e), float(i.value), float(a.value), float(s.value)

    def energy64(self, geno):
        g = _f(geno)
        i, a, s = C.c_double(), C.c_double(), C.c_double()
        e = lib().orc_energy_f64(*self._args(), g.ctypes.data_as(_fp), C.byref(i), C.byref(a), C.byref(s))
        return float(e), float(i.value), float(a.value), float(s.value)

    def energrad32(self, geno, quantize=True):
        g = _f(geno)
        out = np.zeros(64, np.float32)
        e = lib().orc_energrad_f32(*self._args(), g.ctypes.data_as(_fp), int(quantize), out.ctypes.data_as(_fp))
        return float(e), out[: self.ng].copy()

    def energrad64(self, geno, quantize=False):
        g = _f(geno)
        out = np.zeros(64, np.float64)
        e = lib().orc_energrad_f64(*self._args(), g.ctypes.data_as(_fp), int(quantize), out.ctypes.data_as(_dp))
        return float(e), out[: self.ng].copy()

    def sensitivity(self, geno):
        """sum_a |dE/dx_a| |x_a| of the pose (FP64)."""
        g = _f(geno)
        s = C.c_double()
        lib().orc_atom_grad_f64(*self._args(), g.ctypes.data_as(_fp), None, C.byref(s))
        return float(s.value)

    # batches of poses (rows of 64 floats), split over host threads
    @staticmethod
    def _threads(nthreads):
        return int(nthreads or min(16, os.cpu_count() or 1))

    def energy_batch(self, genos, f64=False, nthreads=0):
        """Energies, their intermolecular part and sums of |term| of every pose."""
        g = _f(genos).reshape(-1, 64)
        n = g.shape[0]
        dt, pt, t = (np.float64, _dp, "f64") if f64 else (np.float32, _fp, "f32")
        e, i, a = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
        getattr(lib(), "orc_energy_batch_" + t)(*self._args(), g.ctypes.data_as(_fp), n, e.ctypes.data_as(pt),
                                                i.ctypes.data_as(pt), a.ctypes.data_as(pt), self._threads(nthreads))
        return e, i, a

    def energrad_batch(self, genos, f64=False, quantize=None, nthreads=0):
        """Merged energy and gradient (n, 64) of every pose; FP32 quantizes
        like the device unless told otherwise, FP64 does not."""
        g = _f(genos).reshape(-1, 64)
        n = g.shape[0]
        dt, pt, t = (np.float64, _dp, "f64") if f64 else (np.float32, _fp, "f32")
        q = (not f64) if quantize is None else bool(quantize)
        e, gr = np.zeros(n, dt), np.zeros((n, 64), dt)
        getattr(lib(), "orc_energrad_batch_" + t)(*self._args(), g.ctypes.data_as(_fp), n, int(q),
                                                  e.ctypes.data_as(pt), gr.ctypes.data_as(pt),
                                                  self._threads(nthreads))
        return e, gr

    def gradient_batch(self, genos, f64=False, nthreads=0):
        """Float gradient (FIRE's first step), (n, 64)."""
        g = _f(genos).reshape(-1, 64)
        n = g.shape[0]
        dt, pt, t = (np.float64, _dp, "f64") if f64 else (np.float32, _fp, "f32")
        gr = np.zeros((n, 64), dt)
        getattr(lib(), "orc_gradient_batch_" + t)(*self._args(), g.ctypes.data_as(_fp), n, gr.ctypes.data_as(pt),
                                                  self._threads(nthreads))
        return gr

    def sensitivity_batch(self, genos, nthreads=0):
        """sum_a |dE/dx_a| |x_a| of every pose (FP64)."""
        g = _f(genos).reshape(-1, 64)
        n = g.shape[0]
        s = np.zeros(n, np.float64)
        lib().orc_sensitivity_batch_f64(*self._args(), g.ctypes.data_as(_fp), n, s.ctypes.data_as(_dp),
                                        self._threads(nthreads))
        return s

    def atom_grad64(self, geno):
        """Per-atom dE/dx (kcal/mol/A, unquantized), shape (atoms, 3)."""
        g = _f(geno)
        out = np.zeros(3 * self.n, np.float64)
        lib().orc_atom_grad_f64(*self._args(), g.ctypes.data_as(_fp), out.ctypes.data_as(_dp), None)
        return out.reshape(-1, 3)

    def gradient32(self, geno):
        g = _f(geno)
        out = np.zeros(64, np.float32)
        lib().orc_gradient_f32(*self._args(), g.ctypes.data_as(_fp), out.ctypes.data_as(_fp))
        return out[: self.ng].copy()

    def gradient64(self, geno):
        g = _f(geno)
        out = np.zeros(64, np.float64)
        lib().orc_gradient_f64(*self._args(), g.ctypes.data_as(_fp), out.ctypes.data_as(_dp))
        return out[: self.ng].copy()

    def coords32(self, geno):
        g = _f(geno)
        out = np.zeros(3 * self.n, np.float32)
        lib().orc_coords_f32(C.byref(self.cl.s), g.ctypes.data_as(_fp), out.ctypes.data_as(_fp))
        return out.reshape(-1, 3)

    def coords64(self, geno):
        g = _f(geno)
        out = np.zeros(3 * self.n, np.float64)
        lib().orc_coords_f64(C.byref(self.cl.s), g.ctypes.data_as(_fp), out.ctypes.data_as(_dp))
        return out.reshape(-1, 3)


class OracleJob:
    """An LGA job run by the oracle (reference device control flow on the CPU)."""

    def __init__(self, cgrid, cligand, ep, sp, init_pop, seeds):
        self.keep = (cgrid, cligand, ep, sp)
        self.sp = sp
        self.j = OrcJob()
        self.init_pop = _f(init_pop)
        self.seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        rc = lib().orc_job_init(C.byref(self.j), C.byref(cgrid.s), C.byref(cligand.s), C.byref(ep), C.byref(sp),
                                self.init_pop.ctypes.data_as(_fp), self.seeds.ctypes.data_as(_up))
        if rc:
            raise MemoryError("orc_job_init failed")
        self.ne = sp.num_runs * sp.pop_size

    def __del__(self):
        try:
            lib().orc_job_free(C.byref(self.j))
        except Exception:
            pass

    def _arr(self, ptr, n, dt):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).copy()

    def state(self):
        return dict(pop=self._arr(self.j.pop_cur, self.ne * 64, np.float32).reshape(self.ne, 64),
                    ene=self._arr(self.j.ene_cur, self.ne, np.float32),
                    pop_next=self._arr(self.j.pop_next, self.ne * 64, np.float32).reshape(self.ne, 64),
                    ene_next=self._arr(self.j.ene_next, self.ne, np.float32),
                    evals_runs=self._arr(self.j.evals_runs, self.sp.num_runs, np.int32),
                    evals_new=self._arr(self.j.evals_new, self.ne, np.int32),
                    prng=self._arr(self.j.prng, self.ne * 64, np.uint32))

    def set_state(self, pop=None, ene=None, prng=None):
        if pop is not None:
            np.ctypeslib.as_array(self.j.pop_cur, shape=(self.ne * 64,))[:] = np.asarray(pop, np.float32).reshape(-1)
        if ene is not None:
            np.ctypeslib.as_array(self.j.ene_cur, shape=(self.ne,))[:] = np.asarray(ene, np.float32)
        if prng is not None:
            np.ctypeslib.as_array(self.j.prng, shape=(self.ne * 64,))[:] = np.asarray(prng, np.uint32).reshape(-1)

    def initpop(self):
        lib().orc_initpop(C.byref(self.j), 0, self.sp.num_runs)

    def gen(self):
        lib().orc_gen(C.byref(self.j), 0, self.sp.num_runs)

    def ls(self, generation=0):
        lib().orc_ls(C.byref(self.j), 0, self.sp.num_runs, generation)

    def sum_and_swap(self):
        lib().orc_sum_evals(C.byref(self.j), 0, self.sp.num_runs)
        lib().orc_swap(C.byref(self.j))

    def run(self, nthreads=1):
        tot = C.c_int64()
        g = lib().orc_lga_run(C.byref(self.j), nthreads, C.byref(tot))
        return int(g), int(tot.value)

    def generations(self, r0, r1, ngen, nthreads=1):
        return int(lib().orc_lga_generations(C.byref(self.j), r0, r1, ngen, nthreads))


def cmwc_uints(seed, n):
    s = np.array(list(seed) + [0] * (3 - len(seed)), np.uint32)
    out = np.zeros(n, np.uint32)
    lib().orc_cmwc_uints(s.ctypes.data_as(_up), out.ctypes.data_as(_up), n)
    return out


def initial_population(seed, runs, pop, size_angstr, spacing):
    s = np.array(list(seed) + [0] * (3 - len(seed)), np.uint32)
    sz = np.array(size_angstr, np.float32)
    out = np.zeros(runs * pop * 64, np.float32)
    lib().orc_initial_population(s.ctypes.data_as(_up), runs, pop, sz.ctypes.data_as(_fp), spacing,
                                 out.ctypes.data_as(_fp))
    return out.reshape(runs * pop, 64)


# ---------------------------------------------------------------------------
# the reference's own host code (oracle/_ref/libadref.so)
# ---------------------------------------------------------------------------
_REF = None


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libadref.so"))


def ref_lib():
    global _REF
    if _REF is None:
        L = C.CDLL(os.path.join(HERE, "_ref", "libadref.so"))
        L.ref_open.restype = C.c_void_p
        L.ref_open.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int]
        L.ref_close.argtypes = [C.c_void_p]
        L.ref_info.argtypes = [C.c_void_p, _ip, _dp]
        L.ref_ptr.restype = C.c_void_p
        L.ref_ptr.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_score.argtypes = [C.c_void_p, _fp, _dp, _dp, _dp, _fp]
        L.ref_autostop_new.restype = C.c_void_p
        L.ref_autostop_new.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int]
        L.ref_autostop_free.argtypes = [C.c_void_p]
        L.ref_autostop_check.restype = C.c_int
        L.ref_autostop_check.argtypes = [C.c_void_p, C.c_int, _fp, C.c_ulong, _up]
        L.ref_autostop_final.argtypes = [C.c_void_p, C.c_int, _fp, C.c_ulong, _up]
        _REF = L
    return _REF


class RefSetup:
    """The reference host's view of (fld, pdbqt): device tables, packed maps,
    initial population, PRNG seeds and CPU scoring."""

    def __init__(self, fld, ligand, nrun=2, psize=8, seed=(11, 23), extra=()):
        L = ref_lib()
        args = ["adgpu", "--ffile", fld, "--lfile", ligand, "--nrun", str(nrun), "--psize", str(psize),
                "--seed", ",".join(str(s) for s in seed)] + list(extra)
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        self.h = L.ref_open(len(args), arr, 64)
        if not self.h:
            raise RuntimeError("reference setup failed for %s / %s" % (fld, ligand))
        iv = np.zeros(16, np.int32)
        dv = np.zeros(32, np.float64)
        L.ref_info(self.h, iv.ctypes.data_as(_ip), dv.ctypes.data_as(_dp))
        (self.num_atoms, self.true_ligand_atoms, self.num_atypes, self.num_rotbonds, self.num_pairs,
         self.num_map_atypes, gx, gy, gz, self.num_rotcyc, self.nrun, self.psize) = [int(x) for x in iv[:12]]
        self.size = (gx, gy, gz)
        self.spacing = float(dv[0])
        self.coeff_elec, self.coeff_desolv, self.qasp, self.smooth, self.elec_min = [float(x) for x in dv[1:6]]
        self.size_angstr = tuple(float(x) for x in dv[6:9])
        self.origo = tuple(float(x) for x in dv[9:12])
        self.abs_max_dmov, self.abs_max_dang = float(dv[12]), float(dv[13])
        self.base_dmov_mul_sqrt3, self.base_dang_mul_sqrt3 = float(dv[14]), float(dv[15])

    def __del__(self):
        try:
            ref_lib().ref_close(self.h)
        except Exception:
            pass

    def arr(self, name, dtype, n):
        p = ref_lib().ref_ptr(self.h, name.encode())
        if not p:
            raise KeyError(name)
        ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()

    def tables(self):
        """The reference's device constants as a LigandTables-like namespace
        (prepare_const_fields_for_gpu, calcenergy.cpp:31-345)."""
        from types import SimpleNamespace
        n, T, nt, P = self.num_atoms, self.num_rotbonds, self.num_atypes, self.num_pairs
        rba = self.arr("rotbonds_atoms", np.int32, 256 * 57).reshape(57, 256)
        nra = self.arr("num_rotating_atoms", np.int32, 57)
        ar = np.zeros((n, T), np.uint8)
        for t in range(T):
            ar[rba[t, :nra[t]], t] = 1
        vw_exp_full = self.arr("vw_exp", np.uint16, 32 * 32)
        f = lambda name: self.arr(name, np.float32, 32 * 32)[: nt * nt]
        return SimpleNamespace(
            num_atoms=n, true_ligand_atoms=self.true_ligand_atoms, num_rotbonds=T, num_atypes=nt,
            charges=self.arr("charges", np.float32, 256)[:n], atom_types=self.arr("atom_types", np.int32, 256)[:n],
            atom_map=self.arr("atom_map", np.int32, 256)[:n],
            ignore_inter=self.arr("ignore_inter", np.uint8, 256)[:n],
            ref_coords=self.arr("ref_coords", np.float32, 768)[: 3 * n],
            rotbond_moving=self.arr("rotbond_moving", np.float32, 171)[: 3 * T],
            rotbond_unit=self.arr("rotbond_unit", np.float32, 171)[: 3 * T],
            rotbond_atoms=self.arr("rotbonds", np.int32, 114)[: 2 * T],
            atom_rotbonds=ar.reshape(-1), pairs=self.arr("pairs", np.int32, 2 * P),
            atype_reqm=self.arr("atype_reqm", np.uint32, 32)[:nt], vw_exp=vw_exp_full[: nt * nt],
            reqm_ab=f("reqm_ab"), vw_ac=f("vw_ac"), vw_bd=f("vw_bd"),
            ds_s=self.arr("ds_s", np.float32, 32)[:nt], ds_v=self.arr("ds_v", np.float32, 32)[:nt],
            rotlist=self.arr("rotlist", np.int32, 256 * 57))

    def grids(self):
        gx, gy, gz = self.size
        return self.arr("grids", np.float32, 4 * gx * gy * gz * (self.num_map_atypes + 2))

    def init_pop(self):
        return self.arr("init_pop", np.float32, self.nrun * self.psize * 64).reshape(-1, 64)

    def seeds(self):
        return self.arr("seeds", np.uint32, self.nrun * self.psize * 64)

    def score(self, geno, peratom=False):
        g = np.ascontiguousarray(geno, np.float32)
        i, a = C.c_double(), C.c_double()
        xyz = np.zeros(3 * self.num_atoms, np.float64)
        pa = np.zeros(2 * self.num_atoms, np.float32)
        ref_lib().ref_score(self.h, g.ctypes.data_as(_fp), C.byref(i), C.byref(a), xyz.ctypes.data_as(_dp),
                            pa.ctypes.data_as(_fp) if peratom else None)
        out = dict(inter=i.value, intra=a.value, xyz=xyz.reshape(-1, 3))
        if peratom:
            out["vdw"] = pa[: self.num_atoms].copy()
            out["elec"] = pa[self.num_atoms:].copy()
        return out


# State fields compared between the oracle's AutoStop and the reference's class.
AUTOSTOP_FLOATS = ("threshold", "threshold_used", "thres_stddev", "curr_avg", "curr_std", "delta_energy",
                   "overall_best")
AUTOSTOP_INTS = ("roll_count", "best_n", "first_time", "autostopped")


class OracleAutoStop:
    """autostop_oracle.c: the AutoStop test (host/inc/autostop.hpp:36-244)."""

    def __init__(self, pop_size, num_runs, stopstd=0.15, as_frequency=5):
        self.s = OrcAutostop()
        self.pop, self.nrun = pop_size, num_runs
        lib().orc_autostop_init(C.byref(self.s), pop_size, num_runs, stopstd, as_frequency)
        self.bins = np.zeros(3 * (pop_size + 1), np.float32)

    def check(self, generation, energies, total_evals, min_as_evals=0):
        e = _f(energies).reshape(-1)
        assert e.size == self.pop * self.nrun
        r = AutostopRecord()
        stop = lib().orc_autostop_check(C.byref(self.s), int(generation), e.ctypes.data_as(_fp), int(total_evals),
                                        int(min_as_evals), self.bins.ctypes.data_as(_fp), C.byref(r))
        return bool(stop), r.as_dict()

    def final(self, generation, energies, total_evals):
        e = _f(energies).reshape(-1)
        r = AutostopRecord()
        lib().orc_autostop_final(C.byref(self.s), int(generation), e.ctypes.data_as(_fp), int(total_evals),
                                 self.bins.ctypes.data_as(_fp), C.byref(r))
        return r.as_dict()

    def state(self):
        d = {k: np.float32(getattr(self.s, k)) for k in AUTOSTOP_FLOATS}
        d.update({k: int(getattr(self.s, k)) for k in AUTOSTOP_INTS})
        d["rolling"] = np.array(self.s.rolling[:], np.float32)
        d["sums"] = np.array(self.s.sums[:], np.float32)
        return d


class RefAutoStop:
    """The reference's own AutoStop class (oracle/_ref build)."""

    def __init__(self, pop_size, num_runs, stopstd=0.15, as_frequency=5):
        self.h = ref_lib().ref_autostop_new(pop_size, num_runs, stopstd, as_frequency)
        self.w = np.zeros(33, np.uint32)

    def __del__(self):
        try:
            ref_lib().ref_autostop_free(self.h)
        except Exception:
            pass

    def check(self, generation, energies, total_evals):
        e = _f(energies).reshape(-1)
        return bool(ref_lib().ref_autostop_check(self.h, int(generation), e.ctypes.data_as(_fp), int(total_evals),
                                                 self.w.ctypes.data_as(_up)))

    def final(self, generation, energies, total_evals):
        e = _f(energies).reshape(-1)
        ref_lib().ref_autostop_final(self.h, int(generation), e.ctypes.data_as(_fp), int(total_evals),
                                     self.w.ctypes.data_as(_up))

    def state(self):
        f = self.w.view(np.float32)
        d = dict(zip(AUTOSTOP_FLOATS, [np.float32(x) for x in f[:7]]))
        d["rolling"] = f[7:23].copy()
        d["sums"] = f[23:26].copy()
        d.update(roll_count=int(self.w[26]), best_n=int(self.w[27]), first_time=int(self.w[28]),
                 autostopped=int(self.w[29]))
        return d
