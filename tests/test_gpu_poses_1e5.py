"""Per-pose parity at the full pose-batch size: 1e5 seeded random poses of the
30-atom / 6-torsion ligand on 60^3 FP32 maps (translations uniform in the
box, uniform random quaternions, torsions uniform in [-180, 180) degrees),
energy and merged gradient through adgpu_eval_poses, against the oracle's FP32
restatement of gpu_calc_energy / gpu_calc_energrad
(device/calcenergy.cl:132-554, device/calcMergedEneGra.cl:46-936) and its FP64
evaluation of the same formulas.

Poses fall into three classes, judged on the FP64 oracle:
  * out of grid: some atom outside the maps (each adds 2^24, calcenergy.cl:293-301);
  * clash: in grid, |E| > 1000 kcal/mol;
  * normal: in grid, |E| <= 1000 kcal/mol.
Poses with an atom within 1e-3 grid units of the map boundary can change class
on the last bits of an FP32 coordinate; they are counted and left out.

Bars (normal poses, the docking-relevant class):
  * energy against the FP32 oracle: |dE| / max(1, |E|) <= 1e-4 on at least
    99.5 % of the poses, and every pose within 1e-4 of the sum of |terms| plus
    the coordinate-rounding allowance of helpers.energy_bar;
  * energy against the FP64 oracle: p99 of |dE| / max(1, |E|) <= 1e-4 and max
    <= 2e-3 (FP32 cancellation of large terms);
  * translation and torsion genes against the FP32 oracle: within
    1e-4 |g| + 4 fixed-point quanta on at least 99.8 % of (pose, gene) pairs;
  * orientation genes against FP64: the GPU's error within 2x the FP32
    oracle's own error (+2 quanta) at the median and 99th percentile.
Out-of-grid and clash poses: |dE| / max(1, |E|) <= 1e-4 against the FP32
oracle on at least 99.5 %.
"""
import numpy as np
import pytest

import adgpu
import helpers
from adgpu import synth

pytestmark = pytest.mark.gpu

N_POSES = 100000
OUT_OF_GRID = float(2 ** 24) * 0.99


def _classes(case, poses, e64, inter64):
    xyz = np.stack([case.orc.coords64(p) for p in poses])           # (n, atoms, 3)
    g = case.cg.s
    hi = np.array([g.size_x - 1, g.size_y - 1, g.size_z - 1], np.float64)
    edge = np.min(np.minimum(np.abs(xyz), np.abs(hi - xyz)), axis=(1, 2)) < helpers.EDGE_EPS
    out = inter64 >= OUT_OF_GRID
    clash = ~out & (np.abs(e64) > 1000.0)
    normal = ~out & ~clash
    return out & ~edge, clash & ~edge, normal & ~edge, edge


def _rel(a, b):
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(1.0, np.abs(b.astype(np.float64)))


def test_1e5_random_poses_energy_and_gradient(gpu_ctx):
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    lid = gpu_ctx.ligand_upload(case.tables, case.ep)
    poses = synth.random_poses(N_POSES, case.ng, case.size, seed=2025)
    e_gpu, g_gpu = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    e_en, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
    assert np.all(np.isfinite(e_gpu)) and np.all(np.isfinite(g_gpu))

    e32, _, s32 = case.orc.energy_batch(poses)
    e64, inter64, _ = case.orc.energy_batch(poses, f64=True)
    _, g32 = case.orc.energrad_batch(poses)
    _, g64 = case.orc.energrad_batch(poses, f64=True)
    sens = case.orc.sensitivity_batch(poses)
    out, clash, normal, edge = _classes(case, poses, e64, inter64)
    n = len(poses)
    report = {"out_of_grid": out.mean(), "clash": clash.mean(), "normal": normal.mean(), "edge": edge.mean()}
    assert edge.mean() <= helpers.EDGE_SHARE, report
    assert normal.sum() >= 5000, report

    # energies against the FP32 oracle
    r32 = _rel(e_gpu, e32)
    for name, m in (("normal", normal), ("clash", clash), ("out_of_grid", out)):
        share = float(np.mean(r32[m] <= 1e-4))
        report[name + "_strict_share"] = share
        assert share >= 0.995, (name, report)
    bar = helpers.ENERGY_RTOL * np.maximum(1.0, s32.astype(np.float64)) + helpers.COORD_EPS * sens
    d32 = np.abs(e_gpu.astype(np.float64) - e32.astype(np.float64))
    assert np.all(d32[~edge] <= bar[~edge]), np.flatnonzero((d32 > bar) & ~edge)[:10]
    # the energy-only kernel gives the merged kernel's energy
    assert np.all(np.abs(e_en.astype(np.float64) - e_gpu) <= bar), "energy-only against merged"

    # energies against the FP64 oracle, per class
    r64 = _rel(e_gpu, e64)
    for name, m in (("normal", normal), ("clash", clash), ("out_of_grid", out)):
        if m.any():
            report[name + "_fp64"] = (float(np.median(r64[m])), float(np.percentile(r64[m], 99)),
                                      float(r64[m].max()))
    assert report["normal_fp64"][1] <= 1e-4 and report["normal_fp64"][2] <= 2e-3, report

    # gradient genes (normal poses)
    ng = case.ng
    qv = np.array([helpers.gene_quantum(case, k) for k in range(ng)])
    well = [0, 1, 2] + list(range(6, ng))
    gg, go = g_gpu[normal][:, well].astype(np.float64), g32[normal][:, well].astype(np.float64)
    ok = np.abs(gg - go) <= 1e-4 * np.abs(go) + 4 * qv[well]
    report["grad_well_share"] = float(ok.mean())
    assert ok.mean() >= 0.998, report
    eg = np.abs(g_gpu[normal][:, 3:6] - g64[normal][:, 3:6])
    e3 = np.abs(g32[normal][:, 3:6] - g64[normal][:, 3:6])
    for pct in (50, 99):
        lhs = np.percentile(eg, pct, axis=0)
        rhs = 2 * np.percentile(e3, pct, axis=0) + 2 * qv[3]
        report["orient_p%d" % pct] = (lhs.tolist(), rhs.tolist())
        assert np.all(lhs <= rhs), report
    assert np.all(g_gpu[:, ng:] == 0.0)
    print("1e5 poses:", {k: v for k, v in report.items()}, "of", n)
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)
