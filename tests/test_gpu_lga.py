"""Whole-docking parity over 100 runs: the LGA workload (30-atom / 6-torsion
ligand, 60^3 FP32 maps, nrun 100, pop 150, 2.5 M evaluations per run) with
each local search (ADADELTA, Solis-Wets, FIRE, ADAM) on the GPU against the
oracle's run of the same job, stored in tests/golden/lga_c1*_oracle.npz
(tests/golden/make_lga_golden.py; minutes of CPU each, so they are not
recomputed here).

GPU and oracle start from the same population and LCG seeds, but their FP32
energies differ in the last bits, so their trajectories part after the first
near-tie; the bar is on the outcome:
* the median of the runs' best energies agrees within 0.1 kcal/mol;
* RMSD clustering of the runs' best poses (energy order, 2.0 A, the
  reference's default cluster tolerance) agrees: the lowest-energy poses of
  the two sets lie within 2.0 A of each other, their energies within
  0.1 kcal/mol, and the share of runs in the lowest-energy cluster within
  0.2."""
import os

import numpy as np
import pytest

import adgpu
import oracle as O
from golden import make_lga_golden as G

pytestmark = pytest.mark.gpu

RMSTOL = 2.0   # A


def _clusters(coords, energies):
    """AutoDock-style clustering: poses in energy order; each joins the first
    cluster whose seed (lowest-energy member) is within RMSTOL, else seeds a
    new one.  Returns the cluster of every pose (0 = lowest energy)."""
    order = np.argsort(energies, kind="stable")
    seeds, label = [], np.zeros(len(energies), int)
    for i in order:
        for k, s in enumerate(seeds):
            if np.sqrt(np.mean(np.sum((coords[i] - coords[s]) ** 2, axis=1))) <= RMSTOL:
                label[i] = k
                break
        else:
            label[i] = len(seeds)
            seeds.append(i)
    return label


@pytest.mark.parametrize("ls", ["ad", "sw", "fire", "adam"])
def test_100_run_median_and_clustering_match_the_oracle(gpu_ctx, ls):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", G.fixture_name(ls))
    if not os.path.exists(path):
        pytest.skip("oracle fixture %s not generated (tests/golden/make_lga_golden.py %s)" % (G.fixture_name(ls), ls))
    f = np.load(path)
    c = G.case(ls)
    maps = c["maps"]
    mid = gpu_ctx.maps_upload(c["packed"], maps.size, maps.num_map_atypes, maps.spacing)
    lid = gpu_ctx.ligand_upload(c["tables"], c["ep"])
    job = gpu_ctx.job(mid, lid, c["sp"], c["init"], c["seeds"])
    job.initpop()
    gens, total = job.run()
    st = job.download()
    job.destroy()
    gpu_ctx.ligand_release(lid)
    gpu_ctx.maps_release(mid)
    ene = st["ene"].reshape(G.NRUN, G.POP)
    pop = st["pop"].reshape(G.NRUN, G.POP, 64)
    bi = np.argmin(ene, axis=1)
    e_gpu, g_gpu = ene[np.arange(G.NRUN), bi], pop[np.arange(G.NRUN), bi]
    e_orc, g_orc = f["best_energy"], f["best_genotype"]
    print("%s: generations gpu %d oracle %d; median best gpu %.4f oracle %.4f; min %.4f / %.4f" %
          (ls, gens, int(f["generations"]), np.median(e_gpu), np.median(e_orc), e_gpu.min(), e_orc.min()), flush=True)
    assert total >= G.NRUN * G.EVALS and abs(gens - int(f["generations"])) <= max(3, int(f["generations"]) // 20)
    assert abs(float(np.median(e_gpu)) - float(np.median(e_orc))) <= 0.1

    orc = O.Oracle(c["cg"], c["cl"], c["ep"])
    sp_a = maps.spacing
    x_gpu = np.array([orc.coords64(g) * sp_a for g in g_gpu])
    x_orc = np.array([orc.coords64(g) * sp_a for g in g_orc])
    l_gpu, l_orc = _clusters(x_gpu, e_gpu), _clusters(x_orc, e_orc)
    b_gpu, b_orc = int(np.argmin(e_gpu)), int(np.argmin(e_orc))
    rmsd = float(np.sqrt(np.mean(np.sum((x_gpu[b_gpu] - x_orc[b_orc]) ** 2, axis=1))))
    share_gpu, share_orc = float(np.mean(l_gpu == 0)), float(np.mean(l_orc == 0))
    print("clusters gpu %d oracle %d; best-pose RMSD %.3f A; lowest-cluster share %.2f / %.2f" %
          (l_gpu.max() + 1, l_orc.max() + 1, rmsd, share_gpu, share_orc), flush=True)
    assert rmsd <= RMSTOL
    assert abs(float(e_gpu.min()) - float(e_orc.min())) <= 0.1
    assert abs(share_gpu - share_orc) <= 0.2
