"""Oracle vs the reference's golden docking logs.

examples/output/ocladock_{gpu,cpu}_{1stp,3ce3}_nrun100.dlg list, for each of
100 final poses, every atom's coordinates (3 decimals) and its vdW (+H-bond
+desolvation) and electrostatic intermolecular energies (2 decimals), computed
on the maps of input/<complex>/derived/ (fixtures in tests/golden/, written by
make_golden.py).  The oracle's trilinear map gathers (calcenergy.cl:280-379,
packed layout processgrid.cpp:234-249) must reproduce them.  In these logs the
per-atom vdW column is the atom-type map alone and Elec = q * electrostatic map;
the per-pose "Final Intermolecular Energy" is the full device sum
type + q * e + |q| * d over the atoms.
"""
import numpy as np
import pytest

import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params


def _per_atom(dlg, dev):
    maps = dlg["maps"]
    nm, nz, ny, nx = maps.shape
    cg = CGrid(synth.pack_maps(maps), (nx, ny, nz), nm - 2, float(dlg["spacing"]))
    labels = [str(x) for x in dlg["map_labels"]]
    ep = energy_params()
    xyz = dlg[dev + "_xyz"]
    grid = (xyz - dlg["origo"]) / dlg["spacing"]
    vdw = np.zeros(xyz.shape[:2])
    elec = np.zeros(xyz.shape[:2])
    desolv = np.zeros(xyz.shape[:2])
    for a, (t, q) in enumerate(zip(dlg["types"], dlg["charges"])):
        tix = labels.index(str(t))
        ors = {}
        for qq in (q, -q, 0.0):
            ors[qq] = O.Oracle(cg, CLigand(helpers.single_atom_tables(tix, qq)), ep)
        for m in range(xyz.shape[0]):
            g = np.zeros(64, np.float32)
            g[0:3] = grid[m, a]
            ep_, em_, e0 = (ors[q].energy64(g)[0], ors[-q].energy64(g)[0], ors[0.0].energy64(g)[0])
            elec[m, a] = 0.5 * (ep_ - em_)           # q * e
            desolv[m, a] = 0.5 * (ep_ + em_) - e0     # |q| * d
            vdw[m, a] = e0                            # type map
    return vdw, elec, desolv


@pytest.mark.parametrize("complex_name", ["1stp", "3ce3"])
@pytest.mark.parametrize("dev", ["gpu", "cpu"])
def test_dlg_per_atom_energies(complex_name, dev):
    dlg = helpers.load_dlg(complex_name)
    vdw, elec, desolv = _per_atom(dlg, dev)
    # 2-decimal values; coordinates carry +-0.0005 A, i.e. +-0.0013 grid units
    dv = np.abs(vdw - dlg[dev + "_vdw"])
    de = np.abs(elec - dlg[dev + "_elec"])
    assert np.percentile(dv, 99) <= 0.0101, np.percentile(dv, 99)
    assert np.percentile(de, 99) <= 0.0101, np.percentile(de, 99)
    assert dv.max() < 0.06 and de.max() < 0.06
    # the per-model "Final Intermolecular Energy" is the rounded sum
    tot = (vdw + elec + desolv).sum(axis=1)
    assert np.max(np.abs(tot - dlg[dev + "_inter"])) < 0.05
