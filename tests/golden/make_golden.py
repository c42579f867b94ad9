"""Writes the golden fixtures under tests/golden/ from the reference's own
data files (run in a container where the reference tree is mounted):

  dlg_<complex>.npz -- per-atom vdW / Elec energies, charges, types and
      DOCKED coordinates of the 4 example docking logs
      (examples/output/ocladock_{gpu,cpu}_{1stp,3ce3}_nrun100.dlg, the
      "DOCKED:" blocks), plus the receptor maps of input/<complex>/derived/
      cropped to the box those poses occupy (+2 grid points).
  ref_<name>.npz -- device tables, initial population and PRNG seeds that
      the reference's own host code (oracle/_ref/libadref.so) derives for a
      synthetic ligand and map set (used by GPU tests on boxes without the
      reference tree); ref_flex.npz adds two flexible residues (--flexres).

Usage: python tests/golden/make_golden.py [/root/reference]
"""
from __future__ import annotations

import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "autodock-gpu_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def read_fld(path):
    d = os.path.dirname(path)
    spacing = None
    nel = None
    center = None
    labels, files = [], []
    for line in open(path):
        if line.startswith("#SPACING"):
            spacing = float(line.split()[1])
        elif line.startswith("#NELEMENTS"):
            nel = [int(x) for x in line.split()[1:4]]
        elif line.startswith("#CENTER"):
            center = [float(x) for x in line.split()[1:4]]
        elif line.startswith("label="):
            lab = line[6:].split()[0]
            labels.append(lab.split("-")[0] if "-" in lab else lab[0].lower())
        elif line.startswith("variable"):
            files.append(os.path.join(d, line.split("file=")[1].split()[0]))
    size = [n + 1 for n in nel]
    origo = [c - n * 0.5 * spacing for c, n in zip(center, nel)]
    vals = []
    for fn in files:
        with open(fn) as fh:
            lines = fh.read().split("\n")
        k = next(i for i, l in enumerate(lines) if l.startswith("CENTER")) + 1
        v = np.array([float(x) for x in lines[k:k + size[0] * size[1] * size[2]]], np.float32)
        vals.append(v.reshape(size[2], size[1], size[0]))
    return dict(spacing=spacing, size=size, origo=origo, labels=labels, values=np.stack(vals))


def read_dlg(path):
    models = []
    cur = None
    for line in open(path, encoding="latin-1"):
        if not line.startswith("DOCKED:"):
            continue
        body = line[8:]
        if body.startswith("MODEL"):
            cur = dict(atoms=[], inter=None)
        elif "Final Intermolecular Energy" in body and cur is not None:
            cur["inter"] = float(body.split("=")[1].split()[0])
        elif body.startswith(("ATOM", "HETATM")) and cur is not None:
            x, y, z = float(body[30:38]), float(body[38:46]), float(body[46:54])
            rest = body[54:].split()
            vdw, elec, q, typ = float(rest[0]), float(rest[1]), float(rest[2]), rest[3]
            cur["atoms"].append((x, y, z, vdw, elec, q, typ))
        elif body.startswith("ENDMDL") and cur is not None:
            models.append(cur)
            cur = None
    return models


def dlg_fixture(ref, complex_name):
    fld = [f for f in os.listdir(os.path.join(ref, "input", complex_name, "derived")) if f.endswith(".maps.fld")][0]
    grid = read_fld(os.path.join(ref, "input", complex_name, "derived", fld))
    out = {}
    allxyz = []
    for dev in ("gpu", "cpu"):
        models = read_dlg(os.path.join(ref, "examples", "output", "ocladock_%s_%s_nrun100.dlg" % (dev, complex_name)))
        xyz = np.array([[a[:3] for a in m["atoms"]] for m in models], np.float64)
        out[dev + "_xyz"] = xyz
        out[dev + "_vdw"] = np.array([[a[3] for a in m["atoms"]] for m in models], np.float32)
        out[dev + "_elec"] = np.array([[a[4] for a in m["atoms"]] for m in models], np.float32)
        out[dev + "_inter"] = np.array([m["inter"] for m in models], np.float32)
        out["charges"] = np.array([a[5] for a in models[0]["atoms"]], np.float32)
        out["types"] = np.array([a[6] for a in models[0]["atoms"]])
        allxyz.append(xyz.reshape(-1, 3))
    allxyz = np.concatenate(allxyz)
    sp = grid["spacing"]
    g = (allxyz - np.array(grid["origo"])) / sp
    lo = np.maximum(np.floor(g.min(0)).astype(int) - 2, 0)
    hi = np.minimum(np.ceil(g.max(0)).astype(int) + 3, np.array(grid["size"]))
    v = grid["values"][:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    out["maps"] = np.ascontiguousarray(v)
    out["map_labels"] = np.array(grid["labels"])
    out["spacing"] = np.float64(sp)
    out["origo"] = np.array(grid["origo"]) + lo * sp   # origin of the cropped box
    out["crop_lo"] = lo
    np.savez_compressed(os.path.join(HERE, "dlg_%s.npz" % complex_name), **out)
    print("dlg_%s.npz: maps %s, %d models" % (complex_name, v.shape, len(out["gpu_xyz"])))


def ref_fixture(name, n_atoms, n_tors, lig_seed, map_seed, npts, nrun, psize, seed):
    import oracle as O
    from adgpu import synth
    lig = synth.make_ligand(n_atoms, n_tors, seed=lig_seed)
    maps = synth.make_maps(lig.type_names, npts=npts, seed=map_seed)
    d = tempfile.mkdtemp()
    fld = synth.write_maps(maps, d)
    lp = synth.write_ligand(lig, os.path.join(d, "lig.pdbqt"))
    R = O.RefSetup(fld, lp, nrun=nrun, psize=psize, seed=seed)
    T = R.tables()
    out = {k: np.asarray(getattr(T, k)) for k in ("charges", "atom_types", "atom_map", "ignore_inter", "ref_coords",
                                                   "rotbond_moving", "rotbond_unit", "rotbond_atoms", "atom_rotbonds",
                                                   "pairs", "atype_reqm", "vw_exp", "reqm_ab", "vw_ac", "vw_bd",
                                                   "ds_s", "ds_v")}
    out.update(num_atoms=R.num_atoms, true_ligand_atoms=R.true_ligand_atoms, num_rotbonds=R.num_rotbonds,
               num_atypes=R.num_atypes, init_pop=R.init_pop(), seeds=R.seeds(), nrun=nrun, psize=psize,
               seed=np.array(seed), size=np.array(R.size), spacing=R.spacing, num_map_atypes=R.num_map_atypes,
               size_angstr=np.array(R.size_angstr), coeff_elec=R.coeff_elec, coeff_desolv=R.coeff_desolv,
               qasp=R.qasp, smooth=R.smooth, elec_min=R.elec_min, lig_seed=lig_seed, map_seed=map_seed,
               n_atoms_req=n_atoms, n_tors_req=n_tors, npts=npts)
    np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), **out)
    print("ref_%s.npz: %d atoms, %d torsions, %d pairs" % (name, R.num_atoms, R.num_rotbonds, R.num_pairs))


def flex_fixture(name, lig_seed=7, map_seed=11, flex_seeds=(21, 22), npts=60, nrun=4, psize=16, seed=(5, 9)):
    """Ligand of the 30-atom / 6-torsion case plus two flexible residues of
    12 atoms and 4 torsions each: device tables as the reference's host code
    builds them with --flexres (ligand atoms first, residue atoms after
    them, residue torsions after the ligand's)."""
    import oracle as O
    from adgpu import synth
    lig = synth.make_ligand(30, 6, seed=lig_seed)
    offsets = ((7.0, 11.0, 11.0), (15.0, 11.0, 11.0))
    flex = [synth.make_flexres(12, 4, seed=s, offset=o, resnum=50 + k)
            for k, (s, o) in enumerate(zip(flex_seeds, offsets))]
    types = list(lig.type_names)
    for f in flex:
        types += [t for t in synth.flexres_types(f) if t not in types]
    maps = synth.make_maps(types, npts=npts, seed=map_seed)
    d = tempfile.mkdtemp()
    fld = synth.write_maps(maps, d)
    lp = synth.write_ligand(lig, os.path.join(d, "lig.pdbqt"))
    fp = os.path.join(d, "flex.pdbqt")
    with open(fp, "w") as fh:
        fh.write("".join(flex))
    R = O.RefSetup(fld, lp, nrun=nrun, psize=psize, seed=seed, extra=["--flexres", fp])
    T = R.tables()
    out = {k: np.asarray(getattr(T, k)) for k in ("charges", "atom_types", "atom_map", "ignore_inter", "ref_coords",
                                                   "rotbond_moving", "rotbond_unit", "rotbond_atoms", "atom_rotbonds",
                                                   "pairs", "atype_reqm", "vw_exp", "reqm_ab", "vw_ac", "vw_bd",
                                                   "ds_s", "ds_v")}
    out.update(num_atoms=R.num_atoms, true_ligand_atoms=R.true_ligand_atoms, num_rotbonds=R.num_rotbonds,
               num_atypes=R.num_atypes, init_pop=R.init_pop(), seeds=R.seeds(), nrun=nrun, psize=psize,
               seed=np.array(seed), size=np.array(R.size), spacing=R.spacing, num_map_atypes=R.num_map_atypes,
               size_angstr=np.array(R.size_angstr), coeff_elec=R.coeff_elec, coeff_desolv=R.coeff_desolv,
               qasp=R.qasp, smooth=R.smooth, elec_min=R.elec_min, lig_seed=lig_seed, map_seed=map_seed,
               npts=npts, map_types=np.array(types))
    np.savez_compressed(os.path.join(HERE, "ref_%s.npz" % name), **out)
    print("ref_%s.npz: %d atoms (%d ligand), %d torsions, %d pairs, %d ignored" % (
        name, R.num_atoms, R.true_ligand_atoms, R.num_rotbonds, R.num_pairs, int(np.sum(out["ignore_inter"]))))


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for c in ("1stp", "3ce3"):
        dlg_fixture(ref, c)
    ref_fixture("c1", 30, 6, lig_seed=7, map_seed=11, npts=60, nrun=4, psize=16, seed=(5, 9))
    flex_fixture("flex")
