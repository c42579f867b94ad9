"""The generator's device tables, the oracle's host-side pieces and its
pose -> coordinates / intermolecular energy against the reference's own host
C++ (compiled from the reference sources into oracle/_ref/libadref.so by
oracle/Makefile).  Skipped where that build is absent."""
import os
import tempfile

import numpy as np
import pytest

import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libadref.so not built")

FIELDS = ("charges", "atom_types", "atom_map", "ref_coords", "rotbond_moving", "rotbond_unit", "rotbond_atoms",
          "atom_rotbonds", "pairs", "atype_reqm", "vw_exp", "reqm_ab", "vw_ac", "vw_bd", "ds_s", "ds_v")


def _setup(n_atoms, n_tors, lig_seed, map_seed, npts=40, nrun=2, psize=8, seed=(11, 23)):
    lig = synth.make_ligand(n_atoms, n_tors, seed=lig_seed)
    maps = synth.make_maps(lig.type_names, npts=npts, seed=map_seed, n_wells=60, n_cores=15)
    d = tempfile.mkdtemp()
    fld = synth.write_maps(maps, d)
    lp = synth.write_ligand(lig, os.path.join(d, "lig.pdbqt"))
    R = O.RefSetup(fld, lp, nrun=nrun, psize=psize, seed=seed)
    return lig, maps, R


@pytest.mark.parametrize("n_atoms,n_tors,seed", [(30, 6, 1), (12, 2, 2), (60, 15, 3), (45, 0, 4), (90, 20, 5)])
def test_generator_tables_equal_reference(n_atoms, n_tors, seed):
    """prepare_const_fields_for_gpu (calcenergy.cpp:31-345) on the generator's
    pdbqt gives exactly the tables the generator builds itself."""
    lig, maps, R = _setup(n_atoms, n_tors, seed, seed + 100)
    T = synth.ligand_tables(lig, maps.spacing, maps.type_names)
    RT = R.tables()
    assert (R.num_atoms, R.num_rotbonds, R.num_pairs) == (T.num_atoms, T.num_rotbonds, T.num_pairs // 2)
    for f in FIELDS:
        a = np.asarray(getattr(T, f)).reshape(-1)
        b = np.asarray(getattr(RT, f)).reshape(-1)
        assert a.shape == b.shape, f
        assert np.array_equal(a.astype(b.dtype), b), f
    assert R.num_map_atypes == maps.num_map_atypes
    # packed layout (get_gridvalues) of the same map files, up to map2float's parsing
    assert np.max(np.abs(R.grids() - maps.packed())) < 1e-5


def test_rotation_list_is_per_atom_ascending():
    """gen_rotlist (calcenergy.cpp:387-488): every atom is rotated about its
    rotatable bonds in ascending index, then by the general rotation."""
    lig, maps, R = _setup(60, 15, 3, 103)
    RT = R.tables()
    T = RT.num_rotbonds
    ar = RT.atom_rotbonds.reshape(RT.num_atoms, T)
    seq = {a: [] for a in range(RT.num_atoms)}
    for e in RT.rotlist[: 64 * R.num_rotcyc]:
        if e & 0x40000:                      # RLIST_DUMMY_MASK
            continue
        a = e & 0xFF
        seq[a].append(-1 if e & 0x20000 else (e & 0xFF00) >> 8)
    for a in range(RT.num_atoms):
        assert seq[a] == [t for t in range(T) if ar[a, t]] + [-1]


def test_prng_seeds_and_initial_population():
    """LocalRNG (miscellaneous.h:146-218) seeds and gen_initpop_and_reflig
    (getparameters.cpp:2587-2622) genotypes."""
    lig, maps, R = _setup(30, 6, 1, 101, nrun=3, psize=10, seed=(11, 23))
    assert np.array_equal(R.seeds(), O.cmwc_uints((11, 23), 3 * 10 * 64))
    ip = R.init_pop()
    mine = O.initial_population((11, 23), 3, 10, R.size_angstr, R.spacing)
    assert np.max(np.abs(ip[:, :63] - mine[:, :63])) < 1e-4


def test_coords_and_inter_energy_vs_reference_cpu_path():
    """change_conform_f (FP64, exact pi/180) and calc_interE_f on the same
    genotypes.  The device converts degrees with DEG_TO_RAD = 0.0174533, which
    differs from pi/180 by 4.3e-7 relative; coordinates agree to that order."""
    lig, maps, R = _setup(30, 6, 1, 101, npts=60)
    RT = R.tables()
    ep = energy_params(coeff_elec=R.coeff_elec, coeff_desolv=R.coeff_desolv, qasp=R.qasp, smooth=R.smooth,
                       elec_min_distance=R.elec_min)
    cg = CGrid(R.grids(), R.size, R.num_map_atypes, R.spacing)
    orc = O.Oracle(cg, CLigand(RT), ep)
    case = helpers.synthetic_case(30, 6, 1, 101, npts=60)
    poses = helpers.poses_for(case, 80, seed=31, frac_central=1.0, spread=5.0)
    for g in poses:
        rs = R.score(g)
        lever = np.max(np.abs(orc.coords64(np.zeros(64, np.float32)))) + 1.0
        assert np.max(np.abs(orc.coords64(g) - rs["xyz"])) < 2e-6 * 360 * lever
        # the reference coordinates through the oracle's rigid evaluation
        t, centre = helpers.rigid_tables(RT.atom_map, RT.charges, rs["xyz"])
        o2 = O.Oracle(cg, CLigand(t), ep)
        gg = np.zeros(64, np.float32)
        gg[0:3] = centre
        e_inter = o2.energy64(gg)[1]
        # calc_interE_f works in float32; the rigid re-placement rounds the
        # coordinates to float32 as well
        assert e_inter == pytest.approx(rs["inter"], rel=1e-4, abs=1e-3)


def test_flexible_residues_vs_reference_cpu_path():
    """A ligand plus two flexible residues given with --flexres: residue atoms
    follow the ligand atoms, only ligand atoms take the general rotation and
    translation, the first two atoms of each residue are excluded from the
    intermolecular energy (processligand.cpp:1245-1250).  The oracle's
    coordinates equal change_conform_f and its intermolecular energy equals
    calc_interE_f; the committed fixture tables equal a fresh reference build."""
    lig = synth.make_ligand(30, 6, seed=7)
    offsets = ((7.0, 11.0, 11.0), (15.0, 11.0, 11.0))
    flex = [synth.make_flexres(12, 4, seed=s, offset=o, resnum=50 + k) for k, (s, o) in enumerate(zip((21, 22), offsets))]
    types = list(lig.type_names)
    for f in flex:
        types += [t for t in synth.flexres_types(f) if t not in types]
    maps = synth.make_maps(types, npts=60, seed=11)
    d = tempfile.mkdtemp()
    fld = synth.write_maps(maps, d)
    lp = synth.write_ligand(lig, os.path.join(d, "lig.pdbqt"))
    fp = os.path.join(d, "flex.pdbqt")
    with open(fp, "w") as fh:
        fh.write("".join(flex))
    R = O.RefSetup(fld, lp, nrun=4, psize=16, seed=(5, 9), extra=["--flexres", fp])
    RT = R.tables()
    assert (R.num_atoms, R.true_ligand_atoms, R.num_rotbonds) == (54, 30, 14)
    assert list(np.nonzero(RT.ignore_inter)[0]) == [30, 31, 42, 43]
    fx = helpers.load_ref_fixture("flex")
    for k in FIELDS + ("ignore_inter",):
        assert np.array_equal(np.asarray(getattr(RT, k)), fx[k]), k
    ep = energy_params(coeff_elec=R.coeff_elec, coeff_desolv=R.coeff_desolv, qasp=R.qasp, smooth=R.smooth,
                       elec_min_distance=R.elec_min)
    cg = CGrid(R.grids(), R.size, R.num_map_atypes, R.spacing)
    orc = O.Oracle(cg, CLigand(RT), ep)
    rng = np.random.default_rng(3)
    poses = R.init_pop()[:40].copy()
    poses[:, 6:20] = rng.uniform(0, 360, size=(40, 14)).astype(np.float32)
    lever = np.max(np.abs(orc.coords64(np.zeros(64, np.float32)))) + 1.0
    for g in poses:
        rs = R.score(g)
        xyz = orc.coords64(g)
        assert np.max(np.abs(xyz - rs["xyz"])) < 2e-6 * 360 * lever
        # the reference's coordinates through the oracle's rigid evaluation,
        # with the residue root atoms excluded as in the reference
        t, centre = helpers.rigid_tables(RT.atom_map, RT.charges, rs["xyz"])
        t.ignore_inter = np.asarray(RT.ignore_inter, np.uint8)
        gg = np.zeros(64, np.float32)
        gg[0:3] = centre
        e_inter = O.Oracle(cg, CLigand(t), ep).energy64(gg)[1]
        assert e_inter == pytest.approx(rs["inter"], rel=1e-4, abs=1e-3)
