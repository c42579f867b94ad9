"""Pins of the oracle by what the mathematics fixes (closed forms, invariants,
finite differences), chosen so that a dropped term, a wrong sign or index, or
a transposed operand fails one of them."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params, search_params

DEG = 0.0174533


def _grid_from_field(f, n=16, spacing=0.375, ntypes=1):
    """Maps whose every map is the field f(x, y, z) (grid units)."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = f(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)).astype(np.float32)
    vals = np.stack([v] * (ntypes + 2))
    return CGrid(synth.pack_maps(vals), (n, n, n), ntypes, spacing)


def _pair_tables(types=("C", "C"), charges=(0.0, 0.0), desolv=False):
    """Two atoms 0 and 1 on the x axis, one intramolecular pair, atom 1 on a
    rotatable bond so its x coordinate is set through the ligand frame."""
    reqm, vw_exp, reqm_ab, ac, bd, ds_s, ds_v = synth.vdw_tables(list(dict.fromkeys(types)))
    names = list(dict.fromkeys(types))
    t = helpers.single_atom_tables(0, 0.0)
    t.num_atoms = 2
    t.true_ligand_atoms = 2
    t.num_atypes = len(names)
    t.charges = np.array(charges, np.float32)
    t.atom_types = np.array([names.index(x) for x in types], np.int32)
    t.atom_map = np.zeros(2, np.int32)
    t.ignore_inter = np.ones(2, np.uint8)          # intramolecular terms only
    t.pairs = np.array([0, 1], np.int32)
    t.atype_reqm, t.vw_exp = reqm, vw_exp.reshape(-1)
    t.reqm_ab, t.vw_ac, t.vw_bd = (reqm_ab.astype(np.float32).reshape(-1), ac.astype(np.float32).reshape(-1),
                                   bd.astype(np.float32).reshape(-1))
    t.ds_s, t.ds_v = ds_s.astype(np.float32), ds_v.astype(np.float32)
    if not desolv:
        t.ds_s[:] = 0.0
        t.ds_v[:] = 0.0
    return t


def _pair_energy(r_angstrom, types=("C", "C"), charges=(0.0, 0.0), ep=None, spacing=0.375, f64=True,
                 desolv=False):
    t = _pair_tables(types, charges, desolv)
    t.ref_coords = np.array([0, 0, 0, r_angstrom / spacing, 0, 0], np.float32)
    cg = _grid_from_field(lambda x, y, z: 0 * x, spacing=spacing)
    o = O.Oracle(cg, CLigand(t), ep or energy_params())
    g = np.zeros(64, np.float32)
    g[0:3] = 8.0
    return (o.energy64(g) if f64 else o.energy32(g))[2]    # intra part


# --------------------------------------------------------------------------
# LCG (auxiliary_genetic.cl:29-67)
# --------------------------------------------------------------------------
def test_lcg_known_answers():
    import ctypes as C
    s = C.c_uint32(1)
    assert O.lib().orc_rand(C.byref(s)) == (1103515245 * 1 + 12345) % 2 ** 32
    s = C.c_uint32(0xFFFFFFFF)
    assert O.lib().orc_rand(C.byref(s)) == (1103515245 * 0xFFFFFFFF + 12345) % 2 ** 32
    s = C.c_uint32(0)
    r = O.lib().orc_randf(C.byref(s))
    assert r == np.float32(np.float32(12345) / np.float32(4294967296.0)) * np.float32(0.999999)


# --------------------------------------------------------------------------
# pose -> coordinates
# --------------------------------------------------------------------------
def test_coords_identity_and_translation(c1):
    ref = c1.tables.ref_coords.reshape(-1, 3).astype(np.float64)
    g = np.zeros(64, np.float32)
    assert np.array_equal(c1.orc.coords64(g), ref)
    g[0:3] = (3.25, -1.5, 7.0)
    assert np.allclose(c1.orc.coords64(g), ref + np.array([3.25, -1.5, 7.0]), atol=1e-12)


def test_coords_torsion_180_is_reflection(c1):
    T = c1.tables
    ref = T.ref_coords.reshape(-1, 3).astype(np.float64)
    ar = T.atom_rotbonds.reshape(T.num_atoms, T.num_rotbonds)
    # a leaf-most torsion: rotbond 0 is closed first (ENDBRANCH order)
    g = np.zeros(64, np.float32)
    g[6] = 180.0
    c = c1.orc.coords64(g)
    mv = T.rotbond_moving.reshape(-1, 3)[0].astype(np.float64)
    u = T.rotbond_unit.reshape(-1, 3)[0].astype(np.float64)
    u = u / np.linalg.norm(u)
    for a in range(T.num_atoms):
        x = ref[a] - mv
        proj = np.dot(x, u) * u
        exp = (2 * proj - x) + mv if ar[a, 0] else ref[a]
        # DEG_TO_RAD = 0.0174533 makes 180 "degrees" 3.1415940 rad
        assert np.allclose(c[a], exp, atol=5e-5 * max(1.0, np.linalg.norm(x)))


def test_coords_rigid_body_invariants(c1):
    """Any pose keeps bond lengths and all distances inside a rigid fragment."""
    T = c1.tables
    ref = T.ref_coords.reshape(-1, 3).astype(np.float64)
    rig = c1.lig.rigid_id
    poses = helpers.poses_for(c1, 50, seed=5)
    for g in poses:
        c = c1.orc.coords64(g)
        for (a, b) in c1.lig.bonds:
            # unit vectors are stored in float32, so |q| = 1 holds to ~1e-8
            assert abs(np.linalg.norm(c[a] - c[b]) - np.linalg.norm(ref[a] - ref[b])) < 1e-6
        for f in np.unique(rig):
            idx = np.where(rig == f)[0]
            d0 = np.linalg.norm(ref[idx, None] - ref[None, idx], axis=-1)
            d1 = np.linalg.norm(c[idx, None] - c[None, idx], axis=-1)
            assert np.allclose(d0, d1, atol=1e-6)


def test_intra_energy_invariant_under_rigid_motion(c1):
    poses = helpers.poses_for(c1, 20, seed=9)
    for g in poses:
        e1 = c1.orc.energy64(g)[2]
        g2 = g.copy()
        g2[0:3] += np.array([1.5, -2.0, 0.75], np.float32)
        g2[3:6] = (37.0, 101.0, 211.0)
        e2 = c1.orc.energy64(g2)[2]
        assert abs(e1 - e2) <= 1e-9 * max(1.0, abs(e1)) + 1e-8


# --------------------------------------------------------------------------
# intermolecular (trilinear map) energy
# --------------------------------------------------------------------------
def test_inter_linear_field_is_exact():
    f = lambda x, y, z: 0.5 + 0.25 * x - 0.125 * y + 0.0625 * z
    cg = _grid_from_field(f)
    q = 0.3
    o = O.Oracle(cg, CLigand(helpers.single_atom_tables(0, q)), energy_params())
    rng = np.random.default_rng(1)
    for _ in range(50):
        p = rng.uniform(0, 14.99, size=3)
        g = np.zeros(64, np.float32)
        g[0:3] = p
        pf = p.astype(np.float32).astype(np.float64)
        exp = f(*pf) * (1 + q + abs(q))          # type + q*e + |q|*d on identical maps
        assert abs(o.energy64(g)[0] - exp) < 1e-5


def test_inter_node_values_and_out_of_grid():
    rng = np.random.default_rng(3)
    n = 12
    vals = rng.uniform(-1, 1, size=(3, n, n, n)).astype(np.float32)
    cg = CGrid(synth.pack_maps(vals), (n, n, n), 1, 0.375)
    o = O.Oracle(cg, CLigand(helpers.single_atom_tables(0, 0.0)), energy_params())
    for _ in range(30):
        i, j, k = rng.integers(0, n - 1, size=3)
        g = np.zeros(64, np.float32)
        g[0:3] = (i, j, k)
        assert o.energy64(g)[0] == pytest.approx(float(vals[0, k, j, i]), abs=1e-7)
    for p in [(-0.01, 5, 5), (5, 5, n - 1), (n - 0.5, 1, 1), (3, -2, 3)]:
        g = np.zeros(64, np.float32)
        g[0:3] = p
        assert o.energy64(g)[0] == 16777216.0
    # NaN coordinates (a FIRE step from a zero gradient) count as outside the
    # maps, as on the device: the penalty, no map index formed from NaN
    for p in [(np.nan, 5, 5), (5, np.nan, 5), (5, 5, np.nan)]:
        g = np.zeros(64, np.float32)
        g[0:3] = p
        assert o.energy64(g)[0] == 16777216.0 and o.energy32(g)[0] == 16777216.0
        e, grad = o.energrad32(g)
        assert e == 16777216.0


# --------------------------------------------------------------------------
# intramolecular closed forms (calcenergy.cl:397-536)
# --------------------------------------------------------------------------
def test_vdw_well_depth_and_smoothing_plateau():
    ep = energy_params(qasp=0.0)
    eps_cc = synth.AD4_COEFF_VDW * 0.150          # AD4 C-C well depth, weighted
    for r in (4.0, 3.9, 4.1, 3.76, 4.24):          # |r - 4.0| < smooth/2 = 0.25: plateau at E(reqm)
        assert _pair_energy(r, ep=ep) == pytest.approx(-eps_cc, rel=1e-6)
    # outside the plateau the distance shifts by smooth/2 toward reqm
    r = 5.0
    rs = r - 0.25
    A, B = eps_cc * 4.0 ** 12, 2 * eps_cc * 4.0 ** 6
    assert _pair_energy(r, ep=ep) == pytest.approx(A / rs ** 12 - B / rs ** 6, rel=1e-6)
    # H-bond 12-10 pair (HD-OA): well depth = coeff_hb * 5 * 1 at 1.9 A
    e = _pair_energy(1.9, types=("HD", "OA"), ep=ep)
    assert e == pytest.approx(-synth.AD4_COEFF_HB * 5.0, rel=1e-6)


def test_vdw_cutoff_at_8_angstrom():
    ep = energy_params(qasp=0.0)
    assert _pair_energy(8.05, ep=ep) == 0.0          # no charges: only vdW could contribute
    assert _pair_energy(7.95, ep=ep) != 0.0


def test_electrostatic_fit_matches_mehler_solmajer():
    """The device's rational dielectric fit (calcenergy.cl:519-525) tracks the
    distance-dependent Mehler-Solmajer dielectric it replaces."""
    ep = energy_params(qasp=0.0)
    q1, q2 = 0.4, -0.3
    for r in (3.0, 5.0, 9.0, 15.0, 20.0):
        e = _pair_energy(r, charges=(q1, q2), ep=ep) - _pair_energy(r, ep=ep)
        A, B, lam, k = -8.5525, 78.4 + 8.5525, 0.003627, 7.7839
        eps_r = A + B / (1 + k * math.exp(-lam * B * r))
        exact = synth.AD4_COEFF_ELEC * q1 * q2 / (r * eps_r)
        # the fit's largest relative deviation on [0.5, 20.4] A is 9.0 % (r = 11.4 A)
        assert e == pytest.approx(exact, rel=0.091)
    assert _pair_energy(20.6, charges=(q1, q2), ep=ep) == 0.0   # 20.48 A cutoff


def test_elec_min_distance_clamp():
    ep = energy_params(qasp=0.0, elec_min_distance=0.5)
    a = _pair_energy(0.2, charges=(0.5, 0.5), ep=ep) - _pair_energy(0.2, ep=ep)
    b = _pair_energy(0.5, charges=(0.5, 0.5), ep=ep) - _pair_energy(0.5, ep=ep)
    assert a == pytest.approx(b, rel=1e-6)   # difference of two large vdW values


def test_desolvation_fit_matches_gaussian():
    """Rational desolvation fit (calcenergy.cl:492-500) vs exp(-r^2/(2*3.6^2))."""
    ep = energy_params(qasp=0.0)
    t = _pair_tables(desolv=True)
    S, V = float(t.ds_s[0]), float(t.ds_v[0])
    pref = 2 * S * V * synth.AD4_COEFF_DESOLV
    for r in (2.0, 3.0, 5.0, 8.5, 12.0):
        # desolvation = total minus the same pair without desolvation parameters
        e = _pair_energy(r, ep=ep, desolv=True) - _pair_energy(r, ep=ep, desolv=False)
        # the fit deviates from the Gaussian by at most 0.0029 (r = 2.14 A)
        assert e == pytest.approx(pref * math.exp(-r * r / 25.92), abs=0.003 * abs(pref))


# --------------------------------------------------------------------------
# gradients (calcMergedEneGra.cl) vs central finite differences
# --------------------------------------------------------------------------
def _no_desolv_case(c1):
    """C1 with the desolvation parameters zeroed: every remaining energy term
    has its exact derivative in the merged kernel."""
    t = c1.tables
    t2 = SimpleNamespace(**vars(t))
    t2.ds_s = np.zeros_like(t.ds_s)
    t2.ds_v = np.zeros_like(t.ds_v)
    nm = c1.maps.num_map_atypes
    vals = c1.maps.values.copy()
    vals[nm + 1] = 0.0                       # no receptor desolvation map either
    cg = CGrid(synth.pack_maps(vals), c1.maps.size, nm, c1.spacing)
    return O.Oracle(cg, CLigand(t2), energy_params())


def test_energrad_matches_finite_differences(c1):
    o = _no_desolv_case(c1)
    poses = helpers.poses_for(c1, 300, seed=21, frac_central=1.0, spread=3.0)
    conv = c1.spacing ** 2 / DEG ** 2
    checked = 0
    for g in poses:
        e, an = o.energrad64(g, quantize=False)
        if not (abs(e) < 50):                 # keep to non-clashing, in-grid poses
            continue
        # trilinear interpolation has kinks on cell faces: skip poses with an
        # atom within the finite-difference reach of one
        frac = np.abs(o.coords64(g) - np.round(o.coords64(g)))
        if frac.min() < 2e-3:
            continue
        checked += 1
        for k in list(range(3)) + list(range(6, c1.ng)):
            h = 2e-4 if k < 3 else 2e-3
            gp, gm = g.astype(np.float64).copy(), g.astype(np.float64).copy()
            gp[k] += h
            gm[k] -= h
            fd = (o.energy64(gp.astype(np.float32))[0] - o.energy64(gm.astype(np.float32))[0]) / (
                np.float32(g[k] + h) - np.float32(g[k] - h))
            a = an[k] if k < 3 else an[k] / conv * (1.0 / 1.0)
            # genes >= 3: reported gradient = dE/d(deg) * spacing^2 / DEG_TO_RAD^2
            assert a == pytest.approx(fd, rel=2e-3, abs=2e-3), (k, a, fd)
    assert checked >= 10


@pytest.mark.parametrize("a1024,expected", [(5.2, 3), (-5.2, -3), (1.2, 0), (-1.4, 0), (2.6, 1), (20480.0, 8190)])
def test_fixed_point_rounding_known_answers(a1024, expected):
    """float2int_round (calcMergedEneGra.cl:30-35) applied twice: the atom
    force 1024*dE/dx = a1024 becomes trunc(a1024 -+ 0.5) clamped to +-8192, and
    the translation gene rounds 1024 * force * spacing again (spacing 1 A here).
    Hand-derived: 5.2 -> 4 -> 3;  -5.2 -> -4 -> -3;  1.2 -> 0;  2.6 -> 2 -> 1;
    20480 -> clamp 8192 -> 8191 -> 8190."""
    a = a1024 / 1024.0
    cg = _grid_from_field(lambda x, y, z: a * x, spacing=1.0)
    o = O.Oracle(cg, CLigand(helpers.single_atom_tables(0, 0.0)), energy_params())
    g = np.zeros(64, np.float32)
    g[0:3] = (7.3, 7.6, 7.1)
    e, gr = o.energrad32(g, quantize=True)
    assert gr[0] * 1024.0 == expected
    assert gr[1] == 0.0 and gr[2] == 0.0


def test_rotation_gradient_zero_without_torque():
    """A single atom at the rotation centre feels no torque: orientation genes
    are 0 (calcMergedEneGra.cl:756-861)."""
    f = lambda x, y, z: 0.3 * x + 0.1 * y
    cg = _grid_from_field(f)
    o = O.Oracle(cg, CLigand(helpers.single_atom_tables(0, 0.0)), energy_params())
    g = np.zeros(64, np.float32)
    g[0:3] = 7.3
    g[3:6] = (40.0, 70.0, 120.0)
    e, gr = o.energrad64(g, quantize=True)
    assert np.all(gr[3:6] == 0.0)
    assert gr[0] == pytest.approx(0.3, abs=2e-3) and gr[1] == pytest.approx(0.1, abs=2e-3)


# --------------------------------------------------------------------------
# GA generation (kernel4.cl) and local-search control flow
# --------------------------------------------------------------------------
def _flat_case(n_atoms=1):
    cg = _grid_from_field(lambda x, y, z: 0 * x - 1.0)
    return cg, CLigand(helpers.single_atom_tables(0, 0.0)), energy_params()


def test_ga_step_structure(c1):
    sp = search_params(num_runs=2, pop_size=40, ls="sw", spacing=c1.spacing)
    init = O.initial_population((3, 4), 2, 40, c1.maps.size_angstr, c1.spacing)
    seeds = O.cmwc_uints((3, 4), 2 * 40 * 64)
    j = O.OracleJob(c1.cg, c1.cl, c1.ep, sp, init, seeds)
    j.initpop()
    s0 = j.state()
    j.gen()
    s1 = j.state()
    ng = c1.ng
    for r in range(2):
        e = s0["ene"][r * 40:(r + 1) * 40]
        # elitism: entity 0 of the next generation is the best current entity
        b = int(np.argmin(e))
        assert np.array_equal(s1["pop_next"][r * 40, :ng], s0["pop"][r * 40 + b, :ng])
        assert s1["ene_next"][r * 40] == e[b]
        assert s1["evals_new"][r * 40] == 0
        for k in range(1, 40):
            child = s1["pop_next"][r * 40 + k, :ng]
            # every gene comes from some parent of the run or is a bounded mutation
            par = s0["pop"][r * 40:(r + 1) * 40, :ng]
            for gidx in range(ng):
                d = np.min(np.abs(par[:, gidx] - child[gidx]))
                if gidx >= 3:
                    d = np.min(helpers.angle_diff(par[:, gidx], child[gidx]))
                    assert 0.0 <= child[gidx] < 360.0 or d == 0.0
                bound = sp.abs_max_dmov if gidx < 3 else sp.abs_max_dang
                assert d <= bound + 1e-3
            assert s1["evals_new"][r * 40 + k] == 1


def test_solis_wets_rho_schedule_on_flat_energy():
    """Flat energy: every trial fails, rho halves every cons_limit=4 failures,
    the search stops once rho < 0.01 -> 7 halvings = 28 iterations, 2 evals each."""
    cg, cl, ep = _flat_case()
    sp = search_params(num_runs=1, pop_size=2, ls="sw", lsearch_rate=100.0)
    init = np.zeros((2, 64), np.float32)
    init[:, 0:3] = 5.0
    j = O.OracleJob(cg, cl, ep, sp, init, O.cmwc_uints((1,), 2 * 64))
    j.initpop()
    j.gen()
    j.ls(0)
    s = j.state()
    assert list(s["evals_new"]) == [0 + 56, 1 + 56]


def test_adadelta_autostop_on_flat_energy():
    """Zero gradient: the first evaluation improves on +inf, then 4 failures
    halve varrho each time; varrho < 0.01 after 7 halvings -> 1 + 28 iterations."""
    cg, cl, ep = _flat_case()
    sp = search_params(num_runs=1, pop_size=2, ls="ad", lsearch_rate=100.0)
    init = np.zeros((2, 64), np.float32)
    init[:, 0:3] = 5.0
    j = O.OracleJob(cg, cl, ep, sp, init, O.cmwc_uints((1,), 2 * 64))
    j.initpop()
    j.gen()
    j.ls(0)
    s = j.state()
    assert list(s["evals_new"]) == [29, 30]


def test_adadelta_first_step_closed_form():
    """Linear energy E = c*x: the first ADADELTA step is
    -g * sqrt(eps / ((1-rho) g^2 + eps)) (kernel_ad.cl:434-456)."""
    c = 0.4
    cg = _grid_from_field(lambda x, y, z: c * x)
    cl = CLigand(helpers.single_atom_tables(0, 0.0))
    sp = search_params(num_runs=1, pop_size=2, ls="ad", max_iters=1)
    init = np.zeros((2, 64), np.float32)
    init[:, 0:3] = 6.0
    j = O.OracleJob(cg, cl, energy_params(), sp, init, O.cmwc_uints((2,), 2 * 64))
    j.initpop()
    j.gen()
    pre = j.state()["pop_next"][0, 0]
    j.ls(0)
    # max_iters=1: best genotype is the evaluated one (unchanged); run two iterations
    sp2 = search_params(num_runs=1, pop_size=2, ls="ad", max_iters=2)
    j2 = O.OracleJob(cg, cl, energy_params(), sp2, init, O.cmwc_uints((2,), 2 * 64))
    j2.initpop()
    j2.gen()
    j2.ls(0)
    got = j2.state()["pop_next"][0, 0]
    g = np.float32(c)   # dE/dx in grid units (gradient gene 0 = force * spacing, map gradient / spacing)
    g = np.float32(round(float(g) * 1024) / 1024)
    step = -g * np.sqrt(np.float32(0.01) / (np.float32(0.19999999) * g * g + np.float32(0.01)))
    # second evaluation is lower, so best = first step
    assert got == pytest.approx(pre + step, abs=2e-3)


def test_ls_entity_choice_and_counts(c1):
    """With lsearch_rate 100 %, every entity is searched once per generation."""
    sp = search_params(num_runs=2, pop_size=10, ls="adam", max_iters=7, spacing=c1.spacing)
    init = O.initial_population((9,), 2, 10, c1.maps.size_angstr, c1.spacing)
    j = O.OracleJob(c1.cg, c1.cl, c1.ep, sp, init, O.cmwc_uints((9,), 2 * 10 * 64))
    j.initpop()
    j.gen()
    j.ls(0)
    s = j.state()
    assert list(s["evals_new"]) == [7] + [8] * 9 + [7] + [8] * 9
