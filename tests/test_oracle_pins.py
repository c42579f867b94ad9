"""Pins of the oracle's orientation-gene gradient, its float gradient (FIRE's
first step), and its ADAM and FIRE update rules, by what the mathematics
fixes:

* orientation genes: rotating the current orientation by a small angle about
  the torque axis (space frame, after the current rotation) and reading the
  genotype angles of the result, computed here with rotation matrices instead
  of the oracle's quaternion products (calcMergedEneGra.cl:712-861,
  calcgradient.cl:714-873);
* float gradient: central differences of the analytic AD4 pair energies
  (12-6 / 12-10 with smoothing, Mehler-Solmajer electrostatics, Gaussian
  desolvation; calcgradient.cl:560-620), and of the oracle's FP64 energy where
  both use the same function;
* ADAM: on a linear energy every bias-corrected step is g / (|g| + eps) per
  gene (cuda/kernel_adam.cu:280-298);
* FIRE: closed-form trajectories on linear, V-shaped and kinked fields
  (kernel_fire.cl:40-56, 389-414, 630-760): dt0 = DT_MAX/3, v0 = -ALPHA_START
  g/|g|, the DT_INC / DT_MAX / ALPHA_DEC schedule after more than SUCCESS_MIN
  successes, and the reset on negative power.

tests/test_oracle_mutants.py plants a plausible mistake in each of these
functions of the oracle and checks that a test here fails."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params, search_params

# the reference's float constants (common/calcenergy_basic.h, kernel_fire.cl)
DEG = float(np.float32(0.0174533))
PI_F = float(np.float32(math.pi))
TWO_PI_F = 2.0 * PI_F
INF_RAD = float(np.float32(1e-3))
ROT_STEP = 2.0 * float(np.float32(0.5) * np.float32(1e-3))    # rotation of the large-torque branch
INV_INF_RAD = float(np.float32(1.0) / np.float32(1e-3))


def _grid_from_field(f, n=16, spacing=0.375, ntypes=1):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = f(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)).astype(np.float32)
    return CGrid(synth.pack_maps(np.stack([v] * (ntypes + 2))), (n, n, n), ntypes, spacing)


# --------------------------------------------------------------------------
# orientation genes
# --------------------------------------------------------------------------
def _rotmat(axis, ang):
    """Rodrigues: active rotation by ang about axis (right hand)."""
    n = np.asarray(axis, np.float64)
    n = n / np.linalg.norm(n)
    K = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    return np.eye(3) + math.sin(ang) * K + (1.0 - math.cos(ang)) * (K @ K)


def _genes_rotation(geno):
    """Rotation of genes 3..5: angle alpha about the axis with polar angle
    theta and azimuth phi (the genotype convention the coordinate tests pin
    against the reference's change_conform)."""
    phi, theta, alpha = (float(np.float32(geno[k])) * DEG for k in (3, 4, 5))
    axis = (math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))
    return _rotmat(axis, alpha), phi, theta, alpha


def _wrap(x, two_pi=2 * math.pi, pi=math.pi):
    return (x + pi) % two_pi - pi


def orientation_expected(geno, torque, spacing, acos=math.acos, exact_pi=False):
    """Orientation-gene gradient by its definition: the change of (phi,
    theta, alpha) when the pose is turned about the torque axis, scaled to the
    torque; in the reported units (x spacing^2 / DEG_TO_RAD^2, the unit change
    of calcMergedEneGra.cl:927-935).  acos / exact_pi: the arc cosine and the
    2 pi the angles are reduced with (the reference's are fast_acos and
    2 * (float)pi)."""
    R, phi, theta, alpha = _genes_rotation(geno)
    two_pi = 2 * math.pi if exact_pi else TWO_PI_F
    pi = math.pi if exact_pi else PI_F
    t = float(np.linalg.norm(torque))
    if t < INF_RAD:
        step, scale = t, 1.0
    else:
        step, scale = ROT_STEP, t * INV_INF_RAD
    Rt = _rotmat(torque, step) @ R if t > 0 else R
    # angle and axis of Rt, on the side of the current alpha
    a1 = math.acos(min(1.0, max(-1.0, (np.trace(Rt) - 1.0) / 2.0)))
    n1 = np.array([Rt[2, 1] - Rt[1, 2], Rt[0, 2] - Rt[2, 0], Rt[1, 0] - Rt[0, 1]]) / (2.0 * math.sin(a1))
    ac = alpha % (2 * math.pi)
    if abs(_wrap(a1 - ac)) <= abs(_wrap(2 * math.pi - a1 - ac)):
        at, nt = a1, n1
    else:
        at, nt = 2 * math.pi - a1, -n1
    at = 2.0 * acos(math.cos(at / 2.0))
    if math.sin(theta) >= 0.0:
        tht = acos(nt[2])
        pht = math.atan2(nt[1], nt[0])
    else:
        tht = two_pi - acos(nt[2])
        pht = math.atan2(-nt[1], -nt[0])
    if pht < 0.0:
        pht += two_pi
    cur_phi, cur_theta = phi % two_pi, theta % two_pi
    dphi = _wrap(pht - cur_phi, two_pi, pi)
    dtheta = _wrap(tht - cur_theta, two_pi, pi)
    dalpha = _wrap(at - ac, two_pi, pi)
    corr = 4.0 * math.sin(alpha / 2.0) ** 2
    conv = spacing * spacing / (DEG * DEG)
    return np.array([scale * dphi * math.sin(theta) ** 2 * corr * DEG,
                     scale * dtheta * corr * DEG,
                     scale * dalpha * DEG]) * conv, t


def _torque(orc, geno, spacing):
    """Sum over ligand atoms of (r - centre) x dE/dr (kcal/mol), from the
    oracle's per-atom FP64 gradient."""
    c = orc.coords64(geno)
    F = orc.atom_grad64(geno)
    n = orc.cl.s.true_ligand_atoms
    r = (c[:n] - np.asarray(geno[0:3], np.float32).astype(np.float64)) * spacing
    return np.cross(r, F[:n]).sum(axis=0)


def _orientation_poses(case, n, seed):
    """In-grid poses whose orientation stays clear of alpha = 0, pi and of
    theta = 0, pi (where the angles are degenerate)."""
    rng = np.random.default_rng(seed)
    g = helpers.poses_for(case, n, seed=seed, frac_central=1.0, spread=3.0)
    a = rng.uniform(20.0, 160.0, size=n)
    a[n // 2:] += 180.0                                    # alpha beyond pi too
    th = rng.uniform(15.0, 165.0, size=n)
    th[::3] += 180.0                                       # sin(theta) < 0
    g[:, 3] = rng.uniform(0.0, 360.0, size=n)
    g[:, 4] = th
    g[:, 5] = a
    return g.astype(np.float32)


def test_fast_acos_bound():
    """fast_acos (calcenergy.cl:93-108) is acos within 2.5e-6 rad on [-1, 1]
    (the polynomial fit, plus (float)pi for negative arguments)."""
    xs = np.concatenate([np.linspace(-1.0, 1.0, 20001), [-1.0, 0.0, 1.0, 0.5, -0.5]])
    f = O.lib().orc_fast_acos_f64
    err = max(abs(f(float(x)) - math.acos(x)) for x in xs)
    assert err < 2.5e-6, err
    assert abs(f(1.0)) < 1e-9


@pytest.mark.parametrize("tight", [True, False])
def test_orientation_gradient_energrad(c1, tight):
    """Genes 3..5 of the merged gradient (calcMergedEneGra.cl:756-861),
    unquantized FP64, against the definition on the 30-atom ligand.  tight:
    with the reference's fast_acos and float 2 pi (agreement to FP64 rounding);
    otherwise exact acos and pi: fast_acos's 2.5e-6 rad error over the 1e-3
    rad turn, and (float)pi's 1.8e-7 rad, bound the difference by 0.25 % of
    the torque for theta (0.5 % for alpha, twice the arc cosine) times the
    gene's own factors."""
    facos = O.lib().orc_fast_acos_f64
    for g in _orientation_poses(c1, 40, seed=31):
        e, grad = c1.orc.energrad64(g, quantize=False)
        tq = _torque(c1.orc, g, c1.spacing)
        unit = c1.spacing ** 2 / DEG
        if tight:
            exp, t = orientation_expected(g, tq, c1.spacing, acos=facos)
            tol = 1e-7 * t * unit * np.ones(3)
        else:
            exp, t = orientation_expected(g, tq, c1.spacing, exact_pi=True)
            th, al = float(g[4]) * DEG, float(g[5]) * DEG
            corr = 4.0 * math.sin(al / 2) ** 2
            tol = t * unit * np.array([2e-4 * corr * math.sin(th) ** 2, 2.6e-3 * corr, 5.2e-3])
        assert t > INF_RAD
        assert np.all(np.abs(grad[3:6] - exp) <= tol + 1e-9), (g[3:6], grad[3:6], exp)


def test_orientation_gradient_small_torque():
    """|torque| < 1e-3: the turn is by |torque| itself and the change is not
    rescaled (the small-angle branch of calcMergedEneGra.cl:756-790)."""
    rng = np.random.default_rng(5)
    xyz = np.array([[7.0, 7.5, 8.0], [8.5, 7.0, 7.2], [7.8, 8.9, 8.4], [6.9, 8.1, 6.6]])
    t, centre = helpers.rigid_tables([0, 0, 0, 0], [0.0] * 4, xyz)
    cl = CLigand(t)
    for k in range(12):
        s = 10.0 ** rng.uniform(-7.0, -5.5)
        cx, cy, cz = rng.uniform(-1, 1, size=3) * s
        cg = _grid_from_field(lambda x, y, z: cx * x * y + cy * y * z + cz * z * x + 0.3 * s * x * x)
        orc = O.Oracle(cg, cl, energy_params())
        g = np.zeros(64, np.float32)
        g[0:3] = centre
        g[3:6] = (rng.uniform(0, 360), rng.uniform(20, 160), rng.uniform(30, 150))
        e, grad = orc.energrad64(g, quantize=False)
        tq = _torque(orc, g, 0.375)
        exp, tn = orientation_expected(g, tq, 0.375, acos=O.lib().orc_fast_acos_f64)
        assert 0 < tn < INF_RAD
        assert np.allclose(grad[3:6], exp, rtol=1e-6, atol=1e-12), (grad[3:6], exp)


def test_orientation_gradient_float_path():
    """The float gradient's orientation genes (calcgradient.cl:714-873) on a
    rigid ligand, whose per-atom forces are the map gradient in both paths."""
    rng = np.random.default_rng(8)
    n = 20
    vals = rng.normal(0.0, 0.4, size=(3, n, n, n)).astype(np.float32)
    cg = CGrid(synth.pack_maps(vals), (n, n, n), 1, 0.375)
    xyz = rng.uniform(-2.5, 2.5, size=(6, 3)) + 10.0
    t, centre = helpers.rigid_tables([0] * 6, list(rng.uniform(-0.4, 0.4, size=6)), xyz)
    orc = O.Oracle(cg, CLigand(t), energy_params())
    for k in range(20):
        g = np.zeros(64, np.float32)
        g[0:3] = centre + rng.uniform(-1, 1, size=3)
        g[3:6] = (rng.uniform(0, 360), rng.uniform(15, 165) + (180 if k % 3 == 0 else 0),
                  rng.uniform(20, 160) + (180 if k % 2 else 0))
        grad = orc.gradient64(g)
        tq = _torque(orc, g, 0.375)
        exp, tn = orientation_expected(g, tq, 0.375, acos=O.lib().orc_fast_acos_f64)
        unit = tn * 0.375 ** 2 / DEG
        assert np.all(np.abs(grad[3:6] - exp) <= 1e-7 * unit + 1e-9), (grad[3:6], exp)


# --------------------------------------------------------------------------
# float gradient (calcgradient.cl:109-1003)
# --------------------------------------------------------------------------
MS_A, MS_L, MS_K = -8.5525, 0.003627, 7.7839      # Mehler-Solmajer 1991, eps0 = 78.4
MS_B = 78.4 - MS_A
SIGMA = 3.6                                        # AD4 desolvation Gaussian width (A)


def _pair_case(types, charges, reqm_override=None):
    """Atom 0 is the ligand (moved by genes 0..2), atom 1 a fixed
    flexible-residue atom on the +x axis; one intramolecular pair, no maps."""
    names = list(dict.fromkeys(types))
    reqm, vw_exp, reqm_ab, ac, bd, ds_s, ds_v = synth.vdw_tables(names)
    t = helpers.single_atom_tables(0, 0.0)
    t.num_atoms, t.true_ligand_atoms, t.num_atypes = 2, 1, len(names)
    t.charges = np.array(charges, np.float32)
    t.atom_types = np.array([names.index(x) for x in types], np.int32)
    t.atom_map = np.zeros(2, np.int32)
    t.ignore_inter = np.ones(2, np.uint8)
    t.pairs = np.array([0, 1], np.int32)
    t.atype_reqm = np.asarray(reqm if reqm_override is None else reqm_override, np.uint32)
    t.vw_exp = vw_exp.reshape(-1)
    t.reqm_ab = reqm_ab.astype(np.float32).reshape(-1)
    t.vw_ac, t.vw_bd = ac.astype(np.float32).reshape(-1), bd.astype(np.float32).reshape(-1)
    t.ds_s, t.ds_v = ds_s.astype(np.float32), ds_v.astype(np.float32)
    t.ref_coords = np.array([0, 0, 0, 12.0, 0, 0], np.float32)
    return t


def _analytic_pair_energy(t, ep, r, g_bond=False):
    """AD4 pair energy with the exact functions: smoothed 12-6/12-10 (r < 8 A),
    Mehler-Solmajer electrostatics and Gaussian desolvation (r < 20.48 A,
    distance clamped at elec_min_distance), CG-G0 bond 50 r."""
    nt = t.num_atypes
    t1, t2 = int(t.atom_types[0]), int(t.atom_types[1])
    idx = t1 * nt + t2
    exps = int(t.vw_exp[idx])
    m, n = (exps >> 8) & 0xFF, exps & 0xFF
    reqm = float(t.reqm_ab[idx])
    h = 0.5 * ep.smooth
    sd = r + h if r <= reqm - h else (reqm if r < reqm + h else r - h)
    e = 0.0
    if r < 8.0:
        e += float(t.vw_ac[idx]) / sd ** m - float(t.vw_bd[idx]) / sd ** n
    if r < 20.48:
        d = max(r, ep.elec_min_distance)
        q1, q2 = float(t.charges[0]), float(t.charges[1])
        eps_r = MS_A + MS_B / (1.0 + MS_K * math.exp(-MS_L * MS_B * d))
        e += ep.coeff_elec * q1 * q2 / (d * eps_r)
        w = ((float(t.ds_s[t1]) + ep.qasp * abs(q1)) * float(t.ds_v[t2]) +
             (float(t.ds_s[t2]) + ep.qasp * abs(q2)) * float(t.ds_v[t1]))
        e += w * ep.coeff_desolv * math.exp(-d * d / (2.0 * SIGMA * SIGMA))
    if g_bond:
        e += 50.0 * r
    return e


@pytest.mark.parametrize("types,charges,bond", [(("C", "OA"), (0.4, -0.3), False),
                                                (("HD", "OA"), (0.25, -0.45), False),
                                                (("N", "A"), (-0.2, 0.1), True)])
def test_float_gradient_pair_terms(types, charges, bond):
    """Gene 0 of the float gradient is -spacing dE/dr of the analytic pair
    energy: central differences at distances on every side of the smoothing
    plateau and of the 8 A / 20.48 A cutoffs (calcgradient.cl:560-620)."""
    ep = energy_params()
    spacing = 0.375
    t = _pair_case(types, charges, reqm_override=[24, 25] if bond else None)
    cg = _grid_from_field(lambda x, y, z: 0 * x, n=80, spacing=spacing)
    orc = O.Oracle(cg, CLigand(t), ep)
    for r in (1.6, 2.4, 3.3, 4.6, 6.2, 7.9, 8.3, 11.0, 16.5, 20.3, 20.7):
        g = np.zeros(64, np.float32)
        g[0] = np.float32(12.0 - r / spacing)
        r32 = (12.0 - float(g[0])) * spacing
        hh = 1e-5
        fd = (_analytic_pair_energy(t, ep, r32 + hh, bond) - _analytic_pair_energy(t, ep, r32 - hh, bond)) / (2 * hh)
        got = orc.gradient64(g)[0]
        # 1e-4: the reference's rounded Gaussian constants (0.077160, 0.038580)
        assert got == pytest.approx(-spacing * fd, rel=1e-4, abs=1e-9), (r, got, -spacing * fd)


def test_float_gradient_matches_energy_derivative(c1):
    """Without charges and desolvation both the energy and the float gradient
    use the same functions (maps, smoothed vdW): translation and torsion genes
    are the central differences of the FP64 energy."""
    t = SimpleNamespace(**vars(c1.tables))
    t.charges = np.zeros_like(t.charges)
    t.ds_s = np.zeros_like(t.ds_s)
    t.ds_v = np.zeros_like(t.ds_v)
    orc = O.Oracle(c1.cg, CLigand(t), energy_params())
    conv = c1.spacing ** 2 / DEG ** 2
    checked = 0
    for g in helpers.poses_for(c1, 200, seed=41, frac_central=1.0, spread=3.0):
        e = orc.energy64(g)[0]
        if not abs(e) < 50:
            continue
        frac = np.abs(orc.coords64(g) - np.round(orc.coords64(g)))
        if frac.min() < 2e-3:
            continue
        an = orc.gradient64(g)
        for k in list(range(3)) + list(range(6, c1.ng)):
            h = 2e-4 if k < 3 else 2e-3
            gp, gm = g.astype(np.float64).copy(), g.astype(np.float64).copy()
            gp[k] += h
            gm[k] -= h
            fd = (orc.energy64(gp.astype(np.float32))[0] - orc.energy64(gm.astype(np.float32))[0]) / (
                np.float32(g[k] + h) - np.float32(g[k] - h))
            a = an[k] if k < 3 else an[k] / conv
            assert a == pytest.approx(fd, rel=2e-3, abs=2e-3), (k, a, fd)
        checked += 1
    assert checked >= 10


# --------------------------------------------------------------------------
# ADAM (cuda/kernel_adam.cu:280-298)
# --------------------------------------------------------------------------
def _single_atom_job(field, start, sp, n=16):
    cg = _grid_from_field(field, n=n)
    cl = CLigand(helpers.single_atom_tables(0, 0.0))
    init = np.zeros((sp.pop_size * sp.num_runs, 64), np.float32)
    init[:, 0:6] = start
    j = O.OracleJob(cg, cl, energy_params(), sp, init, O.cmwc_uints((4,), sp.num_runs * sp.pop_size * 64))
    j.initpop()
    j.gen()            # entity 0 of the next population is the elite copy of `start`
    return j, O.Oracle(cg, cl, energy_params())


@pytest.mark.parametrize("eps", [1e-8, 0.25])
def test_adam_linear_field_closed_form(eps):
    """E = a x + b y + c z: the gradient g is constant, so each bias-corrected
    step is exactly g / (|g| + eps) per gene; the best genotype is the last one
    evaluated (iteration N-1) and the search counts N evaluations."""
    N = 9
    sp = search_params(num_runs=1, pop_size=1, ls="adam", max_iters=N, adam_epsilon=eps)
    start = (12.3, 11.6, 12.9, 30.0, 70.0, 100.0)
    field = lambda x, y, z: 0.75 * x + 0.5 * y + 0.25 * z
    j, orc = _single_atom_job(field, start, sp)
    j.ls(0)
    s = j.state()
    g = orc.energrad32(np.array(start + (0,) * 58, np.float32))[1].astype(np.float64)
    assert g[0] > 0.7 and g[1] > 0.45 and g[2] > 0.2 and np.all(g[3:6] == 0.0)
    step = g / (np.abs(g) + eps)
    exp = np.array(start, np.float64) - (N - 1) * step
    assert np.allclose(s["pop_next"][0, 0:6], exp, atol=5e-5), (s["pop_next"][0, 0:6], exp)
    assert s["evals_new"][0] == N
    assert s["ene_next"][0] == pytest.approx(0.75 * exp[0] + 0.5 * exp[1] + 0.25 * exp[2], abs=1e-4)


# --------------------------------------------------------------------------
# FIRE (kernel_fire.cl)
# --------------------------------------------------------------------------
FIRE_DT0, FIRE_DT_INC, FIRE_DT_DEC, FIRE_DT_MAX = 10.0 / 3.0, 1.2, 0.8, 10.0
FIRE_A0, FIRE_A_DEC, FIRE_NMIN = 0.2, 0.99, 5


def _fire_dts(n):
    """Step sizes of n successful iterations: DT_MAX/3 until more than
    SUCCESS_MIN successes, then x DT_INC capped at DT_MAX."""
    dts, dt = [], FIRE_DT0
    for k in range(1, n + 1):
        dts.append(dt)
        if k > FIRE_NMIN:
            dt = min(dt * FIRE_DT_INC, FIRE_DT_MAX)
    return dts


@pytest.mark.parametrize("N", [1, 6, 13, 15])
def test_fire_linear_field_schedule(N):
    """E = a x: v stays -ALPHA_START x^ (the mixing keeps its length on a
    constant gradient), so x moves by -0.2 dt per iteration with dt from the
    DT_INC / DT_MAX schedule; the best genotype is the last one."""
    sp = search_params(num_runs=1, pop_size=1, ls="fire", max_iters=N)
    start = (21.5, 9.25, 10.5, 30.0, 60.0, 90.0)
    j, orc = _single_atom_job(lambda x, y, z: 0.8 * x, start, sp, n=24)
    j.ls(0)
    s = j.state()
    exp_x = start[0] - FIRE_A0 * sum(_fire_dts(N))
    assert s["pop_next"][0, 0] == pytest.approx(exp_x, abs=2e-5)
    assert s["pop_next"][0, 1] == pytest.approx(start[1], abs=1e-6)
    assert s["pop_next"][0, 2] == pytest.approx(start[2], abs=1e-6)
    assert s["evals_new"][0] == N
    assert s["ene_next"][0] == pytest.approx(0.8 * exp_x, abs=1e-4)


def test_fire_v_shaped_field_resets():
    """E = a |x - 8|, starting at 8.5 (hand-derived, kernel_fire.cl:685-760):
    1: x = 8 - 1/6 overshoots: power < 0, so v = -0.2 g_old/|g_old| = -0.2,
       dt = 0.8 dt0, alpha = 0.2; the mix with the new gradient leaves
       v = 0.8 (-0.2) + 0.2 * 0.2 = -0.12;
    2: x = 8 - 1/6 - 0.12 (8/3) = 8 - 0.48667: power < 0 again, v = +0.2 from
       the old gradient, dt = 0.64 dt0; the mix keeps v = +0.2;
    3: x = 8 - 0.06 (power > 0), the lowest energy of the search;
    4: x = 8 + 0.36667: reset, dt = 0.512 dt0, v mixes to +0.12;
    5: x = 8 + 0.57147: reset, dt = 0.4096 dt0, v = -0.2;  6: x = 8 + 0.29840."""
    xs = [8 - 1 / 6, 8 - 1 / 6 - 0.32, 8 - 0.06, 8 + 0.36666667, 8 + 0.36666667 + 0.12 * 0.512 * FIRE_DT0]
    xs.append(xs[-1] - 0.2 * 0.4096 * FIRE_DT0)
    for N in range(1, 7):
        sp = search_params(num_runs=1, pop_size=1, ls="fire", max_iters=N)
        j, orc = _single_atom_job(lambda x, y, z: 0.8 * np.abs(x - 8.0), (8.5, 7.25, 9.5, 30.0, 60.0, 90.0), sp)
        j.ls(0)
        s = j.state()
        best = min(xs[:N], key=lambda x: abs(x - 8.0))
        assert s["pop_next"][0, 0] == pytest.approx(best, abs=2e-5), N
        assert s["evals_new"][0] == N
        assert s["ene_next"][0] == pytest.approx(0.8 * abs(best - 8.0), abs=1e-4)


def test_fire_alpha_decay_at_a_kink():
    """E = a x + a |y - 8| from (13.8, 11.8): v = -0.2 (1,1)/sqrt2 for seven
    successful iterations (dt: 10/3 x 6, then 4); the eighth (dt 4.8) crosses
    y = 8, where power is 0 (a success): count 8 > SUCCESS_MIN gives
    alpha = 0.2 * 0.99^3, dt = 5.76, and the mix turns v to
    -0.2/sqrt2 (1, 1 - 2 alpha); the ninth candidate is the lowest.  x and y
    keep equal fractional bits up to the crossing, so both gradients stay
    bitwise equal and power is exactly 0 there."""
    sp = search_params(num_runs=1, pop_size=1, ls="fire", max_iters=9)
    j, orc = _single_atom_job(lambda x, y, z: 0.75 * x + 0.75 * np.abs(y - 8.0), (13.8, 11.8, 9.4, 30.0, 60.0, 90.0),
                              sp)
    j.ls(0)
    s = j.state()
    u = FIRE_A0 / math.sqrt(2.0)
    d8 = u * (6 * FIRE_DT0 + 4.0 + 4.8)
    alpha = FIRE_A0 * FIRE_A_DEC ** 3
    x9 = 13.8 - d8 - 5.76 * u
    y9 = 11.8 - d8 - 5.76 * u * (1 - 2 * alpha)
    assert 11.8 - u * (6 * FIRE_DT0 + 4.0) > 8.0 > 11.8 - d8
    assert s["pop_next"][0, 0] == pytest.approx(x9, abs=3e-5)
    assert s["pop_next"][0, 1] == pytest.approx(y9, abs=3e-5)
    assert s["evals_new"][0] == 9
