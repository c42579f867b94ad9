"""GPU parity through the C ABI against the oracle.

Bars (stated once, used below):
  * energy: |E_gpu - E_oracle32| <= 1e-4 * max(1, S) for at least 98 % of the
    poses, where S is the oracle's sum of |energy terms| (FP32 error of a sum
    scales with S), and for every pose once the energy's sensitivity to FP32
    coordinate rounding is added (helpers.energy_bar);
  * merged gradient, translation and torsion genes (well conditioned):
    |g_gpu - g_oracle32| <= 1e-4 |g| + 4 quanta of the int32 fixed point
    (a per-term rounding flip moves a gene by one quantum);
  * orientation genes and the float FIRE gradient are ill-conditioned in FP32
    (FP32 and FP64 evaluations of the same formulas already differ by more than
    1e-4 relative): the GPU's error against the FP64 oracle must stay within 2x
    the FP32 oracle's own error against FP64, at the median and 99th percentile.
"""
import numpy as np
import pytest

import adgpu
import helpers
import oracle as O
from adgpu import synth
from adgpu._abi import CGrid, CLigand, energy_params, search_params

pytestmark = pytest.mark.gpu


def _upload(ctx, case, fmt=adgpu.MAP_F32):
    mid = ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing, fmt)
    lid = ctx.ligand_upload(case.tables, case.ep)
    return mid, lid


def _cases():
    return {
        "c1": dict(n_atoms=30, n_tors=6, lig_seed=7, map_seed=11, npts=60),
        "small_rigid": dict(n_atoms=12, n_tors=0, lig_seed=3, map_seed=5, npts=40),
        "a90_t20": dict(n_atoms=90, n_tors=20, lig_seed=5, map_seed=13, npts=64),
        "a150_t30": dict(n_atoms=150, n_tors=30, lig_seed=8, map_seed=17, npts=72),
    }


@pytest.fixture(scope="module", params=list(_cases().keys()))
def case(request):
    kw = _cases()[request.param]
    c = helpers.synthetic_case(kw["n_atoms"], kw["n_tors"], kw["lig_seed"], kw["map_seed"], npts=kw["npts"])
    c.name = request.param
    return c


def test_energy_parity(gpu_ctx, case):
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 1500, seed=101)
    e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
    helpers.check_energies(case.orc, poses, e)
    # energy-only and merged kernels compute the same energy
    e2, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    helpers.check_energies(case.orc, poses, e2, e_ref=e)


def test_energrad_parity(gpu_ctx, case):
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 1500, seed=202)
    e, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    ng = case.ng
    well = [0, 1, 2] + list(range(6, ng))
    qv = np.array([helpers.gene_quantum(case, k) for k in range(ng)])
    err_gpu, err_32, viol = [], [], []
    helpers.check_energies(case.orc, poses, e)
    for i, p in enumerate(poses):
        eo, go = case.orc.energrad32(p, True)
        d = np.abs(g[i, :ng] - go)
        tol = 1e-4 * np.abs(go) + 4 * qv
        viol.extend([(i, k, float(g[i, k]), float(go[k])) for k in well if d[k] > tol[k]])
        if case.tables.true_ligand_atoms:
            _, g64 = case.orc.energrad64(p, True)
            err_gpu.append(np.abs(g[i, 3:6] - g64[3:6]))
            err_32.append(np.abs(go[3:6] - g64[3:6]))
    assert len(viol) <= 0.002 * len(poses) * len(well), viol[:10]
    assert np.all(g[:, ng:] == 0.0)
    if err_gpu:
        eg, e3 = np.array(err_gpu), np.array(err_32)
        q = qv[3]
        for pct in (50, 99):
            assert np.all(np.percentile(eg, pct, axis=0) <= 2 * np.percentile(e3, pct, axis=0) + 2 * q), pct


def test_float_gradient_parity(gpu_ctx, case):
    """gpu_calc_gradient (FIRE's first step, calcgradient.cl:109-1003)."""
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 600, seed=303)
    _, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_GRADIENT)
    ng = case.ng
    eg, e3, scale = [], [], []
    for i, p in enumerate(poses):
        g64 = case.orc.gradient64(p)
        g32 = case.orc.gradient32(p)
        eg.append(np.abs(g[i, :ng] - g64))
        e3.append(np.abs(g32 - g64))
        scale.append(np.abs(g64))
    eg, e3, scale = np.array(eg), np.array(e3), np.array(scale)
    for pct in (50, 99):
        lhs = np.percentile(eg, pct, axis=0)
        rhs = 2 * np.percentile(e3, pct, axis=0) + 1e-4 * np.percentile(scale, pct, axis=0)
        assert np.all(lhs <= rhs), (pct, lhs, rhs)


def test_energy_parity_fp16_and_bf16_maps(gpu_ctx):
    """Reduced-precision map storage with FP32 interpolation: the GPU must
    match the oracle evaluated on the same rounded map values."""
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    poses = helpers.poses_for(case, 500, seed=404)
    for fmt, dt in ((adgpu.MAP_F16, np.float16), (adgpu.MAP_BF16, None)):
        mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing, fmt)
        lid = gpu_ctx.ligand_upload(case.tables, case.ep)
        e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
        if dt is not None:
            # FP16 storage saturates at +-65504 (the maps clip repulsion at 1e5)
            rounded = np.clip(case.packed, -65504.0, 65504.0).astype(np.float16).astype(np.float32)
        else:
            u = case.packed.view(np.uint32).astype(np.uint64)
            u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
            rounded = u.astype(np.uint32).view(np.float32)
        orc = O.Oracle(CGrid(rounded, case.size, case.maps.num_map_atypes, case.spacing), case.cl, case.ep)
        helpers.check_energies(orc, poses, e)
        gpu_ctx.maps_release(mid)
        gpu_ctx.ligand_release(lid)


def test_reference_tables_fixture(gpu_ctx):
    """Device tables made by the reference's own host code (tests/golden/ref_c1.npz)."""
    f = helpers.load_ref_fixture("c1")
    T = helpers.tables_from_fixture(f)
    case = helpers.synthetic_case(int(f["n_atoms_req"]), int(f["n_tors_req"]), int(f["lig_seed"]),
                                  int(f["map_seed"]), npts=int(f["npts"]))
    ep = energy_params(coeff_elec=float(f["coeff_elec"]), coeff_desolv=float(f["coeff_desolv"]),
                       qasp=float(f["qasp"]), smooth=float(f["smooth"]), elec_min_distance=float(f["elec_min"]))
    mid = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing)
    lid = gpu_ctx.ligand_upload(T, ep)
    orc = O.Oracle(case.cg, CLigand(T), ep)
    poses = f["init_pop"]
    e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
    helpers.check_energies(orc, poses, e)


@pytest.mark.parametrize("complex_name", ["1stp", "3ce3"])
def test_golden_dlg_intermolecular_energy(gpu_ctx, complex_name):
    """The 200 docked poses of the reference's example logs: the GPU's
    intermolecular energy of each pose equals the log's "Final Intermolecular
    Energy" (2 decimals)."""
    dlg = helpers.load_dlg(complex_name)
    maps = dlg["maps"]
    nm, nz, ny, nx = maps.shape
    mid = gpu_ctx.maps_upload(synth.pack_maps(maps), (nx, ny, nz), nm - 2, float(dlg["spacing"]))
    labels = [str(x) for x in dlg["map_labels"]]
    tix = [labels.index(str(t)) for t in dlg["types"]]
    for dev in ("gpu", "cpu"):
        xyz = (dlg[dev + "_xyz"] - dlg["origo"]) / dlg["spacing"]
        for m in range(xyz.shape[0]):
            t, centre = helpers.rigid_tables(tix, dlg["charges"], xyz[m])
            lid = gpu_ctx.ligand_upload(t, energy_params())
            g = np.zeros((1, 64), np.float32)
            g[0, 0:3] = centre
            e, _ = gpu_ctx.eval_poses(mid, lid, g, adgpu.EVAL_ENERGY)
            assert abs(e[0] - dlg[dev + "_inter"][m]) < 0.05, (dev, m, e[0], dlg[dev + "_inter"][m])
            gpu_ctx.ligand_release(lid)


def test_million_scale_batch_properties(gpu_ctx):
    """1e5 seeded random poses at the C1 size: every energy is finite, a
    second launch gives identical bits, energies are invariant to batch
    position, and a subsample matches the oracle."""
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    mid, lid = _upload(gpu_ctx, case)
    poses = synth.random_poses(100000, case.ng, case.size, seed=2024)
    e1, g1 = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    e2, g2 = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    assert np.all(np.isfinite(e1)) and np.all(np.isfinite(g1))
    assert np.array_equal(e1, e2) and np.array_equal(g1, g2)
    perm = np.random.default_rng(0).permutation(len(poses))[:5000]
    e3, g3 = gpu_ctx.eval_poses(mid, lid, poses[perm], adgpu.EVAL_ENERGRAD)
    assert np.array_equal(e3, e1[perm]) and np.array_equal(g3, g1[perm])
    helpers.check_energies(case.orc, poses[perm[:400]], e1[perm[:400]])


@pytest.mark.parametrize("width,shape", [(32, (90, 20, 5, 13, 64)), (64, (90, 20, 5, 13, 64)),
                                         (16, (30, 6, 7, 11, 60)), (16, (14, 3, 3, 5, 40))])
def test_team_width_parity(gpu_ctx, width, shape):
    """Every team width (four, two or one individuals per wavefront) meets the
    energy bar and the gradient bar against the oracle: a ligand that needs
    one atom chunk per lane at width 64 and three at width 32, and ligands of
    two chunks and one chunk at width 16."""
    case = helpers.synthetic_case(*shape[:4], npts=shape[4])
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 400, seed=505)
    gpu_ctx.set_team_width(width)
    try:
        e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
        e2, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    finally:
        gpu_ctx.set_team_width(0)
    helpers.check_energies(case.orc, poses, e)
    helpers.check_energies(case.orc, poses, e2, e_ref=e)
    ng = case.ng
    qv = np.array([helpers.gene_quantum(case, k) for k in range(ng)])
    well = [0, 1, 2] + list(range(6, ng))
    viol = 0
    for i, p in enumerate(poses):
        _, go = case.orc.energrad32(p, True)
        d = np.abs(g[i, :ng] - go)
        viol += int(np.sum(d[well] > (1e-4 * np.abs(go) + 4 * qv)[well]))
    assert viol <= 0.002 * len(poses) * len(well), viol


@pytest.mark.parametrize("name", ["c1", "a90_t20"])
def test_fp64_rescoring_matches_fp64_oracle(gpu_ctx, name):
    """FP64 re-scoring on the GPU equals the FP64 oracle to 1e-10 of the sum of
    |terms| (double rounding plus the device's and libm's last-ulp sin/cos/pow
    differences); inter and intra separately."""
    kw = _cases()[name]
    case = helpers.synthetic_case(kw["n_atoms"], kw["n_tors"], kw["lig_seed"], kw["map_seed"], npts=kw["npts"])
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 300, seed=606)
    tot, inter, intra = gpu_ctx.rescore_f64(mid, lid, poses)
    for i, p in enumerate(poses):
        e, ei, ea, s = case.orc.energy64(p)
        tol = 1e-10 * max(1.0, s)
        assert abs(tot[i] - e) <= tol, (i, tot[i], e)
        assert abs(inter[i] - ei) <= tol and abs(intra[i] - ea) <= tol, (i, inter[i], ei, intra[i], ea)


def test_job_best_f64(gpu_ctx):
    """Each run's best individual of a short LGA job, re-scored in FP64: the
    index is the run's lowest FP32 energy and the FP64 energy is the oracle's."""
    case = helpers.synthetic_case(30, 6, 7, 11, npts=60)
    mid, lid = _upload(gpu_ctx, case)
    sp = search_params(num_runs=5, pop_size=30, max_evals=20000, ls="ad", spacing=case.spacing)
    init = O.initial_population((4, 4), 5, 30, case.maps.size_angstr, case.spacing)
    job = gpu_ctx.job(mid, lid, sp, init, O.cmwc_uints((4, 4), 5 * 30 * 64))
    job.initpop()
    job.run()
    st = job.download()
    b = job.best_f64()
    ene = st["ene"].reshape(5, 30)
    assert np.array_equal(b["index"], np.argmin(ene, axis=1))
    for r in range(5):
        e64, _, _, s = case.orc.energy64(b["genotypes"][r])
        assert abs(b["total"][r] - e64) <= 1e-10 * max(1.0, s)
        # the FP32 search energy of the same pose is within the FP32 bar of its FP64 value
        assert abs(b["energy_f32"][r] - e64) <= 1e-4 * max(1.0, s) + 1e-3
    job.destroy()


@pytest.fixture(scope="module")
def flex_case():
    return helpers.fixture_case("flex")


def test_flexres_energy_and_gradient_parity(gpu_ctx, flex_case):
    """Ligand plus two flexible residues, tables made by the reference's host
    code with --flexres (tests/golden/ref_flex.npz): residue atoms take only
    their own torsions, the first two atoms of each residue have no grid
    energy, ligand-residue pairs are intramolecular.  Energy bar and merged
    gradient bar as for the ligand-only cases, FP32 and BF16 maps."""
    case = flex_case
    poses = np.concatenate([case.fixture["init_pop"], helpers.poses_for(case, 800, seed=707)])
    mid, lid = _upload(gpu_ctx, case)
    e, _ = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGY)
    helpers.check_energies(case.orc, poses, e)
    e2, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_ENERGRAD)
    helpers.check_energies(case.orc, poses, e2, e_ref=e)
    ng = case.ng
    qv = np.array([helpers.gene_quantum(case, k) for k in range(ng)])
    well = [0, 1, 2] + list(range(6, ng))
    viol = []
    for i, p in enumerate(poses):
        _, go = case.orc.energrad32(p, True)
        d = np.abs(g[i, :ng] - go)
        tol = 1e-4 * np.abs(go) + 4 * qv
        viol.extend([(i, k, float(g[i, k]), float(go[k])) for k in well if d[k] > tol[k]])
    assert len(viol) <= 0.002 * len(poses) * len(well), viol[:10]
    assert np.all(g[:, ng:] == 0.0)
    # BF16 map storage against the oracle on the same rounded values
    u = case.packed.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    rounded = u.astype(np.uint32).view(np.float32)
    mid16 = gpu_ctx.maps_upload(case.packed, case.size, case.maps.num_map_atypes, case.spacing, adgpu.MAP_BF16)
    e3, _ = gpu_ctx.eval_poses(mid16, lid, poses, adgpu.EVAL_ENERGY)
    orc16 = O.Oracle(CGrid(rounded, case.size, case.maps.num_map_atypes, case.spacing), case.cl, case.ep)
    helpers.check_energies(orc16, poses, e3)
    gpu_ctx.maps_release(mid16)


def test_flexres_float_gradient_and_fp64(gpu_ctx, flex_case):
    """FIRE's float gradient (ligand atoms only in the translation and rotation
    genes, calcgradient.cl) within 2x the FP32 oracle's own error against
    FP64, and FP64 re-scoring equal to the FP64 oracle (1e-10 of the sum of
    |terms|), with flexible residues."""
    case = flex_case
    mid, lid = _upload(gpu_ctx, case)
    poses = helpers.poses_for(case, 300, seed=808)
    _, g = gpu_ctx.eval_poses(mid, lid, poses, adgpu.EVAL_GRADIENT)
    ng = case.ng
    eg, e3, scale = [], [], []
    for i, p in enumerate(poses):
        g64 = case.orc.gradient64(p)
        eg.append(np.abs(g[i, :ng] - g64))
        e3.append(np.abs(case.orc.gradient32(p) - g64))
        scale.append(np.abs(g64))
    eg, e3, scale = np.array(eg), np.array(e3), np.array(scale)
    for pct in (50, 99):
        lhs = np.percentile(eg, pct, axis=0)
        rhs = 2 * np.percentile(e3, pct, axis=0) + 1e-4 * np.percentile(scale, pct, axis=0)
        assert np.all(lhs <= rhs), (pct, lhs, rhs)
    tot, inter, intra = gpu_ctx.rescore_f64(mid, lid, poses)
    for i, p in enumerate(poses):
        e, ei, ea, s = case.orc.energy64(p)
        tol = 1e-10 * max(1.0, s)
        assert abs(tot[i] - e) <= tol and abs(inter[i] - ei) <= tol and abs(intra[i] - ea) <= tol, i


def test_flexres_lga_job(gpu_ctx, flex_case):
    """A short LGA + ADADELTA job on the flexible-residue case from the
    reference's own initial population and seeds: the downloaded energies are
    the oracle's FP32 energies of the downloaded genotypes, and the search
    lowers every run's best energy below its initial best."""
    case = flex_case
    f = case.fixture
    mid, lid = _upload(gpu_ctx, case)
    nrun, psize = int(f["nrun"]), int(f["psize"])
    sp = search_params(num_runs=nrun, pop_size=psize, max_evals=30000, ls="ad", spacing=case.spacing)
    job = gpu_ctx.job(mid, lid, sp, f["init_pop"], f["seeds"])
    job.initpop()
    e0 = job.download()["ene"].reshape(nrun, psize).min(axis=1)
    job.run()
    st = job.download()
    job.destroy()
    pop = st["pop"].reshape(nrun * psize, -1)
    helpers.check_energies(case.orc, pop, st["ene"].reshape(-1))
    assert np.all(st["ene"].reshape(nrun, psize).min(axis=1) < e0)
