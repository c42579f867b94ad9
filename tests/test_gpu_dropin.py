"""The reference's own CLI (host/src/main.cpp and the host sources it needs)
built with the MI355X backend in place of performdocking.cpp.Cuda
(autodock-gpu_amd/dropin/): a docking job from .fld/.map/.pdbqt files to the
dlg and xml outputs, through the C ABI.

The CLI is built where the reference sources are (autodock-gpu_amd/dropin/
Makefile) and travels as a binary; without it these tests are skipped."""
import os
import re
import subprocess

import numpy as np
import pytest

from adgpu import synth

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "autodock-gpu_amd", "dropin", "_build", "autodock_hip_64wi")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(CLI), reason="drop-in CLI not built (its build needs the "
                                                                 "reference's host sources)")]


def _inputs(d):
    lig = synth.make_ligand(30, 6, seed=7)
    maps = synth.make_maps(lig.type_names, npts=60, spacing=0.375, seed=11)
    fld = synth.write_maps(maps, str(d))
    pdbqt = synth.write_ligand(lig, os.path.join(str(d), "lig.pdbqt"))
    return fld, pdbqt


def _dock(d, name, *extra):
    fld, pdbqt = _inputs(d)
    res = os.path.join(str(d), name)
    cmd = [CLI, "--ffile", fld, "--lfile", pdbqt, "--resnam", res] + list(extra)
    print("running:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, (p.returncode, p.stdout[-3000:], p.stderr[-3000:])
    return res, p.stdout


def _docked_energies(dlg_text):
    # "DOCKED: USER    Estimated Free Energy of Binding    =   -7.12 kcal/mol ..."
    vals = re.findall(r"DOCKED: USER\s+Estimated Free Energy of Binding\s+=\s+([-+0-9.eE]+)", dlg_text)
    return np.array([float(v) for v in vals])


def test_cli_adadelta_writes_dlg_and_xml(tmp_path):
    res, out = _dock(tmp_path, "ad", "--nrun", "10", "--psize", "50", "--nev", "200000", "--seed", "3,5,7",
                     "--heuristics", "0", "--autostop", "0", "--lsmet", "ad")
    dlg = open(res + ".dlg", encoding="latin-1").read()
    assert os.path.exists(res + ".xml")
    e = _docked_energies(dlg)
    assert len(e) == 10 and np.all(np.isfinite(e)), e
    assert "CLUSTERING HISTOGRAM" in dlg
    # the search found poses inside the wells of the synthetic maps
    assert e.min() < 0.0, e


def test_cli_is_deterministic(tmp_path):
    args = ("--nrun", "6", "--psize", "40", "--nev", "100000", "--seed", "11,2,9", "--heuristics", "0",
            "--autostop", "0", "--lsmet", "ad")
    r1, _ = _dock(tmp_path, "a", *args)
    r2, _ = _dock(tmp_path, "b", *args)
    e1 = _docked_energies(open(r1 + ".dlg", encoding="latin-1").read())
    e2 = _docked_energies(open(r2 + ".dlg", encoding="latin-1").read())
    assert len(e1) == 6 and np.array_equal(e1, e2)


@pytest.mark.parametrize("lsmet", ["sw", "adam", "fire"])
def test_cli_other_local_searches_with_autostop(tmp_path, lsmet):
    res, out = _dock(tmp_path, lsmet, "--nrun", "6", "--psize", "40", "--nev", "150000", "--seed", "1,2,3",
                     "--heuristics", "0", "--autostop", "1", "--lsmet", lsmet)
    e = _docked_energies(open(res + ".dlg", encoding="latin-1").read())
    assert len(e) == 6 and np.all(np.isfinite(e)), (lsmet, e)


@pytest.mark.parametrize("lsmet", ["ad", "sw"])
def test_cli_device_autostop_equals_reference_host_autostop(tmp_path, lsmet):
    """The AutoStop test on the device against the reference's own AutoStop
    class on the host (ADGPU_HOST_AUTOSTOP=1): same table, same stop, same
    docked poses."""
    args = ("--nrun", "8", "--psize", "60", "--nev", "400000", "--seed", "4,4,4", "--heuristics", "0",
            "--autostop", "1", "--asfreq", "5", "--stopstd", "0.5", "--lsmet", lsmet)
    r1, out1 = _dock(tmp_path, "dev", *args)
    env = dict(os.environ, ADGPU_HOST_AUTOSTOP="1")
    fld, pdbqt = _inputs(tmp_path)
    r2 = os.path.join(str(tmp_path), "host")
    p = subprocess.run([CLI, "--ffile", fld, "--lfile", pdbqt, "--resnam", r2] + list(args), capture_output=True,
                       text=True, timeout=240, env=env)
    assert p.returncode == 0, p.stdout[-3000:]
    table = lambda t: [ln for ln in t.splitlines() if re.match(r"^\s+\d+ \|", ln) or "Finished" in ln
                       or "kcal/mol combined" in ln or "samples, best" in ln or "evaluations. Best" in ln]
    assert table(out1) == table(p.stdout) and len(table(out1)) > 2
    d1 = open(r1 + ".dlg", encoding="latin-1").read()
    d2 = open(r2 + ".dlg", encoding="latin-1").read()
    assert np.array_equal(_docked_energies(d1), _docked_energies(d2))


def _dlg_runs(dlg_text):
    """Per run of the dlg's DOCKED blocks: (run, inter, intra) as printed
    (processresult.cpp ligand_calc_output, %+7.2f)."""
    out = []
    for block in dlg_text.split("DOCKED: USER    Run = ")[1:]:
        run = int(block.split()[0])
        inter = float(re.search(r"\(1\) Final Intermolecular Energy\s+=\s+([-+0-9.eE]+)", block).group(1))
        intra = float(re.search(r"\(2\) Final Total Internal Energy\s+=\s+([-+0-9.eE]+)", block).group(1))
        out.append((run, inter, intra))
    return sorted(out)


def test_cli_fp64_rescoring_of_each_runs_best_pose(tmp_path):
    """The CLI prints the GPU's FP64 re-scoring of each run's best pose.  It
    agrees with the dlg's energies of the same pose, which the reference's
    CPU re-scoring computes (processresult.cpp:430-526): the intermolecular
    energy within the dlg's 2-decimal rounding (same trilinear rule), the
    intramolecular energy within 0.1 kcal/mol (the CPU uses distance tables;
    on 400 docked poses of this ligand the analytic FP64 energy and the
    tables differ by at most 0.036 kcal/mol, median 0.004)."""
    res, out = _dock(tmp_path, "f64", "--nrun", "10", "--psize", "50", "--nev", "200000", "--seed", "3,5,7",
                     "--heuristics", "0", "--autostop", "0", "--lsmet", "ad")
    rows = [ln.split("|") for ln in out.splitlines() if re.match(r"^FP64\s+\d+\s+\|", ln)]
    assert len(rows) == 10, out[-3000:]
    f64 = {int(r[0].split()[1]): [float(x) for x in r[2:6]] for r in rows}
    dlg = _dlg_runs(open(res + ".dlg", encoding="latin-1").read())
    assert [r for r, _, _ in dlg] == list(range(1, 11))
    for run, inter, intra in dlg:
        e32, tot, i64, a64 = f64[run]
        assert abs(tot - (i64 + a64)) <= 2e-4       # printed with 4 decimals
        # the FP32 energy the search kept for the pose
        assert abs(e32 - tot) <= 1e-4 * max(1.0, abs(tot)) + 2e-4, (run, e32, tot)
        assert abs(i64 - inter) <= 0.0051 + 1e-4 * abs(inter), (run, i64, inter)
        assert abs(a64 - intra) <= 0.1, (run, a64, intra)
