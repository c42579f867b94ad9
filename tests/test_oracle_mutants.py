"""The oracle's pins have teeth: each case plants one plausible mistake in a
copy of oracle/adock_oracle.c (a dropped term, a wrong sign or index, a
swapped operand, a wrong constant), builds that copy, and checks that the pin
tests of tests/test_oracle_pins.py fail on it.  Everything runs on the CPU;
the real oracle is not touched."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE = os.path.join(REPO, "oracle")

# (name, text in adock_oracle.c, replacement, pin tests expected to fail)
MUTANTS = [
    ("atan2-swapped",
     "atan2(P->is_theta_gt_pi * target_q.y, P->is_theta_gt_pi * target_q.x)",
     "atan2(P->is_theta_gt_pi * target_q.x, P->is_theta_gt_pi * target_q.y)",
     "orientation"),
    ("body-frame-turn", "V4 target_q = qmul(qt, P->genrot_q);", "V4 target_q = qmul(P->genrot_q, qt);",
     "orientation"),
    ("phi-factor-dropped", "out[0] = gphi * P->sin_theta * P->sin_theta * rot_angle_corr * DEG_TO_RAD;",
     "out[0] = gphi * rot_angle_corr * DEG_TO_RAD;", "orientation"),
    ("theta-corr-dropped", "out[1] = gtheta * rot_angle_corr * DEG_TO_RAD;", "out[1] = gtheta * DEG_TO_RAD;",
     "orientation"),
    ("torque-sign", "V4 tr = vcross(r, F[a]);", "V4 tr = vcross(F[a], r);", "orientation_gradient_energrad"),
    ("half-turn-cos", "cos_half = cos(HALF_INF_RAD);", "cos_half = cos(INF_RAD);", "orientation"),
    ("small-torque-series", "(tl < INF_RAD) ? (R)1.0f - tl * tl * (R)0.125f : cos_half",
     "(tl < INF_RAD) ? (R)1.0f - tl * tl * (R)0.25f : cos_half", "small_torque"),
    ("fast-acos-constant", "const R e = K(2.74084), f = K(0.370388);", "const R e = K(2.74184), f = K(0.370388);",
     "fast_acos"),
    ("adam-bias-correction-dropped", "float mp = mt[k] / b1p, vp = vt[k] / b2p;", "float mp = mt[k], vp = vt[k];",
     "adam"),
    ("adam-eps-inside-sqrt", "geno[k] -= mp / (sqrtf(vp) + eps);", "geno[k] -= mp / sqrtf(vp + eps);", "adam"),
    ("adam-beta2-power-uses-beta1", "float b2p = 1.0f - powf(b2, 1.0f + (float)iters);",
     "float b2p = 1.0f - powf(b1, 1.0f + (float)iters);", "adam"),
    ("fire-dt0", "float alpha = 0.2f, dt = 10.0f / 3.0f;", "float alpha = 0.2f, dt = 10.0f;", "fire"),
    ("fire-alpha-not-decayed", "alpha = alpha * 0.99f;", "alpha = alpha * 1.0f;", "fire_alpha"),
    ("fire-dt-dec", "count_success = 0; alpha = 0.2f; dt = dt * 0.8f;", "count_success = 0; alpha = 0.2f; dt = dt * 0.5f;",
     "fire_v_shaped"),
    ("fire-nmin", "if (count_success > 5) { dt = fminf(dt * 1.2f, 10.0f);",
     "if (count_success > 4) { dt = fminf(dt * 1.2f, 10.0f);", "fire_linear"),
    ("fire-dt-max", "dt = fminf(dt * 1.2f, 10.0f);", "dt = fminf(dt * 1.2f, 20.0f);", "fire_linear"),
    ("fire-reset-new-gradient", "s = 0; for (int k = 0; k < NG; ++k) s += grad[k] * grad[k];\n            inv_gn = 0.2f * (1.0f / sqrtf(s));\n            for (int k = 0; k < NG; ++k) vel[k] = -grad[k] * inv_gn;",
     "s = 0; for (int k = 0; k < NG; ++k) s += cgrad[k] * cgrad[k];\n            inv_gn = 0.2f * (1.0f / sqrtf(s));\n            for (int k = 0; k < NG; ++k) vel[k] = -cgrad[k] * inv_gn;",
     "fire_v_shaped"),
    ("gradient-desolv-distance-dropped", "K(-0.077160) * d * exp(", "K(-0.077160) * exp(", "float_gradient_pair"),
    ("gradient-elec-sign", "pg += -(R)ep->coeff_elec * q1 * q2 * (upper / lower);",
     "pg += (R)ep->coeff_elec * q1 * q2 * (upper / lower);", "float_gradient_pair"),
    ("gradient-vdw-power", "pg += (-(R)m * (R)L->vw_ac[idx]) / pow(sd, (R)(m + 1));",
     "pg += (-(R)m * (R)L->vw_ac[idx]) / pow(sd, (R)(m));", "float_gradient"),
    ("gradient-smoothing-side", "if (d <= opt - delta_distance) sd = d + delta_distance;",
     "if (d <= opt - delta_distance) sd = d - delta_distance;", "float_gradient"),
    ("gradient-translation-over-all-atoms",
     "for (int a = 0; a < L->true_ligand_atoms; ++a) { gg[0] += gin[a][0];",
     "for (int a = 0; a < N; ++a) { gg[0] += gin[a][0];", "float_gradient_pair"),
    ("gradient-torsion-lever-arm", "V4 r = v4((c[a].x - c[at1].x) * spacing, (c[a].y - c[at1].y) * spacing, (c[a].z - c[at1].z) * spacing, 0);",
     "V4 r = v4((c[a].x - c[at2].x) * spacing, (c[a].y - c[at2].y) * spacing, (c[a].z - c[at1].z) * spacing, 0);",
     "float_gradient_matches"),
]


def _build(name, old, new, tmp):
    src = open(os.path.join(ORACLE, "adock_oracle.c")).read()
    assert src.count(old) >= 1, "mutation text not found: %s" % name
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    cfile = os.path.join(d, "adock_oracle.c")
    with open(cfile, "w") as f:
        f.write(src.replace(old, new))
    flags = ["-O2", "-fPIC", "-fopenmp", "-std=c11", "-w", "-I", ORACLE]
    objs = []
    for tag, extra in (("f32", []), ("f64", ["-DORC_DOUBLE"])):
        o = os.path.join(d, "orc_%s.o" % tag)
        subprocess.run(["gcc"] + flags + extra + ["-c", cfile, "-o", o], check=True)
        objs.append(o)
    o = os.path.join(d, "orc_autostop.o")
    subprocess.run(["gcc"] + flags + ["-c", os.path.join(ORACLE, "autostop_oracle.c"), "-o", o], check=True)
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["gcc", "-shared", "-fopenmp"] + objs + [o, "-lm", "-o", lib], check=True)
    return lib


def _pins_fail(lib, select):
    env = dict(os.environ, ADGPU_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(HERE, "test_oracle_pins.py"), "-q", "-x",
                        "-p", "no:cacheprovider", "-k", select], env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-400:]


def test_unmutated_copy_passes(tmp_path):
    """The harness itself: an unchanged copy passes every pin."""
    lib = _build("identity", "orc_fast_acos", "orc_fast_acos", str(tmp_path))
    rc, out = _pins_fail(lib, "orientation or adam or fire or float_gradient or fast_acos")
    assert rc == 0, out


def test_every_mutant_fails_a_pin(tmp_path):
    def one(m):
        name, old, new, select = m
        lib = _build(name, old, new, str(tmp_path))
        rc, out = _pins_fail(lib, select)
        return name, rc, out
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(one, MUTANTS))
    survived = [(n, out) for n, rc, out in res if rc != 1]
    assert not survived, survived
