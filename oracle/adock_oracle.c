/*
 * adock_oracle.c -- CPU restatement of AutoDock-GPU's docking hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see adock_oracle.h).  Compiled twice:
 * default = FP32 (*_f32), -DORC_DOUBLE = FP64 (*_f64).
 *
 * Every function follows the reference file:line given in its comment.
 * Reference = ccsb-scripps/AutoDock-GPU v1.5, OpenCL device code in device/
 * (the build the CPU baseline names), ADAM from cuda/kernel_adam.cu.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   coords      -- identity at zero genes, exact translation, 180 deg torsion
 *                  reflection, rigid-body invariance of bond lengths, and the
 *                  reference's own FP64 change_conform (oracle/_ref build);
 *   inter       -- exact on linear map fields, node values at integer
 *                  coordinates, +2^24 out of grid, and the reference's
 *                  calc_interE_f on the same maps (oracle/_ref);
 *   intra       -- two-atom closed forms of each term, the 8 / 20.48 A cutoffs
 *                  and smoothing plateau, invariance under rigid motion, and
 *                  the reference's table-based calc_intraE_f (oracle/_ref);
 *   energrad    -- central finite differences of the FP64 energy for the
 *                  translation and torsion genes; orientation genes (the
 *                  torque mapping of calcMergedEneGra.cl:712-861, both
 *                  branches) against their definition -- the genotype-angle
 *                  change of a turn about the torque axis, computed with
 *                  rotation matrices -- and fast_acos against acos;
 *   gradient    -- (FIRE's float gradient) central differences of the
 *                  analytic AD4 pair energies (Mehler-Solmajer, Gaussian
 *                  desolvation, smoothed vdW, CG-G0) and of the FP64 energy
 *                  where both use one function; orientation genes as above;
 *   GA / LS     -- injected LCG streams with hand-derived expectations
 *                  (selection, crossover points, mutation bounds, rho schedule,
 *                  ADADELTA first step); ADAM and FIRE by closed-form
 *                  trajectories on linear, V-shaped and kinked fields.
 * tests/test_oracle_mutants.py plants a mistake in each of these functions
 * (a copy of this file) and checks that a pin fails.
 */
#include "adock_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(ORC_DOUBLE)
typedef double R;
#define SFX(n) n##_f64
#else
typedef float R;
#define SFX(n) n##_f32
#endif

/* Constants exactly as the reference spells them (common/calcenergy_basic.h,
 * common/defines.h); float literals converted to R so the FP64 twin differs
 * from the FP32 build only in arithmetic precision. */
#define K(x) ((R)(x##f))
#define DEG_TO_RAD        K(0.0174533)
#define PI_F              ((R)3.14159265358979323846f)   /* (float)M_PI */
#define PI_TIMES_2        ((R)2.0f * PI_F)
#define ATYPE_CG_IDX      24
#define ATYPE_G0_IDX      25
#define G_BOND            K(50.0)
#define TERMSCALE         K(1024.0)
#define MAXTERM           K(8192.0)
#define ONEOVERTERMSCALE  ((R)(1.0f / 1024.0f))
#define INF_RAD           K(1E-3)
#define INV_INF_RAD       ((R)(1.0f / 1E-3f))
#define HALF_INF_RAD      ((R)(0.5f * 1E-3f))

/* ------------------------------------------------------------------------- */
/* Small vector helpers (calcenergy.cl:105-129)                               */
/* ------------------------------------------------------------------------- */
typedef struct { R x, y, z, w; } V4;

static inline V4 v4(R x, R y, R z, R w) { V4 r = {x, y, z, w}; return r; }
static inline V4 vsub(V4 a, V4 b) { return v4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline V4 vadd(V4 a, V4 b) { return v4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline V4 vscale(V4 a, R s) { return v4(a.x * s, a.y * s, a.z * s, a.w * s); }
/* OpenCL cross() of float4: w = 0 */
static inline V4 vcross(V4 a, V4 b) {
    return v4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, (R)0);
}

/* quaternion_rotate, calcenergy.cl:119-127 */
static inline V4 qrotate(V4 v, V4 rot) {
    V4 z = vscale(vcross(rot, v), (R)2.0f);
    return vadd(vadd(v, vscale(z, rot.w)), vcross(rot, z));
}

/* quaternion_multiply, calcenergy.cl:110-117 */
static inline V4 qmul(V4 a, V4 b) {
    return v4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
              a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}

/* fmod_pi2, calcenergy.cl:86-91 */
static inline R fmod_pi2(R x) {
    return x - (R)((int)(((R)1.0f / PI_TIMES_2) * x)) * PI_TIMES_2;
}

/* fast_acos, calcenergy.cl:93-108 */
static inline R fast_acos(R cosine) {
    const R a = K(9.78056e-05), b = K(-0.00104588), c = K(0.00418716), d = K(-0.00314347);
    const R e = K(2.74084), f = K(0.370388);
    const R o = -(a + b + c + d);
    R x = fabs(cosine);
    R x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    R ac = (((o * x4 + a) * x3 + b) * x2 + c) * x + d +
           e * sqrt((R)2.0f - sqrt((R)2.0f + (R)2.0f * x)) - f * sqrt((R)2.0f - (R)2.0f * x);
    return copysign(ac, cosine) + (cosine < (R)0.0f ? PI_F : (R)0.0f);
}

/* fast_acos on its own, so tests can bound it against acos. */
R SFX(orc_fast_acos)(R cosine) { return fast_acos(cosine); }

/* float2int_round, calcMergedEneGra.cl:30-35, applied to the clamped term
 * fmin(MAXTERM, fmax(-MAXTERM, x)) as every call site does.  Note the
 * reference's sign convention: the half is subtracted toward zero before the
 * truncating conversion. */
static inline int32_t quant(R x) {
    R v = fmin(MAXTERM, fmax(-MAXTERM, x));
    int addsub = (int)((v < (R)0.0f) - (v > (R)0.0f));
    return (int32_t)(v + (R)0.5f * (R)addsub);
}

/* ------------------------------------------------------------------------- */
/* Pose -> coordinates (calcenergy.cl:185-275; rotation order from           */
/* gen_rotlist, host/src/calcenergy.cpp:387-488: per atom, the rotatable     */
/* bonds moving it in ascending index, then the general rotation).           */
/* ------------------------------------------------------------------------- */
typedef struct {
    V4 genrot_q, genrot_mv;
    R phi, theta, rotangle, sin_theta, sin_half_rotangle, is_theta_gt_pi;
} Pose;

static void pose_coords(const adgpu_ligand* L, const float* geno, V4* coords, Pose* P) {
    P->phi      = (R)geno[3] * DEG_TO_RAD;
    P->theta    = (R)geno[4] * DEG_TO_RAD;
    P->rotangle = (R)geno[5] * DEG_TO_RAD;
    P->genrot_q = v4(0, 0, 0, 1);
    P->genrot_mv = v4(0, 0, 0, 0);
    P->sin_theta = 0; P->sin_half_rotangle = 0; P->is_theta_gt_pi = 1;
    if (L->true_ligand_atoms) {
        P->genrot_mv = v4((R)geno[0], (R)geno[1], (R)geno[2], (R)0.0f);
        R sin_theta = sin(P->theta);
        R cos_theta = cos(P->theta);
        R s2 = sin(P->rotangle * (R)0.5f);
        P->genrot_q = v4(s2 * sin_theta * cos(P->phi), s2 * sin_theta * sin(P->phi),
                         s2 * cos_theta, cos(P->rotangle * (R)0.5f));
        P->sin_theta = sin_theta;
        P->sin_half_rotangle = s2;
        P->is_theta_gt_pi = (R)1.0f - (R)2.0f * (R)(sin_theta < (R)0.0f);
    }
    const int T = L->num_rotbonds;
    for (int a = 0; a < L->num_atoms; ++a) {
        V4 v = v4((R)L->ref_coords[3 * a], (R)L->ref_coords[3 * a + 1], (R)L->ref_coords[3 * a + 2], 0);
        for (int t = 0; t < T; ++t) {
            if (!L->atom_rotbonds[a * T + t]) continue;
            R ang = (R)geno[6 + t] * DEG_TO_RAD * (R)0.5f;
            R s = sin(ang);
            V4 q = v4(s * (R)L->rotbond_unit[3 * t], s * (R)L->rotbond_unit[3 * t + 1],
                      s * (R)L->rotbond_unit[3 * t + 2], cos(ang));
            V4 mv = v4((R)L->rotbond_moving[3 * t], (R)L->rotbond_moving[3 * t + 1],
                       (R)L->rotbond_moving[3 * t + 2], 0);
            v = vadd(qrotate(vsub(v, mv), q), mv);
        }
        if (a < L->true_ligand_atoms)
            v = vadd(qrotate(v, P->genrot_q), P->genrot_mv);
        /* flexible-residue atoms: identity rotation, zero translation */
        coords[a] = v;
    }
}

void SFX(orc_coords)(const adgpu_ligand* L, const float* genotype, R* xyz) {
    V4 c[ADGPU_MAX_ATOMS];
    Pose P;
    pose_coords(L, genotype, c, &P);
    for (int a = 0; a < L->num_atoms; ++a) {
        xyz[3 * a] = c[a].x; xyz[3 * a + 1] = c[a].y; xyz[3 * a + 2] = c[a].z;
    }
}

/* ------------------------------------------------------------------------- */
/* Trilinear map gathers (calcenergy.cl:280-379; layout processgrid.cpp:     */
/* 234-249).                                                                 */
/* ------------------------------------------------------------------------- */
typedef struct { R c[8]; } Cube;

static inline Cube load_cube(const adgpu_grid* g, int map, long base) {
    const long g3 = (long)g->size_x * g->size_y * g->size_z;
    const float* p = g->packed + ((long)map * g3 << 2) + base;
    Cube cb;
    for (int i = 0; i < 8; ++i) cb.c[i] = (R)p[i];
    return cb;
}

/* weights idx_000..idx_111 (common/defines.h indices enum) */
static inline R cube_interp(const Cube* cb, const R* w) {
    return cb->c[0] * w[0] + cb->c[1] * w[1] + cb->c[2] * w[2] + cb->c[3] * w[3] +
           cb->c[4] * w[4] + cb->c[5] * w[5] + cb->c[6] * w[6] + cb->c[7] * w[7];
}

/* calcenergy.cl:293-301.  Written as "not inside" so that NaN coordinates
 * (a FIRE step from a zero gradient divides by its norm) count as outside,
 * as the device kernels do, instead of forming a map index from NaN. */
static inline int out_of_grid(const adgpu_grid* g, R x, R y, R z) {
    return !((x >= 0) && (y >= 0) && (z >= 0) && (x < (R)(g->size_x - 1)) &&
             (y < (R)(g->size_y - 1)) && (z < (R)(g->size_z - 1)));
}

/* ------------------------------------------------------------------------- */
/* Intramolecular pair energy (calcenergy.cl:397-536)                         */
/* ------------------------------------------------------------------------- */
typedef struct { R e; R dedr; } PairE;

static inline PairE pair_terms(const adgpu_ligand* L, const adgpu_energy_params* ep, int a1, int a2,
                               R atomic_distance, int want_grad, R* absacc) {
    PairE out = {0, 0};
    const int NT = L->num_atypes;
    int t1 = L->atom_types[a1], t2 = L->atom_types[a2];
    unsigned r1 = L->atype_reqm[t1], r2 = L->atype_reqm[t2];
    R vbond = G_BOND * (R)(((r1 == ATYPE_CG_IDX) && (r2 == ATYPE_G0_IDX)) ||
                           ((r1 == ATYPE_G0_IDX) && (r2 == ATYPE_CG_IDX)));
    out.e += vbond * atomic_distance;
    out.dedr += vbond;
    *absacc += fabs(vbond * atomic_distance);
    const R delta_distance = (R)0.5f * (R)ep->smooth;
    if (atomic_distance < (R)8.0f) {
        int idx = t1 * NT + t2;
        unsigned short exps = L->vw_exp[idx];
        signed char m = (signed char)((exps & 0xFF00) >> 8);
        signed char n = (signed char)(exps & 0xFF);
        R opt_distance = (R)L->reqm_ab[idx];
        R opt_dist_delta = opt_distance - atomic_distance;
        R sd;
        if (fabs(opt_dist_delta) >= delta_distance)
            sd = atomic_distance + copysign(delta_distance, opt_dist_delta);
        else
            sd = opt_distance;
        R rmn = pow(sd, (R)(m - n));
        R rm  = pow(sd, (R)(-m));
        R AC = (R)L->vw_ac[idx], BD = (R)L->vw_bd[idx];
        R ev = (AC - rmn * BD) * rm;
        out.e += ev;
        *absacc += fabs(AC * rm) + fabs(rmn * BD * rm);
        if (want_grad)   /* calcMergedEneGra.cl:515 */
            out.dedr += ((R)n * BD * rmn - (R)m * AC) * rm / sd;
    }
    if (atomic_distance < (R)20.48f) {
        R d = atomic_distance;
        if (d < (R)ep->elec_min_distance) d = (R)ep->elec_min_distance;
        R q1 = (R)L->charges[a1], q2 = (R)L->charges[a2];
        R dist2 = d * d;
        R desolv = (((R)L->ds_s[t1] + (R)ep->qasp * fabs(q1)) * (R)L->ds_v[t2] +
                    ((R)L->ds_s[t2] + (R)ep->qasp * fabs(q2)) * (R)L->ds_v[t1]) *
                   (((R)ep->coeff_desolv * (K(12.96) - K(0.1063) * dist2 * ((R)1.0f - K(0.001947) * dist2))) /
                    (K(12.96) + dist2 * (K(0.4137) + dist2 * (K(0.00357) + K(0.000112) * dist2))));
        R dist_shift = d + K(1.26366);
        R ds2 = dist_shift * dist_shift;
        R diel = K(1.10859) / ds2 + K(0.010358);
        R es = (R)ep->coeff_elec * q1 * q2 / d;
        out.e += diel * es + desolv;
        *absacc += fabs(diel * es) + fabs(desolv);
        if (want_grad)   /* calcMergedEneGra.cl:575-582 */
            out.dedr += -es / d * diel - es * K(2.21718) / (ds2 * dist_shift) - K(0.0771605) * d * desolv;
    }
    return out;
}

/* ------------------------------------------------------------------------- */
/* gpu_calc_energy (calcenergy.cl:132-554)                                    */
/* ------------------------------------------------------------------------- */
R SFX(orc_energy)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                  const float* genotype, R* inter_out, R* intra_out, R* abs_out) {
    V4 c[ADGPU_MAX_ATOMS];
    Pose P;
    pose_coords(L, genotype, c, &P);
    const long g1 = g->size_x, g2 = (long)g->size_x * g->size_y;
    R inter = 0, intra = 0, absacc = 0;
    for (int a = 0; a < L->num_atoms; ++a) {
        if (L->ignore_inter[a]) continue;
        R x = c[a].x, y = c[a].y, z = c[a].z;
        R q = (R)L->charges[a];
        if (out_of_grid(g, x, y, z)) { inter += (R)16777216.0f; absacc += (R)16777216.0f; continue; }
        R xl = floor(x), yl = floor(y), zl = floor(z);
        long base = ((long)(xl + yl * (R)g1 + zl * (R)g2)) << 2;
        R dx = x - xl, omdx = (R)1.0f - dx, dy = y - yl, omdy = (R)1.0f - dy, dz = z - zl, omdz = (R)1.0f - dz;
        R w[8] = {omdx * omdy * omdz, omdx * dy * omdz, omdx * omdy * dz, omdx * dy * dz,
                  dx * omdy * omdz,   dx * dy * omdz,   dx * omdy * dz,   dx * dy * dz};
        Cube ct = load_cube(g, L->atom_map[a], base);
        Cube ce = load_cube(g, g->num_map_atypes, base);
        Cube cd = load_cube(g, g->num_map_atypes + 1, base);
        R et = cube_interp(&ct, w), ee = q * cube_interp(&ce, w), ed = fabs(q) * cube_interp(&cd, w);
        inter += et; inter += ee; inter += ed;
        absacc += fabs(et) + fabs(ee) + fabs(ed);
    }
    for (int p = 0; p < L->num_pairs; ++p) {
        int a1 = L->pairs[2 * p], a2 = L->pairs[2 * p + 1];
        R sx = c[a1].x - c[a2].x, sy = c[a1].y - c[a2].y, sz = c[a1].z - c[a2].z;
        R dist = sqrt(sx * sx + sy * sy + sz * sz) * (R)g->spacing;
        PairE pe = pair_terms(L, ep, a1, a2, dist, 0, &absacc);
        intra += pe.e;
    }
    if (inter_out) *inter_out = inter;
    if (intra_out) *intra_out = intra;
    if (abs_out) *abs_out = absacc;
    return inter + intra;
}

/* ------------------------------------------------------------------------- */
/* Rotation-gene gradient from the total torque (calcMergedEneGra.cl:712-861;*/
/* the float gradient of calcgradient.cl:795-873 differs only in using fmod). */
/* ------------------------------------------------------------------------- */
static void torque_to_rotgenes(V4 torque, const Pose* P, int use_fmod_pi2, R out[3]) {
    R tl = sqrt(torque.x * torque.x + torque.y * torque.y + torque.z * torque.z);
    R orientation_scaling = (tl < INF_RAD) ? (R)1.0f : tl * INV_INF_RAD;
    R sin_half = (R)sinf(0.5f * 1E-3f), cos_half = (R)cosf(0.5f * 1E-3f);
#if defined(ORC_DOUBLE)
    sin_half = sin(HALF_INF_RAD); cos_half = cos(HALF_INF_RAD);
#endif
    R sc = (tl < INF_RAD) ? (R)0.5f - tl * tl / (R)48.0f : sin_half / tl;
    V4 qt = v4(torque.x * sc, torque.y * sc, torque.z * sc,
               (tl < INF_RAD) ? (R)1.0f - tl * tl * (R)0.125f : cos_half);
    V4 target_q = qmul(qt, P->genrot_q);
    R current_phi      = fmod_pi2(PI_TIMES_2 + P->phi);
    R current_theta    = fmod_pi2(PI_TIMES_2 + P->theta);
    R current_rotangle = fmod_pi2(PI_TIMES_2 + P->rotangle);
    R target_rotangle = (R)2.0f * fast_acos(target_q.w);
    R inv_sin_ang = (R)1.0f / sqrt((R)1.0f - target_q.w * target_q.w);
    R target_theta = PI_TIMES_2 + P->is_theta_gt_pi * fast_acos(target_q.z * inv_sin_ang);
    R target_phi = fmod_pi2(atan2(P->is_theta_gt_pi * target_q.y, P->is_theta_gt_pi * target_q.x) + PI_TIMES_2);
    R gphi, gtheta, grot;
    if (use_fmod_pi2) {
        gphi   = orientation_scaling * (fmod_pi2(target_phi - current_phi + PI_F) - PI_F);
        gtheta = orientation_scaling * (fmod_pi2(target_theta - current_theta + PI_F) - PI_F);
        grot   = orientation_scaling * (fmod_pi2(target_rotangle - current_rotangle + PI_F) - PI_F);
    } else {
        gphi   = orientation_scaling * (fmod(target_phi - current_phi + PI_F, PI_TIMES_2) - PI_F);
        gtheta = orientation_scaling * (fmod(target_theta - current_theta + PI_F, PI_TIMES_2) - PI_F);
        grot   = orientation_scaling * (fmod(target_rotangle - current_rotangle + PI_F, PI_TIMES_2) - PI_F);
    }
    R rot_angle_corr = (R)4.0f * P->sin_half_rotangle * P->sin_half_rotangle;
    out[0] = gphi * P->sin_theta * P->sin_theta * rot_angle_corr * DEG_TO_RAD;
    out[1] = gtheta * rot_angle_corr * DEG_TO_RAD;
    out[2] = grot * DEG_TO_RAD;
}

/* ------------------------------------------------------------------------- */
/* gpu_calc_energrad (calcMergedEneGra.cl:46-936)                             */
/* ------------------------------------------------------------------------- */
static R energrad_core(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                       const float* genotype, int quantize, R* grad, R* atom_grad);

R SFX(orc_energrad)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                    const float* genotype, int quantize, R* grad) {
    return energrad_core(g, L, ep, genotype, quantize, grad, NULL);
}

/* Per-atom gradient dE/dx (kcal/mol/A) and energy-sensitivity scale
 * sum_a |dE/dx_a| * |x_a| (x in A from the grid origin): the first-order change
 * of E under a relative coordinate perturbation. */
R SFX(orc_atom_grad)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                     const float* genotype, R* atom_grad, R* sensitivity) {
    R grad[ADGPU_GENOTYPE_STRIDE];
    R ag[3 * ADGPU_MAX_ATOMS];
    R e = energrad_core(g, L, ep, genotype, 0, grad, ag);
    if (atom_grad) memcpy(atom_grad, ag, sizeof(R) * 3 * L->num_atoms);
    if (sensitivity) {
        V4 c[ADGPU_MAX_ATOMS];
        Pose P;
        pose_coords(L, genotype, c, &P);
        R s = 0;
        for (int a = 0; a < L->num_atoms; ++a) {
            /* an atom outside the maps carries the constant 2^24 penalty in
             * its gradient entries; only its pair forces measure sensitivity */
            R off = (!L->ignore_inter[a] && out_of_grid(g, c[a].x, c[a].y, c[a].z)) ? (R)16777216.0f : (R)0;
            R fx = ag[3 * a] - off, fy = ag[3 * a + 1] - off, fz = ag[3 * a + 2] - off;
            R fn = sqrt(fx * fx + fy * fy + fz * fz);
            R xn = sqrt(c[a].x * c[a].x + c[a].y * c[a].y + c[a].z * c[a].z) * (R)g->spacing;
            s += fn * xn;
        }
        *sensitivity = s;
    }
    return e;
}

static R energrad_core(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                       const float* genotype, int quantize, R* grad, R* atom_grad) {
    V4 c[ADGPU_MAX_ATOMS];
    Pose P;
    pose_coords(L, genotype, c, &P);
    const int N = L->num_atoms, T = L->num_rotbonds, NG = T + 6;
    const long g1 = g->size_x, g2 = (long)g->size_x * g->size_y;
    const R spacing = (R)g->spacing;
    const R inv_gs = (R)1.0f / spacing;
    /* per-atom forces: int fixed point (quantize) or R */
    int32_t gi[ADGPU_MAX_ATOMS][3];
    R gf[ADGPU_MAX_ATOMS][3];
    memset(gi, 0, sizeof(gi));
    memset(gf, 0, sizeof(gf));
    R energy = 0, absacc = 0;
    for (int a = 0; a < N; ++a) {
        if (L->ignore_inter[a]) continue;
        R x = c[a].x, y = c[a].y, z = c[a].z;
        R q = (R)L->charges[a];
        if (out_of_grid(g, x, y, z)) {
            energy += (R)16777216.0f;
            if (quantize) { gi[a][0] += 16777216; gi[a][1] += 16777216; gi[a][2] += 16777216; }
            else { gf[a][0] += (R)16777216.0f; gf[a][1] += (R)16777216.0f; gf[a][2] += (R)16777216.0f; }
            continue;
        }
        int xl = (int)floor(x), yl = (int)floor(y), zl = (int)floor(z);
        long base = ((long)(xl + yl * g1 + zl * g2)) << 2;
        R dx = x - (R)xl, omdx = (R)1.0f - dx, dy = y - (R)yl, omdy = (R)1.0f - dy, dz = z - (R)zl, omdz = (R)1.0f - dz;
        R w[8] = {omdx * omdy * omdz, omdx * dy * omdz, omdx * omdy * dz, omdx * dy * dz,
                  dx * omdy * omdz,   dx * dy * omdz,   dx * omdy * dz,   dx * dy * dz};
        Cube cb[3];
        cb[0] = load_cube(g, L->atom_map[a], base);
        cb[1] = load_cube(g, g->num_map_atypes, base);
        cb[2] = load_cube(g, g->num_map_atypes + 1, base);
        R scale[3] = {(R)1.0f, q, fabs(q)};
        R gx = 0, gy = 0, gz = 0;
        for (int k = 0; k < 3; ++k) {
            const R* cc = cb[k].c;
            R ek = scale[k] * cube_interp(&cb[k], w);
            energy += ek;
            R s = (k == 0) ? inv_gs : (k == 1 ? q * inv_gs : fabs(q * inv_gs));
            R ax = omdz * (omdy * (cc[4] - cc[0]) + dy * (cc[5] - cc[1])) + dz * (omdy * (cc[6] - cc[2]) + dy * (cc[7] - cc[3]));
            R ay = omdz * (omdx * (cc[1] - cc[0]) + dx * (cc[5] - cc[4])) + dz * (omdx * (cc[3] - cc[2]) + dx * (cc[7] - cc[6]));
            R az = omdy * (omdx * (cc[2] - cc[0]) + dx * (cc[6] - cc[4])) + dy * (omdx * (cc[3] - cc[1]) + dx * (cc[7] - cc[5]));
            if (k == 0) { gx = ax * s; gy = ay * s; gz = az * s; }
            else { gx += s * ax; gy += s * ay; gz += s * az; }
        }
        if (quantize) {
            gi[a][0] += quant(TERMSCALE * gx); gi[a][1] += quant(TERMSCALE * gy); gi[a][2] += quant(TERMSCALE * gz);
        } else {
            gf[a][0] += gx; gf[a][1] += gy; gf[a][2] += gz;
        }
    }
    for (int p = 0; p < L->num_pairs; ++p) {
        int a1 = L->pairs[2 * p], a2 = L->pairs[2 * p + 1];
        R sx = c[a1].x - c[a2].x, sy = c[a1].y - c[a2].y, sz = c[a1].z - c[a2].z;
        R dist = sqrt(sx * sx + sy * sy + sz * sz);
        PairE pe = pair_terms(L, ep, a1, a2, dist * spacing, 1, &absacc);
        energy += pe.e;
        R gdd = -pe.dedr / dist;
        if (quantize) {
            int32_t ix = quant(TERMSCALE * sx * gdd), iy = quant(TERMSCALE * sy * gdd), iz = quant(TERMSCALE * sz * gdd);
            gi[a1][0] -= ix; gi[a1][1] -= iy; gi[a1][2] -= iz;
            gi[a2][0] += ix; gi[a2][1] += iy; gi[a2][2] += iz;
        } else {
            R fx = sx * gdd, fy = sy * gdd, fz = sz * gdd;
            gf[a1][0] -= fx; gf[a1][1] -= fy; gf[a1][2] -= fz;
            gf[a2][0] += fx; gf[a2][1] += fy; gf[a2][2] += fz;
        }
    }
    /* total per-atom force in kcal/mol/A */
    V4 F[ADGPU_MAX_ATOMS];
    for (int a = 0; a < N; ++a) {
        if (quantize) F[a] = v4(ONEOVERTERMSCALE * (R)gi[a][0], ONEOVERTERMSCALE * (R)gi[a][1], ONEOVERTERMSCALE * (R)gi[a][2], 0);
        else F[a] = v4(gf[a][0], gf[a][1], gf[a][2], 0);
        if (atom_grad) { atom_grad[3 * a] = F[a].x; atom_grad[3 * a + 1] = F[a].y; atom_grad[3 * a + 2] = F[a].z; }
    }
    R gg[ADGPU_GENOTYPE_STRIDE];
    int32_t ig[ADGPU_GENOTYPE_STRIDE];
    memset(gg, 0, sizeof(gg));
    memset(ig, 0, sizeof(ig));
    if (L->true_ligand_atoms) {
        /* translation: sum over all atoms (calcMergedEneGra.cl:633-697) */
        R ax = 0, ay = 0, az = 0;
        for (int a = 0; a < N; ++a) { ax += F[a].x; ay += F[a].y; az += F[a].z; }
        if (quantize) {
            ig[0] = quant(TERMSCALE * ax * spacing); ig[1] = quant(TERMSCALE * ay * spacing); ig[2] = quant(TERMSCALE * az * spacing);
        } else { gg[0] = ax * spacing; gg[1] = ay * spacing; gg[2] = az * spacing; }
        /* rotation: torque over ligand atoms (calcMergedEneGra.cl:712-861) */
        V4 torque = v4(0, 0, 0, 0);
        for (int a = 0; a < L->true_ligand_atoms; ++a) {
            V4 r = vscale(vsub(c[a], P.genrot_mv), spacing);
            V4 tr = vcross(r, F[a]);
            torque.x += tr.x; torque.y += tr.y; torque.z += tr.z;
        }
        R rg[3];
        torque_to_rotgenes(torque, &P, 1, rg);
        for (int k = 0; k < 3; ++k) {
            if (quantize) ig[3 + k] = quant(TERMSCALE * rg[k]);
            else gg[3 + k] = rg[k];
        }
    }
    /* torsions (calcMergedEneGra.cl:866-915) */
    for (int t = 0; t < T; ++t) {
        int at1 = L->rotbond_atoms[2 * t], at2 = L->rotbond_atoms[2 * t + 1];
        V4 ref = c[at1];
        V4 ax = vsub(c[at2], ref);
        R len = sqrt(ax.x * ax.x + ax.y * ax.y + ax.z * ax.z);
        ax = v4(ax.x / len, ax.y / len, ax.z / len, 0);
        for (int a = 0; a < N; ++a) {
            if (!L->atom_rotbonds[a * T + t]) continue;
            V4 r = vsub(c[a], ref);
            V4 tq = vcross(r, F[a]);
            R toa = (ax.x * tq.x + ax.y * tq.y + ax.z * tq.z) * spacing;
            if (quantize) ig[6 + t] += quant(TERMSCALE * toa * DEG_TO_RAD);
            else gg[6 + t] += toa * DEG_TO_RAD;
        }
    }
    const R conv = spacing * spacing * (R)1.0f / (DEG_TO_RAD * DEG_TO_RAD);
    for (int k = 0; k < NG; ++k) {
        R v = quantize ? ONEOVERTERMSCALE * (R)ig[k] : gg[k];
        if (k >= 3) v *= conv;   /* CONVERT_INTO_ANGSTROM_RADIAN, calcMergedEneGra.cl:927-935 */
        grad[k] = v;
    }
    return energy;
}

/* ------------------------------------------------------------------------- */
/* gpu_calc_gradient (calcgradient.cl:109-1003): float gradient with exact    */
/* Mehler-Solmajer / exponential desolvation derivatives; FIRE's first step.  */
/* ------------------------------------------------------------------------- */
void SFX(orc_gradient)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                       const float* genotype, R* grad) {
    V4 c[ADGPU_MAX_ATOMS];
    Pose P;
    pose_coords(L, genotype, c, &P);
    const int N = L->num_atoms, T = L->num_rotbonds, NG = T + 6, NT = L->num_atypes;
    const long g1 = g->size_x, g2 = (long)g->size_x * g->size_y;
    const long g3 = g2 * g->size_z;
    const R spacing = (R)g->spacing;
    R gin[ADGPU_MAX_ATOMS][3], gintra[ADGPU_MAX_ATOMS][3];
    memset(gin, 0, sizeof(gin));
    memset(gintra, 0, sizeof(gintra));
    for (int a = 0; a < N; ++a) {
        if (L->ignore_inter[a]) continue;
        R x = c[a].x, y = c[a].y, z = c[a].z;
        R q = (R)L->charges[a];
        if (out_of_grid(g, x, y, z)) {
            gin[a][0] += (R)16777216.0f; gin[a][1] += (R)16777216.0f; gin[a][2] += (R)16777216.0f;
            continue;
        }
        unsigned xl = (unsigned)floor(x), yl = (unsigned)floor(y), zl = (unsigned)floor(z);
        unsigned xh = (unsigned)ceil(x), yh = (unsigned)ceil(y), zh = (unsigned)ceil(z);
        R dx = x - (R)xl, dy = y - (R)yl, dz = z - (R)zl;
        long off[8];   /* idx_000 idx_010 idx_001 idx_011 idx_100 idx_110 idx_101 idx_111 */
        off[0] = (long)(xl + yl * g1 + zl * g2) << 2;
        off[1] = (long)(xl + yh * g1 + zl * g2) << 2;
        off[2] = (long)(xl + yl * g1 + zh * g2) << 2;
        off[3] = (long)(xl + yh * g1 + zh * g2) << 2;
        off[4] = (long)(xh + yl * g1 + zl * g2) << 2;
        off[5] = (long)(xh + yh * g1 + zl * g2) << 2;
        off[6] = (long)(xh + yl * g1 + zh * g2) << 2;
        off[7] = (long)(xh + yh * g1 + zh * g2) << 2;
        int maps[3] = {L->atom_map[a], g->num_map_atypes, g->num_map_atypes + 1};
        R sc[3] = {(R)1.0f, q, fabs(q)};
        for (int k = 0; k < 3; ++k) {
            R cb[8];
            for (int i = 0; i < 8; ++i) cb[i] = (R)g->packed[((long)maps[k] * g3 << 2) + off[i]];
            R x10 = cb[4] - cb[0], x52 = cb[5] - cb[1], x43 = cb[6] - cb[2], x76 = cb[7] - cb[3];
            R vx_z0 = ((R)1.0f - dy) * x10 + dy * x52, vx_z1 = ((R)1.0f - dy) * x43 + dy * x76;
            R y20 = cb[1] - cb[0], y51 = cb[5] - cb[4], y63 = cb[3] - cb[2], y74 = cb[7] - cb[6];
            R vy_z0 = ((R)1.0f - dx) * y20 + dx * y51, vy_z1 = ((R)1.0f - dx) * y63 + dx * y74;
            R z30 = cb[2] - cb[0], z41 = cb[6] - cb[4], z62 = cb[3] - cb[1], z75 = cb[7] - cb[5];
            R vz_y0 = ((R)1.0f - dx) * z30 + dx * z41, vz_y1 = ((R)1.0f - dx) * z62 + dx * z75;
            gin[a][0] += sc[k] * (((R)1.0f - dz) * vx_z0 + dz * vx_z1);
            gin[a][1] += sc[k] * (((R)1.0f - dz) * vy_z0 + dz * vy_z1);
            gin[a][2] += sc[k] * (((R)1.0f - dy) * vz_y0 + dy * vz_y1);
        }
    }
    const R delta_distance = (R)0.5f * (R)ep->smooth;
    const R DIEL_A = K(-8.5525), DIEL_B = K(78.4) - K(-8.5525), DIEL_K = K(7.7839);
    const R DIEL_B_TIMES_H = DIEL_B * K(0.003627);
    const R DIEL_B_TIMES_H_TIMES_K = DIEL_B_TIMES_H * DIEL_K;
    for (int p = 0; p < L->num_pairs; ++p) {
        int a1 = L->pairs[2 * p], a2 = L->pairs[2 * p + 1];
        R sx = c[a1].x - c[a2].x, sy = c[a1].y - c[a2].y, sz = c[a1].z - c[a2].z;
        R dist = sqrt(sx * sx + sy * sy + sz * sz);
        R d = dist * spacing;
        int t1 = L->atom_types[a1], t2 = L->atom_types[a2];
        unsigned r1 = L->atype_reqm[t1], r2 = L->atype_reqm[t2];
        int idx = t1 * NT + t2;
        unsigned short exps = L->vw_exp[idx];
        signed char m = (signed char)((exps & 0xFF00) >> 8), n = (signed char)(exps & 0xFF);
        R opt = (R)L->reqm_ab[idx];
        R sd;
        if (d <= opt - delta_distance) sd = d + delta_distance;
        else if (d < opt + delta_distance) sd = opt;
        else sd = d - delta_distance;
        R pg = 0;
        if (d < (R)8.0f) {
            pg += (-(R)m * (R)L->vw_ac[idx]) / pow(sd, (R)(m + 1));
            pg += ((R)n * (R)L->vw_bd[idx]) / pow(sd, (R)(n + 1));
        }
        if (d < (R)20.48f) {
            if (d < (R)ep->elec_min_distance) d = (R)ep->elec_min_distance;
            R ex = exp(DIEL_B_TIMES_H * d);
            R upper = DIEL_A * pow(ex + DIEL_K, (R)2) + DIEL_B * ex * (DIEL_B_TIMES_H_TIMES_K * d + ex + DIEL_K);
            R lower = pow(d, (R)2) * pow(DIEL_A * (ex + DIEL_K) + DIEL_B * ex, (R)2);
            R q1 = (R)L->charges[a1], q2 = (R)L->charges[a2];
            pg += -(R)ep->coeff_elec * q1 * q2 * (upper / lower);
            pg += (((R)L->ds_s[t1] + (R)ep->qasp * fabs(q1)) * (R)L->ds_v[t2] +
                   ((R)L->ds_s[t2] + (R)ep->qasp * fabs(q2)) * (R)L->ds_v[t1]) *
                  (R)ep->coeff_desolv * K(-0.077160) * d * exp(K(-0.038580) * pow(d, (R)2));
        }
        if (((r1 == ATYPE_CG_IDX) && (r2 == ATYPE_G0_IDX)) || ((r1 == ATYPE_G0_IDX) && (r2 == ATYPE_CG_IDX)))
            pg += G_BOND;
        R ux = -sx / dist, uy = -sy / dist, uz = -sz / dist;
        R fx = pg * ux, fy = pg * uy, fz = pg * uz;
        gintra[a1][0] -= fx; gintra[a1][1] -= fy; gintra[a1][2] -= fz;
        gintra[a2][0] += fx; gintra[a2][1] += fy; gintra[a2][2] += fz;
    }
    for (int a = 0; a < N; ++a)
        for (int k = 0; k < 3; ++k) gin[a][k] = gin[a][k] / spacing + gintra[a][k];
    R gg[ADGPU_GENOTYPE_STRIDE];
    memset(gg, 0, sizeof(gg));
    /* translation over ligand atoms only (calcgradient.cl:690-708) */
    for (int a = 0; a < L->true_ligand_atoms; ++a) { gg[0] += gin[a][0]; gg[1] += gin[a][1]; gg[2] += gin[a][2]; }
    gg[0] *= spacing; gg[1] *= spacing; gg[2] *= spacing;
    if (L->true_ligand_atoms) {   /* calcgradient.cl:714-873 (thread 1) */
        V4 torque = v4(0, 0, 0, 0);
        for (int a = 0; a < L->true_ligand_atoms; ++a) {
            V4 r = v4((c[a].x - (R)genotype[0]) * spacing, (c[a].y - (R)genotype[1]) * spacing,
                      (c[a].z - (R)genotype[2]) * spacing, 0);
            V4 tr = vcross(r, v4(gin[a][0], gin[a][1], gin[a][2], 0));
            torque.x += tr.x; torque.y += tr.y; torque.z += tr.z;
        }
        R rg[3];
        torque_to_rotgenes(torque, &P, 0, rg);
        gg[3] = rg[0]; gg[4] = rg[1]; gg[5] = rg[2];
    }
    for (int t = 0; t < T; ++t) {   /* calcgradient.cl:877-989 */
        int at1 = L->rotbond_atoms[2 * t], at2 = L->rotbond_atoms[2 * t + 1];
        V4 ax = vsub(c[at2], c[at1]);
        R len = sqrt(ax.x * ax.x + ax.y * ax.y + ax.z * ax.z);
        ax = v4(ax.x / len, ax.y / len, ax.z / len, 0);
        V4 tq = v4(0, 0, 0, 0);
        for (int a = 0; a < N; ++a) {
            if (!L->atom_rotbonds[a * T + t]) continue;
            V4 r = v4((c[a].x - c[at1].x) * spacing, (c[a].y - c[at1].y) * spacing, (c[a].z - c[at1].z) * spacing, 0);
            V4 tr = vcross(r, v4(gin[a][0], gin[a][1], gin[a][2], 0));
            tq.x += tr.x; tq.y += tr.y; tq.z += tr.z;
        }
        gg[6 + t] = (ax.x * tq.x + ax.y * tq.y + ax.z * tq.z) * DEG_TO_RAD;
    }
    const R conv = spacing * spacing * (R)1.0f / (DEG_TO_RAD * DEG_TO_RAD);
    for (int k = 0; k < NG; ++k) grad[k] = (k > 2) ? gg[k] * conv : gg[k];
}

/* ------------------------------------------------------------------------- */
/* Batches of genotypes (stride 64), split over nthreads OpenMP threads: the  */
/* per-pose functions above, one call per pose, for parity runs over 1e5      */
/* poses.  Optional outputs may be NULL.                                      */
/* ------------------------------------------------------------------------- */
void SFX(orc_energy_batch)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                           const float* genotypes, int64_t n, R* energies, R* inters, R* abs_sums,
                           int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        R a = 0, it = 0;
        R e = SFX(orc_energy)(g, L, ep, genotypes + i * ADGPU_GENOTYPE_STRIDE, &it, NULL, &a);
        energies[i] = e;
        if (inters) inters[i] = it;
        if (abs_sums) abs_sums[i] = a;
    }
}

void SFX(orc_energrad_batch)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                             const float* genotypes, int64_t n, int quantize, R* energies, R* grads,
                             int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        R gr[ADGPU_GENOTYPE_STRIDE];
        memset(gr, 0, sizeof(gr));
        R e = SFX(orc_energrad)(g, L, ep, genotypes + i * ADGPU_GENOTYPE_STRIDE, quantize, gr);
        if (energies) energies[i] = e;
        if (grads) memcpy(grads + i * ADGPU_GENOTYPE_STRIDE, gr, sizeof(gr));
    }
}

void SFX(orc_gradient_batch)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                             const float* genotypes, int64_t n, R* grads, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        R gr[ADGPU_GENOTYPE_STRIDE];
        memset(gr, 0, sizeof(gr));
        SFX(orc_gradient)(g, L, ep, genotypes + i * ADGPU_GENOTYPE_STRIDE, gr);
        memcpy(grads + i * ADGPU_GENOTYPE_STRIDE, gr, sizeof(gr));
    }
}

void SFX(orc_sensitivity_batch)(const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                                const float* genotypes, int64_t n, R* sensitivity, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i)
        SFX(orc_atom_grad)(g, L, ep, genotypes + i * ADGPU_GENOTYPE_STRIDE, NULL, &sensitivity[i]);
}

#if !defined(ORC_DOUBLE)
/* ========================================================================= */
/* FP32-only: GA, local searches, LGA driver                                  */
/* ========================================================================= */

/* gpu_rand / gpu_randf (auxiliary_genetic.cl:29-67) */
uint32_t orc_rand(uint32_t* s) { *s = 1103515245u * (*s) + 12345u; return *s; }
float orc_randf(uint32_t* s) { return ((float)orc_rand(s) / 4294967296.0f) * 0.999999f; }

/* map_angle (auxiliary_genetic.cl:72-85) */
static inline void map_angle(float* a) {
    /* values the loop cannot reduce (|a| >= 1e6, inf) go through fmodf first;
     * the reference would loop forever on them */
    if (!(fabsf(*a) < 1.0e6f)) *a = fmodf(*a, 360.0f);
    if (isnan(*a) || isinf(*a)) return;
    while (*a >= 360.0f) *a -= 360.0f;
    while (*a < 0.0f) *a += 360.0f;
}

int orc_job_init(orc_job* j, const adgpu_grid* g, const adgpu_ligand* L, const adgpu_energy_params* ep,
                 const adgpu_search_params* sp, const float* init_pop, const uint32_t* seeds) {
    memset(j, 0, sizeof(*j));
    j->grid = g; j->lig = L; j->ep = ep; j->sp = *sp;
    size_t ne = (size_t)sp->num_runs * sp->pop_size;
    j->pop_cur = (float*)calloc(ne * 64, sizeof(float));
    j->pop_next = (float*)calloc(ne * 64, sizeof(float));
    j->ene_cur = (float*)calloc(ne, sizeof(float));
    j->ene_next = (float*)calloc(ne, sizeof(float));
    j->evals_new = (int32_t*)calloc(ne, sizeof(int32_t));
    j->evals_runs = (int32_t*)calloc(sp->num_runs, sizeof(int32_t));
    j->prng = (uint32_t*)calloc(ne * 64, sizeof(uint32_t));
    if (!j->pop_cur || !j->pop_next || !j->ene_cur || !j->ene_next || !j->evals_new || !j->evals_runs || !j->prng)
        return 1;
    memcpy(j->pop_cur, init_pop, ne * 64 * sizeof(float));
    memcpy(j->prng, seeds, ne * 64 * sizeof(uint32_t));
    return 0;
}

void orc_job_free(orc_job* j) {
    free(j->pop_cur); free(j->pop_next); free(j->ene_cur); free(j->ene_next);
    free(j->evals_new); free(j->evals_runs); free(j->prng);
    memset(j, 0, sizeof(*j));
}

void orc_swap(orc_job* j) {
    float* t = j->pop_cur; j->pop_cur = j->pop_next; j->pop_next = t;
    t = j->ene_cur; j->ene_cur = j->ene_next; j->ene_next = t;
}

/* gpu_calc_initpop (kernel1.cl:28-145) + gpu_sum_evals */
void orc_initpop(orc_job* j, int r0, int r1) {
    const int pop = j->sp.pop_size;
    for (int r = r0; r < r1; ++r) {
        for (int e = 0; e < pop; ++e) {
            int b = r * pop + e;
            j->ene_cur[b] = orc_energy_f32(j->grid, j->lig, j->ep, j->pop_cur + (size_t)b * 64, NULL, NULL, NULL);
            j->evals_new[b] = 1;
        }
    }
    orc_sum_evals(j, r0, r1);
}

void orc_sum_evals(orc_job* j, int r0, int r1) {
    const int pop = j->sp.pop_size;
    for (int r = r0; r < r1; ++r) {
        int s = 0;
        for (int e = 0; e < pop; ++e) s += j->evals_new[r * pop + e];
        j->evals_runs[r] += s;
    }
}

/* gpu_gen_and_eval_newpops (kernel4.cl:32-336) with the elitist copy of
 * gpu_perform_elitist_selection (auxiliary_genetic.cl:89-162). */
static void gen_one(orc_job* j, int r, int e) {
    const int pop = j->sp.pop_size;
    const int NG = j->lig->num_rotbonds + 6;
    const float tr = j->sp.tournament_rate / 100.0f, cr = j->sp.crossover_rate / 100.0f;
    const float mr = j->sp.mutation_rate / 100.f;
    {
        {
            const int b = r * pop + e;
            uint32_t* st = j->prng + (size_t)b * ADGPU_NWI;   /* thread tidx uses st[tidx] */
            float* child = j->pop_next + (size_t)b * 64;
            if (e == 0) {
                /* elitism: the NWI-strided scan keeps, per work-item, the first
                 * strict minimum; work-items are then scanned in order. */
                float best[ADGPU_NWI] = {0}; int bid[ADGPU_NWI] = {0};
                int nl = pop < ADGPU_NWI ? pop : ADGPU_NWI;
                for (int t = 0; t < nl; ++t) { best[t] = j->ene_cur[r * pop + t]; bid[t] = t; }
                for (int t = 0; t < nl; ++t)
                    for (int k = ADGPU_NWI + t; k < pop; k += ADGPU_NWI)
                        if (j->ene_cur[r * pop + k] < best[t]) { best[t] = j->ene_cur[r * pop + k]; bid[t] = k; }
                float be = best[0]; int bi = bid[0];
                for (int t = 1; t < nl; ++t) if (best[t] < be) { be = best[t]; bi = bid[t]; }
                j->ene_next[b] = be;
                j->evals_new[b] = 0;
                for (int k = 0; k < NG; ++k) child[k] = j->pop_cur[(size_t)(r * pop + bi) * 64 + k];
                return;
            }
            float rn[10];
            for (int k = 0; k < 10; ++k) rn[k] = orc_randf(&st[k]);
            int cand[4]; float ce[4];
            for (int k = 0; k < 4; ++k) { cand[k] = (int)((float)pop * rn[k]); ce[k] = j->ene_cur[r * pop + cand[k]]; }
            int par[2];
            for (int k = 0; k < 2; ++k) {
                if (ce[2 * k] < ce[2 * k + 1]) par[k] = (rn[4 + k] < tr) ? cand[2 * k] : cand[2 * k + 1];
                else par[k] = (rn[4 + k] < tr) ? cand[2 * k + 1] : cand[2 * k];
            }
            const float* p0 = j->pop_cur + (size_t)(r * pop + par[0]) * 64;
            const float* p1 = j->pop_cur + (size_t)(r * pop + par[1]) * 64;
            if (rn[6] < cr) {
                int cp0 = (int)((float)(NG - 1) * rn[7]), cp1 = (int)((float)(NG - 1) * rn[8]);
                if (cp1 < cp0) { int t = cp1; cp1 = cp0; cp0 = t; }
                for (int k = 0; k < NG; ++k) {
                    if (cp0 != cp1) child[k] = ((k <= cp0) || (k > cp1)) ? p0[k] : p1[k];
                    else child[k] = (k <= cp0) ? p0[k] : p1[k];
                }
            } else {
                for (int k = 0; k < NG; ++k) child[k] = p0[k];
            }
            for (int k = 0; k < NG; ++k) {   /* gene k is handled by work-item k */
                if (orc_randf(&st[k]) < mr) {
                    if (k < 3) child[k] += j->sp.abs_max_dmov * (2 * orc_randf(&st[k]) - 1);
                    else { child[k] += j->sp.abs_max_dang * (2 * orc_randf(&st[k]) - 1); map_angle(&child[k]); }
                }
            }
            j->ene_next[b] = orc_energy_f32(j->grid, j->lig, j->ep, child, NULL, NULL, NULL);
            j->evals_new[b] = 1;
        }
    }
}

void orc_gen(orc_job* j, int r0, int r1) {
    for (int r = r0; r < r1; ++r)
        for (int e = 0; e < j->sp.pop_size; ++e) gen_one(j, r, e);
}

static int ls_entities(const orc_job* j) {
    return (int)(j->sp.lsearch_rate / 100.0 * j->sp.pop_size + 0.5);
}

/* Entity choice shared by all LS kernels (e.g. kernel3.cl:104-117). */
static int ls_pick(orc_job* j, int r, int le, uint32_t* st0) {
    int nls = ls_entities(j);
    int ent = le;
    if (ent == 0 && 100.0f * orc_randf(st0) > j->sp.lsearch_rate) ent = nls;
    (void)r;
    /* the reference may read entity pop_size (see the kernels); clamp */
    return ent < j->sp.pop_size ? ent : j->sp.pop_size - 1;
}

/* perform_LS, Solis-Wets with SWAT3 (kernel3.cl:30-436) */
static void ls_sw(orc_job* j, int r, int le) {
    const adgpu_ligand* L = j->lig;
    const int pop = j->sp.pop_size, nls = ls_entities(j), NG = L->num_rotbonds + 6;
    const int blk = r * nls + le;
    uint32_t* st = j->prng + (size_t)blk * ADGPU_NWI;
    int ent = ls_pick(j, r, le, &st[0]);
    float* geno = j->pop_next + (size_t)(r * pop + ent) * 64;
    float off[64], cand[64], dev[64], bias[64];
    memcpy(off, geno, NG * sizeof(float));
    memset(bias, 0, sizeof(bias));
    memset(dev, 0, sizeof(dev));
    memcpy(cand, off, sizeof(float) * NG);
    float off_e = j->ene_next[r * pop + ent];
    float rho = 1.0f;
    unsigned cons_succ = 0, cons_fail = 0, iters = 0;
    int evals = 0;
    const int gene_start = (L->true_ligand_atoms == 0) * 6;
    const float lig_scale = 1.0f / sqrtf((float)L->num_atoms);
    const float gene_scale = 1.0f / sqrtf((float)NG);
    while (iters < (unsigned)j->sp.max_num_of_iters && rho > j->sp.rho_lower_bound) {
        for (int k = gene_start; k < NG; ++k) {
            float a = rho * (2.0f * orc_randf(&st[k]) - 1.0f);
            float m = (orc_randf(&st[k]) < gene_scale) ? 1.0f : 0.0f;
            dev[k] = a * m;
            if (k < 3) dev[k] *= j->sp.base_dmov_mul_sqrt3;
            else if (k < 6) dev[k] *= j->sp.base_dang_mul_sqrt3 * lig_scale;
            else dev[k] *= j->sp.base_dang_mul_sqrt3 * gene_scale;
        }
        for (int k = gene_start; k < NG; ++k) cand[k] = off[k] + dev[k] + bias[k];
        float ce = orc_energy_f32(j->grid, L, j->ep, cand, NULL, NULL, NULL);
        evals++;
        if (ce < off_e) {
            for (int k = gene_start; k < NG; ++k) { off[k] = cand[k]; bias[k] = 0.6f * bias[k] + 0.4f * dev[k]; }
            off_e = ce; cons_succ++; cons_fail = 0;
        } else {
            for (int k = gene_start; k < NG; ++k) cand[k] = off[k] - dev[k] - bias[k];
            ce = orc_energy_f32(j->grid, L, j->ep, cand, NULL, NULL, NULL);
            evals++;
            if (ce < off_e) {
                for (int k = gene_start; k < NG; ++k) { off[k] = cand[k]; bias[k] = 0.6f * bias[k] - 0.4f * dev[k]; }
                off_e = ce; cons_succ++; cons_fail = 0;
            } else {
                for (int k = gene_start; k < NG; ++k) bias[k] = 0.5f * bias[k];
                cons_succ = 0; cons_fail++;
            }
        }
        iters++;
        if (cons_succ >= (unsigned)j->sp.cons_limit) { rho *= 2.0f; cons_succ = 0; }
        else if (cons_fail >= (unsigned)j->sp.cons_limit) { rho *= 0.5f; cons_fail = 0; }
    }
    j->evals_new[r * pop + ent] += evals;
    j->ene_next[r * pop + ent] = off_e;
    for (int k = 3; k < NG; ++k) map_angle(&off[k]);
    memcpy(geno, off, NG * sizeof(float));
}

/* gradient_minAD (kernel_ad.cl:60-549), ADADELTA_AUTOSTOP on */
static void ls_ad(orc_job* j, int r, int le) {
    const adgpu_ligand* L = j->lig;
    const int pop = j->sp.pop_size, nls = ls_entities(j), NG = L->num_rotbonds + 6;
    uint32_t* st = j->prng + (size_t)(r * nls + le) * ADGPU_NWI;
    int ent = ls_pick(j, r, le, &st[0]);
    float* geno_g = j->pop_next + (size_t)(r * pop + ent) * 64;
    float geno[64], best[64], grad[64], sqg[64], delta[64], sqd[64];
    memcpy(geno, geno_g, NG * sizeof(float));
    for (int k = 0; k < NG; ++k) { grad[k] = 0; sqg[k] = 0; delta[k] = 0; sqd[k] = 0; best[k] = geno[k]; }
    float best_e = INFINITY, varrho = 1.0f;
    int cons_succ = 0, cons_fail = 0;
    unsigned iters = 0;
    do {
        float e = orc_energrad_f32(j->grid, L, j->ep, geno, 1, grad);
        for (int k = 0; k < NG; ++k) {
            if (e < best_e) best[k] = geno[k];
            sqg[k] *= 0.8f;
            sqg[k] += (1.0f - 0.8f) * grad[k] * grad[k];
            delta[k] = -1.0f * grad[k] * sqrtf((sqd[k] + 1e-2f) / (sqg[k] + 1e-2f));
            sqd[k] *= 0.8f;
            sqd[k] += (1.0f - 0.8f) * delta[k] * delta[k];
            geno[k] += delta[k];
        }
        if (e < best_e) { best_e = e; cons_succ++; cons_fail = 0; }
        else { cons_succ = 0; cons_fail++; }
        iters++;
        if (cons_succ >= 4) { varrho *= 2.0f; cons_succ = 0; }
        else if (cons_fail >= 4) { varrho *= 0.5f; cons_fail = 0; }
    } while (iters < (unsigned)j->sp.max_num_of_iters && varrho > 0.01f);
    for (int k = 3; k < NG; ++k) map_angle(&best[k]);
    memcpy(geno_g, best, NG * sizeof(float));
    j->evals_new[r * pop + ent] += iters;
    j->ene_next[r * pop + ent] = best_e;
}

/* gpu_gradient_minAdam (cuda/kernel_adam.cu:53-416), AD_RHO_CRITERION off */
static void ls_adam(orc_job* j, int r, int le) {
    const adgpu_ligand* L = j->lig;
    const int pop = j->sp.pop_size, nls = ls_entities(j), NG = L->num_rotbonds + 6;
    uint32_t* st = j->prng + (size_t)(r * nls + le) * ADGPU_NWI;
    int ent = ls_pick(j, r, le, &st[0]);
    float* geno_g = j->pop_next + (size_t)(r * pop + ent) * 64;
    float geno[64], best[64], grad[64], mt[64], vt[64];
    memcpy(geno, geno_g, NG * sizeof(float));
    for (int k = 0; k < NG; ++k) { grad[k] = 0; mt[k] = 0; vt[k] = 0; best[k] = geno[k]; }
    float best_e = INFINITY;
    const float b1 = j->sp.adam_beta1, b2 = j->sp.adam_beta2, eps = j->sp.adam_epsilon;
    unsigned iters = 0;
    do {
        float e = orc_energrad_f32(j->grid, L, j->ep, geno, 1, grad);
        float b1p = 1.0f - powf(b1, 1.0f + (float)iters);
        float b2p = 1.0f - powf(b2, 1.0f + (float)iters);
        for (int k = 0; k < NG; ++k) {
            if (e < best_e) best[k] = geno[k];
            mt[k] = b1 * mt[k] + (1.0f - b1) * grad[k];
            vt[k] = b2 * vt[k] + (1.0f - b2) * grad[k] * grad[k];
            float mp = mt[k] / b1p, vp = vt[k] / b2p;
            geno[k] -= mp / (sqrtf(vp) + eps);
        }
        iters++;
        if (e < best_e) best_e = e;
    } while (iters < (unsigned)j->sp.max_num_of_iters);
    for (int k = 3; k < NG; ++k) map_angle(&best[k]);
    memcpy(geno_g, best, NG * sizeof(float));
    j->evals_new[r * pop + ent] += iters;
    j->ene_next[r * pop + ent] = best_e;
}

/* gradient_minFire (kernel_fire.cl:75-895) */
static void ls_fire(orc_job* j, int r, int le) {
    const adgpu_ligand* L = j->lig;
    const int pop = j->sp.pop_size, nls = ls_entities(j), NG = L->num_rotbonds + 6;
    uint32_t* st = j->prng + (size_t)(r * nls + le) * ADGPU_NWI;
    int ent = ls_pick(j, r, le, &st[0]);
    float* geno_g = j->pop_next + (size_t)(r * pop + ent) * 64;
    float geno[64], grad[64], cgeno[64], cgrad[64], best[64], vel[64];
    memcpy(geno, geno_g, NG * sizeof(float));
    float energy = j->ene_next[r * pop + ent];
    orc_gradient_f32(j->grid, L, j->ep, geno, grad);
    float alpha = 0.2f, dt = 10.0f / 3.0f;
    unsigned count_success = 0, iters = 0;
    float s = 0; for (int k = 0; k < NG; ++k) s += grad[k] * grad[k];
    float inv_gn = 0.2f * (1.0f / sqrtf(s));
    for (int k = 0; k < NG; ++k) vel[k] = -grad[k] * inv_gn;
    float best_e = energy;
    memcpy(best, geno, NG * sizeof(float));
    do {
        for (int k = 0; k < NG; ++k) cgeno[k] = geno[k] + dt * vel[k];
        float ce = orc_energrad_f32(j->grid, L, j->ep, cgeno, 1, cgrad);
        float power = 0; for (int k = 0; k < NG; ++k) power += -cgrad[k] * vel[k];
        if (power < 0.0f) {
            s = 0; for (int k = 0; k < NG; ++k) s += grad[k] * grad[k];
            inv_gn = 0.2f * (1.0f / sqrtf(s));
            for (int k = 0; k < NG; ++k) vel[k] = -grad[k] * inv_gn;
            count_success = 0; alpha = 0.2f; dt = dt * 0.8f;
        } else {
            count_success++;
            if (count_success > 5) { dt = fminf(dt * 1.2f, 10.0f); alpha = alpha * 0.99f; }
        }
        energy = ce;
        for (int k = 0; k < NG; ++k) { geno[k] = cgeno[k]; grad[k] = cgrad[k]; }
        s = 0; for (int k = 0; k < NG; ++k) s += grad[k] * grad[k];
        inv_gn = 1.0f / sqrtf(s);
        float vn = 0; for (int k = 0; k < NG; ++k) vn += vel[k] * vel[k];
        vn = sqrtf(vn);
        float vdg = alpha * vn * inv_gn;
        for (int k = 0; k < NG; ++k) vel[k] = (1 - alpha) * vel[k] - grad[k] * vdg;
        iters++;
        if (energy < best_e) { best_e = energy; memcpy(best, geno, NG * sizeof(float)); }
    } while (iters < (unsigned)j->sp.max_num_of_iters && dt > 1e-6f);
    j->evals_new[r * pop + ent] += iters;
    j->ene_next[r * pop + ent] = best_e;
    for (int k = 3; k < NG; ++k) map_angle(&best[k]);
    memcpy(geno_g, best, NG * sizeof(float));
}

/* LS dispatch (performdocking.cpp.Cuda:727-785; OpenCL host adds FIRE). */
static int ls_method_of(const orc_job* j, int64_t gen) {
    int method = j->sp.ls_method;
    if (method == ADGPU_LS_AD && gen < j->sp.initial_sw_generations) method = ADGPU_LS_SW;
    return method;
}

static void ls_one(orc_job* j, int r, int le, int method) {
    switch (method) {
        case ADGPU_LS_SW:   ls_sw(j, r, le); break;
        case ADGPU_LS_AD:   ls_ad(j, r, le); break;
        case ADGPU_LS_ADAM: ls_adam(j, r, le); break;
        case ADGPU_LS_FIRE: ls_fire(j, r, le); break;
        default: break;
    }
}

void orc_ls(orc_job* j, int r0, int r1, int64_t gen) {
    if (j->sp.lsearch_rate == 0.0f) return;
    const int nls = ls_entities(j), method = ls_method_of(j, gen);
    for (int r = r0; r < r1; ++r)
        for (int le = 0; le < nls; ++le) ls_one(j, r, le, method);
}

/* One generation of runs [r0, r1) split over nthreads threads by
 * (run, individual) -- the work-groups the reference's OpenCL CPU runtime
 * schedules (one per individual in kernel4, one per LS entity in the local
 * search).  Every individual and every search only touches its own entity,
 * so the result does not depend on the split; the one exception is the
 * clamped entity choice when all individuals are searched at a rate below
 * 100 % (ls_pick), where each run's searches then stay in order. */
static void generation_split(orc_job* j, int r0, int r1, int64_t gen, int nthreads) {
    const int pop = j->sp.pop_size;
    const int64_t ng = (int64_t)(r1 - r0) * pop;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (int64_t k = 0; k < ng; ++k) gen_one(j, r0 + (int)(k / pop), (int)(k % pop));
    if (j->sp.lsearch_rate != 0.0f) {
        const int nls = ls_entities(j), method = ls_method_of(j, gen);
        if (nls >= pop && j->sp.lsearch_rate < 100.0f) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
            for (int r = r0; r < r1; ++r)
                for (int le = 0; le < nls; ++le) ls_one(j, r, le, method);
        } else {
            const int64_t nl = (int64_t)(r1 - r0) * nls;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
            for (int64_t k = 0; k < nl; ++k) ls_one(j, r0 + (int)(k / nls), (int)(k % nls), method);
        }
    }
    orc_sum_evals(j, r0, r1);
}

/* check_progress (performdocking.cpp.Cuda:908-942) */
static double check_progress(const orc_job* j, int64_t gen, int64_t* total) {
    int64_t t = 0;
    for (int r = 0; r < j->sp.num_runs; ++r) t += j->evals_runs[r];
    *total = t;
    double ep = (double)t / (double)j->sp.num_runs / (double)j->sp.max_evals * 100.0;
    double gp = (double)gen / (double)j->sp.max_generations * 100.0;
    return ep > gp ? ep : gp;
}

int64_t orc_lga_run(orc_job* j, int nthreads, int64_t* total_evals) {
    const int R_ = j->sp.num_runs;
    /* runs are independent except for the shared stopping test, which reads
     * the totals after every generation; the generation loop therefore stays
     * global and each generation's work is split over (run, individual). */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int r = 0; r < R_; ++r) orc_initpop(j, r, r + 1);
    int64_t gen = 0, tot = 0;
    while (check_progress(j, gen, &tot) < 100.0) {
        generation_split(j, 0, R_, gen, nthreads);
        gen++;
        orc_swap(j);
    }
    if (total_evals) *total_evals = tot;
    j->generation = gen;
    return gen;
}

int64_t orc_lga_generations(orc_job* j, int r0, int r1, int64_t ngen, int nthreads) {
    (void)nthreads;
    int64_t before = 0, after = 0;
    for (int r = r0; r < r1; ++r) before += j->evals_runs[r];
    for (int64_t g = 0; g < ngen; ++g) {
        generation_split(j, r0, r1, j->generation + g, nthreads);
        orc_swap(j);
    }
    j->generation += ngen;
    for (int r = r0; r < r1; ++r) after += j->evals_runs[r];
    return after - before;
}

/* ---- host CMWC4096 (G. Marsaglia, JMASM 2003) with the reference seeding,
 * host/inc/miscellaneous.h:146-218 ---------------------------------------- */
typedef struct { uint32_t Q[4096], i, c; } Cmwc;
#define CMWC_PHI 0x9e3779b9u

static uint32_t cmwc_next(Cmwc* g) {
    const uint64_t a = 18782ULL;
    const uint32_t m = 0xfffffffeu;
    g->i = (g->i + 1) & 4095;
    uint64_t t = a * g->Q[g->i] + g->c;
    g->c = (uint32_t)(t >> 32);
    uint32_t x = (uint32_t)(t + g->c);
    if (x < g->c) { x++; g->c++; }
    return (g->Q[g->i] = m - x);
}

static void cmwc_init3(Cmwc* g, uint32_t x, uint32_t y, uint32_t z) {
    g->Q[0] = x; g->Q[1] = y + CMWC_PHI; g->Q[2] = z + CMWC_PHI;
    for (uint32_t k = 3; k < 4096; ++k) g->Q[k] = g->Q[k - 3] ^ g->Q[k - 2] ^ CMWC_PHI ^ k;
    g->i = 4095; g->c = 362436;
    do g->c = cmwc_next(g); while (g->c >= 809430660u);
    g->i = 4095;
}

static void cmwc_seed(Cmwc* g, const uint32_t s[3]) {
    if (!s[2]) {
        if (!s[1]) cmwc_init3(g, s[0], s[0] + CMWC_PHI, s[0] + CMWC_PHI + CMWC_PHI);
        else cmwc_init3(g, s[0], s[1], s[1] + CMWC_PHI);
    } else cmwc_init3(g, s[0], s[1], s[2]);
}

static float cmwc_float(Cmwc* g) {
    float r = 0.0f;
    while (r <= 0.0f || r >= 1.0f) r = (float)cmwc_next(g) / 4294967295.0f;
    return r;
}

void orc_cmwc_uints(const uint32_t seed[3], uint32_t* out, int64_t n) {
    Cmwc* g = (Cmwc*)malloc(sizeof(Cmwc));
    cmwc_seed(g, seed);
    for (int64_t k = 0; k < n; ++k) out[k] = cmwc_next(g);
    free(g);
}

void orc_initial_population(const uint32_t seed[3], int runs, int pop, const float size_xyz_angstr[3],
                            float spacing, float* out) {
    Cmwc* g = (Cmwc*)malloc(sizeof(Cmwc));
    cmwc_seed(g, seed);
    const float PI2 = 2.0f * 3.14159265358979323846f;
    for (long e = 0; e < (long)runs * pop; ++e) {
        float* o = out + e * 64;
        for (int k = 0; k < 3; ++k) o[k] = (float)((double)cmwc_float(g) * (double)size_xyz_angstr[k] / (double)spacing);
        float u1 = cmwc_float(g), u2 = cmwc_float(g), u3 = cmwc_float(g);
        /* C++ overload resolution in the reference: float arguments call the
         * float functions, mixed ones the double ones. */
        float qw = (float)(sqrt(1.0 - u1) * sinf(PI2 * u2));
        float qx = (float)(sqrt(1.0 - u1) * cosf(PI2 * u2));
        float qy = (float)(sqrt((double)u1) * sinf(PI2 * u3));
        float qz = (float)(sqrt((double)u1) * cosf(PI2 * u3));
        float rotangle = (float)(2.0 * acosf(qw));
        float s = (float)sqrt(1.0 - (double)(qw * qw));
        float x, y, z;
        if (s < 0.001f) { x = qx; y = qy; z = qz; } else { x = qx / s; y = qy / s; z = qz / s; }
        float theta = acosf(z);
        float phi = atan2f(y, x);
        o[3] = phi / 0.0174533f; o[4] = theta / 0.0174533f; o[5] = rotangle / 0.0174533f;
        for (int k = 6; k < ADGPU_MAX_ROTBONDS + 6; ++k) o[k] = cmwc_float(g) * 360;
        o[63] = 0.0f;
    }
    free(g);
}
#endif /* !ORC_DOUBLE */
