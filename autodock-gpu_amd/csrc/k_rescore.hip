// k_rescore.hip -- FP64 energies of final poses on the GPU.
//
// The reference re-scores its final population on the CPU in FP32
// (process_result, host/src/processresult.cpp:430-466).  Here each run's best
// pose can be re-scored in FP64 on the device: the same pose -> coordinates,
// trilinear inter energy and intramolecular pair terms as gpu_calc_energy
// (calcenergy.cl:132-554), every operation in double, with the reference's
// single-precision constants and map values widened exactly.  One wavefront
// per pose; this runs once per run, so it is written for accuracy, not speed.
#include <type_traits>

#include "adgpu_device.h"

namespace adgpu {

namespace {

struct D3 {
    double x, y, z;
};

DEV D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV D3 dadd(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }

// quaternion_rotate, calcenergy.cl:119-127, in double
DEV D3 qrot_d(D3 v, double qx, double qy, double qz, double qw) {
    const double zx = (qy * v.z - qz * v.y) * 2.0, zy = (qz * v.x - qx * v.z) * 2.0, zz = (qx * v.y - qy * v.x) * 2.0;
    return D3{v.x + zx * qw + (qy * zz - qz * zy), v.y + zy * qw + (qz * zx - qx * zz),
              v.z + zz * qw + (qx * zy - qy * zx)};
}

DEV double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

}  // namespace

template <int MF>
__global__ __launch_bounds__(64) void k_rescore_f64(const LigandDev* L, GridDev G, const float* geno, int n,
                                                    double* out3) {
    __shared__ double cx[ADGPU_MAX_ATOMS], cy[ADGPU_MAX_ATOMS], cz[ADGPU_MAX_ATOMS];
    const int pose = blockIdx.x;
    if (pose >= n) return;
    const int lane = threadIdx.x;
    const char* gb = reinterpret_cast<const char*>(L);
    const float4* ar = reinterpret_cast<const float4*>(gb + L->off_atom_ref);
    const int4* ai = reinterpret_cast<const int4*>(gb + L->off_atom_info);
    const uint2* am = reinterpret_cast<const uint2*>(gb + L->off_atom_mask);
    const float4* tu = reinterpret_cast<const float4*>(gb + L->off_tor_unit);
    const float4* tm = reinterpret_cast<const float4*>(gb + L->off_tor_move);
    const PairA* pa = reinterpret_cast<const PairA*>(gb + L->off_pair_a);
    const PairB* pb = reinterpret_cast<const PairB*>(gb + L->off_pair_b);
    const float2* tds = reinterpret_cast<const float2*>(gb + L->off_type_ds);
    const float* g = geno + (size_t)pose * kGenes;
    const double d2r = (double)kDegToRad;
    const int N = L->num_atoms, P = L->num_pairs;

    // general rotation, calcenergy.cl:201-221
    double gqx = 0.0, gqy = 0.0, gqz = 0.0, gqw = 1.0;
    D3 gmv{0.0, 0.0, 0.0};
    if (L->true_ligand_atoms) {
        const double phi = (double)g[3] * d2r, theta = (double)g[4] * d2r, rot = (double)g[5] * d2r;
        const double st = sin(theta), ct = cos(theta), s2 = sin(rot * 0.5);
        gqx = s2 * st * cos(phi); gqy = s2 * st * sin(phi); gqz = s2 * ct; gqw = cos(rot * 0.5);
        gmv = D3{(double)g[0], (double)g[1], (double)g[2]};
    }
    double e_inter = 0.0, e_intra = 0.0;
    const uint32_t ne = (uint32_t)(G.nmaps - 2) * G.g3, nd = ne + G.g3;
    for (int a = lane; a < N; a += 64) {
        const float4 r = ar[a];
        D3 v{(double)r.x, (double)r.y, (double)r.z};
        const uint2 mk = am[a];
        uint64_t mask = (uint64_t)mk.x | ((uint64_t)mk.y << 32);
        while (mask) {   // rotatable bonds in ascending index (gen_rotlist, calcenergy.cpp:387-488)
            const int t = __builtin_ctzll(mask);
            mask &= mask - 1;
            const double ang = (double)g[6 + t] * d2r * 0.5;
            const double s = sin(ang), c = cos(ang);
            const float4 u = tu[t], m = tm[t];
            const D3 mv{(double)m.x, (double)m.y, (double)m.z};
            v = dadd(qrot_d(dsub(v, mv), s * (double)u.x, s * (double)u.y, s * (double)u.z, c), mv);
        }
        if (a < L->true_ligand_atoms) v = dadd(qrot_d(v, gqx, gqy, gqz, gqw), gmv);
        cx[a] = v.x; cy[a] = v.y; cz[a] = v.z;
        const int4 inf = ai[a];
        if (inf.z) continue;   // ignore_inter
        if ((v.x < 0.0) || (v.y < 0.0) || (v.z < 0.0) || (v.x >= (double)(G.gx - 1)) ||
            (v.y >= (double)(G.gy - 1)) || (v.z >= (double)(G.gz - 1)) || !(v.x == v.x) || !(v.y == v.y) ||
            !(v.z == v.z)) {
            e_inter += 16777216.0;
            continue;
        }
        const double xl = floor(v.x), yl = floor(v.y), zl = floor(v.z);
        const uint32_t p = (uint32_t)xl + (uint32_t)yl * G.g1 + (uint32_t)zl * G.g2;
        const double dx = v.x - xl, omdx = 1.0 - dx, dy = v.y - yl, omdy = 1.0 - dy, dz = v.z - zl, omdz = 1.0 - dz;
        const double w[8] = {omdx * omdy * omdz, omdx * dy * omdz, omdx * omdy * dz, omdx * dy * dz,
                             dx * omdy * omdz,   dx * dy * omdz,   dx * omdy * dz,   dx * dy * dz};
        const uint32_t base[3] = {(uint32_t)inf.x * G.g3, ne, nd};
        double val[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float4 lo = load_pt<MF>(G.data, base[k] + p), hi = load_pt<MF>(G.data, base[k] + p + 1);
            val[k] = (double)lo.x * w[0] + (double)lo.y * w[1] + (double)lo.z * w[2] + (double)lo.w * w[3] +
                     (double)hi.x * w[4] + (double)hi.y * w[5] + (double)hi.z * w[6] + (double)hi.w * w[7];
        }
        const double q = (double)r.w;
        e_inter += val[0];
        e_inter += q * val[1];
        e_inter += fabs(q) * val[2];
    }
    __syncthreads();
    const double spacing = (double)G.spacing;
    const double delta_distance = (double)L->half_smooth;
    const double elec_min = (double)L->elec_min_distance;
    const double coeff_elec = (double)L->coeff_elec, coeff_desolv = (double)L->coeff_desolv, qasp = (double)L->qasp;
    for (int k = lane; k < P; k += 64) {   // pair terms, calcenergy.cl:397-536
        const PairA A = pa[k];
        const PairB B = pb[k];
        const int i = (int)(A.ij & 0xffffu), j = (int)(A.ij >> 16);
        const double sx = cx[i] - cx[j], sy = cy[i] - cy[j], sz = cz[i] - cz[j];
        const double dist = sqrt(sx * sx + sy * sy + sz * sz) * spacing;
        double e = ((A.flags >> 16) & 1u) ? 50.0 * dist : 0.0;   // CG-G0 bond term
        if (dist < 8.0) {
            const int m = (int)(A.flags & 0xffu), nn = (int)((A.flags >> 8) & 0xffu);
            const double opt = (double)B.reqm, od = opt - dist;
            const double sd = fabs(od) >= delta_distance ? dist + copysign(delta_distance, od) : opt;
            const double rmn = pow(sd, (double)(m - nn)), rm = pow(sd, (double)(-m));
            e += ((double)A.ac - rmn * (double)A.bd) * rm;
        }
        if (dist < (double)20.48f) {
            const double d = dist < elec_min ? elec_min : dist;
            const int4 ii = ai[i], ij = ai[j];
            const double q1 = (double)ar[i].w, q2 = (double)ar[j].w;
            const float2 s1 = tds[ii.y], s2 = tds[ij.y];
            const double dist2 = d * d;
            const double desolv =
                (((double)s1.x + qasp * fabs(q1)) * (double)s2.y + ((double)s2.x + qasp * fabs(q2)) * (double)s1.y) *
                ((coeff_desolv * ((double)12.96f - (double)0.1063f * dist2 * (1.0 - (double)0.001947f * dist2))) /
                 ((double)12.96f + dist2 * ((double)0.4137f + dist2 * ((double)0.00357f + (double)0.000112f * dist2))));
            const double shift = d + (double)1.26366f;
            const double diel = (double)1.10859f / (shift * shift) + (double)0.010358f;
            const double es = coeff_elec * q1 * q2 / d;
            e += diel * es + desolv;
        }
        e_intra += e;
    }
    e_inter = wave_sum_d(e_inter);
    e_intra = wave_sum_d(e_intra);
    if (lane == 0) {
        out3[3 * (size_t)pose] = e_inter + e_intra;
        out3[3 * (size_t)pose + 1] = e_inter;
        out3[3 * (size_t)pose + 2] = e_intra;
    }
}

hipError_t launch_rescore_f64(const LaunchArgs& la, const float* genotypes, int n, double* out3) {
    if (n <= 0) return hipSuccess;
    switch (la.grid.format) {
        case ADGPU_MAP_F16:
            hipLaunchKernelGGL(k_rescore_f64<ADGPU_MAP_F16>, dim3(n), dim3(64), 0, la.stream, la.lig_dev, la.grid,
                               genotypes, n, out3);
            break;
        case ADGPU_MAP_BF16:
            hipLaunchKernelGGL(k_rescore_f64<ADGPU_MAP_BF16>, dim3(n), dim3(64), 0, la.stream, la.lig_dev, la.grid,
                               genotypes, n, out3);
            break;
        default:
            hipLaunchKernelGGL(k_rescore_f64<ADGPU_MAP_F32>, dim3(n), dim3(64), 0, la.stream, la.lig_dev, la.grid,
                               genotypes, n, out3);
    }
    return hipGetLastError();
}

}  // namespace adgpu
