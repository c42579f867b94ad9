// adgpu_kernels.hip -- gfx950 kernels of the docking hot path.
//
// Execution model (CDNA, 64-wide wavefronts):
//   * one wavefront evaluates one individual; a workgroup holds
//     kWavesPerGroup individuals that share the ligand tables staged once
//     into LDS (pair parameters, torsion vectors);
//   * lane a owns atom a (atoms 64..255 are handled as extra chunks); each
//     lane walks its own torsion chain, so pose -> coordinates needs no
//     barriers (the reference's rotation list, calcenergy.cl:232-275, is a
//     barrier per cycle);
//   * lane g owns gene g of the genotype (<= 63 genes) for the whole local
//     search, so the minimiser updates are lane-parallel register code;
//   * pairs are strided over lanes; pair forces go to per-atom int32
//     fixed-point accumulators in LDS through ds_add (order-independent, so
//     results do not depend on scheduling); energies and torques are summed
//     with a fixed butterfly, giving identical results in every lane.
// The arithmetic of every device function follows the reference file:line
// given at its definition (ccsb-scripps/AutoDock-GPU v1.5, device/ = OpenCL).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "adgpu_internal.h"

namespace adgpu {

#define DEV __device__ __forceinline__

// common/calcenergy_basic.h
constexpr float kDegToRad = 0.0174533f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kTermScale = 1024.0f;
constexpr float kOneOverTermScale = 1.0f / 1024.0f;
constexpr float kMaxTerm = 8192.0f;
constexpr int kCG = 24, kG0 = 25;   // ATYPE_CG_IDX, ATYPE_G0_IDX (common/defines.h)

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
DEV int lane_id() { return __lane_id(); }

// Orders LDS accesses of the lanes of one wavefront (no workgroup barrier).
DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Butterfly all-reduce: every lane receives the same bits.
DEV float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
DEV int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

DEV float bcast(float v, int src) { return __shfl(v, src); }

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
DEV float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
DEV float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DEV float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// fmod_pi2, calcenergy.cl:86-91
DEV float fmod_pi2(float x) { return x - (float)((int)((1.0f / kTwoPi) * x)) * kTwoPi; }

// fast_acos, calcenergy.cl:93-108
DEV float fast_acos(float cosine) {
    const float a = 9.78056e-05f, b = -0.00104588f, c = 0.00418716f, d = -0.00314347f;
    const float e = 2.74084f, f = 0.370388f;
    const float o = -(a + b + c + d);
    float x = fabsf(cosine);
    float x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    float ac = (((o * x4 + a) * x3 + b) * x2 + c) * x + d +
               e * fast_sqrt(2.0f - fast_sqrt(2.0f + 2.0f * x)) - f * fast_sqrt(2.0f - 2.0f * x);
    return copysignf(ac, cosine) + (cosine < 0.0f ? kPi : 0.0f);
}

// float2int_round of the clamped term, calcMergedEneGra.cl:30-35 (the half is
// moved toward zero before truncation, as the reference does).
DEV int quant(float x) {
    float v = fminf(kMaxTerm, fmaxf(-kMaxTerm, x));
    int addsub = (int)(v < 0.0f) - (int)(v > 0.0f);
    return (int)(v + 0.5f * (float)addsub);
}

// quaternion_rotate, calcenergy.cl:119-127
DEV float3 qrot(float3 v, float4 q) {
    float zx = (q.y * v.z - q.z * v.y) * 2.0f;
    float zy = (q.z * v.x - q.x * v.z) * 2.0f;
    float zz = (q.x * v.y - q.y * v.x) * 2.0f;
    return make_float3(v.x + zx * q.w + (q.y * zz - q.z * zy),
                       v.y + zy * q.w + (q.z * zx - q.x * zz),
                       v.z + zz * q.w + (q.x * zy - q.y * zx));
}

// quaternion_multiply, calcenergy.cl:110-117
DEV float4 qmul(float4 a, float4 b) {
    return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                       a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                       a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                       a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}

// map_angle, auxiliary_genetic.cl:72-85.  Values the subtraction loop could
// not bring into range (|a| >= 1e6, where a - 360 may round back to a, and
// inf) are reduced with fmodf first; the reference loops forever on them.
DEV void map_angle(float& a) {
    if (!(fabsf(a) < 1.0e6f)) a = fmodf(a, 360.0f);
    if (!(a == a) || isinf(a)) return;
    while (a >= 360.0f) a -= 360.0f;
    while (a < 0.0f) a += 360.0f;
}

// LCG, auxiliary_genetic.cl:29-67 (state lives in global memory)
DEV float randf(uint32_t* st) {
    uint32_t s = 1103515245u * (*st) + 12345u;
    *st = s;
    return ((float)s * (1.0f / 4294967296.0f)) * 0.999999f;
}

// ---------------------------------------------------------------------------
// map gathers: one grid point of one map = 4 values (packed-4 layout)
// ---------------------------------------------------------------------------
template <int MF>
DEV float4 load_pt(const void* base, uint32_t idx);

template <>
DEV float4 load_pt<ADGPU_MAP_F32>(const void* base, uint32_t idx) {
    return reinterpret_cast<const float4*>(base)[idx];
}
template <>
DEV float4 load_pt<ADGPU_MAP_F16>(const void* base, uint32_t idx) {
    uint2 r = reinterpret_cast<const uint2*>(base)[idx];
    __half2 lo = *reinterpret_cast<__half2*>(&r.x), hi = *reinterpret_cast<__half2*>(&r.y);
    float2 a = __half22float2(lo), b = __half22float2(hi);
    return make_float4(a.x, a.y, b.x, b.y);
}
template <>
DEV float4 load_pt<ADGPU_MAP_BF16>(const void* base, uint32_t idx) {
    uint2 r = reinterpret_cast<const uint2*>(base)[idx];
    return make_float4(__uint_as_float(r.x << 16), __uint_as_float(r.x & 0xffff0000u),
                       __uint_as_float(r.y << 16), __uint_as_float(r.y & 0xffff0000u));
}

// ---------------------------------------------------------------------------
// LDS layout
// ---------------------------------------------------------------------------
struct SharedView {           // staged once per workgroup
    const PairRec* pairs;     // LDS or global
    const float4* tor_unit;
    const float4* tor_move;
    const int2* tor_atoms;
};

struct WaveView {             // private to one wavefront
    float4* coords;           // [npad]
    int* fx;                  // [npad]
    int* fy;
    int* fz;
    float4* tq;               // [64] torsion quaternions
    float4* tax;              // [64] torsion axes (gradient)
    float4* tref;             // [64] torsion reference atoms (gradient)
    int* gi;                  // [64] int gene gradients
};

DEV size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t shared_bytes_for(int T, int P, bool pairs_in_lds) {
    size_t s = (pairs_in_lds ? (size_t)P * sizeof(PairRec) : 0);
    s += (size_t)T * (sizeof(float4) * 2 + sizeof(int2));
    return (s + 15) & ~size_t(15);
}

__host__ __device__ inline size_t wave_bytes_for(int npad) {
    return (size_t)npad * (16 + 12) + 3 * 64 * 16 + 64 * 4;
}

size_t shared_bytes(const LigandDev& L, bool pairs_in_lds) {
    return shared_bytes_for(L.num_rotbonds, L.num_pairs, pairs_in_lds) +
           kWavesPerGroup * wave_bytes_for(64 * L.nchunk);
}

// Stages the shared tables (all threads of the workgroup) and carves the
// per-wave region.  Contains the only workgroup barrier of every kernel.
template <bool PL>
DEV void stage(const LigandDev* L, char* lds, SharedView& sv, WaveView& wv, int wave) {
    const char* gbase = reinterpret_cast<const char*>(L);
    const int T = L->num_rotbonds, P = L->num_pairs, npad = 64 * L->nchunk;
    char* p = lds;
    const PairRec* gpairs = reinterpret_cast<const PairRec*>(gbase + L->off_pairs);
    if (PL) {
        PairRec* lp = reinterpret_cast<PairRec*>(p);
        const uint4* src = reinterpret_cast<const uint4*>(gpairs);
        uint4* dst = reinterpret_cast<uint4*>(lp);
        for (int k = threadIdx.x; k < 2 * P; k += blockDim.x) dst[k] = src[k];
        sv.pairs = lp;
        p += (size_t)P * sizeof(PairRec);
    } else {
        sv.pairs = gpairs;
    }
    float4* tu = reinterpret_cast<float4*>(p);
    float4* tm = tu + T;
    int2* ta = reinterpret_cast<int2*>(tm + T);
    const float4* gtu = reinterpret_cast<const float4*>(gbase + L->off_tor_unit);
    const float4* gtm = reinterpret_cast<const float4*>(gbase + L->off_tor_move);
    const int2* gta = reinterpret_cast<const int2*>(gbase + L->off_tor_atoms);
    for (int k = threadIdx.x; k < T; k += blockDim.x) { tu[k] = gtu[k]; tm[k] = gtm[k]; ta[k] = gta[k]; }
    sv.tor_unit = tu; sv.tor_move = tm; sv.tor_atoms = ta;
    size_t sh = shared_bytes_for(T, P, PL);
    char* w = lds + sh + (size_t)wave * wave_bytes_for(npad);
    wv.coords = reinterpret_cast<float4*>(w);
    wv.fx = reinterpret_cast<int*>(wv.coords + npad);
    wv.fy = wv.fx + npad;
    wv.fz = wv.fy + npad;
    wv.tq = reinterpret_cast<float4*>(wv.fz + npad);
    wv.tax = wv.tq + 64;
    wv.tref = wv.tax + 64;
    wv.gi = reinterpret_cast<int*>(wv.tref + 64);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// per-lane atom data, loaded once per kernel
// ---------------------------------------------------------------------------
template <int NC>
struct Atoms {
    float3 ref[NC];
    float q[NC];
    uint32_t mlo[NC], mhi[NC];
    uint32_t mapoff[NC];      // map index * g3
    bool valid[NC], ign[NC], lig[NC];
    int32_t type[NC];

    DEV void load(const LigandDev* L, const GridDev& G) {
        const char* gb = reinterpret_cast<const char*>(L);
        const float4* ar = reinterpret_cast<const float4*>(gb + L->off_atom_ref);
        const int4* ai = reinterpret_cast<const int4*>(gb + L->off_atom_info);
        const uint2* am = reinterpret_cast<const uint2*>(gb + L->off_atom_mask);
        const int lane = lane_id();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            int a = lane + 64 * k;
            valid[k] = a < L->num_atoms;
            int aa = valid[k] ? a : 0;
            float4 r = ar[aa];
            int4 inf = ai[aa];
            uint2 m = am[aa];
            ref[k] = make_float3(r.x, r.y, r.z);
            q[k] = r.w;
            mlo[k] = valid[k] ? m.x : 0u;
            mhi[k] = valid[k] ? m.y : 0u;
            mapoff[k] = (uint32_t)inf.x * G.g3;
            type[k] = inf.y;
            ign[k] = !valid[k] || inf.z != 0;
            lig[k] = valid[k] && a < L->true_ligand_atoms;
        }
    }
};

// ---------------------------------------------------------------------------
// Pose -> coordinates (calcenergy.cl:185-275).  Per atom: the rotatable bonds
// moving it in ascending index, then the general rotation (the per-atom order
// gen_rotlist encodes, host/src/calcenergy.cpp:387-488).
// ---------------------------------------------------------------------------
struct Pose {
    float4 q;          // genrot quaternion
    float3 mv;         // translation (grid units)
    float phi, theta, rotangle, sin_theta, sin_half, is_theta_gt_pi;
};

DEV Pose pose_from_genes(float gene, bool has_ligand) {
    Pose P;
    float g0 = bcast(gene, 0), g1 = bcast(gene, 1), g2 = bcast(gene, 2);
    float g3 = bcast(gene, 3), g4 = bcast(gene, 4), g5 = bcast(gene, 5);
    P.phi = g3 * kDegToRad;
    P.theta = g4 * kDegToRad;
    P.rotangle = g5 * kDegToRad;
    P.q = make_float4(0.f, 0.f, 0.f, 1.f);
    P.mv = make_float3(0.f, 0.f, 0.f);
    P.sin_theta = 0.f; P.sin_half = 0.f; P.is_theta_gt_pi = 1.f;
    if (has_ligand) {
        float st, ct, sp, cp, sh, ch;
        sincosf(P.theta, &st, &ct);
        sincosf(P.phi, &sp, &cp);
        sincosf(P.rotangle * 0.5f, &sh, &ch);
        P.mv = make_float3(g0, g1, g2);
        P.q = make_float4(sh * st * cp, sh * st * sp, sh * ct, ch);
        P.sin_theta = st;
        P.sin_half = sh;
        P.is_theta_gt_pi = 1.0f - 2.0f * (float)(st < 0.0f);
    }
    return P;
}

template <int NC>
DEV void compute_coords(const LigandDev* L, const SharedView& sv, const WaveView& wv, const Atoms<NC>& A,
                        const Pose& P, float gene, float3 (&c)[NC]) {
    const int lane = lane_id();
    const int T = L->num_rotbonds;
    // torsion quaternions, lane t -> rotatable bond t
    float gt = __shfl(gene, lane + 6 < 64 ? lane + 6 : 63);
    if (lane < T) {
        float ang = gt * kDegToRad * 0.5f;
        float s, co;
        sincosf(ang, &s, &co);
        float4 u = sv.tor_unit[lane];
        wv.tq[lane] = make_float4(s * u.x, s * u.y, s * u.z, co);
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        float3 v = A.ref[k];
        uint32_t lo = A.mlo[k], hi = A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            float4 q = wv.tq[t];
            float4 m = sv.tor_move[t];
            float3 mv = make_float3(m.x, m.y, m.z);
            float3 r = qrot(make_float3(v.x - mv.x, v.y - mv.y, v.z - mv.z), q);
            v = make_float3(r.x + mv.x, r.y + mv.y, r.z + mv.z);
        }
        if (A.lig[k]) {
            float3 r = qrot(v, P.q);
            v = make_float3(r.x + P.mv.x, r.y + P.mv.y, r.z + P.mv.z);
        }
        c[k] = v;
        if (A.valid[k]) wv.coords[lane + 64 * k] = make_float4(v.x, v.y, v.z, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Intermolecular energy (+ map gradient) of one atom, calcenergy.cl:280-379,
// calcMergedEneGra.cl:222-443.
// ---------------------------------------------------------------------------
template <int MF, bool GRAD>
DEV float inter_atom(const GridDev& G, float3 c, float q, uint32_t mapoff, int3& fi) {
    if ((c.x < 0.f) || (c.y < 0.f) || (c.z < 0.f) || (c.x >= (float)(G.gx - 1)) || (c.y >= (float)(G.gy - 1)) ||
        (c.z >= (float)(G.gz - 1))) {
        if (GRAD) { fi.x += 16777216; fi.y += 16777216; fi.z += 16777216; }
        return 16777216.0f;
    }
    float xl = floorf(c.x), yl = floorf(c.y), zl = floorf(c.z);
    uint32_t p = (uint32_t)xl + (uint32_t)yl * G.g1 + (uint32_t)zl * G.g2;
    float dx = c.x - xl, omdx = 1.0f - dx;
    float dy = c.y - yl, omdy = 1.0f - dy;
    float dz = c.z - zl, omdz = 1.0f - dz;
    float w0 = omdx * omdy * omdz, w1 = omdx * dy * omdz, w2 = omdx * omdy * dz, w3 = omdx * dy * dz;
    float w4 = dx * omdy * omdz, w5 = dx * dy * omdz, w6 = dx * omdy * dz, w7 = dx * dy * dz;
    const uint32_t g3 = G.g3;
    const uint32_t ne = (uint32_t)(G.nmaps - 2) * g3, nd = ne + g3;
    float4 a0 = load_pt<MF>(G.data, mapoff + p), a1 = load_pt<MF>(G.data, mapoff + p + 1);
    float4 e0 = load_pt<MF>(G.data, ne + p), e1 = load_pt<MF>(G.data, ne + p + 1);
    float4 d0 = load_pt<MF>(G.data, nd + p), d1 = load_pt<MF>(G.data, nd + p + 1);
    float et = a0.x * w0 + a0.y * w1 + a0.z * w2 + a0.w * w3 + a1.x * w4 + a1.y * w5 + a1.z * w6 + a1.w * w7;
    float ee = e0.x * w0 + e0.y * w1 + e0.z * w2 + e0.w * w3 + e1.x * w4 + e1.y * w5 + e1.z * w6 + e1.w * w7;
    float ed = d0.x * w0 + d0.y * w1 + d0.z * w2 + d0.w * w3 + d1.x * w4 + d1.y * w5 + d1.z * w6 + d1.w * w7;
    float aq = fabsf(q);
    float e = et + q * ee + aq * ed;
    if (GRAD) {
        const float ig = G.inv_spacing;
        // cube index k: 0 000, 1 010, 2 001, 3 011, 4 100, 5 110, 6 101, 7 111
        auto gxf = [&](float4 l, float4 h) {
            return omdz * (omdy * (h.x - l.x) + dy * (h.y - l.y)) + dz * (omdy * (h.z - l.z) + dy * (h.w - l.w));
        };
        auto gyf = [&](float4 l, float4 h) {
            return omdz * (omdx * (l.y - l.x) + dx * (h.y - h.x)) + dz * (omdx * (l.w - l.z) + dx * (h.w - h.z));
        };
        auto gzf = [&](float4 l, float4 h) {
            return omdy * (omdx * (l.z - l.x) + dx * (h.z - h.x)) + dy * (omdx * (l.w - l.y) + dx * (h.w - h.y));
        };
        float qg = q * ig, aqg = fabsf(qg);
        float gx = gxf(a0, a1) * ig + qg * gxf(e0, e1) + aqg * gxf(d0, d1);
        float gy = gyf(a0, a1) * ig + qg * gyf(e0, e1) + aqg * gyf(d0, d1);
        float gz = gzf(a0, a1) * ig + qg * gzf(e0, e1) + aqg * gzf(d0, d1);
        fi.x += quant(kTermScale * gx);
        fi.y += quant(kTermScale * gy);
        fi.z += quant(kTermScale * gz);
    }
    return e;
}

// ---------------------------------------------------------------------------
// Intramolecular pair energy and dE/dr (calcenergy.cl:397-536,
// calcMergedEneGra.cl:455-583).
// ---------------------------------------------------------------------------
template <bool GRAD>
DEV float pair_energy(const LigandDev* L, const PairRec& pr, float ad, float& dedr) {
    float e = 0.f;
    dedr = 0.f;
    if (pr.flags & (1u << 16)) { e += 50.0f * ad; dedr += 50.0f; }   // CG-G0 bond term, G = 50
    if (ad < 8.0f) {
        float delta = pr.reqm - ad;
        float sd = (fabsf(delta) >= L->half_smooth) ? ad + copysignf(L->half_smooth, delta) : pr.reqm;
        int m = (int)(pr.flags & 0xffu), n = (int)((pr.flags >> 8) & 0xffu);
        float rm, rmn;
        if (pr.flags & (1u << 17)) {        // m = 12, n in {6, 10}
            float r2 = sd * sd, r4 = r2 * r2, r6 = r4 * r2;
            rm = fast_rcp(r6 * r6);
            rmn = (n == 6) ? r6 : r2;
        } else {                            // general exponents (modified pairs)
            float l2 = __builtin_amdgcn_logf(sd);
            rm = __builtin_amdgcn_exp2f(-(float)m * l2);
            rmn = __builtin_amdgcn_exp2f((float)(m - n) * l2);
        }
        e += (pr.ac - rmn * pr.bd) * rm;
        if (GRAD) dedr += ((float)n * pr.bd * rmn - (float)m * pr.ac) * rm * fast_rcp(sd);
    }
    if (ad < 20.48f) {
        float d = ad < L->elec_min_distance ? L->elec_min_distance : ad;
        float d2 = d * d;
        float desolv = pr.dsol * (L->coeff_desolv * (12.96f - 0.1063f * d2 * (1.0f - 0.001947f * d2)) *
                                  fast_rcp(12.96f + d2 * (0.4137f + d2 * (0.00357f + 0.000112f * d2))));
        float ds = d + 1.26366f;
        float rds2 = fast_rcp(ds * ds);
        float diel = 1.10859f * rds2 + 0.010358f;
        float rd = fast_rcp(d);
        float es = pr.cq * rd;
        e += diel * es + desolv;
        if (GRAD) dedr += -es * rd * diel - es * 2.21718f * rds2 * fast_rcp(ds) - 0.0771605f * d * desolv;
    }
    return e;
}

// ---------------------------------------------------------------------------
// One evaluation of one genotype by one wavefront.
//   GRAD = false: gpu_calc_energy (calcenergy.cl:132-554)
//   GRAD = true : gpu_calc_energrad (calcMergedEneGra.cl:46-936); lane g
//                 returns d E / d gene g in grad (0 for g >= genes).
// ---------------------------------------------------------------------------
template <int NC, int MF, bool GRAD>
DEV float evaluate(const LigandDev* L, const GridDev& G, const SharedView& sv, const WaveView& wv,
                   const Atoms<NC>& A, float gene, float& grad) {
    const int lane = lane_id();
    const int T = L->num_rotbonds, P = L->num_pairs;
    const bool has_lig = L->true_ligand_atoms != 0;
    Pose pose = pose_from_genes(gene, has_lig);
    float3 c[NC];
    compute_coords<NC>(L, sv, wv, A, pose, gene, c);
    if (GRAD) {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            wv.fx[lane + 64 * k] = 0; wv.fy[lane + 64 * k] = 0; wv.fz[lane + 64 * k] = 0;
        }
        wv.gi[lane] = 0;
    }
    wave_sync();
    float e = 0.f;
    int3 fi[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        fi[k] = make_int3(0, 0, 0);
        if (!A.ign[k]) e += inter_atom<MF, GRAD>(G, c[k], A.q[k], A.mapoff[k], fi[k]);
    }
    const float spacing = G.spacing;
    for (int p = lane; p < P; p += 64) {
        PairRec pr = sv.pairs[p];
        int i = (int)(pr.ij & 0xffffu), j = (int)(pr.ij >> 16);
        float4 ci = wv.coords[i], cj = wv.coords[j];
        float sx = ci.x - cj.x, sy = ci.y - cj.y, sz = ci.z - cj.z;
        float d2 = sx * sx + sy * sy + sz * sz;
        float dist = fast_sqrt(d2);
        float dedr;
        e += pair_energy<GRAD>(L, pr, dist * spacing, dedr);
        if (GRAD) {
            float gdd = -dedr * fast_rcp(dist);
            int ix = quant(kTermScale * sx * gdd), iy = quant(kTermScale * sy * gdd), iz = quant(kTermScale * sz * gdd);
            atomicSub(&wv.fx[i], ix); atomicSub(&wv.fy[i], iy); atomicSub(&wv.fz[i], iz);
            atomicAdd(&wv.fx[j], ix); atomicAdd(&wv.fy[j], iy); atomicAdd(&wv.fz[j], iz);
        }
    }
    const float energy = wave_sum(e);
    if (!GRAD) return energy;

    wave_sync();
    // total per-atom forces (kcal/mol/A), int fixed point -> float
    float3 F[NC];
    float sfx = 0.f, sfy = 0.f, sfz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        int a = lane + 64 * k;
        F[k] = make_float3(0.f, 0.f, 0.f);
        if (A.valid[k]) {
            int fx = fi[k].x + wv.fx[a], fy = fi[k].y + wv.fy[a], fz = fi[k].z + wv.fz[a];
            F[k] = make_float3(kOneOverTermScale * (float)fx, kOneOverTermScale * (float)fy,
                               kOneOverTermScale * (float)fz);
            sfx += F[k].x; sfy += F[k].y; sfz += F[k].z;
            if (A.lig[k]) {   // torque about the ligand centre, calcMergedEneGra.cl:712-745
                float rx = (c[k].x - pose.mv.x) * spacing, ry = (c[k].y - pose.mv.y) * spacing,
                      rz = (c[k].z - pose.mv.z) * spacing;
                tx += ry * F[k].z - rz * F[k].y;
                ty += rz * F[k].x - rx * F[k].z;
                tz += rx * F[k].y - ry * F[k].x;
            }
        }
    }
    sfx = wave_sum(sfx); sfy = wave_sum(sfy); sfz = wave_sum(sfz);
    tx = wave_sum(tx); ty = wave_sum(ty); tz = wave_sum(tz);
    int ig = 0;
    if (has_lig) {
        // translation genes, calcMergedEneGra.cl:679-697
        int it0 = quant(kTermScale * sfx * spacing), it1 = quant(kTermScale * sfy * spacing),
            it2 = quant(kTermScale * sfz * spacing);
        // rotation genes, calcMergedEneGra.cl:756-861
        float tl = fast_sqrt(tx * tx + ty * ty + tz * tz);
        float oscale = (tl < 1e-3f) ? 1.0f : tl * (1.0f / 1e-3f);
        float sc = (tl < 1e-3f) ? 0.5f - tl * tl * (1.0f / 48.0f) : 0.0005f * fast_rcp(tl);
        float4 qt = make_float4(tx * sc, ty * sc, tz * sc, (tl < 1e-3f) ? 1.0f - tl * tl * 0.125f : 0.99999988f);
        float4 tq = qmul(qt, pose.q);
        float cur_phi = fmod_pi2(kTwoPi + pose.phi);
        float cur_theta = fmod_pi2(kTwoPi + pose.theta);
        float cur_rot = fmod_pi2(kTwoPi + pose.rotangle);
        float tgt_rot = 2.0f * fast_acos(tq.w);
        float inv_sin = fast_rsq(1.0f - tq.w * tq.w);
        float tgt_theta = kTwoPi + pose.is_theta_gt_pi * fast_acos(tq.z * inv_sin);
        float tgt_phi = fmod_pi2(atan2f(pose.is_theta_gt_pi * tq.y, pose.is_theta_gt_pi * tq.x) + kTwoPi);
        float gphi = oscale * (fmod_pi2(tgt_phi - cur_phi + kPi) - kPi);
        float gtheta = oscale * (fmod_pi2(tgt_theta - cur_theta + kPi) - kPi);
        float grot = oscale * (fmod_pi2(tgt_rot - cur_rot + kPi) - kPi);
        float corr = 4.0f * pose.sin_half * pose.sin_half;
        int ir0 = quant(kTermScale * gphi * pose.sin_theta * pose.sin_theta * corr * kDegToRad);
        int ir1 = quant(kTermScale * gtheta * corr * kDegToRad);
        int ir2 = quant(kTermScale * grot * kDegToRad);
        ig = lane == 0 ? it0 : lane == 1 ? it1 : lane == 2 ? it2 : lane == 3 ? ir0 : lane == 4 ? ir1 : ir2;
    }
    // torsion genes, calcMergedEneGra.cl:866-915
    if (lane < T) {
        int2 ab = sv.tor_atoms[lane];
        float4 r0 = wv.coords[ab.x], r1 = wv.coords[ab.y];
        float ax = r1.x - r0.x, ay = r1.y - r0.y, az = r1.z - r0.z;
        float inv = fast_rsq(ax * ax + ay * ay + az * az);
        wv.tax[lane] = make_float4(ax * inv, ay * inv, az * inv, 0.f);
        wv.tref[lane] = r0;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        uint32_t lo = A.mlo[k], hi = A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            float4 ax = wv.tax[t], r0 = wv.tref[t];
            float rx = c[k].x - r0.x, ry = c[k].y - r0.y, rz = c[k].z - r0.z;
            float qx = ry * F[k].z - rz * F[k].y, qy = rz * F[k].x - rx * F[k].z, qz = rx * F[k].y - ry * F[k].x;
            float toa = (ax.x * qx + ax.y * qy + ax.z * qz) * spacing;
            atomicAdd(&wv.gi[6 + t], quant(kTermScale * toa * kDegToRad));
        }
    }
    wave_sync();
    if (lane >= 6) ig = wv.gi[lane];
    float g = kOneOverTermScale * (float)ig;
    if (lane >= 3) g *= spacing * spacing * 1.0f / (kDegToRad * kDegToRad);   // CONVERT_INTO_ANGSTROM_RADIAN
    grad = lane < L->num_genes ? g : 0.f;
    return energy;
}

// ---------------------------------------------------------------------------
// Float gradient with the exact Mehler-Solmajer / exponential desolvation
// derivatives, used once per FIRE search (calcgradient.cl:109-1003).  Forces
// are float; intra forces use LDS float atomics of one wavefront.
// ---------------------------------------------------------------------------
template <int NC, int MF>
DEV void float_gradient(const LigandDev* L, const GridDev& G, const SharedView& sv, const WaveView& wv,
                        const Atoms<NC>& A, float gene, float& grad) {
    const int lane = lane_id();
    const int T = L->num_rotbonds, P = L->num_pairs;
    const bool has_lig = L->true_ligand_atoms != 0;
    Pose pose = pose_from_genes(gene, has_lig);
    float3 c[NC];
    compute_coords<NC>(L, sv, wv, A, pose, gene, c);
    float* ffx = reinterpret_cast<float*>(wv.fx);
    float* ffy = reinterpret_cast<float*>(wv.fy);
    float* ffz = reinterpret_cast<float*>(wv.fz);
#pragma unroll
    for (int k = 0; k < NC; ++k) { ffx[lane + 64 * k] = 0.f; ffy[lane + 64 * k] = 0.f; ffz[lane + 64 * k] = 0.f; }
    wave_sync();
    float3 gin[NC];
    const uint32_t g3 = G.g3, ne = (uint32_t)(G.nmaps - 2) * g3, nd = ne + g3;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        gin[k] = make_float3(0.f, 0.f, 0.f);
        if (A.ign[k]) continue;
        float x = c[k].x, y = c[k].y, z = c[k].z;
        if ((x < 0.f) || (y < 0.f) || (z < 0.f) || (x >= (float)(G.gx - 1)) || (y >= (float)(G.gy - 1)) ||
            (z >= (float)(G.gz - 1))) {
            gin[k] = make_float3(16777216.0f, 16777216.0f, 16777216.0f);
            continue;
        }
        uint32_t xl = (uint32_t)floorf(x), yl = (uint32_t)floorf(y), zl = (uint32_t)floorf(z);
        uint32_t xh = (uint32_t)ceilf(x), yh = (uint32_t)ceilf(y), zh = (uint32_t)ceilf(z);
        float dx = x - (float)xl, dy = y - (float)yl, dz = z - (float)zl;
        uint32_t o[8] = {xl + yl * G.g1 + zl * G.g2, xl + yh * G.g1 + zl * G.g2, xl + yl * G.g1 + zh * G.g2,
                         xl + yh * G.g1 + zh * G.g2, xh + yl * G.g1 + zl * G.g2, xh + yh * G.g1 + zl * G.g2,
                         xh + yl * G.g1 + zh * G.g2, xh + yh * G.g1 + zh * G.g2};
        uint32_t mb[3] = {A.mapoff[k], ne, nd};
        float sc[3] = {1.0f, A.q[k], fabsf(A.q[k])};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float cb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) cb[i] = load_pt<MF>(G.data, mb[m] + o[i]).x;
            float vx0 = (1.0f - dy) * (cb[4] - cb[0]) + dy * (cb[5] - cb[1]);
            float vx1 = (1.0f - dy) * (cb[6] - cb[2]) + dy * (cb[7] - cb[3]);
            float vy0 = (1.0f - dx) * (cb[1] - cb[0]) + dx * (cb[5] - cb[4]);
            float vy1 = (1.0f - dx) * (cb[3] - cb[2]) + dx * (cb[7] - cb[6]);
            float vz0 = (1.0f - dx) * (cb[2] - cb[0]) + dx * (cb[6] - cb[4]);
            float vz1 = (1.0f - dx) * (cb[3] - cb[1]) + dx * (cb[7] - cb[5]);
            gin[k].x += sc[m] * ((1.0f - dz) * vx0 + dz * vx1);
            gin[k].y += sc[m] * ((1.0f - dz) * vy0 + dz * vy1);
            gin[k].z += sc[m] * ((1.0f - dy) * vz0 + dy * vz1);
        }
    }
    const float spacing = G.spacing;
    const float half_smooth = L->half_smooth;
    const float DA = -8.5525f, DB = 78.4f - (-8.5525f), DK = 7.7839f;
    const float DBH = DB * 0.003627f, DBHK = DBH * DK;
    for (int p = lane; p < P; p += 64) {
        PairRec pr = sv.pairs[p];
        int i = (int)(pr.ij & 0xffffu), j = (int)(pr.ij >> 16);
        float4 ci = wv.coords[i], cj = wv.coords[j];
        float sx = ci.x - cj.x, sy = ci.y - cj.y, sz = ci.z - cj.z;
        float dist = sqrtf(sx * sx + sy * sy + sz * sz);
        float d = dist * spacing;
        float sd;
        if (d <= pr.reqm - half_smooth) sd = d + half_smooth;
        else if (d < pr.reqm + half_smooth) sd = pr.reqm;
        else sd = d - half_smooth;
        int m = (int)(pr.flags & 0xffu), n = (int)((pr.flags >> 8) & 0xffu);
        float pg = 0.f;
        if (d < 8.0f) {
            pg += (-(float)m * pr.ac) / powf(sd, (float)(m + 1));
            pg += ((float)n * pr.bd) / powf(sd, (float)(n + 1));
        }
        if (d < 20.48f) {
            if (d < L->elec_min_distance) d = L->elec_min_distance;
            float ex = __expf(DBH * d);
            float upper = DA * (ex + DK) * (ex + DK) + DB * ex * (DBHK * d + ex + DK);
            float lw = DA * (ex + DK) + DB * ex;
            float lower = d * d * lw * lw;
            pg += -pr.cq * (upper / lower);
            pg += pr.dsol * L->coeff_desolv * -0.077160f * d * __expf(-0.038580f * d * d);
        }
        if (pr.flags & (1u << 16)) pg += 50.0f;
        float fx = pg * (-sx / dist), fy = pg * (-sy / dist), fz = pg * (-sz / dist);
        atomicAdd(&ffx[i], -fx); atomicAdd(&ffy[i], -fy); atomicAdd(&ffz[i], -fz);
        atomicAdd(&ffx[j], fx); atomicAdd(&ffy[j], fy); atomicAdd(&ffz[j], fz);
    }
    wave_sync();
    float3 F[NC];
    float sfx = 0.f, sfy = 0.f, sfz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
    const float g0 = bcast(gene, 0), g1 = bcast(gene, 1), g2 = bcast(gene, 2);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        int a = lane + 64 * k;
        F[k] = make_float3(0.f, 0.f, 0.f);
        if (A.valid[k]) {
            F[k] = make_float3(gin[k].x / spacing + ffx[a], gin[k].y / spacing + ffy[a], gin[k].z / spacing + ffz[a]);
            if (A.lig[k]) {
                sfx += F[k].x; sfy += F[k].y; sfz += F[k].z;
                float rx = (c[k].x - g0) * spacing, ry = (c[k].y - g1) * spacing, rz = (c[k].z - g2) * spacing;
                tx += ry * F[k].z - rz * F[k].y;
                ty += rz * F[k].x - rx * F[k].z;
                tz += rx * F[k].y - ry * F[k].x;
            }
        }
    }
    sfx = wave_sum(sfx); sfy = wave_sum(sfy); sfz = wave_sum(sfz);
    tx = wave_sum(tx); ty = wave_sum(ty); tz = wave_sum(tz);
    float gval = 0.f;
    if (lane < 3) gval = (lane == 0 ? sfx : lane == 1 ? sfy : sfz) * spacing;
    if (has_lig && lane >= 3 && lane < 6) {   // calcgradient.cl:714-873 (fmod)
        float tl = sqrtf(tx * tx + ty * ty + tz * tz);
        float oscale = (tl < 1e-3f) ? 1.0f : tl * (1.0f / 1e-3f);
        float sc = (tl < 1e-3f) ? 0.5f - tl * tl / 48.0f : 0.0005f / tl;
        float4 qt = make_float4(tx * sc, ty * sc, tz * sc, (tl < 1e-3f) ? 1.0f - tl * tl * 0.125f : 0.99999988f);
        float4 tq = qmul(qt, pose.q);
        float cur_phi = fmod_pi2(kTwoPi + pose.phi);
        float cur_theta = fmod_pi2(kTwoPi + pose.theta);
        float cur_rot = fmod_pi2(kTwoPi + pose.rotangle);
        float tgt_rot = 2.0f * fast_acos(tq.w);
        float inv_sin = rsqrtf(1.0f - tq.w * tq.w);
        float tgt_theta = kTwoPi + pose.is_theta_gt_pi * fast_acos(tq.z * inv_sin);
        float tgt_phi = fmod_pi2(atan2f(pose.is_theta_gt_pi * tq.y, pose.is_theta_gt_pi * tq.x) + kTwoPi);
        float corr = 4.0f * pose.sin_half * pose.sin_half;
        if (lane == 3) gval = oscale * (fmodf(tgt_phi - cur_phi + kPi, kTwoPi) - kPi) * pose.sin_theta * pose.sin_theta * corr * kDegToRad;
        if (lane == 4) gval = oscale * (fmodf(tgt_theta - cur_theta + kPi, kTwoPi) - kPi) * corr * kDegToRad;
        if (lane == 5) gval = oscale * (fmodf(tgt_rot - cur_rot + kPi, kTwoPi) - kPi) * kDegToRad;
    }
    if (lane < T) {
        int2 ab = sv.tor_atoms[lane];
        float4 r0 = wv.coords[ab.x], r1 = wv.coords[ab.y];
        float ax = r1.x - r0.x, ay = r1.y - r0.y, az = r1.z - r0.z;
        float inv = rsqrtf(ax * ax + ay * ay + az * az);
        wv.tax[lane] = make_float4(ax * inv, ay * inv, az * inv, 0.f);
        wv.tref[lane] = r0;
    }
    float* tacc = reinterpret_cast<float*>(wv.gi);
    tacc[lane] = 0.f;
    wave_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        uint32_t lo = A.mlo[k], hi = A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            float4 ax = wv.tax[t], r0 = wv.tref[t];
            float rx = (c[k].x - r0.x) * spacing, ry = (c[k].y - r0.y) * spacing, rz = (c[k].z - r0.z) * spacing;
            float qx = ry * F[k].z - rz * F[k].y, qy = rz * F[k].x - rx * F[k].z, qz = rx * F[k].y - ry * F[k].x;
            atomicAdd(&tacc[6 + t], (ax.x * qx + ax.y * qy + ax.z * qz) * kDegToRad);
        }
    }
    wave_sync();
    if (lane >= 6) gval = tacc[lane];
    if (lane > 2) gval *= spacing * spacing * 1.0f / (kDegToRad * kDegToRad);
    grad = lane < L->num_genes ? gval : 0.f;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct Launch {
    const LigandDev* L;
    GridDev G;
    SearchDev S;
    JobBuffers B;
    const float* in_geno;
    float* out_e;
    float* out_g;
    int n;
    int mode;
};

#define KERNEL_PROLOGUE(PL)                                                   \
    extern __shared__ __align__(16) char lds[];                               \
    const int wave = threadIdx.x >> 6;                                        \
    const int lane = lane_id();                                               \
    SharedView sv;                                                            \
    WaveView wv;                                                              \
    stage<PL>(a.L, lds, sv, wv, wave);                                        \
    const int idx = blockIdx.x * kWavesPerGroup + wave;                       \
    Atoms<NC> A;                                                              \
    A.load(a.L, a.G);

// Batched pose evaluation: the per-pose parity entry point.
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_eval_poses(Launch a) {
    KERNEL_PROLOGUE(PL)
    if (idx >= a.n) return;
    float gene = a.in_geno[(size_t)idx * kGenes + lane];
    float grad = 0.f;
    if (a.mode == ADGPU_EVAL_ENERGY) {
        float e = evaluate<NC, MF, false>(a.L, a.G, sv, wv, A, gene, grad);
        if (lane == 0 && a.out_e) a.out_e[idx] = e;
    } else if (a.mode == ADGPU_EVAL_ENERGRAD) {
        float e = evaluate<NC, MF, true>(a.L, a.G, sv, wv, A, gene, grad);
        if (lane == 0 && a.out_e) a.out_e[idx] = e;
        if (a.out_g) a.out_g[(size_t)idx * kGenes + lane] = grad;
    } else {
        float_gradient<NC, MF>(a.L, a.G, sv, wv, A, gene, grad);
        if (a.out_g) a.out_g[(size_t)idx * kGenes + lane] = grad;
    }
}

// K1: energies of the initial population (kernel1.cl:28-145).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_initpop(Launch a) {
    KERNEL_PROLOGUE(PL)
    const int ne = a.S.num_runs * a.S.pop_size;
    if (idx >= ne) return;
    float gene = a.B.pop_cur[(size_t)idx * kGenes + lane];
    float g;
    float e = evaluate<NC, MF, false>(a.L, a.G, sv, wv, A, gene, g);
    if (lane == 0) { a.B.ene_cur[idx] = e; a.B.evals_new[idx] = 1; }
}

// K2: per-run eval totals (kernel2.cl:26-62); one wave per run.
__global__ __launch_bounds__(64) void k_sum_evals(SearchDev S, JobBuffers B) {
    const int run = blockIdx.x;
    const int lane = lane_id();
    int s = 0;
    for (int e = lane; e < S.pop_size; e += 64) s += B.evals_new[run * S.pop_size + e];
    s = wave_sum_i(s);
    if (lane == 0) B.evals_runs[run] += s;
}

// K4: elitism, tournament selection, crossover, mutation and offspring
// energies (kernel4.cl:32-336, auxiliary_genetic.cl:89-162).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_gen(Launch a) {
    KERNEL_PROLOGUE(PL)
    const int pop = a.S.pop_size, NG = a.S.num_genes;
    const int ne = a.S.num_runs * pop;
    if (idx >= ne) return;
    const int run = idx / pop, ent = idx - run * pop;
    const float* cur = a.B.pop_cur + (size_t)run * pop * kGenes;
    const float* ecur = a.B.ene_cur + (size_t)run * pop;
    if (ent == 0) {
        // per lane: first strict minimum over the lane's stride; then the
        // lowest lane among equal minima (the reference's serial scan)
        float be = INFINITY;
        int bi = -1;
        for (int k = lane; k < pop; k += 64) {
            float ek = ecur[k];
            if (bi < 0 || ek < be) { be = ek; bi = k; }
        }
        int bl = bi < 0 ? 64 : lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            float oe = __shfl_xor(be, o);
            int ob = __shfl_xor(bi, o), ol = __shfl_xor(bl, o);
            bool take = (ol < 64) && ((bl >= 64) || (oe < be) || (!(be < oe) && !(oe < be) && ol < bl));
            if (take) { be = oe; bi = ob; bl = ol; }
        }
        float gv = cur[(size_t)bi * kGenes + lane];
        if (lane < NG) a.B.pop_next[(size_t)idx * kGenes + lane] = gv;
        if (lane == 0) { a.B.ene_next[idx] = be; a.B.evals_new[idx] = 0; }
        return;
    }
    uint32_t* st = a.B.prng + (size_t)idx * 64 + lane;
    uint32_t state = *st;
    float rn = 0.f;
    if (lane < 10) rn = randf(&state);
    // parent candidates (lanes 0..3), binary tournaments
    int cand = (int)((float)pop * rn);
    float ce = (lane < 4) ? ecur[cand] : 0.f;
    int c0 = __shfl(cand, 0), c1 = __shfl(cand, 1), c2 = __shfl(cand, 2), c3 = __shfl(cand, 3);
    float e0 = bcast(ce, 0), e1 = bcast(ce, 1), e2 = bcast(ce, 2), e3 = bcast(ce, 3);
    float r4 = bcast(rn, 4), r5 = bcast(rn, 5), r6 = bcast(rn, 6), r7 = bcast(rn, 7), r8 = bcast(rn, 8);
    const float tr = a.S.tournament_rate;
    int p0 = (e0 < e1) ? ((r4 < tr) ? c0 : c1) : ((r4 < tr) ? c1 : c0);
    int p1 = (e2 < e3) ? ((r5 < tr) ? c2 : c3) : ((r5 < tr) ? c3 : c2);
    float v0 = cur[(size_t)p0 * kGenes + lane], v1 = cur[(size_t)p1 * kGenes + lane];
    float child;
    if (r6 < a.S.crossover_rate) {
        int cp0 = (int)((float)(NG - 1) * r7), cp1 = (int)((float)(NG - 1) * r8);
        if (cp1 < cp0) { int t = cp1; cp1 = cp0; cp0 = t; }
        if (cp0 != cp1) child = ((lane <= cp0) || (lane > cp1)) ? v0 : v1;
        else child = (lane <= cp0) ? v0 : v1;
    } else {
        child = v0;
    }
    if (lane < NG) {
        // separate rounding of every operation, as the reference source
        // writes it (no fused multiply-add): the new genes are bit-exact
#pragma clang fp contract(off)
        if (randf(&state) < a.S.mutation_rate) {
            if (lane < 3) child += a.S.abs_max_dmov * (2 * randf(&state) - 1);
            else { child += a.S.abs_max_dang * (2 * randf(&state) - 1); map_angle(child); }
        }
    }
    *st = state;
    float g;
    float e = evaluate<NC, MF, false>(a.L, a.G, sv, wv, A, lane < NG ? child : 0.f, g);
    if (lane < NG) a.B.pop_next[(size_t)idx * kGenes + lane] = child;
    if (lane == 0) { a.B.ene_next[idx] = e; a.B.evals_new[idx] = 1; }
}

// Entity choice of every LS kernel (e.g. kernel3.cl:104-117).
DEV int ls_entity(const SearchDev& S, int le, uint32_t* st0, int lane) {
    int ent = le;
    if (le == 0) {
        float r = 0.f;
        if (lane == 0) { uint32_t s = *st0; r = randf(&s); *st0 = s; }
        r = bcast(r, 0);
        if (100.0f * r > S.lsearch_rate) ent = S.num_ls;
    }
    // the reference reads entity num_ls even when num_ls == pop_size (the next
    // run's first entity, or past the buffer for the last run); clamp instead
    return ent < S.pop_size ? ent : S.pop_size - 1;
}

// Solis-Wets with SWAT3 (kernel3.cl:30-436).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_ls_sw(Launch a, float lig_scale, float gene_scale) {
    KERNEL_PROLOGUE(PL)
    const int nls = a.S.num_ls, pop = a.S.pop_size, NG = a.S.num_genes;
    if (idx >= a.S.num_runs * nls) return;
    const int run = idx / nls, le = idx - run * nls;
    uint32_t* st = a.B.prng + (size_t)idx * 64 + lane;
    const int ent = ls_entity(a.S, le, st, lane);
    const size_t gi = (size_t)run * pop + ent;
    uint32_t state = *st;
    const int gene_start = (a.L->true_ligand_atoms == 0) * 6;
    const bool active = lane >= gene_start && lane < NG;
    float off = a.B.pop_next[gi * kGenes + lane];
    float off_e = a.B.ene_next[gi];
    float bias = 0.f, dev = 0.f, cand = off;
    float rho = 1.0f;
    int cons_succ = 0, cons_fail = 0, iters = 0, evals = 0;
    const float sdmov = a.S.base_dmov_mul_sqrt3, sdang_l = a.S.base_dang_mul_sqrt3 * lig_scale,
                sdang_g = a.S.base_dang_mul_sqrt3 * gene_scale;
    float g;
    while (iters < a.S.max_iters && rho > a.S.rho_lower_bound) {
        if (active) {
            float r1 = randf(&state);
            float r2 = randf(&state);
            dev = rho * (2.0f * r1 - 1.0f) * (float)(r2 < gene_scale);
            dev *= (lane < 3) ? sdmov : ((lane < 6) ? sdang_l : sdang_g);
            cand = off + dev + bias;
        }
        float ce = evaluate<NC, MF, false>(a.L, a.G, sv, wv, A, cand, g);
        evals++;
        if (ce < off_e) {
            if (active) { off = cand; bias = 0.6f * bias + 0.4f * dev; }
            off_e = ce; cons_succ++; cons_fail = 0;
        } else {
            if (active) cand = off - dev - bias;
            ce = evaluate<NC, MF, false>(a.L, a.G, sv, wv, A, cand, g);
            evals++;
            if (ce < off_e) {
                if (active) { off = cand; bias = 0.6f * bias - 0.4f * dev; }
                off_e = ce; cons_succ++; cons_fail = 0;
            } else {
                if (active) bias = 0.5f * bias;
                cons_succ = 0; cons_fail++;
            }
        }
        iters++;
        if (cons_succ >= a.S.cons_limit) { rho *= 2.0f; cons_succ = 0; }
        else if (cons_fail >= a.S.cons_limit) { rho *= 0.5f; cons_fail = 0; }
    }
    *st = state;
    if (lane >= 3) map_angle(off);
    if (lane < NG) a.B.pop_next[gi * kGenes + lane] = off;
    if (lane == 0) { a.B.evals_new[gi] += evals; a.B.ene_next[gi] = off_e; }
}

// ADADELTA (kernel_ad.cl:60-549, ADADELTA_AUTOSTOP).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_ls_ad(Launch a) {
    KERNEL_PROLOGUE(PL)
    const int nls = a.S.num_ls, pop = a.S.pop_size, NG = a.S.num_genes;
    if (idx >= a.S.num_runs * nls) return;
    const int run = idx / nls, le = idx - run * nls;
    uint32_t* st = a.B.prng + (size_t)idx * 64 + lane;
    const int ent = ls_entity(a.S, le, st, lane);
    const size_t gi = (size_t)run * pop + ent;
    float geno = a.B.pop_next[gi * kGenes + lane];
    if (lane >= NG) geno = 0.f;
    float best = geno, sqg = 0.f, sqd = 0.f;
    float best_e = INFINITY, varrho = 1.0f;
    int cons_succ = 0, cons_fail = 0, iters = 0;
    const float RHO = 0.8f, EPS = 1e-2f;
    do {
        float g;
        float e = evaluate<NC, MF, true>(a.L, a.G, sv, wv, A, geno, g);
        if (e < best_e) best = geno;
        sqg *= RHO;
        sqg += (1.0f - RHO) * g * g;
        float delta = -1.0f * g * sqrtf((sqd + EPS) / (sqg + EPS));
        sqd *= RHO;
        sqd += (1.0f - RHO) * delta * delta;
        geno += delta;
        if (e < best_e) { best_e = e; cons_succ++; cons_fail = 0; }
        else { cons_succ = 0; cons_fail++; }
        iters++;
        if (cons_succ >= 4) { varrho *= 2.0f; cons_succ = 0; }
        else if (cons_fail >= 4) { varrho *= 0.5f; cons_fail = 0; }
    } while (iters < a.S.max_iters && varrho > 0.01f);
    if (lane >= 3) map_angle(best);
    if (lane < NG) a.B.pop_next[gi * kGenes + lane] = best;
    if (lane == 0) { a.B.evals_new[gi] += iters; a.B.ene_next[gi] = best_e; }
}

// ADAM (cuda/kernel_adam.cu:53-416; AD_RHO_CRITERION off).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_ls_adam(Launch a) {
    KERNEL_PROLOGUE(PL)
    const int nls = a.S.num_ls, pop = a.S.pop_size, NG = a.S.num_genes;
    if (idx >= a.S.num_runs * nls) return;
    const int run = idx / nls, le = idx - run * nls;
    uint32_t* st = a.B.prng + (size_t)idx * 64 + lane;
    const int ent = ls_entity(a.S, le, st, lane);
    const size_t gi = (size_t)run * pop + ent;
    float geno = a.B.pop_next[gi * kGenes + lane];
    if (lane >= NG) geno = 0.f;
    float best = geno, mt = 0.f, vt = 0.f, best_e = INFINITY;
    const float b1 = a.S.adam_beta1, b2 = a.S.adam_beta2, eps = a.S.adam_epsilon;
    int iters = 0;
    do {
        float g;
        float e = evaluate<NC, MF, true>(a.L, a.G, sv, wv, A, geno, g);
        float b1p = 1.0f - powf(b1, 1.0f + (float)iters);
        float b2p = 1.0f - powf(b2, 1.0f + (float)iters);
        if (e < best_e) best = geno;
        mt = b1 * mt + (1.0f - b1) * g;
        vt = b2 * vt + (1.0f - b2) * g * g;
        float mp = mt / b1p, vp = vt / b2p;
        geno -= mp / (sqrtf(vp) + eps);
        iters++;
        if (e < best_e) best_e = e;
    } while (iters < a.S.max_iters);
    if (lane >= 3) map_angle(best);
    if (lane < NG) a.B.pop_next[gi * kGenes + lane] = best;
    if (lane == 0) { a.B.evals_new[gi] += iters; a.B.ene_next[gi] = best_e; }
}

// FIRE (kernel_fire.cl:75-895).
template <int NC, int MF, bool PL>
__global__ __launch_bounds__(256) void k_ls_fire(Launch a) {
    KERNEL_PROLOGUE(PL)
    const int nls = a.S.num_ls, pop = a.S.pop_size, NG = a.S.num_genes;
    if (idx >= a.S.num_runs * nls) return;
    const int run = idx / nls, le = idx - run * nls;
    uint32_t* st = a.B.prng + (size_t)idx * 64 + lane;
    const int ent = ls_entity(a.S, le, st, lane);
    const size_t gi = (size_t)run * pop + ent;
    float geno = a.B.pop_next[gi * kGenes + lane];
    if (lane >= NG) geno = 0.f;
    float energy = a.B.ene_next[gi];
    float grad;
    float_gradient<NC, MF>(a.L, a.G, sv, wv, A, geno, grad);
    float alpha = 0.2f, dt = 10.0f / 3.0f;
    int count_success = 0, iters = 0;
    float inv_gn = 0.2f * (1.0f / sqrtf(wave_sum(grad * grad)));
    float vel = -grad * inv_gn;
    float best_e = energy, best = geno;
    do {
        float cand = geno + dt * vel;
        float cg;
        float ce = evaluate<NC, MF, true>(a.L, a.G, sv, wv, A, cand, cg);
        float power = wave_sum(-cg * vel);
        if (power < 0.0f) {
            inv_gn = 0.2f * (1.0f / sqrtf(wave_sum(grad * grad)));
            vel = -grad * inv_gn;
            count_success = 0; alpha = 0.2f; dt = dt * 0.8f;
        } else {
            count_success++;
            if (count_success > 5) { dt = fminf(dt * 1.2f, 10.0f); alpha = alpha * 0.99f; }
        }
        energy = ce;
        geno = cand;
        grad = cg;
        float ign = 1.0f / sqrtf(wave_sum(grad * grad));
        float vn = sqrtf(wave_sum(vel * vel));
        float vdg = alpha * vn * ign;
        vel = (1 - alpha) * vel - grad * vdg;
        iters++;
        if (energy < best_e) { best_e = energy; best = geno; }
    } while (iters < a.S.max_iters && dt > 1e-6f);
    if (lane >= 3) map_angle(best);
    if (lane < NG) a.B.pop_next[gi * kGenes + lane] = best;
    if (lane == 0) { a.B.evals_new[gi] += iters; a.B.ene_next[gi] = best_e; }
}

// ---------------------------------------------------------------------------
// Launchers: pick the (atom chunks, map format, pair staging) instance.
// ---------------------------------------------------------------------------
constexpr size_t kLdsBudget = 64 * 1024;

static bool pairs_fit(const LigandDev& L) {
    return shared_bytes(L, true) <= kLdsBudget;
}

#define DEFINE_LAUNCH(FN, KNAME, EXTRA_PARAMS, ...)                                                   \
    template <int NC, int MF>                                                                         \
    static hipError_t FN##_nc_mf(bool pl, dim3 grid, size_t sh, hipStream_t s,                       \
                                 const Launch& a EXTRA_PARAMS) {                                      \
        if (pl) hipLaunchKernelGGL((KNAME<NC, MF, true>), grid, dim3(256), sh, s, a __VA_ARGS__);     \
        else hipLaunchKernelGGL((KNAME<NC, MF, false>), grid, dim3(256), sh, s, a __VA_ARGS__);       \
        return hipGetLastError();                                                                     \
    }                                                                                                 \
    template <int NC>                                                                                 \
    static hipError_t FN##_nc(int mf, bool pl, dim3 grid, size_t sh, hipStream_t s,                  \
                              const Launch& a EXTRA_PARAMS) {                                         \
        if (mf == ADGPU_MAP_F16) return FN##_nc_mf<NC, ADGPU_MAP_F16>(pl, grid, sh, s, a __VA_ARGS__);\
        if (mf == ADGPU_MAP_BF16) return FN##_nc_mf<NC, ADGPU_MAP_BF16>(pl, grid, sh, s, a __VA_ARGS__);\
        return FN##_nc_mf<NC, ADGPU_MAP_F32>(pl, grid, sh, s, a __VA_ARGS__);                        \
    }                                                                                                 \
    static hipError_t FN(const LigandDev& Lh, int nwork, hipStream_t s, const Launch& a EXTRA_PARAMS) { \
        if (nwork <= 0) return hipSuccess;                                                            \
        bool pl = pairs_fit(Lh);                                                                      \
        size_t sh = shared_bytes(Lh, pl);                                                             \
        dim3 grid((nwork + kWavesPerGroup - 1) / kWavesPerGroup);                                     \
        int mf = (int)a.G.format;                                                                     \
        if (Lh.nchunk == 1) return FN##_nc<1>(mf, pl, grid, sh, s, a __VA_ARGS__);                   \
        if (Lh.nchunk == 2) return FN##_nc<2>(mf, pl, grid, sh, s, a __VA_ARGS__);                   \
        return FN##_nc<4>(mf, pl, grid, sh, s, a __VA_ARGS__);                                        \
    }

#define COMMA ,
DEFINE_LAUNCH(run_eval, k_eval_poses, )
DEFINE_LAUNCH(run_initpop, k_initpop, )
DEFINE_LAUNCH(run_gen, k_gen, )
DEFINE_LAUNCH(run_sw, k_ls_sw, COMMA float ls_ COMMA float gs_, COMMA ls_ COMMA gs_)
DEFINE_LAUNCH(run_ad, k_ls_ad, )
DEFINE_LAUNCH(run_adam, k_ls_adam, )
DEFINE_LAUNCH(run_fire, k_ls_fire, )

hipError_t launch_eval_poses(hipStream_t s, const LigandDev* lig_dev, const LigandDev& Lh, const GridDev& grid,
                             int mode, const float* genotypes, int n, float* energies, float* grads) {
    Launch a{};
    a.L = lig_dev; a.G = grid; a.in_geno = genotypes; a.out_e = energies; a.out_g = grads; a.n = n; a.mode = mode;
    return run_eval(Lh, n, s, a);
}

hipError_t launch_initpop(hipStream_t s, const LigandDev* lig_dev, const LigandDev& Lh, const GridDev& grid,
                          const SearchDev& sp, const JobBuffers& b) {
    Launch a{};
    a.L = lig_dev; a.G = grid; a.S = sp; a.B = b;
    return run_initpop(Lh, sp.num_runs * sp.pop_size, s, a);
}

hipError_t launch_gen(hipStream_t s, const LigandDev* lig_dev, const LigandDev& Lh, const GridDev& grid,
                      const SearchDev& sp, const JobBuffers& b) {
    Launch a{};
    a.L = lig_dev; a.G = grid; a.S = sp; a.B = b;
    return run_gen(Lh, sp.num_runs * sp.pop_size, s, a);
}

hipError_t launch_ls(hipStream_t s, const LigandDev* lig_dev, const LigandDev& Lh, const GridDev& grid,
                     const SearchDev& sp, const JobBuffers& b, int method) {
    Launch a{};
    a.L = lig_dev; a.G = grid; a.S = sp; a.B = b;
    const int n = sp.num_runs * sp.num_ls;
    switch (method) {
        case ADGPU_LS_SW: {
            float lig_scale = 1.0f / sqrtf((float)Lh.num_atoms);
            float gene_scale = 1.0f / sqrtf((float)Lh.num_genes);
            return run_sw(Lh, n, s, a, lig_scale, gene_scale);
        }
        case ADGPU_LS_AD: return run_ad(Lh, n, s, a);
        case ADGPU_LS_ADAM: return run_adam(Lh, n, s, a);
        case ADGPU_LS_FIRE: return run_fire(Lh, n, s, a);
        default: return hipSuccess;
    }
}

hipError_t launch_sum_evals(hipStream_t s, const SearchDev& sp, const JobBuffers& b) {
    hipLaunchKernelGGL(k_sum_evals, dim3(sp.num_runs), dim3(64), 0, s, sp, b);
    return hipGetLastError();
}

}  // namespace adgpu
