// adgpu_device.h -- gfx950 device code of the docking hot path (included by
// the kernel translation units k_*.hip; each instantiates its own kernels).
//
// Execution model (CDNA, 64-wide wavefronts):
//   * a "team" of TW lanes (TW = 32 or 64) evaluates one individual, so a
//     wavefront carries 64/TW individuals; a 256-thread workgroup stages the
//     ligand tables (pair records, torsion vectors) into LDS once and every
//     team of the workgroup shares them;
//   * team lane l owns atoms l, l+TW, ... (NC chunks) and walks each atom's
//     own torsion chain, so pose -> coordinates needs no barrier (the
//     reference's rotation list is a barrier per cycle, calcenergy.cl:232-275);
//   * team lane l owns genes l and l+TW (when TW = 32) for a whole local
//     search, so minimiser updates are register code; each gene lane also
//     evaluates the sine/cosine of its own angle once per evaluation;
//   * pairs are strided over the team; pair forces go to per-atom int32
//     fixed-point accumulators in LDS through ds_add (order independent);
//   * energies, forces and torques are summed with DPP row reductions plus
//     readlane, which leaves identical bits in every lane of the team;
//   * local searches run as persistent kernels: a team that finishes its
//     entity takes the next one from a global counter, so teams of one
//     wavefront never wait for each other and uneven searches balance out.
// The arithmetic of every device function follows the reference file:line
// given at its definition (ccsb-scripps/AutoDock-GPU v1.5, device/ = OpenCL).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "adgpu_internal.h"

namespace adgpu {

#define DEV __device__ __forceinline__
// 256-thread workgroups; at least 4 wavefronts per SIMD (<= 128 VGPRs)
// Waves per SIMD the register budget is sized for (4: at most 128 VGPRs).
#ifndef ADGPU_WAVES_PER_EU
#define ADGPU_WAVES_PER_EU 4
#endif
#define ADGPU_GROUP_KERNEL __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ADGPU_WAVES_PER_EU)))

// common/calcenergy_basic.h
constexpr float kDegToRad = 0.0174533f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kTermScale = 1024.0f;
constexpr float kOneOverTermScale = 1.0f / 1024.0f;
constexpr float kMaxTerm = 8192.0f;

// ---------------------------------------------------------------------------
// wavefront / team primitives
// ---------------------------------------------------------------------------
DEV int lane_id() { return (int)__lane_id(); }

// Orders LDS accesses of the lanes of one wavefront (no workgroup barrier).
DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL>
DEV float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
DEV float rdl(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
DEV int rdl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <int TW>
struct Team {
    static_assert(TW == 16 || TW == 32 || TW == 64, "team width");
    static constexpr int kPerWave = 64 / TW;
    static constexpr int kGPL = 64 / TW;          // genes per lane
    DEV static int lane() { return lane_id() & (TW - 1); }
    DEV static int slot() { return lane_id() / TW; }
    // Sum over the team; every lane of the team receives the same bits.
    // Rows of 16 lanes: xor 1, xor 2 (quad_perm), half-row mirror, row mirror;
    // then the row sums are combined with gfx950's row swaps: permlane16_swap
    // gives rows 2k and 2k+1 both (row 2k, row 2k+1), so both add
    // row 2k + row 2k+1 in that order; permlane32_swap does the same for the
    // two halves of the wavefront.
    // A 16-lane team is one row: the four DPP steps give its sum.
    DEV static float sum(float v) {
        v += dpp<0xB1>(v);
        v += dpp<0x4E>(v);
        v += dpp<0x141>(v);
        v += dpp<0x140>(v);
        if constexpr (TW == 16) return v;
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        if constexpr (TW == 64) {
            const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
            v = __uint_as_float(h[0]) + __uint_as_float(h[1]);
        }
        return v;
    }
    // Value of team lane k (k uniform).
    DEV static float bcast(float v, int k) {
        if constexpr (TW == 16) {
            const float a = rdl(v, k), b = rdl(v, 16 + k), c = rdl(v, 32 + k), d = rdl(v, 48 + k);
            const int s = lane_id() >> 4;
            return s == 0 ? a : s == 1 ? b : s == 2 ? c : d;
        } else if constexpr (TW == 32) {
            return lane_id() < 32 ? rdl(v, k) : rdl(v, 32 + k);
        } else {
            return rdl(v, k);
        }
    }
    DEV static int bcast_i(int v, int k) {
        if constexpr (TW == 16) {
            const int a = rdl_i(v, k), b = rdl_i(v, 16 + k), c = rdl_i(v, 32 + k), d = rdl_i(v, 48 + k);
            const int s = lane_id() >> 4;
            return s == 0 ? a : s == 1 ? b : s == 2 ? c : d;
        } else if constexpr (TW == 32) {
            return lane_id() < 32 ? rdl_i(v, k) : rdl_i(v, 32 + k);
        } else {
            return rdl_i(v, k);
        }
    }
    // True when any lane of the wavefront has p set.
    DEV static bool any_in_wave(bool p) { return __ballot(p) != 0ull; }
};

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
DEV float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
DEV float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DEV float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// LDS atomic subtract without a returned value (ds_sub_u32; HIP's atomicSub
// would negate and add).
DEV void lds_sub(int* p, int v) {
    (void)__hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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
// v - copysign(0.5, v) equals the reference's v + 0.5 * (v < 0 ? 1 : v > 0 ? -1 : 0)
// for every v, including +-0.
DEV int quant(float x) {
    const float v = fminf(kMaxTerm, fmaxf(-kMaxTerm, x));
    return (int)(v - copysignf(0.5f, v));
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

// LCG, auxiliary_genetic.cl:29-67
DEV float randf(uint32_t& s) {
    s = 1103515245u * s + 12345u;
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
    const PairA* pa;          // LDS (or global when the pairs do not fit)
    const PairB* pb;
    const float4* tor_unit;
    const float4* tor_move;
    const int2* tor_atoms;
};

struct TeamView {             // private to one team
    float* cx;                // [TW*NC] atom coordinates (grid units), one
    float* cy;                //   array per axis: lanes reading different
    float* cz;                //   atoms hit different banks
    int* fx;                  // [TW*NC] int32 fixed-point pair forces
    int* fy;
    int* fz;
    float4* tq;               // [T] torsion quaternions
    float4* tax;              // [T] torsion axes (gradient)
    float4* tref;             // [T] torsion reference atoms (gradient)
    int* gi;                  // [T] torsion gene gradients (int fixed point)
    float* pg;                // [12] genes 0..5, sin/cos of phi, theta, rot/2
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t shared_bytes_for(int T, int P, bool pairs_in_lds) {
    size_t s = pairs_in_lds ? (size_t)P * (sizeof(PairA) + sizeof(PairB)) : 0;
    s += (size_t)T * (sizeof(float4) * 2 + sizeof(int2));
    return align16(s);
}

__host__ __device__ inline size_t team_bytes_for(int natoms_pad, int T) {
    return 6 * align16((size_t)natoms_pad * 4) + (size_t)T * 48 + align16((size_t)T * 4) + 48;
}

// LDS of one workgroup of `threads` threads (threads / TW teams).
__host__ __device__ inline size_t lds_bytes_for(int TW, int NC, int T, int P, bool pairs_in_lds, int threads = 256) {
    return shared_bytes_for(T, P, pairs_in_lds) + (size_t)(threads / TW) * team_bytes_for(TW * NC, T);
}

// Stages the shared tables (all threads of the workgroup) and carves the
// team's region.  Contains the only workgroup barrier of every kernel.
template <int TW, int NC, bool PL>
DEV void stage(const LigandDev* L, char* lds, SharedView& sv, TeamView& tv) {
    const char* gbase = reinterpret_cast<const char*>(L);
    const int T = L->num_rotbonds, P = L->num_pairs;
    char* p = lds;
    const PairA* gpa = reinterpret_cast<const PairA*>(gbase + L->off_pair_a);
    const PairB* gpb = reinterpret_cast<const PairB*>(gbase + L->off_pair_b);
    if (PL) {
        PairA* la = reinterpret_cast<PairA*>(p);
        PairB* lb = reinterpret_cast<PairB*>(p + (size_t)P * sizeof(PairA));
        for (int k = threadIdx.x; k < P; k += blockDim.x) { la[k] = gpa[k]; lb[k] = gpb[k]; }
        sv.pa = la;
        sv.pb = lb;
        p += (size_t)P * (sizeof(PairA) + sizeof(PairB));
    } else {
        sv.pa = gpa;
        sv.pb = gpb;
    }
    float4* tu = reinterpret_cast<float4*>(p);
    float4* tm = tu + T;
    int2* ta = reinterpret_cast<int2*>(tm + T);
    const float4* gtu = reinterpret_cast<const float4*>(gbase + L->off_tor_unit);
    const float4* gtm = reinterpret_cast<const float4*>(gbase + L->off_tor_move);
    const int2* gta = reinterpret_cast<const int2*>(gbase + L->off_tor_atoms);
    for (int k = threadIdx.x; k < T; k += blockDim.x) { tu[k] = gtu[k]; tm[k] = gtm[k]; ta[k] = gta[k]; }
    sv.tor_unit = tu; sv.tor_move = tm; sv.tor_atoms = ta;
    const int npad = TW * NC;
    char* w = lds + shared_bytes_for(T, P, PL) + (size_t)(threadIdx.x / TW) * team_bytes_for(npad, T);
    tv.cx = reinterpret_cast<float*>(w); w += align16((size_t)npad * 4);
    tv.cy = reinterpret_cast<float*>(w); w += align16((size_t)npad * 4);
    tv.cz = reinterpret_cast<float*>(w); w += align16((size_t)npad * 4);
    tv.fx = reinterpret_cast<int*>(w); w += align16((size_t)npad * 4);
    tv.fy = reinterpret_cast<int*>(w); w += align16((size_t)npad * 4);
    tv.fz = reinterpret_cast<int*>(w); w += align16((size_t)npad * 4);
    tv.tq = reinterpret_cast<float4*>(w);
    tv.tax = tv.tq + T;
    tv.tref = tv.tax + T;
    w += (size_t)T * 48;
    tv.gi = reinterpret_cast<int*>(w);
    w += align16((size_t)T * 4);
    tv.pg = reinterpret_cast<float*>(w);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// per-lane atom data, loaded once per kernel
// ---------------------------------------------------------------------------
template <int TW, int NC>
struct Atoms {
    float3 ref[NC];
    float q[NC];
    uint32_t mlo[NC], mhi[NC];
    uint32_t mapoff[NC];      // map index * g3
    bool valid[NC], ign[NC], lig[NC];

    DEV void load(const LigandDev* L, const GridDev& G) {
        const char* gb = reinterpret_cast<const char*>(L);
        const float4* ar = reinterpret_cast<const float4*>(gb + L->off_atom_ref);
        const int4* ai = reinterpret_cast<const int4*>(gb + L->off_atom_info);
        const uint2* am = reinterpret_cast<const uint2*>(gb + L->off_atom_mask);
        const int tl = Team<TW>::lane();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            int a = tl + TW * k;
            valid[k] = a < L->num_atoms;
            int aa = valid[k] ? a : 0;
            float4 r = ar[aa];
            int4 inf = ai[aa];
            uint2 m = am[aa];
            ref[k] = make_float3(r.x, r.y, r.z);
            q[k] = r.w;
            mlo[k] = valid[k] ? m.x : 0u;
            mhi[k] = valid[k] ? m.y : 0u;
            mapoff[k] = valid[k] ? (uint32_t)inf.x * G.g3 : 0u;
            ign[k] = !valid[k] || inf.z != 0;
            lig[k] = valid[k] && a < L->true_ligand_atoms;
        }
    }
};

// Ligand scalars read once per kernel (kept in registers, not re-read from
// global memory inside the loops).
struct LigScalars {
    int T, P, NG;
    bool has_lig, any_general;
    float half_smooth, elec_min, inv_elec_min, coeff_desolv;
    DEV void load(const LigandDev* L) {
        T = L->num_rotbonds; P = L->num_pairs; NG = L->num_genes;
        has_lig = L->true_ligand_atoms != 0;
        any_general = L->any_general != 0;
        half_smooth = L->half_smooth; elec_min = L->elec_min_distance; coeff_desolv = L->coeff_desolv;
        inv_elec_min = L->inv_elec_min;
    }
};

template <int TW, int NC>
struct EvalCtx {
    const LigandDev* L;
    GridDev G;
    SharedView sv;
    TeamView tv;
    Atoms<TW, NC> A;
    LigScalars ls;
};

// ---------------------------------------------------------------------------
// Pose (calcenergy.cl:201-221): genes 0..5 and the sine/cosine of the
// orientation angles, which the gene lanes left in the team's LDS block.
// ---------------------------------------------------------------------------
struct Pose {
    float4 q;          // genrot quaternion
    float3 mv;         // translation (grid units)
    float phi, theta, rotangle, sin_theta, sin_half, is_theta_gt_pi;
};

// Every gene lane computes the trigonometry of its own angle gene once:
// orientation genes into pg, torsion genes into the torsion quaternions
// (calcenergy.cl:201-221, 248-260).  Followed by wave_sync by the caller.
template <int TW>
DEV void gene_trig(const LigScalars& ls, const SharedView& sv, const TeamView& tv, const float (&gene)[64 / TW]) {
    const int tl = Team<TW>::lane();
    const int NG = ls.NG;
#pragma unroll
    for (int j = 0; j < Team<TW>::kGPL; ++j) {
        const int g = tl + TW * j;
        if (j > 0 && NG <= TW) break;
        const float v = gene[j];
        if (g < 6) tv.pg[g] = v;
        if (g >= 3 && g < NG) {
            float ang = v * kDegToRad;
            if (g >= 5) ang = ang * 0.5f;
            float s, c;
            sincosf(ang, &s, &c);
            if (g < 6) {
                tv.pg[6 + 2 * (g - 3)] = s;
                tv.pg[7 + 2 * (g - 3)] = c;
            } else {
                float4 u = sv.tor_unit[g - 6];
                tv.tq[g - 6] = make_float4(s * u.x, s * u.y, s * u.z, c);
            }
        }
    }
}

DEV Pose read_pose(const TeamView& tv, bool has_ligand) {
    Pose P;
    const float4 p0 = *reinterpret_cast<const float4*>(tv.pg);       // g0 g1 g2 g3
    const float4 p1 = *reinterpret_cast<const float4*>(tv.pg + 4);   // g4 g5 sin(phi) cos(phi)
    const float4 p2 = *reinterpret_cast<const float4*>(tv.pg + 8);   // sin/cos theta, sin/cos rot/2
    P.phi = p0.w * kDegToRad;
    P.theta = p1.x * kDegToRad;
    P.rotangle = p1.y * kDegToRad;
    P.q = make_float4(0.f, 0.f, 0.f, 1.f);
    P.mv = make_float3(0.f, 0.f, 0.f);
    P.sin_theta = 0.f; P.sin_half = 0.f; P.is_theta_gt_pi = 1.f;
    if (has_ligand) {
        const float sp = p1.z, cp = p1.w, st = p2.x, ct = p2.y, sh = p2.z, ch = p2.w;
        P.mv = make_float3(p0.x, p0.y, p0.z);
        P.q = make_float4(sh * st * cp, sh * st * sp, sh * ct, ch);
        P.sin_theta = st;
        P.sin_half = sh;
        P.is_theta_gt_pi = 1.0f - 2.0f * (float)(st < 0.0f);
    }
    return P;
}

// Pose -> coordinates (calcenergy.cl:232-275).  Per atom: the rotatable bonds
// moving it in ascending index, then the general rotation (the per-atom order
// gen_rotlist encodes, host/src/calcenergy.cpp:387-488).
template <int TW, int NC>
DEV void compute_coords(const SharedView& sv, const TeamView& tv, const Atoms<TW, NC>& A, const Pose& P,
                        float3 (&c)[NC]) {
    const int tl = Team<TW>::lane();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        float3 v = A.ref[k];
        uint32_t lo = A.mlo[k], hi = A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            float4 q = tv.tq[t];
            float4 m = sv.tor_move[t];
            float3 r = qrot(make_float3(v.x - m.x, v.y - m.y, v.z - m.z), q);
            v = make_float3(r.x + m.x, r.y + m.y, r.z + m.z);
        }
        if (A.lig[k]) {
            float3 r = qrot(v, P.q);
            v = make_float3(r.x + P.mv.x, r.y + P.mv.y, r.z + P.mv.z);
        }
        c[k] = v;
        tv.cx[tl + TW * k] = v.x;
        tv.cy[tl + TW * k] = v.y;
        tv.cz[tl + TW * k] = v.z;
    }
}

// ---------------------------------------------------------------------------
// Intermolecular energy (+ map gradient) of one atom, calcenergy.cl:280-379,
// calcMergedEneGra.cl:222-443.  Split in two so the map loads of all atoms
// are in flight while the pair loop runs.
// ---------------------------------------------------------------------------
template <int MF>
struct Gather {
    float4 a0, a1, e0, e1, d0, d1;
    bool in;
};

template <int MF>
DEV void gather_issue(const GridDev& G, float3 c, uint32_t mapoff, Gather<MF>& g) {
    // out of the grid (calcenergy.cl:293-301); NaN coordinates count as out,
    // so no map index is ever formed from them
    g.in = (c.x >= 0.f) && (c.y >= 0.f) && (c.z >= 0.f) && (c.x < (float)(G.gx - 1)) &&
           (c.y < (float)(G.gy - 1)) && (c.z < (float)(G.gz - 1));
    uint32_t p = 0;
    if (g.in) p = (uint32_t)floorf(c.x) + (uint32_t)floorf(c.y) * G.g1 + (uint32_t)floorf(c.z) * G.g2;
    const uint32_t ne = (uint32_t)(G.nmaps - 2) * G.g3, nd = ne + G.g3;
    g.a0 = load_pt<MF>(G.data, mapoff + p); g.a1 = load_pt<MF>(G.data, mapoff + p + 1);
    g.e0 = load_pt<MF>(G.data, ne + p); g.e1 = load_pt<MF>(G.data, ne + p + 1);
    g.d0 = load_pt<MF>(G.data, nd + p); g.d1 = load_pt<MF>(G.data, nd + p + 1);
}

// The three maps are combined per cube corner first (type + q e + |q| d),
// then interpolated once: the same sum as the reference's three separate
// interpolations (calcenergy.cl:338-379), with one interpolation's work.
template <int MF, bool GRAD>
DEV float gather_finish(const GridDev& G, float3 c, float q, const Gather<MF>& g, int3& fi) {
    if (!g.in) {
        if (GRAD) { fi.x += 16777216; fi.y += 16777216; fi.z += 16777216; }
        return 16777216.0f;
    }
    const float xl = floorf(c.x), yl = floorf(c.y), zl = floorf(c.z);
    const float dx = c.x - xl, omdx = 1.0f - dx;
    const float dy = c.y - yl, omdy = 1.0f - dy;
    const float dz = c.z - zl, omdz = 1.0f - dz;
    const float aq = fabsf(q);
    // cube corner k: 0 000, 1 010, 2 001, 3 011 (low x: .x .y .z .w of a0/e0/d0), 4..7 the same at x+1
    const float4 l = make_float4(fmaf(aq, g.d0.x, fmaf(q, g.e0.x, g.a0.x)), fmaf(aq, g.d0.y, fmaf(q, g.e0.y, g.a0.y)),
                                 fmaf(aq, g.d0.z, fmaf(q, g.e0.z, g.a0.z)), fmaf(aq, g.d0.w, fmaf(q, g.e0.w, g.a0.w)));
    const float4 h = make_float4(fmaf(aq, g.d1.x, fmaf(q, g.e1.x, g.a1.x)), fmaf(aq, g.d1.y, fmaf(q, g.e1.y, g.a1.y)),
                                 fmaf(aq, g.d1.z, fmaf(q, g.e1.z, g.a1.z)), fmaf(aq, g.d1.w, fmaf(q, g.e1.w, g.a1.w)));
    // interpolate along x, then y, then z
    const float x00 = omdx * l.x + dx * h.x, x10 = omdx * l.y + dx * h.y;
    const float x01 = omdx * l.z + dx * h.z, x11 = omdx * l.w + dx * h.w;
    const float y0 = omdy * x00 + dy * x10, y1 = omdy * x01 + dy * x11;
    const float e = omdz * y0 + dz * y1;
    if (GRAD) {
        const float ig = G.inv_spacing;
        const float gx = omdz * (omdy * (h.x - l.x) + dy * (h.y - l.y)) + dz * (omdy * (h.z - l.z) + dy * (h.w - l.w));
        const float gy = omdz * (x10 - x00) + dz * (x11 - x01);
        const float gz = y1 - y0;
        fi.x += quant(kTermScale * (gx * ig));
        fi.y += quant(kTermScale * (gy * ig));
        fi.z += quant(kTermScale * (gz * ig));
    }
    return e;
}

// ---------------------------------------------------------------------------
// Intramolecular pair energy and dE/dr (calcenergy.cl:397-536,
// calcMergedEneGra.cl:455-583).
// ---------------------------------------------------------------------------
// The cutoff tests select terms instead of branching: over a team's pairs
// both ranges are almost always present, so branches would only add mask
// handling.  Adding the zero of an excluded term leaves the sum unchanged.
// rdist = 1 / (distance in grid units), inv_sp = 1 / spacing: 1/r in A is
// their product unless r is clamped at elec_min_distance.
template <bool GRAD>
DEV float pair_energy(const LigScalars& ls, const PairA& pa, const PairB& pb, float ad, float rdist, float inv_sp,
                      float& dedr) {
    // vdW / H-bond with smoothing, r < 8 A: with qm = sd^-m and qn = sd^-n,
    // E = AC qm - BD qn and dE/dsd = -(m AC qm - n BD qn) / sd.  sd^-12 is one
    // reciprocal of sd^12 (a few ulp; the 12th power of 1/sd would carry up to
    // 12 ulp), and sd^-n = sd^-12 sd^(12-n)
    const float delta = pb.reqm - ad;
    const float sd = (fabsf(delta) >= ls.half_smooth) ? ad + copysignf(ls.half_smooth, delta) : pb.reqm;
    const uint32_t fl = pa.flags;
    const uint32_t nexp = (fl >> 8) & 0xffu;
    const float q = fast_rcp(sd);
    const float r2 = sd * sd, r4 = r2 * r2, r6 = r4 * r2;
    float qm = fast_rcp(r6 * r6);                     // m = 12
    float qn = qm * (nexp == 6u ? r6 : r2);           // n = 6 or 10
    float fn = (float)nexp;
    bool general = false;
    if (__builtin_expect(ls.any_general, 0)) {        // other exponents (modified pairs): rare,
        if (!(fl & (1u << 17))) {                     // kept off the common path
            general = true;
            const float l2 = __builtin_amdgcn_logf(sd);
            qm = __builtin_amdgcn_exp2f(-(float)(fl & 0xffu) * l2);
            qn = __builtin_amdgcn_exp2f(-fn * l2);
        }
    }
    const float am = pa.ac * qm, bn = pa.bd * qn;
    const float evdw = am - bn;
    // electrostatics and desolvation, r < 20.48 A, r clamped at elec_min_distance:
    // 1/max(r, r_min) = min(1/r, 1/r_min)
    const float d = fmaxf(ad, ls.elec_min);
    const float d2 = d * d;
    const float desolv = pb.dsol * (ls.coeff_desolv * (12.96f - 0.1063f * d2 * (1.0f - 0.001947f * d2)) *
                                    fast_rcp(12.96f + d2 * (0.4137f + d2 * (0.00357f + 0.000112f * d2))));
    const float ds = d + 1.26366f;
    const float rds = fast_rcp(ds);
    const float rds2 = rds * rds;
    const float diel = 1.10859f * rds2 + 0.010358f;
    const float rd = fminf(rdist * inv_sp, ls.inv_elec_min);
    const float es = pb.cq * rd;
    const float eel = diel * es + desolv;
    const bool in8 = ad < 8.0f, in20 = ad < 20.48f;
    float e = pb.g * ad;                              // CG-G0 bond term, G = 50 (0 for other pairs)
    e += in8 ? evdw : 0.0f;
    e += in20 ? eel : 0.0f;
    if (GRAD) {
        const float mam = general ? (float)(fl & 0xffu) * am : 12.0f * am;
        const float dvdw = (fn * bn - mam) * q;
        const float del = -es * rd * diel - es * 2.21718f * rds2 * rds - 0.0771605f * d * desolv;
        float g = pb.g;
        g += in8 ? dvdw : 0.0f;
        g += in20 ? del : 0.0f;
        dedr = g;
    } else {
        dedr = 0.f;
    }
    return e;
}

// ---------------------------------------------------------------------------
// One evaluation of one genotype by one team.
//   GRAD = false: gpu_calc_energy (calcenergy.cl:132-554)
//   GRAD = true : gpu_calc_energrad (calcMergedEneGra.cl:46-936); grad[j]
//                 receives d E / d gene (lane + TW j), 0 beyond the genes.
// All lanes of the wavefront must call it together (uniform control flow).
// ---------------------------------------------------------------------------
template <int TW, int NC, int MF, bool GRAD>
DEV float evaluate(const EvalCtx<TW, NC>& X, const float (&gene)[64 / TW], float (&grad)[64 / TW]) {
    using TM = Team<TW>;
    const LigandDev* L = X.L;
    const GridDev& G = X.G;
    const TeamView& tv = X.tv;
    const SharedView& sv = X.sv;
    const int tl = TM::lane();
    const int T = X.ls.T, P = X.ls.P, NG = X.ls.NG;
    const bool has_lig = X.ls.has_lig;
    (void)L;

    gene_trig<TW>(X.ls, sv, tv, gene);
    if (GRAD) {
#pragma unroll
        for (int k = 0; k < NC; ++k) { tv.fx[tl + TW * k] = 0; tv.fy[tl + TW * k] = 0; tv.fz[tl + TW * k] = 0; }
        for (int t = tl; t < T; t += TW) tv.gi[t] = 0;
    }
    wave_sync();
    const Pose pose = read_pose(tv, has_lig);
    float3 c[NC];
    compute_coords<TW, NC>(sv, tv, X.A, pose, c);

    // with one atom per lane the map loads go out before the pair loop and
    // land while it runs; with more, each chunk is gathered after it (the
    // loads of all chunks would not fit the register budget)
    Gather<MF> gth0;
    if constexpr (NC == 1) gather_issue<MF>(G, c[0], X.A.mapoff[0], gth0);
    wave_sync();

    const float spacing = G.spacing;
    float e = 0.f;
    for (int p = tl; p < P; p += TW) {
        const PairA pa = sv.pa[p];
        const PairB pb = sv.pb[p];
        const int i = (int)(pa.ij & 0xffffu), j = (int)(pa.ij >> 16);
        const float sx = tv.cx[i] - tv.cx[j], sy = tv.cy[i] - tv.cy[j], sz = tv.cz[i] - tv.cz[j];
        const float d2g = sx * sx + sy * sy + sz * sz;
        const float rdist = fast_rsq(d2g);            // one transcendental for r and 1/r
        const float dist = d2g * rdist;
        float dedr;
        e += pair_energy<GRAD>(X.ls, pa, pb, dist * spacing, rdist, G.inv_spacing, dedr);
        if (GRAD) {
            // TERMSCALE * s * gdd with the power-of-two scale folded into gdd:
            // the same rounded product
            const float gdd = kTermScale * (-dedr * rdist);
            const int ix = quant(sx * gdd), iy = quant(sy * gdd), iz = quant(sz * gdd);
            lds_sub(&tv.fx[i], ix); lds_sub(&tv.fy[i], iy); lds_sub(&tv.fz[i], iz);
            atomicAdd(&tv.fx[j], ix); atomicAdd(&tv.fy[j], iy); atomicAdd(&tv.fz[j], iz);
        }
    }
    int3 fi[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        fi[k] = make_int3(0, 0, 0);
        if constexpr (NC == 1) {
            const float ea = gather_finish<MF, GRAD>(G, c[k], X.A.q[k], gth0, fi[k]);
            if (!X.A.ign[k]) e += ea;
        } else {
            Gather<MF> gk;
            gather_issue<MF>(G, c[k], X.A.mapoff[k], gk);
            const float ea = gather_finish<MF, GRAD>(G, c[k], X.A.q[k], gk, fi[k]);
            if (!X.A.ign[k]) e += ea;
        }
    }
    const float energy = TM::sum(e);
    if (!GRAD) return energy;

    wave_sync();
    // total per-atom forces (kcal/mol/A), int fixed point -> float
    float3 F[NC];
    float sfx = 0.f, sfy = 0.f, sfz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int a = tl + TW * k;
        F[k] = make_float3(0.f, 0.f, 0.f);
        if (X.A.valid[k]) {
            int fx = tv.fx[a], fy = tv.fy[a], fz = tv.fz[a];
            if (!X.A.ign[k]) { fx += fi[k].x; fy += fi[k].y; fz += fi[k].z; }
            F[k] = make_float3(kOneOverTermScale * (float)fx, kOneOverTermScale * (float)fy,
                               kOneOverTermScale * (float)fz);
            sfx += F[k].x; sfy += F[k].y; sfz += F[k].z;
            if (X.A.lig[k]) {   // torque about the ligand centre, calcMergedEneGra.cl:712-745
                const float rx = (c[k].x - pose.mv.x) * spacing, ry = (c[k].y - pose.mv.y) * spacing,
                            rz = (c[k].z - pose.mv.z) * spacing;
                tx += ry * F[k].z - rz * F[k].y;
                ty += rz * F[k].x - rx * F[k].z;
                tz += rx * F[k].y - ry * F[k].x;
            }
        }
    }
    sfx = TM::sum(sfx); sfy = TM::sum(sfy); sfz = TM::sum(sfz);
    tx = TM::sum(tx); ty = TM::sum(ty); tz = TM::sum(tz);
    int ig = 0;
    if (has_lig && tl < 6) {
        // translation genes, calcMergedEneGra.cl:679-697
        const int it0 = quant(kTermScale * sfx * spacing), it1 = quant(kTermScale * sfy * spacing),
                  it2 = quant(kTermScale * sfz * spacing);
        // rotation genes, calcMergedEneGra.cl:756-861
        const float tl2 = tx * tx + ty * ty + tz * tz;
        const float tlen = fast_sqrt(tl2);
        const bool small = tlen < 1e-3f;
        const float oscale = small ? 1.0f : tlen * (1.0f / 1e-3f);
        const float sc = small ? 0.5f - tlen * tlen * (1.0f / 48.0f) : 0.0005f * fast_rcp(tlen);
        const float4 qt = make_float4(tx * sc, ty * sc, tz * sc, small ? 1.0f - tlen * tlen * 0.125f : 0.99999988f);
        const float4 tq = qmul(qt, pose.q);
        const float cur_phi = fmod_pi2(kTwoPi + pose.phi);
        const float cur_theta = fmod_pi2(kTwoPi + pose.theta);
        const float cur_rot = fmod_pi2(kTwoPi + pose.rotangle);
        const float tgt_rot = 2.0f * fast_acos(tq.w);
        const float inv_sin = fast_rsq(1.0f - tq.w * tq.w);
        const float tgt_theta = kTwoPi + pose.is_theta_gt_pi * fast_acos(tq.z * inv_sin);
        const float tgt_phi = fmod_pi2(atan2f(pose.is_theta_gt_pi * tq.y, pose.is_theta_gt_pi * tq.x) + kTwoPi);
        const float gphi = oscale * (fmod_pi2(tgt_phi - cur_phi + kPi) - kPi);
        const float gtheta = oscale * (fmod_pi2(tgt_theta - cur_theta + kPi) - kPi);
        const float grot = oscale * (fmod_pi2(tgt_rot - cur_rot + kPi) - kPi);
        const float corr = 4.0f * pose.sin_half * pose.sin_half;
        const int ir0 = quant(kTermScale * gphi * pose.sin_theta * pose.sin_theta * corr * kDegToRad);
        const int ir1 = quant(kTermScale * gtheta * corr * kDegToRad);
        const int ir2 = quant(kTermScale * grot * kDegToRad);
        ig = tl == 0 ? it0 : tl == 1 ? it1 : tl == 2 ? it2 : tl == 3 ? ir0 : tl == 4 ? ir1 : ir2;
    }
    // torsion genes, calcMergedEneGra.cl:866-915
    for (int t = tl; t < T; t += TW) {
        const int2 ab = sv.tor_atoms[t];
        const float4 r0 = make_float4(tv.cx[ab.x], tv.cy[ab.x], tv.cz[ab.x], 0.f);
        const float ax = tv.cx[ab.y] - r0.x, ay = tv.cy[ab.y] - r0.y, az = tv.cz[ab.y] - r0.z;
        const float inv = fast_rsq(ax * ax + ay * ay + az * az);
        tv.tax[t] = make_float4(ax * inv, ay * inv, az * inv, 0.f);
        tv.tref[t] = r0;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        uint32_t lo = X.A.mlo[k], hi = X.A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            const float4 ax = tv.tax[t], r0 = tv.tref[t];
            const float rx = c[k].x - r0.x, ry = c[k].y - r0.y, rz = c[k].z - r0.z;
            const float qx = ry * F[k].z - rz * F[k].y, qy = rz * F[k].x - rx * F[k].z,
                        qz = rx * F[k].y - ry * F[k].x;
            const float toa = (ax.x * qx + ax.y * qy + ax.z * qz) * spacing;
            atomicAdd(&tv.gi[t], quant(kTermScale * toa * kDegToRad));
        }
    }
    wave_sync();
    const float conv = spacing * spacing * 1.0f / (kDegToRad * kDegToRad);   // CONVERT_INTO_ANGSTROM_RADIAN
#pragma unroll
    for (int j = 0; j < TM::kGPL; ++j) {
        const int g = tl + TW * j;
        int iv = 0;
        if (g < 6) iv = ig;
        else if (g < NG) iv = tv.gi[g - 6];
        float gv = kOneOverTermScale * (float)iv;
        if (g >= 3) gv *= conv;
        grad[j] = g < NG ? gv : 0.f;
    }
    return energy;
}

// ---------------------------------------------------------------------------
// Float gradient with the exact Mehler-Solmajer / exponential desolvation
// derivatives, used once per FIRE search (calcgradient.cl:109-1003).  Forces
// are float; intra forces use LDS float atomics of one team.
// ---------------------------------------------------------------------------
template <int TW, int NC, int MF>
DEV void float_gradient(const EvalCtx<TW, NC>& X, const float (&gene)[64 / TW], float (&grad)[64 / TW]) {
    using TM = Team<TW>;
    const LigandDev* L = X.L;
    const GridDev& G = X.G;
    const TeamView& tv = X.tv;
    const SharedView& sv = X.sv;
    const int tl = TM::lane();
    const int T = X.ls.T, P = X.ls.P, NG = X.ls.NG;
    const bool has_lig = X.ls.has_lig;
    (void)L;
    float* ffx = reinterpret_cast<float*>(tv.fx);
    float* ffy = reinterpret_cast<float*>(tv.fy);
    float* ffz = reinterpret_cast<float*>(tv.fz);
    float* tacc = reinterpret_cast<float*>(tv.gi);

    gene_trig<TW>(X.ls, sv, tv, gene);
#pragma unroll
    for (int k = 0; k < NC; ++k) { ffx[tl + TW * k] = 0.f; ffy[tl + TW * k] = 0.f; ffz[tl + TW * k] = 0.f; }
    for (int t = tl; t < T; t += TW) tacc[t] = 0.f;
    wave_sync();
    const Pose pose = read_pose(tv, has_lig);
    float3 c[NC];
    compute_coords<TW, NC>(sv, tv, X.A, pose, c);
    wave_sync();
    float3 gin[NC];
    const uint32_t g3 = G.g3, ne = (uint32_t)(G.nmaps - 2) * g3, nd = ne + g3;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        gin[k] = make_float3(0.f, 0.f, 0.f);
        if (X.A.ign[k]) continue;
        const float x = c[k].x, y = c[k].y, z = c[k].z;
        if (!((x >= 0.f) && (y >= 0.f) && (z >= 0.f) && (x < (float)(G.gx - 1)) && (y < (float)(G.gy - 1)) &&
              (z < (float)(G.gz - 1)))) {
            gin[k] = make_float3(16777216.0f, 16777216.0f, 16777216.0f);
            continue;
        }
        const uint32_t xl = (uint32_t)floorf(x), yl = (uint32_t)floorf(y), zl = (uint32_t)floorf(z);
        const uint32_t xh = (uint32_t)ceilf(x), yh = (uint32_t)ceilf(y), zh = (uint32_t)ceilf(z);
        const float dx = x - (float)xl, dy = y - (float)yl, dz = z - (float)zl;
        const uint32_t o[8] = {xl + yl * G.g1 + zl * G.g2, xl + yh * G.g1 + zl * G.g2, xl + yl * G.g1 + zh * G.g2,
                               xl + yh * G.g1 + zh * G.g2, xh + yl * G.g1 + zl * G.g2, xh + yh * G.g1 + zl * G.g2,
                               xh + yl * G.g1 + zh * G.g2, xh + yh * G.g1 + zh * G.g2};
        const uint32_t mb[3] = {X.A.mapoff[k], ne, nd};
        const float sc[3] = {1.0f, X.A.q[k], fabsf(X.A.q[k])};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float cb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) cb[i] = load_pt<MF>(G.data, mb[m] + o[i]).x;
            const float vx0 = (1.0f - dy) * (cb[4] - cb[0]) + dy * (cb[5] - cb[1]);
            const float vx1 = (1.0f - dy) * (cb[6] - cb[2]) + dy * (cb[7] - cb[3]);
            const float vy0 = (1.0f - dx) * (cb[1] - cb[0]) + dx * (cb[5] - cb[4]);
            const float vy1 = (1.0f - dx) * (cb[3] - cb[2]) + dx * (cb[7] - cb[6]);
            const float vz0 = (1.0f - dx) * (cb[2] - cb[0]) + dx * (cb[6] - cb[4]);
            const float vz1 = (1.0f - dx) * (cb[3] - cb[1]) + dx * (cb[7] - cb[5]);
            gin[k].x += sc[m] * ((1.0f - dz) * vx0 + dz * vx1);
            gin[k].y += sc[m] * ((1.0f - dz) * vy0 + dz * vy1);
            gin[k].z += sc[m] * ((1.0f - dy) * vz0 + dy * vz1);
        }
    }
    const float spacing = G.spacing;
    const float half_smooth = X.ls.half_smooth;
    const float DA = -8.5525f, DB = 78.4f - (-8.5525f), DK = 7.7839f;
    const float DBH = DB * 0.003627f, DBHK = DBH * DK;
    for (int p = tl; p < P; p += TW) {
        const PairA pa = sv.pa[p];
        const PairB pb = sv.pb[p];
        const int i = (int)(pa.ij & 0xffffu), j = (int)(pa.ij >> 16);
        const float sx = tv.cx[i] - tv.cx[j], sy = tv.cy[i] - tv.cy[j], sz = tv.cz[i] - tv.cz[j];
        const float dist = sqrtf(sx * sx + sy * sy + sz * sz);
        float d = dist * spacing;
        float sd;
        if (d <= pb.reqm - half_smooth) sd = d + half_smooth;
        else if (d < pb.reqm + half_smooth) sd = pb.reqm;
        else sd = d - half_smooth;
        const int m = (int)(pa.flags & 0xffu), n = (int)((pa.flags >> 8) & 0xffu);
        float pg = 0.f;
        if (d < 8.0f) {
            pg += (-(float)m * pa.ac) / powf(sd, (float)(m + 1));
            pg += ((float)n * pa.bd) / powf(sd, (float)(n + 1));
        }
        if (d < 20.48f) {
            if (d < X.ls.elec_min) d = X.ls.elec_min;
            const float ex = __expf(DBH * d);
            const float upper = DA * (ex + DK) * (ex + DK) + DB * ex * (DBHK * d + ex + DK);
            const float lw = DA * (ex + DK) + DB * ex;
            const float lower = d * d * lw * lw;
            pg += -pb.cq * (upper / lower);
            pg += pb.dsol * X.ls.coeff_desolv * -0.077160f * d * __expf(-0.038580f * d * d);
        }
        if (pa.flags & (1u << 16)) pg += 50.0f;
        const float fx = pg * (-sx / dist), fy = pg * (-sy / dist), fz = pg * (-sz / dist);
        atomicAdd(&ffx[i], -fx); atomicAdd(&ffy[i], -fy); atomicAdd(&ffz[i], -fz);
        atomicAdd(&ffx[j], fx); atomicAdd(&ffy[j], fy); atomicAdd(&ffz[j], fz);
    }
    wave_sync();
    float3 F[NC];
    float sfx = 0.f, sfy = 0.f, sfz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
    const float g0 = tv.pg[0], g1 = tv.pg[1], g2 = tv.pg[2];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int a = tl + TW * k;
        F[k] = make_float3(0.f, 0.f, 0.f);
        if (X.A.valid[k]) {
            F[k] = make_float3(gin[k].x / spacing + ffx[a], gin[k].y / spacing + ffy[a], gin[k].z / spacing + ffz[a]);
            if (X.A.lig[k]) {
                sfx += F[k].x; sfy += F[k].y; sfz += F[k].z;
                const float rx = (c[k].x - g0) * spacing, ry = (c[k].y - g1) * spacing, rz = (c[k].z - g2) * spacing;
                tx += ry * F[k].z - rz * F[k].y;
                ty += rz * F[k].x - rx * F[k].z;
                tz += rx * F[k].y - ry * F[k].x;
            }
        }
    }
    sfx = TM::sum(sfx); sfy = TM::sum(sfy); sfz = TM::sum(sfz);
    tx = TM::sum(tx); ty = TM::sum(ty); tz = TM::sum(tz);
    float gval = 0.f;
    if (tl < 3) gval = (tl == 0 ? sfx : tl == 1 ? sfy : sfz) * spacing;
    if (has_lig && tl >= 3 && tl < 6) {   // calcgradient.cl:714-873 (fmod)
        const float tlen = sqrtf(tx * tx + ty * ty + tz * tz);
        const bool small = tlen < 1e-3f;
        const float oscale = small ? 1.0f : tlen * (1.0f / 1e-3f);
        const float sc = small ? 0.5f - tlen * tlen / 48.0f : 0.0005f / tlen;
        const float4 qt = make_float4(tx * sc, ty * sc, tz * sc, small ? 1.0f - tlen * tlen * 0.125f : 0.99999988f);
        const float4 tq = qmul(qt, pose.q);
        const float cur_phi = fmod_pi2(kTwoPi + pose.phi);
        const float cur_theta = fmod_pi2(kTwoPi + pose.theta);
        const float cur_rot = fmod_pi2(kTwoPi + pose.rotangle);
        const float tgt_rot = 2.0f * fast_acos(tq.w);
        const float inv_sin = rsqrtf(1.0f - tq.w * tq.w);
        const float tgt_theta = kTwoPi + pose.is_theta_gt_pi * fast_acos(tq.z * inv_sin);
        const float tgt_phi = fmod_pi2(atan2f(pose.is_theta_gt_pi * tq.y, pose.is_theta_gt_pi * tq.x) + kTwoPi);
        const float corr = 4.0f * pose.sin_half * pose.sin_half;
        if (tl == 3) gval = oscale * (fmodf(tgt_phi - cur_phi + kPi, kTwoPi) - kPi) * pose.sin_theta * pose.sin_theta * corr * kDegToRad;
        if (tl == 4) gval = oscale * (fmodf(tgt_theta - cur_theta + kPi, kTwoPi) - kPi) * corr * kDegToRad;
        if (tl == 5) gval = oscale * (fmodf(tgt_rot - cur_rot + kPi, kTwoPi) - kPi) * kDegToRad;
    }
    for (int t = tl; t < T; t += TW) {
        const int2 ab = sv.tor_atoms[t];
        const float4 r0 = make_float4(tv.cx[ab.x], tv.cy[ab.x], tv.cz[ab.x], 0.f);
        const float ax = tv.cx[ab.y] - r0.x, ay = tv.cy[ab.y] - r0.y, az = tv.cz[ab.y] - r0.z;
        const float inv = rsqrtf(ax * ax + ay * ay + az * az);
        tv.tax[t] = make_float4(ax * inv, ay * inv, az * inv, 0.f);
        tv.tref[t] = r0;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        uint32_t lo = X.A.mlo[k], hi = X.A.mhi[k];
        while (lo | hi) {
            int t;
            if (lo) { t = __builtin_ctz(lo); lo &= lo - 1; }
            else { t = 32 + __builtin_ctz(hi); hi &= hi - 1; }
            const float4 ax = tv.tax[t], r0 = tv.tref[t];
            const float rx = (c[k].x - r0.x) * spacing, ry = (c[k].y - r0.y) * spacing, rz = (c[k].z - r0.z) * spacing;
            const float qx = ry * F[k].z - rz * F[k].y, qy = rz * F[k].x - rx * F[k].z, qz = rx * F[k].y - ry * F[k].x;
            atomicAdd(&tacc[t], (ax.x * qx + ax.y * qy + ax.z * qz) * kDegToRad);
        }
    }
    wave_sync();
    const float conv = spacing * spacing * 1.0f / (kDegToRad * kDegToRad);
#pragma unroll
    for (int j = 0; j < TM::kGPL; ++j) {
        const int g = tl + TW * j;
        float gv = 0.f;
        if (g < 6) gv = gval;
        else if (g < NG) gv = tacc[g - 6];
        if (g > 2) gv *= conv;
        grad[j] = g < NG ? gv : 0.f;
    }
}

// ---------------------------------------------------------------------------
// Kernel arguments and common prologue
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
    float lig_scale;     // SW: 1/sqrt(atoms)
    float gene_scale;    // SW: 1/sqrt(genes)
};

// A passed AutoStop test has ended the job: generation kernels do nothing.
#define JOB_STOPPED(b) ((b).stop != nullptr && *(b).stop != 0)

#define KERNEL_PROLOGUE(TW, NC, PL)                                            \
    extern __shared__ __align__(16) char lds[];                                \
    using TM = Team<TW>;                                                       \
    constexpr int GPL = TM::kGPL;                                              \
    EvalCtx<TW, NC> X;                                                         \
    X.L = a.L;                                                                 \
    X.G = a.G;                                                                 \
    stage<TW, NC, PL>(a.L, lds, X.sv, X.tv);                                   \
    X.A.load(a.L, a.G);                                                        \
    X.ls.load(a.L);                                                            \
    const int tl = TM::lane();                                                 \
    (void)tl;                                                                  \
    const int team_idx = (int)(blockIdx.x * (blockDim.x / TW) + threadIdx.x / TW); \
    (void)team_idx;

// Entity choice of every LS kernel (e.g. kernel3.cl:104-117): team lane 0
// draws from its own stream (state *st0) when le == 0.
template <int TW>
DEV int ls_entity(const SearchDev& S, int le, uint32_t* st0) {
    int ent = le;
    if (le == 0) {
        float r = 0.f;
        if (Team<TW>::lane() == 0) { uint32_t s = *st0; r = randf(s); *st0 = s; }
        r = Team<TW>::bcast(r, 0);
        if (100.0f * r > S.lsearch_rate) ent = S.num_ls;
    }
    // the reference reads entity num_ls even when num_ls == pop_size (the next
    // run's first entity, or past the buffer for the last run); clamp instead
    return ent < S.pop_size ? ent : S.pop_size - 1;
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
struct Shape {
    int tw, nc;
    bool pl;
    size_t lds;
    int threads;   // workgroup size (64, 128 or 256)
};

// Pair records are staged in LDS only while at least three workgroups still
// fit in a compute unit's 160 KB: beyond that, reading them through L1/L2
// keeps the four workgroups the registers allow.  Measured (M evals/s, FIRE /
// ADADELTA): 60 atoms in 64-lane teams, 57 KB in LDS (2 workgroups) 145 / 139
// against 172 / 167 from memory; 50 atoms in 32-lane teams, 49 KB in LDS (3
// workgroups) 236 / 204 against 206 / 192 from memory.
constexpr size_t kLdsBudget = 160 * 1024 / 3;

// Chooses the team width and atom chunks for a ligand: the requested width,
// else 32-lane teams (two individuals per wavefront) up to 128 atoms and
// 64-lane teams beyond; 16-lane teams (four per wavefront) only for ligands
// of at most 32 atoms.  Jobs pass the width fixed when they start.
inline Shape choose_shape(const LigandDev& L, int team_width_hint, int threads = 256) {
    Shape s{};
    s.threads = threads;
    const int N = L.num_atoms;
    const int tw = team_lanes_for(N, team_width_hint);
    s.tw = tw;
    const int per = (N + tw - 1) / tw;
    s.nc = per <= 1 ? 1 : (per <= 2 ? 2 : 4);
    s.pl = lds_bytes_for(s.tw, s.nc, L.num_rotbonds, L.num_pairs, true, threads) <= kLdsBudget;
    s.lds = lds_bytes_for(s.tw, s.nc, L.num_rotbonds, L.num_pairs, s.pl, threads);
    return s;
}

// Dispatches a kernel template over (TW, NC, MF, PL).  KT must be a
// template<int TW, int NC, int MF, bool PL> __global__ function.
#define ADGPU_DISPATCH(KT, shape, mf, body)                                                     \
    do {                                                                                        \
        auto go_ = [&](auto tw_, auto nc_, auto mf_, auto pl_) {                                 \
            constexpr int TW_ = decltype(tw_)::value, NC_ = decltype(nc_)::value;               \
            constexpr int MF_ = decltype(mf_)::value;                                           \
            constexpr bool PL_ = decltype(pl_)::value;                                          \
            auto kfn = KT<TW_, NC_, MF_, PL_>;                                                  \
            body;                                                                               \
        };                                                                                      \
        auto mf_sel = [&](auto tw_, auto nc_, auto pl_) {                                       \
            if ((mf) == ADGPU_MAP_F16) go_(tw_, nc_, std::integral_constant<int, ADGPU_MAP_F16>{}, pl_); \
            else if ((mf) == ADGPU_MAP_BF16) go_(tw_, nc_, std::integral_constant<int, ADGPU_MAP_BF16>{}, pl_); \
            else go_(tw_, nc_, std::integral_constant<int, ADGPU_MAP_F32>{}, pl_);             \
        };                                                                                      \
        auto pl_sel = [&](auto tw_, auto nc_) {                                                 \
            if ((shape).pl) mf_sel(tw_, nc_, std::true_type{});                                 \
            else mf_sel(tw_, nc_, std::false_type{});                                           \
        };                                                                                      \
        using I16 = std::integral_constant<int, 16>;                                            \
        using I32 = std::integral_constant<int, 32>;                                            \
        using I64 = std::integral_constant<int, 64>;                                            \
        using N1 = std::integral_constant<int, 1>;                                              \
        using N2 = std::integral_constant<int, 2>;                                              \
        using N4 = std::integral_constant<int, 4>;                                              \
        if ((shape).tw == 16) {                                                                 \
            if ((shape).nc == 1) pl_sel(I16{}, N1{});                                           \
            else pl_sel(I16{}, N2{});                                                           \
        } else if ((shape).tw == 32) {                                                          \
            if ((shape).nc == 1) pl_sel(I32{}, N1{});                                           \
            else if ((shape).nc == 2) pl_sel(I32{}, N2{});                                      \
            else pl_sel(I32{}, N4{});                                                           \
        } else {                                                                                \
            if ((shape).nc == 1) pl_sel(I64{}, N1{});                                           \
            else if ((shape).nc == 2) pl_sel(I64{}, N2{});                                      \
            else pl_sel(I64{}, N4{});                                                           \
        }                                                                                       \
    } while (0)

}  // namespace adgpu
