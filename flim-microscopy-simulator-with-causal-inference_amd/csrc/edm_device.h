// Device-side helpers shared by every ED-MRDT kernel (gfx950).
//
// * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
//   "Parallel random numbers: as easy as 1, 2, 3", SC'11): one call per
//   (step, particle/pair, stream tag) so results never depend on launch
//   shape or on which other jobs share the GPU.
// * The reference's numerics that the parity path must keep bit-for-bit:
//   Python float modulo (dynamics.py:390 `x % L`), int() truncation of cell
//   indices (dynamics.py:184), the single-wrap minimum image
//   (dynamics.py:200-216), the Hill factor (dynamics.py:966-976) and the
//   LCG-seeded Poisson draw (photophysics.py:487-521).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edm {

// ---------------------------------------------------------------- RNG
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Stream tags: the third counter word separates independent draw streams.
enum Tag : uint32_t {
    kTagNoise = 1,
    kTagBind = 2,
    kTagDissoc = 3,
    kTagCat = 4,
    kTagRecruitAcc = 5,
    kTagRecruitPos = 6,
    kTagDetach = 7,
    kTagConvert = 8,
    kTagPhotonPois = 9,
    kTagPhotonTrans = 10,
    kTagArrival = 11,
    kTagQE = 12,
    kTagPixel = 13,
    kTagJitter = 14,
    kTagDark = 15,
    kTagAfterpulse = 16,
    kTagRecruitLoc = 17,
    kTagRecruitOff = 18,
};

// Step number of a launch's draw keys.  Launches recorded into a graph that
// is replayed read the replay's first step from base[0] (written by the host
// before every replay) and add their offset in the recording; other launches
// carry the step itself (base == nullptr).
struct StepRef {
    const uint32_t* base;
    uint32_t off;
    __device__ __forceinline__ uint32_t get() const { return base ? *base + off : off; }
};

struct Rng {
    uint32_t k0, k1;
    __device__ __forceinline__ U4 draw(uint32_t step, uint32_t a, uint32_t b,
                                       uint32_t tag) const {
        return philox4x32(U4{step, a, b, tag}, k0, k1);
    }
};

// 53-bit double in [0, 1) from two 32-bit words.
__device__ __forceinline__ double u01d(uint32_t a, uint32_t b) {
    const uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)v * (1.0 / 9007199254740992.0);
}
// 24-bit float in [0, 1).
__device__ __forceinline__ float u01f(uint32_t a) {
    return (float)(a >> 8) * (1.0f / 16777216.0f);
}

template <typename Real>
struct Uni;
template <>
struct Uni<double> {
    __device__ __forceinline__ static double first(const U4& r) { return u01d(r.x, r.y); }
    __device__ __forceinline__ static double second(const U4& r) { return u01d(r.z, r.w); }
};
template <>
struct Uni<float> {
    __device__ __forceinline__ static float first(const U4& r) { return u01f(r.x); }
    __device__ __forceinline__ static float second(const U4& r) { return u01f(r.y); }
};

// Two standard normals (Box-Muller on one Philox block).
template <typename Real>
__device__ __forceinline__ void normal2(const U4& r, Real& z0, Real& z1) {
    // u1 in (0, 1] so log never sees 0
    const double u1 = 1.0 - u01d(r.x, r.y);
    const double u2 = u01d(r.z, r.w);
    if constexpr (sizeof(Real) == 8) {
        const double m = sqrt(-2.0 * log(u1));
        double s, c;
        sincospi(2.0 * u2, &s, &c);
        z0 = (Real)(m * c);
        z1 = (Real)(m * s);
    } else {
        const float m = sqrtf(-2.0f * logf((float)u1));
        float s, c;
        sincospif(2.0f * (float)u2, &s, &c);
        z0 = m * c;
        z1 = m * s;
    }
}

// Waves per SIMD some FP32 kernels are compiled for (a register cap, the
// minimum of amdgpu_waves_per_eu): they are bound by memory latency, and the
// extra waves hide more of it than the few spills the cap causes cost
// (measured per kernel: the single-membrane force kernel at 7, catalysis and
// the fused passes at 6, the ensemble's force, catalysis and passes at 7).
// FP64 instantiations are left uncapped.
template <typename Real, int kWaves>
inline constexpr int kFp32Waves = sizeof(Real) == 4 ? kWaves : 1;

// --------------------------------------------------------- reference math
template <typename Real>
__device__ __forceinline__ Real py_mod(Real x, Real L) {
    // a moved coordinate is within one box length of the box: there the
    // result is x, or x - L which is exact for L <= x < 2L (Sterbenz), as
    // fmod's; everything else (and +-0) takes the general path
    if (x > (Real)0 && x < L) return x;
    if (x >= L && x < (Real)2 * L) return x - L;
    Real r = fmod(x, L);
    if (r != (Real)0) {
        if ((L < (Real)0) != (r < (Real)0)) r += L;
    } else {
        r = copysign((Real)0, L);
    }
    return r;
}

// sin and cos of x in FP64 by Cody-Waite reduction modulo pi/2 (pi/2 split
// into two 33-bit parts and a 53-bit tail, so k * part is exact for
// |k| < 2^20) and Taylor polynomials on [-pi/4, pi/4] (through r^17 and
// r^18; the first omitted terms are below 5e-17 relative) in Horner form.
// Only IEEE add, multiply and rint, so a host build at -ffp-contract=off
// computes the same bits (the sinusoidal D(x) field of the FP64 parity
// path).
__device__ __forceinline__ void det_sincos(double x, double& s, double& c) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
    const double r = ((x - k * 0x1.921fb54400000p+0) - k * 0x1.0b4611a600000p-34) -
                     k * 0x1.3198a2e037073p-69;
    const double r2 = r * r;
    double ps = 0x1.952c77030ad4ap-49;  // 1/17!
    ps = -0x1.ae7f3e733b81fp-41 + r2 * ps;
    ps = 0x1.6124613a86d09p-33 + r2 * ps;
    ps = -0x1.ae64567f544e4p-26 + r2 * ps;
    ps = 0x1.71de3a556c734p-19 + r2 * ps;
    ps = -0x1.a01a01a01a01ap-13 + r2 * ps;
    ps = 0x1.1111111111111p-7 + r2 * ps;
    ps = -0x1.5555555555555p-3 + r2 * ps;
    const double sr = r + r * (r2 * ps);
    double pc = -0x1.6827863b97d97p-53;  // -1/18!
    pc = 0x1.ae7f3e733b81fp-45 + r2 * pc;
    pc = -0x1.93974a8c07c9dp-37 + r2 * pc;
    pc = 0x1.1eed8eff8d898p-29 + r2 * pc;
    pc = -0x1.27e4fb7789f5cp-22 + r2 * pc;
    pc = 0x1.a01a01a01a01ap-16 + r2 * pc;
    pc = -0x1.6c16c16c16c17p-10 + r2 * pc;
    pc = 0x1.5555555555555p-5 + r2 * pc;
    pc = -0x1.0000000000000p-1 + r2 * pc;
    const double cr = 1.0 + r2 * pc;
    switch (((long long)k) & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

__device__ __forceinline__ int clamp_cell(int v, int n) {
    if (v >= n) v = n - 1;
    if (v < 0) v = 0;
    return v;
}

template <typename Real>
__device__ __forceinline__ int trunc_cell(Real x, Real s, int n) {
    return clamp_cell((int)(x / s), n);
}

template <typename Real>
__device__ __forceinline__ Real min_image(Real x1, Real y1, Real x2, Real y2, Real Lx,
                                          Real Ly, bool px, bool py, Real& dx,
                                          Real& dy) {
    // single image shift by +-L beyond L/2, written as selects so that the
    // wavefront does not branch per candidate
    dx = x1 - x2;
    dy = y1 - y2;
    if (px) {
        const Real h = (Real)0.5 * Lx;
        const Real lo = dx + Lx, hi = dx - Lx;
        dx = dx > h ? hi : (dx < -h ? lo : dx);
    }
    if (py) {
        const Real h = (Real)0.5 * Ly;
        const Real lo = dy + Ly, hi = dy - Ly;
        dy = dy > h ? hi : (dy < -h ? lo : dy);
    }
    return dx * dx + dy * dy;
}

template <typename Real>
__device__ __forceinline__ Real hill_factor(Real rho, Real K, Real n, bool inhibition,
                                            Real basal) {
    if (K > (Real)0) {
        const Real ratio = rho / K;
        const Real rn = pow(ratio, n);
        if (inhibition) return (Real)1 / ((Real)1 + rn);
        const Real h = rn / ((Real)1 + rn);
        return basal + ((Real)1 - basal) * h;
    }
    return (Real)1;
}

// photophysics.py:487-521, restated with exact 64-bit integer arithmetic.
__device__ __forceinline__ int poisson_lcg(double mean, double u_seed) {
    if (mean <= 0.0) return 0;
    uint64_t s = (uint64_t)(int64_t)(u_seed * 9.007199254740992e15);
    const uint64_t A = 6364136223846793005ull, C = 1442695040888963407ull;
    if (mean < 30.0) {
        const double L = exp(-mean);
        int k = 0;
        double p = 1.0;
        for (;;) {
            s = s * A + C;
            const double u = (double)(s >> 11) / 9.007199254740992e15;
            p *= u;
            if (p <= L) break;
            ++k;
            if (k > 10000) break;
        }
        return k;
    }
    s = s * A + C;
    double u1 = (double)(s >> 11) / 9.007199254740992e15;
    s = s * A + C;
    const double u2 = (double)(s >> 11) / 9.007199254740992e15;
    if (u1 < 1e-30) u1 = 1e-30;
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double v = mean + sqrt(mean) * z + 0.5;
    const long long iv = (long long)v;
    return iv > 0 ? (int)iv : 0;
}

// --------------------------------------------------- wavefront helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Add `val` to counter[idx] for every active lane with `take`, issuing one
// atomic per distinct idx in the wavefront (lanes of a cell-sorted sweep
// mostly share the same coarse-grid cell).
__device__ __forceinline__ void wave_aggregated_add(int* counter, int idx, bool take,
                                                    int val) {
    uint64_t pending = __ballot(take);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lidx = __shfl(idx, leader);
        const bool mine = take && idx == lidx;
        const uint64_t same = __ballot(mine);
        if (lane_id() == leader) atomicAdd(&counter[lidx], val * __popcll(same));
        if (mine) take = false;
        pending &= ~same;
    }
}

__device__ __forceinline__ void wave_count(unsigned long long* counter, bool hit) {
    const uint64_t m = __ballot(hit);
    if (m && lane_id() == __ffsll((unsigned long long)m) - 1)
        atomicAdd(counter, (unsigned long long)__popcll(m));
}

}  // namespace edm
