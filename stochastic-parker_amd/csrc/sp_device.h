// sp_device.h -- device helpers shared by the mover and particle kernels.
#pragma once

#include "sp_internal.h"

namespace sp {

constexpr int kWave = 64;

// Philox4x32-10 (Salmon et al., SC'11).  One block of four 32-bit outputs
// feeds one Parker step: ran1, ran2, ran3 and the momentum draw.
struct U4 {
    uint32_t a, b, c, d;
};

// a ^ b ^ k in one gfx950 three-input bitwise op (truth table 0x96 = xor of
// all three; k wave-uniform); the compiler emits two v_xor_b32 for the same
// expression
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
    return r;
#else
    return a ^ b ^ k;   // host pass of the translation unit: never executed
#endif
}

// Each round's two 32 x 32 -> 64-bit products are one v_mad_u64_u32 each
// (both halves at once) instead of a v_mul_lo_u32 / v_mul_hi_u32 pair.
__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
    // A scalar zero the compiler cannot see through makes the round keys
    // depend on the call: they are then formed by scalar adds where they are
    // used instead of being hoisted out of the mover's loop as twenty
    // constants, which overflow the scalar registers and come back one
    // v_readlane per use.
    uint32_t z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    k0 += z;
    k1 += z;
#endif
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    return U4{c0, c1, c2, c3};
}

// genrand_double1 mapping: closed [0, 1]
// Index of the XCD (accelerator die) the calling wave runs on.
__device__ __forceinline__ unsigned xcc_id() {
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xFu;
}

__device__ __forceinline__ double u01(uint32_t r) { return (double)r * (1.0 / 4294967295.0); }

// Particle boundary condition for one rank (particle_module.f90:1984-2129):
// periodic axes wrap by (hi - lo) of the bounds passed in, open axes mark the
// particle as escaped.  Returns the weight leaked through an open face.
template <int NDIM>
__device__ __forceinline__ double boundary_condition(int flags, double& x, double& y, double& z,
                                                     int& cf, double w, double lo0, double lo1,
                                                     double lo2, double hi0, double hi1,
                                                     double hi2) {
    double leaked = 0.0;
    if (x < lo0 && cf == SP_COUNT_FLAG_INBOX) {
        if (flags & MF_OPEN_X) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_LX; }
        else x = x - lo0 + hi0;
    } else if (x > hi0 && cf == SP_COUNT_FLAG_INBOX) {
        if (flags & MF_OPEN_X) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_HX; }
        else x = x - hi0 + lo0;
    }
    if (y < lo1 && cf == SP_COUNT_FLAG_INBOX) {
        if (flags & MF_OPEN_Y) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_LY; }
        else y = y - lo1 + hi1;
    } else if (y > hi1 && cf == SP_COUNT_FLAG_INBOX) {
        if (flags & MF_OPEN_Y) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_HY; }
        else y = y - hi1 + lo1;
    }
    if (NDIM == 3) {
        if (z < lo2 && cf == SP_COUNT_FLAG_INBOX) {
            if (flags & MF_OPEN_Z) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_LZ; }
            else z = z - lo2 + hi2;
        } else if (z > hi2 && cf == SP_COUNT_FLAG_INBOX) {
            if (flags & MF_OPEN_Z) { leaked += w; cf = SP_COUNT_FLAG_ESCAPE_HZ; }
            else z = z - hi2 + lo2;
        }
    }
    return leaked;
}

// Weight -> fixed point (2^-32 units).  Particles carry weight 1 or a split
// weight 2^-k; the split refuses to go past k = kMaxSplitTimes and uploads
// reject weights that are not multiples of 2^-32, so the conversion is exact
// and every sum is exact and independent of the order of the additions (up to
// a total weight of 2^32, more particles than one GPU holds).
__device__ __forceinline__ unsigned long long weight_fx(double w) {
    return (unsigned long long)(w * kFxScale);
}

// Cell index of a position along one axis, clamped into the node array so
// that a corrupt or out-of-range position can never address memory outside
// it.  For in-range positions the clamp is inactive and the value equals
// the reference's floor(p).
__device__ __forceinline__ int clamp_cell(double fl, int ncell_max) {
    // fmax/fmin also map NaN to the lower bound
    const double c = fmin(fmax(fl, -2.0), (double)ncell_max);
    return (int)c;
}

// One particle's entry in every particle array.  Copies read the whole row
// before writing any of it, so the 18 loads go out together (source and
// destination may alias as far as the compiler knows, and interleaved loads
// and stores would each wait for the one before).
struct PtlRow {
    double x, y, z, p, v, mu, w, t, dt, rdt;
    int8_t split_times, count_flag;
    int32_t origin, nsteps_tracked, nsteps_pushed, tag_injected, tag_splitted;
    uint32_t rng_step;
};

__device__ __forceinline__ PtlRow load_row(const Ptl& S, long long i) {
    PtlRow r;
    r.x = S.x[i]; r.y = S.y[i]; r.z = S.z[i]; r.p = S.p[i]; r.v = S.v[i];
    r.mu = S.mu[i]; r.w = S.w[i]; r.t = S.t[i]; r.dt = S.dt[i]; r.rdt = S.rdt[i];
    r.split_times = S.split_times[i];
    r.count_flag = S.count_flag[i];
    r.origin = S.origin[i];
    r.nsteps_tracked = S.nsteps_tracked[i];
    r.nsteps_pushed = S.nsteps_pushed[i];
    r.tag_injected = S.tag_injected[i];
    r.tag_splitted = S.tag_splitted[i];
    r.rng_step = S.rng_step[i];
    return r;
}

// RDT: also the pause record (rdt), which only the spatial sort carries
template <bool RDT>
__device__ __forceinline__ void store_row(const Ptl& D, long long o, const PtlRow& r) {
    D.x[o] = r.x; D.y[o] = r.y; D.z[o] = r.z; D.p[o] = r.p; D.v[o] = r.v;
    D.mu[o] = r.mu; D.w[o] = r.w; D.t[o] = r.t; D.dt[o] = r.dt;
    if (RDT) D.rdt[o] = r.rdt;
    D.split_times[o] = r.split_times;
    D.count_flag[o] = r.count_flag;
    D.origin[o] = r.origin;
    D.nsteps_tracked[o] = r.nsteps_tracked;
    D.nsteps_pushed[o] = r.nsteps_pushed;
    D.tag_injected[o] = r.tag_injected;
    D.tag_splitted[o] = r.tag_splitted;
    D.rng_step[o] = r.rng_step;
}

// Per-particle dynamic state carried through the steps of one interval.
struct State {
    double x, y, z, p, t, dt;
    double mu;                    // read only by NLGC steps (kperp ~ mu^2)
};

}  // namespace sp
