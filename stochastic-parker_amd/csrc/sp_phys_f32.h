// sp_phys_f32.h -- FP32 gather, diffusion tensor and Parker step.
//
// Same physics as sp_phys_f64.h (particle_module.f90:2208-2450, :3358-3606,
// :4625-4907) with the gather, the coefficient arithmetic and the increments
// in FP32; position, momentum and elapsed time stay FP64 accumulators.
// Arithmetic is regrouped for the hardware: reciprocals and square roots use
// the single-instruction v_rcp_f32 / v_rsq_f32 / v_sqrt_f32, powers are
// exp2(y * log2(x)) on v_log_f32 / v_exp_f32, and a multiply and an add of
// one expression contract into an FMA (never across statements, so every
// kernel that runs this step rounds it alike).
#pragma once

#include "sp_device.h"

namespace sp {
namespace f32 {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^y = 2^(y log2 x), x > 0, on v_log_f32 / v_exp_f32 (about 1 ulp each): three
// instructions where the library powf spends about 120 on double-float
// arithmetic
__device__ __forceinline__ float fpow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

struct Interp {
    int ix, iy, iz;
    float rx, ry, rz;   // position inside the cell, [0, 1)
    float w[8];         // corner weights (unused by the coefficient records)
};

// The diffusion parameters a batched run varies per particle (KappaVar's
// FP32 forms): one copy per particle a lane carries.
struct KF {
    float kpara0, kret, pindex, gm2, perp_mag, perp_mom;
};

__device__ __forceinline__ KF kf_of(const MoverParams& m) {
    return KF{m.f_kpara0, m.f_kret, m.f_pindex, m.f_gm2, m.f_perp_mag, m.f_perp_mom};
}

__device__ __forceinline__ KF kf_of(const KappaVar& v) {
    return KF{v.f_kpara0, v.f_kret, v.f_pindex, v.f_gm2, v.f_perp_mag, v.f_perp_mom};
}

template <int NDIM>
__device__ __forceinline__ Interp interp_params(const MoverParams& m, double x, double y, double z) {
    Interp I;
    const double px = (x - m.xmin) * m.idx;
    const double py = (y - m.ymin) * m.idy;
    const double fx = floor(px), fy = floor(py);
    const float rx = (float)(px - fx), ry = (float)(py - fy);
    I.ix = clamp_cell(fx, m.nxg - 4) + 2;
    I.iy = clamp_cell(fy, m.nyg - 4) + 2;
    I.iz = 0;
    float rz = 0.0f;
    if (NDIM == 3) {
        const double pz = (z - m.zmin) * m.idz;
        const double fz = floor(pz);
        rz = (float)(pz - fz);
        I.iz = clamp_cell(fz, m.nzg - 4) + 2;
    }
    I.rx = rx;
    I.ry = ry;
    I.rz = rz;
    const float rx1 = 1.0f - rx, ry1 = 1.0f - ry, rz1 = 1.0f - rz;
    I.w[0] = rx1 * ry1 * rz1;
    I.w[1] = rx * ry1 * rz1;
    I.w[2] = rx1 * ry * rz1;
    I.w[3] = rx * ry * rz1;
    I.w[4] = rx1 * ry1 * rz;
    I.w[5] = rx * ry1 * rz;
    I.w[6] = rx1 * ry * rz;
    I.w[7] = rx * ry * rz;
    return I;
}

template <int NDIM, bool DPP>
struct RecShape {
    static constexpr int W = NDIM == 2 ? (DPP ? kRec2DX : kRec2D) : (DPP ? kRec3DX : kRec3D);
    static constexpr int NC = NDIM == 2 ? (DPP ? 18 : 15) : (DPP ? 28 : 21);
};

template <int NDIM, bool TINTERP, bool DPP = false>
__device__ __forceinline__ void gather(const MoverParams& m, const Interp& I, float rt,
                                       float* __restrict__ f) {
    constexpr int W = RecShape<NDIM, DPP>::W;
    constexpr int NC = RecShape<NDIM, DPP>::NC;
    constexpr int NCORN = NDIM == 2 ? 4 : 8;
    const long long sx = 1, sy = m.nxg, sz = (long long)m.nxg * m.nyg;
    const long long base = (long long)I.ix * sx + (long long)I.iy * sy + (long long)I.iz * sz;
    float g[TINTERP ? NC : 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        f[c] = 0.0f;
        if (TINTERP) g[c] = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < NCORN; ++q) {
        const int i = q & 1, j = (q >> 1) & 1, k = (q >> 2) & 1;
        const long long node = NDIM == 3 ? brick_record(I.ix + i, I.iy + j, I.iz + k, m.bnx, m.bny)
                                         : base + i * sx + j * sy + k * sz;
        const float4* r4 = reinterpret_cast<const float4*>(m.rec1 + node * W);
        float v[W];
#pragma unroll
        for (int u = 0; u < W / 4; ++u) {
            const float4 t = r4[u];
            v[4 * u] = t.x; v[4 * u + 1] = t.y; v[4 * u + 2] = t.z; v[4 * u + 3] = t.w;
        }
        const float wq = I.w[q];
#pragma unroll
        for (int c = 0; c < NC; ++c) f[c] += v[c] * wq;
        if (TINTERP) {
            const float4* s4 = reinterpret_cast<const float4*>(m.rec2 + node * W);
#pragma unroll
            for (int u = 0; u < W / 4; ++u) {
                const float4 t = s4[u];
                v[4 * u] = t.x; v[4 * u + 1] = t.y; v[4 * u + 2] = t.z; v[4 * u + 3] = t.w;
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) g[c] += v[c] * wq;
        }
    }
    if (TINTERP) {
#pragma unroll
        for (int c = 0; c < NC; ++c) f[c] = f[c] * (1.0f - rt) + g[c] * rt;
    }
}


// FP16 value `t` of a record's halves starting at word w0.
__device__ __forceinline__ float half_at(const uint32_t* wd, int w0, int t) {
    const uint32_t word = wd[w0 + t / 2];
    const uint16_t bits = (uint16_t)((t & 1) ? (word >> 16) : (word & 0xFFFFu));
    return (float)__builtin_bit_cast(_Float16, bits);
}

// Records a lane's particle last gathered, kept in registers across its
// steps: fields do not change within a mover launch, so records that are the
// same at the next step are not loaded again and only the lanes that need new
// ones issue loads (the texture path's cost follows the active lanes of a
// load).  2D: the records of the last cell; 3D: the record of each corner
// parity class; one set per field frame.  Used by the steps of one particle
// per lane (3D FP32 records take 160 registers per frame, most of them
// accumulation registers at one wave per SIMD); empty elsewhere.
template <int NDIM, bool TINTERP, bool HALF, bool DPP, int PPL>
struct GatherCache {
    static constexpr bool kOn = false;
    __device__ __forceinline__ void invalidate() {}
};
template <bool TINTERP, bool HALF>
struct GatherCache<2, TINTERP, HALF, false, 1> {
    static constexpr bool kOn = true;
    // FP32: the cell's coefficient record; FP16 gradients: its two pair
    // records; one set per field frame
    static constexpr int kN = HALF ? 2 * Pair2Shape::W16 / 4 : kCell2D / 4;
    uint4 raw[TINTERP ? 2 : 1][kN];
    unsigned key;    // node index of the cell's corner 0; ~0u: none
    __device__ __forceinline__ void invalidate() { key = ~0u; }
};

// 3D: the record of each corner parity class (gather_node3) with the node
// coordinates of the classes along each axis.  A step into a neighbouring
// cell keeps the classes whose node the two cells share (four of eight for
// a move along one axis) and loads only the others.
template <bool TINTERP, bool HALF>
struct GatherCache<3, TINTERP, HALF, false, 1> {
    static constexpr bool kOn = true;
    uint4 raw[TINTERP ? 2 : 1][8][(HALF ? kRec3H : kRec3F) / 4];   // per field frame
    int nx[2], ny[2], nz[2];   // -1: none
    __device__ __forceinline__ void invalidate() {
        nx[0] = nx[1] = ny[0] = ny[1] = nz[0] = nz[1] = -1;
    }
};

// 2D gather from FP16-gradient x-pair records (Pair2Shape): the record of
// node (ix, iy + j) holds corners (0, j) and (1, j).  The sums run over the
// corners in the order of gather() above; FP16 gradients are summed as floats
// and the frame's scale is undone after the sum (scaling by a power of two is
// exact).  Output in Rec2 slot order, divV in R2_DVXDX and zero in R2_DVYDY
// (the step only adds the two).
template <bool TINTERP, class GC>
__device__ __forceinline__ void gather_pair2(const MoverParams& m, const Interp& I, float rt,
                                             float* __restrict__ f, GC& gc) {
    using S = Pair2Shape;
    constexpr int W = S::W16;
    float a[S::N], g[TINTERP ? S::N : 1];
#pragma unroll
    for (int c = 0; c < S::N; ++c) {
        a[c] = 0.0f;
        if (TINTERP) g[c] = 0.0f;
    }
    // node indices in 32 bits (the host refuses 2D grids of 2^32 nodes)
    const unsigned node0 = (unsigned)I.ix + (unsigned)I.iy * (unsigned)m.nxg;
    bool miss = false;
    if constexpr (GC::kOn) miss = gc.key != node0;
#pragma unroll
    for (int frame = 0; frame < (TINTERP ? 2 : 1); ++frame) {
        const float* rec = frame == 0 ? m.rec1 : m.rec2;
        float* acc = frame == 0 ? a : g;
        uint4 raw[2][W / 4];   // both rows' loads go out before any is used
        if constexpr (GC::kOn) {
            if (miss) {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)(node0 + j * (unsigned)m.nxg) * W);
#pragma unroll
                    for (int u = 0; u < W / 4; ++u) gc.raw[frame][j * (W / 4) + u] = r4[u];
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int u = 0; u < W / 4; ++u) raw[j][u] = gc.raw[frame][j * (W / 4) + u];
        } else {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)(node0 + j * (unsigned)m.nxg) * W);
#pragma unroll
                for (int u = 0; u < W / 4; ++u) raw[j][u] = r4[u];
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint32_t wd[W];
#pragma unroll
            for (int u = 0; u < W / 4; ++u) {
                wd[4 * u] = raw[j][u].x; wd[4 * u + 1] = raw[j][u].y;
                wd[4 * u + 2] = raw[j][u].z; wd[4 * u + 3] = raw[j][u].w;
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float wq = I.w[2 * j + h];
#pragma unroll
                for (int c = 0; c < S::N; ++c) {
                    const float v = c < S::NF ? __uint_as_float(wd[h * S::NF + c])
                                              : half_at(wd, 2 * S::NF + h * S::NGW, c - S::NF);
                    acc[c] += v * wq;
                }
            }
        }
        const float gs = frame == 0 ? m.f_gs1 : m.f_gs2;
#pragma unroll
        for (int c = S::NF; c < S::N; ++c) acc[c] *= gs;
    }
    if constexpr (GC::kOn) gc.key = node0;
    if (TINTERP) {
#pragma unroll
        for (int c = 0; c < S::N; ++c) a[c] = a[c] * (1.0f - rt) + g[c] * rt;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] = a[c];   // VX .. BZ
    f[R2_DVXDX] = a[5];
    f[R2_DVYDY] = 0.0f;
#pragma unroll
    for (int c = R2_DBXDX; c <= R2_DBDY; ++c) f[c] = a[c - 1];
}

// 2D gather from FP32 cell records (kCell2D): f = c0 + rx c1 + ry (c2 + rx c3)
// for each value, three FMAs and no corner weights; the bilinear interpolant
// of the four corners, rounded differently from their weighted sum.  Output in
// Rec2 slot order, divV in R2_DVXDX and zero in R2_DVYDY.
template <bool TINTERP, class GC>
__device__ __forceinline__ void gather_cell2(const MoverParams& m, float rx, float ry, int ix,
                                             int iy, float rt, float* __restrict__ f, GC& gc) {
    constexpr int N = Pair2Shape::N;
    constexpr int NU = kCell2D / 4;
    // node index in 32 bits (the host refuses 2D grids of 2^32 nodes)
    const unsigned cell = (unsigned)ix + (unsigned)iy * (unsigned)m.nxg;
    bool miss = false;
    if constexpr (GC::kOn) miss = gc.key != cell;
    float a[N], g[TINTERP ? N : 1];
#pragma unroll
    for (int frame = 0; frame < (TINTERP ? 2 : 1); ++frame) {
        const float* rec = frame == 0 ? m.rec1 : m.rec2;
        float* acc = frame == 0 ? a : g;
        uint4 raw[NU];
        if constexpr (GC::kOn) {
            if (miss) {
                const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)cell * kCell2D);
#pragma unroll
                for (int u = 0; u < NU; ++u) gc.raw[frame][u] = r4[u];
            }
#pragma unroll
            for (int u = 0; u < NU; ++u) raw[u] = gc.raw[frame][u];
        } else {
            const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)cell * kCell2D);
#pragma unroll
            for (int u = 0; u < NU; ++u) raw[u] = r4[u];
        }
        uint32_t wd[kCell2D];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            wd[4 * u] = raw[u].x; wd[4 * u + 1] = raw[u].y;
            wd[4 * u + 2] = raw[u].z; wd[4 * u + 3] = raw[u].w;
        }
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const float c0 = __uint_as_float(wd[c]), c1 = __uint_as_float(wd[N + c]);
            const float c2 = __uint_as_float(wd[2 * N + c]), c3 = __uint_as_float(wd[3 * N + c]);
            acc[c] = __builtin_fmaf(ry, __builtin_fmaf(rx, c3, c2), __builtin_fmaf(rx, c1, c0));
        }
    }
    if constexpr (GC::kOn) gc.key = cell;
    if (TINTERP) {
#pragma unroll
        for (int c = 0; c < N; ++c) a[c] = a[c] * (1.0f - rt) + g[c] * rt;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] = a[c];   // VX .. BZ
    f[R2_DVXDX] = a[5];
    f[R2_DVYDY] = 0.0f;
#pragma unroll
    for (int c = R2_DBXDX; c <= R2_DBDY; ++c) f[c] = a[c - 1];
}

// 3D gather from the FP32 movers' bricked node records (kRec3F, kRec3H).  The
// eight corners are visited by parity class: class q = (a, b, c) is the
// corner whose node coordinates have parities a, b, c, i.e. the node
// (ix + (a ^ ix&1), iy + (b ^ iy&1), iz + (c ^ iz&1)).  Every FP32 3D gather
// sums in class order, so the cached and uncached kernels round alike.  FP16
// values (V with the velocity scale, divV and the gradients with the gradient
// scale) are summed as floats and scaled after the sum.  Output in Rec3 slot
// order, divV in R3_DVXDX and zeros in R3_DVYDY, R3_DVZDZ.
template <bool TINTERP, bool HALF, class GC>
__device__ __forceinline__ void gather_node3(const MoverParams& m, const Interp& I, float rt,
                                             float* __restrict__ f, GC& gc) {
    constexpr int W = HALF ? kRec3H : kRec3F;
    constexpr int N = 19;
    const int px = I.ix & 1, py = I.iy & 1, pz = I.iz & 1;
    // node coordinate of each parity class along each axis
    const int nx[2] = {I.ix + px, I.ix + (px ^ 1)};
    const int ny[2] = {I.iy + py, I.iy + (py ^ 1)};
    const int nz[2] = {I.iz + pz, I.iz + (pz ^ 1)};
    // weight factor of each parity class along each axis
    const float rx1 = 1.0f - I.rx, ry1 = 1.0f - I.ry, rz1 = 1.0f - I.rz;
    const float wx[2] = {px ? I.rx : rx1, px ? rx1 : I.rx};
    const float wy[2] = {py ? I.ry : ry1, py ? ry1 : I.ry};
    const float wz[2] = {pz ? I.rz : rz1, pz ? rz1 : I.rz};
    // brick_record in 32-bit arithmetic (the host refuses grids of 2^32 or
    // more bricked nodes), its per-axis parts formed once per class
    unsigned bx[2], by[2], bz[2], lo_x[2], lo_y[2], lo_z[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        bx[q] = (unsigned)nx[q] >> 3; lo_x[q] = (unsigned)nx[q] & 7u;
        by[q] = (unsigned)ny[q] >> 3; lo_y[q] = ((unsigned)ny[q] & 7u) << 3;
        bz[q] = (unsigned)nz[q] >> 3; lo_z[q] = ((unsigned)nz[q] & 7u) << 6;
    }
    auto record = [&](int q) -> unsigned {
        const int qa = q & 1, qb = (q >> 1) & 1, qc = (q >> 2) & 1;
        const unsigned b = __umul24(__umul24(bz[qc], (unsigned)m.bny) + by[qb], (unsigned)m.bnx) + bx[qa];
        return (b << 9) | lo_z[qc] | lo_y[qb] | lo_x[qa];
    };
    float a[N], g[TINTERP ? N : 1];
#pragma unroll
    for (int c = 0; c < N; ++c) {
        a[c] = 0.0f;
        if (TINTERP) g[c] = 0.0f;
    }
    // a class keeps its node when the cell moved along no axis on which that
    // class's node coordinate changed
    bool keep_x[2] = {false, false}, keep_y[2] = {false, false}, keep_z[2] = {false, false};
    if constexpr (GC::kOn) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            keep_x[q] = nx[q] == gc.nx[q];
            keep_y[q] = ny[q] == gc.ny[q];
            keep_z[q] = nz[q] == gc.nz[q];
        }
    }
#pragma unroll
    for (int frame = 0; frame < (TINTERP ? 2 : 1); ++frame) {
        const float* rec = frame == 0 ? m.rec1 : m.rec2;
        float* acc = frame == 0 ? a : g;
        uint4 raw[8][W / 4];   // all eight classes' loads go out before any is used
        if constexpr (GC::kOn) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int qa = q & 1, qb = (q >> 1) & 1, qc = (q >> 2) & 1;
                if (!(keep_x[qa] && keep_y[qb] && keep_z[qc])) {
                    const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)record(q) * W);
#pragma unroll
                    for (int u = 0; u < W / 4; ++u) gc.raw[frame][q][u] = r4[u];
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int u = 0; u < W / 4; ++u) raw[q][u] = gc.raw[frame][q][u];
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint4* r4 = reinterpret_cast<const uint4*>(rec + (size_t)record(q) * W);
#pragma unroll
                for (int u = 0; u < W / 4; ++u) raw[q][u] = r4[u];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            uint32_t wd[W];
#pragma unroll
            for (int u = 0; u < W / 4; ++u) {
                wd[4 * u] = raw[q][u].x; wd[4 * u + 1] = raw[q][u].y;
                wd[4 * u + 2] = raw[q][u].z; wd[4 * u + 3] = raw[q][u].w;
            }
            const float wq = wx[q & 1] * wy[(q >> 1) & 1] * wz[(q >> 2) & 1];
#pragma unroll
            for (int c = 0; c < N; ++c) {
                float v;
                if (!HALF) v = __uint_as_float(wd[c]);
                else if (c >= 3 && c < 6) v = __uint_as_float(wd[c - 3]);   // B
                else v = half_at(wd, 3, c < 3 ? c : c - 3);               // V, divV, gradients
                acc[c] += v * wq;
            }
        }
        if (HALF) {
            const float gs = frame == 0 ? m.f_gs1 : m.f_gs2;
            const float vs = frame == 0 ? m.f_vs1 : m.f_vs2;
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] *= vs;
#pragma unroll
            for (int c = 6; c < N; ++c) acc[c] *= gs;
        }
    }
    if constexpr (GC::kOn) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            gc.nx[q] = nx[q];
            gc.ny[q] = ny[q];
            gc.nz[q] = nz[q];
        }
    }
    if (TINTERP) {
#pragma unroll
        for (int c = 0; c < N; ++c) a[c] = a[c] * (1.0f - rt) + g[c] * rt;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) f[c] = a[c];   // VX .. BZ
    f[R3_DVXDX] = a[6];
    f[R3_DVYDY] = 0.0f;
    f[R3_DVZDZ] = 0.0f;
#pragma unroll
    for (int c = R3_DBXDX; c <= R3_DBDZ; ++c) f[c] = a[c - 2];
}

// Momentum diffusion in FP32 (f64::dpp_terms, particle_module.f90:2918-2979):
// va^2 = B^2 / rho; p^(pindex-1) p0^(2-pindex) = p0^2 (p/p0)^pindex / p.
template <int NDIM>
__device__ __forceinline__ void dpp_terms(const MoverParams& m, const KF& kf, float pf, const float* f,
                                          float kpara, float knorm_para, float b2, float bx,
                                          float by, float bz, float ib, float divv, float& dp_dt,
                                          float& dpp) {
    if (m.flags & MF_DPP_WAVE) {
        const float rho = NDIM == 2 ? f[R2_RHO] : f[R3_RHO];
        const float va2 = b2 * frcp(rho);
        const float c = (m.flags & MF_MOMENTUM_DEP) ? (8.0f / 27.0f) : (4.0f / 9.0f);
        const float pk = pf * frcp(kpara) * va2;
        dp_dt += c * pk;
        dpp += (1.0f / 9.0f) * pf * pk;
    }
    if (m.flags & MF_DPP_SHEAR) {
        const float d3 = divv * (1.0f / 3.0f);
        float sxx, syy, szz, sxy, sxz = 0.f, syz = 0.f;
        if (NDIM == 2) {
            sxx = f[R2_DVXDX] - d3;
            syy = f[R2_DVYDY] - d3;
            szz = -d3;
            sxy = 0.5f * (f[R2_DVXDY] + f[R2_DVYDX]);
        } else {
            sxx = f[R3_DVXDX] - d3;
            syy = f[R3_DVYDY] - d3;
            szz = f[R3_DVZDZ] - d3;
            sxy = 0.5f * (f[R3_DVXDY] + f[R3_DVYDX]);
            sxz = 0.5f * (f[R3_DVXDZ] + f[R3_DVZDX]);
            syz = 0.5f * (f[R3_DVYDZ] + f[R3_DVZDY]);
        }
        float gshear;
        if (m.flags & MF_WEAK_SCATTERING) {
            const float bbs = (sxx * bx * bx + syy * by * by + szz * bz * bz +
                               2.0f * (sxy * bx * by + sxz * bx * bz + syz * by * bz)) * ib * ib;
            gshear = bbs * bbs * 0.2f;
        } else {
            gshear = (2.0f / 15.0f) * (sxx * sxx + syy * syy + szz * szz +
                                       2.0f * (sxy * sxy + sxz * sxz + syz * syz));
        }
        if (gshear > 0.0f) {
            const float q = fpow(pf * (float)m.inv_p0, kf.pindex);
            const float dd = gshear * m.f_tau0 * knorm_para * m.f_p02 * q;
            dp_dt += (2.0f + kf.pindex) * dd * frcp(pf);
            dpp += dd;
        }
    }
}

// One Parker step in FP32 coefficients.  Semantics match f64::push: same dt
// rule (including the dt_min fallback when a rate is exactly zero), same
// increments, same momentum floor; d = (dx, dy, dz, dp) as doubles.  The
// configuration's flags and the per-particle choices (the dt rule's branches,
// the acceleration gate, the floor) are selects rather than branches, so
// the step of several particles per lane is one straight-line block.
// With FX_LEAN (no magnetic-field dependence of kappa, no 2D out-of-plane
// drift) the terms those flags zero are left out of the arithmetic.
template <int NDIM, int FX>
__device__ __forceinline__ void kappa_push(const MoverParams& m, const KF& kf, State& s,
                                           const float* f, bool fixed_dt, const float r[4],
                                           double d[4]) {
    constexpr bool GEN = (FX & FX_LEAN) == 0;
    constexpr bool ACC = (FX & FX_ACC) != 0;
    constexpr bool DPP = (FX & FX_DPP) != 0;
    constexpr bool NLGC = (FX & FX_NLGC) != 0;
    const float bx = NDIM == 2 ? f[R2_BX] : f[R3_BX];
    const float by = NDIM == 2 ? f[R2_BY] : f[R3_BY];
    const float bz = NDIM == 2 ? f[R2_BZ] : f[R3_BZ];
    const float b2 = bx * bx + by * by + bz * bz;
    const float b = fsqrt(b2);
    const bool bzero = b < (float)__DBL_EPSILON__;
    const float ibk = bzero ? 1.0f : frcp(b);   // kappa's 1/B (1 when B = 0)
    const float ib = bzero ? 0.0f : ibk;         // pusher's 1/B (0 when B = 0)
    const float ib2 = ibk * ibk, ib3 = ibk * ib2;
    const float pf = (float)s.p;
    const bool mag = GEN && (m.flags & MF_MAG_DEP) != 0;

    const bool momdep = (m.flags & MF_MOMENTUM_DEP) != 0;
    const float pnorm = (float)(s.p * m.inv_p0);

    // --- kappa ---------------------------------------------------------
    float knorm = mag ? fpow(b, kf.gm2) : 1.0f;
    const float knorm_para = knorm;
    knorm = momdep ? knorm * fpow(pnorm, kf.pindex) : knorm;
    const float kpara = kf.kpara0 * knorm;
    float kperp, kpp;
    if constexpr (NLGC) {
        // kpara0 * kperp_kpara * B^((gamma-2)/3) * (p/p0)^((5-gamma)/3) * mu^2
        float kn_perp = mag ? fpow(b, kf.perp_mag) : 1.0f;
        kn_perp = momdep ? kn_perp * fpow(pnorm, kf.perp_mom) : kn_perp;
        const float muf = (float)s.mu;
        kperp = kf.kpara0 * m.f_kperp_kpara * kn_perp * (muf * muf);
        kpp = kpara - kperp;
    } else {
        kperp = kpara * kf.kret;
        // kpara - kperp as a product of non-negative factors: a contracted
        // kpara - kpara * kret can come out as a negative rounding error when
        // kret = 1, and its square root as NaN
        kpp = kpara * (1.0f - kf.kret);
    }
    const float skpara = fsqrt(2.0f * kpara);
    const float skperp = fsqrt(2.0f * kperp);
    const float skpp = fsqrt(2.0f * kpp);

    float dkxx_dx, dkyy_dy, dkxy_dx, dkxy_dy;
    float dkzz_dz = 0.f, dkxz_dx = 0.f, dkxz_dz = 0.f, dkyz_dy = 0.f, dkyz_dz = 0.f;
    float dx_dt, dy_dt, dz_dt, divv;
    // drift prefactor vdp = pre / sqrt((d1 p0/p)^2 + (d2 p0^2/p^2)^2)
    const float ip = frcp(pf);
    const float t1 = m.f_d1p0 * ip;
    const float t2 = m.f_d2p02 * ip * ip;
    const float vdp = m.f_vdp_pre * frsq(t1 * t1 + t2 * t2);
    float bxn = 0.f, byn = 0.f, bzn = 0.f, bxyn = 0.f, ibxyn = 0.f;
    if (NDIM == 2) {
        const float dbx_dx = f[R2_DBXDX], dbx_dy = f[R2_DBXDY];
        const float dby_dx = f[R2_DBYDX], dby_dy = f[R2_DBYDY];
        const float dbz_dx = f[R2_DBZDX], dbz_dy = f[R2_DBZDY];
        const float db_dx = f[R2_DBDX], db_dy = f[R2_DBDY];
        const float dkdx = mag ? db_dx * ibk * kf.gm2 : 0.f;
        const float dkdy = mag ? db_dy * ibk * kf.gm2 : 0.f;
        const float kb2 = kpp * ib2, kb3 = 2.0f * kpp * ib3;
        if constexpr (NLGC) {
            // d ln kperp = d ln kpara / 3 along B's magnitude
            const float ex = dkdx * (1.0f / 3.0f), ey = dkdy * (1.0f / 3.0f);
            const float ax = (kpara * dkdx - kperp * ex) * ib2, ay = (kpara * dkdy - kperp * ey) * ib2;
            dkxx_dx = kperp * ex + ax * bx * bx + kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kperp * ey + ay * by * by + kb3 * by * (dby_dy * b - by * db_dy);
            dkxy_dx = ax * bx * by + kb2 * (dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = ay * bx * by + kb2 * (dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
        } else if constexpr (GEN) {
            dkxx_dx = kperp * dkdx + kb2 * dkdx * bx * bx + kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kperp * dkdy + kb2 * dkdy * by * by + kb3 * by * (dby_dy * b - by * db_dy);
            dkxy_dx = kb2 * (dkdx * bx * by + dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = kb2 * (dkdy * bx * by + dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
        } else {
            dkxx_dx = kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kb3 * by * (dby_dy * b - by * db_dy);
            dkxy_dx = kb2 * (dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = kb2 * (dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
        }
        const float ibp2 = ib * ib, ibp3 = ib * ibp2;
        const float vdx = vdp * (dbz_dy * ibp2 - 2.0f * bz * db_dy * ibp3);
        const float vdy = vdp * (-(dbz_dx * ibp2) + 2.0f * bz * db_dx * ibp3);
        const float vdz = (GEN && (m.flags & MF_DRIFT_2D))
            ? vdp * ((dby_dx - dbx_dy) * ibp2 - 2.0f * (by * db_dx - bx * db_dy) * ibp3) : 0.f;
        dx_dt = f[R2_VX] + vdx + dkxx_dx + dkxy_dy;
        dy_dt = f[R2_VY] + vdy + dkxy_dx + dkyy_dy;
        dz_dt = vdz;
        divv = f[R2_DVXDX] + f[R2_DVYDY];
    } else {
        const float dbx_dx = f[R3_DBXDX], dbx_dy = f[R3_DBXDY], dbx_dz = f[R3_DBXDZ];
        const float dby_dx = f[R3_DBYDX], dby_dy = f[R3_DBYDY], dby_dz = f[R3_DBYDZ];
        const float dbz_dx = f[R3_DBZDX], dbz_dy = f[R3_DBZDY], dbz_dz = f[R3_DBZDZ];
        const float db_dx = f[R3_DBDX], db_dy = f[R3_DBDY], db_dz = f[R3_DBDZ];
        const float kb2 = kpp * ib2, kb3 = 2.0f * kpp * ib3;
        if constexpr (NLGC) {
            // NLGC keeps 1/B in the 3D logarithmic gradients (:2708-2715)
            const float dkdx = mag ? db_dx * ibk * kf.gm2 : 0.f;
            const float dkdy = mag ? db_dy * ibk * kf.gm2 : 0.f;
            const float dkdz = mag ? db_dz * ibk * kf.gm2 : 0.f;
            const float ex = dkdx * (1.0f / 3.0f), ey = dkdy * (1.0f / 3.0f), ez = dkdz * (1.0f / 3.0f);
            const float ax = (kpara * dkdx - kperp * ex) * ib2;
            const float ay = (kpara * dkdy - kperp * ey) * ib2;
            const float az = (kpara * dkdz - kperp * ez) * ib2;
            dkxx_dx = kperp * ex + ax * bx * bx + kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kperp * ey + ay * by * by + kb3 * by * (dby_dy * b - by * db_dy);
            dkzz_dz = kperp * ez + az * bz * bz + kb3 * bz * (dbz_dz * b - bz * db_dz);
            dkxy_dx = ax * bx * by + kb2 * (dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = ay * bx * by + kb2 * (dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
            dkxz_dx = ax * bx * bz + kb2 * (dbx_dx * bz + bx * dbz_dx) - kb3 * bx * bz * db_dx;
            dkxz_dz = az * bx * bz + kb2 * (dbx_dz * bz + bx * dbz_dz) - kb3 * bx * bz * db_dz;
            dkyz_dy = ay * by * bz + kb2 * (dby_dy * bz + by * dbz_dy) - kb3 * by * bz * db_dy;
            dkyz_dz = az * by * bz + kb2 * (dby_dz * bz + by * dbz_dz) - kb3 * by * bz * db_dz;
        } else if constexpr (GEN) {
            const float dkdx = mag ? db_dx * kf.gm2 : 0.f;
            const float dkdy = mag ? db_dy * kf.gm2 : 0.f;
            const float dkdz = mag ? db_dz * kf.gm2 : 0.f;
            dkxx_dx = kperp * dkdx + kb2 * dkdx * bx * bx + kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kperp * dkdy + kb2 * dkdy * by * by + kb3 * by * (dby_dy * b - by * db_dy);
            dkzz_dz = kperp * dkdz + kb2 * dkdz * bz * bz + kb3 * bz * (dbz_dz * b - bz * db_dz);
            dkxy_dx = kb2 * (dkdx * bx * by + dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = kb2 * (dkdy * bx * by + dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
            dkxz_dx = kb2 * (dkdx * bx * bz + dbx_dx * bz + bx * dbz_dx) - kb3 * bx * bz * db_dx;
            dkxz_dz = kb2 * (dkdz * bx * bz + dbx_dz * bz + bx * dbz_dz) - kb3 * bx * bz * db_dz;
            dkyz_dy = kb2 * (dkdy * by * bz + dby_dy * bz + by * dbz_dy) - kb3 * by * bz * db_dy;
            dkyz_dz = kb2 * (dkdz * by * bz + dby_dz * bz + by * dbz_dz) - kb3 * by * bz * db_dz;
        } else {
            dkxx_dx = kb3 * bx * (dbx_dx * b - bx * db_dx);
            dkyy_dy = kb3 * by * (dby_dy * b - by * db_dy);
            dkzz_dz = kb3 * bz * (dbz_dz * b - bz * db_dz);
            dkxy_dx = kb2 * (dbx_dx * by + bx * dby_dx) - kb3 * bx * by * db_dx;
            dkxy_dy = kb2 * (dbx_dy * by + bx * dby_dy) - kb3 * bx * by * db_dy;
            dkxz_dx = kb2 * (dbx_dx * bz + bx * dbz_dx) - kb3 * bx * bz * db_dx;
            dkxz_dz = kb2 * (dbx_dz * bz + bx * dbz_dz) - kb3 * bx * bz * db_dz;
            dkyz_dy = kb2 * (dby_dy * bz + by * dbz_dy) - kb3 * by * bz * db_dy;
            dkyz_dz = kb2 * (dby_dz * bz + by * dbz_dz) - kb3 * by * bz * db_dz;
        }
        bxn = bx * ib; byn = by * ib; bzn = bz * ib;
        const float bxy2 = bxn * bxn + byn * byn;
        bxyn = fsqrt(bxy2);
        ibxyn = bxyn < (float)__DBL_EPSILON__ ? 0.0f : frcp(bxyn);
        const float ibp2 = ib * ib, ibp3 = ib * ibp2;
        const float vdx = vdp * ((dbz_dy - dby_dz) * ibp2 - 2.0f * (bz * db_dy - by * db_dz) * ibp3);
        const float vdy = vdp * ((dbx_dz - dbz_dx) * ibp2 - 2.0f * (bx * db_dz - bz * db_dx) * ibp3);
        const float vdz = vdp * ((dby_dx - dbx_dy) * ibp2 - 2.0f * (by * db_dx - bx * db_dy) * ibp3);
        dx_dt = f[R3_VX] + vdx + dkxx_dx + dkxy_dy + dkxz_dz;
        dy_dt = f[R3_VY] + vdy + dkxy_dx + dkyy_dy + dkyz_dz;
        dz_dt = f[R3_VZ] + vdz + dkxz_dx + dkyz_dy + dkzz_dz;
        divv = f[R3_DVXDX] + f[R3_DVYDY] + f[R3_DVZDZ];
    }
    float dp_dt = -pf * divv * (1.0f / 3.0f);
    float dpp = 0.0f;
    if constexpr (DPP) dpp_terms<NDIM>(m, kf, pf, f, kpara, knorm_para, b2, bx, by, bz, ib, divv, dp_dt, dpp);

    // the adaptive dt (the dt_min fallback when a rate is exactly zero), or
    // the fixed dt of a redo step
    const float sk = skperp > 0.0f ? skperp : skpara;
    const float isk = frcp(skpara);
    float a = m.f_half_dx * isk;
    float mm = a * a;
    a = m.f_half_dy * isk; mm = fminf(mm, a * a);
    if (NDIM == 3) { a = m.f_half_dz * isk; mm = fminf(mm, a * a); }
    a = sk * frcp(dx_dt); mm = fminf(mm, a * a);
    a = sk * frcp(dy_dt); mm = fminf(mm, a * a);
    if (NDIM == 3) { a = sk * frcp(dz_dt); mm = fminf(mm, a * a); }
    mm = fminf(mm, 0.1f * pf * frcp(fabsf(dp_dt)));
    const bool rates = dx_dt != 0.0f && dy_dt != 0.0f && dp_dt != 0.0f;
    mm = rates ? mm : m.f_dt_min;
    mm = fminf(fmaxf(mm, m.f_dt_min), m.f_dt_max);
    s.dt = fixed_dt ? s.dt : (double)mm;
    const float dtf32 = fixed_dt ? (float)s.dt : mm;
    const float sdt = fsqrt(dtf32);
    float deltax, deltay, deltaz;
    if (NDIM == 2) {
        const float a = r[2] * skpp * ib;
        deltax = dx_dt * dtf32 + (r[0] * skperp + a * bx) * sdt;
        deltay = dy_dt * dtf32 + (r[1] * skperp + a * by) * sdt;
        deltaz = dz_dt * dtf32;
    } else {
        const float sp = skperp * ibxyn;
        deltax = dx_dt * dtf32 + (bxn * skpara * r[0] - bxn * bzn * sp * r[1] - byn * sp * r[2]) * sdt;
        deltay = dy_dt * dtf32 + (byn * skpara * r[0] - byn * bzn * sp * r[1] + bxn * sp * r[2]) * sdt;
        deltaz = dz_dt * dtf32 + (bzn * skpara * r[0] + bxyn * skperp * r[1]) * sdt;
    }
    s.x += (double)deltax;
    s.y += (double)deltay;
    if (NDIM == 3 || GEN) s.z += (double)deltaz;   // 2D without the drift: deltaz = 0
    s.t += s.dt;
    double dp = DPP ? (double)(dp_dt * dtf32 + r[3] * fsqrt(2.0f * dpp) * sdt) : (double)(dp_dt * dtf32);
    if (ACC) {
        const double xn = (s.x - m.xmin) / m.lx;
        const double yn = (s.y - m.ymin) / m.ly;
        bool in = (xn >= m.acc_region[0]) && (xn <= m.acc_region[1]) && (yn >= m.acc_region[2]) &&
                  (yn <= m.acc_region[3]);
        if (NDIM == 3) {
            const double zn = (s.z - m.zmin) / m.lz;
            in = in && (zn >= m.acc_region[4]) && (zn <= m.acc_region[5]);
        }
        s.p = in ? s.p + dp : s.p;
        dp = in ? dp : 0.0;
    } else {
        s.p += dp;
    }
    // momentum floor: back out the increment, land on 0.25 p0
    const bool low = s.p < m.p_floor;
    const double pback = s.p - dp;
    dp = low ? m.p_floor - pback : dp;
    s.p = low ? m.p_floor : s.p;
    d[0] = (double)deltax;
    d[1] = (double)deltay;
    d[2] = NDIM == 2 ? 0.0 : (double)deltaz;
    d[3] = dp;
}

// (2u - 1) sqrt(3) in FP32 from one Philox block: u = r / (2^32 - 1).
__device__ __forceinline__ void step_rands(const MoverParams& m, uint32_t tag_i, uint32_t tag_s,
                                           uint32_t step, float r[4]) {
    const U4 u = philox4x32_10(tag_i, tag_s, step, m.stream, m.key0, m.key1);
    const float sc = (float)(2.0 * 1.7320508075688772 / 4294967295.0);
    const float of = -1.7320508075688772f;
    r[0] = (float)u.a * sc + of;
    r[1] = (float)u.b * sc + of;
    r[2] = (float)u.c * sc + of;
    r[3] = (float)u.d * sc + of;
}

template <int NDIM, bool TINTERP, int FX, bool HALF, class GC>
__device__ __forceinline__ void gather_kappa_push(const MoverParams& m, const KF& kf, State& s,
                                                  bool fixed_dt, const float r[4], double d[4],
                                                  GC& gc) {
    constexpr bool DPP = (FX & FX_DPP) != 0;
    static_assert(!(HALF && DPP), "FP16-gradient records carry no momentum-diffusion slots");
    float f[RecShape<NDIM, DPP>::NC];
    const Interp I = interp_params<NDIM>(m, s.x, s.y, s.z);
    const float rt = TINTERP ? (float)((s.t - m.t0) / m.dtf) : 0.0f;
    if constexpr (DPP) gather<NDIM, TINTERP, DPP>(m, I, rt, f);
    else if constexpr (NDIM == 2 && HALF) gather_pair2<TINTERP>(m, I, rt, f, gc);
    else if constexpr (NDIM == 2) gather_cell2<TINTERP>(m, I.rx, I.ry, I.ix, I.iy, rt, f, gc);
    else gather_node3<TINTERP, HALF>(m, I, rt, f, gc);
    kappa_push<NDIM, FX>(m, kf, s, f, fixed_dt, r, d);
}

}  // namespace f32
}  // namespace sp
