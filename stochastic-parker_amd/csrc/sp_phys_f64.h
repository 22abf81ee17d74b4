// sp_phys_f64.h -- FP64 gather, diffusion tensor and Parker step in the
// reference's operation order.  Included only by translation units built
// with -ffp-contract=off, so every +, -, *, / is one IEEE-rounded operation,
// as in the reference Fortran evaluated without contraction.  Constants
// precomputed on the host (MoverParams) are the first operation of the
// reference expression they replace, so the rounding sequence is unchanged.
#pragma once

#include "sp_device.h"

namespace sp {
namespace f64 {

// get_interp_paramters (particle_module.f90:642-674) on a uniform grid:
// 0-based node index of the lower corner and the corner weights.
struct Interp {
    int ix, iy, iz;
    double w[8];
};

template <int NDIM>
__device__ __forceinline__ Interp interp_params(const MoverParams& m, double x, double y, double z) {
    Interp I;
    const double px = (x - m.xmin) / m.dx;
    const double py = (y - m.ymin) / m.dy;
    const double fx = floor(px), fy = floor(py);
    const double rx = (px - (fx + 1.0)) + 1.0;  // px - pos + 1 with pos = floor(px) + 1
    const double ry = (py - (fy + 1.0)) + 1.0;
    double rz = 0.0;
    I.ix = clamp_cell(fx, m.nxg - 4) + 2;
    I.iy = clamp_cell(fy, m.nyg - 4) + 2;
    I.iz = 0;
    if (NDIM == 3) {
        const double pz = (z - m.zmin) / m.dz;
        const double fz = floor(pz);
        rz = (pz - (fz + 1.0)) + 1.0;
        I.iz = clamp_cell(fz, m.nzg - 4) + 2;
    }
    const double rx1 = 1.0 - rx, ry1 = 1.0 - ry, rz1 = 1.0 - rz;
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

// interp_fields (mhd_data_parallel.f90:1751-1793) over node records:
// fields = 0; for corners (k, j, i; i fastest) fields += f_node * w; then the
// optional (1-rt, rt) blend of the two frames.
template <int NDIM, bool DPP>
struct RecShape {
    static constexpr int W = NDIM == 2 ? (DPP ? kRec2DX : kRec2D) : (DPP ? kRec3DX : kRec3D);
    static constexpr int NC = NDIM == 2 ? (DPP ? 18 : 15) : (DPP ? 28 : 21);
};

template <int NDIM, bool TINTERP, bool DPP = false>
__device__ __forceinline__ void gather(const MoverParams& m, const Interp& I, double rt,
                                       double* __restrict__ f) {
    constexpr int W = RecShape<NDIM, DPP>::W;
    constexpr int NC = RecShape<NDIM, DPP>::NC;
    constexpr int NCORN = NDIM == 2 ? 4 : 8;
    double g[TINTERP ? NC : 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        f[c] = 0.0;
        if (TINTERP) g[c] = 0.0;
    }
    const long long sy = m.nxg, sz = (long long)m.nxg * m.nyg;
    const long long base = (long long)I.ix + (long long)I.iy * sy + (long long)I.iz * sz;
#pragma unroll
    for (int q = 0; q < NCORN; ++q) {
        const int i = q & 1, j = (q >> 1) & 1, k = (q >> 2) & 1;
        const long long node = NDIM == 3 ? brick_record(I.ix + i, I.iy + j, I.iz + k, m.bnx, m.bny)
                                         : base + i + j * sy + k * sz;
        const float4* r4 = reinterpret_cast<const float4*>(m.rec1 + node * W);
        float v[W];
#pragma unroll
        for (int u = 0; u < W / 4; ++u) {
            const float4 t = r4[u];
            v[4 * u] = t.x; v[4 * u + 1] = t.y; v[4 * u + 2] = t.z; v[4 * u + 3] = t.w;
        }
        const double wq = I.w[q];
#pragma unroll
        for (int c = 0; c < NC; ++c) f[c] = f[c] + (double)v[c] * wq;
        if (TINTERP) {
            const float4* s4 = reinterpret_cast<const float4*>(m.rec2 + node * W);
#pragma unroll
            for (int u = 0; u < W / 4; ++u) {
                const float4 t = s4[u];
                v[4 * u] = t.x; v[4 * u + 1] = t.y; v[4 * u + 2] = t.z; v[4 * u + 3] = t.w;
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) g[c] = g[c] + (double)v[c] * wq;
        }
    }
    if (TINTERP) {
#pragma unroll
        for (int c = 0; c < NC; ++c) f[c] = f[c] * (1.0 - rt) + g[c] * rt;
    }
}

struct Kappa {
    double knorm_para, kpara, kperp, skpara, skperp, skpara_perp;
    double dkxx_dx, dkyy_dy, dkzz_dz, dkxy_dx, dkxy_dy, dkxz_dx, dkxz_dz, dkyz_dy, dkyz_dz;
};

// calc_spatial_diffusion_coefficients (particle_module.f90:2208-2450),
// Cartesian Parker transport without delta-B / correlation-length maps.
// Only the quantities the pusher reads are kept (the diagonal/off-diagonal
// tensor entries feed the spherical and 3D-FT branches only).
template <int NDIM>
__device__ __forceinline__ Kappa kappa(const MoverParams& m, double p, const double* f) {
    Kappa k;
    const double bx = NDIM == 2 ? f[R2_BX] : f[R3_BX];
    const double by = NDIM == 2 ? f[R2_BY] : f[R3_BY];
    const double bz = NDIM == 2 ? f[R2_BZ] : f[R3_BZ];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    double ib1;
    if (b < __DBL_EPSILON__) ib1 = 1.0; else ib1 = 1.0 / b;
    const double ib2 = ib1 * ib1;
    const double ib3 = ib1 * ib2;
    const bool mag = (m.flags & MF_MAG_DEP) != 0;
    double knorm_para = 1.0;
    if (mag) knorm_para = knorm_para * pow(b, m.gamma_m2);
    double knorm;
    if (m.flags & MF_MOMENTUM_DEP) knorm = knorm_para * pow(p / m.p0, m.pindex);
    else knorm = knorm_para;
    k.knorm_para = knorm_para;
    k.kpara = m.kpara0 * knorm;
    k.kperp = k.kpara * m.kret;
    k.skpara = sqrt(2.0 * k.kpara);
    k.skperp = sqrt(2.0 * k.kperp);
    k.skpara_perp = sqrt(2.0 * (k.kpara - k.kperp));
    const double kpp = k.kpara - k.kperp;
    const double kperp = k.kperp;
    if (NDIM == 2) {
        const double dbx_dx = f[R2_DBXDX], dbx_dy = f[R2_DBXDY];
        const double dby_dx = f[R2_DBYDX], dby_dy = f[R2_DBYDY];
        const double db_dx = f[R2_DBDX], db_dy = f[R2_DBDY];
        double dkdx = 0.0, dkdy = 0.0;
        if (mag) {
            dkdx = db_dx * ib1 * m.gamma_m2;
            dkdy = db_dy * ib1 * m.gamma_m2;
        }
        k.dkxx_dx = kperp * dkdx + kpp * dkdx * (bx * bx) * ib2 +
                    2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k.dkyy_dy = kperp * dkdy + kpp * dkdy * (by * by) * ib2 +
                    2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k.dkxy_dx = kpp * dkdx * bx * by * ib2 +
                    kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k.dkxy_dy = kpp * dkdy * bx * by * ib2 +
                    kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k.dkzz_dz = k.dkxz_dx = k.dkxz_dz = k.dkyz_dy = k.dkyz_dz = 0.0;
    } else {
        const double dbx_dx = f[R3_DBXDX], dbx_dy = f[R3_DBXDY], dbx_dz = f[R3_DBXDZ];
        const double dby_dx = f[R3_DBYDX], dby_dy = f[R3_DBYDY], dby_dz = f[R3_DBYDZ];
        const double dbz_dx = f[R3_DBZDX], dbz_dy = f[R3_DBZDY], dbz_dz = f[R3_DBZDZ];
        const double db_dx = f[R3_DBDX], db_dy = f[R3_DBDY], db_dz = f[R3_DBDZ];
        double dkdx = 0.0, dkdy = 0.0, dkdz = 0.0;
        if (mag) {  // no 1/B in the 3D branch (particle_module.f90:2405-2409)
            dkdx = db_dx * m.gamma_m2;
            dkdy = db_dy * m.gamma_m2;
            dkdz = db_dz * m.gamma_m2;
        }
        k.dkxx_dx = kperp * dkdx + kpp * dkdx * (bx * bx) * ib2 +
                    2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k.dkyy_dy = kperp * dkdy + kpp * dkdy * (by * by) * ib2 +
                    2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k.dkzz_dz = kperp * dkdz + kpp * dkdz * (bz * bz) * ib2 +
                    2.0 * kpp * bz * (dbz_dz * b - bz * db_dz) * ib3;
        k.dkxy_dx = kpp * dkdx * bx * by * ib2 +
                    kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k.dkxy_dy = kpp * dkdy * bx * by * ib2 +
                    kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k.dkxz_dx = kpp * dkdx * bx * bz * ib2 +
                    kpp * ((dbx_dx * bz + bx * dbz_dx) * ib2 - 2.0 * bx * bz * db_dx * ib3);
        k.dkxz_dz = kpp * dkdz * bx * bz * ib2 +
                    kpp * ((dbx_dz * bz + bx * dbz_dz) * ib2 - 2.0 * bx * bz * db_dz * ib3);
        k.dkyz_dy = kpp * dkdy * by * bz * ib2 +
                    kpp * ((dby_dy * bz + by * dbz_dy) * ib2 - 2.0 * by * bz * db_dy * ib3);
        k.dkyz_dz = kpp * dkdz * by * bz * ib2 +
                    kpp * ((dby_dz * bz + by * dbz_dz) * ib2 - 2.0 * by * bz * db_dz * ib3);
    }
    return k;
}

// calc_spatial_diffusion_coefficients_nlgc (particle_module.f90:2464-2771),
// Cartesian Parker transport without delta-B / correlation-length maps: kperp
// from the NLGC scaling, kpara0 * kperp_kpara * B^((gamma-2)/3) *
// (p/p0)^((5-gamma)/3) * mu^2, and separate logarithmic gradients of kpara and
// kperp in the divergence terms (both with 1/B, in 2D and 3D).
template <int NDIM>
__device__ __forceinline__ Kappa kappa_nlgc(const MoverParams& m, double p, double mu,
                                            const double* f) {
    Kappa k;
    const double bx = NDIM == 2 ? f[R2_BX] : f[R3_BX];
    const double by = NDIM == 2 ? f[R2_BY] : f[R3_BY];
    const double bz = NDIM == 2 ? f[R2_BZ] : f[R3_BZ];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    double ib1;
    if (b < __DBL_EPSILON__) ib1 = 1.0; else ib1 = 1.0 / b;
    const double ib2 = ib1 * ib1;
    const double ib3 = ib1 * ib2;
    const bool mag = (m.flags & MF_MAG_DEP) != 0;
    double knorm_para = 1.0, knorm_perp = 1.0;
    if (mag) {
        knorm_para = knorm_para * pow(b, m.gamma_m2);
        knorm_perp = knorm_perp * pow(b, m.perp_mag);
    }
    double kn_para, kn_perp;
    if (m.flags & MF_MOMENTUM_DEP) {
        kn_para = knorm_para * pow(p / m.p0, m.pindex);
        kn_perp = knorm_perp * pow(p / m.p0, m.perp_mom);
    } else {
        kn_para = knorm_para;
        kn_perp = knorm_perp;
    }
    k.knorm_para = knorm_para;
    k.kpara = m.kpara0 * kn_para;
    k.kperp = m.kpara0 * m.kperp_kpara * kn_perp * (mu * mu);
    k.skpara = sqrt(2.0 * k.kpara);
    k.skperp = sqrt(2.0 * k.kperp);
    k.skpara_perp = sqrt(2.0 * (k.kpara - k.kperp));
    const double kpara = k.kpara, kperp = k.kperp;
    const double kpp = kpara - kperp;
    if (NDIM == 2) {
        const double dbx_dx = f[R2_DBXDX], dbx_dy = f[R2_DBXDY];
        const double dby_dx = f[R2_DBYDX], dby_dy = f[R2_DBYDY];
        const double db_dx = f[R2_DBDX], db_dy = f[R2_DBDY];
        double dpa_x = 0.0, dpa_y = 0.0, dpe_x = 0.0, dpe_y = 0.0;
        if (mag) {
            dpa_x = db_dx * ib1 * m.gamma_m2;
            dpa_y = db_dy * ib1 * m.gamma_m2;
            dpe_x = db_dx * ib1 * m.gamma_m2 / 3.0;
            dpe_y = db_dy * ib1 * m.gamma_m2 / 3.0;
        }
        const double ax = kpara * dpa_x - kperp * dpe_x;
        const double ay = kpara * dpa_y - kperp * dpe_y;
        k.dkxx_dx = kperp * dpe_x + ax * (bx * bx) * ib2 + 2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k.dkyy_dy = kperp * dpe_y + ay * (by * by) * ib2 + 2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k.dkxy_dx = ax * bx * by * ib2 + kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k.dkxy_dy = ay * bx * by * ib2 + kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k.dkzz_dz = k.dkxz_dx = k.dkxz_dz = k.dkyz_dy = k.dkyz_dz = 0.0;
    } else {
        const double dbx_dx = f[R3_DBXDX], dbx_dy = f[R3_DBXDY], dbx_dz = f[R3_DBXDZ];
        const double dby_dx = f[R3_DBYDX], dby_dy = f[R3_DBYDY], dby_dz = f[R3_DBYDZ];
        const double dbz_dx = f[R3_DBZDX], dbz_dy = f[R3_DBZDY], dbz_dz = f[R3_DBZDZ];
        const double db_dx = f[R3_DBDX], db_dy = f[R3_DBDY], db_dz = f[R3_DBDZ];
        double dpa_x = 0.0, dpa_y = 0.0, dpa_z = 0.0, dpe_x = 0.0, dpe_y = 0.0, dpe_z = 0.0;
        if (mag) {
            dpa_x = db_dx * ib1 * m.gamma_m2;
            dpa_y = db_dy * ib1 * m.gamma_m2;
            dpa_z = db_dz * ib1 * m.gamma_m2;
            dpe_x = db_dx * ib1 * m.gamma_m2 / 3.0;
            dpe_y = db_dy * ib1 * m.gamma_m2 / 3.0;
            dpe_z = db_dz * ib1 * m.gamma_m2 / 3.0;
        }
        const double ax = kpara * dpa_x - kperp * dpe_x;
        const double ay = kpara * dpa_y - kperp * dpe_y;
        const double az = kpara * dpa_z - kperp * dpe_z;
        k.dkxx_dx = kperp * dpe_x + ax * (bx * bx) * ib2 + 2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k.dkyy_dy = kperp * dpe_y + ay * (by * by) * ib2 + 2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k.dkzz_dz = kperp * dpe_z + az * (bz * bz) * ib2 + 2.0 * kpp * bz * (dbz_dz * b - bz * db_dz) * ib3;
        k.dkxy_dx = ax * bx * by * ib2 + kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k.dkxy_dy = ay * bx * by * ib2 + kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k.dkxz_dx = ax * bx * bz * ib2 + kpp * ((dbx_dx * bz + bx * dbz_dx) * ib2 - 2.0 * bx * bz * db_dx * ib3);
        k.dkxz_dz = az * bx * bz * ib2 + kpp * ((dbx_dz * bz + bx * dbz_dz) * ib2 - 2.0 * bx * bz * db_dz * ib3);
        k.dkyz_dy = ay * by * bz * ib2 + kpp * ((dby_dy * bz + by * dbz_dy) * ib2 - 2.0 * by * bz * db_dy * ib3);
        k.dkyz_dz = az * by * bz * ib2 + kpp * ((dby_dz * bz + by * dbz_dz) * ib2 - 2.0 * by * bz * db_dz * ib3);
    }
    return k;
}

// particle_in_acceleration_region (particle_module.f90:2885-2906)
template <int NDIM>
__device__ __forceinline__ bool in_acc_region(const MoverParams& m, double x, double y, double z) {
    const double xn = (x - m.xmin) / m.lx;
    const double yn = (y - m.ymin) / m.ly;
    bool in = (xn >= m.acc_region[0]) && (xn <= m.acc_region[1]) && (yn >= m.acc_region[2]) &&
              (yn <= m.acc_region[3]);
    if (NDIM == 3) {
        const double zn = (z - m.zmin) / m.lz;
        in = in && (zn >= m.acc_region[4]) && (zn <= m.acc_region[5]);
    }
    return in;
}

__device__ __forceinline__ double fmin_seq(double a, double b) { return b < a ? b : a; }

// Momentum diffusion added to dp_dt and dpp: calc_dpp_wave_scattering
// (particle_module.f90:2918-2933) with va = B / sqrt(rho), and
// calc_dpp_flow_shear (:2947-2979) on the shear tensor of push_particle_2d
// (:3487-3500; sigma_zz = -divv/3, no xz / yz terms) or push_particle_3d
// (:4772-4787).
template <int NDIM>
__device__ __forceinline__ void dpp_terms(const MoverParams& m, double p, const double* f,
                                          const Kappa& k, double b, double bx, double by,
                                          double bz, double divv, double& dp_dt, double& dpp) {
    if (m.flags & MF_DPP_WAVE) {
        const double rho = NDIM == 2 ? f[R2_RHO] : f[R3_RHO];
        const double va = b / sqrt(rho);
        if (m.flags & MF_MOMENTUM_DEP) dp_dt = dp_dt + (8.0 * p / (27.0 * k.kpara)) * (va * va);
        else dp_dt = dp_dt + (4.0 * p / (9.0 * k.kpara)) * (va * va);
        dpp = dpp + ((p * va) * (p * va)) / (9.0 * k.kpara);
    }
    if (m.flags & MF_DPP_SHEAR) {
        double sxx, syy, szz, sxy, sxz, syz;
        if (NDIM == 2) {
            sxx = f[R2_DVXDX] - divv / 3.0;
            syy = f[R2_DVYDY] - divv / 3.0;
            szz = -(divv / 3.0);
            sxy = (f[R2_DVXDY] + f[R2_DVYDX]) / 2.0;
            sxz = 0.0;
            syz = 0.0;
        } else {
            sxx = f[R3_DVXDX] - divv / 3.0;
            syy = f[R3_DVYDY] - divv / 3.0;
            szz = f[R3_DVZDZ] - divv / 3.0;
            sxy = (f[R3_DVXDY] + f[R3_DVYDX]) / 2.0;
            sxz = (f[R3_DVXDZ] + f[R3_DVZDX]) / 2.0;
            syz = (f[R3_DVYDZ] + f[R3_DVZDY]) / 2.0;
        }
        double gshear;
        if (m.flags & MF_WEAK_SCATTERING) {
            double ib;
            if (b < __DBL_EPSILON__) ib = 0.0; else ib = 1.0 / b;
            double bbs = sxx * (bx * bx) + syy * (by * by) + szz * (bz * bz) +
                         2.0 * (sxy * bx * by + sxz * bx * bz + syz * by * bz);
            bbs = bbs * ib * ib;
            gshear = bbs * bbs / 5.0;
        } else {
            gshear = 2.0 * (sxx * sxx + syy * syy + szz * szz +
                            2.0 * (sxy * sxy + sxz * sxz + syz * syz)) / 15.0;
        }
        if (gshear > 0.0) {
            dp_dt = dp_dt + (2.0 + m.pindex) * gshear * m.tau0 * k.knorm_para *
                                pow(p, m.pindex - 1.0) * m.p0_2mpi;
            dpp = dpp + gshear * m.tau0 * k.knorm_para * pow(p, m.pindex) * m.p0_2mpi;
        }
    }
}

// push_particle_2d (particle_module.f90:3358-3606) / push_particle_3d
// (:4625-4907), Cartesian, uniform grid; the momentum-diffusion terms when FX
// has FX_DPP.  r = (ran1, ran2, ran3, ran_p).  d receives (deltax, deltay,
// deltaz, deltap); the 2D pusher leaves deltaz at zero like the reference (its
// z drift is not undone).
template <int NDIM, int FX>
__device__ __forceinline__ void push(const MoverParams& m, State& s, const double* f,
                                     const Kappa& k, bool fixed_dt, const double r[4],
                                     double d[4]) {
    constexpr bool ACC = (FX & FX_ACC) != 0;
    const double vx = NDIM == 2 ? f[R2_VX] : f[R3_VX];
    const double vy = NDIM == 2 ? f[R2_VY] : f[R3_VY];
    const double bx = NDIM == 2 ? f[R2_BX] : f[R3_BX];
    const double by = NDIM == 2 ? f[R2_BY] : f[R3_BY];
    const double bz = NDIM == 2 ? f[R2_BZ] : f[R3_BZ];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    double ib;
    if (b < __DBL_EPSILON__) ib = 0.0; else ib = 1.0 / b;
    const double ib2 = ib * ib;
    const double ib3 = ib * ib2;
    // vdp = 1.0/(3*pcharge) / sqrt((drift1*p0/p)**2 + (drift2*p0**2/p**2)**2)
    const double t1 = m.d1p0 / s.p;
    const double t2 = m.d2p02 / (s.p * s.p);
    const double vdp = m.vdp_pre / sqrt(t1 * t1 + t2 * t2);

    double dx_dt, dy_dt, dz_dt, divv;
    double bxn = 0.0, byn = 0.0, bzn = 0.0, bxyn = 0.0, ibxyn = 0.0;
    if (NDIM == 2) {
        const double dbz_dx = f[R2_DBZDX], dbz_dy = f[R2_DBZDY];
        const double db_dx = f[R2_DBDX], db_dy = f[R2_DBDY];
        const double vdx = vdp * (dbz_dy * ib2 - 2.0 * bz * db_dy * ib3);
        const double vdy = vdp * (-(dbz_dx * ib2) + 2.0 * bz * db_dx * ib3);
        double vdz = 0.0;
        if (m.flags & MF_DRIFT_2D) {
            const double dbx_dy = f[R2_DBXDY], dby_dx = f[R2_DBYDX];
            vdz = vdp * ((dby_dx - dbx_dy) * ib2 - 2.0 * (by * db_dx - bx * db_dy) * ib3);
        }
        dx_dt = vx + vdx + k.dkxx_dx + k.dkxy_dy;
        dy_dt = vy + vdy + k.dkxy_dx + k.dkyy_dy;
        dz_dt = vdz;
        divv = f[R2_DVXDX] + f[R2_DVYDY];
    } else {
        const double vz = f[R3_VZ];
        bxn = bx * ib; byn = by * ib; bzn = bz * ib;
        bxyn = sqrt(bxn * bxn + byn * byn);
        if (bxyn < __DBL_EPSILON__) ibxyn = 0.0; else ibxyn = 1.0 / bxyn;
        const double dbx_dy = f[R3_DBXDY], dbx_dz = f[R3_DBXDZ];
        const double dby_dx = f[R3_DBYDX], dby_dz = f[R3_DBYDZ];
        const double dbz_dx = f[R3_DBZDX], dbz_dy = f[R3_DBZDY];
        const double db_dx = f[R3_DBDX], db_dy = f[R3_DBDY], db_dz = f[R3_DBDZ];
        const double vdx = vdp * ((dbz_dy - dby_dz) * ib2 - 2.0 * (bz * db_dy - by * db_dz) * ib3);
        const double vdy = vdp * ((dbx_dz - dbz_dx) * ib2 - 2.0 * (bx * db_dz - bz * db_dx) * ib3);
        const double vdz = vdp * ((dby_dx - dbx_dy) * ib2 - 2.0 * (by * db_dx - bx * db_dy) * ib3);
        dx_dt = vx + vdx + k.dkxx_dx + k.dkxy_dy + k.dkxz_dz;
        dy_dt = vy + vdy + k.dkxy_dx + k.dkyy_dy + k.dkyz_dz;
        dz_dt = vz + vdz + k.dkxz_dx + k.dkyz_dy + k.dkzz_dz;
        divv = f[R3_DVXDX] + f[R3_DVYDY] + f[R3_DVZDZ];
    }
    double dp_dt = -(s.p * divv / 3.0);
    double dpp = 0.0;
    if constexpr ((FX & FX_DPP) != 0) dpp_terms<NDIM>(m, s.p, f, k, b, bx, by, bz, divv, dp_dt, dpp);

    if (!fixed_dt) {
        // the zero test omits dz_dt in 3D (particle_module.f90:4792-4794)
        if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0) {
            const double sk = k.skperp > 0.0 ? k.skperp : k.skpara;
            double a = m.half_dx / k.skpara;
            double mm = a * a;
            a = m.half_dy / k.skpara; mm = fmin_seq(mm, a * a);
            if (NDIM == 3) { a = m.half_dz / k.skpara; mm = fmin_seq(mm, a * a); }
            a = sk / dx_dt; mm = fmin_seq(mm, a * a);
            a = sk / dy_dt; mm = fmin_seq(mm, a * a);
            if (NDIM == 3) { a = sk / dz_dt; mm = fmin_seq(mm, a * a); }
            mm = fmin_seq(mm, (double)0.1f * s.p / fabs(dp_dt));  // 0.1 is an FP32 literal
            s.dt = mm;
        } else {
            s.dt = m.dt_min;
        }
        if (s.dt < m.dt_min) s.dt = m.dt_min;
        if (s.dt > m.dt_max) s.dt = m.dt_max;
    }

    const double sdt = sqrt(s.dt);
    double deltax, deltay, deltaz;
    if (NDIM == 2) {
        deltax = dx_dt * s.dt + r[0] * k.skperp * sdt + r[2] * k.skpara_perp * sdt * bx * ib;
        deltay = dy_dt * s.dt + r[1] * k.skperp * sdt + r[2] * k.skpara_perp * sdt * by * ib;
        deltaz = dz_dt * s.dt;
    } else {
        deltax = dx_dt * s.dt + (bxn * k.skpara * r[0] - bxn * bzn * k.skperp * ibxyn * r[1] -
                                 byn * k.skperp * ibxyn * r[2]) * sdt;
        deltay = dy_dt * s.dt + (byn * k.skpara * r[0] - byn * bzn * k.skperp * ibxyn * r[1] +
                                 bxn * k.skperp * ibxyn * r[2]) * sdt;
        deltaz = dz_dt * s.dt + (bzn * k.skpara * r[0] + bxyn * k.skperp * r[1]) * sdt;
    }
    s.x = s.x + deltax;
    s.y = s.y + deltay;
    s.z = s.z + deltaz;
    s.t = s.t + s.dt;
    double dp = dp_dt * s.dt + r[3] * sqrt(2.0 * dpp) * sdt;
    if (ACC) {
        if (in_acc_region<NDIM>(m, s.x, s.y, s.z)) s.p = s.p + dp;
        else dp = 0.0;
    } else {
        s.p = s.p + dp;
    }
    if (s.p < m.p_floor) {
        s.p = s.p - dp;
        dp = m.p_floor - s.p;
        s.p = m.p_floor;
    }
    d[0] = deltax;
    d[1] = deltay;
    d[2] = NDIM == 2 ? 0.0 : deltaz;
    d[3] = dp;
}

// Steps' increments from one Philox block: (2u - 1) * sqrt(3) for each
// of the four words (particle_module.f90:3548-3550, :3589).
__device__ __forceinline__ void step_rands(const MoverParams& m, uint32_t tag_i, uint32_t tag_s,
                                           uint32_t step, double r[4]) {
    const U4 u = philox4x32_10(tag_i, tag_s, step, m.stream, m.key0, m.key1);
    const double sqrt3 = sqrt(3.0);
    r[0] = (2.0 * u01(u.a) - 1.0) * sqrt3;
    r[1] = (2.0 * u01(u.b) - 1.0) * sqrt3;
    r[2] = (2.0 * u01(u.c) - 1.0) * sqrt3;
    r[3] = (2.0 * u01(u.d) - 1.0) * sqrt3;
}

// The full per-step chain: interpolation parameters, gather, kappa, push.
template <int NDIM, bool TINTERP, int FX>
__device__ __forceinline__ void gather_kappa_push(const MoverParams& m, State& s, bool fixed_dt,
                                                  const double r[4], double d[4]) {
    constexpr bool DPP = (FX & FX_DPP) != 0;
    double f[RecShape<NDIM, DPP>::NC];
    const Interp I = interp_params<NDIM>(m, s.x, s.y, s.z);
    const double rt = TINTERP ? (s.t - m.t0) / m.dtf : 0.0;
    gather<NDIM, TINTERP, DPP>(m, I, rt, f);
    Kappa k;
    if constexpr ((FX & FX_NLGC) != 0) k = kappa_nlgc<NDIM>(m, s.p, s.mu, f);
    else k = kappa<NDIM>(m, s.p, f);
    push<NDIM, FX>(m, s, f, k, fixed_dt, r, d);
}

}  // namespace f64
}  // namespace sp
