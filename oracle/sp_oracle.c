/*
 * sp_oracle.c -- CPU restatement of the stochastic-parker Parker pusher.
 *
 * TEST INFRASTRUCTURE ONLY (see sp_oracle.h).  Built with gcc
 * -ffp-contract=off; never linked into the product library.
 *
 * Every function states the reference lines it follows.  Reference paths are
 * relative to the xiaocanli/stochastic-parker repository root.  Expressions
 * keep the Fortran evaluation order (left to right for operators of equal
 * precedence, ** before * and /) and its literal kinds: a literal written
 * without _dp in the Fortran (e.g. 0.1, 1.0/(3*pcharge)) is a single-precision
 * value and is reproduced as such.
 */
#include "sp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_sizeof_particle(void) { return (int)sizeof(orc_particle); }

/* ======================================================================
 * Random numbers
 * ====================================================================== */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with
 * multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
 * 0xBB67AE85.  Replaces mt_stream (random_number_generator.f90:28-101). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* genrand_double1 / genrand_real1: u32 * (1/(2^32-1)), closed [0,1]
 * (random_number_generator.f90:97-101). */
double orc_u01(uint32_t r) { return (double)r * (1.0 / 4294967295.0); }

/* MT19937 (Matsumoto & Nishimura 1998, mt19937ar.c reference algorithm). */
void orc_mt_init_genrand(orc_mt19937* s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->mti = 624;
}

void orc_mt_init_by_array(orc_mt19937* s, const uint32_t* key, int len)
{
    orc_mt_init_genrand(s, 19650218u);
    int i = 1, j = 0;
    int k = 624 > len ? 624 : len;
    for (; k; --k) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        ++i; ++j;
        if (i >= 624) { s->mt[0] = s->mt[623]; i = 1; }
        if (j >= len) j = 0;
    }
    for (k = 623; k; --k) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        ++i;
        if (i >= 624) { s->mt[0] = s->mt[623]; i = 1; }
    }
    s->mt[0] = 0x80000000u;
    s->mti = 624;
}

uint32_t orc_mt_genrand_int32(orc_mt19937* s)
{
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    uint32_t y;
    if (s->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; ++kk) {
            y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
            s->mt[kk] = s->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < 623; ++kk) {
            y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
            s->mt[kk] = s->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (s->mt[623] & 0x80000000u) | (s->mt[0] & 0x7fffffffu);
        s->mt[623] = s->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        s->mti = 0;
    }
    y = s->mt[s->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* Per-particle draw stream for injection: block b of Philox with counter
 * (tag_injected, 0, b, stream).  Step draws use counter word 1 =
 * tag_splitted >= 1, so the two families never collide. */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used;
} orc_drawstream;

static void ds_init(orc_drawstream* d, const orc_params* prm, uint32_t tag)
{
    d->key[0] = (uint32_t)prm->seed;
    d->key[1] = (uint32_t)(prm->seed >> 32);
    d->ctr[0] = tag; d->ctr[1] = 0u; d->ctr[2] = 0u; d->ctr[3] = prm->stream;
    d->used = 4;
}

static double ds_next(orc_drawstream* d)
{
    if (d->used == 4) {
        orc_philox4x32_10(d->ctr, d->key, d->buf);
        d->ctr[2] += 1u;
        d->used = 0;
    }
    return orc_u01(d->buf[d->used++]);
}

/* ======================================================================
 * Fields
 * ====================================================================== */

static int nxg_of(const orc_params* prm) { return prm->nx + 4; }
static int nyg_of(const orc_params* prm) { return prm->ndim > 1 ? prm->ny + 4 : prm->ny; }
static int nzg_of(const orc_params* prm) { return prm->ndim > 2 ? prm->nz + 4 : prm->nz; }

int64_t orc_field_nodes(const orc_params* prm)
{
    return (int64_t)nxg_of(prm) * nyg_of(prm) * nzg_of(prm);
}

#define FA(arr, c, i, j, k) (arr)[(size_t)(c) + 32u * ((size_t)(i) + (size_t)nxg * ((size_t)(j) + (size_t)nyg * (size_t)(k)))]

/* calc_fields_gradients, mhd_data_parallel.f90:504-605 (uniform grid).
 * Central differences on the ghost-inclusive interior, one-sided second
 * order at the two outermost nodes.  The difference is formed in FP32, scaled
 * by the FP64 factor 0.5/dx and stored back as FP32.  Component layout:
 * gradient of field c along axis a goes to 8 + 3c + a. */
void orc_calc_fields_gradients(const orc_params* prm, float* farray)
{
    const int nxg = nxg_of(prm), nyg = nyg_of(prm), nzg = nzg_of(prm);
    const double idxh = 0.5 / prm->dx, idyh = 0.5 / prm->dy, idzh = 0.5 / prm->dz;
    for (int k = 0; k < nzg; ++k)
        for (int j = 0; j < nyg; ++j)
            for (int c = 0; c < 8; ++c) {
                for (int i = 1; i < nxg - 1; ++i) {
                    float d = FA(farray, c, i + 1, j, k) - FA(farray, c, i - 1, j, k);
                    FA(farray, 8 + 3 * c, i, j, k) = (float)((double)d * idxh);
                }
                {
                    float d = (-3.0f * FA(farray, c, 0, j, k) + 4.0f * FA(farray, c, 1, j, k)) - FA(farray, c, 2, j, k);
                    FA(farray, 8 + 3 * c, 0, j, k) = (float)((double)d * idxh);
                }
                {
                    const int u = nxg - 1;
                    float d = (3.0f * FA(farray, c, u, j, k) - 4.0f * FA(farray, c, u - 1, j, k)) + FA(farray, c, u - 2, j, k);
                    FA(farray, 8 + 3 * c, u, j, k) = (float)((double)d * idxh);
                }
            }
    if (nyg > 1) {
        for (int k = 0; k < nzg; ++k)
            for (int i = 0; i < nxg; ++i)
                for (int c = 0; c < 8; ++c) {
                    for (int j = 1; j < nyg - 1; ++j) {
                        float d = FA(farray, c, i, j + 1, k) - FA(farray, c, i, j - 1, k);
                        FA(farray, 9 + 3 * c, i, j, k) = (float)((double)d * idyh);
                    }
                    {
                        float d = (-3.0f * FA(farray, c, i, 0, k) + 4.0f * FA(farray, c, i, 1, k)) - FA(farray, c, i, 2, k);
                        FA(farray, 9 + 3 * c, i, 0, k) = (float)((double)d * idyh);
                    }
                    {
                        const int u = nyg - 1;
                        float d = (3.0f * FA(farray, c, i, u, k) - 4.0f * FA(farray, c, i, u - 1, k)) + FA(farray, c, i, u - 2, k);
                        FA(farray, 9 + 3 * c, i, u, k) = (float)((double)d * idyh);
                    }
                }
    }
    if (nzg > 1) {
        for (int j = 0; j < nyg; ++j)
            for (int i = 0; i < nxg; ++i)
                for (int c = 0; c < 8; ++c) {
                    for (int k = 1; k < nzg - 1; ++k) {
                        float d = FA(farray, c, i, j, k + 1) - FA(farray, c, i, j, k - 1);
                        FA(farray, 10 + 3 * c, i, j, k) = (float)((double)d * idzh);
                    }
                    {
                        float d = (-3.0f * FA(farray, c, i, j, 0) + 4.0f * FA(farray, c, i, j, 1)) - FA(farray, c, i, j, 2);
                        FA(farray, 10 + 3 * c, i, j, 0) = (float)((double)d * idzh);
                    }
                    {
                        const int u = nzg - 1;
                        float d = (3.0f * FA(farray, c, i, j, u) - 4.0f * FA(farray, c, i, j, u - 1)) + FA(farray, c, i, j, u - 2);
                        FA(farray, 10 + 3 * c, i, j, u) = (float)((double)d * idzh);
                    }
                }
    }
}

/* get_interp_paramters, particle_module.f90:642-674.  pos is the Fortran
 * (1-based, ghost index -1 ... ) lower corner; the weight order is i fastest,
 * then j, then k.  r = p - pos + 1 is kept in that operation order. */
void orc_get_interp_paramters(const orc_params* prm, double px, double py, double pz,
                              int pos[3], double weights[8])
{
    double rx, ry, rz;
    if (prm->ndim == 2) {
        pos[0] = (int)floor(px) + 1;
        pos[1] = (int)floor(py) + 1;
        pos[2] = 1;
        ry = (py - (double)pos[1]) + 1.0;
        rz = 0.0;
    } else {
        pos[0] = (int)floor(px) + 1;
        pos[1] = (int)floor(py) + 1;
        pos[2] = (int)floor(pz) + 1;
        ry = (py - (double)pos[1]) + 1.0;
        rz = (pz - (double)pos[2]) + 1.0;
    }
    rx = (px - (double)pos[0]) + 1.0;
    const double rx1 = 1.0 - rx, ry1 = 1.0 - ry, rz1 = 1.0 - rz;
    weights[0] = rx1 * ry1 * rz1;
    weights[1] = rx * ry1 * rz1;
    weights[2] = rx1 * ry * rz1;
    weights[3] = rx * ry * rz1;
    weights[4] = rx1 * ry1 * rz;
    weights[5] = rx * ry1 * rz;
    weights[6] = rx1 * ry * rz;
    weights[7] = rx * ry * rz;
}

/* interp_fields, mhd_data_parallel.f90:1751-1793: sum over corners (k, j, i
 * loops, i fastest) of FP32 node values times FP64 weights, then the optional
 * linear time interpolation between the two frames. */
void orc_interp_fields(const orc_params* prm, const float* farray1, const float* farray2,
                       const int pos[3], const double weights[8], double rt,
                       double fields[32])
{
    const int nxg = nxg_of(prm), nyg = nyg_of(prm);
    const int ye = prm->ndim > 1 ? 1 : 0, ze = prm->ndim > 2 ? 1 : 0;
    /* Fortran index -1 maps to 0 */
    const int ix = pos[0] + 1, iy = prm->ndim > 1 ? pos[1] + 1 : 0;
    const int iz = prm->ndim > 2 ? pos[2] + 1 : 0;
    double f2[32];
    for (int c = 0; c < 32; ++c) { fields[c] = 0.0; f2[c] = 0.0; }
    for (int k = 0; k <= ze; ++k)
        for (int j = 0; j <= ye; ++j)
            for (int i = 0; i <= 1; ++i) {
                const int w = k * 4 + j * 2 + i;
                for (int c = 0; c < 32; ++c) {
                    fields[c] = fields[c] + (double)FA(farray1, c, ix + i, iy + j, iz + k) * weights[w];
                    if (prm->time_interp)
                        f2[c] = f2[c] + (double)FA(farray2, c, ix + i, iy + j, iz + k) * weights[w];
                }
            }
    if (prm->time_interp)
        for (int c = 0; c < 32; ++c) fields[c] = fields[c] * (1.0 - rt) + f2[c] * rt;
}

/* ======================================================================
 * Diffusion tensor: calc_spatial_diffusion_coefficients,
 * particle_module.f90:2208-2450 (Cartesian, Parker, no delta-B or
 * correlation-length maps).
 * ====================================================================== */
void orc_kappa_standard(const orc_params* prm, double p, const double f[32], orc_kappa* k)
{
    const double bx = f[4], by = f[5], bz = f[6];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    double ib1;
    if (b < DBL_EPSILON) ib1 = 1.0; else ib1 = 1.0 / b;
    const double ib2 = ib1 * ib1;
    const double ib3 = ib1 * ib2;

    k->knorm_para = 1.0;
    k->knorm_perp = 1.0;
    if (prm->mag_dependency == 1) k->knorm_para = k->knorm_para * pow(b, prm->gamma_turb - 2.0);
    double knorm;
    if (prm->momentum_dependency == 1) knorm = k->knorm_para * pow(p / prm->p0, prm->pindex);
    else knorm = k->knorm_para;
    k->knorm_perp = k->knorm_para;
    k->kpara = prm->kpara0 * knorm;
    k->kperp = k->kpara * prm->kret;
    k->skpara = sqrt(2.0 * k->kpara);
    k->skperp = sqrt(2.0 * k->kperp);
    k->skpara_perp = sqrt(2.0 * (k->kpara - k->kperp));

    const double kpp = k->kpara - k->kperp;
    const double kperp = k->kperp;
    if (prm->ndim == 2) {
        const double dbx_dx = f[20], dbx_dy = f[21];
        const double dby_dx = f[23], dby_dy = f[24];
        const double db_dx = f[29], db_dy = f[30];
        double dkdx = 0.0, dkdy = 0.0;
        if (prm->mag_dependency == 1) {
            dkdx = db_dx * ib1 * (prm->gamma_turb - 2.0);
            dkdy = db_dy * ib1 * (prm->gamma_turb - 2.0);
        }
        k->dkxx_dx = kperp * dkdx + kpp * dkdx * (bx * bx) * ib2 +
                     2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k->dkyy_dy = kperp * dkdy + kpp * dkdy * (by * by) * ib2 +
                     2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k->dkxy_dx = kpp * dkdx * bx * by * ib2 +
                     kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k->dkxy_dy = kpp * dkdy * bx * by * ib2 +
                     kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k->kxx = kperp + kpp * bx * bx * ib2;
        k->kyy = kperp + kpp * by * by * ib2;
        k->kxy = kpp * bx * by * ib2;
        k->kzz = k->kxz = k->kyz = 0.0;
        k->dkzz_dz = k->dkxz_dx = k->dkxz_dz = k->dkyz_dy = k->dkyz_dz = 0.0;
    } else {
        const double dbx_dx = f[20], dbx_dy = f[21], dbx_dz = f[22];
        const double dby_dx = f[23], dby_dy = f[24], dby_dz = f[25];
        const double dbz_dx = f[26], dbz_dy = f[27], dbz_dz = f[28];
        const double db_dx = f[29], db_dy = f[30], db_dz = f[31];
        double dkdx = 0.0, dkdy = 0.0, dkdz = 0.0;
        if (prm->mag_dependency == 1) {
            /* the 3D branch carries no 1/B factor (particle_module.f90:2405-2409) */
            dkdx = db_dx * (prm->gamma_turb - 2.0);
            dkdy = db_dy * (prm->gamma_turb - 2.0);
            dkdz = db_dz * (prm->gamma_turb - 2.0);
        }
        k->dkxx_dx = kperp * dkdx + kpp * dkdx * (bx * bx) * ib2 +
                     2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
        k->dkyy_dy = kperp * dkdy + kpp * dkdy * (by * by) * ib2 +
                     2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
        k->dkzz_dz = kperp * dkdz + kpp * dkdz * (bz * bz) * ib2 +
                     2.0 * kpp * bz * (dbz_dz * b - bz * db_dz) * ib3;
        k->dkxy_dx = kpp * dkdx * bx * by * ib2 +
                     kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
        k->dkxy_dy = kpp * dkdy * bx * by * ib2 +
                     kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
        k->dkxz_dx = kpp * dkdx * bx * bz * ib2 +
                     kpp * ((dbx_dx * bz + bx * dbz_dx) * ib2 - 2.0 * bx * bz * db_dx * ib3);
        k->dkxz_dz = kpp * dkdz * bx * bz * ib2 +
                     kpp * ((dbx_dz * bz + bx * dbz_dz) * ib2 - 2.0 * bx * bz * db_dz * ib3);
        k->dkyz_dy = kpp * dkdy * by * bz * ib2 +
                     kpp * ((dby_dy * bz + by * dbz_dy) * ib2 - 2.0 * by * bz * db_dy * ib3);
        k->dkyz_dz = kpp * dkdz * by * bz * ib2 +
                     kpp * ((dby_dz * bz + by * dbz_dz) * ib2 - 2.0 * by * bz * db_dz * ib3);
        k->kxx = kperp + kpp * bx * bx * ib2;
        k->kyy = kperp + kpp * by * by * ib2;
        k->kzz = kperp + kpp * bz * bz * ib2;
        k->kxy = kpp * bx * by * ib2;
        k->kxz = kpp * bx * bz * ib2;
        k->kyz = kpp * by * bz * ib2;
    }
}

/* calc_spatial_diffusion_coefficients_nlgc, particle_module.f90:2464-2771
 * (Cartesian, Parker, no delta-B or correlation-length maps, no third
 * dimension in 2D): kperp = kpara0 * kperp_kpara * B^((gamma-2)/3) *
 * (p/p0)^((5-gamma)/3) * mu^2, and separate gradients of ln kpara and
 * ln kperp (the latter a third of the former) in the divergence terms. */
void orc_kappa_nlgc(const orc_params* prm, double p, double mu, const double f[32], orc_kappa* k)
{
    const double bx = f[4], by = f[5], bz = f[6];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    const double gamma_turb = prm->gamma_turb;
    double ib1;
    if (b < DBL_EPSILON) ib1 = 1.0; else ib1 = 1.0 / b;
    const double ib2 = ib1 * ib1;
    const double ib3 = ib1 * ib2;

    k->knorm_para = 1.0;
    k->knorm_perp = 1.0;
    if (prm->mag_dependency == 1) {
        k->knorm_para = k->knorm_para * pow(b, gamma_turb - 2.0);
        k->knorm_perp = k->knorm_perp * pow(b, (gamma_turb - 2.0) / 3.0);
    }
    double knorm_para, knorm_perp;
    if (prm->momentum_dependency == 1) {
        knorm_para = k->knorm_para * pow(p / prm->p0, prm->pindex);
        knorm_perp = k->knorm_perp * pow(p / prm->p0, (5.0 - gamma_turb) / 3.0);
    } else {
        knorm_para = k->knorm_para;
        knorm_perp = k->knorm_perp;
    }
    k->kpara = prm->kpara0 * knorm_para;
    k->kperp = prm->kpara0 * prm->kperp_kpara * knorm_perp * (mu * mu);
    k->skpara = sqrt(2.0 * k->kpara);
    k->skperp = sqrt(2.0 * k->kperp);
    k->skpara_perp = sqrt(2.0 * (k->kpara - k->kperp));

    const double kpara = k->kpara, kperp = k->kperp;
    const double kpp = kpara - kperp;
    const int nd = prm->ndim;
    const double dbx_dx = f[20], dbx_dy = f[21], dbx_dz = nd == 3 ? f[22] : 0.0;
    const double dby_dx = f[23], dby_dy = f[24], dby_dz = nd == 3 ? f[25] : 0.0;
    const double dbz_dx = f[26], dbz_dy = f[27], dbz_dz = nd == 3 ? f[28] : 0.0;
    const double db_dx = f[29], db_dy = f[30], db_dz = nd == 3 ? f[31] : 0.0;
    double dkpara_dx = 0.0, dkpara_dy = 0.0, dkpara_dz = 0.0;
    double dkperp_dx = 0.0, dkperp_dy = 0.0, dkperp_dz = 0.0;
    if (prm->mag_dependency == 1) {
        dkpara_dx = db_dx * ib1 * (gamma_turb - 2.0);
        dkpara_dy = db_dy * ib1 * (gamma_turb - 2.0);
        dkperp_dx = db_dx * ib1 * (gamma_turb - 2.0) / 3.0;
        dkperp_dy = db_dy * ib1 * (gamma_turb - 2.0) / 3.0;
        if (nd == 3) {
            dkpara_dz = db_dz * ib1 * (gamma_turb - 2.0);
            dkperp_dz = db_dz * ib1 * (gamma_turb - 2.0) / 3.0;
        }
    }
    k->dkxx_dx = kperp * dkperp_dx + (kpara * dkpara_dx - kperp * dkperp_dx) * (bx * bx) * ib2 +
                 2.0 * kpp * bx * (dbx_dx * b - bx * db_dx) * ib3;
    k->dkyy_dy = kperp * dkperp_dy + (kpara * dkpara_dy - kperp * dkperp_dy) * (by * by) * ib2 +
                 2.0 * kpp * by * (dby_dy * b - by * db_dy) * ib3;
    k->dkxy_dx = (kpara * dkpara_dx - kperp * dkperp_dx) * bx * by * ib2 +
                 kpp * ((dbx_dx * by + bx * dby_dx) * ib2 - 2.0 * bx * by * db_dx * ib3);
    k->dkxy_dy = (kpara * dkpara_dy - kperp * dkperp_dy) * bx * by * ib2 +
                 kpp * ((dbx_dy * by + bx * dby_dy) * ib2 - 2.0 * bx * by * db_dy * ib3);
    k->kxx = kperp + kpp * bx * bx * ib2;
    k->kyy = kperp + kpp * by * by * ib2;
    k->kxy = kpp * bx * by * ib2;
    if (nd == 3) {
        k->dkzz_dz = kperp * dkperp_dz + (kpara * dkpara_dz - kperp * dkperp_dz) * (bz * bz) * ib2 +
                     2.0 * kpp * bz * (dbz_dz * b - bz * db_dz) * ib3;
        k->dkxz_dx = (kpara * dkpara_dx - kperp * dkperp_dx) * bx * bz * ib2 +
                     kpp * ((dbx_dx * bz + bx * dbz_dx) * ib2 - 2.0 * bx * bz * db_dx * ib3);
        k->dkxz_dz = (kpara * dkpara_dz - kperp * dkperp_dz) * bx * bz * ib2 +
                     kpp * ((dbx_dz * bz + bx * dbz_dz) * ib2 - 2.0 * bx * bz * db_dz * ib3);
        k->dkyz_dy = (kpara * dkpara_dy - kperp * dkperp_dy) * by * bz * ib2 +
                     kpp * ((dby_dy * bz + by * dbz_dy) * ib2 - 2.0 * by * bz * db_dy * ib3);
        k->dkyz_dz = (kpara * dkpara_dz - kperp * dkperp_dz) * by * bz * ib2 +
                     kpp * ((dby_dz * bz + by * dbz_dz) * ib2 - 2.0 * by * bz * db_dz * ib3);
        k->kzz = kperp + kpp * bz * bz * ib2;
        k->kxz = kpp * bx * bz * ib2;
        k->kyz = kpp * by * bz * ib2;
    } else {
        k->kzz = k->kxz = k->kyz = 0.0;
        k->dkzz_dz = k->dkxz_dx = k->dkxz_dz = k->dkyz_dy = k->dkyz_dz = 0.0;
    }
}

/* particle_in_acceleration_region, particle_module.f90:2885-2906 */
static int in_acc_region(const orc_params* prm, const orc_particle* ptl)
{
    const double xn = (ptl->x - prm->xmin) / prm->lx;
    int inx = (xn >= prm->acc_region[0]) && (xn <= prm->acc_region[1]);
    int iny = 1, inz = 1;
    if (prm->ndim > 1) {
        const double yn = (ptl->y - prm->ymin) / prm->ly;
        iny = (yn >= prm->acc_region[2]) && (yn <= prm->acc_region[3]);
    }
    if (prm->ndim == 3) {
        const double zn = (ptl->z - prm->zmin) / prm->lz;
        inz = (zn >= prm->acc_region[4]) && (zn <= prm->acc_region[5]);
    }
    return inx && iny && inz;
}

/* Drift-speed prefactor, particle_module.f90:3436 / :4727.
 * 1.0/(3*pcharge) is a default-real (FP32) division in the Fortran. */
static double drift_vdp(const orc_params* prm, double p)
{
    const float pre = 1.0f / (float)(3 * prm->pcharge);
    const double t1 = prm->drift1 * prm->p0 / p;
    const double t2 = prm->drift2 * (prm->p0 * prm->p0) / (p * p);
    return (double)pre / sqrt(t1 * t1 + t2 * t2);
}

static double fmin_seq(double a, double b) { return b < a ? b : a; }

/* calc_dpp_wave_scattering, particle_module.f90:2918-2933: va = b/sqrt(rho) */
static void dpp_wave_scattering(const orc_params* prm, double rho, double b, double kpara,
                                double p, double* dp_dt, double* dpp)
{
    const double va = b / sqrt(rho);
    if (prm->momentum_dependency == 1)
        *dp_dt = *dp_dt + (8 * p / (27 * kpara)) * (va * va);
    else
        *dp_dt = *dp_dt + (4 * p / (9 * kpara)) * (va * va);
    *dpp = *dpp + ((p * va) * (p * va)) / (9 * kpara);
}

/* calc_dpp_flow_shear, particle_module.f90:2947-2979 */
static void dpp_flow_shear(const orc_params* prm, double b, double bx, double by, double bz,
                           double knorm_para, double sxx, double syy, double szz, double sxy,
                           double sxz, double syz, double p, double* dp_dt, double* dpp)
{
    double gshear;
    if (prm->weak_scattering) {
        double ib;
        if (b < DBL_EPSILON) ib = 0.0; else ib = 1.0 / b;
        double bbsigma = sxx * (bx * bx) + syy * (by * by) + szz * (bz * bz) +
                         2.0 * (sxy * bx * by + sxz * bx * bz + syz * by * bz);
        bbsigma = bbsigma * ib * ib;
        gshear = (bbsigma * bbsigma) / 5;
    } else {
        gshear = 2 * (sxx * sxx + syy * syy + szz * szz +
                      2 * (sxy * sxy + sxz * sxz + syz * syz)) / 15;
    }
    if (gshear > 0.0) {
        const double pindex = prm->pindex;
        *dp_dt = *dp_dt + (2 + pindex) * gshear * prm->tau0 * knorm_para *
                 pow(p, pindex - 1) * pow(prm->p0, 2.0 - pindex);
        *dpp = *dpp + gshear * prm->tau0 * knorm_para * pow(p, pindex) * pow(prm->p0, 2.0 - pindex);
    }
}

/* momentum update shared by both pushers (particle_module.f90:3589-3605,
 * :4885-4906): the increment with the momentum-diffusion term, the
 * acceleration-region gate and the 0.25*p0 floor. */
static void update_momentum(const orc_params* prm, orc_particle* ptl, double dp_dt, double dpp,
                            double sdt, double ran_p, double* deltap)
{
    double dp = dp_dt * ptl->dt + ran_p * sqrt(2 * dpp) * sdt;
    if (prm->acc_region_flag == 1) {
        if (in_acc_region(prm, ptl)) ptl->p = ptl->p + dp;
        else dp = 0.0;
    } else {
        ptl->p = ptl->p + dp;
    }
    if (ptl->p < 0.25 * prm->p0) {
        ptl->p = ptl->p - dp;
        dp = 0.25 * prm->p0 - ptl->p;
        ptl->p = 0.25 * prm->p0;
    }
    *deltap = dp;
}

/* push_particle_2d, particle_module.f90:3358-3606 (Cartesian, uniform grid). */
void orc_push_particle_2d(const orc_params* prm, const double dtmm[2], orc_particle* ptl,
                          const double f[32], const orc_kappa* k, int fixed_dt,
                          const double rands[4], double delta[4])
{
    const double vx = f[0], vy = f[1];
    const double bx = f[4], by = f[5], bz = f[6];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    const double dvx_dx = f[8], dvy_dy = f[12];
    const double dxm = prm->dx, dym = prm->dy;
    delta[0] = delta[1] = delta[2] = delta[3] = 0.0;
    double ib;
    if (b < DBL_EPSILON) ib = 0.0; else ib = 1.0 / b;

    const double dbz_dx = f[26], dbz_dy = f[27];
    const double db_dx = f[29], db_dy = f[30];
    const double ib2 = ib * ib;
    const double ib3 = ib * ib2;
    const double vdp = drift_vdp(prm, ptl->p);
    const double vdx = vdp * (dbz_dy * ib2 - 2.0 * bz * db_dy * ib3);
    const double vdy = vdp * (-(dbz_dx * ib2) + 2.0 * bz * db_dx * ib3);
    double vdz = 0.0;
    if (prm->check_drift_2d) {
        const double dbx_dy = f[21], dby_dx = f[23];
        vdz = vdp * ((dby_dx - dbx_dy) * ib2 - 2.0 * (by * db_dx - bx * db_dy) * ib3);
    }
    const double dx_dt = vx + vdx + k->dkxx_dx + k->dkxy_dy;
    const double dy_dt = vy + vdy + k->dkxy_dx + k->dkyy_dy;
    const double dz_dt = vdz;
    const double divv = dvx_dx + dvy_dy;
    double dp_dt = -(ptl->p * divv / 3.0);
    double dpp = 0.0;
    /* momentum diffusion, :3479-3500 */
    if (prm->dpp_wave) dpp_wave_scattering(prm, f[3], b, k->kpara, ptl->p, &dp_dt, &dpp);
    if (prm->dpp_shear) {
        const double dvx_dy = f[9], dvy_dx = f[11];
        const double sxx = dvx_dx - divv / 3;
        const double syy = dvy_dy - divv / 3;
        const double szz = -(divv / 3);
        const double sxy = (dvx_dy + dvy_dx) / 2;
        dpp_flow_shear(prm, b, bx, by, bz, k->knorm_para, sxx, syy, szz, sxy, 0.0, 0.0, ptl->p,
                       &dp_dt, &dpp);
    }

    if (!fixed_dt) {
        if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0) {
            const double sk = k->skperp > 0.0 ? k->skperp : k->skpara;
            double a = 0.5 * dxm / k->skpara;
            double m = a * a;
            a = 0.5 * dym / k->skpara; m = fmin_seq(m, a * a);
            a = sk / dx_dt;            m = fmin_seq(m, a * a);
            a = sk / dy_dt;            m = fmin_seq(m, a * a);
            m = fmin_seq(m, (double)0.1f * ptl->p / fabs(dp_dt));
            ptl->dt = m;
        } else {
            ptl->dt = dtmm[0];
        }
        if (ptl->dt < dtmm[0]) ptl->dt = dtmm[0];
        if (ptl->dt > dtmm[1]) ptl->dt = dtmm[1];
    }

    const double sdt = sqrt(ptl->dt);
    const double ran1 = rands[0], ran2 = rands[1], ran3 = rands[2];
    const double deltax = dx_dt * ptl->dt + ran1 * k->skperp * sdt +
                          ran3 * k->skpara_perp * sdt * bx * ib;
    const double deltay = dy_dt * ptl->dt + ran2 * k->skperp * sdt +
                          ran3 * k->skpara_perp * sdt * by * ib;
    const double deltaz = dz_dt * ptl->dt;
    ptl->x = ptl->x + deltax;
    ptl->y = ptl->y + deltay;
    ptl->z = ptl->z + deltaz;
    ptl->t = ptl->t + ptl->dt;
    delta[0] = deltax;
    delta[1] = deltay;
    update_momentum(prm, ptl, dp_dt, dpp, sdt, rands[3], &delta[3]);
}

/* push_particle_3d, particle_module.f90:4625-4907 (Cartesian, uniform grid). */
void orc_push_particle_3d(const orc_params* prm, const double dtmm[2], orc_particle* ptl,
                          const double f[32], const orc_kappa* k, int fixed_dt,
                          const double rands[4], double delta[4])
{
    const double vx = f[0], vy = f[1], vz = f[2];
    const double bx = f[4], by = f[5], bz = f[6];
    const double b = sqrt(bx * bx + by * by + bz * bz);
    double ib;
    if (b < DBL_EPSILON) ib = 0.0; else ib = 1.0 / b;
    const double bxn = bx * ib, byn = by * ib, bzn = bz * ib;
    const double bxyn = sqrt(bxn * bxn + byn * byn);
    double ibxyn;
    if (bxyn < DBL_EPSILON) ibxyn = 0.0; else ibxyn = 1.0 / bxyn;
    const double dvx_dx = f[8], dvy_dy = f[12], dvz_dz = f[16];
    const double dxm = prm->dx, dym = prm->dy, dzm = prm->dz;
    delta[0] = delta[1] = delta[2] = delta[3] = 0.0;

    const double dbx_dy = f[21], dbx_dz = f[22];
    const double dby_dx = f[23], dby_dz = f[25];
    const double dbz_dx = f[26], dbz_dy = f[27];
    const double db_dx = f[29], db_dy = f[30], db_dz = f[31];
    const double ib2 = ib * ib;
    const double ib3 = ib * ib2;
    const double vdp = drift_vdp(prm, ptl->p);
    const double vdx = vdp * ((dbz_dy - dby_dz) * ib2 - 2.0 * (bz * db_dy - by * db_dz) * ib3);
    const double vdy = vdp * ((dbx_dz - dbz_dx) * ib2 - 2.0 * (bx * db_dz - bz * db_dx) * ib3);
    const double vdz = vdp * ((dby_dx - dbx_dy) * ib2 - 2.0 * (by * db_dx - bx * db_dy) * ib3);

    const double dx_dt = vx + vdx + k->dkxx_dx + k->dkxy_dy + k->dkxz_dz;
    const double dy_dt = vy + vdy + k->dkxy_dx + k->dkyy_dy + k->dkyz_dz;
    const double dz_dt = vz + vdz + k->dkxz_dx + k->dkyz_dy + k->dkzz_dz;
    const double divv = dvx_dx + dvy_dy + dvz_dz;
    double dp_dt = -(ptl->p * divv / 3.0);
    double dpp = 0.0;
    /* momentum diffusion, :4764-4787 */
    if (prm->dpp_wave) dpp_wave_scattering(prm, f[3], b, k->kpara, ptl->p, &dp_dt, &dpp);
    if (prm->dpp_shear) {
        const double dvx_dy = f[9], dvx_dz = f[10], dvy_dx = f[11];
        const double dvy_dz = f[13], dvz_dx = f[14], dvz_dy = f[15];
        const double sxx = dvx_dx - divv / 3;
        const double syy = dvy_dy - divv / 3;
        const double szz = dvz_dz - divv / 3;
        const double sxy = (dvx_dy + dvy_dx) / 2;
        const double sxz = (dvx_dz + dvz_dx) / 2;
        const double syz = (dvy_dz + dvz_dy) / 2;
        dpp_flow_shear(prm, b, bx, by, bz, k->knorm_para, sxx, syy, szz, sxy, sxz, syz, ptl->p,
                       &dp_dt, &dpp);
    }

    if (!fixed_dt) {
        /* the zero test omits dz_dt (particle_module.f90:4792-4794) */
        if (dx_dt != 0.0 && dy_dt != 0.0 && dp_dt != 0.0) {
            const double sk = k->skperp > 0.0 ? k->skperp : k->skpara;
            double a = 0.5 * dxm / k->skpara;
            double m = a * a;
            a = 0.5 * dym / k->skpara; m = fmin_seq(m, a * a);
            a = 0.5 * dzm / k->skpara; m = fmin_seq(m, a * a);
            a = sk / dx_dt;            m = fmin_seq(m, a * a);
            a = sk / dy_dt;            m = fmin_seq(m, a * a);
            a = sk / dz_dt;            m = fmin_seq(m, a * a);
            m = fmin_seq(m, (double)0.1f * ptl->p / fabs(dp_dt));
            ptl->dt = m;
        } else {
            ptl->dt = dtmm[0];
        }
        if (ptl->dt < dtmm[0]) ptl->dt = dtmm[0];
        if (ptl->dt > dtmm[1]) ptl->dt = dtmm[1];
    }

    const double sdt = sqrt(ptl->dt);
    const double ran1 = rands[0], ran2 = rands[1], ran3 = rands[2];
    const double skpara = k->skpara, skperp = k->skperp;
    const double deltax = dx_dt * ptl->dt +
        (bxn * skpara * ran1 - bxn * bzn * skperp * ibxyn * ran2 - byn * skperp * ibxyn * ran3) * sdt;
    const double deltay = dy_dt * ptl->dt +
        (byn * skpara * ran1 - byn * bzn * skperp * ibxyn * ran2 + bxn * skperp * ibxyn * ran3) * sdt;
    const double deltaz = dz_dt * ptl->dt + (bzn * skpara * ran1 + bxyn * skperp * ran2) * sdt;
    ptl->x = ptl->x + deltax;
    ptl->y = ptl->y + deltay;
    ptl->z = ptl->z + deltaz;
    ptl->t = ptl->t + ptl->dt;
    delta[0] = deltax;
    delta[1] = deltay;
    delta[2] = deltaz;
    update_momentum(prm, ptl, dp_dt, dpp, sdt, rands[3], &delta[3]);
}

/* set_dt_min_max, particle_module.f90:5519-5524 */
static void dt_min_max(const orc_params* prm, double dtf, double dtmm[2])
{
    dtmm[0] = prm->dt_min_rel * dtf;
    dtmm[1] = prm->dt_max_rel * dtf;
}

/* One gather + kappa + push, as in particle_module.f90:1615-1690. */
static void gather_kappa_push(const orc_params* prm, const float* farray1, const float* farray2,
                              double t0, double dtf, const double dtmm[2], orc_particle* ptl,
                              int fixed_dt, const double rands[4], double delta[4])
{
    const double px = (ptl->x - prm->xmin) / prm->dx;
    const double py = (ptl->y - prm->ymin) / prm->dy;
    const double pz = (ptl->z - prm->zmin) / prm->dz;
    const double rt = (ptl->t - t0) / dtf;
    int pos[3];
    double w[8], fields[32];
    orc_kappa kap;
    orc_get_interp_paramters(prm, px, py, pz, pos, w);
    orc_interp_fields(prm, farray1, farray2, pos, w, rt, fields);
    if (prm->nlgc) orc_kappa_nlgc(prm, ptl->p, ptl->mu, fields, &kap);
    else orc_kappa_standard(prm, ptl->p, fields, &kap);
    if (prm->ndim == 2) {
        double d[4];
        orc_push_particle_2d(prm, dtmm, ptl, fields, &kap, fixed_dt, rands, d);
        delta[0] = d[0]; delta[1] = d[1]; delta[3] = d[3];  /* 2D leaves deltaz untouched */
    } else {
        orc_push_particle_3d(prm, dtmm, ptl, fields, &kap, fixed_dt, rands, delta);
    }
}

void orc_step_once(const orc_params* prm, const float* farray1, const float* farray2,
                   double t0, double dtf, orc_particle* ptl, int fixed_dt,
                   const double rands[4], double delta[4])
{
    double dtmm[2];
    dt_min_max(prm, dtf, dtmm);
    delta[0] = delta[1] = delta[2] = delta[3] = 0.0;
    gather_kappa_push(prm, farray1, farray2, t0, dtf, dtmm, ptl, fixed_dt, rands, delta);
}

/* particle_boundary_condition, particle_module.f90:1984-2129, single rank:
 * a periodic axis has the rank itself as neighbour, an open axis has -1.
 * The periodic shift uses the bounds passed in (the half-cell slack bounds
 * while stepping). */
static void boundary_condition(const orc_params* prm, orc_particle* ptl,
                               const double lo[3], const double hi[3], double* leak)
{
    if (ptl->x < lo[0] && ptl->count_flag == ORC_INBOX) {
        if (prm->pbcx == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_LX; }
        else ptl->x = ptl->x - lo[0] + hi[0];
    } else if (ptl->x > hi[0] && ptl->count_flag == ORC_INBOX) {
        if (prm->pbcx == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_HX; }
        else ptl->x = ptl->x - hi[0] + lo[0];
    }
    if (prm->ndim > 1) {
        if (ptl->y < lo[1] && ptl->count_flag == ORC_INBOX) {
            if (prm->pbcy == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_LY; }
            else ptl->y = ptl->y - lo[1] + hi[1];
        } else if (ptl->y > hi[1] && ptl->count_flag == ORC_INBOX) {
            if (prm->pbcy == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_HY; }
            else ptl->y = ptl->y - hi[1] + lo[1];
        }
    }
    if (prm->ndim == 3) {
        if (ptl->z < lo[2] && ptl->count_flag == ORC_INBOX) {
            if (prm->pbcz == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_LZ; }
            else ptl->z = ptl->z - lo[2] + hi[2];
        } else if (ptl->z > hi[2] && ptl->count_flag == ORC_INBOX) {
            if (prm->pbcz == 1) { *leak = *leak + ptl->weight; ptl->count_flag = ORC_ESC_HZ; }
            else ptl->z = ptl->z - hi[2] + lo[2];
        }
    }
}

typedef struct {
    int mode;
    orc_mt19937* mt;
    uint32_t key[2];
    uint32_t stream;
} orc_rng;

static void step_rands(orc_rng* r, orc_particle* ptl, double rands[4])
{
    const double sqrt3 = sqrt(3.0);
    uint32_t u[4];
    if (r->mode == 0) {
        const uint32_t ctr[4] = {(uint32_t)ptl->tag_injected, (uint32_t)ptl->tag_splitted,
                                 ptl->rng_step, r->stream};
        orc_philox4x32_10(ctr, r->key, u);
    } else {
        for (int i = 0; i < 4; ++i) u[i] = orc_mt_genrand_int32(r->mt);
    }
    ptl->rng_step += 1u;
    for (int i = 0; i < 4; ++i) rands[i] = (2.0 * orc_u01(u[i]) - 1.0) * sqrt3;
}

/* One particle through one MHD interval: particle_mover_one_cycle body,
 * particle_module.f90:1561-1830, followed by the exact-bounds boundary pass
 * of particle_mover, :1959-1968. */
static void move_one(const orc_params* prm, const float* farray1, const float* farray2,
                     orc_particle* ptl, double t0, double dtf, int num_fine_steps,
                     int nsteps_interval, const double dtmm[2], const double slo[3],
                     const double shi[3], const double elo[3], const double ehi[3],
                     orc_rng* rng, double* leak, double* leak_negp, int64_t* steps)
{
    const double dt_fine = dtf / num_fine_steps;
    double delta[4] = {0.0, 0.0, 0.0, 0.0};
    double rands[4];

    const int step = (int)ceil((ptl->t - t0) / dt_fine);
    double dt_target;
    if (step <= 0) dt_target = dt_fine; else dt_target = step * dt_fine;
    if (dt_target > dtf) dt_target = dtf;

    if (ptl->p < 0.0 && ptl->count_flag == ORC_INBOX) {
        ptl->count_flag = ORC_OTHERS;
        *leak_negp = *leak_negp + ptl->weight;
    } else {
        boundary_condition(prm, ptl, slo, shi, leak);
    }
    if (ptl->count_flag != ORC_INBOX) return;

    while (dt_target < (dtf + dt_fine * (double)0.1f)) {
        if (ptl->count_flag != ORC_INBOX) break;
        while ((ptl->t - t0) < dt_target && ptl->count_flag == ORC_INBOX) {
            if (ptl->p < 0.0) {
                ptl->count_flag = ORC_OTHERS;
                *leak_negp = *leak_negp + ptl->weight;
            } else {
                boundary_condition(prm, ptl, slo, shi, leak);
            }
            if (ptl->count_flag != ORC_INBOX) break;
            step_rands(rng, ptl, rands);
            gather_kappa_push(prm, farray1, farray2, t0, dtf, dtmm, ptl, 0, rands, delta);
            ptl->nsteps_pushed = (ptl->nsteps_pushed + 1) % nsteps_interval;
            *steps += 1;
        }
        if ((ptl->t - t0) > dt_target && ptl->count_flag == ORC_INBOX) {
            ptl->x = ptl->x - delta[0];
            ptl->y = ptl->y - delta[1];
            ptl->z = ptl->z - delta[2];
            ptl->p = ptl->p - delta[3];
            ptl->t = ptl->t - ptl->dt;
            const double dt_old = ptl->dt;
            ptl->dt = t0 + dt_target - ptl->t;
            if (ptl->dt > 0) {
                ptl->nsteps_pushed = ptl->nsteps_pushed - 1;
                step_rands(rng, ptl, rands);
                gather_kappa_push(prm, farray1, farray2, t0, dtf, dtmm, ptl, 1, rands, delta);
                ptl->nsteps_pushed = (ptl->nsteps_pushed + 1) % nsteps_interval;
                *steps += 1;
            }
            ptl->dt = dt_old;
            if (ptl->p < 0.0) {
                ptl->count_flag = ORC_OTHERS;
                *leak_negp = *leak_negp + ptl->weight;
            } else {
                boundary_condition(prm, ptl, slo, shi, leak);
            }
        }
        dt_target = dt_target + dt_fine;
    }
    /* particle_mover's final pass with the exact bounds (:1959-1968); the
     * p<0 branch there requires count_flag /= INBOX and never fires for the
     * in-box particles that reach it. */
    if (ptl->count_flag == ORC_INBOX) boundary_condition(prm, ptl, elo, ehi, leak);
}

/* particle_mover, particle_module.f90:1846-1974, on one rank: one push cycle
 * over all particles, removal, exact-bounds boundary pass, removal. */
void orc_particle_mover(const orc_params* prm, const float* farray1, const float* farray2,
                        orc_particle* ptls, int64_t* nptl_current,
                        double t0, double dtf, int num_fine_steps, int nsteps_interval,
                        int rng_mode, int nthreads, orc_mover_acc* acc)
{
    double dtmm[2];
    dt_min_max(prm, dtf, dtmm);
    const double slo[3] = {prm->xmin - prm->dx * 0.5, prm->ymin - prm->dy * 0.5, prm->zmin - prm->dz * 0.5};
    const double shi[3] = {prm->xmax + prm->dx * 0.5, prm->ymax + prm->dy * 0.5, prm->zmax + prm->dz * 0.5};
    const double elo[3] = {prm->xmin, prm->ymin, prm->zmin};
    const double ehi[3] = {prm->xmax, prm->ymax, prm->zmax};
    const int64_t n = *nptl_current;
    double leak = 0.0, leak_negp = 0.0;
    int64_t steps = 0;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : leak, leak_negp, steps)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_mt19937 mt;
        orc_rng rng;
        rng.mode = rng_mode;
        rng.mt = &mt;
        rng.key[0] = (uint32_t)prm->seed;
        rng.key[1] = (uint32_t)(prm->seed >> 32);
        rng.stream = prm->stream;
        if (rng_mode == 1) {
            const uint32_t base[5] = {0x123u, 0x234u, 0x345u, 0x456u, (uint32_t)tid};
            orc_mt_init_by_array(&mt, base, tid == 0 ? 4 : 5);
        }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t i = 0; i < n; ++i)
            move_one(prm, farray1, farray2, &ptls[i], t0, dtf, num_fine_steps, nsteps_interval,
                     dtmm, slo, shi, elo, ehi, &rng, &leak, &leak_negp, &steps);
    }
    int64_t nesc = 0;
    orc_remove_particles(ptls, nptl_current, &nesc);
    if (acc) {
        acc->leak += leak;
        acc->leak_negp += leak_negp;
        acc->steps += steps;
        acc->nptl_escaped += nesc;
    }
}

/* remove_particles, particle_module.f90:5365-5403: swap non-INBOX particles
 * to the tail (not order preserving); escaped ones (count_flag < 0) are
 * counted. */
void orc_remove_particles(orc_particle* ptls, int64_t* nptl_current, int64_t* nptl_escaped)
{
    int64_t n = *nptl_current;
    if (n <= 0) return;
    int64_t nremoved = 0;
    int64_t i = 1;  /* 1-based as in the Fortran */
    while (i <= n) {
        if ((n - i) == (nremoved - 1)) break;
        orc_particle* pi = &ptls[i - 1];
        if (pi->count_flag == ORC_INBOX) {
            i = i + 1;
        } else {
            if (pi->count_flag < 0) *nptl_escaped += 1;
            orc_particle tmp = ptls[n - nremoved - 1];
            ptls[n - nremoved - 1] = *pi;
            *pi = tmp;
            nremoved = nremoved + 1;
        }
    }
    *nptl_current = n - nremoved;
}

static void inject_one_particle(const orc_params* prm, orc_particle* q, orc_drawstream* ds,
                                double xt, double yt, double zt, double mu, int dist_flag,
                                double particle_v0, double t0, double dtf, double dt,
                                double power_index, int32_t* tag_max);

/* inject_particles_spatial_uniform + inject_one_particle,
 * particle_module.f90:385-530 (whole domain on one rank).  Draw order per
 * particle: x, y, z, mu, [momentum draws], t, from the particle's own Philox
 * draw stream.  An injection past nptl_max overwrites the last slot, as the
 * reference does (:477). */
void orc_inject_particles_spatial_uniform(const orc_params* prm, orc_particle* ptls,
                                          int64_t* nptl_current, int64_t nptl_max,
                                          int64_t nptl, double dt, int dist_flag,
                                          double particle_v0, double t0, double dtf,
                                          const double part_box[6], double power_index,
                                          int32_t* tag_max)
{
    const double mu_max = (double)0.99f;
    for (int64_t i = 0; i < nptl; ++i) {
        int64_t slot = *nptl_current + 1;
        if (slot > nptl_max) slot = nptl_max;
        *nptl_current = slot;
        orc_particle* q = &ptls[slot - 1];
        orc_drawstream ds;
        ds_init(&ds, prm, (uint32_t)*tag_max);
        const double xt = ds_next(&ds) * (part_box[3] - part_box[0]) + part_box[0];
        const double yt = ds_next(&ds) * (part_box[4] - part_box[1]) + part_box[1];
        const double zt = ds_next(&ds) * (part_box[5] - part_box[2]) + part_box[2];
        const double mu = mu_max * (2.0 * ds_next(&ds) - 1.0);
        inject_one_particle(prm, q, &ds, xt, yt, zt, mu, dist_flag, particle_v0, t0, dtf, dt,
                            power_index, tag_max);
    }
}

/* inject_one_particle, particle_module.f90:385-441: momentum by the
 * distribution law, then the time stamp, from the particle's draw stream. */
static void inject_one_particle(const orc_params* prm, orc_particle* q, orc_drawstream* ds_,
                                double xt, double yt, double zt, double mu, int dist_flag,
                                double particle_v0, double t0, double dtf, double dt,
                                double power_index, int32_t* tag_max)
{
    orc_drawstream ds = *ds_;
    {   /* draws continue the caller's stream */
        memset(q, 0, sizeof(*q));
        q->x = xt; q->y = yt; q->z = zt;
        if (dist_flag == 0) {
            double ftest = 1.0, fxp = 0.5, ptmp = 0.0;
            while (ftest > fxp) {
                ptmp = (ds_next(&ds) * (prm->pmax - prm->pmin) + prm->pmin) / prm->p0;
                fxp = (ptmp * ptmp) * exp(-(ptmp * ptmp));
                ftest = ds_next(&ds) * (double)0.37f;
            }
            q->p = ptmp * prm->p0;
        } else if (dist_flag == 1) {
            q->p = prm->p0;
        } else if (dist_flag == 2) {
            const double r01 = ds_next(&ds);
            if ((int)power_index == 1) {
                q->p = pow(prm->pmax / prm->p0, r01) * prm->p0;
            } else {
                const double norm = pow(prm->pmax, -power_index + 1) - pow(prm->p0, -power_index + 1);
                q->p = pow(r01 * norm + pow(prm->p0, -power_index + 1), 1.0 / (-power_index + 1));
            }
        }
        q->v = particle_v0 * q->p / prm->p0;
        q->mu = mu;
        q->weight = 1.0;
        q->t = t0 + ds_next(&ds) * dtf;
        q->dt = dt;
        q->split_times = 0;
        q->count_flag = ORC_INBOX;
        q->origin = 0;
        q->nsteps_tracked = 0;
        q->nsteps_pushed = 0;
        q->tag_injected = *tag_max;
        *tag_max = *tag_max + 1;
        q->tag_splitted = 1;
        q->rng_step = 0u;
    }
    *ds_ = ds;
}

/* The interior-node quantity of get_ncells_large_jz / _absj / _divv / _rho
 * (mhd_data_parallel.f90:2211-2499), in the precision the Fortran evaluates it:
 * |jz| and |J| in FP32 (farray1 is real(sp)), divV summed in FP64.  (gi, gj,
 * gk) is the 0-based ghost-inclusive node. */
static double node_quantity(const orc_params* prm, const float* fa, int quantity, int gi, int gj,
                            int gk)
{
    const int nxg = nxg_of(prm), nyg = nyg_of(prm);
    switch (quantity) {
    case ORC_THR_JZ: {
        const float d = FA(fa, 23, gi, gj, gk) - FA(fa, 21, gi, gj, gk);
        return (double)fabsf(d);
    }
    case ORC_THR_ABSJ: {
        const float a = FA(fa, 25, gi, gj, gk) - FA(fa, 27, gi, gj, gk);
        const float b = FA(fa, 26, gi, gj, gk) - FA(fa, 22, gi, gj, gk);
        const float c = FA(fa, 21, gi, gj, gk) - FA(fa, 23, gi, gj, gk);
        return (double)sqrtf(a * a + b * b + c * c);
    }
    case ORC_THR_DIVV: {
        double d = (double)FA(fa, 8, gi, gj, gk);
        if (prm->ndim > 1) {
            d = d + (double)FA(fa, 12, gi, gj, gk);
            if (prm->ndim > 2) d = d + (double)FA(fa, 16, gi, gj, gk);
        }
        return -d;
    }
    default:
        return (double)FA(fa, 3, gi, gj, gk);
    }
}

/* get_ncells_large_jz / _absj / _divv / _rho, mhd_data_parallel.f90:2211-2499,
 * uniform Cartesian grid, one rank: interior nodes strictly inside part_box
 * (axes beyond ndim not tested) whose quantity exceeds the threshold (-divV for
 * divV).  The divV count reads node (i-2, j-2, k-2): the routine assigns the
 * whole ghosted farray1 slice to an array allocated without ghosts, which
 * Fortran's reallocation on assignment rebounds to start at 1 (:2423-2429);
 * in 2D the z extent has no ghosts and no shift. */
int64_t orc_ncells_above_threshold(const orc_params* prm, const float* farray1, int quantity,
                                   double threshold, const double part_box[6])
{
    const int nz = prm->ndim > 2 ? prm->nz : 1, ny = prm->ndim > 1 ? prm->ny : 1;
    int64_t n = 0;
    for (int fk = 1; fk <= nz; ++fk) {
        const double zp = prm->zmin + (fk - 1) * prm->dz;
        const int inz = prm->ndim > 2 ? (zp > part_box[2] && zp < part_box[5]) : 1;
        for (int fj = 1; fj <= ny; ++fj) {
            const double yp = prm->ymin + (fj - 1) * prm->dy;
            const int iny = prm->ndim > 1 ? (yp > part_box[1] && yp < part_box[4]) : 1;
            for (int fi = 1; fi <= prm->nx; ++fi) {
                const double xp = prm->xmin + (fi - 1) * prm->dx;
                const int inx = xp > part_box[0] && xp < part_box[3];
                if (!(inx && iny && inz)) continue;
                int gi = fi + 1, gj = prm->ndim > 1 ? fj + 1 : 0, gk = prm->ndim > 2 ? fk + 1 : 0;
                if (quantity == ORC_THR_DIVV) {
                    gi -= 2;
                    if (prm->ndim > 1) gj -= 2;
                    if (prm->ndim > 2) gk -= 2;
                }
                if (node_quantity(prm, farray1, quantity, gi, gj, gk) > threshold) n++;
            }
        }
    }
    return n;
}

/* The quantity the threshold injections test at a trial position, from the
 * interpolated fields (particle_module.f90:883-890 jz, :1035-1043 |J|,
 * :1318-1323 divV, :1456 rho), in FP64. */
static double point_quantity(const orc_params* prm, int quantity, const double f[32])
{
    switch (quantity) {
    case ORC_THR_JZ:
        return fabs(f[23] - f[21]);
    case ORC_THR_ABSJ: {
        const double a = f[25] - f[27], b = f[26] - f[22], c = f[21] - f[23];
        return sqrt(a * a + b * b + c * c);
    }
    case ORC_THR_DIVV: {
        double d = f[8];
        if (prm->ndim > 1) {
            d = d + f[12];
            if (prm->ndim > 2) d = d + f[16];
        }
        return d;
    }
    default:
        return f[3];
    }
}

/* inject_particles_at_large_jz / _absj / _divv / _rho,
 * particle_module.f90:785-1470, uniform Cartesian grid, one rank.  The count
 * is int(nptl * ncells / ncells_g) with inject_same_nptl (one rank: nptl when
 * some cell passes), else int(nptl * ncells / ncells_norm); no cell: none.
 * Each particle draws x, y, z over the domain until the point lies in
 * part_box (inclusive) and its interpolated quantity (frame 1) passes; a
 * point outside the box scores -3 (jz, |J|), divV = 3 or rho = 0.  The loop
 * tests its initial value first (jz, |J|: -2, divV: 2, rho: 0): when that
 * already passes no point is drawn and the particle sits at the box's low
 * corner.  Then mu, then inject_one_particle.  Returns the number injected. */
int64_t orc_inject_particles_threshold(const orc_params* prm, const float* farray1,
                                       orc_particle* ptls, int64_t* nptl_current,
                                       int64_t nptl_max, int quantity, int64_t nptl, double dt,
                                       int dist_flag, double particle_v0, double t0, double dtf,
                                       int inject_same_nptl, double threshold,
                                       int64_t ncells_norm, const double part_box[6],
                                       double power_index, int32_t* tag_max)
{
    const double mu_max = (double)0.99f;
    const int64_t ncells = orc_ncells_above_threshold(prm, farray1, quantity, threshold, part_box);
    if (ncells == 0) return 0;
    int64_t nptl_inject;
    if (inject_same_nptl) nptl_inject = (int64_t)((double)nptl * ((double)ncells / (double)ncells));
    else nptl_inject = (int64_t)((double)nptl * ((double)ncells / (double)ncells_norm));
    const int divv = quantity == ORC_THR_DIVV;
    const double v_init = divv ? 2.0 : (quantity == ORC_THR_RHO ? 0.0 : -2.0);
    const double v_out = divv ? 3.0 : (quantity == ORC_THR_RHO ? 0.0 : -3.0);
    const double zlo = prm->zmin, zhi = prm->zmax;
    for (int64_t i = 0; i < nptl_inject; ++i) {
        int64_t slot = *nptl_current + 1;
        if (slot > nptl_max) slot = nptl_max;
        *nptl_current = slot;
        orc_drawstream ds;
        ds_init(&ds, prm, (uint32_t)*tag_max);
        double xt = part_box[0], yt = part_box[1], zt = part_box[2];
        double v = v_init;
        while ((divv ? -v : v) < threshold) {
            xt = ds_next(&ds) * (prm->xmax - prm->xmin) + prm->xmin;
            yt = ds_next(&ds) * (prm->ymax - prm->ymin) + prm->ymin;
            zt = ds_next(&ds) * (zhi - zlo) + zlo;
            if (xt >= part_box[0] && xt <= part_box[3] && yt >= part_box[1] && yt <= part_box[4] &&
                zt >= part_box[2] && zt <= part_box[5]) {
                int pos[3];
                double w[8], f[32];
                orc_get_interp_paramters(prm, (xt - prm->xmin) / prm->dx, (yt - prm->ymin) / prm->dy,
                                         (zt - prm->zmin) / prm->dz, pos, w);
                orc_interp_fields(prm, farray1, farray1, pos, w, 0.0, f);
                v = point_quantity(prm, quantity, f);
            } else {
                v = v_out;
            }
        }
        const double mu = mu_max * (2.0 * ds_next(&ds) - 1.0);
        inject_one_particle(prm, &ptls[slot - 1], &ds, xt, yt, zt, mu, dist_flag, particle_v0, t0,
                            dtf, dt, power_index, tag_max);
    }
    return nptl_inject;
}

/* x**n for an integer n, by right-to-left binary powering (the method of
 * libgcc's __powidf2, which gfortran uses for real**integer). */
static double powi_rl(double x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    double y = (n % 2u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* split_particle, particle_module.f90:5430-5480 (no particle tracking). */
void orc_split_particle(const orc_params* prm, orc_particle* ptls, int64_t* nptl_current,
                        int64_t nptl_max, double split_ratio, double pmin_split,
                        int64_t* nptl_split)
{
    const int64_t nptl = *nptl_current;
    for (int64_t i = 0; i < nptl; ++i) {
        orc_particle ptl = ptls[i];
        const double thr = pmin_split * prm->p0 * powi_rl(split_ratio, ptl.split_times);
        if (ptl.p > thr && ptl.p <= prm->pmax) {
            *nptl_current = *nptl_current + 1;
            if (*nptl_current > nptl_max) {
                *nptl_current = nptl_max;
                return;
            }
            *nptl_split = *nptl_split + 1;
            ptl.weight = ldexp(1.0, -(1 + ptl.split_times));  /* 0.5**(1.0 + split_times) */
            ptl.split_times = (int8_t)(ptl.split_times + 1);
            ptls[*nptl_current - 1] = ptl;
            ptls[*nptl_current - 1].tag_splitted = ptl.tag_splitted + (int32_t)(1u << (ptl.split_times - 1));
            ptls[i] = ptl;
        }
    }
}

/* ======================================================================
 * Distributions
 * ====================================================================== */

static int64_t floor_i64(double v)
{
    const double f = floor(v);
    if (!(f > -9.0e18 && f < 9.0e18)) return INT64_MIN;
    return (int64_t)f;
}

static void local_geometry(const orc_params* prm, const orc_dist_config* dc, int k,
                           int64_t nr[3], double dd[3])
{
    const int rx = dc->local[k][3], ry = dc->local[k][4], rz = dc->local[k][5];
    nr[0] = (prm->nx + rx - 1) / rx;
    nr[1] = (prm->ny + ry - 1) / ry;
    nr[2] = (prm->nz + rz - 1) / rz;
    dd[0] = prm->lx / (double)nr[0];
    dd[1] = prm->ly / (double)nr[1];
    dd[2] = prm->lz / (double)nr[2];
}

int64_t orc_local_dist_size(const orc_params* prm, const orc_dist_config* dc, int k)
{
    if (!dc->local_dist || !dc->local[k][0]) return 0;
    int64_t nr[3];
    double dd[3];
    local_geometry(prm, dc, k, nr, dd);
    return (int64_t)dc->local[k][2] * dc->local[k][1] * nr[0] * nr[1] * nr[2];
}

/* calc_particle_distributions, diagnostics.f90:738-906, with the bin set-up
 * of init_particle_distributions (:178-226) and
 * init_local_particle_distributions (:263-284).  Bins are accumulated into
 * the caller's zeroed arrays (Fortran order, mu fastest).  An index outside
 * the array (p == pmax landing in bin npp+1) is skipped. */
void orc_calc_particle_distributions(const orc_params* prm, const orc_dist_config* dc,
                                     const orc_particle* ptls, int64_t nptl,
                                     double* fglobal, double* flocal[4])
{
    const int npp = dc->npp_global, nmu = dc->nmu_global;
    const double pmin_log = log10(prm->pmin);
    const double pmax_log = log10(prm->pmax);
    const double dp_log = (pmax_log - pmin_log) / npp;
    const double dmu = (double)(2.0f / (float)nmu);
    int64_t nr[4][3];
    double dd[4][3], lp0[4], ldp[4], ldmu[4];
    for (int k = 0; k < 4; ++k) {
        if (!dc->local_dist || !dc->local[k][0]) continue;
        local_geometry(prm, dc, k, nr[k], dd[k]);
        lp0[k] = log10(dc->local_p[k][0]);
        ldp[k] = (log10(dc->local_p[k][1]) - lp0[k]) / dc->local[k][1];
        ldmu[k] = (double)(2.0f / (float)dc->local[k][2]);
    }
    for (int64_t i = 0; i < nptl; ++i) {
        const orc_particle* q = &ptls[i];
        const double p = q->p, mu = q->mu, weight = q->weight;
        if (p > prm->pmin && p <= prm->pmax && mu >= -1.0 && mu <= 1.0) {
            const int64_t ip = floor_i64((log10(p) - pmin_log) / dp_log) + 1;
            const int64_t imu = floor_i64((mu + 1.0) / dmu) + 1;
            if (ip >= 1 && ip <= npp && imu >= 1 && imu <= nmu)
                fglobal[(imu - 1) + (int64_t)nmu * (ip - 1)] += q->weight;
        }
        if (!dc->local_dist) continue;
        for (int k = 0; k < 4; ++k) {
            if (!dc->local[k][0]) continue;
            const int npb = dc->local[k][1], nmk = dc->local[k][2];
            const int64_t ix = floor_i64((q->x - prm->xmin) / dd[k][0]) + 1;
            const int64_t iy = floor_i64((q->y - prm->ymin) / dd[k][1]) + 1;
            const int64_t iz = floor_i64((q->z - prm->zmin) / dd[k][2]) + 1;
            const int64_t ip = floor_i64((log10(p) - lp0[k]) / ldp[k]) + 1;
            const int64_t imu = floor_i64((mu + 1.0) / ldmu[k]) + 1;
            const int condx = ix >= 1 && ix <= nr[k][0];
            const int condy = iy >= 1 && iy <= nr[k][1];
            const int condz = iz >= 1 && iz <= nr[k][2];
            const int condp = ip > 0 && ip < npb;
            const int condmu = imu >= 1 && imu <= nmk;
            if (condx && condy && condz && condp && condmu) {
                const int64_t idx = (imu - 1) + nmk * ((ip - 1) + npb * ((ix - 1) + nr[k][0] * ((iy - 1) + nr[k][1] * (iz - 1))));
                flocal[k][idx] += weight;
            }
        }
    }
}

/* Sizes of calc_escaped_distributions' arrays: k < 0 the global fescaped
 * (nmu_global, npp_global, 2 ndim), k = 0..3 local set k+1's face arrays
 * x (nmu, npbins, nry, nrz, 2), y (nmu, npbins, nrx, nrz, 2) and in 3D
 * z (nmu, npbins, nrx, nry, 2) back to back (diagnostics.f90:188-191,359-395). */
int64_t orc_escaped_dist_size(const orc_params* prm, const orc_dist_config* dc, int k)
{
    if (k < 0) return (int64_t)dc->nmu_global * dc->npp_global * 2 * prm->ndim;
    if (!dc->local_dist || !dc->local[k][0]) return 0;
    int64_t nr[3];
    double dd[3];
    local_geometry(prm, dc, k, nr, dd);
    const int64_t per = (int64_t)dc->local[k][2] * dc->local[k][1] * 2;
    return per * (nr[1] * nr[2] + nr[0] * nr[2] + (prm->ndim > 2 ? nr[0] * nr[1] : 0));
}

/* calc_escaped_distributions, diagnostics.f90:913-1173, over the particles
 * remove_particles copied out (:5382-5390), Cartesian grid.  The global bins
 * take p in (pmin, pmax] and mu in [-1, 1] with face index |count_flag|;
 * local set k bins the two coordinates along the escape face and drops the
 * top momentum bin (condp = ip > 0 and ip < npbins).  Accumulates into the
 * caller's zeroed arrays (sizes from orc_escaped_dist_size). */
void orc_calc_escaped_distributions(const orc_params* prm, const orc_dist_config* dc,
                                    const orc_particle* escaped, int64_t nesc,
                                    double* fescaped, double* fesc[4])
{
    const int npp = dc->npp_global, nmu = dc->nmu_global;
    const double pmin_log = log10(prm->pmin);
    const double dp_log = (log10(prm->pmax) - pmin_log) / npp;
    const double dmu = (double)(2.0f / (float)nmu);
    int64_t nr[4][3];
    double dd[4][3], lp0[4], ldp[4], ldmu[4];
    for (int k = 0; k < 4; ++k) {
        if (!dc->local_dist || !dc->local[k][0]) continue;
        local_geometry(prm, dc, k, nr[k], dd[k]);
        lp0[k] = log10(dc->local_p[k][0]);
        ldp[k] = (log10(dc->local_p[k][1]) - lp0[k]) / dc->local[k][1];
        ldmu[k] = (double)(2.0f / (float)dc->local[k][2]);
    }
    for (int64_t i = 0; i < nesc; ++i) {
        const orc_particle* q = &escaped[i];
        const double x = q->x, y = q->y, z = q->z, p = q->p, mu = q->mu, weight = q->weight;
        if (q->count_flag >= 0) continue;
        if (p > prm->pmin && p <= prm->pmax && mu >= -1.0 && mu <= 1.0) {
            const int64_t ip = floor_i64((log10(p) - pmin_log) / dp_log) + 1;
            const int64_t imu = floor_i64((mu + 1.0) / dmu) + 1;
            const int ibc = -q->count_flag;
            if (ip >= 1 && ip <= npp && imu >= 1 && imu <= nmu && ibc <= 2 * prm->ndim)
                fescaped[(imu - 1) + (int64_t)nmu * ((ip - 1) + (int64_t)npp * (ibc - 1))] += q->weight;
        }
        if (!dc->local_dist) continue;
        for (int k = 0; k < 4; ++k) {
            if (!dc->local[k][0]) continue;
            const int npb = dc->local[k][1], nmk = dc->local[k][2];
            const int64_t ix = floor_i64((x - prm->xmin) / dd[k][0]) + 1;
            const int64_t iy = floor_i64((y - prm->ymin) / dd[k][1]) + 1;
            const int64_t iz = floor_i64((z - prm->zmin) / dd[k][2]) + 1;
            const int64_t ip = floor_i64((log10(p) - lp0[k]) / ldp[k]) + 1;
            const int64_t imu = floor_i64((mu + 1.0) / ldmu[k]) + 1;
            const int condx = ix >= 1 && ix <= nr[k][0];
            const int condy = iy >= 1 && iy <= nr[k][1];
            const int condz = iz >= 1 && iz <= nr[k][2];
            const int condp = ip > 0 && ip < npb;
            const int condmu = imu >= 1 && imu <= nmk;
            if (!(condp && condmu)) continue;
            const int64_t per = (int64_t)nmk * npb;   /* one (mu, p) plane */
            const int64_t nx = nr[k][0], ny = nr[k][1], nz = nr[k][2];
            const int64_t off_y = per * ny * nz * 2, off_z = off_y + per * nx * nz * 2;
            const int64_t ipm = (imu - 1) + nmk * (ip - 1);
            double* f = fesc[k];
            switch (q->count_flag) {
            case ORC_ESC_LX: if (condy && condz) f[ipm + per * ((iy - 1) + ny * (iz - 1))] += weight; break;
            case ORC_ESC_HX: if (condy && condz) f[ipm + per * ((iy - 1) + ny * ((iz - 1) + nz))] += weight; break;
            case ORC_ESC_LY: if (condx && condz) f[off_y + ipm + per * ((ix - 1) + nx * (iz - 1))] += weight; break;
            case ORC_ESC_HY: if (condx && condz) f[off_y + ipm + per * ((ix - 1) + nx * ((iz - 1) + nz))] += weight; break;
            case ORC_ESC_LZ: if (condx && condy) f[off_z + ipm + per * ((ix - 1) + nx * (iy - 1))] += weight; break;
            case ORC_ESC_HZ: if (condx && condy) f[off_z + ipm + per * ((ix - 1) + nx * ((iy - 1) + ny))] += weight; break;
            default: break;
            }
        }
    }
}

void orc_quick_check(const orc_particle* ptls, int64_t nptl, int64_t nptl_split,
                     double leak, double leak_negp, double out[9])
{
    double ntot = 0.0, dtsum = 0.0, pdt_min = 1.0, pdt_max = 0.0, pmax = 0.0;
    for (int64_t i = 0; i < nptl; ++i) {
        ntot = ntot + ptls[i].weight;
        if (ptls[i].dt < pdt_min) pdt_min = ptls[i].dt;
        if (ptls[i].dt > pdt_max) pdt_max = ptls[i].dt;
        dtsum = dtsum + ptls[i].dt;
        if (ptls[i].p > pmax) pmax = ptls[i].p;
    }
    out[0] = (double)nptl;
    out[1] = (double)nptl_split;
    out[2] = ntot;
    out[3] = leak;
    out[4] = leak_negp;
    out[5] = pdt_min;
    out[6] = pdt_max;
    out[7] = nptl > 0 ? dtsum / (double)nptl : 0.0;
    out[8] = pmax;
}
