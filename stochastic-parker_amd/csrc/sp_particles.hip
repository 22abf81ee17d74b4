// sp_particles.hip -- particle population kernels: injection, removal of
// non-INBOX particles (stable stream compaction), splitting, distribution
// histograms and the quick_check reductions.
//
// Compaction and splitting share one three-pass scheme: a per-tile count, a
// single-workgroup exclusive scan of the tile counts, and a per-tile scatter
// that recomputes its in-tile offsets from wave ballots.  Orders are stable,
// so the outcome is deterministic; the reference's remove_particles swaps
// removed particles to the tail instead, which changes only the array order.
#include "sp_device.h"
#include "sp_nodes.h"
#include "sp_particles.h"

#include <algorithm>

namespace sp {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = kBlock * kTileItems;

// ---------------------------------------------------------------------------
// Injection: inject_particles_spatial_uniform + inject_one_particle
// (particle_module.f90:385-530).  Particle i of the batch takes tag
// tag0 + i and draws x, y, z, mu, [momentum draws], t from its own Philox
// stream (counter = tag, 0, block, stream), 4 draws per block.
struct DrawStream {
    uint32_t tag, blk, k0, k1, stream;
    U4 buf;
    int used;
    __device__ double next() {
        if (used == 4) {
            buf = philox4x32_10(tag, 0u, blk, stream, k0, k1);
            ++blk;
            used = 0;
        }
        const uint32_t r = used == 0 ? buf.a : used == 1 ? buf.b : used == 2 ? buf.c : buf.d;
        ++used;
        return u01(r);
    }
};

__device__ void inject_one(const DevParams& prm, const Ptl& P, long long slot, DrawStream& ds,
                           double xt, double yt, double zt, double mu, const InjectArgs& ia,
                           int tag);

__global__ void __launch_bounds__(kBlock) inject_kernel(DevParams prm, Ptl P, long long start,
                                                        long long nptl, long long nmax, int tag0,
                                                        InjectArgs ia) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nptl) return;
    long long slot = start + i;
    if (slot >= nmax - 1) {
        // past nptl_max every new particle lands in the last slot, so that slot
        // keeps the batch's last particle; no other thread writes it
        if (i != nptl - 1) return;
        slot = nmax - 1;
    }
    DrawStream ds{(uint32_t)(tag0 + (int)i), 0u, prm.key0, prm.key1, prm.stream, U4{0, 0, 0, 0}, 4};
    const double mu_max = (double)0.99f;
    const double xt = ds.next() * (ia.box[3] - ia.box[0]) + ia.box[0];
    const double yt = ds.next() * (ia.box[4] - ia.box[1]) + ia.box[1];
    const double zt = ds.next() * (ia.box[5] - ia.box[2]) + ia.box[2];
    const double mu = mu_max * (2.0 * ds.next() - 1.0);
    inject_one(prm, P, slot, ds, xt, yt, zt, mu, ia, tag0 + (int)i);
}

// inject_one_particle (particle_module.f90:385-441): momentum by the
// distribution law, then the time stamp, continuing the particle's draws.
__device__ void inject_one(const DevParams& prm, const Ptl& P, long long slot, DrawStream& ds,
                           double xt, double yt, double zt, double mu, const InjectArgs& ia,
                           int tag) {
    double p = prm.p0;
    if (ia.dist_flag == 0) {
        double ftest = 1.0, fxp = 0.5, ptmp = 0.0;
        int guard = 0;
        while (ftest > fxp && guard < 100000) {
            ptmp = (ds.next() * (prm.pmax - prm.pmin) + prm.pmin) / prm.p0;
            fxp = (ptmp * ptmp) * exp(-(ptmp * ptmp));
            ftest = ds.next() * (double)0.37f;
            ++guard;
        }
        p = ptmp * prm.p0;
    } else if (ia.dist_flag == 1) {
        p = prm.p0;
    } else if (ia.dist_flag == 2) {
        const double r01 = ds.next();
        const double pi = ia.power_index;
        if ((int)pi == 1) {
            p = pow(prm.pmax / prm.p0, r01) * prm.p0;
        } else {
            const double norm = pow(prm.pmax, -pi + 1) - pow(prm.p0, -pi + 1);
            p = pow(r01 * norm + pow(prm.p0, -pi + 1), 1.0 / (-pi + 1));
        }
    }
    P.x[slot] = xt;
    P.y[slot] = yt;
    P.z[slot] = zt;
    P.p[slot] = p;
    P.v[slot] = ia.v0 * p / prm.p0;
    P.mu[slot] = mu;
    P.w[slot] = 1.0;
    P.t[slot] = ia.t0 + ds.next() * ia.dtf;
    P.dt[slot] = ia.dt;
    P.split_times[slot] = 0;
    P.count_flag[slot] = SP_COUNT_FLAG_INBOX;
    P.origin[slot] = ia.origin;
    P.nsteps_tracked[slot] = 0;
    P.nsteps_pushed[slot] = 0;
    P.tag_injected[slot] = tag;
    P.tag_splitted[slot] = 1;
    P.rng_step[slot] = 0u;
}

// ---------------------------------------------------------------------------
// Threshold injections.  The node quantity of the cell counters, in the
// precision the Fortran evaluates it (farray1 is real(sp)): |jz| and |J| in
// FP32, divV summed in FP64 (mhd_data_parallel.f90:2236-2238, :2306-2312,
// :2423-2429, :2488).
__device__ double node_quantity(const DevParams& prm, const ThreshArgs& ta, long long node) {
    float f[32];
    node_components(prm, ta.idxh, ta.idyh, ta.idzh, ta.n8, node, f);
    switch (ta.quantity) {
    case SP_INJECT_JZ:
        return (double)fabsf(f[23] - f[21]);
    case SP_INJECT_ABSJ: {
        const float a = f[25] - f[27], b = f[26] - f[22], c = f[21] - f[23];
        return (double)sqrtf(a * a + b * b + c * c);
    }
    case SP_INJECT_DIVV: {
        double d = (double)f[8];
        if (prm.ndim > 1) {
            d = d + (double)f[12];
            if (prm.ndim > 2) d = d + (double)f[16];
        }
        return -d;
    }
    default:
        return (double)f[3];
    }
}

// get_ncells_large_* (mhd_data_parallel.f90:2211-2499): interior nodes
// strictly inside the box whose quantity exceeds the threshold.  The divV
// counter reads the node two below in each ghosted axis: that routine assigns
// the whole ghosted farray1 slice to an array allocated without ghosts, and
// reallocation on assignment rebounds it to start at 1 (:2423-2446).
__global__ void __launch_bounds__(kBlock) threshold_count_kernel(DevParams prm, ThreshArgs ta,
                                                                 unsigned long long* count) {
    const int nx = prm.nxg - 4;
    const int ny = prm.ndim > 1 ? prm.nyg - 4 : 1;
    const int nz = prm.ndim > 2 ? prm.nzg - 4 : 1;
    const long long total = (long long)nx * ny * nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned c = 0;
    if (t < total) {
        const int fi = (int)(t % nx) + 1;
        const int fj = (int)((t / nx) % ny) + 1;
        const int fk = (int)(t / ((long long)nx * ny)) + 1;
        const double xp = prm.xmin + (fi - 1) * prm.dx;
        const double yp = prm.ymin + (fj - 1) * prm.dy;
        const double zp = prm.zmin + (fk - 1) * prm.dz;
        bool in = xp > ta.box[0] && xp < ta.box[3];
        if (prm.ndim > 1) in = in && yp > ta.box[1] && yp < ta.box[4];
        if (prm.ndim > 2) in = in && zp > ta.box[2] && zp < ta.box[5];
        if (in) {
            const int sh = ta.quantity == SP_INJECT_DIVV ? 2 : 0;
            const int gi = fi + 1 - sh;
            const int gj = prm.ndim > 1 ? fj + 1 - sh : 0;
            const int gk = prm.ndim > 2 ? fk + 1 - sh : 0;
            const long long node = gi + (long long)prm.nxg * (gj + (long long)prm.nyg * gk);
            c = node_quantity(prm, ta, node) > ta.threshold ? 1u : 0u;
        }
    }
    const unsigned long long m = __ballot(c != 0u);
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

// The quantity at a trial point from the frame-1 fields interpolated as
// interp_fields does with rt = 0 (corners k, j, i with i fastest, FP32 node
// values times FP64 weights), in FP64 (particle_module.f90:883-890 jz,
// :1035-1043 |J|, :1318-1323 divV, :1456 rho).
__device__ double point_quantity(const DevParams& prm, const ThreshArgs& ta, double x, double y,
                                 double z) {
    const double px = (x - prm.xmin) / prm.dx;
    const double py = (y - prm.ymin) / prm.dy;
    const double fx = floor(px), fy = floor(py);
    const double rx = (px - (fx + 1.0)) + 1.0, ry = (py - (fy + 1.0)) + 1.0;
    const int ix = clamp_cell(fx, prm.nxg - 4) + 2, iy = clamp_cell(fy, prm.nyg - 4) + 2;
    int iz = 0;
    double rz = 0.0;
    if (prm.ndim > 2) {
        const double pz = (z - prm.zmin) / prm.dz;
        const double fz = floor(pz);
        rz = (pz - (fz + 1.0)) + 1.0;
        iz = clamp_cell(fz, prm.nzg - 4) + 2;
    }
    const double rx1 = 1.0 - rx, ry1 = 1.0 - ry, rz1 = 1.0 - rz;
    const double w[8] = {rx1 * ry1 * rz1, rx * ry1 * rz1, rx1 * ry * rz1, rx * ry * rz1,
                         rx1 * ry1 * rz,  rx * ry1 * rz,  rx1 * ry * rz,  rx * ry * rz};
    // components used: rho, dvx/dx, dvy/dy, dvz/dz and the six off-diagonal dB_i/dx_j
    const int comp[10] = {3, 8, 12, 16, 21, 22, 23, 25, 26, 27};
    double acc[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) acc[c] = 0.0;
    const int nk = prm.ndim > 2 ? 2 : 1;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const long long node = (ix + i) + (long long)prm.nxg * ((iy + j) + (long long)prm.nyg * (iz + k));
                float f[32];
                node_components(prm, ta.idxh, ta.idyh, ta.idzh, ta.n8, node, f);
                const double wq = w[k * 4 + j * 2 + i];
#pragma unroll
                for (int c = 0; c < 10; ++c) acc[c] = acc[c] + (double)f[comp[c]] * wq;
            }
    const double rho = acc[0], dvx_dx = acc[1], dvy_dy = acc[2], dvz_dz = acc[3];
    const double dbx_dy = acc[4], dbx_dz = acc[5], dby_dx = acc[6], dby_dz = acc[7];
    const double dbz_dx = acc[8], dbz_dy = acc[9];
    switch (ta.quantity) {
    case SP_INJECT_JZ:
        return fabs(dby_dx - dbx_dy);
    case SP_INJECT_ABSJ: {
        const double a = dby_dz - dbz_dy, b = dbz_dx - dbx_dz, c = dbx_dy - dby_dx;
        return sqrt(a * a + b * b + c * c);
    }
    case SP_INJECT_DIVV: {
        double d = dvx_dx;
        if (prm.ndim > 1) {
            d = d + dvy_dy;
            if (prm.ndim > 2) d = d + dvz_dz;
        }
        return d;
    }
    default:
        return rho;
    }
}

// inject_particles_at_large_{jz,absj,divv,rho} (particle_module.f90:785-1470),
// one rank: particle i (tag tag0 + i) draws x, y, z over the domain until the
// point lies in the box (inclusive) and its quantity passes; a point outside
// the box scores -3 (jz, |J|), divV = 3 or rho = 0.  The loop tests its initial
// value first (jz, |J|: -2, divV: 2, rho: 0); when that passes, no point is
// drawn and the particle sits at the box's low corner.  Then mu and
// inject_one_particle.  A particle still failing after max_trials draws
// counts in *fails; with commit = 0 nothing is stored.
__global__ void __launch_bounds__(kBlock) threshold_inject_kernel(DevParams prm, Ptl P,
                                                                  long long start, long long nptl,
                                                                  long long nmax, int tag0,
                                                                  InjectArgs ia, ThreshArgs ta,
                                                                  unsigned long long* fails) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nptl) return;
    long long slot = start + i;
    bool store = ta.commit != 0;
    if (slot >= nmax - 1) {
        // the last slot keeps the batch's last particle (as in inject_kernel)
        if (i != nptl - 1) store = false;
        slot = nmax - 1;
    }
    DrawStream ds{(uint32_t)(tag0 + (int)i), 0u, prm.key0, prm.key1, prm.stream, U4{0, 0, 0, 0}, 4};
    const bool divv = ta.quantity == SP_INJECT_DIVV;
    const double v_init = divv ? 2.0 : (ta.quantity == SP_INJECT_RHO ? 0.0 : -2.0);
    const double v_out = divv ? 3.0 : (ta.quantity == SP_INJECT_RHO ? 0.0 : -3.0);
    double xt = ta.box[0], yt = ta.box[1], zt = ta.box[2];
    double v = v_init;
    long long trials = 0;
    while ((divv ? -v : v) < ta.threshold && trials < ta.max_trials) {
        xt = ds.next() * (prm.xmax - prm.xmin) + prm.xmin;
        yt = ds.next() * (prm.ymax - prm.ymin) + prm.ymin;
        zt = ds.next() * (prm.zmax - prm.zmin) + prm.zmin;
        if (xt >= ta.box[0] && xt <= ta.box[3] && yt >= ta.box[1] && yt <= ta.box[4] &&
            zt >= ta.box[2] && zt <= ta.box[5])
            v = point_quantity(prm, ta, xt, yt, zt);
        else
            v = v_out;
        ++trials;
    }
    if ((divv ? -v : v) < ta.threshold) {
        atomicAdd(fails, 1ull);
        return;
    }
    if (!store) return;
    const double mu = (double)0.99f * (2.0 * ds.next() - 1.0);
    inject_one(prm, P, slot, ds, xt, yt, zt, mu, ia, tag0 + (int)i);
}

// ---------------------------------------------------------------------------
// Predicates for the scan passes.
// mode 0: keep (count_flag == INBOX); mode 1: split candidate.
__device__ __forceinline__ double powi_rl(double x, int m) {
    // right-to-left binary powering, the method of libgcc's __powidf2 that
    // gfortran uses for real ** integer (particle_module.f90:5441)
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    double y = (n % 2u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

__device__ __forceinline__ bool predicate(const Ptl& P, long long i, int mode, const SplitArgs& sa) {
    if (mode == 0) return P.count_flag[i] == SP_COUNT_FLAG_INBOX;
    const double thr = sa.pmin_split * sa.p0 * powi_rl(sa.split_ratio, (int)P.split_times[i]);
    const double p = P.p[i];
    return p > thr && p <= sa.pmax;
}

// In-block exclusive prefix of a per-thread count (kBlock threads, 4 waves).
__device__ __forceinline__ unsigned block_exclusive(unsigned v, unsigned* lds_wave, unsigned& total) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const unsigned y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) lds_wave[wv] = x;
    __syncthreads();
    unsigned wbase = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) {
        const unsigned t = lds_wave[k];
        if (k < wv) wbase += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

__device__ __forceinline__ long long floor_ll(double v) {
    const double f = floor(v);
    if (!(f > -9.0e18 && f < 9.0e18)) return LLONG_MIN;
    return (long long)f;
}

// calc_escaped_distributions (diagnostics.f90:913-1173) for one escaped
// particle, face ibc = |count_flag| (1..6 = low x, high x, low y, ...).  The
// global bins take p in (pmin, pmax] and mu in [-1, 1]; the local face maps
// drop the top momentum bin and bin the two coordinates along the face.
__device__ void escaped_hist(const Ptl& P, long long i, int ibc, const DistArgs& da,
                             const EscArgs& ea) {
    const double p = P.p[i], mu = P.mu[i];
    const unsigned long long wfx = weight_fx(P.w[i]);
    const double lp = log10(p);
    if (p > da.pmin && p <= da.pmax && mu >= -1.0 && mu <= 1.0) {
        const long long ip = floor_ll((lp - da.pmin_log) / da.dp_log) + 1;
        const long long imu = floor_ll((mu + 1.0) / da.dmu) + 1;
        if (ip >= 1 && ip <= da.npp && imu >= 1 && imu <= da.nmu && ibc <= 2 * ea.ndim)
            atomicAdd(&ea.glob[(imu - 1) + (long long)da.nmu * ((ip - 1) + (long long)da.npp * (ibc - 1))],
                      wfx);
    }
    if (!da.local_dist) return;
    const int axis = (ibc - 1) >> 1, side = (ibc - 1) & 1;
    if (axis >= ea.ndim) return;
    const double x = P.x[i], y = P.y[i], z = P.z[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const LocalArgs& L = da.loc[k];
        unsigned long long* f = ea.face[k][axis];
        if (!L.enabled || !f) continue;
        const long long ix = floor_ll((x - da.xmin) / L.dd[0]) + 1;
        const long long iy = floor_ll((y - da.ymin) / L.dd[1]) + 1;
        const long long iz = floor_ll((z - da.zmin) / L.dd[2]) + 1;
        const long long ip = floor_ll((lp - L.pmin_log) / L.dp_log) + 1;
        const long long imu = floor_ll((mu + 1.0) / L.dmu) + 1;
        if (!(ip > 0 && ip < L.npbins && imu >= 1 && imu <= L.nmu)) continue;
        const bool cx = ix >= 1 && ix <= L.nr[0];
        const bool cy = iy >= 1 && iy <= L.nr[1];
        const bool cz = iz >= 1 && iz <= L.nr[2];
        long long ib, ic, nb, nc;
        bool ok;
        if (axis == 0) { ok = cy && cz; ib = iy; ic = iz; nb = L.nr[1]; nc = L.nr[2]; }
        else if (axis == 1) { ok = cx && cz; ib = ix; ic = iz; nb = L.nr[0]; nc = L.nr[2]; }
        else { ok = cx && cy; ib = ix; ic = iy; nb = L.nr[0]; nc = L.nr[1]; }
        if (!ok) continue;
        const long long idx = (imu - 1) + (long long)L.nmu * ((ip - 1) + (long long)L.npbins *
                              ((ib - 1) + nb * ((ic - 1) + nc * side)));
        atomicAdd(&f[idx], wfx);
    }
}

__global__ void __launch_bounds__(kBlock) tile_count_kernel(Ptl P, long long n, int mode,
                                                            SplitArgs sa, unsigned* tile_cnt,
                                                            unsigned* tile_esc, DistArgs da,
                                                            EscArgs ea) {
    __shared__ unsigned lds[kBlock / kWave];
    const long long t0 = (long long)blockIdx.x * kTile;
    unsigned c = 0, e = 0;
#pragma unroll
    for (int it = 0; it < kTileItems; ++it) {
        const long long i = t0 + (long long)it * kBlock + threadIdx.x;
        if (i < n) {
            const bool hit = predicate(P, i, mode, sa);
            c += hit ? 1u : 0u;
            if (mode == 0) {
                const int cf = P.count_flag[i];
                e += cf < 0 ? 1u : 0u;
                if (cf < 0 && ea.on) escaped_hist(P, i, -cf, da, ea);
            } else {
                // split candidates whose copy would weigh 2^-(k+1) < 2^-32
                e += hit && P.split_times[i] >= kMaxSplitTimes ? 1u : 0u;
            }
        }
    }
    unsigned tc = 0, te = 0;
    block_exclusive(c, lds, tc);
    block_exclusive(e, lds, te);
    if (threadIdx.x == 0) {
        tile_cnt[blockIdx.x] = tc;
        if (tile_esc) tile_esc[blockIdx.x] = te;
    }
}

// Exclusive scan of ntiles counts by one workgroup; writes offsets and the
// grand totals (cnt -> totals[0], esc -> totals[1]).
__global__ void __launch_bounds__(1024) tile_scan_kernel(const unsigned* tile_cnt,
                                                         const unsigned* tile_esc, long long ntiles,
                                                         unsigned long long* tile_off,
                                                         unsigned long long* totals) {
    __shared__ unsigned long long wsum[1024 / kWave];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    unsigned long long esc = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (long long b = 0; b < ntiles; b += 1024) {
        const long long i = b + threadIdx.x;
        const unsigned long long v = i < ntiles ? tile_cnt[i] : 0ull;
        if (tile_esc && i < ntiles) esc += tile_esc[i];
        unsigned long long x = v;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const unsigned long long y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == kWave - 1) wsum[wv] = x;
        __syncthreads();
        unsigned long long wbase = 0, tot = 0;
        for (int k = 0; k < 1024 / kWave; ++k) {
            if (k < wv) wbase += wsum[k];
            tot += wsum[k];
        }
        if (i < ntiles) tile_off[i] = carry + wbase + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    // escaped total: block reduction
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) esc += __shfl_xor(esc, off);
    if (lane == 0) wsum[wv] = esc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long e = 0;
        for (int k = 0; k < 1024 / kWave; ++k) e += wsum[k];
        totals[0] = carry;
        totals[1] = e;
    }
}

__device__ __forceinline__ void copy_particle(const Ptl& S, long long i, const Ptl& D, long long o) {
    store_row<false>(D, o, load_row(S, i));
}

__global__ void __launch_bounds__(kBlock) compact_scatter_kernel(Ptl S, Ptl D, long long n,
                                                                 const unsigned long long* tile_off) {
    __shared__ unsigned lds[kBlock / kWave];
    const long long t0 = (long long)blockIdx.x * kTile;
    unsigned long long base = tile_off[blockIdx.x];
    const SplitArgs none{};
#pragma unroll 1
    for (int it = 0; it < kTileItems; ++it) {
        const long long i = t0 + (long long)it * kBlock + threadIdx.x;
        const bool keep = i < n && predicate(S, i, 0, none);
        unsigned tot = 0;
        const unsigned ex = block_exclusive(keep ? 1u : 0u, lds, tot);
        if (keep) copy_particle(S, i, D, (long long)(base + ex));
        base += tot;
    }
}

// split_particle (particle_module.f90:5430-5480): a candidate with global
// exclusive rank r is split iff n + r < nmax (the reference returns at the
// first candidate that would overflow nptl_max); its copy goes to n + r.
__global__ void __launch_bounds__(kBlock) split_apply_kernel(Ptl P, long long n, long long nmax,
                                                             SplitArgs sa,
                                                             const unsigned long long* tile_off) {
    __shared__ unsigned lds[kBlock / kWave];
    const long long t0 = (long long)blockIdx.x * kTile;
    unsigned long long base = tile_off[blockIdx.x];
#pragma unroll 1
    for (int it = 0; it < kTileItems; ++it) {
        const long long i = t0 + (long long)it * kBlock + threadIdx.x;
        const bool cand = i < n && predicate(P, i, 1, sa);
        unsigned tot = 0;
        const unsigned ex = block_exclusive(cand ? 1u : 0u, lds, tot);
        if (cand) {
            const long long o = n + (long long)(base + ex);
            if (o < nmax) {
                // k < kMaxSplitTimes: the host refuses the call otherwise
                const int k = (int)P.split_times[i];
                const double wnew = ldexp(1.0, -(1 + k));  // 0.5**(1.0 + split_times)
                P.w[i] = wnew;
                P.split_times[i] = (int8_t)(k + 1);
                copy_particle(P, i, P, o);
                // tag_splitted + 2**k in wrapping 32-bit arithmetic
                P.tag_splitted[o] = (int32_t)((uint32_t)P.tag_splitted[i] + (1u << k));
            }
        }
        base += tot;
    }
}

// ---------------------------------------------------------------------------
// Distributions (diagnostics.f90:738-906).  Weights are accumulated as
// unsigned 2^-32 fixed point, so every sum is exact and independent of the
// order of the atomics.
//
// A block takes one contiguous slice of the particle array; after the
// spatial sort its particles occupy a few map cells, so most of their
// local-map updates fall on a few hundred bins.  The block sums those in an
// LDS hash table (open addressing, keys = 4 * bin + set) and adds each
// occupied slot to global memory once; an update that finds no free slot
// within kProbe probes goes straight to global memory.  f(p) has its own
// LDS bins.
constexpr int kMapSlotsLog2 = 11;
constexpr int kMapSlots = 1 << kMapSlotsLog2;
constexpr int kProbe = 8;
constexpr unsigned kEmptyKey = 0xffffffffu;

__device__ __forceinline__ bool lds_hash_add(unsigned* tk, unsigned long long* tv, unsigned key,
                                             unsigned long long w) {
    const unsigned h = (key * 2654435761u) >> (32 - kMapSlotsLog2);
#pragma unroll 1
    for (int probe = 0; probe < kProbe; ++probe) {
        const unsigned s = (h + (unsigned)probe) & (kMapSlots - 1);
        unsigned cur = tk[s];
        if (cur == kEmptyKey) cur = atomicCAS(&tk[s], kEmptyKey, key);
        if (cur == kEmptyKey || cur == key) {
            atomicAdd(&tv[s], w);
            return true;
        }
    }
    return false;
}

__global__ void __launch_bounds__(kBlock) dist_kernel(Ptl P, long long n, long long slice,
                                                      DistArgs da, int use_hash,
                                                      unsigned long long* gglob,
                                                      unsigned long long* gloc0,
                                                      unsigned long long* gloc1,
                                                      unsigned long long* gloc2,
                                                      unsigned long long* gloc3) {
    extern __shared__ unsigned long long sh[];
    __shared__ unsigned tkey[kMapSlots];
    __shared__ unsigned long long tval[kMapSlots];
    const int nglob = da.npp * da.nmu;
    for (int b = threadIdx.x; b < nglob; b += blockDim.x) sh[b] = 0ull;
    for (int s = threadIdx.x; s < kMapSlots; s += blockDim.x) {
        tkey[s] = kEmptyKey;
        tval[s] = 0ull;
    }
    __syncthreads();
    unsigned long long* gl[4] = {gloc0, gloc1, gloc2, gloc3};
    const long long i0 = (long long)blockIdx.x * slice;
    const long long i1 = i0 + slice < n ? i0 + slice : n;
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double p = P.p[i], mu = P.mu[i];
        const unsigned long long wfx = weight_fx(P.w[i]);
        const double lp = log10(p);
        if (p > da.pmin && p <= da.pmax && mu >= -1.0 && mu <= 1.0) {
            const long long ip = floor_ll((lp - da.pmin_log) / da.dp_log) + 1;
            const long long imu = floor_ll((mu + 1.0) / da.dmu) + 1;
            if (ip >= 1 && ip <= da.npp && imu >= 1 && imu <= da.nmu)
                atomicAdd(&sh[(imu - 1) + (long long)da.nmu * (ip - 1)], wfx);
        }
        if (!da.local_dist) continue;
        const double x = P.x[i], y = P.y[i], z = P.z[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const LocalArgs& L = da.loc[k];
            if (!L.enabled) continue;
            const long long ix = floor_ll((x - da.xmin) / L.dd[0]) + 1;
            const long long iy = floor_ll((y - da.ymin) / L.dd[1]) + 1;
            const long long iz = floor_ll((z - da.zmin) / L.dd[2]) + 1;
            const long long ip = floor_ll((lp - L.pmin_log) / L.dp_log) + 1;
            const long long imu = floor_ll((mu + 1.0) / L.dmu) + 1;
            const bool ok = ix >= 1 && ix <= L.nr[0] && iy >= 1 && iy <= L.nr[1] && iz >= 1 &&
                            iz <= L.nr[2] && ip > 0 && ip < L.npbins && imu >= 1 && imu <= L.nmu;
            if (ok) {
                const long long idx =
                    (imu - 1) + (long long)L.nmu * ((ip - 1) + (long long)L.npbins *
                                 ((ix - 1) + L.nr[0] * ((iy - 1) + L.nr[1] * (iz - 1))));
                if (!use_hash || !lds_hash_add(tkey, tval, (unsigned)(idx * 4 + k), wfx))
                    atomicAdd(&gl[k][idx], wfx);
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nglob; b += blockDim.x)
        if (sh[b]) atomicAdd(&gglob[b], sh[b]);
    for (int s = threadIdx.x; s < kMapSlots; s += blockDim.x) {
        const unsigned key = tkey[s];
        if (key != kEmptyKey) atomicAdd(&gl[key & 3u][key >> 2], tval[s]);
    }
}

// f(p) per parameter variant: the global momentum bins of dist_kernel
// (summed over mu) for the particles whose origin is each variant, plus the
// particle count of each variant.  A block owns LDS bins for the chunk of
// variants blockIdx.y * vchunk .. + vchunk - 1 and reads its contiguous
// slice of particles, skipping other variants; it adds its bins and counts
// to global memory once, non-zero entries only.  Fixed-point weights, so the
// sums do not depend on the order of the atomics.
constexpr int kVarBlock = 1024;
constexpr int kVarBins = 16384;        // 128 KiB of u64 bins per block
constexpr int kVarCounts = 1024;

__global__ void __launch_bounds__(kVarBlock) variant_dist_kernel(Ptl P, long long n,
                                                                 long long slice, DistArgs da,
                                                                 int nvar, int vchunk,
                                                                 unsigned long long* hist,
                                                                 unsigned long long* counts) {
    __shared__ unsigned long long bins[kVarBins];
    __shared__ unsigned cnt[kVarCounts];
    const int v0 = (int)blockIdx.y * vchunk;
    const int nv = nvar - v0 < vchunk ? nvar - v0 : vchunk;
    const int nb = nv * da.npp;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) bins[b] = 0ull;
    for (int b = threadIdx.x; b < nv; b += blockDim.x) cnt[b] = 0u;
    __syncthreads();
    const long long i0 = (long long)blockIdx.x * slice;
    const long long i1 = i0 + slice < n ? i0 + slice : n;
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int v = P.origin[i] - v0;
        if ((unsigned)v >= (unsigned)nv) continue;
        atomicAdd(&cnt[v], 1u);
        const double p = P.p[i];
        if (!(p > da.pmin && p <= da.pmax)) continue;
        const long long ip = floor_ll((log10(p) - da.pmin_log) / da.dp_log) + 1;
        if (ip >= 1 && ip <= da.npp) atomicAdd(&bins[(ip - 1) + (long long)da.npp * v], weight_fx(P.w[i]));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (bins[b]) atomicAdd(&hist[(long long)da.npp * v0 + b], bins[b]);
    for (int b = threadIdx.x; b < nv; b += blockDim.x)
        if (cnt[b]) atomicAdd(&counts[v0 + b], (unsigned long long)cnt[b]);
}

// quick_check / get_pmax_global partials, one row per block:
// {sum w (fixed point, as double bits), sum dt, min dt, max dt, max p}
__global__ void __launch_bounds__(kBlock) stats_kernel(Ptl P, long long n, double* part) {
    __shared__ double s_sum[kBlock / kWave], s_min[kBlock / kWave], s_max[kBlock / kWave],
        s_pmx[kBlock / kWave];
    __shared__ unsigned long long s_w[kBlock / kWave];
    unsigned long long wsum = 0;
    double dsum = 0.0, dmin = 1.0e300, dmax = -1.0e300, pmx = -1.0e300;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        wsum += weight_fx(P.w[i]);
        const double dt = P.dt[i];
        dsum += dt;
        dmin = fmin(dmin, dt);
        dmax = fmax(dmax, dt);
        pmx = fmax(pmx, P.p[i]);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        wsum += __shfl_xor(wsum, off);
        dsum += __shfl_xor(dsum, off);
        dmin = fmin(dmin, __shfl_xor(dmin, off));
        dmax = fmax(dmax, __shfl_xor(dmax, off));
        pmx = fmax(pmx, __shfl_xor(pmx, off));
    }
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (lane == 0) {
        s_w[wv] = wsum; s_sum[wv] = dsum; s_min[wv] = dmin; s_max[wv] = dmax; s_pmx[wv] = pmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long W = 0;
        double S = 0.0, MN = 1.0e300, MX = -1.0e300, PM = -1.0e300;
        for (int k = 0; k < kBlock / kWave; ++k) {
            W += s_w[k]; S += s_sum[k];
            MN = fmin(MN, s_min[k]); MX = fmax(MX, s_max[k]); PM = fmax(PM, s_pmx[k]);
        }
        double* row = part + 5 * (long long)blockIdx.x;
        row[0] = __longlong_as_double((long long)W);
        row[1] = S; row[2] = MN; row[3] = MX; row[4] = PM;
    }
}

// Plain streaming copy (16 B per lane) used to calibrate memory counters.
__global__ void __launch_bounds__(kBlock) copy_kernel(const float4* __restrict__ a,
                                                      float4* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace

hipError_t launch_copy(const void* a, void* b, long long n16, hipStream_t s) {
    const long long grid = std::min<long long>((n16 + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s,
                       static_cast<const float4*>(a), static_cast<float4*>(b), n16);
    return hipGetLastError();
}

hipError_t launch_inject(const DevParams& prm, const Ptl& P, long long start, long long nptl,
                         long long nmax, int tag0, const InjectArgs& ia, hipStream_t s) {
    if (nptl <= 0) return hipSuccess;
    const long long grid = (nptl + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(inject_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s, prm, P, start, nptl,
                       nmax, tag0, ia);
    return hipGetLastError();
}

long long num_tiles(long long n) { return (n + kTile - 1) / kTile; }

hipError_t launch_threshold_count(const DevParams& prm, const ThreshArgs& ta,
                                  unsigned long long* count, hipStream_t s) {
    const long long nx = prm.nxg - 4;
    const long long ny = prm.ndim > 1 ? prm.nyg - 4 : 1;
    const long long nz = prm.ndim > 2 ? prm.nzg - 4 : 1;
    const long long total = nx * ny * nz;
    const long long grid = (total + kBlock - 1) / kBlock;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(threshold_count_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s, prm, ta,
                       count);
    return hipGetLastError();
}

hipError_t launch_threshold_inject(const DevParams& prm, const Ptl& P, long long start,
                                   long long nptl, long long nmax, int tag0, const InjectArgs& ia,
                                   const ThreshArgs& ta, unsigned long long* fails, hipStream_t s) {
    if (nptl <= 0) return hipSuccess;
    const long long grid = (nptl + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(threshold_inject_kernel, dim3((unsigned)grid), dim3(kBlock), 0, s, prm, P,
                       start, nptl, nmax, tag0, ia, ta, fails);
    return hipGetLastError();
}

hipError_t launch_tile_count(const Ptl& P, long long n, int mode, const SplitArgs& sa,
                             unsigned* tile_cnt, unsigned* tile_esc, const DistArgs* da,
                             const EscArgs* ea, hipStream_t s) {
    const long long nt = num_tiles(n);
    if (nt == 0) return hipSuccess;
    const DistArgs d = da ? *da : DistArgs{};
    EscArgs e = ea ? *ea : EscArgs{};
    if (!da || mode != 0) e.on = 0;
    hipLaunchKernelGGL(tile_count_kernel, dim3((unsigned)nt), dim3(kBlock), 0, s, P, n, mode, sa,
                       tile_cnt, tile_esc, d, e);
    return hipGetLastError();
}

hipError_t launch_tile_scan(const unsigned* tile_cnt, const unsigned* tile_esc, long long ntiles,
                            unsigned long long* tile_off, unsigned long long* totals, hipStream_t s) {
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, s, tile_cnt, tile_esc, ntiles,
                       tile_off, totals);
    return hipGetLastError();
}

hipError_t launch_compact_scatter(const Ptl& S, const Ptl& D, long long n,
                                  const unsigned long long* tile_off, hipStream_t s) {
    const long long nt = num_tiles(n);
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(compact_scatter_kernel, dim3((unsigned)nt), dim3(kBlock), 0, s, S, D, n,
                       tile_off);
    return hipGetLastError();
}

hipError_t launch_split_apply(const Ptl& P, long long n, long long nmax, const SplitArgs& sa,
                              const unsigned long long* tile_off, hipStream_t s) {
    const long long nt = num_tiles(n);
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(split_apply_kernel, dim3((unsigned)nt), dim3(kBlock), 0, s, P, n, nmax, sa,
                       tile_off);
    return hipGetLastError();
}

hipError_t launch_dist(const Ptl& P, long long n, const DistArgs& da, unsigned long long* gglob,
                       unsigned long long* const gloc[4], const long long loc_size[4], int grid,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    // a hash key holds 4 * bin + set in 32 bits (all ones is the empty slot)
    int use_hash = 1;
    for (int k = 0; k < 4; ++k)
        if (da.local_dist && da.loc[k].enabled && loc_size[k] >= (1ll << 30)) use_hash = 0;
    const long long slice = ((n + grid - 1) / grid + kBlock - 1) / kBlock * kBlock;
    const int nblk = (int)((n + slice - 1) / slice);
    const size_t shm = sizeof(unsigned long long) * (size_t)da.npp * da.nmu;
    hipLaunchKernelGGL(dist_kernel, dim3(nblk), dim3(kBlock), shm, s, P, n, slice, da, use_hash,
                       gglob, gloc[0], gloc[1], gloc[2], gloc[3]);
    return hipGetLastError();
}

hipError_t launch_variant_dist(const Ptl& P, long long n, const DistArgs& da, int nvar,
                               unsigned long long* hist, unsigned long long* counts, int ncu,
                               hipStream_t s) {
    if (n <= 0 || nvar <= 0) return hipSuccess;
    if (da.npp > kVarBins) return hipErrorInvalidValue;
    const int vchunk = std::min(std::min(nvar, kVarBins / da.npp), kVarCounts);
    const int nchunk = (nvar + vchunk - 1) / vchunk;
    // one block per CU per variant chunk, each over a contiguous slice
    const long long want = std::max<long long>(1, ncu);
    const long long slice = ((n + want - 1) / want + kVarBlock - 1) / kVarBlock * kVarBlock;
    const unsigned gx = (unsigned)((n + slice - 1) / slice);
    hipLaunchKernelGGL(variant_dist_kernel, dim3(gx, (unsigned)nchunk), dim3(kVarBlock), 0, s, P,
                       n, slice, da, nvar, vchunk, hist, counts);
    return hipGetLastError();
}

hipError_t launch_stats(const Ptl& P, long long n, double* part, int grid, hipStream_t s) {
    hipLaunchKernelGGL(stats_kernel, dim3(grid), dim3(kBlock), 0, s, P, n, part);
    return hipGetLastError();
}

}  // namespace sp
