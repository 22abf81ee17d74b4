// sp_particles.h -- argument blocks and launchers of the particle population
// kernels (sp_particles.hip).
#pragma once

#include "sp_internal.h"

namespace sp {

constexpr int kTileItems = 8;  // particles per thread in one scan tile

struct InjectArgs {
    double box[6];
    double dt, v0, t0, dtf, power_index;
    int dist_flag;
    int origin;                 // stored in each injected particle (variant index)
};

// Threshold injections (particle_module.f90:785-1470) and their cell
// counters (mhd_data_parallel.f90:2211-2499).  quantity: SP_INJECT_*.
struct ThreshArgs {
    int quantity;
    double threshold;
    double box[6];                // part_box: xmin, ymin, zmin, xmax, ymax, zmax
    double idxh, idyh, idzh;      // 0.5 / dx ... of the gradient stencil
    const float* n8;              // frame-1 node fields (8 per node)
    long long max_trials;         // per particle, then the particle counts as failed
    int commit;                   // 0: only count failures; 1: store the particles
};

struct SplitArgs {
    double split_ratio, pmin_split, p0, pmax;
};

struct LocalArgs {
    int enabled, npbins, nmu;
    long long nr[3];
    double dd[3];
    double pmin_log, dp_log, dmu;
};

struct DistArgs {
    int npp, nmu, local_dist;
    double pmin, pmax, pmin_log, dp_log, dmu;
    double xmin, ymin, zmin;
    LocalArgs loc[4];
};

// Distributions of escaped particles (diagnostics.f90:913-1231), filled while
// the removal pass counts them.  glob is (nmu, npp, 2 ndim); face[k][a] is
// set k's array for the faces normal to axis a: (nmu, npbins, n_b, n_c, 2)
// with (b, c) the other two axes in x, y, z order; nullptr where absent.
struct EscArgs {
    int on, ndim;
    unsigned long long* glob;
    unsigned long long* face[4][3];
};

struct SortArgs {
    int ndim, nx, ny, nz;       // interior cells per axis
    int shift;                  // cells per sort block = 2^shift along each axis
    int passes;                 // 4-bit radix passes over the Morton key
    int done_bit;               // >= 0: particles not paused in the mover (rdt == 0)
                                // get this key bit and sort after the paused ones
    double xmin, ymin, zmin, idx, idy, idz;
};

long long num_tiles(long long n);
hipError_t launch_copy(const void* a, void* b, long long n16, hipStream_t s);
hipError_t launch_spatial_sort(const Ptl& S, const Ptl& D, long long n, const SortArgs& sa,
                               uint32_t* keyA, uint32_t* idxA, uint32_t* keyB, uint32_t* idxB,
                               unsigned* hist, unsigned long long* off, unsigned long long* totals,
                               bool copy_back, hipStream_t s);
hipError_t launch_inject(const DevParams& prm, const Ptl& P, long long start, long long nptl,
                         long long nmax, int tag0, const InjectArgs& ia, hipStream_t s);
hipError_t launch_threshold_count(const DevParams& prm, const ThreshArgs& ta,
                                  unsigned long long* count, hipStream_t s);
hipError_t launch_threshold_inject(const DevParams& prm, const Ptl& P, long long start,
                                   long long nptl, long long nmax, int tag0, const InjectArgs& ia,
                                   const ThreshArgs& ta, unsigned long long* fails, hipStream_t s);
hipError_t launch_tile_count(const Ptl& P, long long n, int mode, const SplitArgs& sa,
                             unsigned* tile_cnt, unsigned* tile_esc, const DistArgs* da,
                             const EscArgs* ea, hipStream_t s);
hipError_t launch_tile_scan(const unsigned* tile_cnt, const unsigned* tile_esc, long long ntiles,
                            unsigned long long* tile_off, unsigned long long* totals, hipStream_t s);
hipError_t launch_compact_scatter(const Ptl& S, const Ptl& D, long long n,
                                  const unsigned long long* tile_off, hipStream_t s);
hipError_t launch_split_apply(const Ptl& P, long long n, long long nmax, const SplitArgs& sa,
                              const unsigned long long* tile_off, hipStream_t s);
hipError_t launch_dist(const Ptl& P, long long n, const DistArgs& da, unsigned long long* gglob,
                       unsigned long long* const gloc[4], const long long loc_size[4], int grid,
                       hipStream_t s);
hipError_t launch_variant_dist(const Ptl& P, long long n, const DistArgs& da, int nvar,
                               unsigned long long* hist, unsigned long long* counts, int ncu,
                               hipStream_t s);
hipError_t launch_stats(const Ptl& P, long long n, double* part, int grid, hipStream_t s);

}  // namespace sp
