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

long long num_tiles(long long n);
hipError_t launch_inject(const DevParams& prm, const Ptl& P, long long start, long long nptl,
                         long long nmax, int tag0, const InjectArgs& ia, hipStream_t s);
hipError_t launch_tile_count(const Ptl& P, long long n, int mode, const SplitArgs& sa,
                             unsigned* tile_cnt, unsigned* tile_esc, hipStream_t s);
hipError_t launch_tile_scan(const unsigned* tile_cnt, const unsigned* tile_esc, long long ntiles,
                            unsigned long long* tile_off, unsigned long long* totals, hipStream_t s);
hipError_t launch_compact_scatter(const Ptl& S, const Ptl& D, long long n,
                                  const unsigned long long* tile_off, hipStream_t s);
hipError_t launch_split_apply(const Ptl& P, long long n, long long nmax, const SplitArgs& sa,
                              const unsigned long long* tile_off, hipStream_t s);
hipError_t launch_dist(const Ptl& P, long long n, const DistArgs& da, unsigned long long* gglob,
                       unsigned long long* const gloc[4], int grid, hipStream_t s);
hipError_t launch_stats(const Ptl& P, long long n, double* part, int grid, hipStream_t s);

}  // namespace sp
