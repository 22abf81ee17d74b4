// sp_internal.h -- types shared by the host side of libsp_hip.so and its
// kernels.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sp_hip.h"

namespace sp {

// Node-record widths (floats).  A 2D record holds the 15 components the 2D
// Parker step consumes plus one pad (64 B = four dwordx4 loads); a 3D record
// holds 21 components plus three pads (96 B = six dwordx4 loads).
constexpr int kRec2D = 16;
constexpr int kRec3D = 24;

// 2D record slots
enum Rec2 { R2_VX, R2_VY, R2_BX, R2_BY, R2_BZ, R2_DVXDX, R2_DVYDY, R2_DBXDX, R2_DBXDY,
            R2_DBYDX, R2_DBYDY, R2_DBZDX, R2_DBZDY, R2_DBDX, R2_DBDY, R2_PAD };
// 3D record slots
enum Rec3 { R3_VX, R3_VY, R3_VZ, R3_BX, R3_BY, R3_BZ, R3_DVXDX, R3_DVYDY, R3_DVZDZ,
            R3_DBXDX, R3_DBXDY, R3_DBXDZ, R3_DBYDX, R3_DBYDY, R3_DBYDZ,
            R3_DBZDX, R3_DBZDY, R3_DBZDZ, R3_DBDX, R3_DBDY, R3_DBDZ };

// farray component (0-based) stored in each record slot
__host__ __device__ inline int rec2_src(int s) {
    const int m[15] = {0, 1, 4, 5, 6, 8, 12, 20, 21, 23, 24, 26, 27, 29, 30};
    return s < 15 ? m[s] : -1;
}
__host__ __device__ inline int rec3_src(int s) {
    const int m[21] = {0, 1, 2, 4, 5, 6, 8, 12, 16, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31};
    return s < 21 ? m[s] : -1;
}

// Kernel-side copy of the configuration (passed by value).
struct DevParams {
    int ndim;
    int nxg, nyg, nzg;
    double dx, dy, dz;
    double idx, idy, idz;          // 1/dx ... (FP32 path only)
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double lx, ly, lz;
    int pbc[3];
    int momentum_dependency, mag_dependency, pcharge;
    double p0, pmin, pmax;
    double gamma_turb, pindex, kpara0, kret;
    double drift1, drift2;
    int check_drift_2d, acc_region_flag, time_interp, precision;
    double acc_region[6];
    uint32_t key0, key1, stream;
};

// Device particle arrays (structure of arrays).
struct Ptl {
    double *x, *y, *z, *p, *v, *mu, *w, *t, *dt;
    int8_t *split_times, *count_flag;
    int32_t *origin, *nsteps_tracked, *nsteps_pushed, *tag_injected, *tag_splitted;
    uint32_t *rng_step;
};

// Device-side counters of one context.
struct Counters {
    unsigned long long work;       // next particle index handed out by the mover
    unsigned long long steps;      // particle-steps of the current mover call
    unsigned long long leak_fx;    // leak in units of 2^-32 weight
    unsigned long long negp_fx;    // leak_negp in units of 2^-32 weight
    unsigned long long nkeep;      // compaction / split results
    unsigned long long nesc;
    unsigned long long pad[2];
};

constexpr double kFxScale = 4294967296.0;  // 2^32 weight units per unit weight

struct MoverArgs {
    DevParams prm;
    Ptl P;
    const float* rec1;
    const float* rec2;
    Counters* ctr;
    long long n;
    double t0, dtf, dt_fine, dt_limit, dt_min, dt_max;
    int nsteps_interval;
};

// Launchers (defined in the .hip translation units)
hipError_t launch_gradients(const DevParams& prm, const float* nodes8, float* rec, float* farray32,
                            hipStream_t s);
int mover_blocks_per_cu_f64(int ndim, bool ti);
int mover_blocks_per_cu_f32(int ndim, bool ti);
hipError_t launch_mover_f64(const MoverArgs& a, int grid, int block, hipStream_t s);
hipError_t launch_mover_f32(const MoverArgs& a, int grid, int block, hipStream_t s);
hipError_t launch_step_once_f64(const DevParams& prm, const float* rec1, const float* rec2,
                                long long n, double t0, double dtf, double dt_min, double dt_max,
                                double* state, const double* rands, const int8_t* fixed,
                                double* deltas, hipStream_t s);
hipError_t launch_step_once_f32(const DevParams& prm, const float* rec1, const float* rec2,
                                long long n, double t0, double dtf, double dt_min, double dt_max,
                                double* state, const double* rands, const int8_t* fixed,
                                double* deltas, hipStream_t s);

}  // namespace sp
