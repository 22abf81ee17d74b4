// sp_internal.h -- types shared by the host side of libsp_hip.so and its
// kernels.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sp_hip.h"

namespace sp {

// Node-record widths (floats).  A 2D record holds the 15 components the 2D
// Parker step consumes plus one pad (64 B = four dwordx4 loads); a 3D record
// holds 21 components plus three pads (96 B = six dwordx4 loads).
constexpr int kRec2D = 16;
constexpr int kRec3D = 24;
// Records for steps with momentum diffusion (particle_module.f90:2918-2979):
// the 2D / 3D slots above, then rho and the off-diagonal velocity gradients
// of the flow-shear tensor.  2D: 18 components + 2 pads (80 B); 3D: 28
// components (112 B).
constexpr int kRec2DX = 20;
constexpr int kRec3DX = 28;

// 2D record slots
enum Rec2 { R2_VX, R2_VY, R2_BX, R2_BY, R2_BZ, R2_DVXDX, R2_DVYDY, R2_DBXDX, R2_DBXDY,
            R2_DBYDX, R2_DBYDY, R2_DBZDX, R2_DBZDY, R2_DBDX, R2_DBDY, R2_PAD,
            R2_RHO = R2_PAD, R2_DVXDY, R2_DVYDX };
// 3D record slots
enum Rec3 { R3_VX, R3_VY, R3_VZ, R3_BX, R3_BY, R3_BZ, R3_DVXDX, R3_DVYDY, R3_DVZDZ,
            R3_DBXDX, R3_DBXDY, R3_DBXDZ, R3_DBYDX, R3_DBYDY, R3_DBYDZ,
            R3_DBZDX, R3_DBZDY, R3_DBZDZ, R3_DBDX, R3_DBDY, R3_DBDZ,
            R3_RHO, R3_DVXDY, R3_DVXDZ, R3_DVYDX, R3_DVYDZ, R3_DVZDX, R3_DVZDY };

// Records of the FP32 movers (steps without momentum diffusion).  The two
// or three velocity gradients are stored as their sum divV, the only form
// the step reads.  The texture path's cost follows the number of load
// instructions rather than their width, so the layouts minimise dwordx4
// loads per cell.  FP16 values are stored times a power-of-two scale of the
// frame (one for the gradients, one for the velocity).
//
// 2D values per node, in the order of pair2_src: NF = 5 field values (V, B)
// and NG = 9 gradients (divV, the B-component gradients, grad |B|).
// 2D FP16 gradients: x-pair records.  The record of node (i, j) holds that
// node and node i + 1 of the same row, so the two x-corners of a cell arrive
// in one run of loads (the 2D field set stays cache resident at twice the
// size):
//          [node i: 5 floats][node i + 1: 5 floats]
//          [node i: 9 halves + 1 pad][node i + 1: same]         (80 B, 5 loads)
// 2D FP32: cell coefficient records (kCell2D below).
// 3D: bricked node records (the 3D gather is sensitive to its cache
// footprint, so nodes are not duplicated), in the order of node3_src:
//   FP32:  V, B, divV and 12 gradients, 19 floats + 1 pad      (80 B, 5 loads)
//   FP16:  B as 3 floats, then V, divV and the 12 gradients as
//          16 halves, + 1 pad word                              (48 B, 3 loads)
struct Pair2Shape {
    static constexpr int NF = 5;
    static constexpr int NG = 9;
    static constexpr int N = NF + NG;
    static constexpr int NGW = (NG + 1) / 2;   // words of halves per node
    static constexpr int W16 = 20;             // record width in words
};
// 2D FP32: cell records of bilinear coefficients.  The record of cell (i, j)
// holds, per slot s of pair2_src, c0 = v00, c1 = v10 - v00, c2 = v01 - v00,
// c3 = (v11 - v10) - (v01 - v00) (v_ab the value at node (i + a, j + b)), as
// [c0 x 14][c1 x 14][c2 x 14][c3 x 14] (224 B, 14 loads), so the step
// interpolates with three FMAs per value and no corner weights.
constexpr int kCell2D = 56;
constexpr int kRec3F = 20;   // 3D FP32 record (floats)
constexpr int kRec3H = 12;   // 3D FP16 record (words)
// farray component of each slot; -1 marks divV (dvx/dx + dvy/dy [+ dvz/dz])
__host__ __device__ inline int pair2_src(int s) {
    const int m[14] = {0, 1, 4, 5, 6, -1, 20, 21, 23, 24, 26, 27, 29, 30};
    return m[s];
}
// 3D slots: VX VY VZ BX BY BZ DIVV, then dBx/dx .. d|B|/dz (farray 20 .. 31)
__host__ __device__ inline int node3_src(int s) {
    const int m[19] = {0, 1, 2, 4, 5, 6, -1, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31};
    return m[s];
}

// farray component (0-based) stored in each record slot
// (slots past 14 / 20 belong to the momentum-diffusion records)
__host__ __device__ inline int rec2_src(int s) {
    const int m[18] = {0, 1, 4, 5, 6, 8, 12, 20, 21, 23, 24, 26, 27, 29, 30, 3, 9, 11};
    return s < 18 ? m[s] : -1;
}
__host__ __device__ inline int rec3_src(int s) {
    const int m[28] = {0, 1, 2, 4, 5, 6, 8, 12, 16, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31,
                       3, 9, 10, 11, 13, 14, 15};
    return s < 28 ? m[s] : -1;
}

// Step features chosen at compile time (template argument of the mover):
// the acceleration-region gate, the momentum-diffusion terms and the NLGC
// perpendicular diffusion coefficient.  FX_LEAN marks a configuration
// without kappa's magnetic-field dependence and without the 2D out-of-plane
// drift; the FP32 movers then build the step without the terms those flags
// zero (the FP64 mover ignores it).
enum StepFeatures { FX_ACC = 1, FX_DPP = 2, FX_NLGC = 4, FX_LEAN = 8 };

// 3D node records are stored in bricks of 8 x 8 x 8 nodes (bricks x-fastest,
// nodes x-fastest inside a brick), 32-56 KB each: the 8 corners of a cell share
// one brick unless the cell straddles a brick face, so a gather touches one or
// a few neighbouring pages instead of four rows up to a plane apart.
__host__ __device__ inline long long brick_record(int ix, int iy, int iz, int bnx, int bny) {
    const long long b = ((long long)(iz >> 3) * bny + (iy >> 3)) * bnx + (ix >> 3);
    return (b << 9) | (long long)(((iz & 7) << 6) | ((iy & 7) << 3) | (ix & 7));
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
    int ext_rec;                   // records carry the momentum-diffusion slots
    int bnx, bny, bnz;             // 3D record bricks per axis
    double acc_region[6];
    uint32_t key0, key1, stream;
};

// Device particle arrays (structure of arrays).  `rdt` is internal to the
// mover: the interval target time of a particle paused at the step budget
// (0 = not paused); it is not part of the exchanged particle record.
struct Ptl {
    double *x, *y, *z, *p, *v, *mu, *w, *t, *dt, *rdt;
    int8_t *split_times, *count_flag;
    int32_t *origin, *nsteps_tracked, *nsteps_pushed, *tag_injected, *tag_splitted;
    uint32_t *rng_step;
};

// The mover splits its particle range into one contiguous part per XCD (the
// chip's eight accelerator dies, each with its own L2); a block takes work
// from the part of the die it runs on first, so each L2 serves a compact
// window of the sorted particles.  Part counters sit 128 B apart.
constexpr int kParts = 8;
constexpr int kPartStride = 16;

// Device-side counters of one context.
struct Counters {
    unsigned long long part[kParts * kPartStride];  // next index inside each part
    unsigned long long steps;      // particle-steps of the current mover launch
    unsigned long long paused;     // particles that stopped at the step budget
    unsigned long long leak_fx;    // leak in units of 2^-32 weight
    unsigned long long negp_fx;    // leak_negp in units of 2^-32 weight
    unsigned long long nkeep;      // compaction / split results
    unsigned long long nesc;
    unsigned long long pad[2];
};

constexpr double kFxScale = 4294967296.0;  // 2^32 weight units per unit weight
// Largest split_times whose weight 2^-split_times the fixed point holds.
constexpr int kMaxSplitTimes = 32;

// Particle arrays of one buffer set live in one allocation with a common
// capacity `cap` (a multiple of 256): the 10 FP64 arrays first, then the two
// int8 arrays, then the six 32-bit arrays, in the order of Ptl (106 bytes per
// particle).  The mover addresses them from (base, cap) alone, which keeps
// its kernel arguments small enough to stay in scalar registers.
enum PtlD { PD_X, PD_Y, PD_Z, PD_P, PD_V, PD_MU, PD_W, PD_T, PD_DT, PD_RDT };
enum PtlB { PB_SPLIT, PB_FLAG };
enum PtlI { PI_ORIGIN, PI_TRACKED, PI_PUSHED, PI_TAGI, PI_TAGS, PI_RSTEP };
constexpr int kPtlBytes = 106;

struct PtlBase {
    char* base;
    long long cap;
    __host__ __device__ double* d(int k) const { return reinterpret_cast<double*>(base + 8 * cap * k); }
    __host__ __device__ int8_t* b(int k) const { return reinterpret_cast<int8_t*>(base + 80 * cap + cap * k); }
    __host__ __device__ int32_t* i(int k) const {
        return reinterpret_cast<int32_t*>(base + 82 * cap + 4 * cap * k);
    }
};

// Diffusion parameters of one variant, in the forms the two step paths use.
struct KappaVar {
    double kpara0, kret, pindex, gamma_m2;
    double p0_2mpi;               // p0**(2 - pindex), flow-shear momentum diffusion
    double perp_mag, perp_mom;    // NLGC kperp exponents (gamma-2)/3 of B, (5-gamma)/3 of p/p0
    float f_kpara0, f_kret, f_pindex, f_gm2;
    float f_perp_mag, f_perp_mom;
};

// Bit flags of MoverParams::flags
enum MoverFlags {
    MF_OPEN_X = 1, MF_OPEN_Y = 2, MF_OPEN_Z = 4,
    MF_MOMENTUM_DEP = 8, MF_MAG_DEP = 16, MF_DRIFT_2D = 32,
    MF_DPP_WAVE = 64, MF_DPP_SHEAR = 128, MF_WEAK_SCATTERING = 256
};

// Kernel arguments of the mover and of the single-step kernel: only the
// values the step reads, precomputed on the host in the operation order of
// the reference expressions they stand for.
struct MoverParams {
    PtlBase P;
    long long n;
    long long part_len;           // particles per XCD part (the last part may be shorter)
    unsigned step_budget;         // steps per particle per launch before it pauses (0 = none)
    const KappaVar* var;          // per-variant diffusion parameters (nvar > 0)
    int nvar;                     // 0: every particle uses the values below
    const float* rec1;
    const float* rec2;
    Counters* ctr;
    int nxg, nyg, nzg, nsteps_interval;
    int bnx, bny;                 // 3D record bricks along x and y
    int flags;
    uint32_t key0, key1, stream;
    double t0, dtf, dt_fine, dt_limit;
    double xmin, ymin, zmin;
    double dx, dy, dz;            // FP64 path: p = (x - xmin) / dx
    double idx, idy, idz;         // FP32 path: p = (x - xmin) * idx
    double slo[3], shi[3];        // slack (half-cell) bounds used while stepping
    double elo[3], ehi[3];        // exact bounds of particle_mover's final pass
    double p0, p_floor;           // p_floor = 0.25 * p0
    // FP64 physics
    double kpara0, kret, pindex, gamma_m2;  // gamma_m2 = gamma_turb - 2
    double vdp_pre;               // (double)(1.0f / (3 * charge)), an FP32 quotient
    double d1p0, d2p02;           // drift1 * p0, drift2 * (p0 * p0)
    double half_dx, half_dy, half_dz;
    double dt_min, dt_max;
    // FP32 physics
    float f_kpara0, f_kret, f_pindex, f_gm2, f_vdp_pre, f_d1p0, f_d2p02;
    float f_half_dx, f_half_dy, f_half_dz, f_dt_min, f_dt_max;
    float f_gs1, f_gs2;           // FP16 records: 1 / gradient scale of frames 1 and 2
    float f_vs1, f_vs2;           // FP16 3D records: 1 / velocity scale of frames 1 and 2
    double inv_p0;
    // acceleration region (template-enabled)
    double acc_region[6];
    double lx, ly, lz;
    // momentum diffusion (template-enabled)
    double tau0, p0_2mpi;         // p0_2mpi = p0**(2 - pindex)
    float f_tau0, f_p02;          // FP32 path: tau0, p0 * p0
    // NLGC kperp (template-enabled): kpara0 * kperp_kpara * B^perp_mag * (p/p0)^perp_mom * mu^2
    double kperp_kpara, perp_mag, perp_mom;
    float f_kperp_kpara, f_perp_mag, f_perp_mom;
};

// Records the calling thread's error message (sp_last_error) and returns code.
int set_error(int code, const std::string& msg);

// Launchers (defined in the .hip translation units)
hipError_t launch_gradients(const DevParams& prm, const float* nodes8, float* rec, float* farray32,
                            hipStream_t s);
// Records of the FP32 movers.  FP16 records: max |gradient| and max |V| of
// the frame (as float bits, into maxbits[0] and maxbits[1]) set its
// power-of-two scales, then the packing applies them.
hipError_t launch_value_max(const DevParams& prm, const float* nodes8, unsigned* maxbits,
                            hipStream_t s);
hipError_t launch_pack_fast(const DevParams& prm, const float* nodes8, bool half, float gscale,
                            float vscale, float* rec, float* farray32, hipStream_t s);
// ppl: particles per lane of the FP32 movers (1 or 2); the FP64 mover takes one
int mover_blocks_per_cu_f32h(int ndim, bool ti, int fx, int ppl);
hipError_t launch_mover_f32h(const MoverParams& m, bool ti, int fx, int ndim, int ppl, int grid,
                             hipStream_t s);
hipError_t launch_step_once_f32h(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                 const double* rands, const int8_t* fixed, double* deltas,
                                 hipStream_t s);
int mover_blocks_per_cu_f64(int ndim, bool ti, int fx);
int mover_blocks_per_cu_f32(int ndim, bool ti, int fx, int ppl);
hipError_t launch_mover_f64(const MoverParams& m, bool ti, int fx, int ndim, int grid,
                            hipStream_t s);
hipError_t launch_mover_f32(const MoverParams& m, bool ti, int fx, int ndim, int ppl, int grid,
                            hipStream_t s);
hipError_t launch_step_once_f64(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                const double* rands, const int8_t* fixed, double* deltas,
                                hipStream_t s);
hipError_t launch_step_once_f32(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                const double* rands, const int8_t* fixed, double* deltas,
                                hipStream_t s);

}  // namespace sp
