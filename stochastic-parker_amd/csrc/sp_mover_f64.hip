// sp_mover_f64.hip -- mover and single-step kernels in the reference's FP64
// arithmetic.  Built with -ffp-contract=off (see the package Makefile).
#include "sp_mover_impl.h"
#include "sp_phys_f64.h"

namespace sp {

struct PolicyF64 {
    template <int NDIM, bool TI>
    __device__ __forceinline__ static void step(const MoverArgs& a, State& s, bool fixed,
                                                uint32_t tagi, uint32_t tags, uint32_t rstep,
                                                double d[4]) {
        double r[4];
        f64::step_rands(a.prm, tagi, tags, rstep, r);
        f64::gather_kappa_push<NDIM, TI>(a.prm, a.rec1, a.rec2, a.t0, a.dtf, a.dt_min, a.dt_max,
                                         s, fixed, r, d);
    }
    template <int NDIM, bool TI>
    __device__ __forceinline__ static void step_given(const DevParams& prm, const float* rec1,
                                                      const float* rec2, double t0, double dtf,
                                                      double dt_min, double dt_max, State& s,
                                                      bool fixed, const double* r, double d[4]) {
        f64::gather_kappa_push<NDIM, TI>(prm, rec1, rec2, t0, dtf, dt_min, dt_max, s, fixed, r, d);
    }
};

int mover_blocks_per_cu_f64(int ndim, bool ti) { return mover_occupancy<PolicyF64>(ndim, ti); }

hipError_t launch_mover_f64(const MoverArgs& a, int grid, int block, hipStream_t s) {
    return launch_mover<PolicyF64>(a, grid, block, s);
}

hipError_t launch_step_once_f64(const DevParams& prm, const float* rec1, const float* rec2,
                                long long n, double t0, double dtf, double dt_min, double dt_max,
                                double* state, const double* rands, const int8_t* fixed,
                                double* deltas, hipStream_t s) {
    return launch_step_once<PolicyF64>(prm, rec1, rec2, n, t0, dtf, dt_min, dt_max, state, rands,
                                       fixed, deltas, s);
}

}  // namespace sp
