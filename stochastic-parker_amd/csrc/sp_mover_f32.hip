// sp_mover_f32.hip -- mover and single-step kernels with FP32 gather and
// coefficients, FP64 position / momentum / time accumulators.
#include "sp_mover_impl.h"
#include "sp_phys_f32.h"

namespace sp {

struct PolicyF32 {
    template <int NDIM, bool TI>
    __device__ __forceinline__ static void step(const MoverArgs& a, State& s, bool fixed,
                                                uint32_t tagi, uint32_t tags, uint32_t rstep,
                                                double d[4]) {
        float r[4];
        f32::step_rands(a.prm, tagi, tags, rstep, r);
        f32::gather_kappa_push<NDIM, TI>(a.prm, a.rec1, a.rec2, a.t0, a.dtf, (float)a.dt_min,
                                         (float)a.dt_max, s, fixed, r, d);
    }
    template <int NDIM, bool TI>
    __device__ __forceinline__ static void step_given(const DevParams& prm, const float* rec1,
                                                      const float* rec2, double t0, double dtf,
                                                      double dt_min, double dt_max, State& s,
                                                      bool fixed, const double* r, double d[4]) {
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        f32::gather_kappa_push<NDIM, TI>(prm, rec1, rec2, t0, dtf, (float)dt_min, (float)dt_max, s,
                                         fixed, rf, d);
    }
};

int mover_blocks_per_cu_f32(int ndim, bool ti) { return mover_occupancy<PolicyF32>(ndim, ti); }

hipError_t launch_mover_f32(const MoverArgs& a, int grid, int block, hipStream_t s) {
    return launch_mover<PolicyF32>(a, grid, block, s);
}

hipError_t launch_step_once_f32(const DevParams& prm, const float* rec1, const float* rec2,
                                long long n, double t0, double dtf, double dt_min, double dt_max,
                                double* state, const double* rands, const int8_t* fixed,
                                double* deltas, hipStream_t s) {
    return launch_step_once<PolicyF32>(prm, rec1, rec2, n, t0, dtf, dt_min, dt_max, state, rands,
                                       fixed, deltas, s);
}

}  // namespace sp
