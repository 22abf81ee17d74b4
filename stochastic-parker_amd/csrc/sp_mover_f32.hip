// sp_mover_f32.hip -- mover and single-step kernels with FP32 gather and
// coefficients, FP64 position / momentum / time accumulators.
#include "sp_mover_impl.h"
#include "sp_phys_f32.h"

namespace sp {

struct PolicyF32 {
    static constexpr bool kOnly3D = false;
    static constexpr bool kDpp = true;
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step(const MoverParams& m, State& s, bool fixed,
                                                uint32_t tagi, uint32_t tags, uint32_t rstep,
                                                double d[4]) {
        float r[4];
        f32::step_rands(m, tagi, tags, rstep, r);
        f32::gather_kappa_push<NDIM, TI, FX>(m, s, fixed, r, d);
    }
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step_given(const MoverParams& m, State& s, bool fixed,
                                                      const double* r, double d[4]) {
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        f32::gather_kappa_push<NDIM, TI, FX>(m, s, fixed, rf, d);
    }
};

// FP32 arithmetic on 3D records with FP16 gradients.
struct PolicyF32H {
    static constexpr bool kOnly3D = true;
    static constexpr bool kDpp = false;   // no momentum-diffusion slots in these records
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step(const MoverParams& m, State& s, bool fixed,
                                                uint32_t tagi, uint32_t tags, uint32_t rstep,
                                                double d[4]) {
        float r[4];
        f32::step_rands(m, tagi, tags, rstep, r);
        f32::gather_kappa_push<NDIM, TI, FX, true>(m, s, fixed, r, d);
    }
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step_given(const MoverParams& m, State& s, bool fixed,
                                                      const double* r, double d[4]) {
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        f32::gather_kappa_push<NDIM, TI, FX, true>(m, s, fixed, rf, d);
    }
};

int mover_blocks_per_cu_f32h(bool ti, int fx) {
    return mover_occupancy<PolicyF32H>(3, ti, fx);
}

hipError_t launch_mover_f32h(const MoverParams& m, bool ti, int fx, int grid, hipStream_t s) {
    return launch_mover<PolicyF32H>(m, ti, fx, 3, grid, s);
}

hipError_t launch_step_once_f32h(const MoverParams& m, bool ti, int fx, double* state,
                                 const double* rands, const int8_t* fixed, double* deltas,
                                 hipStream_t s) {
    return launch_step_once<PolicyF32H>(m, ti, fx, 3, state, rands, fixed, deltas, s);
}

int mover_blocks_per_cu_f32(int ndim, bool ti, int fx) {
    return mover_occupancy<PolicyF32>(ndim, ti, fx);
}

hipError_t launch_mover_f32(const MoverParams& m, bool ti, int fx, int ndim, int grid,
                            hipStream_t s) {
    return launch_mover<PolicyF32>(m, ti, fx, ndim, grid, s);
}

hipError_t launch_step_once_f32(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                const double* rands, const int8_t* fixed, double* deltas,
                                hipStream_t s) {
    return launch_step_once<PolicyF32>(m, ti, fx, ndim, state, rands, fixed, deltas, s);
}

}  // namespace sp
