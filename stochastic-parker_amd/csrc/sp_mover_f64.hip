// sp_mover_f64.hip -- mover and single-step kernels in the reference's FP64
// arithmetic.  Built with -ffp-contract=off (see the package Makefile).
#include "sp_mover_impl.h"
#include "sp_phys_f64.h"

namespace sp {

// One particle per lane.  A batched run's per-particle diffusion parameters
// overwrite the lane's own copy of the kernel arguments when it takes a
// particle, so Var carries nothing.
struct PolicyF64 {
    static constexpr bool kOnly3D = false;
    static constexpr bool kDpp = true;
    static constexpr int kMaxPPL = 1;
    static constexpr bool kThreeWave2D = false;
    static constexpr bool kLean = false;
    struct Var {};
    __device__ __forceinline__ static Var var_of(const MoverParams&) { return Var{}; }
    __device__ __forceinline__ static void set_var(MoverParams& m, Var&, const KappaVar& kv) {
        m.kpara0 = kv.kpara0; m.kret = kv.kret;
        m.pindex = kv.pindex; m.gamma_m2 = kv.gamma_m2;
        m.p0_2mpi = kv.p0_2mpi;
        m.perp_mag = kv.perp_mag; m.perp_mom = kv.perp_mom;
    }
    // no records are kept across steps
    template <int NDIM, bool TI, int FX, int PPL>
    struct Cache {
        __device__ __forceinline__ void invalidate() {}
    };
    template <int NDIM, bool TI, int FX, class GC>
    __device__ __forceinline__ static void step(const MoverParams& m, const Var&, State& s,
                                                bool fixed, uint32_t tagi, uint32_t tags,
                                                uint32_t rstep, double d[4], GC&) {
        double r[4];
        f64::step_rands(m, tagi, tags, rstep, r);
        f64::gather_kappa_push<NDIM, TI, FX>(m, s, fixed, r, d);
    }
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step_given(const MoverParams& m, const Var&, State& s,
                                                      bool fixed, const double* r, double d[4]) {
        f64::gather_kappa_push<NDIM, TI, FX>(m, s, fixed, r, d);
    }
};

int mover_blocks_per_cu_f64(int ndim, bool ti, int fx) {
    return mover_occupancy<PolicyF64>(ndim, ti, fx, 1);
}

hipError_t launch_mover_f64(const MoverParams& m, bool ti, int fx, int ndim, int grid,
                            hipStream_t s) {
    return launch_mover<PolicyF64>(m, ti, fx, ndim, 1, grid, s);
}

hipError_t launch_step_once_f64(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                const double* rands, const int8_t* fixed, double* deltas,
                                hipStream_t s) {
    return launch_step_once<PolicyF64>(m, ti, fx, ndim, state, rands, fixed, deltas, s);
}

}  // namespace sp
