// sp_mover_f32.hip -- mover and single-step kernels with FP32 gather and
// coefficients, FP64 position / momentum / time accumulators.
#include "sp_mover_impl.h"
#include "sp_phys_f32.h"

namespace sp {

// FP32 gather and coefficients; one or two particles per lane, each with its
// own copy of the per-particle diffusion parameters (KF).  Steps without
// momentum diffusion gather from compact records (2D: FP32 cell coefficient
// records, or FP16-gradient x-pair records when HALF; 3D: bricked node
// records, FP16 values when HALF), those with it from the node records that
// carry its slots.  Configurations marked FX_LEAN get their own kernels.
template <bool HALF>
struct PolicyF32Base {
    static constexpr int kMaxPPL = 2;
    static constexpr bool kThreeWave2D = true;
    static constexpr bool kLean = true;
    using Var = f32::KF;
    __device__ __forceinline__ static Var var_of(const MoverParams& m) { return f32::kf_of(m); }
    __device__ __forceinline__ static void set_var(MoverParams&, Var& v, const KappaVar& kv) {
        v = f32::kf_of(kv);
    }
    template <int NDIM, bool TI, int FX, int PPL>
    using Cache = f32::GatherCache<NDIM, TI, HALF, (FX & FX_DPP) != 0, PPL>;
    template <int NDIM, bool TI, int FX, class GC>
    __device__ __forceinline__ static void step(const MoverParams& m, const Var& kf, State& s,
                                                bool fixed, uint32_t tagi, uint32_t tags,
                                                uint32_t rstep, double d[4], GC& gc) {
        float r[4];
        f32::step_rands(m, tagi, tags, rstep, r);
        f32::gather_kappa_push<NDIM, TI, FX, HALF>(m, kf, s, fixed, r, d, gc);
    }
    // PPL particles: the Philox blocks, the gathers and the steps of all
    // slots in one block, for the scheduler to interleave
    template <int NDIM, bool TI, int FX, int PPL>
    __device__ __forceinline__ static void step_n(const MoverParams& m, const Var* kf, State* s,
                                                  const bool* fixed, const Slot* sl,
                                                  double (*d)[4]) {
        float r[PPL][4];
#pragma unroll
        for (int k = 0; k < PPL; ++k) f32::step_rands(m, sl[k].tagi, sl[k].tags, sl[k].rstep, r[k]);
        Cache<NDIM, TI, FX, PPL> none;   // no records kept across steps at PPL > 1
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            f32::gather_kappa_push<NDIM, TI, FX, HALF>(m, kf[k], s[k], fixed[k], r[k], d[k], none);
    }
    template <int NDIM, bool TI, int FX>
    __device__ __forceinline__ static void step_given(const MoverParams& m, const Var& kf,
                                                      State& s, bool fixed, const double* r,
                                                      double d[4]) {
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        Cache<NDIM, TI, FX, 2> none;   // one step: nothing to keep
        f32::gather_kappa_push<NDIM, TI, FX, HALF>(m, kf, s, fixed, rf, d, none);
    }
};

// FP32 arithmetic on FP32 records.
struct PolicyF32 : PolicyF32Base<false> {
    static constexpr bool kOnly3D = false;
    static constexpr bool kDpp = true;
};

// FP32 arithmetic on x-pair records with FP16 gradients.
struct PolicyF32H : PolicyF32Base<true> {
    static constexpr bool kOnly3D = false;
    static constexpr bool kDpp = false;   // no momentum-diffusion slots in these records
};

int mover_blocks_per_cu_f32h(int ndim, bool ti, int fx, int ppl) {
    return mover_occupancy<PolicyF32H>(ndim, ti, fx, ppl);
}

hipError_t launch_mover_f32h(const MoverParams& m, bool ti, int fx, int ndim, int ppl, int grid,
                             hipStream_t s) {
    return launch_mover<PolicyF32H>(m, ti, fx, ndim, ppl, grid, s);
}

hipError_t launch_step_once_f32h(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                 const double* rands, const int8_t* fixed, double* deltas,
                                 hipStream_t s) {
    return launch_step_once<PolicyF32H>(m, ti, fx, ndim, state, rands, fixed, deltas, s);
}

int mover_blocks_per_cu_f32(int ndim, bool ti, int fx, int ppl) {
    return mover_occupancy<PolicyF32>(ndim, ti, fx, ppl);
}

hipError_t launch_mover_f32(const MoverParams& m, bool ti, int fx, int ndim, int ppl, int grid,
                            hipStream_t s) {
    return launch_mover<PolicyF32>(m, ti, fx, ndim, ppl, grid, s);
}

hipError_t launch_step_once_f32(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                                const double* rands, const int8_t* fixed, double* deltas,
                                hipStream_t s) {
    return launch_step_once<PolicyF32>(m, ti, fx, ndim, state, rands, fixed, deltas, s);
}

}  // namespace sp
