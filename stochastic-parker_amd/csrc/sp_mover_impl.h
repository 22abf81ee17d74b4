// sp_mover_impl.h -- the persistent particle-mover kernel.
//
// One lane owns one particle at a time and walks it through the reference's
// per-particle interval loop (particle_mover_one_cycle,
// particle_module.f90:1561-1830): the fine-step loop, the adaptive steps, the
// interval-end undo + fixed-dt redo, the slack-bound boundary checks and,
// when the particle is done, particle_mover's exact-bounds pass
// (:1959-1968).  Each loop iteration is one particle-step for every lane
// that has one to take; a lane whose particle finished takes the next
// particle index from a global counter (one atomic per wave and refill), so
// differing step counts never idle a wave.  Random increments come from
// Philox keyed on the particle's (tag_injected, tag_splitted, rng_step), so
// results do not depend on which lane or wave pushes a particle.
//
// The reference loop is flattened into a per-lane action chosen at the top of
// each iteration, so the boundary check, the step and the store each appear
// once in the kernel body:
//   pending PRE   after a load: safety check, finish if out of the box
//   inner loop    t - t0 < target: boundary check, then an adaptive step
//   overshoot     t - t0 > target: undo the last step, fixed-dt redo step;
//                 the redo's boundary check and the advance of the target run
//                 at the top of the next iteration (pending REDO)
//   otherwise     advance the target (next fine step, or done)
//
// With a step budget, a particle that has taken that many steps in this
// launch stops at the top of its inner loop (before the boundary check of its
// next step): its state and interval target are stored and the launch counts
// it as paused.  The next launch resumes it at the same point; the increments
// depend only on the particle's counter, so pausing never changes a result.
#pragma once

#include "sp_device.h"

namespace sp {

enum : int { PEND_NONE = 0, PEND_PRE = 1, PEND_REDO = 2 };
enum : int { ACT_NONE = 0, ACT_PRE = 1, ACT_PUSH = 2, ACT_PUSH_FIXED = 3, ACT_ADVANCE = 4,
              ACT_PAUSE = 5 };

// Keeps a uniform floating-point value in vector registers: the hot loop
// reads dozens of FP64/FP32 constants, more than the scalar register file
// holds, and constants spilled from scalar registers cost one v_readlane per
// 32-bit half at every use.  An operand already in a VGPR costs nothing.
template <class T>
__device__ __forceinline__ void pin_vgpr(T& v) {
    asm volatile("" : "+v"(v));
}

__device__ __forceinline__ void pin_constants(MoverParams& m) {
    pin_vgpr(m.t0); pin_vgpr(m.dtf); pin_vgpr(m.dt_fine); pin_vgpr(m.dt_limit);
    pin_vgpr(m.xmin); pin_vgpr(m.ymin); pin_vgpr(m.zmin);
    pin_vgpr(m.dx); pin_vgpr(m.dy); pin_vgpr(m.dz);
    pin_vgpr(m.idx); pin_vgpr(m.idy); pin_vgpr(m.idz);
    for (int k = 0; k < 3; ++k) {
        pin_vgpr(m.slo[k]); pin_vgpr(m.shi[k]); pin_vgpr(m.elo[k]); pin_vgpr(m.ehi[k]);
    }
    pin_vgpr(m.p0); pin_vgpr(m.p_floor);
    pin_vgpr(m.kpara0); pin_vgpr(m.kret); pin_vgpr(m.pindex); pin_vgpr(m.gamma_m2);
    pin_vgpr(m.vdp_pre); pin_vgpr(m.d1p0); pin_vgpr(m.d2p02);
    pin_vgpr(m.half_dx); pin_vgpr(m.half_dy); pin_vgpr(m.half_dz);
    pin_vgpr(m.dt_min); pin_vgpr(m.dt_max);
    pin_vgpr(m.f_kpara0); pin_vgpr(m.f_kret); pin_vgpr(m.f_pindex); pin_vgpr(m.f_gm2);
    pin_vgpr(m.f_vdp_pre); pin_vgpr(m.f_d1p0); pin_vgpr(m.f_d2p02);
    pin_vgpr(m.f_half_dx); pin_vgpr(m.f_half_dy); pin_vgpr(m.f_half_dz);
    pin_vgpr(m.f_dt_min); pin_vgpr(m.f_dt_max);
    pin_vgpr(m.inv_p0);
    for (int k = 0; k < 6; ++k) pin_vgpr(m.acc_region[k]);
    pin_vgpr(m.lx); pin_vgpr(m.ly); pin_vgpr(m.lz);
}

// At most two resident waves per SIMD: the mover runs one or two 256-lane
// blocks per CU (fewer particles in flight keep their field records in L2),
// so the compiler may spend registers on keeping the gather's loads in flight
// rather than on a higher occupancy that is never used (3D: +13 %).
template <class Pol, int NDIM, bool TINTERP, int FX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2)))
mover_kernel(MoverParams m0) {
    MoverParams m = m0;
    // within the two-wave register budget the constants fit in VGPRs for 2D
    // and 3D alike (2D: +5 %, 3D: +1 %)
    pin_constants(m);
    const int lane = threadIdx.x & (kWave - 1);
    double* const px = m.P.d(PD_X);
    double* const py = m.P.d(PD_Y);
    double* const pz = m.P.d(PD_Z);
    double* const pp = m.P.d(PD_P);
    double* const pt = m.P.d(PD_T);
    double* const pdt = m.P.d(PD_DT);
    int8_t* const pflag = m.P.b(PB_FLAG);
    int32_t* const ppushed = m.P.i(PI_PUSHED);
    uint32_t* const prstep = reinterpret_cast<uint32_t*>(m.P.i(PI_RSTEP));
    double* const prdt = m.P.d(PD_RDT);

    long long idx = -1;
    bool exhausted = false;
    // start on this die's part of the range, then help the other parts
    int part = (int)(xcc_id() % (unsigned)kParts), parts_done = 0;
    State s{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double w = 0.0, dtarget = 0.0, dt_old = 0.0;
    double dl[4] = {0.0, 0.0, 0.0, 0.0};
    int cf = 0, nsp = 0, pending = PEND_NONE;
    uint32_t tagi = 0, tags = 0, rstep = 0;
    unsigned ksteps = 0;          // steps of the current particle in this launch
    bool resumed = false;
    unsigned long long my_steps = 0, my_paused = 0;

    while (true) {
        // ---- refill idle lanes (wave-aggregated work fetch) ----
        const bool need = idx < 0 && !exhausted;
        const unsigned long long need_mask = __ballot(need);
        if (need_mask) {
            const int leader = __ffsll((unsigned long long)need_mask) - 1;
            const unsigned long long cnt = (unsigned long long)__popcll(need_mask);
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&m.ctr->part[part * kPartStride], cnt);
            base = __shfl(base, leader);
            const long long pb = (long long)part * m.part_len;
            const long long pe = pb + m.part_len < m.n ? pb + m.part_len : m.n;
            if (pb + (long long)(base + cnt) >= pe) {
                // this part is used up: move on; after all parts, stop fetching
                part = (part + 1) % kParts;
                if (++parts_done == kParts) exhausted = true;
            }
            if (need) {
                const long long j =
                    pb + (long long)(base + (unsigned long long)__popcll(need_mask & ((1ull << lane) - 1ull)));
                if (j < pe) {
                    idx = j;
                    s.x = px[idx];
                    s.y = py[idx];
                    s.z = pz[idx];
                    s.p = pp[idx];
                    s.t = pt[idx];
                    s.dt = pdt[idx];
                    if constexpr ((FX & FX_NLGC) != 0) s.mu = m.P.d(PD_MU)[idx];
                    w = m.P.d(PD_W)[idx];
                    cf = pflag[idx];
                    nsp = ppushed[idx];
                    if (m0.nvar > 0) {
                        // this particle's variant: its diffusion parameters
                        int v = m.P.i(PI_ORIGIN)[idx];
                        v = v < 0 ? 0 : (v >= m0.nvar ? m0.nvar - 1 : v);
                        const KappaVar kv = m0.var[v];
                        m.kpara0 = kv.kpara0; m.kret = kv.kret;
                        m.pindex = kv.pindex; m.gamma_m2 = kv.gamma_m2;
                        m.p0_2mpi = kv.p0_2mpi;
                        m.perp_mag = kv.perp_mag; m.perp_mom = kv.perp_mom;
                        m.f_perp_mag = kv.f_perp_mag; m.f_perp_mom = kv.f_perp_mom;
                        m.f_kpara0 = kv.f_kpara0; m.f_kret = kv.f_kret;
                        m.f_pindex = kv.f_pindex; m.f_gm2 = kv.f_gm2;
                    }
                    tagi = (uint32_t)m.P.i(PI_TAGI)[idx];
                    tags = (uint32_t)m.P.i(PI_TAGS)[idx];
                    rstep = prstep[idx];
                    dl[0] = dl[1] = dl[2] = dl[3] = 0.0;
                    ksteps = 0;
                    const double rd = prdt[idx];
                    resumed = rd > 0.0;
                    if (resumed) {
                        // paused at the top of the inner loop by an earlier launch
                        dtarget = rd;
                        pending = PEND_NONE;
                    } else {
                        // target time from the particle's own time stamp (:1570-1578)
                        const double stp = ceil((s.t - m.t0) / m.dt_fine);
                        if (stp <= 0.0) dtarget = m.dt_fine; else dtarget = stp * m.dt_fine;
                        if (dtarget > m.dtf) dtarget = m.dtf;
                        pending = PEND_PRE;
                    }
                }
            }
        }
        if (__ballot(idx >= 0 || !exhausted) == 0ull) break;

        // ---- choose this lane's action ----
        int act = ACT_NONE;
        bool need_bc = false;
        if (idx >= 0) {
            if (pending == PEND_PRE) {
                need_bc = true;  // safety check before the loop (:1581-1592)
                act = ACT_PRE;
            } else if (pending == PEND_REDO) {
                s.dt = dt_old;   // the redo step is done; restore dt, check, advance (:1804-1826)
                need_bc = true;
                act = ACT_ADVANCE;
            } else if ((s.t - m.t0) < dtarget && cf == SP_COUNT_FLAG_INBOX) {
                if (m.step_budget != 0u && ksteps >= m.step_budget) {
                    act = ACT_PAUSE;
                } else {
                    need_bc = true;  // top of the inner loop (:1600-1612)
                    act = ACT_PUSH;
                }
            } else if ((s.t - m.t0) > dtarget && cf == SP_COUNT_FLAG_INBOX) {
                // land exactly on the target time (:1707-1720)
                s.x = s.x - dl[0];
                s.y = s.y - dl[1];
                s.z = s.z - dl[2];
                s.p = s.p - dl[3];
                s.t = s.t - s.dt;
                dt_old = s.dt;
                s.dt = m.t0 + dtarget - s.t;
                if (s.dt > 0) {
                    nsp = nsp - 1;
                    act = ACT_PUSH_FIXED;
                } else {
                    s.dt = dt_old;
                    need_bc = true;
                    act = ACT_ADVANCE;
                }
            } else {
                act = ACT_ADVANCE;
            }
            pending = PEND_NONE;
        }

        // ---- the slack-bound check: momentum sign first, then the faces ----
        if (need_bc) {
            if (s.p < 0.0 && cf == SP_COUNT_FLAG_INBOX) {
                cf = SP_COUNT_FLAG_OTHERS;
                atomicAdd(&m.ctr->negp_fx, weight_fx(w));
            } else {
                const double lk = boundary_condition<NDIM>(m.flags, s.x, s.y, s.z, cf, w, m.slo[0],
                                                           m.slo[1], m.slo[2], m.shi[0], m.shi[1],
                                                           m.shi[2]);
                if (lk != 0.0) atomicAdd(&m.ctr->leak_fx, weight_fx(lk));
            }
        }

        // ---- transitions ----
        bool fin = false, do_push = false;
        if (act == ACT_PUSH) {
            if (cf == SP_COUNT_FLAG_INBOX) do_push = true;
            else fin = true;  // the inner loop exits and the outer loop ends on count_flag
        } else if (act == ACT_PUSH_FIXED) {
            do_push = true;
        } else if (act == ACT_ADVANCE) {
            dtarget = dtarget + m.dt_fine;
            fin = !(dtarget < m.dt_limit) || cf != SP_COUNT_FLAG_INBOX;
        } else if (act == ACT_PRE) {
            fin = cf != SP_COUNT_FLAG_INBOX || !(dtarget < m.dt_limit);
        }
        const bool pause = act == ACT_PAUSE;
        if (fin || pause) {
            if (fin && cf == SP_COUNT_FLAG_INBOX) {
                const double lk = boundary_condition<NDIM>(m.flags, s.x, s.y, s.z, cf, w, m.elo[0],
                                                           m.elo[1], m.elo[2], m.ehi[0], m.ehi[1],
                                                           m.ehi[2]);
                if (lk != 0.0) atomicAdd(&m.ctr->leak_fx, weight_fx(lk));
            }
            if (pause) {
                prdt[idx] = dtarget;
                ++my_paused;
            } else if (resumed) {
                prdt[idx] = 0.0;
            }
            px[idx] = s.x;
            py[idx] = s.y;
            pz[idx] = s.z;
            pp[idx] = s.p;
            pt[idx] = s.t;
            pdt[idx] = s.dt;
            pflag[idx] = (int8_t)cf;
            ppushed[idx] = nsp;
            prstep[idx] = rstep;
            idx = -1;
        }

        // ---- one particle-step for every lane that takes one ----
        if (do_push) {
            const bool fixed = act == ACT_PUSH_FIXED;
            Pol::template step<NDIM, TINTERP, FX>(m, s, fixed, tagi, tags, rstep, dl);
            rstep += 1u;
            int v = nsp + 1;  // mod(nsteps_pushed + 1, nsteps_interval)
            if ((unsigned)v >= (unsigned)m.nsteps_interval) v %= m.nsteps_interval;
            nsp = v;
            ++my_steps;
            ++ksteps;
            if (fixed) pending = PEND_REDO;
        }
    }

    // ---- per-wave step and pause totals: one atomic each per wave ----
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        my_steps += __shfl_xor(my_steps, off);
        my_paused += __shfl_xor(my_paused, off);
    }
    if (lane == 0 && my_steps) atomicAdd(&m.ctr->steps, my_steps);
    if (lane == 0 && my_paused) atomicAdd(&m.ctr->paused, my_paused);
}

// Single-step kernel for parity checks: caller-supplied states and increments.
template <class Pol, int NDIM, bool TINTERP, int FX>
__global__ void __launch_bounds__(256) step_once_kernel(MoverParams m, double* state,
                                                        const double* rands, const int8_t* fixed,
                                                        double* deltas) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    State s{state[6 * i + 0], state[6 * i + 1], state[6 * i + 2],
            state[6 * i + 3], state[6 * i + 4], state[6 * i + 5], 0.0};
    if constexpr ((FX & FX_NLGC) != 0) s.mu = m.P.d(PD_MU)[i];
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    Pol::template step_given<NDIM, TINTERP, FX>(m, s, fixed[i] != 0, rands + 4 * i, d);
    state[6 * i + 0] = s.x;
    state[6 * i + 1] = s.y;
    state[6 * i + 2] = s.z;
    state[6 * i + 3] = s.p;
    state[6 * i + 4] = s.t;
    state[6 * i + 5] = s.dt;
    for (int c = 0; c < 4; ++c) deltas[4 * i + c] = d[c];
}

template <int N, bool T, int X>
struct Variant {
    static constexpr int ndim = N;
    static constexpr bool ti = T;
    static constexpr int fx = X;
};

template <bool DPP, int N, int BASE, class F>
auto dispatch_acc(bool ti, int fx, F&& f) {
    const bool with_acc = (fx & FX_ACC) != 0;
    if (ti) return with_acc ? f(Variant<N, true, BASE | FX_ACC>{}) : f(Variant<N, true, BASE>{});
    return with_acc ? f(Variant<N, false, BASE | FX_ACC>{}) : f(Variant<N, false, BASE>{});
}

template <bool DPP, int N, class F>
auto dispatch_fx(bool ti, int fx, F&& f) {
    if constexpr (DPP) {
        if (fx & FX_DPP) {
            if (fx & FX_NLGC) return dispatch_acc<DPP, N, FX_DPP | FX_NLGC>(ti, fx, f);
            return dispatch_acc<DPP, N, FX_DPP>(ti, fx, f);
        }
    }
    if (fx & FX_NLGC) return dispatch_acc<DPP, N, FX_NLGC>(ti, fx, f);
    return dispatch_acc<DPP, N, 0>(ti, fx, f);
}

// Calls f(Variant<...>{}) for the runtime (ndim, time_interp, step features)
// combination; a policy with kOnly3D builds (and serves) 3D variants only, one
// without kDpp no momentum-diffusion variants.
template <bool ONLY3D, bool DPP, class F>
auto dispatch_variant(int ndim, bool ti, int fx, F&& f) {
    if constexpr (!ONLY3D) {
        if (ndim == 2) return dispatch_fx<DPP, 2>(ti, fx, f);
    }
    return dispatch_fx<DPP, 3>(ti, fx, f);
}

template <class Pol>
int mover_occupancy(int ndim, bool ti, int fx) {
    return dispatch_variant<Pol::kOnly3D, Pol::kDpp>(ndim, ti, fx, [](auto v) {
        using V = decltype(v);
        int nb = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, mover_kernel<Pol, V::ndim, V::ti, V::fx>, 256, 0);
        return e == hipSuccess ? nb : 1;
    });
}

template <class Pol>
hipError_t launch_mover(const MoverParams& m, bool ti, int fx, int ndim, int grid, hipStream_t st) {
    dispatch_variant<Pol::kOnly3D, Pol::kDpp>(ndim, ti, fx, [&](auto v) {
        using V = decltype(v);
        hipLaunchKernelGGL((mover_kernel<Pol, V::ndim, V::ti, V::fx>), dim3(grid), dim3(256), 0,
                           st, m);
        return 0;
    });
    return hipGetLastError();
}

template <class Pol>
hipError_t launch_step_once(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                            const double* rands, const int8_t* fixed, double* deltas,
                            hipStream_t st) {
    if (m.n <= 0) return hipSuccess;
    const int grid = (int)((m.n + 255) / 256);
    dispatch_variant<Pol::kOnly3D, Pol::kDpp>(ndim, ti, fx, [&](auto v) {
        using V = decltype(v);
        hipLaunchKernelGGL((step_once_kernel<Pol, V::ndim, V::ti, V::fx>), dim3(grid), dim3(256),
                           0, st, m, state, rands, fixed, deltas);
        return 0;
    });
    return hipGetLastError();
}

}  // namespace sp
