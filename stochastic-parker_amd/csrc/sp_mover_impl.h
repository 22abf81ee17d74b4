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

#include <type_traits>

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

// Per-slot state of the particle a lane is pushing.
struct Slot {
    long long idx;                // particle index (-1: none)
    State s;
    double w, dtarget, dt_old;
    double dl[4];                 // the last step's increments (for the undo)
    int cf, nsp, pending;
    uint32_t tagi, tags, rstep;
    unsigned ksteps;              // steps of this particle in this launch
    bool resumed;
};

// At most two resident waves per SIMD: the mover runs one or two 256-lane
// blocks per CU (fewer particles in flight keep their field records in L2),
// so the compiler may spend registers on keeping the gather's loads in flight
// rather than on a higher occupancy that is never used (3D: +13 %).
//
// PPL particles per lane: each lane carries PPL independent particles
// ("slots") through the loop above.  Their steps are computed in one
// straight-line block (every slot steps, a slot that has no step to take
// discards its result), so the compiler can interleave two particles' loads,
// Philox rounds and arithmetic and fill the latency one particle alone leaves
// idle.  Every slot draws its increments from its own particle's counter, so
// the particles per lane never change a result.  Two 2D particles per lane
// fit the two-wave register budget (233 VGPRs) once the constants are left to
// the compiler; two 3D particles need the one-wave budget.
template <int NDIM, int PPL>
constexpr bool kTwoWavePair = NDIM == 2 && PPL == 2;
// The one-frame 2D step of the FP32 movers with one particle per lane runs
// three waves per SIMD (129-168 VGPRs, the constants left to the compiler):
// with most gathers served by the register cache, a third wave hides more of
// the step's latency than pinning every constant saves (-9 % kernel time
// against two pinned waves).  Two-frame steps would spill at that budget.
template <class Pol, int NDIM, bool TINTERP, int PPL>
constexpr bool kThreeWave = Pol::kThreeWave2D && NDIM == 2 && !TINTERP && PPL == 1;

template <class Pol, int NDIM, bool TINTERP, int FX, int PPL>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(kThreeWave<Pol, NDIM, TINTERP, PPL> ? 3 : (kTwoWavePair<NDIM, PPL> ? 2 : 1),
                                   kThreeWave<Pol, NDIM, TINTERP, PPL> ? 3 : 2)))
mover_kernel(MoverParams m0) {
    static_assert(PPL >= 1 && PPL <= Pol::kMaxPPL, "particles per lane of this policy");
    MoverParams m = m0;
    // within the two-wave register budget the constants fit in VGPRs for 2D
    // and 3D alike (2D: +5 %, 3D: +1 %)
    if constexpr (!kTwoWavePair<NDIM, PPL> && !kThreeWave<Pol, NDIM, TINTERP, PPL>) pin_constants(m);
    // at three waves only the FP64 constants of every iteration fit (interval
    // times, the interpolation origin and scale, the slack bounds): -6 to -9 %
    // kernel time; the acceleration-gated NLGC step has no room for them
    constexpr bool kPinHot = kThreeWave<Pol, NDIM, TINTERP, PPL> &&
                             !((FX & FX_ACC) != 0 && (FX & FX_NLGC) != 0);
    if constexpr (kPinHot) {
        pin_vgpr(m.t0); pin_vgpr(m.dt_fine); pin_vgpr(m.dt_limit);
        pin_vgpr(m.xmin); pin_vgpr(m.ymin); pin_vgpr(m.idx); pin_vgpr(m.idy);
        pin_vgpr(m.slo[0]); pin_vgpr(m.slo[1]); pin_vgpr(m.shi[0]); pin_vgpr(m.shi[1]);
    }
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

    bool exhausted = false;
    // start on this die's part of the range, then help the other parts
    int part = (int)(xcc_id() % (unsigned)kParts), parts_done = 0;
    Slot sl[PPL];
    typename Pol::Var var[PPL];
    typename Pol::template Cache<NDIM, TINTERP, FX, PPL> gc;   // slot 0's last records
    gc.invalidate();
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        sl[k].idx = -1;
        sl[k].s = State{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        sl[k].w = sl[k].dtarget = sl[k].dt_old = 0.0;
        sl[k].dl[0] = sl[k].dl[1] = sl[k].dl[2] = sl[k].dl[3] = 0.0;
        sl[k].cf = sl[k].nsp = 0;
        sl[k].pending = PEND_NONE;
        sl[k].tagi = sl[k].tags = sl[k].rstep = 0u;
        sl[k].ksteps = 0u;
        sl[k].resumed = false;
        var[k] = Pol::var_of(m);
    }
    unsigned long long my_steps = 0, my_paused = 0;

    while (true) {
        // ---- refill idle slots (wave-aggregated work fetch, one slot index at a time) ----
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            Slot& q = sl[k];
            const bool need = q.idx < 0 && !exhausted;
            const unsigned long long need_mask = __ballot(need);
            if (!need_mask) continue;
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
                    q.idx = j;
                    if (k == 0) gc.invalidate();
                    q.s.x = px[j];
                    q.s.y = py[j];
                    q.s.z = pz[j];
                    q.s.p = pp[j];
                    q.s.t = pt[j];
                    q.s.dt = pdt[j];
                    if constexpr ((FX & FX_NLGC) != 0) q.s.mu = m.P.d(PD_MU)[j];
                    q.w = m.P.d(PD_W)[j];
                    q.cf = pflag[j];
                    q.nsp = ppushed[j];
                    if (m0.nvar > 0) {
                        // this particle's variant: its diffusion parameters
                        int v = m.P.i(PI_ORIGIN)[j];
                        v = v < 0 ? 0 : (v >= m0.nvar ? m0.nvar - 1 : v);
                        Pol::set_var(m, var[k], m0.var[v]);
                    }
                    q.tagi = (uint32_t)m.P.i(PI_TAGI)[j];
                    q.tags = (uint32_t)m.P.i(PI_TAGS)[j];
                    q.rstep = prstep[j];
                    q.dl[0] = q.dl[1] = q.dl[2] = q.dl[3] = 0.0;
                    q.ksteps = 0;
                    const double rd = prdt[j];
                    q.resumed = rd > 0.0;
                    if (q.resumed) {
                        // paused at the top of the inner loop by an earlier launch
                        q.dtarget = rd;
                        q.pending = PEND_NONE;
                    } else {
                        // target time from the particle's own time stamp (:1570-1578)
                        const double stp = ceil((q.s.t - m.t0) / m.dt_fine);
                        if (stp <= 0.0) q.dtarget = m.dt_fine; else q.dtarget = stp * m.dt_fine;
                        if (q.dtarget > m.dtf) q.dtarget = m.dtf;
                        q.pending = PEND_PRE;
                    }
                }
            }
        }
        bool busy = false;
#pragma unroll
        for (int k = 0; k < PPL; ++k) busy = busy || sl[k].idx >= 0;
        if (__ballot(busy || !exhausted) == 0ull) break;

        bool do_push[PPL], fixed[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            Slot& q = sl[k];
            do_push[k] = false;
            fixed[k] = false;
            if (q.idx < 0) continue;
            // ---- the common case: an inner-loop step of a particle inside the
            // slack bounds with a non-negative momentum, for which the action
            // choice, the check and the transitions below reduce to "push"
            // with nothing changed.  Tested with the same comparisons the
            // check makes, so NaN positions take this path as they pass there.
            {
                bool inside = !(q.s.p < 0.0) && !(q.s.x < m.slo[0]) && !(q.s.x > m.shi[0]) &&
                              !(q.s.y < m.slo[1]) && !(q.s.y > m.shi[1]);
                if (NDIM == 3) inside = inside && !(q.s.z < m.slo[2]) && !(q.s.z > m.shi[2]);
                const bool running = q.pending == PEND_NONE && (q.s.t - m.t0) < q.dtarget &&
                                     q.cf == SP_COUNT_FLAG_INBOX &&
                                     !(m.step_budget != 0u && q.ksteps >= m.step_budget);
                if (running && inside) {
                    do_push[k] = true;
                    continue;
                }
            }
            // ---- choose this slot's action ----
            int act = ACT_NONE;
            bool need_bc = false;
            if (q.pending == PEND_PRE) {
                need_bc = true;  // safety check before the loop (:1581-1592)
                act = ACT_PRE;
            } else if (q.pending == PEND_REDO) {
                q.s.dt = q.dt_old;  // the redo step is done; restore dt, check, advance (:1804-1826)
                need_bc = true;
                act = ACT_ADVANCE;
            } else if ((q.s.t - m.t0) < q.dtarget && q.cf == SP_COUNT_FLAG_INBOX) {
                if (m.step_budget != 0u && q.ksteps >= m.step_budget) {
                    act = ACT_PAUSE;
                } else {
                    need_bc = true;  // top of the inner loop (:1600-1612)
                    act = ACT_PUSH;
                }
            } else if ((q.s.t - m.t0) > q.dtarget && q.cf == SP_COUNT_FLAG_INBOX) {
                // land exactly on the target time (:1707-1720)
                q.s.x = q.s.x - q.dl[0];
                q.s.y = q.s.y - q.dl[1];
                q.s.z = q.s.z - q.dl[2];
                q.s.p = q.s.p - q.dl[3];
                q.s.t = q.s.t - q.s.dt;
                q.dt_old = q.s.dt;
                q.s.dt = m.t0 + q.dtarget - q.s.t;
                if (q.s.dt > 0) {
                    q.nsp = q.nsp - 1;
                    act = ACT_PUSH_FIXED;
                } else {
                    q.s.dt = q.dt_old;
                    need_bc = true;
                    act = ACT_ADVANCE;
                }
            } else {
                act = ACT_ADVANCE;
            }
            q.pending = PEND_NONE;

            // ---- the slack-bound check: momentum sign first, then the faces ----
            if (need_bc) {
                if (q.s.p < 0.0 && q.cf == SP_COUNT_FLAG_INBOX) {
                    q.cf = SP_COUNT_FLAG_OTHERS;
                    atomicAdd(&m.ctr->negp_fx, weight_fx(q.w));
                } else {
                    const double lk = boundary_condition<NDIM>(m.flags, q.s.x, q.s.y, q.s.z, q.cf,
                                                               q.w, m.slo[0], m.slo[1], m.slo[2],
                                                               m.shi[0], m.shi[1], m.shi[2]);
                    if (lk != 0.0) atomicAdd(&m.ctr->leak_fx, weight_fx(lk));
                }
            }

            // ---- transitions ----
            bool fin = false;
            if (act == ACT_PUSH) {
                if (q.cf == SP_COUNT_FLAG_INBOX) do_push[k] = true;
                else fin = true;  // the inner loop exits and the outer loop ends on count_flag
            } else if (act == ACT_PUSH_FIXED) {
                do_push[k] = true;
                fixed[k] = true;
            } else if (act == ACT_ADVANCE) {
                q.dtarget = q.dtarget + m.dt_fine;
                fin = !(q.dtarget < m.dt_limit) || q.cf != SP_COUNT_FLAG_INBOX;
            } else if (act == ACT_PRE) {
                fin = q.cf != SP_COUNT_FLAG_INBOX || !(q.dtarget < m.dt_limit);
            }
            const bool pause = act == ACT_PAUSE;
            if (fin || pause) {
                if (fin && q.cf == SP_COUNT_FLAG_INBOX) {
                    const double lk = boundary_condition<NDIM>(m.flags, q.s.x, q.s.y, q.s.z, q.cf,
                                                               q.w, m.elo[0], m.elo[1], m.elo[2],
                                                               m.ehi[0], m.ehi[1], m.ehi[2]);
                    if (lk != 0.0) atomicAdd(&m.ctr->leak_fx, weight_fx(lk));
                }
                const long long j = q.idx;
                if (pause) {
                    prdt[j] = q.dtarget;
                    ++my_paused;
                } else if (q.resumed) {
                    prdt[j] = 0.0;
                }
                px[j] = q.s.x;
                py[j] = q.s.y;
                pz[j] = q.s.z;
                pp[j] = q.s.p;
                pt[j] = q.s.t;
                pdt[j] = q.s.dt;
                pflag[j] = (int8_t)q.cf;
                ppushed[j] = q.nsp;
                prstep[j] = q.rstep;
                q.idx = -1;
            }
        }

        // ---- one particle-step for every slot that takes one ----
        bool any_push = false;
#pragma unroll
        for (int k = 0; k < PPL; ++k) any_push = any_push || do_push[k];
        if (any_push) {
            if constexpr (PPL == 1) {
                Pol::template step<NDIM, TINTERP, FX>(m, var[0], sl[0].s, fixed[0], sl[0].tagi,
                                                      sl[0].tags, sl[0].rstep, sl[0].dl, gc);
            } else {
                // every slot steps (one straight-line block); a slot without a
                // step to take keeps its state
                State ns[PPL];
                double nd[PPL][4];
#pragma unroll
                for (int k = 0; k < PPL; ++k) ns[k] = sl[k].s;
                Pol::template step_n<NDIM, TINTERP, FX, PPL>(m, var, ns, fixed, sl, nd);
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    if (do_push[k]) {
                        sl[k].s = ns[k];
                        sl[k].dl[0] = nd[k][0];
                        sl[k].dl[1] = nd[k][1];
                        sl[k].dl[2] = nd[k][2];
                        sl[k].dl[3] = nd[k][3];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                if (!do_push[k]) continue;
                Slot& q = sl[k];
                q.rstep += 1u;
                int v = q.nsp + 1;  // mod(nsteps_pushed + 1, nsteps_interval)
                if ((unsigned)v >= (unsigned)m.nsteps_interval) v %= m.nsteps_interval;
                q.nsp = v;
                ++my_steps;
                ++q.ksteps;
                if (fixed[k]) q.pending = PEND_REDO;
            }
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
    Pol::template step_given<NDIM, TINTERP, FX>(m, Pol::var_of(m), s, fixed[i] != 0, rands + 4 * i, d);
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

template <bool DPP, int N, int LEAN, class F>
auto dispatch_fx_base(bool ti, int fx, F&& f) {
    if constexpr (DPP) {
        if (fx & FX_DPP) {
            if (fx & FX_NLGC) return dispatch_acc<DPP, N, LEAN | FX_DPP | FX_NLGC>(ti, fx, f);
            return dispatch_acc<DPP, N, LEAN | FX_DPP>(ti, fx, f);
        }
    }
    if (fx & FX_NLGC) return dispatch_acc<DPP, N, LEAN | FX_NLGC>(ti, fx, f);
    return dispatch_acc<DPP, N, LEAN>(ti, fx, f);
}

template <bool DPP, bool LEAN, int N, class F>
auto dispatch_fx(bool ti, int fx, F&& f) {
    if constexpr (LEAN) {
        if (fx & FX_LEAN) return dispatch_fx_base<DPP, N, FX_LEAN>(ti, fx, f);
    }
    return dispatch_fx_base<DPP, N, 0>(ti, fx, f);
}

// Calls f(Variant<...>{}) for the runtime (ndim, time_interp, step features)
// combination; a policy with kOnly3D builds (and serves) 3D variants only, one
// without kDpp no momentum-diffusion variants, one without kLean ignores
// FX_LEAN.
template <class Pol, class F>
auto dispatch_variant(int ndim, bool ti, int fx, F&& f) {
    if constexpr (!Pol::kOnly3D) {
        if (ndim == 2) return dispatch_fx<Pol::kDpp, Pol::kLean, 2>(ti, fx, f);
    }
    return dispatch_fx<Pol::kDpp, Pol::kLean, 3>(ti, fx, f);
}

// Calls f(v, std::integral_constant<int, PPL>{}) for the runtime particles
// per lane: 1, or 2 where the policy supports it and the step gathers one
// frame (two frames per particle and two particles spill registers).
template <class Pol, class V, class F>
auto dispatch_ppl(int ppl, V v, F&& f) {
    if constexpr (Pol::kMaxPPL >= 2 && !V::ti) {
        if (ppl == 2) return f(v, std::integral_constant<int, 2>{});
    }
    return f(v, std::integral_constant<int, 1>{});
}

template <class Pol>
int mover_occupancy(int ndim, bool ti, int fx, int ppl) {
    return dispatch_variant<Pol>(ndim, ti, fx, [&](auto v) {
        return dispatch_ppl<Pol>(ppl, v, [](auto vv, auto pc) {
            using V = decltype(vv);
            int nb = 0;
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, mover_kernel<Pol, V::ndim, V::ti, V::fx, decltype(pc)::value>, 256, 0);
            return e == hipSuccess ? nb : 1;
        });
    });
}

template <class Pol>
hipError_t launch_mover(const MoverParams& m, bool ti, int fx, int ndim, int ppl, int grid,
                        hipStream_t st) {
    dispatch_variant<Pol>(ndim, ti, fx, [&](auto v) {
        return dispatch_ppl<Pol>(ppl, v, [&](auto vv, auto pc) {
            using V = decltype(vv);
            hipLaunchKernelGGL((mover_kernel<Pol, V::ndim, V::ti, V::fx, decltype(pc)::value>),
                               dim3(grid), dim3(256), 0, st, m);
            return 0;
        });
    });
    return hipGetLastError();
}

template <class Pol>
hipError_t launch_step_once(const MoverParams& m, bool ti, int fx, int ndim, double* state,
                            const double* rands, const int8_t* fixed, double* deltas,
                            hipStream_t st) {
    if (m.n <= 0) return hipSuccess;
    const int grid = (int)((m.n + 255) / 256);
    dispatch_variant<Pol>(ndim, ti, fx, [&](auto v) {
        using V = decltype(v);
        hipLaunchKernelGGL((step_once_kernel<Pol, V::ndim, V::ti, V::fx>), dim3(grid), dim3(256),
                           0, st, m, state, rands, fixed, deltas);
        return 0;
    });
    return hipGetLastError();
}

}  // namespace sp
