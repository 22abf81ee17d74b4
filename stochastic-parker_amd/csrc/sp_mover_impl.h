// sp_mover_impl.h -- the persistent particle-mover kernel.
//
// One lane owns one particle at a time and walks it through the reference's
// per-particle interval loop (particle_mover_one_cycle,
// particle_module.f90:1561-1830): the fine-step loop, the adaptive steps, the
// interval-end undo + fixed-dt redo, the slack-bound boundary checks and,
// when the particle is done, particle_mover's exact-bounds pass
// (:1959-1968).  The loop body is one particle-step for every lane that has
// one to take; lanes whose particle finished pick up the next particle index
// from a global counter (one atomic per wave), so step-count differences
// between particles never idle a wave.  Random increments come from Philox
// keyed on the particle's (tag_injected, tag_splitted, rng_step), so results
// do not depend on which lane or wave pushes a particle.
#pragma once

#include "sp_device.h"

namespace sp {

template <class Pol, int NDIM, bool TINTERP>
__global__ void __launch_bounds__(256) mover_kernel(MoverArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const DevParams& prm = a.prm;
    const Bounds slack{prm.xmin - prm.dx * 0.5, prm.ymin - prm.dy * 0.5, prm.zmin - prm.dz * 0.5,
                       prm.xmax + prm.dx * 0.5, prm.ymax + prm.dy * 0.5, prm.zmax + prm.dz * 0.5};
    const Bounds exact{prm.xmin, prm.ymin, prm.zmin, prm.xmax, prm.ymax, prm.zmax};

    long long idx = -1;
    bool exhausted = false;
    State s{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double w = 0.0, dtarget = 0.0;
    double dl[4] = {0.0, 0.0, 0.0, 0.0};
    int cf = 0, nsp = 0;
    uint32_t tagi = 0, tags = 0, rstep = 0;
    unsigned long long my_steps = 0;

    auto bc_slack = [&]() {
        const double lk = boundary_condition(prm, s.x, s.y, s.z, cf, w, slack);
        if (lk != 0.0) atomicAdd(&a.ctr->leak_fx, weight_fx(lk));
    };
    auto check_p_or_bc = [&]() {
        if (s.p < 0.0) {
            cf = SP_COUNT_FLAG_OTHERS;
            atomicAdd(&a.ctr->negp_fx, weight_fx(w));
        } else {
            bc_slack();
        }
    };
    auto finalize = [&]() {
        if (cf == SP_COUNT_FLAG_INBOX) {
            const double lk = boundary_condition(prm, s.x, s.y, s.z, cf, w, exact);
            if (lk != 0.0) atomicAdd(&a.ctr->leak_fx, weight_fx(lk));
        }
        a.P.x[idx] = s.x;
        a.P.y[idx] = s.y;
        a.P.z[idx] = s.z;
        a.P.p[idx] = s.p;
        a.P.t[idx] = s.t;
        a.P.dt[idx] = s.dt;
        a.P.count_flag[idx] = (int8_t)cf;
        a.P.nsteps_pushed[idx] = nsp;
        a.P.rng_step[idx] = rstep;
        idx = -1;
    };
    auto advance = [&]() {
        dtarget = dtarget + a.dt_fine;
        if (!(dtarget < a.dt_limit) || cf != SP_COUNT_FLAG_INBOX) finalize();
    };

    while (true) {
        // ---- refill idle lanes (wave-aggregated work fetch) ----
        const bool need = idx < 0 && !exhausted;
        const unsigned long long need_mask = __ballot(need);
        if (need_mask) {
            const int leader = __ffsll((unsigned long long)need_mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&a.ctr->work, (unsigned long long)__popcll(need_mask));
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long j =
                    base + (unsigned long long)__popcll(need_mask & ((1ull << lane) - 1ull));
                if (j >= (unsigned long long)a.n) {
                    exhausted = true;
                } else {
                    idx = (long long)j;
                    s.x = a.P.x[idx];
                    s.y = a.P.y[idx];
                    s.z = a.P.z[idx];
                    s.p = a.P.p[idx];
                    s.t = a.P.t[idx];
                    s.dt = a.P.dt[idx];
                    w = a.P.w[idx];
                    cf = a.P.count_flag[idx];
                    nsp = a.P.nsteps_pushed[idx];
                    tagi = (uint32_t)a.P.tag_injected[idx];
                    tags = (uint32_t)a.P.tag_splitted[idx];
                    rstep = a.P.rng_step[idx];
                    dl[0] = dl[1] = dl[2] = dl[3] = 0.0;
                    // target time from the particle's own time stamp (:1570-1578)
                    const double stp = ceil((s.t - a.t0) / a.dt_fine);
                    if (stp <= 0.0) dtarget = a.dt_fine; else dtarget = stp * a.dt_fine;
                    if (dtarget > a.dtf) dtarget = a.dtf;
                    // safety check before the loop (:1581-1592)
                    if (s.p < 0.0 && cf == SP_COUNT_FLAG_INBOX) {
                        cf = SP_COUNT_FLAG_OTHERS;
                        atomicAdd(&a.ctr->negp_fx, weight_fx(w));
                    } else {
                        bc_slack();
                    }
                    if (cf != SP_COUNT_FLAG_INBOX || !(dtarget < a.dt_limit)) finalize();
                }
            }
        }
        if (__ballot(idx >= 0 || !exhausted) == 0ull) break;

        // ---- decide this lane's action for the iteration ----
        bool do_push = false, fixed = false;
        double dt_old = 0.0;
        if (idx >= 0) {
            if ((s.t - a.t0) < dtarget && cf == SP_COUNT_FLAG_INBOX) {
                // top of the inner while loop (:1600-1612)
                check_p_or_bc();
                if (cf == SP_COUNT_FLAG_INBOX) do_push = true;
                else finalize();  // inner loop exits; the outer loop ends on count_flag
            } else {
                // inner loop finished: land exactly on the target time (:1707-1826)
                if ((s.t - a.t0) > dtarget && cf == SP_COUNT_FLAG_INBOX) {
                    s.x = s.x - dl[0];
                    s.y = s.y - dl[1];
                    s.z = s.z - dl[2];
                    s.p = s.p - dl[3];
                    s.t = s.t - s.dt;
                    dt_old = s.dt;
                    s.dt = a.t0 + dtarget - s.t;
                    if (s.dt > 0) {
                        nsp = nsp - 1;
                        do_push = true;
                        fixed = true;
                    } else {
                        s.dt = dt_old;
                        check_p_or_bc();
                        advance();
                    }
                } else {
                    advance();
                }
            }
        }
        // ---- one particle-step for every lane that takes one ----
        if (do_push) {
            Pol::template step<NDIM, TINTERP>(a, s, fixed, tagi, tags, rstep, dl);
            rstep += 1u;
            nsp = (nsp + 1) % a.nsteps_interval;
            ++my_steps;
            if (fixed) {
                s.dt = dt_old;
                check_p_or_bc();
                advance();
            }
        }
    }

    // ---- per-wave step total: one atomic per wave ----
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) my_steps += __shfl_xor(my_steps, off);
    if (lane == 0 && my_steps) atomicAdd(&a.ctr->steps, my_steps);
}

// Single-step kernel for parity checks: caller-supplied states and increments.
template <class Pol, int NDIM, bool TINTERP>
__global__ void __launch_bounds__(256) step_once_kernel(DevParams prm, const float* rec1,
                                                        const float* rec2, long long n, double t0,
                                                        double dtf, double dt_min, double dt_max,
                                                        double* state, const double* rands,
                                                        const int8_t* fixed, double* deltas) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    State s{state[6 * i + 0], state[6 * i + 1], state[6 * i + 2],
            state[6 * i + 3], state[6 * i + 4], state[6 * i + 5]};
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    Pol::template step_given<NDIM, TINTERP>(prm, rec1, rec2, t0, dtf, dt_min, dt_max, s,
                                            fixed[i] != 0, rands + 4 * i, d);
    state[6 * i + 0] = s.x;
    state[6 * i + 1] = s.y;
    state[6 * i + 2] = s.z;
    state[6 * i + 3] = s.p;
    state[6 * i + 4] = s.t;
    state[6 * i + 5] = s.dt;
    for (int c = 0; c < 4; ++c) deltas[4 * i + c] = d[c];
}

template <class Pol>
hipError_t launch_mover(const MoverArgs& a, int grid, int block, hipStream_t st) {
    const bool ti = a.prm.time_interp != 0;
    if (a.prm.ndim == 2) {
        if (ti) hipLaunchKernelGGL((mover_kernel<Pol, 2, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((mover_kernel<Pol, 2, false>), dim3(grid), dim3(block), 0, st, a);
    } else {
        if (ti) hipLaunchKernelGGL((mover_kernel<Pol, 3, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((mover_kernel<Pol, 3, false>), dim3(grid), dim3(block), 0, st, a);
    }
    return hipGetLastError();
}

template <class Pol>
int mover_occupancy(int ndim, bool ti) {
    int nb = 0;
    hipError_t e;
    if (ndim == 2)
        e = ti ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mover_kernel<Pol, 2, true>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mover_kernel<Pol, 2, false>, 256, 0);
    else
        e = ti ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mover_kernel<Pol, 3, true>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mover_kernel<Pol, 3, false>, 256, 0);
    return e == hipSuccess ? nb : 1;
}

template <class Pol>
hipError_t launch_step_once(const DevParams& prm, const float* rec1, const float* rec2,
                            long long n, double t0, double dtf, double dt_min, double dt_max,
                            double* state, const double* rands, const int8_t* fixed,
                            double* deltas, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    const bool ti = prm.time_interp != 0;
#define SP_STEP_ONCE(ND, TI)                                                                     \
    hipLaunchKernelGGL((step_once_kernel<Pol, ND, TI>), dim3(grid), dim3(block), 0, st, prm, rec1, \
                       rec2, n, t0, dtf, dt_min, dt_max, state, rands, fixed, deltas)
    if (prm.ndim == 2) {
        if (ti) SP_STEP_ONCE(2, true); else SP_STEP_ONCE(2, false);
    } else {
        if (ti) SP_STEP_ONCE(3, true); else SP_STEP_ONCE(3, false);
    }
#undef SP_STEP_ONCE
    return hipGetLastError();
}

}  // namespace sp
