/*
 * sp_hip.h -- C ABI of the MI355X (gfx950) Parker pseudo-particle pusher.
 *
 * The reference keeps this path behind Fortran module procedures operating on
 * module-global state (particle_module, mhd_data_parallel, diagnostics).  Each
 * entry point below replaces one of those procedures; the reference interface
 * it stands in for is cited as path:line in the xiaocanli/stochastic-parker
 * repository.  A Fortran caller binds these with iso_c_binding (interfaces in
 * stochastic-parker_amd/fortran/sp_hip_binding.f90).
 *
 * Conventions
 *  - Every function returns SP_OK (0) or a negative sp_status; the message of
 *    the last failure on the calling host thread is sp_last_error().
 *  - Host buffers are borrowed for the duration of the call only.
 *  - Fields and particles live in device memory owned by the context.
 *  - One context per host thread; calls on one context are serialised on the
 *    context's own HIP stream and are synchronous on return.
 *  - Integer/byte results (counts, histograms, tags, flags) are exact; the
 *    FP64 precision mode evaluates the reference's FP64 arithmetic in the
 *    reference's operation order.
 */
#ifndef SP_HIP_H
#define SP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_HIP_ABI_VERSION 4

typedef enum sp_status {
    SP_OK = 0,
    SP_ERR_ARG = -1,         /* invalid argument / configuration */
    SP_ERR_HIP = -2,         /* HIP runtime error */
    SP_ERR_STATE = -3,       /* call out of order (e.g. mover before fields) */
    SP_ERR_CAPACITY = -4,    /* request exceeds nptl_max or a host buffer */
    SP_ERR_UNSUPPORTED = -5, /* flag combination outside this path */
    SP_ERR_NOMEM = -6        /* device allocation failed */
} sp_status;

typedef enum sp_precision {
    SP_PREC_F64 = 0, /* reference arithmetic: FP32 fields, FP64 gather/coefficients/step */
    SP_PREC_F32 = 1, /* FP32 gather and coefficients; FP64 position, momentum and time */
    SP_PREC_F32H = 2 /* SP_PREC_F32 arithmetic on records whose gradients (divV, the
                        B-component gradients, grad |B|) are stored as FP16, and in 3D
                        also V, each scaled by a power of two per frame; B (and V in 2D)
                        stay FP32.  Not with momentum diffusion. */
} sp_precision;

/* count_flag values, particle_module.f90:55-62 */
#define SP_COUNT_FLAG_INBOX 1
#define SP_COUNT_FLAG_OTHERS 0
#define SP_COUNT_FLAG_ESCAPE_LX (-1)
#define SP_COUNT_FLAG_ESCAPE_HX (-2)
#define SP_COUNT_FLAG_ESCAPE_LY (-3)
#define SP_COUNT_FLAG_ESCAPE_HY (-4)
#define SP_COUNT_FLAG_ESCAPE_LZ (-5)
#define SP_COUNT_FLAG_ESCAPE_HZ (-6)

/* Everything the pusher reads from mhd_config.dat (mhd_config.f90:16-29),
 * conf.dat (particle_module.f90:2778-2878, simulation_setup.f90:104-150) and
 * the command line (stochastic-mhd.f90:572-946).  pindex is derived as
 * 3 - gamma_turb (particle_module.f90:2798). */
typedef struct sp_config {
    int32_t ndim;               /* ndim_field (-nd): 2 or 3 */
    int32_t nx, ny, nz;         /* interior grid nodes (mhd_config nx, ny, nz) */
    double dx, dy, dz;
    double xmin, ymin, zmin;
    double xmax, ymax, zmax;
    double lx, ly, lz;
    int32_t pbcx, pbcy, pbcz;   /* particle BCs: 0 periodic, 1 open */
    int32_t momentum_dependency;
    int32_t mag_dependency;
    int32_t charge;             /* -ch */
    double p0, pmin, pmax, b0;
    double gamma_turb;
    double kpara0, kret;
    double dt_min_rel, dt_max_rel;
    double drift1, drift2;      /* -dp1, -dp2 */
    int32_t check_drift_2d;     /* -cd */
    int32_t acc_region_flag;
    double acc_region[6];       /* acc_xmin, acc_xmax, acc_ymin, ... (0-1 normalised) */
    int32_t time_interp;        /* -ti: two-frame linear interpolation */
    int32_t precision;          /* sp_precision */
    uint64_t seed;              /* Philox key */
    uint32_t stream;            /* Philox stream word (run / variant id) */
    int32_t sort_every;         /* order particles spatially before every k-th mover call (0: never) */
    int32_t sort_steps;         /* with sort_every > 0: re-sort the particles still moving after
                                   every sort_steps steps of each (0: once per call); results
                                   do not depend on it */
    int32_t reserved0;          /* must be 0 */
    int64_t nptl_max;           /* -nm */
    /* momentum diffusion in the push (set_dpp_params, particle_module.f90:260-272;
     * stochastic-mhd.f90 flags -dw, -ds, -ws, -t0) */
    int32_t dpp_wave;           /* 1: wave scattering (calc_dpp_wave_scattering, :2918-2933) */
    int32_t dpp_shear;          /* 1: flow shear (calc_dpp_flow_shear, :2947-2979) */
    int32_t weak_scattering;    /* 1: the weak-scattering form of the flow-shear rate */
    int32_t nlgc;               /* 1: kperp from NLGC (calc_spatial_diffusion_coefficients_nlgc,
                                   particle_module.f90:2464-2771; -nl) */
    double tau0;                /* scattering time of particles with p0 (flow shear) */
    double kperp_kpara;         /* NLGC kperp / kpara at p0 (-kk) */
} sp_config;

/* Structure-of-arrays view of particle_type (particle_module.f90:38-50) plus
 * the counter-based RNG step index.  Any pointer may be NULL on download to
 * skip that member; all must be non-NULL on upload. */
typedef struct sp_particles {
    double *x, *y, *z, *p, *v, *mu, *weight, *t, *dt;
    int8_t *split_times, *count_flag;
    int32_t *origin, *nsteps_tracked, *nsteps_pushed, *tag_injected, *tag_splitted;
    uint32_t *rng_step;
} sp_particles;

typedef struct sp_mover_stats {
    int64_t steps;          /* particle-steps executed by this call (redo steps included) */
    int64_t nptl_current;   /* particles left after removal */
    int64_t nptl_escaped;   /* cumulative count of escaped particles */
    double leak;            /* cumulative weight leaked through open boundaries */
    double leak_negp;       /* cumulative weight removed for p < 0 */
    double kernel_ms;       /* device time of the push kernel (HIP events, context stream) */
    double total_ms;        /* device time of the whole call (push + boundary pass + removal) */
} sp_mover_stats;

/* Local distribution set k (diagnostics.f90:2049-2222): bins over
 * (mu, p, x, y, z) with npbins momentum bins on [pmin, pmax] and rx*ry*rz
 * grid cells per spatial bin. */
typedef struct sp_local_dist {
    int32_t enabled, npbins, nmu, rx, ry, rz;
    double pmin, pmax;
} sp_local_dist;

typedef struct sp_dist_config {
    int32_t npp_global, nmu_global;
    int32_t local_dist;           /* -ld */
    int32_t escaped_dist;         /* -ded: accumulate distributions of escaped particles */
    sp_local_dist local[4];
} sp_dist_config;

/* One parameter variant of a batched run: the diffusion parameters of
 * read_particle_params (particle_module.f90:2778-2878) that a sweep varies. */
typedef struct sp_variant {
    double kpara0;          /* kappa_parallel at p0, B0 */
    double kret;            /* kappa_perp / kappa_parallel */
    double gamma_turb;      /* turbulence index: kappa ~ p^(3 - gamma_turb) */
    double reserved0;       /* must be 0 */
} sp_variant;

/* ---- context --------------------------------------------------------- */
/* Replaces the module initialisation chain: set_mhd_config / set_field_configuration
 * (mhd_config.f90:43-74, simulation_setup.f90:171-253), init_field_data
 * (mhd_data_parallel.f90:66-101), init_particles (particle_module.f90:171-241),
 * read_particle_params (particle_module.f90:2778-2878) and the set_* calls
 * (particle_module.f90:260-335). */
int sp_ctx_create(const sp_config* cfg, void** out_ctx);
/* free_particles / free_field_data (particle_module.f90:246-251,
 * mhd_data_parallel.f90:187-196) */
int sp_ctx_destroy(void* ctx);
const char* sp_last_error(void);
int sp_abi_version(void);

/* ---- fields (mhd_data_parallel) --------------------------------------- */
/* read_field_data_parallel (mhd_data_parallel.f90:224-297): the 8 node fields
 * vx, vy, vz, rho, bx, by, bz, |B| as Fortran farray(1:8, -1:nx+2, -1:ny+2,
 * nz-range), FP32, ghost cells included.  slot 0 = farray1, 1 = farray2. */
int sp_fields_upload(void* ctx, int32_t slot, const float* nodes8, int64_t nvals);
/* calc_fields_gradients (mhd_data_parallel.f90:504-605) on the device.  When
 * farray32 is non-NULL the full 32-component farray (8 fields + 24
 * gradients) is also copied back. */
int sp_fields_gradients(void* ctx, int32_t slot, float* farray32);
/* copy_fields (mhd_data_parallel.f90:1920-1923): farray1 = farray2 */
int sp_fields_copy(void* ctx);

/* ---- particles (particle_module) -------------------------------------- */
int sp_particles_upload(void* ctx, const sp_particles* host, int64_t n);
int sp_particles_download(void* ctx, sp_particles* host, int64_t capacity, int64_t* n);
/* nptl_current, nptl_split, nptl_escaped, tag_max (module globals,
 * particle_module.f90:73-85) */
int sp_particles_counts(void* ctx, int64_t* nptl_current, int64_t* nptl_split,
                        int64_t* nptl_escaped, int32_t* tag_max);
/* inject_particles_spatial_uniform (particle_module.f90:454-530): the
 * reference takes the frame index ct_mhd and reads t0 = tstamps_mhd(ct_mhd),
 * dtf = tstamps_mhd(ct_mhd+1) - t0; they are passed directly here. */
int sp_inject_particles_spatial_uniform(void* ctx, int64_t nptl, double dt, int32_t dist_flag,
                                        double particle_v0, double t0, double dtf,
                                        const double part_box[6], double power_index);
/* Quantities of the threshold injections */
#define SP_INJECT_JZ 1      /* |dBy/dx - dBx/dy|  (inject_particles_at_large_jz) */
#define SP_INJECT_ABSJ 2    /* |curl B|           (inject_particles_at_large_absj) */
#define SP_INJECT_DB2 3     /* not supported: needs the turbulence-amplitude map */
#define SP_INJECT_DIVV 4    /* -div V             (inject_particles_at_large_divv) */
#define SP_INJECT_RHO 5     /* rho                (inject_particles_at_large_rho) */
/* get_ncells_large_jz / _absj / _divv / _rho (mhd_data_parallel.f90:2211-2499):
 * interior nodes of frame 0 strictly inside part_box (axes beyond ndim
 * untested) whose quantity exceeds threshold, in the reference's precision
 * (|jz|, |J| in FP32).  The divV count reads each node's value two nodes
 * below along every ghosted axis, as the reference's does. */
int sp_ncells_above_threshold(void* ctx, int32_t quantity, double threshold,
                              const double part_box[6], int64_t* ncells);
/* inject_particles_at_large_jz / _absj / _divv / _rho (particle_module.f90:
 * 785-1470), one rank: nptl_inject = int(nptl * ncells / ncells_norm), or nptl
 * with inject_same_nptl (none when no cell passes).  Each particle draws
 * positions over the domain until one lies in part_box and its quantity
 * (frame 0, interpolated) reaches threshold; then mu and the momentum law as
 * sp_inject_particles_spatial_uniform.  A particle needing more than 2^24 draws
 * fails the call (SP_ERR_ARG) before anything is stored.  *nptl_injected
 * (may be NULL) receives nptl_inject. */
int sp_inject_particles_threshold(void* ctx, int32_t quantity, int64_t nptl, double dt,
                                  int32_t dist_flag, double particle_v0, double t0, double dtf,
                                  int32_t inject_same_nptl, double threshold,
                                  int64_t ncells_norm, const double part_box[6],
                                  double power_index, int64_t* nptl_injected);
/* particle_mover (particle_module.f90:1846-1974) for Parker transport with the
 * standard kappa (nlgc = .false.), one rank: push every particle to t0 + dtf,
 * apply the exact-bounds boundary pass and remove non-INBOX particles. */
int sp_particle_mover(void* ctx, double t0, double dtf, int32_t num_fine_steps,
                      int32_t nsteps_interval, sp_mover_stats* stats);
/* No reference counterpart (the reference never reorders particles; its
 * quicksort_particle, particle_module.f90:5485-5512, is unused): orders the
 * particle arrays by a Morton key of coarse grid blocks so that particles
 * pushed together gather from neighbouring field records.  Changes the array
 * order only; every particle's result is unchanged. */
int sp_sort_particles(void* ctx);
/* split_particle (particle_module.f90:5430-5480) */
int sp_split_particle(void* ctx, double split_ratio, double pmin_split);
/* One gather + kappa + push per particle from caller-supplied states and
 * increments (ran1, ran2, ran3, ran_p per particle, interleaved), as one
 * iteration of the hot loop (particle_module.f90:1615-1690).  state rows are
 * (x, y, z, p, t, dt) per particle and are updated in place; deltas receives
 * (dx, dy, dz, dp) per particle.  fixed_dt is per particle (0/1).  With nlgc,
 * row i takes its pitch-angle cosine mu from the context's particle i
 * (n <= nptl_current). */
int sp_step_once(void* ctx, int64_t n, double t0, double dtf, double* state,
                 const double* rands, const int8_t* fixed_dt, double* deltas);
/* No counterpart in the reference.  sp_step_once without device allocation,
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernel and
 * results, with the device staging of the rows in a fixed per-GPU array of
 * the library (7.06 MiB, set up by sp_ctx_create).  At most 65536 rows per
 * call (SP_ERR_CAPACITY beyond).  Calls on one GPU share that array: calls
 * that overlap on one GPU (from several host threads or contexts) race. */
int sp_step_once_noalloc(void* ctx, int64_t n, double t0, double dtf, double* state,
                         const double* rands, const int8_t* fixed_dt, double* deltas);

/* ---- launch autotuning (no reference counterpart) -------------------------
 * Times the mover over [t0, t0 + dtf] from the current particles for every
 * triple of particles per lane (1, or 1 and 2 for the FP32 precisions),
 * resident 256-lane blocks per CU (1 .. occupancy, at most 4) and step budget
 * between re-sorts (0, 100, 200, 400 when sort_every > 0), keeps the fastest
 * for this context's later mover calls, and restores the particles and
 * counters, so the call changes no result.  Pick dtf so that a trial takes a
 * fraction of a second.  Launch settings never change a particle's result. */
typedef struct sp_tune_result {
    int32_t blocks_per_cu;        /* chosen resident mover blocks per CU */
    int32_t sort_steps;           /* chosen step budget between re-sorts */
    int32_t ntrials;
    int32_t particles_per_lane;   /* chosen particles each mover lane carries */
    double steps_per_s;           /* mover throughput of the chosen setting */
    double default_steps_per_s;   /* mover throughput of the untuned setting */
} sp_tune_result;
int sp_autotune_mover(void* ctx, double t0, double dtf, int32_t num_fine_steps,
                      int32_t nsteps_interval, sp_tune_result* out);

/* ---- batched parameter sweeps (no reference counterpart) ----------------
 * The reference runs one parameter set per job.  Here one context can push
 * many variants at once: each particle's `origin` holds its variant index and
 * the mover takes that variant's kpara0, kret and gamma_turb; every other
 * setting is the context's.  Each particle's result is the one a
 * single-variant run with that variant's parameters gives.
 * sp_set_variants installs nvar variants (nvar = 0 returns to the
 * configuration's values). */
int sp_set_variants(void* ctx, int32_t nvar, const sp_variant* variants);
/* inject_particles_spatial_uniform for each variant v in turn: nptl[v]
 * particles in part_boxes[6 v .. 6 v + 5], origin = v. */
int sp_inject_variants(void* ctx, int32_t nvar, const int64_t* nptl, double dt, int32_t dist_flag,
                       double particle_v0, double t0, double dtf, const double* part_boxes,
                       double power_index);
/* f(p) per variant: fglobal is (npp_global, nvar) in Fortran order (the
 * momentum bins of sp_distributions_init, summed over mu); counts[v] is the
 * number of particles of variant v (NULL skips).  Needs sp_distributions_init. */
int sp_variant_distributions(void* ctx, int32_t nvar, double* fglobal, int64_t* counts);

/* ---- f(p) surrogate network (no reference counterpart) -------------------
 * A dense network nin -> nhidden (x nlayers, SiLU) -> nout, trained and run on
 * the GPU with BF16 matrix-core GEMMs that accumulate in FP32; AdamW keeps
 * FP32 master weights.  Parameters are one FP32 vector: for each linear layer
 * l its weights W_l (K_l x N_l, row-major, Z = X W + b) then its bias b_l.
 * Initial values: U(-1/sqrt(K_l), 1/sqrt(K_l)) from a splitmix64 stream of
 * `seed`.  Training minimises the mean squared error over all outputs. */
typedef struct sp_mlp_config {
    int32_t nin, nhidden, nlayers, nout;
    int32_t batch_max;          /* rows per device pass (forward splits larger inputs) */
    int32_t reserved0;          /* must be 0 */
    uint64_t seed;
    double beta1, beta2, eps, weight_decay;   /* AdamW (decoupled decay, bias correction) */
} sp_mlp_config;
int sp_mlp_create(const sp_mlp_config* cfg, void** out_mlp);
int sp_mlp_destroy(void* mlp);
int sp_mlp_num_params(void* mlp, int64_t* nparams);
int sp_mlp_get_params(void* mlp, float* params);
/* Sets the parameters and clears the AdamW moments and step count. */
int sp_mlp_set_params(void* mlp, const float* params);
/* y (rows x nout) = network(x (rows x nin)), row-major host arrays. */
int sp_mlp_forward(void* mlp, const float* x, int64_t rows, float* y);
/* Device time (HIP events) of the layer kernels of the last sp_mlp_forward on
 * this network, or of the forward launches of the last ensemble call whose
 * first member it is. */
int sp_mlp_last_forward_ms(void* mlp, double* ms);
/* Mean squared error of network(x) against t and its gradient with respect
 * to every parameter (rows <= batch_max); no update.  grads may be NULL. */
int sp_mlp_gradients(void* mlp, const float* x, const float* t, int64_t rows, float* grads,
                     double* loss);
/* One AdamW step at learning rate lr on the rows given. */
int sp_mlp_train_step(void* mlp, const float* x, const float* t, int64_t rows, double lr,
                      double* loss);
/* Ensemble of networks with equal nin, nout whose outputs are log10(f + floor):
 * spec (rows x nout) = mean over members of max(10^y - floor, 0); spread
 * (rows) = mean over members of the L1 distance of their spectrum from spec. */
int sp_mlp_ensemble_spectra(void* const* nets, int32_t members, const float* x, int64_t rows,
                            double floor, float* spec, float* spread);
/* As sp_mlp_ensemble_spectra with x, spec and spread in device memory (HBM;
 * other pointers give SP_ERR_ARG); returns once the results are written. */
int sp_mlp_ensemble_spectra_device(void* const* nets, int32_t members, const float* x,
                                   int64_t rows, double floor, float* spec, float* spread);
/* Copies a training set (rows x nin inputs, rows x nout targets) to the device. */
int sp_mlp_set_data(void* mlp, const float* x, const float* t, int64_t rows);
/* nsteps AdamW steps on device-drawn mini-batches of `batch` rows (with
 * replacement, Philox keyed on seed and step), learning rate decaying from
 * lr0 to lr_min on a cosine; losses[s] receives each step's batch loss (NULL
 * skips).  Every kernel of a step reads the step number from device memory,
 * so one step is captured as a HIP graph and replayed nsteps times, without
 * host synchronisation (SP_MLP_EAGER=1 launches the same kernels one by one). */
int sp_mlp_train(void* mlp, int64_t nsteps, int32_t batch, double lr0, double lr_min,
                 uint64_t seed, float* losses);

/* ---- diagnostics ------------------------------------------------------ */
/* init_particle_distributions (diagnostics.f90:178-226) */
int sp_distributions_init(void* ctx, const sp_dist_config* dc);
/* calc_particle_distributions (diagnostics.f90:738-906): fglobal is
 * (nmu_global, npp_global); flocal[k] is (nmu, npbins, nrx, nry, nrz), all
 * Fortran order; NULL skips a set.  The device arrays are cleared afterwards
 * (clean_particle_distributions). */
int sp_distributions(void* ctx, double* fglobal, double* flocal1, double* flocal2,
                     double* flocal3, double* flocal4);
int sp_local_dist_dims(void* ctx, int32_t set, int64_t dims[5]);
/* remove_particles' copy of escaped particles (particle_module.f90:5382-5390),
 * calc_escaped_distributions (diagnostics.f90:913-1231),
 * clean_escaped_distributions (:492-584) and reset_escaped_particles
 * (:2227-2246) in one: every particle that sp_particle_mover removed with an
 * escape flag since the previous call is binned by the mover's removal pass;
 * this returns the bins and clears them.  Needs escaped_dist in
 * sp_distributions_init.  fescaped is (nmu_global, npp_global, 2 ndim), the
 * last index the face |count_flag|.  fescK is set K's face arrays back to back:
 * x faces (nmu, npbins, nry, nrz, 2), y faces (nmu, npbins, nrx, nrz, 2) and in
 * 3D z faces (nmu, npbins, nrx, nry, 2), dimensions as sp_local_dist_dims;
 * NULL skips an array. */
int sp_escaped_distributions(void* ctx, double* fescaped, double* fesc1, double* fesc2,
                             double* fesc3, double* fesc4);
/* No reference counterpart: a plain device-to-device streaming copy of
 * `bytes` (multiple of 16) on the context stream, timed with HIP events.  It
 * calibrates hardware memory counters against a kernel of known traffic. */
int sp_calibration_copy(void* ctx, int64_t bytes, double* ms);
/* quick_check + get_pmax_global (diagnostics.f90:116-170,1691-1719):
 * out = {nptl_current, nptl_split, ntot, leak, leak_negp, pdt_min, pdt_max,
 *        pdt_avg, pmax} */
int sp_quick_check(void* ctx, double out[9]);

#ifdef __cplusplus
}
#endif
#endif
