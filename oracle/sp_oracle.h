/*
 * sp_oracle.h -- CPU restatement of the stochastic-parker Parker pusher.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libsp_hip.so) never links or calls it.
 *
 * Arithmetic follows the reference Fortran (FP32 field storage, FP64 gather
 * and coefficient math, operation order of the Fortran expressions) and is
 * compiled with -ffp-contract=off so that every +,-,*,/,sqrt is a single
 * IEEE-rounded operation.  Each function cites the reference lines it
 * restates (paths relative to the reference repository root).
 */
#ifndef SP_ORACLE_H
#define SP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* count_flag values, src/modules/particle_module.f90:55-62 */
enum {
    ORC_INBOX = 1,
    ORC_OTHERS = 0,
    ORC_ESC_LX = -1, ORC_ESC_HX = -2,
    ORC_ESC_LY = -3, ORC_ESC_HY = -4,
    ORC_ESC_LZ = -5, ORC_ESC_HZ = -6
};

/* particle_type, particle_module.f90:38-50, plus the per-particle
 * counter-based RNG step index (rng_step) that replaces the per-thread
 * Mersenne-Twister stream. */
typedef struct {
    int8_t split_times;
    int8_t count_flag;
    int32_t origin;
    int32_t nsteps_tracked;
    int32_t nsteps_pushed;
    int32_t tag_injected;
    int32_t tag_splitted;
    uint32_t rng_step;
    double x, y, z, p;
    double v, mu;
    double weight, t, dt;
    double padding;
} orc_particle;

/* The configuration the pusher reads from mhd_config.dat, conf.dat and the
 * command line (mhd_config.f90:16-29, particle_module.f90:2778-2878,
 * simulation_setup.f90:104-150, stochastic-mhd.f90:572-946). */
typedef struct {
    int32_t ndim;                 /* ndim_field: 2 or 3 */
    int32_t nx, ny, nz;           /* interior nodes */
    double dx, dy, dz;
    double xmin, ymin, zmin;
    double xmax, ymax, zmax;
    double lx, ly, lz;
    int32_t pbcx, pbcy, pbcz;     /* 0 periodic, 1 open */
    double p0, pmin, pmax;
    int32_t momentum_dependency, mag_dependency;
    double gamma_turb, pindex;    /* pindex = 3 - gamma_turb */
    double kpara0, kret;
    double dt_min_rel, dt_max_rel;
    double drift1, drift2;
    int32_t pcharge;
    int32_t check_drift_2d;
    int32_t acc_region_flag;
    double acc_region[6];
    int32_t time_interp;
    uint64_t seed;                /* Philox key */
    uint32_t stream;              /* Philox counter word 3 (variant id) */
    /* momentum diffusion (set_dpp_params, particle_module.f90:260-272) */
    int32_t dpp_wave, dpp_shear, weak_scattering;
    double tau0;
    /* NLGC kperp (calc_spatial_diffusion_coefficients_nlgc) */
    int32_t nlgc;
    double kperp_kpara;
} orc_params;

/* kappa_type, particle_module.f90:93-104 */
typedef struct {
    double knorm_para, knorm_perp;
    double kpara, kperp;
    double skpara, skperp, skpara_perp;
    double kxx, kyy, kzz, kxy, kxz, kyz;
    double dkxx_dx, dkyy_dy, dkzz_dz;
    double dkxy_dx, dkxy_dy, dkxz_dx, dkxz_dz, dkyz_dy, dkyz_dz;
} orc_kappa;

/* ---- random numbers -------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* u32 -> closed [0,1] double, the mapping of genrand_double1 / real1 */
double orc_u01(uint32_t r);

typedef struct { uint32_t mt[624]; int mti; } orc_mt19937;
void orc_mt_init_genrand(orc_mt19937* s, uint32_t seed);
void orc_mt_init_by_array(orc_mt19937* s, const uint32_t* key, int len);
uint32_t orc_mt_genrand_int32(orc_mt19937* s);

/* ---- fields ---------------------------------------------------------- */
/* farray layout: Fortran farray(32, -1:nx+2, -1:ny+2, 1 | -1:nz+2) as a flat
 * float array, component fastest (mhd_data_parallel.f90:66-101). */
int64_t orc_field_nodes(const orc_params* prm);
void orc_calc_fields_gradients(const orc_params* prm, float* farray);
void orc_get_interp_paramters(const orc_params* prm, double px, double py, double pz,
                              int pos[3], double weights[8]);
void orc_interp_fields(const orc_params* prm, const float* farray1, const float* farray2,
                       const int pos[3], const double weights[8], double rt,
                       double fields[32]);

/* ---- physics --------------------------------------------------------- */
void orc_kappa_standard(const orc_params* prm, double p, const double fields[32],
                        orc_kappa* kappa);
void orc_kappa_nlgc(const orc_params* prm, double p, double mu, const double fields[32],
                    orc_kappa* k);
/* One Parker step.  rands = (ran1, ran2, ran3, ran_p) already mapped to the
 * increment law; dtmm = (dt_min, dt_max).  Deltas are written to delta[4]
 * (x, y, z, p); for 2D delta[2] is left at 0 as in the reference. */
void orc_push_particle_2d(const orc_params* prm, const double dtmm[2], orc_particle* ptl,
                          const double fields[32], const orc_kappa* kappa, int fixed_dt,
                          const double rands[4], double delta[4]);
void orc_push_particle_3d(const orc_params* prm, const double dtmm[2], orc_particle* ptl,
                          const double fields[32], const orc_kappa* kappa, int fixed_dt,
                          const double rands[4], double delta[4]);

/* gather + kappa + push from a particle state, as in one iteration of the
 * reference hot loop.  Used for single-step parity tests. */
void orc_step_once(const orc_params* prm, const float* farray1, const float* farray2,
                   double t0, double dtf, orc_particle* ptl, int fixed_dt,
                   const double rands[4], double delta[4]);

/* ---- particle population --------------------------------------------- */
typedef struct {
    double leak, leak_negp;
    int64_t steps;
    int64_t nptl_escaped;
} orc_mover_acc;

/* rng_mode: 0 = Philox keyed on (tag_injected, tag_splitted, rng_step,
 * stream); 1 = per-thread MT19937 streams seeded by init_by_array with
 * {0x123,0x234,0x345,0x456} + thread id (reference-like; results depend on
 * the thread count, as in the reference). */
void orc_particle_mover(const orc_params* prm, const float* farray1, const float* farray2,
                        orc_particle* ptls, int64_t* nptl_current,
                        double t0, double dtf, int num_fine_steps, int nsteps_interval,
                        int rng_mode, int nthreads, orc_mover_acc* acc);

void orc_remove_particles(orc_particle* ptls, int64_t* nptl_current, int64_t* nptl_escaped);

/* quantities of the threshold injections */
enum { ORC_THR_JZ = 1, ORC_THR_ABSJ = 2, ORC_THR_DIVV = 4, ORC_THR_RHO = 5 };
int64_t orc_ncells_above_threshold(const orc_params* prm, const float* farray1, int quantity,
                                   double threshold, const double part_box[6]);
int64_t orc_inject_particles_threshold(const orc_params* prm, const float* farray1,
                                       orc_particle* ptls, int64_t* nptl_current,
                                       int64_t nptl_max, int quantity, int64_t nptl, double dt,
                                       int dist_flag, double particle_v0, double t0, double dtf,
                                       int inject_same_nptl, double threshold,
                                       int64_t ncells_norm, const double part_box[6],
                                       double power_index, int32_t* tag_max);
void orc_inject_particles_spatial_uniform(const orc_params* prm, orc_particle* ptls,
                                          int64_t* nptl_current, int64_t nptl_max,
                                          int64_t nptl, double dt, int dist_flag,
                                          double particle_v0, double t0, double dtf,
                                          const double part_box[6], double power_index,
                                          int32_t* tag_max);

void orc_split_particle(const orc_params* prm, orc_particle* ptls, int64_t* nptl_current,
                        int64_t nptl_max, double split_ratio, double pmin_split,
                        int64_t* nptl_split);

/* Distribution diagnostics, diagnostics.f90:178-226,263-284,738-906.
 * local[k] = {enabled, npbins, nmu, rx, ry, rz}; local_p[k] = {pmin, pmax}. */
typedef struct {
    int32_t npp_global, nmu_global;
    int32_t local_dist;
    int32_t local[4][6];
    double local_p[4][2];
} orc_dist_config;
int64_t orc_local_dist_size(const orc_params* prm, const orc_dist_config* dc, int k);
int64_t orc_escaped_dist_size(const orc_params* prm, const orc_dist_config* dc, int k);
void orc_calc_escaped_distributions(const orc_params* prm, const orc_dist_config* dc,
                                    const orc_particle* escaped, int64_t nesc,
                                    double* fescaped, double* fesc[4]);
void orc_calc_particle_distributions(const orc_params* prm, const orc_dist_config* dc,
                                     const orc_particle* ptls, int64_t nptl,
                                     double* fglobal, double* flocal[4]);

/* quick_check + get_pmax_global, diagnostics.f90:116-170,1691-1719.
 * out = {nptl_current, nptl_split, ntot, leak, leak_negp, pdt_min, pdt_max,
 *        pdt_avg, pmax} */
void orc_quick_check(const orc_particle* ptls, int64_t nptl, int64_t nptl_split,
                     double leak, double leak_negp, double out[9]);

int orc_sizeof_particle(void);

#ifdef __cplusplus
}
#endif
#endif
