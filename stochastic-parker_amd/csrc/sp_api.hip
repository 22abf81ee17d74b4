// sp_api.hip -- the extern "C" entry points of libsp_hip.so (include/sp_hip.h).
//
// A context owns one HIP stream (default priority), the field records of one
// or two frames, two particle buffer sets (compaction writes the other set
// and swaps), the device counters and the diagnostics arrays.  Every entry
// point validates its arguments, enqueues its kernels on the context stream
// and synchronises before returning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sp_internal.h"
#include "sp_particles.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

int sp::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define SP_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(SP_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

constexpr double kInvFx = 1.0 / sp::kFxScale;
// Resident 256-lane mover blocks per CU, capped by the kernel's occupancy:
// 2D FP32 movers three (their one-frame kernels are built for three waves per
// SIMD), the 2D FP64 mover two; 3D FP16 records two (their register cache
// leaves the die's L2 a smaller window to hold), other 3D movers one.
constexpr int kMoverBlocksPerCU2D = 3;
constexpr int kMoverBlocksPerCU2D64 = 2;
constexpr int kMoverBlocksPerCU3DH = 2;
constexpr int kMoverBlocksPerCU3D = 1;
// Particles per lane of the FP32 movers unless tuned or overridden.
constexpr int kMoverPPLDefault = 1;

struct Ctx {
    sp_config cfg{};
    sp::DevParams dp{};
    int device = 0;
    int ncu = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    long long nnodes = 0;
    int recw = 0;
    long long nrec = 0;      // records per frame (3D: whole 8^3 bricks)
    float* n8[2] = {nullptr, nullptr};
    float* rec[2] = {nullptr, nullptr};
    bool have_fields[2] = {false, false};
    sp::Ptl P[2]{};
    void* pbuf[2] = {nullptr, nullptr};
    int cur = 0;
    long long nptl = 0, nmax = 0, nsplit = 0, nesc = 0;
    int tag_max = 0;
    // cumulative leak / leak_negp in 2^-32 weight units; the device counters
    // are folded in and cleared after every mover call
    unsigned __int128 leak_acc = 0, negp_acc = 0;
    sp::Counters* ctr = nullptr;
    unsigned* tile_cnt = nullptr;
    unsigned* tile_esc = nullptr;
    unsigned long long* tile_off = nullptr;
    unsigned long long* totals = nullptr;
    double* stats_part = nullptr;
    int stats_grid = 1024;
    // spatial sort scratch
    uint32_t* skey[2] = {nullptr, nullptr};
    uint32_t* sidx[2] = {nullptr, nullptr};
    unsigned* shist = nullptr;
    unsigned long long* soff = nullptr;
    long long mover_calls = 0;
    // autotuned launch settings (0 / -1: not tuned) and the tuner's state copy
    int tuned_bpc = 0, tuned_sort_steps = -1;
    void* pbuf_save = nullptr;
    // batched parameter variants
    sp::KappaVar* dvar = nullptr;
    int nvar = 0, nvar_cap = 0;
    unsigned long long* vhist = nullptr;   // per-variant f(p) bins, then per-variant counts
    size_t vhist_len = 0;
    // occupancy of the mover per precision (F64, F32, F32H) and particles per lane (1, 2)
    int mover_bpc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    int tuned_ppl = 0;                        // autotuned particles per lane (0: not tuned)
    float gscale[2] = {1.0f, 1.0f};           // FP16 records: gradient scale per frame
    float vscale[2] = {1.0f, 1.0f};           // FP16 3D records: velocity scale per frame
    unsigned* gmax = nullptr;
    // distributions
    bool dist_ready = false;
    sp::DistArgs da{};
    long long nglob = 0;
    long long loc_size[4] = {0, 0, 0, 0};
    long long loc_dims[4][5] = {};
    unsigned long long* dglob = nullptr;
    unsigned long long* dloc[4] = {nullptr, nullptr, nullptr, nullptr};
    // distributions of escaped particles: global, then per set the x, y, z
    // face arrays back to back
    sp::EscArgs ea{};
    long long esc_size[5] = {0, 0, 0, 0, 0};
    unsigned long long* desc[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

Ctx* as_ctx(void* p) { return static_cast<Ctx*>(p); }

// Staging memory of sp_step_once_noalloc, one copy per GPU for the life of
// the process: up to kNoAllocRows rows of state (6 doubles), increments (4
// doubles), deltas (4 doubles) and fixed-dt flags (1 byte), 7.06 MiB.  In
// host code `scratch` names a host-side stand-in; kernels and copies use the
// device address that hipGetSymbolAddress returns for the current GPU, kept
// per device in g_scratch.
constexpr long long kNoAllocRows = 65536;
constexpr size_t kScratchBytes = (size_t)kNoAllocRows * (14 * sizeof(double) + 1);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Looks up this GPU's copy of `scratch` once (the device must be current).
int init_scratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(SP_ERR_UNSUPPORTED, "device index beyond the staging table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return SP_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SP_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return SP_OK;
}

void* device_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

double fx_weight(unsigned __int128 v) { return (double)v * kInvFx; }

void free_dev(void* p) {
    if (p) (void)hipFree(p);
}

// One allocation per particle buffer set: capacity rounded up to 256 so every
// array starts 256-byte aligned; array order and offsets are those of
// sp::PtlBase (10 FP64 arrays, 2 int8 arrays, 6 int32 arrays).
long long particle_cap(long long nmax) { return (nmax + 255) / 256 * 256; }

int alloc_particles(long long nmax, sp::Ptl& P, void*& base) {
    const long long cap = particle_cap(nmax);
    const size_t total = (size_t)sp::kPtlBytes * (size_t)cap;
    if (hipMalloc(&base, total) != hipSuccess)
        return fail(SP_ERR_NOMEM, "device allocation of " + std::to_string(total) + " bytes for particles failed");
    const sp::PtlBase B{static_cast<char*>(base), cap};
    P.x = B.d(sp::PD_X); P.y = B.d(sp::PD_Y); P.z = B.d(sp::PD_Z); P.p = B.d(sp::PD_P);
    P.v = B.d(sp::PD_V); P.mu = B.d(sp::PD_MU); P.w = B.d(sp::PD_W); P.t = B.d(sp::PD_T);
    P.dt = B.d(sp::PD_DT);
    P.rdt = B.d(sp::PD_RDT);
    P.split_times = B.b(sp::PB_SPLIT);
    P.count_flag = B.b(sp::PB_FLAG);
    P.origin = B.i(sp::PI_ORIGIN);
    P.nsteps_tracked = B.i(sp::PI_TRACKED);
    P.nsteps_pushed = B.i(sp::PI_PUSHED);
    P.tag_injected = B.i(sp::PI_TAGI);
    P.tag_splitted = B.i(sp::PI_TAGS);
    P.rng_step = reinterpret_cast<uint32_t*>(B.i(sp::PI_RSTEP));
    return SP_OK;
}

int validate(const sp_config* c) {
    if (!c) return fail(SP_ERR_ARG, "config is NULL");
    if (c->ndim != 2 && c->ndim != 3) return fail(SP_ERR_UNSUPPORTED, "ndim must be 2 or 3");
    if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail(SP_ERR_ARG, "nx, ny, nz must be >= 1");
    if (c->ndim == 2 && c->nz != 1) return fail(SP_ERR_ARG, "a 2D grid has nz == 1");
    if (!(c->dx > 0) || !(c->dy > 0) || !(c->dz > 0)) return fail(SP_ERR_ARG, "dx, dy, dz must be > 0");
    if (!(c->lx > 0) || !(c->ly > 0) || !(c->lz > 0)) return fail(SP_ERR_ARG, "lx, ly, lz must be > 0");
    if (!(c->p0 > 0) || !(c->pmin > 0) || !(c->pmax > c->pmin)) return fail(SP_ERR_ARG, "need p0 > 0 and 0 < pmin < pmax");
    if (!(c->kpara0 > 0)) return fail(SP_ERR_ARG, "kpara0 must be > 0");
    if (!(c->kret >= 0) || c->kret > 1) return fail(SP_ERR_ARG, "kret must be in [0, 1]");
    if (!(c->dt_min_rel > 0) || !(c->dt_max_rel >= c->dt_min_rel)) return fail(SP_ERR_ARG, "need 0 < dt_min_rel <= dt_max_rel");
    if (c->charge == 0) return fail(SP_ERR_ARG, "charge must be non-zero (the drift prefactor divides by it)");
    if (c->drift1 == 0 && c->drift2 == 0) return fail(SP_ERR_ARG, "drift1 and drift2 are both zero (the drift speed would be infinite)");
    if (c->pbcx < 0 || c->pbcx > 1 || c->pbcy < 0 || c->pbcy > 1 || c->pbcz < 0 || c->pbcz > 1) return fail(SP_ERR_ARG, "pbc must be 0 or 1");
    if (c->precision != SP_PREC_F64 && c->precision != SP_PREC_F32 && c->precision != SP_PREC_F32H)
        return fail(SP_ERR_ARG, "precision must be SP_PREC_F64, SP_PREC_F32 or SP_PREC_F32H");
    if (c->nptl_max < 1 || c->nptl_max > (1ll << 40)) return fail(SP_ERR_ARG, "nptl_max out of range");
    if (c->sort_steps < 0 || c->reserved0 != 0) return fail(SP_ERR_ARG, "need sort_steps >= 0 and reserved0 == 0");
    if (c->sort_every < 0) return fail(SP_ERR_ARG, "sort_every must be >= 0");
    if (c->dpp_wave < 0 || c->dpp_wave > 1 || c->dpp_shear < 0 || c->dpp_shear > 1 ||
        c->weak_scattering < 0 || c->weak_scattering > 1 || c->nlgc < 0 || c->nlgc > 1)
        return fail(SP_ERR_ARG, "dpp_wave, dpp_shear, weak_scattering and nlgc must be 0 or 1");
    if (c->nlgc && !(c->kperp_kpara >= 0)) return fail(SP_ERR_ARG, "kperp_kpara must be >= 0");
    if (c->dpp_shear && !(c->tau0 >= 0)) return fail(SP_ERR_ARG, "tau0 must be >= 0");
    if ((c->dpp_wave || c->dpp_shear) && c->precision == SP_PREC_F32H)
        return fail(SP_ERR_UNSUPPORTED, "momentum diffusion needs SP_PREC_F64 or SP_PREC_F32 records");
    return SP_OK;
}

// Step features of a configuration (template argument of the mover).  The
// lean mark goes with exactly the configurations whose flags (mover_params)
// hold neither MF_MAG_DEP nor MF_DRIFT_2D.
int step_features(const sp_config& c) {
    const bool lean = c.mag_dependency != 1 && !(c.ndim == 2 && c.check_drift_2d);
    return (c.acc_region_flag == 1 ? sp::FX_ACC : 0) | (c.dpp_wave || c.dpp_shear ? sp::FX_DPP : 0) |
           (c.nlgc ? sp::FX_NLGC : 0) | (lean ? sp::FX_LEAN : 0);
}

sp::DevParams make_params(const sp_config& c) {
    sp::DevParams d{};
    d.ndim = c.ndim;
    d.nxg = c.nx + 4;
    d.nyg = c.ny + 4;
    d.nzg = c.ndim == 3 ? c.nz + 4 : c.nz;
    d.dx = c.dx; d.dy = c.dy; d.dz = c.dz;
    d.idx = 1.0 / c.dx; d.idy = 1.0 / c.dy; d.idz = 1.0 / c.dz;
    d.xmin = c.xmin; d.ymin = c.ymin; d.zmin = c.zmin;
    d.xmax = c.xmax; d.ymax = c.ymax; d.zmax = c.zmax;
    d.lx = c.lx; d.ly = c.ly; d.lz = c.lz;
    d.pbc[0] = c.pbcx; d.pbc[1] = c.pbcy; d.pbc[2] = c.pbcz;
    d.momentum_dependency = c.momentum_dependency;
    d.mag_dependency = c.mag_dependency;
    d.pcharge = c.charge;
    d.p0 = c.p0; d.pmin = c.pmin; d.pmax = c.pmax;
    d.gamma_turb = c.gamma_turb;
    d.pindex = 3.0 - c.gamma_turb;  // particle_module.f90:2798
    d.kpara0 = c.kpara0; d.kret = c.kret;
    d.drift1 = c.drift1; d.drift2 = c.drift2;
    d.check_drift_2d = c.check_drift_2d;
    d.acc_region_flag = c.acc_region_flag;
    d.time_interp = c.time_interp;
    d.precision = c.precision;
    d.ext_rec = (c.dpp_wave || c.dpp_shear) ? 1 : 0;
    d.bnx = (d.nxg + 7) / 8;
    d.bny = (d.nyg + 7) / 8;
    d.bnz = (d.nzg + 7) / 8;
    for (int i = 0; i < 6; ++i) d.acc_region[i] = c.acc_region[i];
    d.key0 = (uint32_t)c.seed;
    d.key1 = (uint32_t)(c.seed >> 32);
    d.stream = c.stream;
    return d;
}

Ctx* live_ctx(void* p, int& rc) {
    if (!p) {
        rc = fail(SP_ERR_ARG, "context is NULL");
        return nullptr;
    }
    rc = SP_OK;
    return as_ctx(p);
}

// Particles each mover lane carries: 1 for the FP64 mover and for two-frame
// time interpolation; the FP32 movers take 1 or 2.  Precedence, last wins:
// the default, the autotuner's choice, SP_MOVER_PPL, an explicit trial value.
int mover_ppl(const Ctx* c, int ppl_trial = 0) {
    if (c->cfg.precision == SP_PREC_F64 || c->cfg.time_interp) return 1;
    int ppl = kMoverPPLDefault;
    if (c->tuned_ppl > 0) ppl = c->tuned_ppl;
    if (const char* e = std::getenv("SP_MOVER_PPL")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2) ppl = v;
    }
    if (ppl_trial == 1 || ppl_trial == 2) ppl = ppl_trial;
    return ppl;
}

// Occupancy-sized persistent grid for the mover (bpc > 0: that many blocks
// per CU, within the occupancy limit).
int mover_grid(Ctx* c, long long n, int bpc = 0, int ppl = 1) {
    const int prec = c->cfg.precision;
    const int occ = c->mover_bpc[prec][ppl - 1];
    // Resident 256-lane blocks per CU.  Particles in flight are, on each XCD,
    // a window of consecutive (spatially sorted) indices; a smaller window
    // keeps its field records in that XCD's L2.  Measured on MI355X: three
    // blocks per CU for the FP32 movers on the 2048^2 snapshot (two for
    // FP64), two for the FP16-record mover on the 512^3 snapshot (-38 %
    // kernel time against one); SP_MOVER_BLOCKS_PER_CU overrides.
    const int per_cu_default =
        c->cfg.ndim == 3 ? (prec == SP_PREC_F32H ? kMoverBlocksPerCU3DH : kMoverBlocksPerCU3D)
                         : (prec == SP_PREC_F64 ? kMoverBlocksPerCU2D64 : kMoverBlocksPerCU2D);
    int per_cu = std::min(std::max(1, occ), per_cu_default);
    if (c->tuned_bpc > 0) per_cu = std::min(std::max(1, occ), c->tuned_bpc);
    if (const char* e = std::getenv("SP_MOVER_BLOCKS_PER_CU")) {
        const int v = std::atoi(e);
        if (v > 0) per_cu = std::min(std::max(1, occ), v);
    }
    if (bpc > 0) per_cu = std::min(std::max(1, occ), bpc);
    const long long lanes = 256ll * ppl;
    const long long need = (n + lanes - 1) / lanes;
    const long long cap = (long long)c->ncu * per_cu;
    return (int)std::max<long long>(1, std::min(need, cap));
}

// Orders the current particle set by the Morton key of 2^shift-cell blocks
// (at most 64 blocks per axis) into the spare buffer set and swaps.  With
// paused_first, only the first `n` particles are ordered, in place: those the
// mover paused (rdt > 0) first, in key order, then the others.
int spatial_sort(Ctx* c, bool paused_first = false, long long n = -1) {
    if (n < 0) n = c->nptl;
    if (n <= 1) return SP_OK;
    if (c->nptl >= (1ll << 32)) return fail(SP_ERR_UNSUPPORTED, "sorting needs fewer than 2^32 particles");
    const long long cap = particle_cap(c->nmax);
    if (!c->skey[0]) {
        const long long nt = std::max<long long>(1, sp::num_tiles(c->nmax));
        for (int i = 0; i < 2; ++i) {
            SP_HIP(hipMalloc(&c->skey[i], sizeof(uint32_t) * cap));
            SP_HIP(hipMalloc(&c->sidx[i], sizeof(uint32_t) * cap));
        }
        SP_HIP(hipMalloc(&c->shist, sizeof(unsigned) * 16 * nt));
        SP_HIP(hipMalloc(&c->soff, sizeof(unsigned long long) * 16 * nt));
    }
    const sp_config& f = c->cfg;
    sp::SortArgs sa{};
    sa.ndim = f.ndim;
    sa.nx = f.nx; sa.ny = f.ny; sa.nz = f.ndim == 3 ? f.nz : 1;
    int nmaxc = std::max(sa.nx, std::max(sa.ny, sa.nz)), bits = 0;
    while ((1 << bits) < nmaxc) ++bits;
    // blocks of 2^shift cells per axis; SP_SORT_SHIFT overrides the default
    sa.shift = std::max(0, bits - 6);
    if (const char* e = std::getenv("SP_SORT_SHIFT")) sa.shift = std::max(0, std::atoi(e));
    const int axis_bits = std::max(1, bits - sa.shift);
    if (axis_bits > (f.ndim == 3 ? 10 : 15)) return fail(SP_ERR_UNSUPPORTED, "sort key too wide");
    const int kbits = f.ndim * axis_bits + (paused_first ? 1 : 0);
    sa.done_bit = paused_first ? f.ndim * axis_bits : -1;
    sa.passes = (kbits + 3) / 4;
    sa.xmin = f.xmin; sa.ymin = f.ymin; sa.zmin = f.zmin;
    sa.idx = 1.0 / f.dx; sa.idy = 1.0 / f.dy; sa.idz = 1.0 / f.dz;
    const bool prefix = n < c->nptl;
    SP_HIP(sp::launch_spatial_sort(c->P[c->cur], c->P[1 - c->cur], n, sa, c->skey[0],
                                   c->sidx[0], c->skey[1], c->sidx[1], c->shist, c->soff,
                                   c->totals, prefix, c->st));
    if (!prefix) c->cur = 1 - c->cur;
    return SP_OK;
}

// Diffusion parameters of one variant in the forms both step paths use, by
// the same expressions as the configuration's (pindex = 3 - gamma_turb,
// particle_module.f90:2798).
sp::KappaVar kappa_var(double p0, double kpara0, double kret, double gamma_turb) {
    sp::KappaVar k{};
    k.kpara0 = kpara0;
    k.kret = kret;
    k.pindex = 3.0 - gamma_turb;
    k.gamma_m2 = gamma_turb - 2.0;
    k.p0_2mpi = std::pow(p0, 2.0 - k.pindex);   // p0**(2.0-pindex) (:2975-2977)
    k.perp_mag = (gamma_turb - 2.0) / 3.0;      // NLGC kperp exponents (:2499-2525)
    k.perp_mom = (5.0 - gamma_turb) / 3.0;
    k.f_perp_mag = (float)k.perp_mag;
    k.f_perp_mom = (float)k.perp_mom;
    k.f_kpara0 = (float)kpara0;
    k.f_kret = (float)kret;
    k.f_pindex = (float)k.pindex;
    k.f_gm2 = (float)k.gamma_m2;
    return k;
}

// Kernel arguments of the mover / single-step kernels for one interval.
sp::MoverParams mover_params(Ctx* c, double t0, double dtf, int num_fine_steps, int nsteps_interval) {
    const sp_config& f = c->cfg;
    const sp::DevParams& d = c->dp;
    sp::MoverParams m{};
    m.P = sp::PtlBase{static_cast<char*>(c->pbuf[c->cur]), particle_cap(c->nmax)};
    m.n = c->nptl;
    m.part_len = (c->nptl + sp::kParts - 1) / sp::kParts;
    m.rec1 = c->rec[0];
    m.rec2 = f.time_interp ? c->rec[1] : c->rec[0];
    m.ctr = c->ctr;
    m.nxg = d.nxg; m.nyg = d.nyg; m.nzg = d.nzg;
    m.bnx = d.bnx; m.bny = d.bny;
    m.nsteps_interval = nsteps_interval;
    m.flags = (f.pbcx == 1 ? sp::MF_OPEN_X : 0) | (f.pbcy == 1 ? sp::MF_OPEN_Y : 0) |
              (f.ndim == 3 && f.pbcz == 1 ? sp::MF_OPEN_Z : 0) |
              (f.momentum_dependency == 1 ? sp::MF_MOMENTUM_DEP : 0) |
              (f.mag_dependency == 1 ? sp::MF_MAG_DEP : 0) |
              (f.ndim == 2 && f.check_drift_2d ? sp::MF_DRIFT_2D : 0) |
              (f.dpp_wave ? sp::MF_DPP_WAVE : 0) | (f.dpp_shear ? sp::MF_DPP_SHEAR : 0) |
              (f.weak_scattering ? sp::MF_WEAK_SCATTERING : 0);
    m.key0 = d.key0; m.key1 = d.key1; m.stream = d.stream;
    m.t0 = t0;
    m.dtf = dtf;
    m.dt_fine = dtf / num_fine_steps;
    m.dt_limit = dtf + m.dt_fine * (double)0.1f;  // 0.1 is a default-real literal (:1596)
    m.xmin = f.xmin; m.ymin = f.ymin; m.zmin = f.zmin;
    m.dx = f.dx; m.dy = f.dy; m.dz = f.dz;
    m.idx = d.idx; m.idy = d.idy; m.idz = d.idz;
    // slack bounds xmin - dx*0.5 ... (particle_module.f90:1527-1532)
    m.slo[0] = f.xmin - f.dx * 0.5; m.slo[1] = f.ymin - f.dy * 0.5; m.slo[2] = f.zmin - f.dz * 0.5;
    m.shi[0] = f.xmax + f.dx * 0.5; m.shi[1] = f.ymax + f.dy * 0.5; m.shi[2] = f.zmax + f.dz * 0.5;
    m.elo[0] = f.xmin; m.elo[1] = f.ymin; m.elo[2] = f.zmin;
    m.ehi[0] = f.xmax; m.ehi[1] = f.ymax; m.ehi[2] = f.zmax;
    m.p0 = f.p0;
    m.p_floor = 0.25 * f.p0;
    const sp::KappaVar kv = kappa_var(f.p0, f.kpara0, f.kret, f.gamma_turb);
    m.kpara0 = kv.kpara0;
    m.kret = kv.kret;
    m.pindex = kv.pindex;
    m.gamma_m2 = kv.gamma_m2;
    m.var = c->dvar;
    m.nvar = c->nvar;
    m.vdp_pre = (double)(1.0f / (float)(3 * f.charge));
    m.d1p0 = f.drift1 * f.p0;
    m.d2p02 = f.drift2 * (f.p0 * f.p0);
    m.half_dx = 0.5 * f.dx; m.half_dy = 0.5 * f.dy; m.half_dz = 0.5 * f.dz;
    m.dt_min = f.dt_min_rel * dtf;  // set_dt_min_max (:5519-5524)
    m.dt_max = f.dt_max_rel * dtf;
    m.f_gs1 = 1.0f / c->gscale[0];
    m.f_gs2 = 1.0f / (f.time_interp ? c->gscale[1] : c->gscale[0]);
    m.f_vs1 = 1.0f / c->vscale[0];
    m.f_vs2 = 1.0f / (f.time_interp ? c->vscale[1] : c->vscale[0]);
    m.f_kpara0 = kv.f_kpara0;
    m.f_kret = kv.f_kret;
    m.f_pindex = kv.f_pindex;
    m.f_gm2 = kv.f_gm2;
    m.f_vdp_pre = 1.0f / (float)(3 * f.charge);
    m.f_d1p0 = (float)m.d1p0;
    m.f_d2p02 = (float)m.d2p02;
    m.f_half_dx = (float)m.half_dx; m.f_half_dy = (float)m.half_dy; m.f_half_dz = (float)m.half_dz;
    m.f_dt_min = (float)m.dt_min;
    m.f_dt_max = (float)m.dt_max;
    m.inv_p0 = 1.0 / f.p0;
    for (int i = 0; i < 6; ++i) m.acc_region[i] = f.acc_region[i];
    m.lx = f.lx; m.ly = f.ly; m.lz = f.lz;
    m.tau0 = f.tau0;
    m.p0_2mpi = kv.p0_2mpi;
    m.kperp_kpara = f.kperp_kpara;
    m.perp_mag = kv.perp_mag;
    m.perp_mom = kv.perp_mom;
    m.f_kperp_kpara = (float)f.kperp_kpara;
    m.f_perp_mag = kv.f_perp_mag;
    m.f_perp_mom = kv.f_perp_mom;
    m.f_tau0 = (float)f.tau0;
    m.f_p02 = (float)(f.p0 * f.p0);
    return m;
}

void destroy(Ctx* c) {
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (int i = 0; i < 2; ++i) {
        free_dev(c->n8[i]);
        free_dev(c->rec[i]);
        free_dev(c->pbuf[i]);
    }
    free_dev(c->ctr);
    free_dev(c->tile_cnt);
    free_dev(c->tile_esc);
    free_dev(c->tile_off);
    free_dev(c->totals);
    free_dev(c->stats_part);
    for (int i = 0; i < 2; ++i) {
        free_dev(c->skey[i]);
        free_dev(c->sidx[i]);
    }
    free_dev(c->shist);
    free_dev(c->soff);
    free_dev(c->dglob);
    for (auto* q : c->dloc) free_dev(q);
    for (auto* q : c->desc) free_dev(q);
    free_dev(c->dvar);
    free_dev(c->vhist);
    free_dev(c->gmax);
    free_dev(c->pbuf_save);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

}  // namespace

extern "C" {

const char* sp_last_error(void) { return g_err.c_str(); }

int sp_abi_version(void) { return SP_HIP_ABI_VERSION; }

int sp_ctx_create(const sp_config* cfg, void** out_ctx) {
    if (!out_ctx) return fail(SP_ERR_ARG, "out_ctx is NULL");
    *out_ctx = nullptr;
    int rc = validate(cfg);
    if (rc != SP_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SP_ERR_HIP, "no HIP device is available");
    Ctx* c = new Ctx();
    c->cfg = *cfg;
    c->dp = make_params(*cfg);
    c->nmax = cfg->nptl_max;
    c->nnodes = (long long)c->dp.nxg * c->dp.nyg * c->dp.nzg;
    if (c->dp.ext_rec) c->recw = cfg->ndim == 2 ? sp::kRec2DX : sp::kRec3DX;
    else if (cfg->precision == SP_PREC_F64) c->recw = cfg->ndim == 2 ? sp::kRec2D : sp::kRec3D;
    else if (cfg->precision == SP_PREC_F32) c->recw = cfg->ndim == 2 ? sp::kCell2D : sp::kRec3F;
    else c->recw = cfg->ndim == 2 ? sp::Pair2Shape::W16 : sp::kRec3H;
    c->nrec = cfg->ndim == 3 ? (long long)c->dp.bnx * c->dp.bny * c->dp.bnz * 512 : c->nnodes;
    auto bail = [&](int code) {
        destroy(c);
        return code;
    };
    // the FP32 movers index node records in 32 bits
    if (!c->dp.ext_rec && cfg->precision != SP_PREC_F64 && c->nrec >= 0xFFFFFFFFLL)
        return bail(fail(SP_ERR_UNSUPPORTED,
                         "grid has 2^32 or more (bricked) nodes: use SP_PREC_F64"));
#define SP_HIP_C(expr)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return bail(fail(SP_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_))); \
    } while (0)
    SP_HIP_C(hipGetDevice(&c->device));
    rc = init_scratch(c->device);
    if (rc != SP_OK) return bail(rc);
    hipDeviceProp_t prop;
    SP_HIP_C(hipGetDeviceProperties(&prop, c->device));
    c->ncu = prop.multiProcessorCount;
    SP_HIP_C(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    for (auto& e : c->ev) SP_HIP_C(hipEventCreate(&e));
    SP_HIP_C(hipMalloc(&c->ctr, sizeof(sp::Counters)));
    SP_HIP_C(hipMemsetAsync(c->ctr, 0, sizeof(sp::Counters), c->st));
    for (int i = 0; i < 2; ++i) {
        rc = alloc_particles(c->nmax, c->P[i], c->pbuf[i]);
        if (rc != SP_OK) return bail(rc);
    }
    const long long ntiles = std::max<long long>(1, sp::num_tiles(c->nmax));
    SP_HIP_C(hipMalloc(&c->tile_cnt, sizeof(unsigned) * ntiles));
    SP_HIP_C(hipMalloc(&c->tile_esc, sizeof(unsigned) * ntiles));
    SP_HIP_C(hipMalloc(&c->tile_off, sizeof(unsigned long long) * ntiles));
    SP_HIP_C(hipMalloc(&c->totals, sizeof(unsigned long long) * 2));
    SP_HIP_C(hipMalloc(&c->stats_part, sizeof(double) * 5 * c->stats_grid));
    // occupancy of the persistent mover decides its grid
    {
        const bool ti = cfg->time_interp != 0;
        const int fx = step_features(*cfg);
        c->mover_bpc[0][0] = c->mover_bpc[0][1] = std::max(1, sp::mover_blocks_per_cu_f64(cfg->ndim, ti, fx));
        for (int ppl = 1; ppl <= 2; ++ppl) {
            c->mover_bpc[1][ppl - 1] = std::max(1, sp::mover_blocks_per_cu_f32(cfg->ndim, ti, fx, ppl));
            c->mover_bpc[2][ppl - 1] =
                cfg->dpp_wave || cfg->dpp_shear ? 1 : std::max(1, sp::mover_blocks_per_cu_f32h(cfg->ndim, ti, fx, ppl));
        }
    }
    SP_HIP_C(hipStreamSynchronize(c->st));
#undef SP_HIP_C
    *out_ctx = c;
    return SP_OK;
}

int sp_ctx_destroy(void* ctx) {
    if (!ctx) return SP_OK;
    destroy(as_ctx(ctx));
    return SP_OK;
}

int sp_fields_upload(void* ctx, int32_t slot, const float* nodes8, int64_t nvals) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (slot < 0 || slot > 1) return fail(SP_ERR_ARG, "slot must be 0 or 1");
    if (slot == 1 && !c->cfg.time_interp) return fail(SP_ERR_STATE, "slot 1 needs time_interp");
    if (!nodes8) return fail(SP_ERR_ARG, "nodes8 is NULL");
    if (nvals != c->nnodes * 8)
        return fail(SP_ERR_ARG, "nodes8 must hold 8 * nxg * nyg * nzg = " + std::to_string(c->nnodes * 8) + " floats");
    if (!c->n8[slot]) SP_HIP(hipMalloc(&c->n8[slot], sizeof(float) * 8 * c->nnodes));
    SP_HIP(hipMemcpyAsync(c->n8[slot], nodes8, sizeof(float) * 8 * c->nnodes, hipMemcpyHostToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->have_fields[slot] = false;
    return SP_OK;
}

int sp_fields_gradients(void* ctx, int32_t slot, float* farray32) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (slot < 0 || slot > 1) return fail(SP_ERR_ARG, "slot must be 0 or 1");
    if (!c->n8[slot]) return fail(SP_ERR_STATE, "fields of this slot were not uploaded");
    if (!c->rec[slot]) SP_HIP(hipMalloc(&c->rec[slot], sizeof(float) * c->recw * c->nrec));
    float* fa = nullptr;
    if (farray32) SP_HIP(hipMalloc(&fa, sizeof(float) * 32 * c->nnodes));
    hipError_t e = hipSuccess;
    if (c->cfg.precision == SP_PREC_F32H) {
        // FP16 values: the largest |gradient| and the largest |V| of the frame
        // set power-of-two scales that map them into (2^13, 2^14] (exact to
        // undo; no FP16 overflow)
        if (!c->gmax) e = hipMalloc(&c->gmax, 2 * sizeof(unsigned));
        unsigned bits[2] = {0u, 0u};
        if (e == hipSuccess) e = hipMemsetAsync(c->gmax, 0, 2 * sizeof(unsigned), c->st);
        if (e == hipSuccess) e = sp::launch_value_max(c->dp, c->n8[slot], c->gmax, c->st);
        if (e == hipSuccess) e = hipMemcpyAsync(bits, c->gmax, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        float sc[2] = {1.0f, 1.0f};
        for (int q = 0; q < 2; ++q) {
            float mx = 0.0f;
            std::memcpy(&mx, &bits[q], sizeof(float));
            if (mx > 0.0f && std::isfinite(mx)) {
                const float r = 16384.0f / mx;
                if (std::isfinite(r)) {
                    int ex = 0;
                    std::frexp(r, &ex);   // 16384 / max in [2^(ex-1), 2^ex)
                    sc[q] = std::ldexp(1.0f, ex - 1);
                } else {
                    sc[q] = std::ldexp(1.0f, 126);   // values below ~1e-34 everywhere
                }
            }
        }
        c->gscale[slot] = sc[0];
        c->vscale[slot] = sc[1];
        if (e == hipSuccess) e = sp::launch_pack_fast(c->dp, c->n8[slot], true, sc[0], sc[1], c->rec[slot], fa, c->st);
    } else if (c->cfg.precision == SP_PREC_F32 && !c->dp.ext_rec) {
        e = sp::launch_pack_fast(c->dp, c->n8[slot], false, 1.0f, 1.0f, c->rec[slot], fa, c->st);
    } else {
        e = sp::launch_gradients(c->dp, c->n8[slot], c->rec[slot], fa, c->st);
    }
    if (e == hipSuccess && farray32)
        e = hipMemcpyAsync(farray32, fa, sizeof(float) * 32 * c->nnodes, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    free_dev(fa);
    if (e != hipSuccess) return fail(SP_ERR_HIP, std::string("field gradients failed: ") + hipGetErrorString(e));
    c->have_fields[slot] = true;
    return SP_OK;
}

int sp_fields_copy(void* ctx) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->cfg.time_interp) return fail(SP_ERR_STATE, "copy_fields needs time_interp");
    if (!c->have_fields[1]) return fail(SP_ERR_STATE, "frame 2 is not prepared");
    if (!c->rec[0]) SP_HIP(hipMalloc(&c->rec[0], sizeof(float) * c->recw * c->nrec));
    if (!c->n8[0]) SP_HIP(hipMalloc(&c->n8[0], sizeof(float) * 8 * c->nnodes));
    SP_HIP(hipMemcpyAsync(c->rec[0], c->rec[1], sizeof(float) * c->recw * c->nrec, hipMemcpyDeviceToDevice, c->st));
    SP_HIP(hipMemcpyAsync(c->n8[0], c->n8[1], sizeof(float) * 8 * c->nnodes, hipMemcpyDeviceToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->gscale[0] = c->gscale[1];
    c->vscale[0] = c->vscale[1];
    c->have_fields[0] = true;
    return SP_OK;
}

int sp_particles_upload(void* ctx, const sp_particles* h, int64_t n) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!h) return fail(SP_ERR_ARG, "particles is NULL");
    if (n < 0 || n > c->nmax) return fail(SP_ERR_CAPACITY, "n exceeds nptl_max");
    const sp::Ptl& P = c->P[c->cur];
    const void* src[17] = {h->x, h->y, h->z, h->p, h->v, h->mu, h->weight, h->t, h->dt,
                           h->split_times, h->count_flag, h->origin, h->nsteps_tracked,
                           h->nsteps_pushed, h->tag_injected, h->tag_splitted, h->rng_step};
    void* dst[17] = {P.x, P.y, P.z, P.p, P.v, P.mu, P.w, P.t, P.dt, P.split_times, P.count_flag,
                     P.origin, P.nsteps_tracked, P.nsteps_pushed, P.tag_injected, P.tag_splitted,
                     P.rng_step};
    const size_t sz[17] = {8, 8, 8, 8, 8, 8, 8, 8, 8, 1, 1, 4, 4, 4, 4, 4, 4};
    for (int k = 0; k < 17; ++k)
        if (!src[k] && n > 0) return fail(SP_ERR_ARG, "every particle array is required on upload");
    // weights are summed exactly as integers of 2^-32 (split weights 2^-k, k <= 32)
    for (int64_t i = 0; i < n; ++i) {
        const double w = h->weight[i] * sp::kFxScale;
        if (!(w >= 0.0 && w < 9.2e18) || w != std::floor(w))
            return fail(SP_ERR_ARG, "particle " + std::to_string(i) +
                                        ": weight must be a multiple of 2^-32 in [0, 2^31)");
        if (h->split_times[i] < 0 || h->split_times[i] > sp::kMaxSplitTimes)
            return fail(SP_ERR_ARG, "particle " + std::to_string(i) + ": split_times must be in [0, " +
                                        std::to_string(sp::kMaxSplitTimes) + "]");
    }
    for (int k = 0; k < 17 && n > 0; ++k)
        SP_HIP(hipMemcpyAsync(dst[k], src[k], sz[k] * (size_t)n, hipMemcpyHostToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->nptl = n;
    for (int64_t i = 0; i < n; ++i) c->tag_max = std::max(c->tag_max, h->tag_injected[i] + 1);
    return SP_OK;
}

int sp_particles_download(void* ctx, sp_particles* h, int64_t capacity, int64_t* n) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!h) return fail(SP_ERR_ARG, "particles is NULL");
    if (n) *n = c->nptl;
    if (capacity < c->nptl) return fail(SP_ERR_CAPACITY, "host capacity is smaller than nptl_current");
    const sp::Ptl& P = c->P[c->cur];
    void* dst[17] = {h->x, h->y, h->z, h->p, h->v, h->mu, h->weight, h->t, h->dt,
                     h->split_times, h->count_flag, h->origin, h->nsteps_tracked,
                     h->nsteps_pushed, h->tag_injected, h->tag_splitted, h->rng_step};
    const void* src[17] = {P.x, P.y, P.z, P.p, P.v, P.mu, P.w, P.t, P.dt, P.split_times,
                           P.count_flag, P.origin, P.nsteps_tracked, P.nsteps_pushed,
                           P.tag_injected, P.tag_splitted, P.rng_step};
    const size_t sz[17] = {8, 8, 8, 8, 8, 8, 8, 8, 8, 1, 1, 4, 4, 4, 4, 4, 4};
    for (int k = 0; k < 17 && c->nptl > 0; ++k)
        if (dst[k]) SP_HIP(hipMemcpyAsync(dst[k], src[k], sz[k] * (size_t)c->nptl, hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    return SP_OK;
}

int sp_particles_counts(void* ctx, int64_t* nptl_current, int64_t* nptl_split,
                        int64_t* nptl_escaped, int32_t* tag_max) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (nptl_current) *nptl_current = c->nptl;
    if (nptl_split) *nptl_split = c->nsplit;
    if (nptl_escaped) *nptl_escaped = c->nesc;
    if (tag_max) *tag_max = c->tag_max;
    return SP_OK;
}

int sp_inject_particles_spatial_uniform(void* ctx, int64_t nptl, double dt, int32_t dist_flag,
                                        double particle_v0, double t0, double dtf,
                                        const double part_box[6], double power_index) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (nptl < 0) return fail(SP_ERR_ARG, "nptl must be >= 0");
    if (dist_flag < 0 || dist_flag > 2) return fail(SP_ERR_ARG, "dist_flag must be 0, 1 or 2");
    if (!part_box) return fail(SP_ERR_ARG, "part_box is NULL");
    if ((long long)c->tag_max + nptl > 2147483647ll) return fail(SP_ERR_CAPACITY, "tag_max would overflow int32");
    sp::InjectArgs ia{};
    for (int i = 0; i < 6; ++i) ia.box[i] = part_box[i];
    ia.dt = dt; ia.v0 = particle_v0; ia.t0 = t0; ia.dtf = dtf; ia.power_index = power_index;
    ia.dist_flag = dist_flag;
    SP_HIP(sp::launch_inject(c->dp, c->P[c->cur], c->nptl, nptl, c->nmax, c->tag_max, ia, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->nptl = std::min<long long>(c->nptl + nptl, c->nmax);
    c->tag_max += (int)nptl;
    return SP_OK;
}

sp::ThreshArgs thresh_args(Ctx* c, int quantity, double threshold, const double part_box[6]) {
    sp::ThreshArgs ta{};
    ta.quantity = quantity;
    ta.threshold = threshold;
    for (int i = 0; i < 6; ++i) ta.box[i] = part_box[i];
    ta.idxh = 0.5 / c->cfg.dx;
    ta.idyh = 0.5 / c->cfg.dy;
    ta.idzh = 0.5 / c->cfg.dz;
    ta.n8 = c->n8[0];
    ta.max_trials = 1ll << 24;
    return ta;
}

int check_threshold_call(Ctx* c, int quantity, const double* part_box) {
    if (quantity == SP_INJECT_DB2)
        return fail(SP_ERR_UNSUPPORTED, "db2 injection needs the turbulence-amplitude map");
    if (quantity != SP_INJECT_JZ && quantity != SP_INJECT_ABSJ && quantity != SP_INJECT_DIVV &&
        quantity != SP_INJECT_RHO)
        return fail(SP_ERR_ARG, "quantity must be SP_INJECT_JZ, _ABSJ, _DIVV or _RHO");
    if (!part_box) return fail(SP_ERR_ARG, "part_box is NULL");
    if (!c->have_fields[0] || !c->n8[0]) return fail(SP_ERR_STATE, "fields (slot 0) are not uploaded");
    return SP_OK;
}

long long ncells_above(Ctx* c, const sp::ThreshArgs& ta, int& rc) {
    rc = SP_OK;
    unsigned long long* d = c->totals;
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), c->st);
    if (e == hipSuccess) e = sp::launch_threshold_count(c->dp, ta, d, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) rc = fail(SP_ERR_HIP, std::string("cell count failed: ") + hipGetErrorString(e));
    return (long long)h;
}

int sp_ncells_above_threshold(void* ctx, int32_t quantity, double threshold, const double part_box[6],
                              int64_t* ncells) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if ((rc = check_threshold_call(c, quantity, part_box)) != SP_OK) return rc;
    if (!ncells) return fail(SP_ERR_ARG, "ncells is NULL");
    const long long n = ncells_above(c, thresh_args(c, quantity, threshold, part_box), rc);
    if (rc != SP_OK) return rc;
    *ncells = n;
    return SP_OK;
}

int sp_inject_particles_threshold(void* ctx, int32_t quantity, int64_t nptl, double dt,
                                  int32_t dist_flag, double particle_v0, double t0, double dtf,
                                  int32_t inject_same_nptl, double threshold, int64_t ncells_norm,
                                  const double part_box[6], double power_index,
                                  int64_t* nptl_injected) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if ((rc = check_threshold_call(c, quantity, part_box)) != SP_OK) return rc;
    if (nptl < 0) return fail(SP_ERR_ARG, "nptl must be >= 0");
    if (dist_flag < 0 || dist_flag > 2) return fail(SP_ERR_ARG, "dist_flag must be 0, 1 or 2");
    if (!inject_same_nptl && ncells_norm < 1) return fail(SP_ERR_ARG, "ncells_norm must be >= 1");
    sp::ThreshArgs ta = thresh_args(c, quantity, threshold, part_box);
    const long long ncells = ncells_above(c, ta, rc);
    if (rc != SP_OK) return rc;
    // nptl_inject = int(nptl * mpi_sub_size * (dble(ncells) / dble(ncells_g | norm)))
    long long ninj = 0;
    if (ncells > 0) {
        const double ratio = inject_same_nptl ? (double)ncells / (double)ncells
                                              : (double)ncells / (double)ncells_norm;
        const double v = (double)nptl * ratio;
        if (!(v < 9.0e18)) return fail(SP_ERR_ARG, "nptl_inject overflows");
        ninj = (long long)v;
    }
    if (nptl_injected) *nptl_injected = ninj;
    if (ninj == 0) return SP_OK;
    if ((long long)c->tag_max + ninj > 2147483647ll) return fail(SP_ERR_CAPACITY, "tag_max would overflow int32");
    sp::InjectArgs ia{};
    for (int i = 0; i < 6; ++i) ia.box[i] = part_box[i];
    ia.dt = dt; ia.v0 = particle_v0; ia.t0 = t0; ia.dtf = dtf; ia.power_index = power_index;
    ia.dist_flag = dist_flag;
    unsigned long long* d = c->totals;
    // pass 1 only draws; the particles are stored once every one has a position
    for (int commit = 0; commit < 2; ++commit) {
        ta.commit = commit;
        unsigned long long fails = 0;
        SP_HIP(hipMemsetAsync(d, 0, sizeof(fails), c->st));
        SP_HIP(sp::launch_threshold_inject(c->dp, c->P[c->cur], c->nptl, ninj, c->nmax, c->tag_max, ia,
                                           ta, d, c->st));
        SP_HIP(hipMemcpyAsync(&fails, d, sizeof(fails), hipMemcpyDeviceToHost, c->st));
        SP_HIP(hipStreamSynchronize(c->st));
        if (fails)
            return fail(SP_ERR_ARG, std::to_string(fails) + " particles found no point above the threshold in " +
                                    std::to_string(ta.max_trials) + " draws");
    }
    c->nptl = std::min<long long>(c->nptl + ninj, c->nmax);
    c->tag_max += (int)ninj;
    return SP_OK;
}

// The push of one interval: launches of the persistent mover over the
// current particles, with re-sorts between launches when a step budget is set.
// bpc > 0 / sort_steps >= 0 override the launch settings (autotuner trials).
int mover_run(Ctx* c, double t0, double dtf, int num_fine_steps, int nsteps_interval, int bpc,
              int sort_steps, int ppl_trial, unsigned long long& steps, float& ms_push) {
    const int ppl = mover_ppl(c, ppl_trial);
    int rc = SP_OK;
    const bool ti = c->cfg.time_interp != 0;
    const int fx = step_features(c->cfg);
    // With sort_steps, the interval runs as launches of at most sort_steps
    // steps per particle; between launches the paused particles are sorted
    // to the front in spatial order and the next launch takes only them.
    // Once fewer than 1/8 of the particles are still moving, the last launch
    // runs them to the end of the interval: re-sorting a sparse remainder
    // costs more than its locality gains.
    // Budget precedence, last wins: the configuration's sort_steps, the
    // autotuner's choice, SP_SORT_STEPS, an explicit trial value.
    unsigned budget = (unsigned)c->cfg.sort_steps;
    if (c->tuned_sort_steps >= 0) budget = (unsigned)c->tuned_sort_steps;
    if (const char* e = std::getenv("SP_SORT_STEPS")) budget = (unsigned)std::max(0, std::atoi(e));
    if (sort_steps >= 0) budget = (unsigned)sort_steps;
    if (c->cfg.sort_every <= 0) budget = 0u;
    steps = 0;
    ms_push = 0.f;
    long long nrun = c->nptl;
    if (nrun > 0) SP_HIP(hipMemsetAsync(c->P[c->cur].rdt, 0, sizeof(double) * nrun, c->st));
    while (nrun > 0) {
        sp::MoverParams m = mover_params(c, t0, dtf, num_fine_steps, nsteps_interval);
        m.n = nrun;
        m.part_len = (nrun + sp::kParts - 1) / sp::kParts;
        m.step_budget = nrun * 8 < c->nptl ? 0u : budget;
        // clear the work, step and pause counters (leak counters are cumulative)
        SP_HIP(hipMemsetAsync(c->ctr, 0, offsetof(sp::Counters, leak_fx), c->st));
        SP_HIP(hipEventRecord(c->ev[1], c->st));
        const int grid = mover_grid(c, nrun, bpc, ppl);
        if (c->cfg.precision == SP_PREC_F32H) SP_HIP(sp::launch_mover_f32h(m, ti, fx, c->cfg.ndim, ppl, grid, c->st));
        else if (c->cfg.precision == SP_PREC_F32) SP_HIP(sp::launch_mover_f32(m, ti, fx, c->cfg.ndim, ppl, grid, c->st));
        else SP_HIP(sp::launch_mover_f64(m, ti, fx, c->cfg.ndim, grid, c->st));
        SP_HIP(hipEventRecord(c->ev[3], c->st));
        unsigned long long sp_[2] = {0, 0};
        SP_HIP(hipMemcpyAsync(sp_, &c->ctr->steps, sizeof(sp_), hipMemcpyDeviceToHost, c->st));
        SP_HIP(hipStreamSynchronize(c->st));
        float ms = 0.f;
        SP_HIP(hipEventElapsedTime(&ms, c->ev[1], c->ev[3]));
        ms_push += ms;
        steps += sp_[0];
        const long long paused = (long long)sp_[1];
        if (paused == 0) break;
        if (paused > nrun) return fail(SP_ERR_STATE, "mover pause count exceeds the particles run");
        rc = spatial_sort(c, true, nrun);
        if (rc != SP_OK) return rc;
        nrun = paused;
    }
    return SP_OK;
}

int sp_particle_mover(void* ctx, double t0, double dtf, int32_t num_fine_steps,
                      int32_t nsteps_interval, sp_mover_stats* stats) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->have_fields[0]) return fail(SP_ERR_STATE, "fields (slot 0) are not prepared");
    if (c->cfg.time_interp && !c->have_fields[1]) return fail(SP_ERR_STATE, "time_interp needs slot 1 prepared");
    if (num_fine_steps < 1 || nsteps_interval < 1) return fail(SP_ERR_ARG, "num_fine_steps and nsteps_interval must be >= 1");
    if (!(dtf > 0)) return fail(SP_ERR_ARG, "dtf must be > 0");
    const bool sorting = c->cfg.sort_every > 0;
    SP_HIP(hipEventRecord(c->ev[0], c->st));
    if (sorting && c->mover_calls % c->cfg.sort_every == 0) {
        rc = spatial_sort(c);
        if (rc != SP_OK) return rc;
    }
    ++c->mover_calls;
    unsigned long long steps = 0;
    float ms_push = 0.f;
    rc = mover_run(c, t0, dtf, num_fine_steps, nsteps_interval, 0, -1, 0, steps, ms_push);
    if (rc != SP_OK) return rc;
    // remove_particles: stable compaction into the other buffer set
    long long kept = 0, esc = 0;
    if (c->nptl > 0) {
        const long long nt = sp::num_tiles(c->nptl);
        const sp::SplitArgs none{};
        SP_HIP(sp::launch_tile_count(c->P[c->cur], c->nptl, 0, none, c->tile_cnt, c->tile_esc,
                                     c->ea.on ? &c->da : nullptr, &c->ea, c->st));
        SP_HIP(sp::launch_tile_scan(c->tile_cnt, c->tile_esc, nt, c->tile_off, c->totals, c->st));
        SP_HIP(sp::launch_compact_scatter(c->P[c->cur], c->P[1 - c->cur], c->nptl, c->tile_off, c->st));
    }
    SP_HIP(hipEventRecord(c->ev[2], c->st));
    sp::Counters hc{};
    SP_HIP(hipMemcpyAsync(&hc, c->ctr, sizeof(hc), hipMemcpyDeviceToHost, c->st));
    // the leak counters of this call move to the host totals (a u64 of 2^-32
    // units would wrap after a cumulative weight of 2^32)
    SP_HIP(hipMemsetAsync(&c->ctr->leak_fx, 0, 2 * sizeof(unsigned long long), c->st));
    unsigned long long tot[2] = {0, 0};
    if (c->nptl > 0) SP_HIP(hipMemcpyAsync(tot, c->totals, sizeof(tot), hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    if (c->nptl > 0) {
        kept = (long long)tot[0];
        esc = (long long)tot[1];
        c->cur = 1 - c->cur;
        c->nptl = kept;
        c->nesc += esc;
    }
    c->leak_acc += hc.leak_fx;
    c->negp_acc += hc.negp_fx;
    float ms_all = 0.f;
    SP_HIP(hipEventElapsedTime(&ms_all, c->ev[0], c->ev[2]));
    if (stats) {
        stats->steps = (int64_t)steps;
        stats->nptl_current = c->nptl;
        stats->nptl_escaped = c->nesc;
        stats->leak = fx_weight(c->leak_acc);
        stats->leak_negp = fx_weight(c->negp_acc);
        stats->kernel_ms = ms_push;
        stats->total_ms = ms_all;
    }
    return SP_OK;
}

int sp_split_particle(void* ctx, double split_ratio, double pmin_split) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (c->nptl == 0) return SP_OK;
    if (!(split_ratio > 1.0)) return fail(SP_ERR_ARG, "split_ratio must be > 1");
    sp::SplitArgs sa{split_ratio, pmin_split, c->cfg.p0, c->cfg.pmax};
    const long long nt = sp::num_tiles(c->nptl);
    // pass 1 counts the candidates and those already split kMaxSplitTimes times
    SP_HIP(sp::launch_tile_count(c->P[c->cur], c->nptl, 1, sa, c->tile_cnt, c->tile_esc, nullptr,
                                 nullptr, c->st));
    SP_HIP(sp::launch_tile_scan(c->tile_cnt, c->tile_esc, nt, c->tile_off, c->totals, c->st));
    unsigned long long tot[2] = {0, 0};
    SP_HIP(hipMemcpyAsync(tot, c->totals, sizeof(tot), hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    if (tot[1] > 0)
        return fail(SP_ERR_UNSUPPORTED,
                    std::to_string(tot[1]) + " split candidates have split_times = " +
                        std::to_string(sp::kMaxSplitTimes) + ": their copies would weigh 2^-" +
                        std::to_string(sp::kMaxSplitTimes + 1) +
                        ", below the 2^-32 resolution of the exact weight sums (nothing was split)");
    SP_HIP(sp::launch_split_apply(c->P[c->cur], c->nptl, c->nmax, sa, c->tile_off, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    const long long ncand = (long long)tot[0];
    const long long room = c->nmax - c->nptl;
    const long long done = std::min(ncand, room);
    c->nsplit += done;
    c->nptl += done;
    return SP_OK;
}

int sp_sort_particles(void* ctx) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    rc = spatial_sort(c);
    if (rc != SP_OK) return rc;
    SP_HIP(hipStreamSynchronize(c->st));
    return SP_OK;
}

int sp_step_once(void* ctx, int64_t n, double t0, double dtf, double* state, const double* rands,
                 const int8_t* fixed_dt, double* deltas) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (n < 0) return fail(SP_ERR_ARG, "n must be >= 0");
    if (n == 0) return SP_OK;
    if (!state || !rands || !fixed_dt || !deltas) return fail(SP_ERR_ARG, "NULL buffer");
    if (!c->have_fields[0]) return fail(SP_ERR_STATE, "fields (slot 0) are not prepared");
    if (c->cfg.time_interp && !c->have_fields[1]) return fail(SP_ERR_STATE, "time_interp needs slot 1 prepared");
    if (!(dtf > 0)) return fail(SP_ERR_ARG, "dtf must be > 0");
    if (c->cfg.nlgc && n > c->nptl)
        return fail(SP_ERR_ARG, "with nlgc, row i takes mu of particle i: need n <= nptl_current");
    double *ds = nullptr, *dr = nullptr, *dd = nullptr;
    int8_t* df = nullptr;
    hipError_t e = hipMalloc(&ds, sizeof(double) * 6 * n);
    if (e == hipSuccess) e = hipMalloc(&dr, sizeof(double) * 4 * n);
    if (e == hipSuccess) e = hipMalloc(&dd, sizeof(double) * 4 * n);
    if (e == hipSuccess) e = hipMalloc(&df, (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, state, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rands, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(df, fixed_dt, (size_t)n, hipMemcpyHostToDevice, c->st);
    sp::MoverParams m = mover_params(c, t0, dtf, 1, 1);
    m.n = n;
    const bool ti = c->cfg.time_interp != 0;
    const int fx = step_features(c->cfg);
    if (e == hipSuccess) {
        if (c->cfg.precision == SP_PREC_F32H)
            e = sp::launch_step_once_f32h(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st);
        else if (c->cfg.precision == SP_PREC_F32)
            e = sp::launch_step_once_f32(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st);
        else
            e = sp::launch_step_once_f64(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(state, ds, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(deltas, dd, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    free_dev(ds);
    free_dev(dr);
    free_dev(dd);
    free_dev(df);
    if (e != hipSuccess) return fail(SP_ERR_HIP, std::string("step_once failed: ") + hipGetErrorString(e));
    return SP_OK;
}

int sp_step_once_noalloc(void* ctx, int64_t n, double t0, double dtf, double* state,
                         const double* rands, const int8_t* fixed_dt, double* deltas) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (n < 0) return fail(SP_ERR_ARG, "n must be >= 0");
    if (n > kNoAllocRows)
        return fail(SP_ERR_CAPACITY, "sp_step_once_noalloc takes at most " + std::to_string(kNoAllocRows) +
                                         " rows per call");
    if (n == 0) return SP_OK;
    if (!state || !rands || !fixed_dt || !deltas) return fail(SP_ERR_ARG, "NULL buffer");
    if (!c->have_fields[0]) return fail(SP_ERR_STATE, "fields (slot 0) are not prepared");
    if (c->cfg.time_interp && !c->have_fields[1]) return fail(SP_ERR_STATE, "time_interp needs slot 1 prepared");
    if (!(dtf > 0)) return fail(SP_ERR_ARG, "dtf must be > 0");
    if (c->cfg.nlgc && n > c->nptl)
        return fail(SP_ERR_ARG, "with nlgc, row i takes mu of particle i: need n <= nptl_current");
    char* base = static_cast<char*>(device_scratch(c->device));
    if (!base) return fail(SP_ERR_STATE, "the staging memory of this GPU was not set up");
    // regions sized for kNoAllocRows rows; every byte a call reads is written
    // by that call first (copies in, then the kernel's deltas)
    double* ds = reinterpret_cast<double*>(base);
    double* dr = ds + 6 * kNoAllocRows;
    double* dd = dr + 4 * kNoAllocRows;
    int8_t* df = reinterpret_cast<int8_t*>(dd + 4 * kNoAllocRows);
    SP_HIP(hipMemcpyAsync(ds, state, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->st));
    SP_HIP(hipMemcpyAsync(dr, rands, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->st));
    SP_HIP(hipMemcpyAsync(df, fixed_dt, (size_t)n, hipMemcpyHostToDevice, c->st));
    sp::MoverParams m = mover_params(c, t0, dtf, 1, 1);
    m.n = n;
    const bool ti = c->cfg.time_interp != 0;
    const int fx = step_features(c->cfg);
    if (c->cfg.precision == SP_PREC_F32H)
        SP_HIP(sp::launch_step_once_f32h(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st));
    else if (c->cfg.precision == SP_PREC_F32)
        SP_HIP(sp::launch_step_once_f32(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st));
    else
        SP_HIP(sp::launch_step_once_f64(m, ti, fx, c->cfg.ndim, ds, dr, df, dd, c->st));
    SP_HIP(hipMemcpyAsync(state, ds, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipMemcpyAsync(deltas, dd, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    return SP_OK;
}

int sp_distributions_init(void* ctx, const sp_dist_config* dc) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!dc) return fail(SP_ERR_ARG, "dist config is NULL");
    if (dc->npp_global < 1 || dc->nmu_global < 1) return fail(SP_ERR_ARG, "npp_global and nmu_global must be >= 1");
    if ((long long)dc->npp_global * dc->nmu_global > 8192)
        return fail(SP_ERR_UNSUPPORTED, "npp_global * nmu_global is limited to 8192 bins");
    sp::DistArgs& d = c->da;
    d = sp::DistArgs{};
    const sp_config& cf = c->cfg;
    d.npp = dc->npp_global;
    d.nmu = dc->nmu_global;
    d.local_dist = dc->local_dist;
    d.pmin = cf.pmin;
    d.pmax = cf.pmax;
    // init_particle_distributions (diagnostics.f90:196-207)
    d.pmin_log = std::log10(cf.pmin);
    d.dp_log = (std::log10(cf.pmax) - d.pmin_log) / dc->npp_global;
    d.dmu = (double)(2.0f / (float)dc->nmu_global);
    d.xmin = cf.xmin; d.ymin = cf.ymin; d.zmin = cf.zmin;
    for (auto*& q : c->dloc) { free_dev(q); q = nullptr; }
    for (auto*& q : c->desc) { free_dev(q); q = nullptr; }
    c->ea = sp::EscArgs{};
    std::memset(c->esc_size, 0, sizeof(c->esc_size));
    c->dist_ready = false;
    free_dev(c->dglob);
    c->dglob = nullptr;
    c->nglob = (long long)d.npp * d.nmu;
    SP_HIP(hipMalloc(&c->dglob, sizeof(unsigned long long) * c->nglob));
    for (int k = 0; k < 4; ++k) {
        const sp_local_dist& L = dc->local[k];
        sp::LocalArgs& A = d.loc[k];
        c->loc_size[k] = 0;
        std::memset(c->loc_dims[k], 0, sizeof(c->loc_dims[k]));
        A.enabled = dc->local_dist && L.enabled;
        if (!A.enabled) continue;
        if (L.npbins < 1 || L.nmu < 1 || L.rx < 1 || L.ry < 1 || L.rz < 1 || !(L.pmin > 0) || !(L.pmax > L.pmin))
            return fail(SP_ERR_ARG, "invalid local distribution set " + std::to_string(k + 1));
        // init_local_particle_distributions (diagnostics.f90:263-284) and
        // read_diagnostics_params (:2121-2126)
        A.npbins = L.npbins;
        A.nmu = L.nmu;
        A.nr[0] = (cf.nx + L.rx - 1) / L.rx;
        A.nr[1] = (cf.ny + L.ry - 1) / L.ry;
        A.nr[2] = (cf.nz + L.rz - 1) / L.rz;
        A.dd[0] = cf.lx / (double)A.nr[0];
        A.dd[1] = cf.ly / (double)A.nr[1];
        A.dd[2] = cf.lz / (double)A.nr[2];
        A.pmin_log = std::log10(L.pmin);
        A.dp_log = (std::log10(L.pmax) - A.pmin_log) / L.npbins;
        A.dmu = (double)(2.0f / (float)L.nmu);
        c->loc_dims[k][0] = L.nmu;
        c->loc_dims[k][1] = L.npbins;
        c->loc_dims[k][2] = A.nr[0];
        c->loc_dims[k][3] = A.nr[1];
        c->loc_dims[k][4] = A.nr[2];
        c->loc_size[k] = (long long)L.nmu * L.npbins * A.nr[0] * A.nr[1] * A.nr[2];
        SP_HIP(hipMalloc(&c->dloc[k], sizeof(unsigned long long) * c->loc_size[k]));
    }
    if (dc->escaped_dist) {
        // init_particle_distributions (diagnostics.f90:188-191) and
        // init_local_escaped_distributions (:359-395)
        const int nd = cf.ndim;
        c->esc_size[0] = c->nglob * 2 * nd;
        for (int k = 0; k < 4; ++k) {
            const sp::LocalArgs& A = d.loc[k];
            if (!A.enabled) continue;
            const long long per = (long long)A.nmu * A.npbins * 2;
            c->esc_size[k + 1] = per * (A.nr[1] * A.nr[2] + A.nr[0] * A.nr[2] + (nd > 2 ? A.nr[0] * A.nr[1] : 0));
        }
        for (int j = 0; j < 5; ++j) {
            if (c->esc_size[j] == 0) continue;
            SP_HIP(hipMalloc(&c->desc[j], sizeof(unsigned long long) * c->esc_size[j]));
            SP_HIP(hipMemsetAsync(c->desc[j], 0, sizeof(unsigned long long) * c->esc_size[j], c->st));
        }
        SP_HIP(hipStreamSynchronize(c->st));
        sp::EscArgs& e = c->ea;
        e.on = 1;
        e.ndim = nd;
        e.glob = c->desc[0];
        for (int k = 0; k < 4; ++k) {
            const sp::LocalArgs& A = d.loc[k];
            unsigned long long* f = c->desc[k + 1];
            if (!f) continue;
            const long long per = (long long)A.nmu * A.npbins * 2;
            e.face[k][0] = f;
            e.face[k][1] = f + per * A.nr[1] * A.nr[2];
            if (nd > 2) e.face[k][2] = e.face[k][1] + per * A.nr[0] * A.nr[2];
        }
    }
    c->dist_ready = true;
    return SP_OK;
}

int sp_escaped_distributions(void* ctx, double* fescaped, double* fesc1, double* fesc2,
                             double* fesc3, double* fesc4) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->dist_ready || !c->ea.on)
        return fail(SP_ERR_STATE, "escaped distributions are not initialised (escaped_dist = 0)");
    double* outs[5] = {fescaped, fesc1, fesc2, fesc3, fesc4};
    for (int j = 0; j < 5; ++j) {
        if (!c->desc[j]) continue;
        const size_t bytes = sizeof(unsigned long long) * (size_t)c->esc_size[j];
        if (outs[j]) {
            std::vector<unsigned long long> h((size_t)c->esc_size[j]);
            SP_HIP(hipMemcpyAsync(h.data(), c->desc[j], bytes, hipMemcpyDeviceToHost, c->st));
            SP_HIP(hipStreamSynchronize(c->st));
            for (long long i = 0; i < c->esc_size[j]; ++i) outs[j][i] = (double)h[(size_t)i] * kInvFx;
        }
        SP_HIP(hipMemsetAsync(c->desc[j], 0, bytes, c->st));
    }
    SP_HIP(hipStreamSynchronize(c->st));
    return SP_OK;
}

int sp_local_dist_dims(void* ctx, int32_t set, int64_t dims[5]) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->dist_ready) return fail(SP_ERR_STATE, "distributions are not initialised");
    if (set < 1 || set > 4 || !dims) return fail(SP_ERR_ARG, "set must be 1..4");
    for (int i = 0; i < 5; ++i) dims[i] = c->loc_dims[set - 1][i];
    return SP_OK;
}

int sp_distributions(void* ctx, double* fglobal, double* flocal1, double* flocal2, double* flocal3,
                     double* flocal4) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->dist_ready) return fail(SP_ERR_STATE, "distributions are not initialised");
    SP_HIP(hipMemsetAsync(c->dglob, 0, sizeof(unsigned long long) * c->nglob, c->st));
    unsigned long long* gl[4];
    for (int k = 0; k < 4; ++k) {
        gl[k] = c->dloc[k];
        if (c->dloc[k]) SP_HIP(hipMemsetAsync(c->dloc[k], 0, sizeof(unsigned long long) * c->loc_size[k], c->st));
    }
    const int grid = (int)std::max<long long>(1, std::min<long long>((c->nptl + 255) / 256, (long long)c->ncu * 8));
    SP_HIP(sp::launch_dist(c->P[c->cur], c->nptl, c->da, c->dglob, gl, c->loc_size, grid, c->st));
    std::vector<unsigned long long> hb((size_t)c->nglob);
    SP_HIP(hipMemcpyAsync(hb.data(), c->dglob, sizeof(unsigned long long) * c->nglob, hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    if (fglobal)
        for (long long i = 0; i < c->nglob; ++i) fglobal[i] = (double)hb[i] * kInvFx;
    double* outs[4] = {flocal1, flocal2, flocal3, flocal4};
    for (int k = 0; k < 4; ++k) {
        if (!outs[k] || !c->dloc[k]) continue;
        std::vector<unsigned long long> hl((size_t)c->loc_size[k]);
        SP_HIP(hipMemcpy(hl.data(), c->dloc[k], sizeof(unsigned long long) * c->loc_size[k], hipMemcpyDeviceToHost));
        for (long long i = 0; i < c->loc_size[k]; ++i) outs[k][i] = (double)hl[i] * kInvFx;
    }
    return SP_OK;
}

int sp_autotune_mover(void* ctx, double t0, double dtf, int32_t num_fine_steps,
                      int32_t nsteps_interval, sp_tune_result* out) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->have_fields[0]) return fail(SP_ERR_STATE, "fields (slot 0) are not prepared");
    if (c->cfg.time_interp && !c->have_fields[1]) return fail(SP_ERR_STATE, "time_interp needs slot 1 prepared");
    if (num_fine_steps < 1 || nsteps_interval < 1) return fail(SP_ERR_ARG, "num_fine_steps and nsteps_interval must be >= 1");
    if (!(dtf > 0)) return fail(SP_ERR_ARG, "dtf must be > 0");
    if (c->nptl < 1) return fail(SP_ERR_STATE, "no particles to tune on");
    const long long cap = particle_cap(c->nmax);
    const size_t bytes = (size_t)sp::kPtlBytes * (size_t)cap;
    if (!c->pbuf_save) SP_HIP(hipMalloc(&c->pbuf_save, bytes));
    // trials start from the order the next mover call would push
    if (c->cfg.sort_every > 0) {
        rc = spatial_sort(c);
        if (rc != SP_OK) return rc;
    }
    const int cur0 = c->cur;
    sp::Counters saved{};
    SP_HIP(hipMemcpyAsync(&saved, c->ctr, sizeof(saved), hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipMemcpyAsync(c->pbuf_save, c->pbuf[cur0], bytes, hipMemcpyDeviceToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    const int prec = c->cfg.precision;
    std::vector<int> budgets = {0};
    if (c->cfg.sort_every > 0) budgets = {0, 100, 200, 400};
    // (particles per lane, blocks per CU, step budget) triples
    struct Trial { int ppl, bpc, ss; };
    std::vector<Trial> trials;
    for (int ppl = 1; ppl <= mover_ppl(c, 2); ++ppl) {
        const int max_bpc = std::max(1, std::min(4, c->mover_bpc[prec][ppl - 1]));
        for (int bpc = 1; bpc <= max_bpc; ++bpc)
            for (int ss : budgets) trials.push_back(Trial{ppl, bpc, ss});
    }
    const int saved_bpc = c->tuned_bpc, saved_ss = c->tuned_sort_steps, saved_ppl = c->tuned_ppl;
    c->tuned_bpc = 0;
    c->tuned_sort_steps = -1;
    c->tuned_ppl = 0;
    double best = -1.0, dflt = -1.0;
    int best_bpc = 1, best_ss = 0, best_ppl = 1, ntrials = 0;
    for (int t = -1; t < (int)trials.size(); ++t) {
        // trial -1: the untuned defaults; then every triple
        const Trial tr = t < 0 ? Trial{0, 0, -1} : trials[(size_t)t];
        if (t >= 0) {
            SP_HIP(hipMemcpyAsync(c->pbuf[cur0], c->pbuf_save, bytes, hipMemcpyDeviceToDevice, c->st));
            c->cur = cur0;
        }
        unsigned long long steps = 0;
        float ms = 0.f;
        rc = mover_run(c, t0, dtf, num_fine_steps, nsteps_interval, tr.bpc, tr.ss, tr.ppl, steps, ms);
        if (rc != SP_OK) break;
        const double rate = ms > 0.f ? (double)steps / (1e-3 * (double)ms) : 0.0;
        if (t < 0) {
            dflt = rate;
            continue;
        }
        ++ntrials;
        if (rate > best) {
            best = rate;
            best_bpc = tr.bpc;
            best_ss = tr.ss;
            best_ppl = tr.ppl;
        }
    }
    // restore the particles and counters whatever the trials did
    SP_HIP(hipMemcpyAsync(c->pbuf[cur0], c->pbuf_save, bytes, hipMemcpyDeviceToDevice, c->st));
    SP_HIP(hipMemcpyAsync(c->ctr, &saved, sizeof(saved), hipMemcpyHostToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->cur = cur0;
    if (rc != SP_OK) {
        c->tuned_bpc = saved_bpc;
        c->tuned_sort_steps = saved_ss;
        c->tuned_ppl = saved_ppl;
        return rc;
    }
    c->tuned_bpc = best_bpc;
    c->tuned_sort_steps = c->cfg.sort_every > 0 ? best_ss : -1;
    c->tuned_ppl = best_ppl;
    if (out) {
        out->blocks_per_cu = best_bpc;
        out->sort_steps = best_ss;
        out->ntrials = ntrials;
        out->particles_per_lane = best_ppl;
        out->steps_per_s = best;
        out->default_steps_per_s = dflt;
    }
    return SP_OK;
}

int sp_set_variants(void* ctx, int32_t nvar, const sp_variant* variants) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (nvar < 0 || nvar > (1 << 24)) return fail(SP_ERR_ARG, "nvar must be in [0, 2^24]");
    if (nvar > 0 && !variants) return fail(SP_ERR_ARG, "variants is NULL");
    std::vector<sp::KappaVar> hv((size_t)nvar);
    for (int v = 0; v < nvar; ++v) {
        const sp_variant& q = variants[v];
        if (!(q.kpara0 > 0)) return fail(SP_ERR_ARG, "variant " + std::to_string(v) + ": kpara0 must be > 0");
        if (!(q.kret >= 0) || q.kret > 1) return fail(SP_ERR_ARG, "variant " + std::to_string(v) + ": kret must be in [0, 1]");
        if (!std::isfinite(q.gamma_turb)) return fail(SP_ERR_ARG, "variant " + std::to_string(v) + ": gamma_turb must be finite");
        if (q.reserved0 != 0) return fail(SP_ERR_ARG, "variant " + std::to_string(v) + ": reserved0 must be 0");
        hv[(size_t)v] = kappa_var(c->cfg.p0, q.kpara0, q.kret, q.gamma_turb);
    }
    if (nvar > c->nvar_cap) {
        free_dev(c->dvar);
        c->dvar = nullptr;
        c->nvar_cap = 0;
        SP_HIP(hipMalloc(&c->dvar, sizeof(sp::KappaVar) * (size_t)nvar));
        c->nvar_cap = nvar;
    }
    if (nvar > 0)
        SP_HIP(hipMemcpyAsync(c->dvar, hv.data(), sizeof(sp::KappaVar) * (size_t)nvar,
                              hipMemcpyHostToDevice, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    c->nvar = nvar;
    return SP_OK;
}

int sp_inject_variants(void* ctx, int32_t nvar, const int64_t* nptl, double dt, int32_t dist_flag,
                       double particle_v0, double t0, double dtf, const double* part_boxes,
                       double power_index) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (nvar < 1 || nvar > c->nvar) return fail(SP_ERR_ARG, "nvar must be in [1, number of installed variants]");
    if (!nptl || !part_boxes) return fail(SP_ERR_ARG, "nptl or part_boxes is NULL");
    if (dist_flag < 0 || dist_flag > 2) return fail(SP_ERR_ARG, "dist_flag must be 0, 1 or 2");
    long long total = 0;
    for (int v = 0; v < nvar; ++v) {
        if (nptl[v] < 0) return fail(SP_ERR_ARG, "nptl must be >= 0");
        total += nptl[v];
    }
    if ((long long)c->tag_max + total > 2147483647ll) return fail(SP_ERR_CAPACITY, "tag_max would overflow int32");
    if (c->nptl + total > c->nmax) return fail(SP_ERR_CAPACITY, "the variants' particles exceed nptl_max");
    for (int v = 0; v < nvar; ++v) {
        if (nptl[v] == 0) continue;
        sp::InjectArgs ia{};
        for (int i = 0; i < 6; ++i) ia.box[i] = part_boxes[6 * v + i];
        ia.dt = dt; ia.v0 = particle_v0; ia.t0 = t0; ia.dtf = dtf; ia.power_index = power_index;
        ia.dist_flag = dist_flag;
        ia.origin = v;
        SP_HIP(sp::launch_inject(c->dp, c->P[c->cur], c->nptl, nptl[v], c->nmax, c->tag_max, ia, c->st));
        c->nptl += nptl[v];
        c->tag_max += (int)nptl[v];
    }
    SP_HIP(hipStreamSynchronize(c->st));
    return SP_OK;
}

int sp_variant_distributions(void* ctx, int32_t nvar, double* fglobal, int64_t* counts) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!c->dist_ready) return fail(SP_ERR_STATE, "distributions are not initialised");
    if (nvar < 1 || nvar > c->nvar) return fail(SP_ERR_ARG, "nvar must be in [1, number of installed variants]");
    const long long npp = c->da.npp;
    const size_t nbins = (size_t)npp * (size_t)nvar;
    if (c->vhist_len < nbins + (size_t)nvar) {
        free_dev(c->vhist);
        c->vhist = nullptr;
        c->vhist_len = 0;
        SP_HIP(hipMalloc(&c->vhist, sizeof(unsigned long long) * (nbins + (size_t)nvar)));
        c->vhist_len = nbins + (size_t)nvar;
    }
    unsigned long long* hist = c->vhist;
    unsigned long long* cnt = c->vhist + nbins;
    SP_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * (nbins + (size_t)nvar), c->st));
    SP_HIP(sp::launch_variant_dist(c->P[c->cur], c->nptl, c->da, nvar, hist, cnt, c->ncu, c->st));
    std::vector<unsigned long long> hb(nbins + (size_t)nvar);
    SP_HIP(hipMemcpyAsync(hb.data(), hist, sizeof(unsigned long long) * hb.size(), hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    if (fglobal)
        for (size_t i = 0; i < nbins; ++i) fglobal[i] = (double)hb[i] * kInvFx;
    if (counts)
        for (int v = 0; v < nvar; ++v) counts[v] = (int64_t)hb[nbins + (size_t)v];
    return SP_OK;
}

int sp_calibration_copy(void* ctx, int64_t bytes, double* ms) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (bytes < 16 || bytes % 16 != 0) return fail(SP_ERR_ARG, "bytes must be a positive multiple of 16");
    void *a = nullptr, *b = nullptr;
    hipError_t e = hipMalloc(&a, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, (size_t)bytes, c->st);
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->st);
    if (e == hipSuccess) e = sp::launch_copy(a, b, bytes / 16, c->st);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
    free_dev(a);
    free_dev(b);
    if (e != hipSuccess) return fail(SP_ERR_HIP, std::string("calibration copy failed: ") + hipGetErrorString(e));
    if (ms) *ms = t;
    return SP_OK;
}

int sp_quick_check(void* ctx, double out[9]) {
    int rc;
    Ctx* c = live_ctx(ctx, rc);
    if (!c) return rc;
    if (!out) return fail(SP_ERR_ARG, "out is NULL");
    const int grid = (int)std::max<long long>(1, std::min<long long>((c->nptl + 255) / 256, c->stats_grid));
    SP_HIP(sp::launch_stats(c->P[c->cur], c->nptl, c->stats_part, grid, c->st));
    std::vector<double> part((size_t)5 * grid);
    sp::Counters hc{};
    SP_HIP(hipMemcpyAsync(part.data(), c->stats_part, sizeof(double) * 5 * grid, hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipMemcpyAsync(&hc, c->ctr, sizeof(hc), hipMemcpyDeviceToHost, c->st));
    SP_HIP(hipStreamSynchronize(c->st));
    unsigned long long W = 0;
    double S = 0.0, mn = 1.0, mx = 0.0, pm = 0.0;  // quick_check starts from 1 and 0
    for (int b = 0; b < grid; ++b) {
        unsigned long long w;
        std::memcpy(&w, &part[5 * b], sizeof(w));
        W += w;
        S += part[5 * b + 1];
        mn = std::min(mn, part[5 * b + 2]);
        mx = std::max(mx, part[5 * b + 3]);
        pm = std::max(pm, part[5 * b + 4]);
    }
    out[0] = (double)c->nptl;
    out[1] = (double)c->nsplit;
    out[2] = (double)W * kInvFx;
    out[3] = fx_weight(c->leak_acc + hc.leak_fx);
    out[4] = fx_weight(c->negp_acc + hc.negp_fx);
    out[5] = mn;
    out[6] = mx;
    out[7] = c->nptl > 0 ? S / (double)c->nptl : 0.0;
    out[8] = pm;
    return SP_OK;
}

}  // extern "C"
