"""Throughput of the Parker pusher on one MI355X.

Workloads:
  2d (default)  2D snapshot on a 2048 x 2048 grid (FP32 fields), 1e8
                pseudo-particles injected uniformly with p = p0, MHD interval
                sized for 2000 diffusion-limited steps per particle.
  3d            3D turbulence snapshot on a 512^3 grid with a compression
                slab, 1e8 particles, 1000 diffusion-limited steps per interval,
                particles sorted once per interval; FP32 arithmetic on node
                records with FP16-stored velocity and gradients (--precision
                f32 for FP32 records).
  batch         64 diffusion-parameter variants over the 512^2 2D snapshot in
                one mover launch: kpara0, kret, momentum exponent and injection
                region per variant, 1e5 .. 1e7 particles each (1.39e8 total).
  surrogate     the f(p) surrogate: 2e4 sweep runs of 8192 particles (eight
                spectra each) train a 6-256x4-64 SiLU network ensemble on the
                GPU; accuracy against held-out 131072-particle runs and the
                time of a 1e5-query sweep (query rows in HBM, spectra left
                there; the host-array rate is reported beside it) against
                the pusher's.
One bench "step" is one MHD interval of the hot path: the mover over every
particle plus the f(p) and spatial-map histograms.  Inputs are resident in
HBM before timing starts.

Printed JSON line (one):
  value          particle-steps/s over the K timed intervals (FP32 gather path)
  roofline       mover kernel against its binding roof, the VALU issue rate
                 or the texture-data (TD) path, whichever is closer: VALU
                 instructions (or TD busy cycles) per particle-step from the
                 committed rocprofv3 counter run of the same configuration x
                 particle-steps per second of mover kernel time (HIP events on
                 the library's stream) against 256 CUs x 4 SIMDs x 32 lanes x
                 2.4 GHz (or 256 TD units x 2.4 GHz); beside it the HBM view
                 (algorithmic gather bytes and the counter-measured HBM bytes
                 per step against 8 TB/s)
  c2_f32h        the same 2D workload on records with FP16-stored gradients:
                 rate, roofline, f(p) / spatial-map L1 against FP64
  c3             the 3D 512^3 snapshot (1e8 particles) on the FP16 and FP32
                 node records, each with its rate, roofline and f(p) /
                 spatial-map L1 against FP64 (--no-extras skips c2_f32h and c3)
  cpu_baseline   the FP64 CPU oracle (OpenMP, per-thread MT19937 increments) on
                 a bounded sample of the same workload, on this host
  accuracy       f(p) and spatial-map L1 of the FP32 path against the FP64
                 path on a particle sample of the same workload
--time-interp runs the workload on two field frames with time interpolation
between them (the step of real MHD runs), without the CPU baseline and the
extra entries; its roofline takes the counter run of the two-frame
configuration, and its algorithmic bytes count both frames' records.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stochastic-parker_amd"))

from sp_hip import ParkerPusher, workloads  # noqa: E402

METRIC = "particle-steps/s (2D/3D snapshot, 1e8 ptcl) at f(p) L1 err ≤1% vs FP64 SDE"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30


def bytes_per_step(ndim, precision, frames=1):
    """Algorithmic gather bytes of one particle-step: the stored bytes of the
    node values the step consumes x the corners of its interpolation cell x
    the field frames it interpolates between (two under time interpolation).
    FP32 records hold V, B, divV and the B and |B| gradients: 2D 14 FP32
    values per corner (224 B), 3D 19 (608 B).  FP16 records: 2D V and B as 5
    FP32 and 9 FP16 gradients (152 B); 3D B as 3 FP32 and V, divV and the 12
    gradients as 16 FP16 (352 B).  The FP64 path reads 15 / 21 FP32 node
    values (240 / 672 B)."""
    corners = (4 if ndim == 2 else 8) * frames
    if precision == "f64":
        return corners * (15 if ndim == 2 else 21) * 4
    if precision == "f32h":
        return corners * ((5 * 4 + 9 * 2) if ndim == 2 else (3 * 4 + 16 * 2))
    return corners * (14 if ndim == 2 else 19) * 4


DEFAULTS = {
    "2d": dict(grid=2048, nsteps=2000, nptl=1e8, cpu_nptl=1e6, precision="f32"),
    "3d": dict(grid=512, nsteps=1000, nptl=1e8, cpu_nptl=2e5, precision="f32h"),
    "batch": dict(grid=512, nsteps=1000, nptl=0, cpu_nptl=1e6, precision="f32"),
}
BOX = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Workload:
    """Snapshot, configuration, interval and injection of one workload."""

    def __init__(self, kind, grid, nsteps, nptl, ti=False):
        self.kind, self.grid, self.nsteps = kind, grid, nsteps
        self.variants = None
        self.ti = ti
        self.nodes2 = None
        t0 = time.perf_counter()
        if kind == "3d":
            self.ndim = 3
            self.cfg = workloads.config_3d(grid)
            self.dist = workloads.dist_config_3d(grid)
            self.dtf = workloads.interval_3d(grid, nsteps)
            self.nodes = workloads.fields_3d(grid)
            self.desc = f"3D turbulence snapshot {grid}^3 FP32 fields with a compression slab"
        else:
            self.ndim = 2
            self.cfg = workloads.config_2d(grid)
            self.dist = workloads.dist_config_2d(grid)
            self.nodes = workloads.fields_2d(grid, grid)
            self.desc = f"2D snapshot {grid}x{grid} FP32 fields"
            if kind == "batch":
                v = workloads.sweep_variants()
                if nptl:
                    v["nptl"] = np.maximum(1, np.round(v["nptl"] * nptl / v["nptl"].sum())).astype(np.int64)
                self.variants = v
                self.dtf = workloads.sweep_interval(grid, float(v["kpara0"].max()), nsteps)
                self.desc = (f"{len(v['nptl'])} diffusion-parameter variants in one launch over the "
                             f"{self.desc}")
            else:
                self.dtf = workloads.interval_2d(grid, nsteps)
        if ti:
            # a second frame of the same kind (another seed): the mover
            # interpolates between the two in time
            seed = workloads.FIELD_SEED + 1
            self.nodes2 = (workloads.fields_3d(grid, seed=seed) if kind == "3d"
                           else workloads.fields_2d(grid, grid, seed=seed))
            self.desc += ", two frames with time interpolation"
        self.nptl = int(self.variants["nptl"].sum()) if self.variants is not None else int(nptl)
        log(f"generated {self.desc} in {time.perf_counter() - t0:.1f} s")

    def config(self, precision, n):
        """sp_config of a run on this workload.  In 3D the FP32 movers run
        fastest sorted once per interval; the FP64 path re-sorts every 200
        steps (results do not depend on it)."""
        c = dict(self.cfg, precision=precision, nptl_max=n, time_interp=1 if self.ti else 0)
        if self.kind == "3d" and precision == "f64":
            c["sort_steps"] = 200
        return c

    def scaled_counts(self, n):
        c = self.variants["nptl"].astype(np.float64) * (n / self.variants["nptl"].sum())
        return np.maximum(1, np.round(c)).astype(np.int64)

    def setup(self, pz, n):
        """Upload fields, set diagnostics and inject n particles (all variants)."""
        pz.upload_fields(self.nodes)
        pz.calc_fields_gradients()
        if self.ti:
            pz.upload_fields(self.nodes2, slot=1)
            pz.calc_fields_gradients(slot=1)
        pz.init_particle_distributions(**self.dist)
        if self.variants is not None:
            v = self.variants
            pz.set_variants(list(zip(v["kpara0"], v["kret"], v["gamma_turb"])))
            pz.inject_variants(self.scaled_counts(n), 1e-7, 1, 5.0, 0.0, self.dtf, v["boxes"])
        else:
            pz.inject_particles_spatial_uniform(n, 1e-7, 1, 5.0, 0.0, self.dtf, BOX)

    def histograms(self, pz):
        """(f(p) per variant or global, spatial map or None)."""
        if self.variants is not None:
            f, _ = pz.variant_distributions()
            return f, None
        fg, maps = pz.calc_particle_distributions()
        return fg, maps[0].sum(axis=(3, 4))


def run_intervals(wl, pz, warmup, steps):
    for w in range(warmup):
        st = pz.particle_mover(w * wl.dtf, wl.dtf, 1, 10)
        wl.histograms(pz)
        log(f"warmup interval {w}: {st['steps']:.3e} steps, kernel {st['kernel_ms']:.1f} ms")
    tot_steps = 0
    kern_ms = 0.0
    fp = fmap = None
    t0 = time.perf_counter()
    for k in range(steps):
        it = warmup + k
        st = pz.particle_mover(it * wl.dtf, wl.dtf, 1, 10)
        fp, fmap = wl.histograms(pz)
        tot_steps += st["steps"]
        kern_ms += st["kernel_ms"]
        log(f"timed interval {it}: {st['steps']:.3e} steps, kernel {st['kernel_ms']:.1f} ms, "
            f"nptl {st['nptl_current']}")
    t1 = time.perf_counter()
    return dict(steps=tot_steps, seconds=t1 - t0, kernel_ms=kern_ms, fp=fp, fmap=fmap)


def l1(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def accuracy(wl, n, intervals, fasts=("f32",)):
    """f(p) and map L1 of each fast path vs FP64 on the same particles and
    increments: {"f64_steps_per_s": .., path: {fp_l1, map_l1, steps_per_s}}."""
    out = {}
    res = {}
    for prec in ("f64",) + tuple(fasts):
        with ParkerPusher(**wl.config(prec, n)) as pz:
            wl.setup(pz, n)
            res[prec] = run_intervals(wl, pz, intervals - 1, 1)
    out["f64_steps_per_s"] = res["f64"]["steps"] / res["f64"]["seconds"]
    for fast in fasts:
        e = {"steps_per_s": res[fast]["steps"] / res[fast]["seconds"]}
        f32, f64 = res[fast]["fp"], res["f64"]["fp"]
        if wl.variants is not None:
            per = [l1(a, b) for a, b in zip(f32, f64) if b.sum() > 0]
            e["fp_l1"] = max(per)
            e["fp_l1_mean"] = float(np.mean(per))
            e["map_l1"] = None
        else:
            e["fp_l1"] = l1(f32, f64)
            e["map_l1"] = l1(res[fast]["fmap"], res["f64"]["fmap"])
        out[fast] = e
    return out


def cpu_baseline(wl, n, nthreads):
    """The FP64 oracle over n particles of the workload for one full interval."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cfg = dict(wl.cfg, seed=wl.cfg.get("seed", 20261202))
    prm = oracle.params(cfg)
    fa = oracle.gradients(prm, wl.nodes)
    jobs = []
    if wl.variants is not None:
        v = wl.variants
        for k, cnt in enumerate(wl.scaled_counts(n)):
            pk = oracle.params(dict(cfg, kpara0=v["kpara0"][k], kret=v["kret"][k],
                                    gamma_turb=v["gamma_turb"][k]))
            ptl, _ = oracle.inject(pk, np.zeros(0, oracle.PARTICLE_DTYPE), int(cnt), int(cnt),
                                   1e-7, 1, 5.0, 0.0, wl.dtf, v["boxes"][k], 7.0)
            jobs.append((pk, ptl))
    else:
        ptl, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), n, n, 1e-7, 1, 5.0, 0.0,
                               wl.dtf, BOX, 7.0)
        jobs.append((prm, ptl))
    steps = 0
    t0 = time.perf_counter()
    for pk, ptl in jobs:
        # every particle starts at the interval start: the sample does full intervals
        ptl["t"] = 0.0
        _, acc = oracle.mover(pk, fa, ptl, 0.0, wl.dtf, rng_mode=1, nthreads=nthreads)
        steps += acc.steps
    dt = time.perf_counter() - t0
    return steps / dt, steps, dt


# MI355X VALU issue rate: the published FP32 vector rate (157.3 TFLOP/s, an
# FMA being one lane-instruction and two flops) = 256 CUs x 4 SIMDs x 32 lanes
# x 2.4 GHz.  Half of it (16 lanes per SIMD per clock) is not a ceiling: the
# FP16-gradient 2D mover issues 4.3e13 lane-instructions/s (counter run).
VALU_LANE_INSTR_PEAK = 256 * 4 * 32 * 2.4e9
# texture-data (TD) unit cycles: one unit per CU, 256 CUs x 2.4 GHz
TD_CYCLES_PEAK = 256 * 2.4e9
COUNTERS = os.path.join(ROOT, "profiles", "mover_counters.json")


def counter_run(wl, n, precision):
    """The committed rocprofv3 counter figures of the mover on exactly this
    configuration (workload, grid, particles, steps per interval, precision,
    one or two field frames), or None."""
    try:
        with open(COUNTERS) as f:
            t = json.load(f).get(f"{wl.kind}_{'ti_' if wl.ti else ''}{precision}")
    except (OSError, ValueError):
        return None
    if not t:
        return None
    want = dict(grid=wl.grid, nptl=n, nsteps=wl.nsteps, precision=precision,
                time_interp=bool(wl.ti))
    if any(t.get(k) != v for k, v in want.items()):
        return None
    return t


def roofline(wl, n, precision, steps, kernel_ms, launches):
    """The mover against its binding roof: the VALU issue rate (its
    arithmetic ceiling) or the texture-data path that returns the gathered
    records to the registers, whichever this run comes closer to; the other
    and the HBM figures beside it (the gathers are served by L2 / MALL, so the
    algorithmic bytes overstate what reaches HBM).  Instructions and TD busy
    cycles per particle-step come from the committed counter run of this exact
    configuration and are scaled by this run's rate; without one the line
    falls back to the HBM roof."""
    rate = steps / (kernel_ms * 1e-3)
    alg = bytes_per_step(wl.ndim, precision, 2 if wl.ti else 1)
    t = counter_run(wl, n, precision)
    hbm = {"algorithmic_bytes_per_step": alg, "algorithmic_GBs": alg * rate / 1e9,
           "algorithmic_frac": alg * rate / 1e9 / HBM_PEAK_GBS, "peak_GBs": HBM_PEAK_GBS,
           "measured_bytes_per_step": t.get("hbm_bytes_per_step") if t else None,
           "measured_GBs": t["hbm_bytes_per_step"] * rate / 1e9 if t and t.get("hbm_bytes_per_step") else None}
    kern = "mover_kernel (persistent; one launch per MHD interval)"
    if t and t.get("valu_instr_per_step"):
        lane_rate = t["valu_instr_per_step"] * rate
        out = {"bound": "valu_issue", "achieved": lane_rate / 1e12,
               "peak": VALU_LANE_INSTR_PEAK / 1e12, "unit": "T lane-instr/s",
               "frac": lane_rate / VALU_LANE_INSTR_PEAK,
               "traffic": t.get("traffic_bytes_per_launch"), "kernel": kern,
               "valu_instr_per_step": t["valu_instr_per_step"], "launches": launches,
               "launch_ms": kernel_ms / max(launches, 1), "source": t.get("source"), "hbm": hbm}
        if t.get("td_busy") is not None:
            # the texture data path returning the gathered records to the
            # vector registers: busy fractions of the counter run
            out["vmem_path"] = {"ta_busy": t.get("ta_busy"), "td_busy": t["td_busy"],
                                "vmem_instr_per_step": t.get("vmem_instr_per_step")}
        if t.get("td_cycles_per_step"):
            td_rate = t["td_cycles_per_step"] * rate
            if td_rate / TD_CYCLES_PEAK > out["frac"]:
                valu = {k: out[k] for k in ("achieved", "peak", "unit", "frac")}
                out.update({"bound": "texture_data", "achieved": td_rate / 1e9,
                            "peak": TD_CYCLES_PEAK / 1e9, "unit": "G TD busy-cycles/s",
                            "frac": td_rate / TD_CYCLES_PEAK,
                            "td_cycles_per_step": t["td_cycles_per_step"], "valu_issue": valu})
        return out
    return {"bound": "hbm", "achieved": hbm["algorithmic_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": hbm["algorithmic_frac"], "traffic": None, "kernel": kern, "launches": launches,
            "launch_ms": kernel_ms / max(launches, 1), "hbm": hbm,
            "note": "no committed counter run of this configuration"}


def c2h_extra(a, wl, n, acc):
    """The same 2D workload on FP16-gradient records: rate, roofline and the
    accuracy of the shared particle sample."""
    with ParkerPusher(**wl.config("f32h", n)) as pz:
        wl.setup(pz, n)
        r = run_intervals(wl, pz, 1, 1)
    e = {"value": r["steps"] / r["seconds"], "unit": "particle-steps/s",
         "particle_steps_timed": r["steps"], "seconds": r["seconds"],
         "mover_kernel_ms": r["kernel_ms"], "precision": "f32h",
         "roofline": roofline(wl, n, "f32h", r["steps"], r["kernel_ms"], 1)}
    if acc is not None:
        e["accuracy"] = {"sample_nptl": int(a.accuracy_nptl), "fp_l1_vs_f64": acc["f32h"]["fp_l1"],
                         "map_l1_vs_f64": acc["f32h"]["map_l1"]}
    log(f"2d f32h: {e['value']:.3e} particle-steps/s")
    return e


def c3_extra(a):
    """The 3D snapshot on both record formats: rate, roofline and accuracy."""
    d = DEFAULTS["3d"]
    wl = Workload("3d", d["grid"], d["nsteps"], d["nptl"])
    n = wl.nptl
    out = {"workload": f"{wl.desc}, {n:.3g} particles, {wl.nsteps} diffusion-limited steps "
                       "per MHD interval", "grid": [wl.grid] * 3, "nptl": n}
    for prec in ("f32h", "f32"):
        with ParkerPusher(**wl.config(prec, n)) as pz:
            wl.setup(pz, n)
            r = run_intervals(wl, pz, 1, 1)
        e = {"value": r["steps"] / r["seconds"], "unit": "particle-steps/s",
             "particle_steps_timed": r["steps"], "seconds": r["seconds"],
             "mover_kernel_ms": r["kernel_ms"],
             "sort_steps": wl.config(prec, n).get("sort_steps", 0),
             "roofline": roofline(wl, n, prec, r["steps"], r["kernel_ms"], 1)}
        out[prec] = e
        log(f"3d {prec}: {e['value']:.3e} particle-steps/s")
    if not a.no_accuracy:
        na = int(a.accuracy_nptl)
        res = {}
        for prec in ("f64", "f32h", "f32"):
            with ParkerPusher(**wl.config(prec, na)) as pz:
                wl.setup(pz, na)
                res[prec] = run_intervals(wl, pz, 1, 1)
        out["f64_sample_steps_per_s"] = res["f64"]["steps"] / res["f64"]["seconds"]
        for prec in ("f32h", "f32"):
            out[prec]["accuracy"] = {
                "sample_nptl": na, "fp_l1_vs_f64": l1(res[prec]["fp"], res["f64"]["fp"]),
                "map_l1_vs_f64": l1(res[prec]["fmap"], res["f64"]["fmap"])}
    return out


def surrogate_bench(a):
    """Trains the surrogate on pusher spectra and measures a 1e5-query sweep."""
    from sp_hip import surrogate as S

    t0 = time.perf_counter()
    sde = S.SweepSDE(grid=512, nptl=8192)
    log(f"sweep snapshot 512^2 in {time.perf_counter() - t0:.1f} s")
    runs = S.sample_queries(a.sur_runs, 1)
    t0 = time.perf_counter()
    x, y, _ = S.training_set(sde, runs)
    t_data = time.perf_counter() - t0
    steps_data = sde.steps
    log(f"training data: {a.sur_runs} runs, {steps_data:.3e} particle-steps, {t_data:.1f} s")
    ref = S.SweepSDE(grid=512, nptl=131072, nodes=sde.nodes)
    val, test = S.sample_queries(128, 2), S.sample_queries(128, 3)
    f_val = ref.run(val["kpara0"], val["kret"], val["pindex"], val["region"])
    f_test = ref.run(test["kpara0"], test["kret"], test["pindex"], test["region"])
    f_sde = sde.run(test["kpara0"], test["kret"], test["pindex"], test["region"])

    def rows(q):
        rep = lambda v: np.repeat(np.asarray(v), sde.ntimes)
        return S.features(rep(q["kpara0"]), rep(q["kret"]), rep(q["pindex"]), rep(q["region"]),
                          np.tile(sde.times, len(q["kpara0"])), sde.tmax)

    with S.Surrogate(members=2) as sur:
        t0 = time.perf_counter()
        losses = sur.fit(x, y, nsteps=a.sur_steps)
        t_train = time.perf_counter() - t0
        log(f"training: {a.sur_steps} steps x 2 members in {t_train:.1f} s, final loss "
            f"{float(losses[0][-20:].mean()):.4f}")
        sur.calibrate(rows(val), f_val.reshape(-1, sde.nbins))
        f_sur, est = sur.predict(rows(test))
        err = S.l1(f_sur, f_test.reshape(-1, sde.nbins))
        err_sde = S.l1(f_sde, f_test).reshape(-1)
        q_err = err.reshape(128, sde.ntimes).max(axis=1)
        q_ok = est.reshape(128, sde.ntimes).max(axis=1) <= 0.05
        sweep = rows(S.sample_queries(100000, 9))
        # host arrays in and out (PCIe included)
        sur.predict(sweep[:4096])
        t0 = time.perf_counter()
        f_host, _ = sur.predict(sweep)
        t_host = time.perf_counter() - t0
        # the measured sweep: query rows resident in HBM, spectra left there
        from sp_hip.device import DeviceArray
        xd = DeviceArray.from_host(sweep)
        spec_d, spread_d = DeviceArray((len(sweep), sde.nbins)), DeviceArray(len(sweep))
        sur.predict_device(xd, spec_d, spread_d, rows=4096)
        t0 = time.perf_counter()
        sur.predict_device(xd, spec_d, spread_d)
        t_sweep = time.perf_counter() - t0
        t_fwd = 1e-3 * sur.nets[0].last_forward_ms()   # fused launches, both members
        same = bool(np.array_equal(spec_d.to_host(), f_host))
        for d in (xd, spec_d, spread_d):
            d.free()
    ms_sde = 1e3 * t_data / a.sur_runs
    members = 2
    flops = 2.0 * sum(k * n for k, n in zip([6, 256, 256, 256, 256], [256, 256, 256, 256, 64]))
    tflops = flops * members * len(sweep) / t_fwd / 1e12
    # HBM bytes per row and member: the fp32 input row and the fp32 output row
    # (activations stay in LDS, the 0.43 MB of weights in L2)
    row_bytes = 6 * 4 + 64 * 4
    result = {
        "metric": "spectrum queries/s (f(p) surrogate, 1e5-query sweep) at f(p) L1 <=5% "
                  "vs 131072-particle pusher runs",
        "value": 1e5 / t_sweep, "unit": "queries/s", "n_gpus": 1, "steps": a.sur_steps,
        "warmup": 0, "ms_per_step": 1e3 * t_train / (2 * a.sur_steps), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "6-256x4-64 SiLU ensemble of 2 on spectra of the 512^2 2D snapshot: "
                               f"{a.sur_runs} runs x 8192 particles x 8 times",
                   "train_rows": int(len(x)), "train_data_seconds": t_data,
                   "train_data_particle_steps": int(steps_data),
                   "train_seconds": t_train, "batch": 4096},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None,
                     "kernel": "mlp_infer_kernel: all five layers of both members over the "
                               "8e5-row sweep (1.6e6 row passes), HIP-event time",
                     "flop_per_byte": flops / row_bytes,
                     "hbm_ceiling_tflops": flops / row_bytes * 8.0},
        "accuracy": {"l1_mean": float(err.mean()), "l1_p90": float(np.quantile(err, 0.9)),
                     "l1_max": float(err.max()), "pusher_8192_l1_mean": float(err_sde.mean()),
                     "accepted_queries": float(q_ok.mean()),
                     "accepted_query_l1_max_mean": float(q_err[q_ok].mean()) if q_ok.any() else None,
                     "accepted_query_l1_max_max": float(q_err[q_ok].max()) if q_ok.any() else None,
                     "accepted_within_tol": float((q_err[q_ok] <= 0.05).mean()) if q_ok.any() else None,
                     "calibration_quantile": 0.95},
        "speed": {"pusher_ms_per_query": ms_sde, "surrogate_us_per_query": 1e6 * t_sweep / 1e5,
                  "speedup": ms_sde * 1e-3 * 1e5 / t_sweep,
                  "host_arrays_queries_per_s": 1e5 / t_host,
                  "host_arrays_speedup": ms_sde * 1e-3 * 1e5 / t_host,
                  "device_equals_host_arrays": same},
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed MHD intervals")
    ap.add_argument("--warmup", type=int, default=1, help="untimed MHD intervals")
    ap.add_argument("--workload", choices=sorted(DEFAULTS) + ["surrogate"], default="2d")
    ap.add_argument("--sur-runs", type=int, default=20000, help="surrogate training runs")
    ap.add_argument("--sur-steps", type=int, default=8000, help="surrogate AdamW steps")
    ap.add_argument("--grid", type=int, default=0, help="cells per axis (default per workload)")
    ap.add_argument("--nptl", type=float, default=-1, help="particles (default per workload)")
    ap.add_argument("--nsteps", type=int, default=0,
                    help="diffusion-limited steps per interval (default per workload)")
    ap.add_argument("--precision", choices=("f32", "f32h", "f64"), default=None,
                    help="mover arithmetic (default per workload: f32; f32h for 3d)")
    ap.add_argument("--accuracy-nptl", type=float, default=1e7)
    ap.add_argument("--cpu-nptl", type=float, default=0, help="CPU sample (default per workload)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-interp", action="store_true",
                    help="two field frames with time interpolation between them (the step "
                         "of real MHD runs); no CPU baseline or extra entries")
    ap.add_argument("--no-extras", "--no-3d", dest="no_extras", action="store_true",
                    help="skip the extra entries of the 2D run's line (FP16-record 2D path, "
                         "3D snapshot)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the pusher shards by "
              "particles and the benchmark measures one GPU's share", file=sys.stderr)
        sys.exit(2)
    if a.workload == "surrogate":
        surrogate_bench(a)
        return

    d = DEFAULTS[a.workload]
    a.precision = a.precision or d["precision"]
    if a.time_interp:
        a.no_cpu = a.no_extras = True
    wl = Workload(a.workload, a.grid or d["grid"], a.nsteps or d["nsteps"],
                  d["nptl"] if a.nptl < 0 else a.nptl, ti=a.time_interp)
    n = wl.nptl
    with ParkerPusher(**wl.config(a.precision, n)) as pz:
        wl.setup(pz, n)
        # a streaming copy of known size, outside the timed region: the same
        # run then carries the reference point for the memory counters
        copy_ms = pz.calibration_copy(COPY_BYTES)
        log(f"calibration copy {COPY_BYTES} B: {copy_ms:.3f} ms")
        log(f"{n} particles, dtf={wl.dtf:.5g}")
        r = run_intervals(wl, pz, a.warmup, a.steps)
    value = r["steps"] / r["seconds"]
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * r["seconds"] / max(a.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.precision,
        "data": "synthetic",
        "config": {
            "workload": f"{wl.desc}, {n:.3g} particles, "
                        f"{wl.nsteps} diffusion-limited steps per MHD interval",
            "grid": [wl.grid] * wl.ndim, "nptl": n, "precision": a.precision,
            "position_time_dtype": "f64", "particle_steps_timed": r["steps"],
            "mover_kernel_ms": r["kernel_ms"],
        },
        "roofline": roofline(wl, n, a.precision, r["steps"], r["kernel_ms"], a.steps),
    }
    extra2 = a.workload == "2d" and not a.no_extras and a.precision == "f32"
    acc = None
    if not a.no_accuracy:
        na = int(a.accuracy_nptl)
        fast = a.precision if a.precision != "f64" else "f32"
        fasts = (fast, "f32h") if extra2 else (fast,)
        log(f"accuracy sample: {na} particles, {', '.join(fasts)} vs f64")
        acc = accuracy(wl, na, a.warmup + 1, fasts)
        result["accuracy"] = dict(sample_nptl=na, path=fast, fp_l1_vs_f64=acc[fast]["fp_l1"],
                                  map_l1_vs_f64=acc[fast]["map_l1"],
                                  f64_steps_per_s=acc["f64_steps_per_s"],
                                  fast_steps_per_s=acc[fast]["steps_per_s"])
        if "fp_l1_mean" in acc[fast]:
            result["accuracy"]["fp_l1_mean_over_variants"] = acc[fast]["fp_l1_mean"]
    if not a.no_cpu:
        nt = a.cpu_threads or min(16, os.cpu_count() or 1)
        nc = int(a.cpu_nptl or d["cpu_nptl"])
        log(f"cpu baseline: {nc} particles, {nt} threads")
        v, s, secs = cpu_baseline(wl, nc, nt)
        result["cpu_baseline"] = {
            "value": v, "unit": "particle-steps/s", "cores": nt, "kind": "port",
            "sample": f"{nc} particles x one full interval of the same workload "
                      f"({s} steps, {secs:.1f} s), "
                      "FP64 oracle, AoS particles, 32-component node records, MT19937 per thread",
        }
    if extra2:
        result["c2_f32h"] = c2h_extra(a, wl, n, acc)
    if a.workload == "2d" and not a.no_extras:
        result["c3"] = c3_extra(a)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
