"""Throughput of the Parker pusher on one MI355X.

Workloads:
  2d (default)  2D snapshot on a 2048 x 2048 grid (FP32 fields), 1e8
                pseudo-particles injected uniformly with p = p0, MHD interval
                sized for 2000 diffusion-limited steps per particle.
  3d            3D turbulence snapshot on a 512^3 grid with a compression
                slab, 1e8 particles, 1000 diffusion-limited steps per interval.
One bench "step" is one MHD interval of the hot path: the mover over every
particle plus the f(p) and spatial-map histograms.  Inputs are resident in
HBM before timing starts.

Printed JSON line (one):
  value          particle-steps/s over the K timed intervals (FP32 gather path)
  roofline       mover kernel: algorithmic HBM bytes / kernel time (HIP events
                 on the library's stream) against 8 TB/s
  cpu_baseline   the FP64 CPU oracle (OpenMP, per-thread MT19937 increments) on
                 a bounded sample of the same workload, on this host
  accuracy       f(p) and spatial-map L1 of the FP32 path against the FP64
                 path on a particle sample of the same workload
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
# node components the step consumes x corners of the interpolation cell
BYTES_PER_STEP = {2: 4 * 15 * 4, 3: 8 * 21 * 4}
COPY_BYTES = 1 << 30
WORKLOADS = {
    "2d": dict(grid=2048, nsteps=2000, cpu_nptl=1e6),
    "3d": dict(grid=512, nsteps=1000, cpu_nptl=2e5),
}


def traffic_file(kind, grid, nptl):
    return os.path.join(ROOT, "profiles", f"mover_{kind}_{grid}_{nptl:.0e}_traffic.json"
                        .replace("+0", ""))


def measured_traffic(path, workload: dict):
    """HBM bytes per mover launch from the committed rocprofv3 counter runs,
    when they were taken on this workload; None otherwise."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if any(t.get("workload", {}).get(k) != v for k, v in workload.items()):
        return None
    return t.get("traffic_bytes_per_launch")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


BOX = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)


def run_intervals(pz, nptl, dtf, warmup, steps, inject=True):
    box = BOX
    if inject:
        pz.inject_particles_spatial_uniform(nptl, 1e-7, 1, 5.0, 0.0, dtf, box)
    for w in range(warmup):
        st = pz.particle_mover(w * dtf, dtf, 1, 10)
        pz.calc_particle_distributions()
        log(f"warmup interval {w}: {st['steps']:.3e} steps, kernel {st['kernel_ms']:.1f} ms")
    tot_steps = 0
    kern_ms = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        it = warmup + k
        st = pz.particle_mover(it * dtf, dtf, 1, 10)
        fg, maps = pz.calc_particle_distributions()
        tot_steps += st["steps"]
        kern_ms += st["kernel_ms"]
        log(f"timed interval {it}: {st['steps']:.3e} steps, kernel {st['kernel_ms']:.1f} ms, "
            f"nptl {st['nptl_current']}")
    t1 = time.perf_counter()
    return dict(steps=tot_steps, seconds=t1 - t0, kernel_ms=kern_ms, fglobal=fg, maps=maps)


def l1(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def accuracy(cfg, nodes, dist, n, dtf, intervals):
    """f(p) and map L1 of FP32 vs FP64 on the same particles and increments."""
    out = {}
    res = {}
    for prec in ("f64", "f32"):
        with ParkerPusher(**dict(cfg, precision=prec, nptl_max=n)) as pz:
            pz.upload_fields(nodes)
            pz.calc_fields_gradients()
            pz.init_particle_distributions(**dist)
            r = run_intervals(pz, n, dtf, intervals - 1, 1)
            res[prec] = r
            out[f"{prec}_steps_per_s"] = r["steps"] / r["seconds"]
    out["fp_l1"] = l1(res["f32"]["fglobal"], res["f64"]["fglobal"])
    m32 = res["f32"]["maps"][0].sum(axis=(3, 4))
    m64 = res["f64"]["maps"][0].sum(axis=(3, 4))
    out["map_l1"] = l1(m32, m64)
    return out


def cpu_baseline(cfg, nodes, n, dtf, nthreads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    prm = oracle.params(dict(cfg, seed=cfg.get("seed", 20261202)))
    fa = oracle.gradients(prm, nodes)
    box = BOX
    ptl, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), n, n, 1e-7, 1, 5.0, 0.0, dtf,
                           box, 7.0)
    # start every particle at the interval start so the sample does full intervals
    ptl["t"] = 0.0
    t0 = time.perf_counter()
    _, acc = oracle.mover(prm, fa, ptl, 0.0, dtf, rng_mode=1, nthreads=nthreads)
    dt = time.perf_counter() - t0
    return acc.steps / dt, acc.steps, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed MHD intervals")
    ap.add_argument("--warmup", type=int, default=1, help="untimed MHD intervals")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="2d")
    ap.add_argument("--grid", type=int, default=0, help="cells per axis (default per workload)")
    ap.add_argument("--nptl", type=float, default=1e8)
    ap.add_argument("--nsteps", type=int, default=0,
                    help="diffusion-limited steps per interval (default per workload)")
    ap.add_argument("--precision", choices=("f32", "f64"), default="f32")
    ap.add_argument("--accuracy-nptl", type=float, default=1e7)
    ap.add_argument("--cpu-nptl", type=float, default=0, help="CPU sample (default per workload)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the pusher shards by "
              "particles and this build measures one GPU's share", file=sys.stderr)
        sys.exit(2)

    n = int(a.nptl)
    wl = WORKLOADS[a.workload]
    g = a.grid or wl["grid"]
    nsteps = a.nsteps or wl["nsteps"]
    ndim = 3 if a.workload == "3d" else 2
    t0 = time.perf_counter()
    if ndim == 2:
        cfg = workloads.config_2d(g)
        dist = workloads.dist_config_2d(g)
        dtf = workloads.interval_2d(g, nsteps)
        nodes = workloads.fields_2d(g, g)
        desc = f"2D snapshot {g}x{g} FP32 fields"
    else:
        cfg = workloads.config_3d(g)
        dist = workloads.dist_config_3d(g)
        dtf = workloads.interval_3d(g, nsteps)
        nodes = workloads.fields_3d(g)
        desc = f"3D turbulence snapshot {g}^3 FP32 fields with a compression slab"
    log(f"generated {desc} in {time.perf_counter() - t0:.1f} s")
    with ParkerPusher(**dict(cfg, precision=a.precision, nptl_max=n)) as pz:
        pz.upload_fields(nodes)
        pz.calc_fields_gradients()
        pz.init_particle_distributions(**dist)
        # a streaming copy of known size, outside the timed region: the same
        # run then carries the reference point for the memory counters
        copy_ms = pz.calibration_copy(COPY_BYTES)
        log(f"calibration copy {COPY_BYTES} B: {copy_ms:.3f} ms")
        log(f"injecting {n} particles, dtf={dtf:.5g}")
        r = run_intervals(pz, n, dtf, a.warmup, a.steps)
    value = r["steps"] / r["seconds"]
    bytes_per_step = BYTES_PER_STEP[ndim]
    achieved = bytes_per_step * r["steps"] / (r["kernel_ms"] * 1e-3) / 1e9
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
        "dtype": "f32" if a.precision == "f32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{desc}, {n:.0e} particles, "
                        f"{nsteps} diffusion-limited steps per MHD interval",
            "grid": [g] * ndim, "nptl": n, "precision": a.precision,
            "position_time_dtype": "f64", "particle_steps_timed": r["steps"],
            "mover_kernel_ms": r["kernel_ms"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_file(a.workload, g, n),
                                        dict(grid=g, nptl=n, precision=a.precision,
                                             nsteps=nsteps)),
            "kernel": "mover_kernel", "bytes_per_step": bytes_per_step,
        },
    }
    if not a.no_accuracy:
        na = int(a.accuracy_nptl)
        log(f"accuracy sample: {na} particles, FP32 vs FP64")
        acc = accuracy(cfg, nodes, dist, na, dtf, a.warmup + 1)
        result["accuracy"] = dict(sample_nptl=na, fp_l1_vs_f64=acc["fp_l1"],
                                  map_l1_vs_f64=acc["map_l1"],
                                  f64_steps_per_s=acc["f64_steps_per_s"],
                                  f32_steps_per_s=acc["f32_steps_per_s"])
    if not a.no_cpu:
        nt = a.cpu_threads or min(16, os.cpu_count() or 1)
        nc = int(a.cpu_nptl or wl["cpu_nptl"])
        log(f"cpu baseline: {nc} particles, {nt} threads")
        v, s, secs = cpu_baseline(cfg, nodes, nc, dtf, nt)
        result["cpu_baseline"] = {
            "value": v, "unit": "particle-steps/s", "cores": nt, "kind": "port",
            "sample": f"{nc} particles x one full interval of the same snapshot "
                      f"({s} steps, {secs:.1f} s), "
                      "FP64 oracle, AoS particles, 32-component node records, MT19937 per thread",
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
