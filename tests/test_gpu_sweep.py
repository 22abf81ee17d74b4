"""The batched sweep workload: 64 diffusion-parameter variants (kpara0,
kret, momentum exponent, one of four injection regions) over the 512^2 2D
snapshot in one mover launch, particle counts log-uniform in [1e5, 1e7] per
variant (1.39e8 in total).

* Counts scaled by 1e-3 (FP64 path): every variant's particles against the
  oracle run with that variant's parameters (identical integers, floats
  within 1e-9), and the per-variant f(p) and counts bit for bit against
  numpy over the downloaded particles.
* Full counts (FP32 path, one interval): every injected particle is kept or
  counted as escaped, and each variant's f(p) holds exactly its remaining
  particles (unit weights; no momentum leaves (pmin, pmax] in one interval).
"""
import os

import numpy as np
import pytest

import oracle
from sp_hip import ParkerPusher, workloads

pytestmark = pytest.mark.gpu

NGRID = 512


def by_id(arrs):
    key = (np.asarray(arrs["tag_injected"], np.int64) << 32) + np.asarray(arrs["tag_splitted"], np.int64)
    o = np.argsort(key)
    return {k: np.asarray(v)[o] for k, v in arrs.items()}


def sweep_ctx(counts, precision, v, dtf, nodes):
    cfg = workloads.config_2d(NGRID)
    n = int(np.sum(counts))
    pz = ParkerPusher(**dict(cfg, precision=precision, nptl_max=n))
    pz.upload_fields(nodes)
    pz.calc_fields_gradients()
    pz.init_particle_distributions(**workloads.dist_config_2d(NGRID))
    pz.set_variants(list(zip(v["kpara0"], v["kret"], v["gamma_turb"])))
    pz.inject_variants(counts, 1e-7, 1, 5.0, 0.0, dtf, v["boxes"])
    return cfg, pz


@pytest.fixture(scope="module")
def sweep():
    v = workloads.sweep_variants()
    dtf = workloads.sweep_interval(NGRID, float(v["kpara0"].max()))
    nodes = workloads.fields_2d(NGRID, NGRID)
    return v, dtf, nodes


def test_sweep_scaled_matches_oracle_per_variant(sweep):
    v, dtf, nodes = sweep
    counts = np.maximum(1, np.round(v["nptl"] * 1e-3)).astype(np.int64)
    assert len(counts) == 64 and counts.sum() > 5e4
    cfg, pz = sweep_ctx(counts, "f64", v, dtf, nodes)
    start = pz.get_particles()
    st = pz.particle_mover(0.0, dtf, 1, 10)
    f, cnt = pz.variant_distributions()
    got = by_id(pz.get_particles())
    pz.close()
    np.testing.assert_array_equal(np.bincount(start["origin"], minlength=64), counts)
    prm0 = oracle.params(cfg)
    fa = oracle.gradients(prm0, nodes)  # gradients do not depend on the variant
    refs, steps = [], 0
    nth = min(16, os.cpu_count() or 1)
    for k in range(64):
        prm = oracle.params(dict(cfg, kpara0=v["kpara0"][k], kret=v["kret"][k],
                                 gamma_turb=v["gamma_turb"][k]))
        sel = start["origin"] == k
        sub = oracle.particles_from_soa({key: a[sel] for key, a in start.items()})
        out, acc = oracle.mover(prm, fa, sub, 0.0, dtf, nthreads=nth)
        refs.append(oracle.soa_from_particles(out))
        steps += acc.steps
    ref = by_id({key: np.concatenate([r[key] for r in refs]) for key in refs[0]})
    assert st["steps"] == steps
    for key in ("count_flag", "tag_injected", "origin", "rng_step", "nsteps_pushed", "weight"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
    for key in ("x", "y"):
        assert np.max(np.abs(got[key] - ref[key])) < 1e-9, key
    for key in ("p", "t", "dt"):
        assert np.max(np.abs(got[key] - ref[key]) / np.abs(ref[key])) < 1e-9, key
    # per-variant f(p) and counts against numpy over the downloaded particles
    edges = np.linspace(np.log10(cfg["pmin"]), np.log10(cfg["pmax"]), 257)
    for k in range(64):
        sel = got["origin"] == k
        assert cnt[k] == sel.sum()
        p = got["p"][sel]
        ib = np.floor((np.log10(p) - edges[0]) / (edges[1] - edges[0])).astype(np.int64)
        ok = (p > cfg["pmin"]) & (p <= cfg["pmax"]) & (ib >= 0) & (ib < 256)
        want = np.bincount(ib[ok], weights=got["weight"][sel][ok], minlength=256)
        np.testing.assert_array_equal(f[k], want)
    # the variants differ: kpara0 spans three decades, and so does the spread
    # of the y displacements over the interval (y is open: no wrap)
    s0 = by_id(start)
    keep = np.isin(s0["tag_injected"], got["tag_injected"])
    dy = got["y"] - s0["y"][keep]
    spread = np.array([np.std(dy[got["origin"] == k]) for k in range(64)])
    assert spread.max() > 10 * spread.min()
    order = np.argsort(v["kpara0"])
    assert spread[order[-8:]].mean() > 5 * spread[order[:8]].mean()


def test_sweep_full_size_conserves_particles(sweep):
    v, dtf, nodes = sweep
    counts = v["nptl"]
    total = int(counts.sum())
    assert 1.3e8 < total < 1.5e8
    _, pz = sweep_ctx(counts, "f32", v, dtf, nodes)
    f0, cnt0 = pz.variant_distributions()
    np.testing.assert_array_equal(cnt0, counts)
    np.testing.assert_array_equal(f0.sum(axis=1), counts.astype(np.float64))
    st = pz.particle_mover(0.0, dtf, 1, 10)
    f, cnt = pz.variant_distributions()
    c = pz.counts()
    pz.close()
    assert st["steps"] > 100 * total
    assert c["nptl_current"] == cnt.sum() == st["nptl_current"]
    assert c["nptl_current"] + c["nptl_escaped"] == total
    assert np.all(cnt <= counts) and np.all(cnt > 0)
    # unit weights: each variant's f(p) total is its number of particles
    np.testing.assert_array_equal(f.sum(axis=1), cnt.astype(np.float64))
