"""Batched parameter variants: a batched run pushes each particle exactly as a
single-variant run with that variant's parameters does."""
import numpy as np
import pytest

import oracle
from helpers import cfg2d, nodes_for
from sp_hip import ParkerPusher, SpError

pytestmark = pytest.mark.gpu

VARIANTS = [(1e-3, 0.01, 5.0 / 3.0), (3e-3, 0.1, 1.5), (5e-4, 1.0, 2.0), (2e-3, 0.001, 1.2)]
COUNTS = [800, 1200, 500, 1500]
BOXES = [(0.0, 0.0, 0.0, 0.5, 0.5, 0.0), (0.5, 0.0, 0.0, 1.0, 0.5, 0.0),
         (0.0, 0.5, 0.0, 0.5, 1.0, 0.0), (0.25, 0.25, 0.0, 0.75, 0.75, 0.0)]
DTF = 0.2


def by_id(arrs):
    key = np.asarray(arrs["tag_injected"], np.int64) * (1 << 32) + np.asarray(arrs["tag_splitted"], np.int64)
    o = np.argsort(key)
    return {k: np.asarray(v)[o] for k, v in arrs.items()}


def _ctx(c, nodes, precision, nmax=20000, **over):
    pz = ParkerPusher(**dict(c, precision=precision, nptl_max=nmax, **over))
    pz.upload_fields(nodes)
    pz.calc_fields_gradients()
    return pz


def _batched(c, nodes, precision, intervals=2):
    pz = _ctx(c, nodes, precision)
    pz.set_variants(VARIANTS)
    pz.inject_variants(COUNTS, 1e-7, 1, 5.0, 0.0, DTF, BOXES)
    start = pz.get_particles()
    steps = 0
    for it in range(intervals):
        steps += pz.particle_mover(it * DTF, DTF, 1, 10)["steps"]
    return pz, start, steps


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batched_run_equals_single_variant_runs(precision):
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    nodes = nodes_for(c)
    pz, start, steps = _batched(c, nodes, precision)
    got = by_id(pz.get_particles())
    pz.close()
    assert np.array_equal(np.bincount(start["origin"], minlength=4), COUNTS)
    steps_single = 0
    parts = []
    for v, (k0, kr, gt) in enumerate(VARIANTS):
        sel = start["origin"] == v
        one = _ctx(c, nodes, precision, kpara0=k0, kret=kr, gamma_turb=gt)
        one.set_particles({k: a[sel] for k, a in start.items()})
        for it in range(2):
            steps_single += one.particle_mover(it * DTF, DTF, 1, 10)["steps"]
        parts.append(one.get_particles())
        one.close()
    ref = by_id({k: np.concatenate([p[k] for p in parts]) for k in parts[0]})
    assert steps == steps_single
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_batched_f64_matches_oracle():
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    nodes = nodes_for(c)
    pz, start, _ = _batched(c, nodes, "f64", intervals=1)
    got = by_id(pz.get_particles())
    pz.close()
    refs = []
    for v, (k0, kr, gt) in enumerate(VARIANTS):
        prm = oracle.params(dict(c, kpara0=k0, kret=kr, gamma_turb=gt))
        fa = oracle.gradients(prm, nodes)
        sub = oracle.particles_from_soa({k: a[start["origin"] == v] for k, a in start.items()})
        out, _ = oracle.mover(prm, fa, sub, 0.0, DTF, nthreads=8)
        refs.append(oracle.soa_from_particles(out))
    ref = by_id({k: np.concatenate([r[k] for r in refs]) for k in refs[0]})
    for k in ("count_flag", "tag_injected", "origin", "rng_step", "nsteps_pushed"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    for k in ("x", "y", "p", "t", "dt"):
        err = np.abs(got[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1.0)
        assert np.max(err) < 1e-9, k
    # the variants differ: their momentum spreads are not all equal
    spreads = [np.std(np.log(ref["p"][ref["origin"] == v])) for v in range(4)]
    assert max(spreads) > 1.5 * min(spreads)


def test_variant_distributions_match_numpy():
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    pz, _, _ = _batched(c, nodes_for(c), "f64", intervals=1)
    pz.split_particle(1.1, 1.02)
    pz.init_particle_distributions(npp_global=64, nmu_global=1)
    f, cnt = pz.variant_distributions()
    g = pz.get_particles()
    pz.close()
    edges_log = np.linspace(np.log10(c["pmin"]), np.log10(c["pmax"]), 65)
    for v in range(4):
        sel = g["origin"] == v
        assert cnt[v] == sel.sum()
        lp = np.log10(g["p"][sel])
        ib = np.floor((lp - edges_log[0]) / (edges_log[1] - edges_log[0])).astype(np.int64)
        ok = (g["p"][sel] > c["pmin"]) & (g["p"][sel] <= c["pmax"]) & (ib >= 0) & (ib < 64)
        ref = np.bincount(ib[ok], weights=g["weight"][sel][ok], minlength=64)
        np.testing.assert_array_equal(f[v], ref)
    assert f.sum() > 0 and np.any(g["weight"] < 1.0)


def test_variant_argument_errors():
    c = cfg2d(16)
    pz = _ctx(c, nodes_for(c), "f64")
    with pytest.raises(SpError):
        pz.set_variants([(1e-3, 2.0, 1.6)])            # kret > 1
    pz.set_variants(VARIANTS[:2])
    with pytest.raises(SpError):
        pz.inject_variants(COUNTS, 1e-7, 1, 5.0, 0.0, DTF, BOXES)   # 4 > 2 installed
    with pytest.raises(SpError):
        pz.inject_variants([30000, 1], 1e-7, 1, 5.0, 0.0, DTF, BOXES[:2])  # capacity
    pz.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_autotune_keeps_results_and_state(precision):
    """The tuner's trials leave the particles as they were, and the tuned
    launch settings give the same per-particle results as the defaults."""
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    nodes = nodes_for(c)
    runs = {}
    for tune in (False, True):
        pz, start, _ = _batched(c, nodes, precision, intervals=0)
        if tune:
            before = by_id(pz.get_particles())
            q0 = pz.quick_check()
            r = pz.autotune_mover(0.0, 0.05)
            assert r["ntrials"] >= 4 and 1 <= r["blocks_per_cu"] <= 4
            assert r["sort_steps"] in (0, 100, 200, 400)
            if precision == "f64":
                assert r["particles_per_lane"] == 1   # the FP64 mover takes one
            else:
                # one and two particles per lane were both tried
                assert r["particles_per_lane"] in (1, 2) and r["ntrials"] >= 8
            assert r["steps_per_s"] > 0 and r["steps_per_s"] >= 0.5 * r["default_steps_per_s"]
            after = by_id(pz.get_particles())
            for k in before:
                np.testing.assert_array_equal(before[k], after[k], err_msg=k)
            q1 = pz.quick_check()
            # pdt_avg sums over the particle order, which the tuner's sort changes
            assert q1.pop("pdt_avg") == pytest.approx(q0.pop("pdt_avg"), rel=1e-14)
            assert q1 == q0
        st = pz.particle_mover(0.0, DTF, 1, 10)
        runs[tune] = (by_id(pz.get_particles()), st["steps"], st["leak"])
        pz.close()
    (g0, s0, l0), (g1, s1, l1) = runs[False], runs[True]
    assert s0 == s1 and l0 == l1
    for k in g0:
        np.testing.assert_array_equal(g0[k], g1[k], err_msg=k)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_isotropic_diffusion_kret_one(precision):
    """kret = 1 (kperp = kpara): the parallel-minus-perpendicular noise term is
    exactly zero and every momentum stays finite."""
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2, kpara0=5e-4, kret=1.0, gamma_turb=2.0)
    pz = _ctx(c, nodes_for(c), precision, nmax=4000)
    pz.inject_particles_spatial_uniform(1000, 1e-7, 1, 5.0, 0.0, DTF, (0, 0, 0, 1, 1, 0))
    st = pz.particle_mover(0.0, DTF, 1, 10)
    g = pz.get_particles()
    pz.close()
    assert np.all(np.isfinite(g["p"])) and np.all(np.isfinite(g["x"]))
    assert st["steps"] < 200 * 1000  # dt stays near dt_max = dtf / 100
