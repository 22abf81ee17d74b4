"""GPU parity of libsp_hip.so against the CPU oracle, through the C ABI.

Integer/byte results (counts, flags, tags, step counters, histogram bins)
must be identical.  FP64-mode floating point must agree with the oracle to
1e-12 relative per step (both evaluate the same unfused IEEE operations;
pow/exp/log10 come from different math libraries) and to 1e-9 over whole
intervals of a few hundred steps."""
import numpy as np
import pytest

import oracle
from helpers import cfg2d, cfg3d, nodes_for, random_states
from sp_hip import ParkerPusher

pytestmark = pytest.mark.gpu

SQ3 = np.sqrt(3.0)
STATE_COLS = ("x", "y", "z", "p", "t", "dt")


def make(c, nodes, precision="f64", nptl_max=20000, nodes2=None):
    cc = dict(c)
    cc["precision"] = precision
    cc["nptl_max"] = nptl_max
    pz = ParkerPusher(**cc)
    pz.upload_fields(nodes, 0)
    fa_gpu = pz.calc_fields_gradients(0, return_farray=True)
    if nodes2 is not None:
        pz.upload_fields(nodes2, 1)
        pz.calc_fields_gradients(1)
    return pz, fa_gpu


def rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def by_id(arrs):
    key = np.asarray(arrs["tag_injected"], np.int64) * (1 << 32) + np.asarray(arrs["tag_splitted"], np.int64)
    o = np.argsort(key)
    return {k: np.asarray(v)[o] for k, v in arrs.items()}


@pytest.mark.parametrize("ndim", [2, 3])
def test_gradients_bit_exact(ndim):
    c = cfg2d(48) if ndim == 2 else cfg3d(12)
    nodes = nodes_for(c)
    pz, fa_gpu = make(c, nodes)
    fa = oracle.gradients(oracle.params(c), nodes)
    np.testing.assert_array_equal(fa_gpu, fa)
    pz.close()


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("law", ["uniform", "gauss"])
def test_step_once_f64_matches_oracle(ndim, law):
    c = cfg2d(32) if ndim == 2 else cfg3d(12)
    c = dict(c, check_drift_2d=1, mag_dependency=1 if ndim == 2 else 0)
    nodes = nodes_for(c)
    pz, _ = make(c, nodes)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    n = 3000
    a = random_states(c, n, seed=17)
    rng = np.random.default_rng(18)
    if law == "uniform":
        r = rng.uniform(-SQ3, SQ3, size=(n, 4))
    else:
        r = rng.standard_normal(size=(n, 4))
    fixed = rng.integers(0, 2, n).astype(np.int8)
    t0, dtf = 0.0, 0.05
    state = np.stack([a[k] for k in STATE_COLS], 1)
    st_gpu, d_gpu = pz.step_once(state, r, fixed, t0, dtf)
    st_ref = np.zeros_like(state)
    d_ref = np.zeros((n, 4))
    for i in range(n):
        row, d = oracle.step_once(prm, fa, a[i], t0, dtf, fixed[i], r[i])
        st_ref[i] = [row[k] for k in STATE_COLS]
        d_ref[i] = d
    assert np.max(rel_err(st_gpu, st_ref)) < 1e-12
    scale = np.abs(d_ref).max(axis=0)
    assert np.all(np.abs(d_gpu - d_ref) <= 1e-12 * np.maximum(np.abs(d_ref), 1e-6 * scale) + 1e-300)
    pz.close()


@pytest.mark.parametrize("dist_flag", [0, 1, 2])
def test_injection_matches_oracle(dist_flag):
    c = cfg2d(32, pmin=1e-2, pmax=10.0, p0=0.1)
    pz, _ = make(c, nodes_for(c), nptl_max=6000)
    box = (0.1, 0.2, 0.0, 0.7, 0.8, 1.0)
    pz.inject_particles_spatial_uniform(5000, 1e-7, dist_flag, 5.0, 0.5, 0.25, box, 3.0)
    g = pz.get_particles()
    ref, tag_max = oracle.inject(oracle.params(c), np.zeros(0, oracle.PARTICLE_DTYPE), 6000, 5000,
                                 1e-7, dist_flag, 5.0, 0.5, 0.25, box, 3.0)
    assert pz.counts()["tag_max"] == tag_max == 5000
    for k in ("x", "y", "z", "mu", "t", "dt", "weight"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)
    for k in ("count_flag", "tag_injected", "tag_splitted", "split_times", "rng_step"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)
    # momentum laws use exp/pow from the two math libraries
    np.testing.assert_allclose(g["p"], ref["p"], rtol=1e-14)
    np.testing.assert_allclose(g["v"], ref["v"], rtol=1e-14)
    pz.close()


def _mover_case(c, nodes, n, dtf, t0=0.0, nodes2=None, mixed=False, intervals=1, split=None,
                nfine=1, prep=None):
    pz, _ = make(c, nodes, nptl_max=4 * n, nodes2=nodes2)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    fa2 = oracle.gradients(prm, nodes2) if nodes2 is not None else None
    box = (c["xmin"], c["ymin"], c["zmin"], c["xmax"], c["ymax"], c["zmax"])
    pz.inject_particles_spatial_uniform(n, 1e-7, 1, 5.0, t0, dtf, box)
    ref, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), 4 * n, n, 1e-7, 1, 5.0, t0,
                           dtf, box, 7.0)
    if mixed:
        g = pz.get_particles()
        rng = np.random.default_rng(3)
        flags = rng.choice([1] * 8 + [0, -1], size=n).astype(np.int8)
        neg = rng.random(n) < 0.02
        g["count_flag"] = flags
        ref["count_flag"] = flags
        g["p"] = np.where(neg, -0.5, g["p"])
        ref["p"] = np.where(neg, -0.5, ref["p"])
        pz.set_particles(g)
    if prep is not None:
        # the same change to both particle sets before the run
        g = pz.get_particles()
        prep(g)
        prep(ref)
        pz.set_particles(g)
    steps_ref = 0
    leak_ref = negp_ref = 0.0
    nsplit_ref = 0
    for it in range(intervals):
        t_a = t0 + it * dtf
        st = pz.particle_mover(t_a, dtf, nfine, 10)
        ref, acc = oracle.mover(prm, fa, ref, t_a, dtf, num_fine_steps=nfine, nthreads=8, fa2=fa2)
        steps_ref += acc.steps
        leak_ref += acc.leak
        negp_ref += acc.leak_negp
        assert st["steps"] == acc.steps
        if split is not None:
            pz.split_particle(*split)
            ref, ns = oracle.split(prm, ref, 4 * n, *split)
            nsplit_ref += ns
    g = by_id(pz.get_particles())
    r = by_id(oracle.soa_from_particles(ref))
    assert len(g["x"]) == len(r["x"])
    for k in ("count_flag", "tag_injected", "tag_splitted", "split_times", "rng_step",
              "nsteps_pushed", "weight"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)
    for k in STATE_COLS:
        err = rel_err(g[k], r[k]) if k not in ("x", "y", "z") else np.abs(g[k] - r[k])
        assert np.max(err, initial=0.0) < 1e-9, (k, np.max(err))
    q = pz.quick_check()
    assert q["leak"] == leak_ref and q["leak_negp"] == negp_ref
    assert q["nptl_split"] == nsplit_ref
    pz.close()
    return steps_ref


# Test intervals: dt_max_rel = 1e-2 caps a step at dtf / 100, so one interval
# of dtf = 0.2 is ~100-200 steps per particle on these coarse grids.
@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_f64_matches_oracle(ndim):
    if ndim == 2:
        c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    else:
        c = cfg3d(12, pbcx=1, dt_max_rel=1e-2)
    steps = _mover_case(c, nodes_for(c), 2000, 0.2)
    assert steps > 2000 * 50


def test_mover_f64_flags_negative_p_and_fine_steps():
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2)
    _mover_case(c, nodes_for(c), 1500, 0.2, mixed=True, nfine=3)


def test_mover_f64_multi_interval_with_split():
    # the snapshot's compressions move momenta by ~10-20 % per interval, enough
    # to cross split thresholds at 1.05 p0 * 1.2^k
    c = cfg2d(32, pbcy=1, p0=1.0, dt_max_rel=1e-2)
    _mover_case(c, nodes_for(c), 1500, 0.2, intervals=3, split=(1.2, 1.05))


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_f64_time_interpolation(ndim):
    c = (cfg2d(32, pbcy=1, time_interp=1, dt_max_rel=1e-2) if ndim == 2
         else cfg3d(12, pbcx=1, time_interp=1, dt_max_rel=1e-2))
    n1 = nodes_for(c, seed=1)
    n2 = nodes_for(c, seed=2)
    _mover_case(c, n1, 1500, 0.2, nodes2=n2)


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_paused_launches_match_oracle(ndim):
    # every particle stops after 7 steps per launch and resumes after a sort
    if ndim == 2:
        c = cfg2d(32, pbcy=1, dt_max_rel=1e-2, sort_steps=7)
        _mover_case(c, nodes_for(c), 1500, 0.2, mixed=True, nfine=3)
    else:
        c = cfg3d(12, pbcx=1, dt_max_rel=1e-2, sort_steps=7)
        _mover_case(c, nodes_for(c), 1500, 0.2)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_paused_launches_are_bit_identical(precision):
    """Re-sorting inside an interval never changes a particle's trajectory."""
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2, time_interp=1)
    n1, n2 = nodes_for(c, seed=1), nodes_for(c, seed=2)
    out = {}
    for ss in (0, 5):
        pz, _ = make(dict(c, sort_steps=ss), n1, precision=precision, nptl_max=8000, nodes2=n2)
        pz.inject_particles_spatial_uniform(4000, 1e-7, 1, 5.0, 0.0, 0.2,
                                            (0.0, 0.0, 0.0, 1.0, 1.0, 0.0))
        steps = 0
        for it in range(2):
            steps += pz.particle_mover(it * 0.2, 0.2, 2, 10)["steps"]
            pz.split_particle(1.2, 1.05)
        out[ss] = (by_id(pz.get_particles()), steps, pz.quick_check())
        pz.close()
    (g0, s0, q0), (g1, s1, q1) = out[0], out[5]
    assert s0 == s1
    # pdt_avg is an FP64 sum over the particle order, which the sorts change
    assert q0.pop("pdt_avg") == pytest.approx(q1.pop("pdt_avg"), rel=1e-14)
    assert q0 == q1
    for k in g0:
        np.testing.assert_array_equal(g0[k], g1[k], err_msg=k)


def test_distributions_and_quick_check_bit_exact():
    c = cfg2d(64)
    pz, _ = make(c, nodes_for(c), nptl_max=50000)
    a = random_states(c, 40000, seed=31, slack=False)
    rng = np.random.default_rng(31)
    a["p"] = np.exp(rng.uniform(np.log(5e-3), np.log(2e4), len(a)))
    a["weight"] = 0.5 ** rng.integers(0, 6, len(a))
    a["mu"] = rng.uniform(-0.99, 0.99, len(a))
    pz.set_particles(oracle.soa_from_particles(a))
    dcfg = dict(npp_global=256, nmu_global=1, local_dist=1,
                local=[dict(enabled=1, npbins=16, nmu=1, rx=4, ry=4, rz=1, pmin=1e-2, pmax=1e4),
                       dict(enabled=1, npbins=8, nmu=2, rx=8, ry=16, rz=1, pmin=0.1, pmax=1e3),
                       None, None])
    pz.init_particle_distributions(**dcfg)
    fg, locs = pz.calc_particle_distributions()
    fg_ref, locs_ref = oracle.distributions(oracle.params(c), dcfg, a)
    np.testing.assert_array_equal(fg, fg_ref)
    np.testing.assert_array_equal(locs[0].ravel(), locs_ref[0])
    np.testing.assert_array_equal(locs[1].ravel(), locs_ref[1])
    # arrays are cleared after each call: a second call gives the same result
    fg2, _ = pz.calc_particle_distributions()
    np.testing.assert_array_equal(fg2, fg_ref)
    q = pz.quick_check()
    qr = oracle.quick_check(a, 0, 0.0, 0.0)
    for k in ("nptl_current", "ntot", "pdt_min", "pdt_max", "pmax"):
        assert q[k] == qr[k], k
    assert q["pdt_avg"] == pytest.approx(qr["pdt_avg"], rel=1e-12)
    pz.close()


def test_split_matches_oracle_including_capacity():
    c = cfg2d(16)
    a = random_states(c, 5000, seed=40)
    rng = np.random.default_rng(41)
    k = rng.integers(0, 3, len(a)).astype(np.int8)
    a["split_times"] = k
    a["weight"] = 0.5 ** k
    a["p"] = np.exp(rng.uniform(np.log(0.3), np.log(40.0), len(a)))
    for cap in (20000, 5000 + 123):
        pz, _ = make(c, nodes_for(c), nptl_max=cap)
        pz.set_particles(oracle.soa_from_particles(a))
        pz.split_particle(2.72, 2.0)
        ref, ns = oracle.split(oracle.params(c), a, cap, 2.72, 2.0)
        assert pz.counts()["nptl_split"] == ns
        g = by_id(pz.get_particles())
        r = by_id(oracle.soa_from_particles(ref))
        for key in r:
            np.testing.assert_array_equal(g[key], r[key], err_msg=key)
        pz.close()


def test_empty_population_and_state_errors():
    c = cfg2d(16)
    pz = ParkerPusher(**dict(c, nptl_max=10))
    with pytest.raises(Exception, match="not prepared"):
        pz.particle_mover(0.0, 1.0)
    pz.upload_fields(nodes_for(c))
    pz.calc_fields_gradients()
    st = pz.particle_mover(0.0, 1.0)
    assert st["steps"] == 0 and st["nptl_current"] == 0
    pz.split_particle(2.0, 2.0)
    assert pz.counts()["nptl_current"] == 0
    q = pz.quick_check()
    assert q["nptl_current"] == 0 and q["pdt_min"] == 1.0 and q["pdt_max"] == 0.0
    pz.close()


def _morton2(bx, by):
    k = np.zeros_like(bx, dtype=np.uint64)
    for b in range(16):
        k |= ((bx >> b) & 1).astype(np.uint64) << np.uint64(2 * b)
        k |= ((by >> b) & 1).astype(np.uint64) << np.uint64(2 * b + 1)
    return k


def test_spatial_sort_is_a_stable_permutation():
    c = cfg2d(128)
    pz, _ = make(c, nodes_for(c), nptl_max=30000)
    a = random_states(c, 25000, seed=50)
    pz.set_particles(oracle.soa_from_particles(a))
    pz.sort_particles()
    g = pz.get_particles()
    gs = by_id(g)
    rs = by_id(oracle.soa_from_particles(a))
    for key in rs:
        np.testing.assert_array_equal(gs[key], rs[key], err_msg=key)
    # non-decreasing Morton key of 2-cell blocks (128 cells -> at most 64 blocks per axis)
    ix = np.clip(np.floor((g["x"] - c["xmin"]) / c["dx"]), 0, c["nx"] - 1).astype(np.int64) >> 1
    iy = np.clip(np.floor((g["y"] - c["ymin"]) / c["dy"]), 0, c["ny"] - 1).astype(np.int64) >> 1
    k = _morton2(ix, iy)
    assert np.all(np.diff(k.astype(np.int64)) >= 0)
    # stable: equal keys keep their previous (tag) order
    same = np.diff(k.astype(np.int64)) == 0
    assert np.all(np.diff(g["tag_injected"])[same] > 0)
    pz.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_f32h_steps_close_to_f64(ndim):
    """FP16-stored gradients change only the drift and kappa-divergence terms,
    by about the FP16 rounding (2^-11 relative) of the gradients; the FP32
    x-pair records (divV stored as one sum) stay at FP32 rounding."""
    c = cfg2d(16, mag_dependency=0) if ndim == 2 else cfg3d(12, mag_dependency=0)
    nodes = nodes_for(c)
    n = 3000
    a = random_states(c, n, seed=21)
    rng = np.random.default_rng(22)
    r = rng.uniform(-SQ3, SQ3, size=(n, 4))
    fixed = np.zeros(n, dtype=np.int8)
    state = np.stack([a[k] for k in STATE_COLS], 1)
    out = {}
    for prec in ("f64", "f32", "f32h"):
        pz, _ = make(c, nodes, precision=prec)
        out[prec] = pz.step_once(state, r, fixed, 0.0, 0.05)
        pz.close()
    d64, d32, d16 = out["f64"][1], out["f32"][1], out["f32h"][1]
    for d in (d32, d16):
        scale = np.abs(d64[:, :ndim]).max(axis=1, keepdims=True)
        rel = np.abs(d[:, :ndim] - d64[:, :ndim]) / scale
        assert np.median(rel) < 2e-4 and rel.max() < 2e-2
    # the FP16 records move the result by more than FP32 rounding, but not much
    e32 = np.abs(d32[:, :ndim] - d64[:, :ndim]).mean()
    e16 = np.abs(d16[:, :ndim] - d64[:, :ndim]).mean()
    assert e16 > e32 and e16 < 1e3 * e32 + 1e-12
    # momentum: dp = -(p/3) divV dt from the interpolated divV
    rel_p = np.abs(d32[:, 3] - d64[:, 3]) / np.maximum(np.abs(d64[:, 3]), 1e-300)
    assert np.median(rel_p) < 1e-5


@pytest.mark.parametrize("ndim", [2, 3])
def test_f32h_mover_distribution_close_to_f64(ndim):
    c = cfg2d(32, pbcx=1, dt_max_rel=1e-2) if ndim == 2 else cfg3d(12, pbcx=1, dt_max_rel=1e-2)
    nodes = nodes_for(c)
    res = {}
    for prec in ("f64", "f32h"):
        pz, _ = make(c, nodes, precision=prec, nptl_max=8000)
        pz.inject_particles_spatial_uniform(6000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 1))
        pz.particle_mover(0.0, 0.2, 1, 10)
        res[prec] = by_id(pz.get_particles())
        pz.close()
    a, b = res["f64"], res["f32h"]
    common = np.intersect1d(a["tag_injected"], b["tag_injected"])
    assert len(common) > 0.98 * len(a["p"])
    ia = np.searchsorted(a["tag_injected"], common)
    ib = np.searchsorted(b["tag_injected"], common)
    dp = np.abs(a["p"][ia] - b["p"][ib]) / a["p"][ia]
    assert np.median(dp) < 1e-4
    h64, _ = np.histogram(np.log(a["p"]), bins=40, range=(-1, 1))
    h16, _ = np.histogram(np.log(b["p"]), bins=40, range=(-1, 1))
    assert np.abs(h64 / h64.sum() - h16 / h16.sum()).sum() < 0.01


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("precision", ["f32", "f32h"])
def test_fast_paths_with_time_interpolation_close_to_f64(precision, ndim):
    """Two field frames blended in time: the FP32 and FP16-gradient gathers
    read both frames (each with its own gradient scale) as the FP64 path does."""
    c = (cfg2d(32, pbcx=1, dt_max_rel=1e-2, time_interp=1) if ndim == 2
         else cfg3d(12, pbcx=1, dt_max_rel=1e-2, time_interp=1))
    n1, n2 = nodes_for(c, seed=1), nodes_for(c, seed=2)
    res = {}
    for prec in ("f64", precision):
        pz, _ = make(c, n1, precision=prec, nptl_max=8000, nodes2=n2)
        pz.inject_particles_spatial_uniform(6000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 1))
        pz.particle_mover(0.0, 0.2, 1, 10)
        res[prec] = by_id(pz.get_particles())
        pz.close()
    a, b = res["f64"], res[precision]
    common = np.intersect1d(a["tag_injected"], b["tag_injected"])
    assert len(common) > 0.98 * len(a["p"])
    ia = np.searchsorted(a["tag_injected"], common)
    ib = np.searchsorted(b["tag_injected"], common)
    dp = np.abs(a["p"][ia] - b["p"][ib]) / a["p"][ia]
    assert np.median(dp) < 1e-4
    h64, _ = np.histogram(np.log(a["p"]), bins=40, range=(-1, 1))
    hb, _ = np.histogram(np.log(b["p"]), bins=40, range=(-1, 1))
    assert np.abs(h64 / h64.sum() - hb / hb.sum()).sum() < 0.01


def test_f32h_paused_launches_are_bit_identical():
    c = cfg3d(12, pbcx=1, dt_max_rel=1e-2, time_interp=1)
    n1, n2 = nodes_for(c, seed=1), nodes_for(c, seed=2)
    out = {}
    for ss in (0, 5):
        pz, _ = make(dict(c, sort_steps=ss), n1, precision="f32h", nptl_max=8000, nodes2=n2)
        pz.inject_particles_spatial_uniform(4000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 1))
        steps = pz.particle_mover(0.0, 0.2, 1, 10)["steps"]
        out[ss] = (by_id(pz.get_particles()), steps)
        pz.close()
    (g0, s0), (g1, s1) = out[0], out[5]
    assert s0 == s1
    for k in g0:
        np.testing.assert_array_equal(g0[k], g1[k], err_msg=k)


# ---------------------------------------------------------------------------
# Acceleration-region gate (particle_in_acceleration_region,
# particle_module.f90:2885-2906; applied at :3592-3600 and :4888-4900): a box
# covering part of a domain that is neither at the origin nor of unit length.
ACC_BOX = (0.2, 0.65, 0.3, 0.9, 0.1, 0.7)


def _acc_cfg(ndim, n, **over):
    base = dict(xmin=-0.5, xmax=1.5, lx=2.0, ymin=0.25, ymax=1.75, ly=1.5,
                acc_region_flag=1, acc_region=ACC_BOX, dt_max_rel=1e-2)
    if ndim == 2:
        c = cfg2d(n, dx=2.0 / n, dy=1.5 / n, pbcy=1, **base)
    else:
        c = cfg3d(n, dx=2.0 / n, dy=1.5 / n, dz=0.5 / n, zmin=-1.0, zmax=-0.5, lz=0.5, pbcx=1,
                  **base)
    c.update(over)
    return c


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_acc_region_f64_matches_oracle(ndim):
    c = _acc_cfg(ndim, 32 if ndim == 2 else 12)
    steps = _mover_case(c, nodes_for(c), 2000, 0.2)
    assert steps > 2000 * 50


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_step_once_acc_region_gate(ndim, precision):
    """Both step paths: the gate changes only the momentum, keeps the ungated
    increment inside the box and zeroes it outside (the 0.25 p0 floor still
    applies); membership is decided by numpy on the positions returned."""
    c1 = _acc_cfg(ndim, 32 if ndim == 2 else 12)
    c0 = dict(c1, acc_region_flag=0)
    nodes = nodes_for(c1)
    n = 4000
    a = random_states(c1, n, seed=81)
    rng = np.random.default_rng(82)
    low = rng.random(n) < 0.2
    a["p"] = np.where(low, rng.uniform(0.05, 0.24, n), a["p"])
    r = rng.uniform(-SQ3, SQ3, size=(n, 4))
    fixed = rng.integers(0, 2, n).astype(np.int8)
    state = np.stack([a[k] for k in STATE_COLS], 1)
    out = {}
    for name, cc in (("off", c0), ("on", c1)):
        pz, _ = make(cc, nodes, precision=precision)
        out[name] = pz.step_once(state, r, fixed, 0.0, 0.05)
        pz.close()
    (s0, d0), (s1, d1) = out["off"], out["on"]
    np.testing.assert_array_equal(np.delete(s1, 3, axis=1), np.delete(s0, 3, axis=1))
    np.testing.assert_array_equal(d1[:, :3], d0[:, :3])
    xn = (s0[:, 0] - c1["xmin"]) / c1["lx"]
    yn = (s0[:, 1] - c1["ymin"]) / c1["ly"]
    inside = (xn >= ACC_BOX[0]) & (xn <= ACC_BOX[1]) & (yn >= ACC_BOX[2]) & (yn <= ACC_BOX[3])
    if ndim == 3:
        zn = (s0[:, 2] - c1["zmin"]) / c1["lz"]
        inside &= (zn >= ACC_BOX[4]) & (zn <= ACC_BOX[5])
    assert 0.1 * n < inside.sum() < 0.6 * n
    np.testing.assert_array_equal(s1[inside, 3], s0[inside, 3])
    np.testing.assert_array_equal(d1[inside, 3], d0[inside, 3])
    p_in = state[:, 3]
    keep = ~inside & (p_in >= 0.25 * c1["p0"])
    floor = ~inside & (p_in < 0.25 * c1["p0"])
    assert keep.sum() > 0 and floor.sum() > 0
    np.testing.assert_array_equal(s1[keep, 3], p_in[keep])
    assert np.all(d1[keep, 3] == 0.0)
    assert np.all(s1[floor, 3] == 0.25 * c1["p0"])
    np.testing.assert_array_equal(d1[floor, 3], 0.25 * c1["p0"] - p_in[floor])


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_acc_region_f32_close_to_f64(ndim):
    """Whole interval with the gate on: the FP32 path's momenta stay within
    FP32 rounding of the FP64 path's, particle by particle."""
    c = _acc_cfg(ndim, 32 if ndim == 2 else 12)
    nodes = nodes_for(c)
    res = {}
    for prec in ("f64", "f32"):
        pz, _ = make(c, nodes, precision=prec, nptl_max=8000)
        box = (c["xmin"], c["ymin"], c["zmin"], c["xmax"], c["ymax"], c["zmax"])
        pz.inject_particles_spatial_uniform(6000, 1e-7, 1, 5.0, 0.0, 0.2, box)
        pz.particle_mover(0.0, 0.2, 1, 10)
        res[prec] = by_id(pz.get_particles())
        pz.close()
    a, b = res["f64"], res["f32"]
    common = np.intersect1d(a["tag_injected"], b["tag_injected"])
    assert len(common) > 0.98 * len(a["p"])
    ia = np.searchsorted(a["tag_injected"], common)
    ib = np.searchsorted(b["tag_injected"], common)
    dp = np.abs(a["p"][ia] - b["p"][ib]) / a["p"][ia]
    assert np.median(dp) < 1e-4
    # particles that never entered the box keep their injected momentum
    # exactly; both paths agree on which particles those are, except where a
    # trajectory grazes a face of the box
    never_a, never_b = a["p"][ia] == 1.0, b["p"][ib] == 1.0
    assert never_a.sum() > 0.05 * len(common)
    assert np.sum(never_a != never_b) <= 0.002 * len(common)


# ---------------------------------------------------------------------------
# Interval-end undo and fixed-dt redo against the closed form (uniform B,
# constant divV, dt_min = dt_max, dtf not a multiple of dt): p after the
# interval is p0 (1 - divV dt / 3)^N (1 - divV (dtf - N dt) / 3) after N + 2
# executed steps (particle_module.f90:1600-1803).
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fixed_dt_redo_matches_closed_form(precision):
    from test_oracle_physics import redo_case, redo_closed_form, redo_particles
    c, nodes, dtf, dt, ndt, a = redo_case()
    n = 4000
    pin = redo_particles(n)
    pz, _ = make(c, nodes, precision=precision, nptl_max=n)
    pz.set_particles(oracle.soa_from_particles(pin))
    st = pz.particle_mover(0.0, dtf, 1, 10)
    g = by_id(pz.get_particles())
    pz.close()
    assert st["steps"] == n * (ndt + 2) and st["nptl_current"] == n
    np.testing.assert_allclose(g["t"], dtf, rtol=1e-13)
    # the FP32 path's dt rule rounds dt to FP32; its time stays FP64
    dt_used = dt if precision == "f64" else float(np.float32(dt))
    assert np.all(g["dt"] == dt_used)
    want = redo_closed_form(pin["p"], a, dt_used, dtf, ndt)
    # FP32 step arithmetic: dp = FP32(dp_dt * dt) per step, ~1e-7 of dp
    tol = 1e-12 if precision == "f64" else 2e-8
    np.testing.assert_allclose(g["p"], want, rtol=tol)
    assert np.all(np.abs(g["p"] / want - 1) < 1e-3 * abs(a * dt / 3.0))


def test_split_stops_at_the_fixed_point_limit():
    """Weights are summed exactly in units of 2^-32.  A split that would give a
    copy split_times 33 (weight 2^-33) is refused before anything changes;
    up to split_times 32 the sums of weights stay exact."""
    from sp_hip import SpError
    c = cfg2d(16)
    n = 2000
    a = random_states(c, n, seed=90)
    k = np.where(np.arange(n) % 2 == 0, 30, 31).astype(np.int8)
    a["split_times"] = k
    a["weight"] = 0.5 ** k.astype(np.float64)
    a["p"] = 6.0          # above pmin_split p0 1.05^k for every k <= 36
    a["mu"] = 0.5
    pz, _ = make(c, nodes_for(c), nptl_max=4 * n)
    pz.set_particles(oracle.soa_from_particles(a))
    pz.init_particle_distributions(npp_global=64, nmu_global=1, local_dist=0)
    pz.split_particle(1.05, 1.0)
    ref, ns = oracle.split(oracle.params(c), a, 4 * n, 1.05, 1.0)
    assert ns == n and pz.counts()["nptl_split"] == n
    g = by_id(pz.get_particles())
    r = by_id(oracle.soa_from_particles(ref))
    for key in r:
        np.testing.assert_array_equal(g[key], r[key], err_msg=key)
    assert g["split_times"].max() == 32 and g["weight"].min() == 2.0 ** -32
    # the weight sum is conserved exactly: 1000 (2^-30 + 2^-31)
    want = (n // 2) * (2.0 ** -30 + 2.0 ** -31)
    assert pz.quick_check()["ntot"] == want
    fg, _ = pz.calc_particle_distributions()
    assert fg.sum() == want
    # the next split would create weights 2^-33: refused, particles unchanged
    with pytest.raises(SpError, match="split_times"):
        pz.split_particle(1.05, 1.0)
    g2 = by_id(pz.get_particles())
    for key in g:
        np.testing.assert_array_equal(g2[key], g[key], err_msg=key)
    assert pz.counts()["nptl_split"] == n
    # uploads must hold weights the exact sums can represent
    bad = oracle.soa_from_particles(a)
    bad["weight"] = bad["weight"].copy()
    bad["weight"][7] = 0.3
    with pytest.raises(SpError, match="multiple of 2\\^-32"):
        pz.set_particles(bad)
    pz.close()
