"""Distribution-level and analytic checks of the GPU pusher.

* Independent RNG streams: the GPU mover (Philox4x32-10) and the oracle's
  mover with the reference's generator (MT19937, four (2u - 1) sqrt(3) draws
  per step, random_number_generator.f90:97-101) push the same population
  through the same snapshot; the final momenta and positions must agree in
  distribution (two-sample KS) and the first three moments of log p within
  3 sigma.
* Pure diffusion (V = 0, uniform B): var(dx) = 2 kpara t, var(dy) = 2 kperp t.
* Diffusive shock acceleration: a planar shock of compression ratio r gives
  f(p) ~ p^-q with q = 3 r / (r - 1) (Parker equation, test-particle limit).
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from helpers import cfg2d
from sp_hip import ParkerPusher, workloads

pytestmark = pytest.mark.gpu


def _moments_within(a, b, nsig=3.0):
    """First three moments of a and b (mean, variance, third central moment)
    agree within nsig standard errors of their difference."""
    out = []
    for k in (1, 2, 3):
        if k == 1:
            fa, fb = a, b
        else:
            fa, fb = (a - a.mean()) ** k, (b - b.mean()) ** k
        se = np.sqrt(fa.var() / len(fa) + fb.var() / len(fb))
        out.append((abs(fa.mean() - fb.mean()), nsig * se))
    return out


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_independent_streams_agree_in_distribution(precision):
    n_grid, npart = 128, 20000
    c = cfg2d(n_grid, pbcx=0, pbcy=1, dt_max_rel=2e-3)
    nodes = workloads.fields_2d(n_grid, n_grid)
    dtf = workloads.interval_2d(n_grid, 400)
    box = (0.0, 0.0, 0.0, 1.0, 1.0, 0.0)
    cc = dict(c, precision=precision, nptl_max=npart)
    with ParkerPusher(**cc) as pz:
        pz.upload_fields(nodes)
        pz.calc_fields_gradients()
        pz.inject_particles_spatial_uniform(npart, 1e-7, 1, 1.0, 0.0, dtf, box)
        start = pz.get_particles()
        st = pz.particle_mover(0.0, dtf, 1, 100)
        gpu = pz.get_particles()
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    ref, acc = oracle.mover(prm, fa, oracle.particles_from_soa(start), 0.0, dtf, rng_mode=1,
                            nthreads=8)
    # both remove escaped particles; the escaped weights agree in distribution
    n_gpu, n_ref = len(gpu["p"]), len(ref)
    assert abs(n_gpu - n_ref) < 5.0 * np.sqrt(npart - min(n_gpu, n_ref)) + 10
    assert st["steps"] > 100 * npart  # the interval holds hundreds of steps
    for key in ("p", "x", "y"):
        res = stats.ks_2samp(gpu[key], ref[key])
        assert res.pvalue > 1e-3, (key, res)
    lg, lr = np.log(gpu["p"]), np.log(ref["p"])
    for diff, bound in _moments_within(lg, lr):
        assert diff < bound
    # the momentum spread is real (diffusion and compression both act)
    assert lg.std() > 1e-3


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_pure_diffusion_variance(precision):
    # uniform B along x, V = 0: every rate is zero, so dt = dt_min = dt_max
    n_grid, npart, dtf = 64, 20000, 0.25
    c = cfg2d(n_grid, pbcx=1, pbcy=1, dt_min_rel=1e-3, dt_max_rel=1e-3, kret=0.1)
    nodes = np.zeros((1, n_grid + 4, n_grid + 4, 8), dtype=np.float32)
    nodes[..., 3] = 1.0
    nodes[..., 4] = 1.0
    nodes[..., 7] = 1.0
    cc = dict(c, precision=precision, nptl_max=npart)
    with ParkerPusher(**cc) as pz:
        pz.upload_fields(nodes)
        pz.calc_fields_gradients()
        pz.inject_particles_spatial_uniform(npart, 1e-7, 1, 1.0, 0.0, 0.0,
                                            (0.5, 0.5, 0.0, 0.5, 0.5, 0.0))
        st = pz.particle_mover(0.0, dtf, 1, 100)
        out = pz.get_particles()
    assert len(out["x"]) == npart and st["leak"] == 0.0
    np.testing.assert_allclose(out["t"], dtf, rtol=1e-12)
    assert np.all(out["p"] == 1.0)
    sig = np.sqrt(2.0 / npart)
    kpar = c["kpara0"]
    assert abs(np.var(out["x"] - 0.5) / (2 * kpar * dtf) - 1) < 4 * sig
    assert abs(np.var(out["y"] - 0.5) / (2 * kpar * c["kret"] * dtf) - 1) < 4 * sig
    assert abs(np.mean(out["x"] - 0.5)) < 4 * np.sqrt(2 * kpar * dtf / npart)
    # dtf / dt steps per particle (the last one possibly the fixed-dt redo)
    assert st["steps"] == pytest.approx(npart * 1000, rel=1e-3)


def _shock_nodes(nx, ny, dx, xs, width, u1, r):
    """Vx steps from u1 to u1 / r across a tanh layer at xs; B = x (parallel shock)."""
    x = (np.arange(nx + 4) - 2.0) * dx
    vx = u1 - 0.5 * (u1 - u1 / r) * (1.0 + np.tanh((x - xs) / width))
    out = np.zeros((1, ny + 4, nx + 4, 8), dtype=np.float64)
    out[0, :, :, 0] = vx[None, :]
    out[0, ..., 3] = 1.0
    out[0, ..., 4] = 1.0
    out[0, ..., 7] = 1.0
    return out.astype(np.float32)


@pytest.mark.parametrize("r", [2.0, 3.0, 4.0])
def test_shock_power_law_index(r):
    """Particles injected at the shock at p0 are followed until they leave the
    box; the time integral of f(p) is the steady-state spectrum of continuous
    injection, whose slope in log bins is 3 - q.  Constant kappa (no momentum
    dependence) with kappa / (U1 width) = 20, upstream and downstream escape
    lengths >= 6.7 kappa / U."""
    nx, ny, dx = 1600, 8, 1e-3
    lx, ly, xs, width, kappa = nx * dx, ny * dx, 0.4, 0.003, 0.06
    npart, dtf, nint = 20000, 0.2, 40
    dtmin = (0.5 * dx) ** 2 / (2 * kappa)  # the diffusive limit of the dt rule
    c = cfg2d(8, ndim=2, nx=nx, ny=ny, dx=dx, dy=dx, xmax=lx, ymax=ly, lx=lx, ly=ly,
              pbcx=1, pbcy=0, momentum_dependency=0, kpara0=kappa, kret=0.01,
              dt_min_rel=dtmin / dtf, dt_max_rel=10 * dtmin / dtf, precision="f32",
              nptl_max=npart, sort_every=0)
    with ParkerPusher(**c) as pz:
        pz.upload_fields(_shock_nodes(nx, ny, dx, xs, width, 1.0, r))
        pz.calc_fields_gradients()
        pz.init_particle_distributions(npp_global=64, nmu_global=1, local_dist=0)
        pz.inject_particles_spatial_uniform(npart, dtmin, 1, 1.0, 0.0, dtf,
                                            (xs, 0.0, 0.0, xs, ly, 0.0))
        acc = np.zeros(64)
        for k in range(nint):
            st = pz.particle_mover(k * dtf, dtf, 1, 100)
            fg, _ = pz.calc_particle_distributions()
            acc += fg
            if st["nptl_current"] == 0:
                break
        left = pz.counts()["nptl_current"]
    assert left < 0.1 * npart
    edges = np.logspace(-2, 4, 65)
    pc = np.sqrt(edges[1:] * edges[:-1])
    sel = (pc > 1.5) & (pc < 15.0)
    assert np.all(acc[sel] > 0)
    slope = np.polyfit(np.log(pc[sel]), np.log(acc[sel]), 1)[0]
    q = 3.0 - slope
    assert abs(q - 3.0 * r / (r - 1.0)) < 0.2, q
    # no particle loses momentum below p0 except by rounding: compression only
    assert np.all(acc[pc < 0.95] == 0)
