"""NLGC perpendicular diffusion (calc_spatial_diffusion_coefficients_nlgc,
particle_module.f90:2464-2771) on the GPU through the C ABI: single steps and
whole intervals of the FP64 path against the oracle, the FP32 path against
the FP64 path."""
import numpy as np
import pytest

import oracle
from helpers import cfg2d, cfg3d, nodes_for, random_states
from test_gpu_parity import STATE_COLS, SQ3, _mover_case, by_id, make, rel_err

pytestmark = pytest.mark.gpu

NLGC = dict(nlgc=1, kperp_kpara=0.05)


def pitch_away_from_zero(a):
    """|mu| into [0.3, 0.99], sign kept.  NLGC's kperp ~ mu^2 makes the dt rule's
    (sqrt(2 kperp) / dx_dt)^2 term, and so the step, vanish as mu -> 0: a
    particle injected with |mu| ~ 1e-3 steps at dt_min, 1e7 steps per interval,
    in the reference as here.  The tests keep such particles out."""
    mu = np.asarray(a["mu"])
    a["mu"] = np.sign(mu) * (0.3 + 0.69 * np.abs(mu) / 0.99)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("mag", [0, 1])
def test_step_once_nlgc_matches_oracle(ndim, mag):
    c = dict(cfg2d(32) if ndim == 2 else cfg3d(12), mag_dependency=mag, **NLGC)
    nodes = nodes_for(c)
    pz, _ = make(c, nodes)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    n = 3000
    a = random_states(c, n, seed=37)
    rng = np.random.default_rng(38)
    a["mu"] = rng.uniform(-0.99, 0.99, n)
    pz.set_particles(oracle.soa_from_particles(a))   # step i takes mu of particle i
    r = rng.uniform(-SQ3, SQ3, size=(n, 4))
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
    # kperp carries one more pow() than the standard model, from the device
    # math library here and glibc in the oracle; one ulp of kperp moves a 3D
    # increment by up to 2e-12 of its size where its terms cancel (measured
    # with the oracle alone), hence 1e-11 instead of the 1e-12 of the
    # standard model's test
    scale = np.abs(d_ref).max(axis=0)
    assert np.all(np.abs(d_gpu - d_ref) <= 1e-11 * np.maximum(np.abs(d_ref), 1e-6 * scale) + 1e-300)
    with pytest.raises(Exception):
        pz.step_once(np.vstack([state, state[:1]]), np.vstack([r, r[:1]]),
                     np.concatenate([fixed, fixed[:1]]), t0, dtf)
    pz.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_nlgc_matches_oracle(ndim):
    if ndim == 2:
        c = dict(cfg2d(32, pbcy=1, dt_max_rel=1e-2, mag_dependency=1), **NLGC)
    else:
        c = dict(cfg3d(12, pbcx=1, dt_max_rel=1e-2, mag_dependency=1), **NLGC)
    _mover_case(c, nodes_for(c), 1500, 0.2, nfine=2, prep=pitch_away_from_zero)


def test_mover_nlgc_with_momentum_diffusion_and_time_interpolation():
    c = dict(cfg2d(32, pbcy=1, dt_max_rel=1e-2, time_interp=1), dpp_shear=1, tau0=0.01, **NLGC)
    _mover_case(c, nodes_for(c, seed=1), 1500, 0.2, nodes2=nodes_for(c, seed=2),
                prep=pitch_away_from_zero)


@pytest.mark.parametrize("ndim,prec", [(2, "f32"), (3, "f32"), (3, "f32h")])
def test_fast_paths_nlgc_close_to_f64(ndim, prec):
    if ndim == 2:
        c = dict(cfg2d(32, pbcy=1, dt_max_rel=1e-2, mag_dependency=1), **NLGC)
    else:
        c = dict(cfg3d(12, pbcx=1, dt_max_rel=1e-2, mag_dependency=1), **NLGC)
    nodes = nodes_for(c)
    res = {}
    for p in ("f64", prec):
        pz, _ = make(c, nodes, precision=p, nptl_max=8000)
        pz.inject_particles_spatial_uniform(6000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 1))
        g = pz.get_particles()
        pitch_away_from_zero(g)
        pz.set_particles(g)
        pz.particle_mover(0.0, 0.2, 1, 10)
        res[p] = by_id(pz.get_particles())
        pz.close()
    a, b = res["f64"], res[prec]
    common = np.intersect1d(a["tag_injected"], b["tag_injected"])
    assert len(common) > 0.98 * len(a["p"])
    ia = np.searchsorted(a["tag_injected"], common)
    ib = np.searchsorted(b["tag_injected"], common)
    dp = np.abs(a["p"][ia] - b["p"][ib]) / a["p"][ia]
    assert np.median(dp) < 1e-4
    h64, _ = np.histogram(np.log(a["p"]), bins=40, range=(-1, 1))
    h32, _ = np.histogram(np.log(b["p"]), bins=40, range=(-1, 1))
    assert np.abs(h64 / h64.sum() - h32 / h32.sum()).sum() < 0.01
