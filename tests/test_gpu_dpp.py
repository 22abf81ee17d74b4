"""Momentum diffusion in the push (calc_dpp_wave_scattering and
calc_dpp_flow_shear, particle_module.f90:2918-2979), on the GPU through the
C ABI: the FP64 path against the oracle step by step and over whole
intervals, the FP32 path against the FP64 path, and the configuration
checks."""
import numpy as np
import pytest

import oracle
from helpers import cfg2d, cfg3d, nodes_for, random_states
from sp_hip import ParkerPusher
from test_gpu_parity import STATE_COLS, SQ3, _mover_case, by_id, make, rel_err

pytestmark = pytest.mark.gpu

# rho ~ 1e3 (Alfven speed ~ 0.03) and tau0 = 0.01 keep both rates at a few
# per cent of p per interval on the test snapshots (whose |B| ~ 1, |grad V| ~ 1)
DPP = dict(dpp_wave=1, dpp_shear=1, tau0=0.01)


def dpp_nodes(c, seed=7):
    """The test snapshot with a density that varies in space (rho enters the
    Alfven speed), 600 to 1400 everywhere."""
    nodes = nodes_for(c, seed=seed).astype(np.float64)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in nodes.shape[:3]], indexing="ij")
    nodes[..., 3] = 1e3 * (1.0 + 0.4 * np.sin(0.37 * x + 0.21 * y + 0.13 * z))
    return nodes.astype(np.float32)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("weak", [0, 1])
def test_step_once_dpp_matches_oracle(ndim, weak):
    c = cfg2d(32) if ndim == 2 else cfg3d(12)
    c = dict(c, mag_dependency=1 if ndim == 2 else 0, weak_scattering=weak, **DPP)
    nodes = dpp_nodes(c)
    pz, _ = make(c, nodes)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    n = 3000
    a = random_states(c, n, seed=27)
    rng = np.random.default_rng(28)
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
    scale = np.abs(d_ref).max(axis=0)
    assert np.all(np.abs(d_gpu - d_ref) <= 1e-12 * np.maximum(np.abs(d_ref), 1e-6 * scale) + 1e-300)
    # the terms are active: the momentum increments differ from a run without them
    pz0, _ = make(dict(c, dpp_wave=0, dpp_shear=0), nodes)
    _, d0 = pz0.step_once(state, r, fixed, t0, dtf)
    assert np.median(np.abs(d0[:, 3] - d_gpu[:, 3]) / np.abs(d_gpu[:, 3])) > 1e-3
    pz.close()
    pz0.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_dpp_matches_oracle(ndim):
    if ndim == 2:
        c = dict(cfg2d(32, pbcy=1, dt_max_rel=1e-2), **DPP)
    else:
        c = dict(cfg3d(12, pbcx=1, dt_max_rel=1e-2), **DPP)
    _mover_case(c, dpp_nodes(c), 1500, 0.2)


@pytest.mark.parametrize("ndim", [2, 3])
def test_mover_dpp_weak_scattering_with_time_interpolation(ndim):
    c = (cfg2d(32, pbcy=1, dt_max_rel=1e-2, time_interp=1) if ndim == 2
         else cfg3d(12, pbcx=1, dt_max_rel=1e-2, time_interp=1))
    c = dict(c, weak_scattering=1, **DPP)
    _mover_case(c, dpp_nodes(c, seed=1), 1500, 0.2, nodes2=dpp_nodes(c, seed=2))


@pytest.mark.parametrize("ndim", [2, 3])
def test_f32_dpp_close_to_f64(ndim):
    if ndim == 2:
        c = dict(cfg2d(32, pbcy=1, dt_max_rel=1e-2), **DPP)
    else:
        c = dict(cfg3d(12, pbcx=1, dt_max_rel=1e-2), **DPP)
    nodes = dpp_nodes(c)
    res = {}
    for prec in ("f64", "f32"):
        pz, _ = make(c, nodes, precision=prec, nptl_max=8000)
        pz.inject_particles_spatial_uniform(6000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 1))
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
    h64, _ = np.histogram(np.log(a["p"]), bins=40, range=(-1.5, 1.5))
    h32, _ = np.histogram(np.log(b["p"]), bins=40, range=(-1.5, 1.5))
    assert np.abs(h64 / h64.sum() - h32 / h32.sum()).sum() < 0.01


def test_dpp_configuration_checks():
    with pytest.raises(Exception) as ei:
        ParkerPusher(**dict(cfg3d(8), precision="f32h", dpp_wave=1))
    assert "momentum diffusion" in str(ei.value)
    with pytest.raises(Exception):
        ParkerPusher(**dict(cfg2d(8), dpp_shear=2))
    with pytest.raises(Exception):
        ParkerPusher(**dict(cfg2d(8), dpp_shear=1, tau0=-1.0))
