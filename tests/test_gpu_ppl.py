"""Particles per mover lane (1 or 2, FP32 movers): a lane carrying two
particles steps each from its own Philox counter, so every particle's
trajectory, step count and the leak sums are bit-identical to one particle per
lane, whatever the features, the batched variants or the paused launches.  One
particle per lane keeps its cell's records in registers across steps and two
do not, so these tests also check that the register cache changes no bit."""
import os

import numpy as np
import pytest

from helpers import cfg2d, cfg3d, nodes_for
from sp_hip import ParkerPusher

pytestmark = pytest.mark.gpu

VARIANTS = [(1e-3, 0.01, 5.0 / 3.0), (3e-3, 0.1, 1.5), (5e-4, 1.0, 2.0), (2e-3, 0.001, 1.2)]


def by_id(arrs):
    key = (np.asarray(arrs["tag_injected"], np.int64) << 32) + np.asarray(arrs["tag_splitted"], np.int64)
    o = np.argsort(key)
    return {k: np.asarray(v)[o] for k, v in arrs.items()}


def run(c, precision, ppl, batched=False, intervals=2):
    old = os.environ.get("SP_MOVER_PPL")
    os.environ["SP_MOVER_PPL"] = str(ppl)
    try:
        with ParkerPusher(**dict(c, precision=precision, nptl_max=20000)) as pz:
            pz.upload_fields(nodes_for(c))
            pz.calc_fields_gradients()
            box = (c["xmin"], c["ymin"], c["zmin"], c["xmax"], c["ymax"], c["zmax"])
            if batched:
                pz.set_variants(VARIANTS)
                pz.inject_variants([1500, 2000, 1000, 2500], 1e-7, 1, 5.0, 0.0, 0.2, [box] * 4)
            else:
                pz.inject_particles_spatial_uniform(7000, 1e-7, 1, 5.0, 0.0, 0.2, box)
            steps = []
            for it in range(intervals):
                steps.append(pz.particle_mover(it * 0.2, 0.2, 2, 10)["steps"])
                pz.split_particle(1.2, 1.05)
            return by_id(pz.get_particles()), steps, pz.quick_check()
    finally:
        if old is None:
            os.environ.pop("SP_MOVER_PPL", None)
        else:
            os.environ["SP_MOVER_PPL"] = old


CASES = {
    "2d": (lambda: cfg2d(32, pbcy=1, dt_max_rel=1e-2), "f32", False),
    "3d": (lambda: cfg3d(12, pbcx=1, dt_max_rel=1e-2), "f32", False),
    "3d_f32h": (lambda: cfg3d(12, pbcx=1, dt_max_rel=1e-2), "f32h", False),
    "2d_f32h": (lambda: cfg2d(32, pbcy=1, dt_max_rel=1e-2), "f32h", False),
    "2d_batched": (lambda: cfg2d(32, pbcy=1, dt_max_rel=1e-2), "f32", True),
    "2d_acc_nlgc_mag": (lambda: cfg2d(32, pbcy=1, dt_max_rel=1e-2, acc_region_flag=1,
                                      acc_region=(0.1, 0.7, 0.2, 0.9, 0.0, 1.0), nlgc=1,
                                      kperp_kpara=0.05, mag_dependency=1, check_drift_2d=1),
                        "f32", False),
    "2d_paused": (lambda: cfg2d(32, pbcy=1, dt_max_rel=1e-2, sort_steps=9), "f32", False),
    "3d_batched_paused": (lambda: cfg3d(12, pbcx=1, dt_max_rel=1e-2, sort_steps=11), "f32h", True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_two_particles_per_lane_bit_identical(case):
    make_cfg, precision, batched = CASES[case]
    c = make_cfg()
    print(f"ppl: {case} {precision}", flush=True)
    g1, s1, q1 = run(c, precision, 1, batched)
    g2, s2, q2 = run(c, precision, 2, batched)
    assert s1 == s2 and s1[0] > 7000 * 20
    # pdt_avg sums over the particle order, which the lane layout changes
    assert q1.pop("pdt_avg") == pytest.approx(q2.pop("pdt_avg"), rel=1e-14)
    assert q1 == q2
    for k in g1:
        np.testing.assert_array_equal(g2[k], g1[k], err_msg=k)
    assert q1["nptl_split"] > 0


def test_time_interpolation_keeps_one_particle_per_lane():
    """Two frames per particle take one particle per lane (two would spill
    registers); asking for two changes nothing."""
    c = cfg2d(32, pbcy=1, dt_max_rel=1e-2, time_interp=1)
    out = {}
    for ppl in (1, 2):
        old = os.environ.get("SP_MOVER_PPL")
        os.environ["SP_MOVER_PPL"] = str(ppl)
        try:
            with ParkerPusher(**dict(c, precision="f32", nptl_max=8000)) as pz:
                pz.upload_fields(nodes_for(c, seed=1), 0)
                pz.calc_fields_gradients(0)
                pz.upload_fields(nodes_for(c, seed=2), 1)
                pz.calc_fields_gradients(1)
                pz.inject_particles_spatial_uniform(4000, 1e-7, 1, 5.0, 0.0, 0.2, (0, 0, 0, 1, 1, 0))
                st = pz.particle_mover(0.0, 0.2, 1, 10)
                out[ppl] = (by_id(pz.get_particles()), st["steps"])
        finally:
            if old is None:
                os.environ.pop("SP_MOVER_PPL", None)
            else:
                os.environ["SP_MOVER_PPL"] = old
    assert out[1][1] == out[2][1]
    for k in out[1][0]:
        np.testing.assert_array_equal(out[2][0][k], out[1][0][k], err_msg=k)
