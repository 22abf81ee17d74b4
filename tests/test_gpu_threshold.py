"""Threshold injections (inject_particles_at_large_jz / _absj / _divv / _rho,
particle_module.f90:785-1470) and their cell counters on the GPU, through the
C ABI, against the oracle: counts equal, injected particles bit for bit."""
import numpy as np
import pytest

import oracle
from sp_hip import ParkerPusher
from test_oracle_population import THRESHOLDS, TBOX, threshold_case

pytestmark = pytest.mark.gpu


def pusher(c, nodes, nptl_max=5000):
    pz = ParkerPusher(**dict(c, nptl_max=nptl_max))
    pz.upload_fields(nodes, 0)
    pz.calc_fields_gradients(0)
    return pz


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("quantity", ["jz", "absj", "divv", "rho"])
def test_threshold_counts_and_injection_match_oracle(ndim, quantity):
    c, nodes, prm, fa = threshold_case(ndim)
    thr = THRESHOLDS[quantity]
    with pusher(c, nodes) as pz:
        assert pz.ncells_above_threshold(quantity, thr, TBOX) == \
            oracle.ncells_above_threshold(prm, fa, quantity, thr, TBOX)
        ref = np.zeros(0, oracle.PARTICLE_DTYPE)
        tag_max = 0
        for dist_flag in (0, 1, 2):
            n = pz.inject_particles_threshold(quantity, 1500, 1e-7, dist_flag, 5.0, 0.2, 0.1, True,
                                              thr, 1, TBOX, 3.0)
            g = pz.get_particles()
            ref, tag_max, ninj = oracle.inject_threshold(
                prm, fa, ref, 5000, quantity, 1500, 1e-7, dist_flag, 5.0, 0.2, 0.1, True, thr, 1,
                TBOX, 3.0, tag_max=tag_max)
            assert n == ninj == 1500
            assert pz.counts()["tag_max"] == tag_max and len(g["x"]) == len(ref)
            for k in ("x", "y", "z", "mu", "t", "dt", "weight", "count_flag", "tag_injected",
                      "tag_splitted", "split_times", "rng_step"):
                np.testing.assert_array_equal(g[k], ref[k], err_msg=f"{k} dist_flag {dist_flag}")
            # momentum laws use exp / pow from the two math libraries
            np.testing.assert_allclose(g["p"], ref["p"], rtol=1e-14)


def test_threshold_injection_edge_cases():
    c, nodes, prm, fa = threshold_case(2)
    with pusher(c, nodes, nptl_max=120) as pz:
        # no node passes: nothing is injected
        assert pz.inject_particles_threshold("jz", 50, 1e-7, 1, 5.0, 0.0, 0.1, True, 1e9, 1,
                                             TBOX) == 0
        assert pz.counts()["nptl_current"] == 0
        # the loop's initial value already passes: particles at the box corner
        assert pz.inject_particles_threshold("rho", 20, 1e-7, 1, 5.0, 0.0, 0.1, True, -1.0, 1,
                                             TBOX) == 20
        g = pz.get_particles()
        assert np.all(g["x"] == TBOX[0]) and np.all(g["y"] == TBOX[1])
        # past nptl_max the last slot is overwritten, as for uniform injection
        n = pz.inject_particles_threshold("divv", 150, 1e-7, 1, 5.0, 0.0, 0.1, True,
                                          THRESHOLDS["divv"], 1, TBOX)
        assert n == 150 and pz.counts()["nptl_current"] == 120
        ref, _, _ = oracle.inject_threshold(prm, fa, oracle.particles_from_soa(g), 120, "divv", 150,
                                            1e-7, 1, 5.0, 0.0, 0.1, True, THRESHOLDS["divv"], 1,
                                            TBOX, 7.0, tag_max=20)
        g2 = pz.get_particles()
        for k in ("x", "y", "tag_injected"):
            np.testing.assert_array_equal(g2[k], ref[k], err_msg=k)
        with pytest.raises(Exception) as ei:
            pz.inject_particles_threshold("db2", 10, 1e-7, 1, 5.0, 0.0, 0.1, True, 0.1, 1, TBOX)
        assert "turbulence" in str(ei.value)
