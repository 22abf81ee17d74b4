"""Distributions of escaped particles on the GPU, through the C ABI: the
removal pass bins every particle it removes with an escape flag; the bins
match the oracle's calc_escaped_distributions bit for bit, both for
particles handed in already escaped and for those the mover flags."""
import numpy as np
import pytest

import oracle
from helpers import ESC_DIST, cfg2d, cfg3d, escaped_particles, face_layout, nodes_for
from test_gpu_parity import by_id, make

pytestmark = pytest.mark.gpu


def _gpu_sets(c, sets):
    """Flatten the binding's per-face arrays in the C ABI's order."""
    out = []
    for k, faces in enumerate(sets):
        if faces is None:
            out.append(None)
            continue
        layout, _ = face_layout(c, ESC_DIST, k)
        out.append(np.concatenate([faces[name].ravel() for name, _ in layout]))
    return out


@pytest.mark.parametrize("ndim", [2, 3])
def test_escaped_particles_binned_bit_exact(ndim):
    c = cfg2d(32) if ndim == 2 else cfg3d(12)
    pz, _ = make(c, nodes_for(c), nptl_max=40000)
    a = escaped_particles(c, 30000, seed=30 + ndim)
    a["count_flag"][::50] = 0          # removed, not escaped
    pz.set_particles(oracle.soa_from_particles(a))
    pz.init_particle_distributions(**ESC_DIST, escaped_dist=1)
    st = pz.particle_mover(0.0, 0.05, 1, 10)
    assert st["nptl_current"] == 0
    assert st["nptl_escaped"] == int(np.sum(a["count_flag"] < 0))
    fe, sets = pz.calc_escaped_distributions()
    fe_ref, locs_ref = oracle.escaped_distributions(oracle.params(c), ESC_DIST, a)
    np.testing.assert_array_equal(fe.ravel(), fe_ref)
    got = _gpu_sets(c, sets)
    for k in (0, 3):
        np.testing.assert_array_equal(got[k], locs_ref[k], err_msg=f"set {k + 1}")
    assert sets[1] is None and sets[2] is None
    # the call clears the bins
    fe2, sets2 = pz.calc_escaped_distributions()
    assert not fe2.any() and not sets2[0]["x"].any()
    pz.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_escaped_distributions_follow_the_mover(ndim):
    if ndim == 2:
        c = cfg2d(32, pbcx=1, pbcy=1, dt_max_rel=1e-2)
    else:
        c = cfg3d(12, pbcx=1, pbcz=1, dt_max_rel=1e-2)
    n = 3000
    pz, _ = make(c, nodes_for(c), nptl_max=4 * n)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes_for(c))
    box = (c["xmin"], c["ymin"], c["zmin"], c["xmax"], c["ymax"], c["zmax"])
    pz.inject_particles_spatial_uniform(n, 1e-7, 1, 5.0, 0.0, 0.2, box)
    ref, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), 4 * n, n, 1e-7, 1, 5.0, 0.0,
                           0.2, box, 7.0)
    pz.init_particle_distributions(**ESC_DIST, escaped_dist=1)
    esc_total = None
    leak = 0.0
    for it in range(2):
        st = pz.particle_mover(0.2 * it, 0.2, 1, 10)
        ref, acc, removed = oracle.mover(prm, fa, ref, 0.2 * it, 0.2, nthreads=8,
                                         return_removed=True)
        fe, sets = pz.calc_escaped_distributions()
        fe_ref, locs_ref = oracle.escaped_distributions(prm, ESC_DIST, removed)
        np.testing.assert_array_equal(fe.ravel(), fe_ref, err_msg=f"interval {it}")
        got = _gpu_sets(c, sets)
        for k in (0, 3):
            np.testing.assert_array_equal(got[k], locs_ref[k], err_msg=f"interval {it} set {k + 1}")
        # every escaped particle is in range here, so the bins hold the leak
        assert fe.sum() == st["leak"] - leak
        leak = st["leak"]
        esc_total = (esc_total or 0) + int(np.sum(removed["count_flag"] < 0))
    assert esc_total > 20 and st["nptl_escaped"] == esc_total
    assert len(by_id(pz.get_particles())["x"]) == len(ref)
    pz.close()


def test_escaped_distributions_need_the_flag():
    c = cfg2d(16)
    pz, _ = make(c, nodes_for(c), nptl_max=100)
    pz.init_particle_distributions(npp_global=16)
    with pytest.raises(RuntimeError):
        pz.calc_escaped_distributions()
    pz.close()
