"""sp_step_once_noalloc (no device allocation; rows staged in the library's
fixed per-GPU array) against sp_step_once through the C ABI: the same kernel
on the same rows, so the states and deltas are bit-identical.  Two calls in a
row on one context, the second with fewer rows, then a call on a second
context of the same GPU (which shares the array), and the row limit."""
import numpy as np
import pytest

from helpers import cfg2d, cfg3d, nodes_for, random_states
from sp_hip import ParkerPusher, SpError

pytestmark = pytest.mark.gpu

SQ3 = np.sqrt(3.0)
STATE_COLS = ("x", "y", "z", "p", "t", "dt")
MAX_ROWS = 65536


def ctx(c, precision):
    pz = ParkerPusher(**dict(c, precision=precision, nptl_max=1000))
    pz.upload_fields(nodes_for(c))
    pz.calc_fields_gradients()
    return pz


def rows(c, n, seed):
    a = random_states(c, n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    r = rng.uniform(-SQ3, SQ3, size=(n, 4))
    fixed = rng.integers(0, 2, n).astype(np.int8)
    return np.stack([a[k] for k in STATE_COLS], 1), r, fixed


@pytest.mark.parametrize("ndim,precision", [(2, "f64"), (2, "f32"), (3, "f64"), (3, "f32h")])
def test_noalloc_step_equals_step_once(ndim, precision):
    c = cfg2d(32, check_drift_2d=1) if ndim == 2 else cfg3d(12)
    pz = ctx(c, precision)
    for n, seed in ((MAX_ROWS, 11), (1237, 12)):
        print(f"noalloc: ndim {ndim} {precision} rows {n}", flush=True)
        st, r, fixed = rows(c, n, seed)
        s1, d1 = pz.step_once(st, r, fixed, 0.0, 0.05)
        s2, d2 = pz.step_once(st, r, fixed, 0.0, 0.05, noalloc=True)
        np.testing.assert_array_equal(s2, s1)
        np.testing.assert_array_equal(d2, d1)
        assert np.any(s2 != st)   # the rows moved
    # a second context on the same GPU stages its rows in the same array
    other = ctx(c, precision)
    st, r, fixed = rows(c, 500, 13)
    s1, d1 = pz.step_once(st, r, fixed, 0.0, 0.05)
    s2, d2 = other.step_once(st, r, fixed, 0.0, 0.05, noalloc=True)
    np.testing.assert_array_equal(s2, s1)
    np.testing.assert_array_equal(d2, d1)
    other.close()
    pz.close()


def test_noalloc_row_limit_and_empty_call():
    c = cfg2d(16)
    pz = ctx(c, "f64")
    st, r, fixed = rows(c, MAX_ROWS + 1, 21)
    with pytest.raises(SpError, match="at most 65536 rows"):
        pz.step_once(st, r, fixed, 0.0, 0.05, noalloc=True)
    s, d = pz.step_once(st[:0], r[:0], fixed[:0], 0.0, 0.05, noalloc=True)
    assert s.shape == (0, 6) and d.shape == (0, 4)
    pz.close()
