"""Pins the oracle's field preparation and gather with closed forms:
central and one-sided differences are exact on quadratics, bilinear and
trilinear interpolation reproduce affine fields and node values."""
import ctypes as C

import numpy as np
import pytest

import oracle
from helpers import cfg2d, cfg3d, coords, grid_shape


def _grid_index_fields(c, funcs):
    """Node fields as functions of the Fortran grid index (exact integers)."""
    shp = grid_shape(c)
    k = np.arange(shp[0])[:, None, None] - (2 if c["ndim"] == 3 else 0)
    j = np.arange(shp[1])[None, :, None] - 2
    i = np.arange(shp[2])[None, None, :] - 2
    out = np.zeros(shp + (8,), dtype=np.float64)
    for q, f in enumerate(funcs):
        out[..., q] = np.broadcast_to(f(i, j, k), shp)
    return out.astype(np.float32), (i, j, k)


@pytest.mark.parametrize("ndim", [2, 3])
def test_gradients_exact_on_integer_quadratics(ndim):
    # dx a power of two and integer-valued quadratics of the grid index: every
    # FP32 difference and the FP64 scaling are exact, so the stencil output
    # must equal the analytic derivative exactly -- interior (central) and
    # both edges (second-order one-sided).
    n = 16
    c = cfg2d(n) if ndim == 2 else cfg3d(n)
    funcs = [
        lambda i, j, k: i * i + 0 * j,
        lambda i, j, k: 3 * j * j - i,
        lambda i, j, k: i * j + k * k,
        lambda i, j, k: 5 + 0 * i,
        lambda i, j, k: 2 * i + 7 * j - 3 * k,
        lambda i, j, k: i * i - j * j + 2 * k,
        lambda i, j, k: j * k + i,
        lambda i, j, k: (i + j) * (i - k),
    ]
    # analytic d/di, d/dj, d/dk of each
    dfun = [
        (lambda i, j, k: 2 * i, lambda i, j, k: 0 * j, lambda i, j, k: 0 * k),
        (lambda i, j, k: -1 + 0 * i, lambda i, j, k: 6 * j, lambda i, j, k: 0 * k),
        (lambda i, j, k: j + 0 * i, lambda i, j, k: i + 0 * j, lambda i, j, k: 2 * k),
        (lambda i, j, k: 0 * i, lambda i, j, k: 0 * j, lambda i, j, k: 0 * k),
        (lambda i, j, k: 2 + 0 * i, lambda i, j, k: 7 + 0 * j, lambda i, j, k: -3 + 0 * k),
        (lambda i, j, k: 2 * i, lambda i, j, k: -2 * j, lambda i, j, k: 2 + 0 * k),
        (lambda i, j, k: 1 + 0 * i, lambda i, j, k: k + 0 * j, lambda i, j, k: j + 0 * k),
        (lambda i, j, k: 2 * i + j - k, lambda i, j, k: i - k, lambda i, j, k: -(i + j)),
    ]
    nodes, (i, j, k) = _grid_index_fields(c, funcs)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes)
    shp = grid_shape(c)
    inv = (1.0 / c["dx"], 1.0 / c["dy"], 1.0 / c["dz"])
    for q in range(8):
        for a in range(3):
            got = fa[..., 8 + 3 * q + a]
            if a == 2 and ndim == 2:
                exp = np.zeros(shp)
            else:
                exp = np.broadcast_to(dfun[q][a](i, j, k), shp) * inv[a]
            np.testing.assert_array_equal(got, exp.astype(np.float32), err_msg=f"field {q} axis {a}")
    # the node fields themselves are untouched
    np.testing.assert_array_equal(fa[..., :8], nodes)


def test_gradients_fp32_rounding_path():
    # generic (non-dyadic) data: the stencil is the FP32 difference scaled in
    # FP64 and rounded once to FP32, so the error to the exact derivative of
    # the stored values is at most 0.5 ulp of the result plus the FP32
    # difference rounding.
    c = cfg2d(24)
    rng = np.random.default_rng(5)
    nodes = rng.uniform(-1, 1, size=grid_shape(c) + (8,)).astype(np.float32)
    fa = oracle.gradients(oracle.params(c), nodes)
    f = nodes.astype(np.float64)
    exact = (f[:, :, 2:, :] - f[:, :, :-2, :]) * (0.5 / c["dx"])
    got = fa[:, :, 1:-1, 8::3].astype(np.float64)
    tol = 2 * np.finfo(np.float32).eps * (np.abs(exact) + 2 * np.abs(f[:, :, 1:-1]) * 0.5 / c["dx"])
    assert np.all(np.abs(got - exact) <= tol)


def _interp(prm, fa, x, y, z, rt=0.0):
    lib = oracle.load()
    pos = (C.c_int * 3)()
    w = np.zeros(8)
    nd = prm.ndim
    px = (x - prm.xmin) / prm.dx
    py = (y - prm.ymin) / prm.dy
    pz = (z - prm.zmin) / prm.dz
    lib.orc_get_interp_paramters(C.byref(prm), px, py, pz, pos, oracle.dptr(w))
    out = np.zeros(32)
    lib.orc_interp_fields(C.byref(prm), oracle.fptr(fa), oracle.fptr(fa), pos, oracle.dptr(w), rt,
                          oracle.dptr(out))
    return list(pos), w, out


@pytest.mark.parametrize("ndim", [2, 3])
def test_interpolation_reproduces_affine_fields(ndim):
    c = cfg2d(16) if ndim == 2 else cfg3d(8)
    funcs = [lambda i, j, k, a=a: (a + 1) * i + (a + 3) * j + (2 * a + 1) * k + a for a in range(8)]
    nodes, _ = _grid_index_fields(c, funcs)
    fa = oracle.farray_from_nodes8(nodes)
    prm = oracle.params(c)
    rng = np.random.default_rng(11)
    for _ in range(200):
        x = rng.uniform(-0.5 * c["dx"], 1 + 0.49 * c["dx"])
        y = rng.uniform(-0.5 * c["dy"], 1 + 0.49 * c["dy"])
        z = rng.uniform(-0.5 * c["dz"], 1 + 0.49 * c["dz"]) if ndim == 3 else 0.0
        pos, w, out = _interp(prm, fa, x, y, z)
        assert abs(w.sum() - 1.0) < 1e-14
        gi, gj = x / c["dx"], y / c["dy"]
        gk = z / c["dz"] if ndim == 3 else 0.0
        for a in range(8):
            exp = (a + 1) * gi + (a + 3) * gj + (2 * a + 1) * gk + a
            assert out[a] == pytest.approx(exp, rel=1e-13, abs=1e-11)


def test_interpolation_at_nodes_and_slack_cells():
    c = cfg2d(16)
    rng = np.random.default_rng(2)
    nodes = rng.uniform(-1, 1, size=grid_shape(c) + (8,)).astype(np.float32)
    fa = oracle.farray_from_nodes8(nodes)
    prm = oracle.params(c)
    # exactly on Fortran node (i, j) = (5, 9): weight 1 on that node
    pos, w, out = _interp(prm, fa, 4 * c["dx"], 8 * c["dy"], 0.0)
    assert pos[:2] == [5, 9]
    assert w[0] == 1.0 and np.all(w[1:] == 0.0)
    np.testing.assert_array_equal(out[:8], nodes[0, 9 + 1, 5 + 1, :].astype(np.float64))
    # inside the lower half-cell slack band the ghost node 0 is used
    pos, w, _ = _interp(prm, fa, -0.25 * c["dx"], 0.5 * c["dy"], 0.0)
    assert pos[0] == 0
    assert w[1] / (w[0] + w[1]) == pytest.approx(0.75, abs=1e-12)


def test_time_interpolation_blend():
    c = cfg2d(8, time_interp=1)
    rng = np.random.default_rng(4)
    n1 = rng.uniform(-1, 1, size=grid_shape(c) + (8,)).astype(np.float32)
    n2 = rng.uniform(-1, 1, size=grid_shape(c) + (8,)).astype(np.float32)
    f1 = oracle.farray_from_nodes8(n1)
    f2 = oracle.farray_from_nodes8(n2)
    prm = oracle.params(c)
    lib = oracle.load()
    pos = (C.c_int * 3)(3, 4, 1)
    w = np.array([0.1, 0.2, 0.3, 0.4, 0, 0, 0, 0])
    out = np.zeros(32)
    lib.orc_interp_fields(C.byref(prm), oracle.fptr(f1), oracle.fptr(f2), pos, oracle.dptr(w), 0.25,
                          oracle.dptr(out))
    a = np.zeros(32)
    b = np.zeros(32)
    prm0 = oracle.params(cfg2d(8))
    lib.orc_interp_fields(C.byref(prm0), oracle.fptr(f1), None, pos, oracle.dptr(w), 0.0, oracle.dptr(a))
    lib.orc_interp_fields(C.byref(prm0), oracle.fptr(f2), None, pos, oracle.dptr(w), 0.0, oracle.dptr(b))
    np.testing.assert_allclose(out, 0.75 * a + 0.25 * b, rtol=1e-15, atol=1e-15)
