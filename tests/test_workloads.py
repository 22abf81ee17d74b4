"""Properties of the synthetic snapshots the benchmarks run on."""
import numpy as np
import pytest

from sp_hip import workloads


def _inner(a):
    return a[2:-2, 2:-2, 2:-2].astype(np.float64)


def test_2d_snapshot_construction():
    n = 128
    a = workloads.fields_2d(n, n)[0].astype(np.float64)
    assert a.shape == (n + 4, n + 4, 8)
    # periodic ghosts: node i and i + n coincide
    np.testing.assert_allclose(a[:, 0:4], a[:, n:n + 4], atol=1e-5)
    np.testing.assert_allclose(a[0:4], a[n:n + 4], atol=1e-5)
    inner = a[2:-2, 2:-2]
    v = inner[..., 0:2]
    assert np.sqrt((v ** 2).sum(-1).mean()) == pytest.approx(0.1, rel=1e-5)
    assert np.all(inner[..., 6] == np.float32(0.1))
    btot = np.sqrt(inner[..., 4] ** 2 + inner[..., 5] ** 2 + inner[..., 6] ** 2)
    np.testing.assert_allclose(inner[..., 7], btot, rtol=1e-6)
    # current sheets: Bx reverses across y = 0.25 and y = 0.75
    y = workloads.node_coords(n, 1.0 / n, 0.0)[2:-2]
    bx = inner[..., 4].mean(axis=1)
    assert bx[np.argmin(abs(y - 0.5))] > 0.9 and bx[np.argmin(abs(y - 0.0))] < -0.9


@pytest.fixture(scope="module")
def snap3d():
    n = 48
    return n, workloads.fields_3d(n, kmax=20)


def _spectral(comp3):
    return [np.fft.fftn(c) for c in comp3]


def test_3d_turbulence_amplitudes_and_ghosts(snap3d):
    n, a = snap3d
    assert a.shape == (n + 4,) * 3 + (8,) and a.dtype == np.float32
    inner = _inner(a)
    db = inner[..., 4:7].copy()
    db[..., 0] -= 1.0
    assert np.sqrt((db ** 2).sum(-1).mean()) == pytest.approx(0.5, rel=1e-5)
    np.testing.assert_allclose(inner[..., 7], np.sqrt((inner[..., 4:7] ** 2).sum(-1)), rtol=1e-6)
    assert np.all(inner[..., 3] == 1.0)
    # B and the transverse velocity are periodic on every axis (ghosts wrap)
    for c in (1, 2, 4, 5, 6):
        f = a[..., c]
        np.testing.assert_array_equal(f[:, :, 0:4], f[:, :, n:n + 4])
        np.testing.assert_array_equal(f[:, 0:4], f[:, n:n + 4])
        np.testing.assert_array_equal(f[0:4], f[n:n + 4])


def test_3d_magnetic_turbulence_is_solenoidal_with_kolmogorov_spectrum(snap3d):
    n, a = snap3d
    inner = _inner(a)
    bk = _spectral([inner[..., 4] - 1.0, inner[..., 5], inner[..., 6]])
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    div = kx * bk[0] + ky * bk[1] + kz * bk[2]
    kabs = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    mag = kabs * np.sqrt(sum(np.abs(b) ** 2 for b in bk))
    # k . B_k vanishes up to the FP32 storage rounding
    assert np.abs(div).max() < 1e-5 * mag.max()
    # band limits and the shell spectrum E(k) ~ k^-5/3
    power = sum(np.abs(b) ** 2 for b in bk)
    assert power[(kabs < 2 - 1e-9) | (kabs > 20 + 1e-9)].max() < 1e-8 * power.max()
    shells = np.arange(2, 21)
    e = np.array([power[(kabs >= s - 0.5) & (kabs < s + 0.5)].sum() for s in shells])
    slope = np.polyfit(np.log(shells[1:-1]), np.log(e[1:-1]), 1)[0]
    assert slope == pytest.approx(-5.0 / 3.0, abs=0.25)


def test_3d_velocity_is_compressive_plus_slab(snap3d):
    n, a = snap3d
    inner = _inner(a)
    x = workloads.node_coords(n, 1.0 / n, 0.0)[2:-2]
    u1, r, w = workloads.SLAB_U1, workloads.SLAB_RATIO, workloads.SLAB_WIDTH
    step = u1 - 0.5 * u1 * (1.0 - 1.0 / r) * (1.0 + np.tanh((x - 0.5) / w))
    vt = [inner[..., 0] - step[None, None, :], inner[..., 1], inner[..., 2]]
    assert np.sqrt(sum(np.mean(c ** 2) for c in vt)) == pytest.approx(0.05, rel=1e-4)
    vk = _spectral(vt)
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    curl = [ky * vk[2] - kz * vk[1], kz * vk[0] - kx * vk[2], kx * vk[1] - ky * vk[0]]
    kabs = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    mag = kabs * np.sqrt(sum(np.abs(v) ** 2 for v in vk))
    assert max(np.abs(c).max() for c in curl) < 1e-4 * mag.max()
    # the slab: upstream U1, downstream U1 / r, divV < 0 inside
    assert step[0] == pytest.approx(u1, rel=1e-6) and step[-1] == pytest.approx(u1 / r, rel=1e-6)
    # the ghost layers along x carry the step continued, not wrapped
    xg = workloads.node_coords(n, 1.0 / n, 0.0)
    stepg = u1 - 0.5 * u1 * (1.0 - 1.0 / r) * (1.0 + np.tanh((xg - 0.5) / w))
    vx_mean = a[2:-2, 2:-2, :, 0].astype(np.float64).mean(axis=(0, 1))
    np.testing.assert_allclose(vx_mean, stepg, atol=0.02)
