"""Synthetic MHD snapshots and run settings for the benchmark configurations.

Field layout is the reference's on-disk node format: for each node the eight
fields (vx, vy, vz, rho, bx, by, bz, |B|) in FP32, x fastest, two ghost nodes
on each side of every resolved axis (reorganize_fields.py:57,193 ==
Fortran farray(1:8, -1:nx+2, -1:ny+2, ...)).  Node (Fortran) index i sits at
x = xmin + (i - 1) dx; the fields are evaluated analytically at every node,
ghosts included, and are periodic with period lx, ly (lz).

2D reconnection-like snapshot (512^2 and 2048^2 configurations)
  Bx = tanh((y - 0.25)/0.02) - tanh((y - 0.75)/0.02) - 1 + dA/dy
  By = -dA/dx,  Bz = 0.1,  A = sum of 8 Fourier modes (|k| <= 4 per axis)
  scaled to an rms |dB| of 0.01 of the guide scale;  V = 16 Fourier modes
  with random in-plane directions (compressive and solenoidal parts) scaled
  to rms |V| = 0.1.  Fields are generated in FP64 and stored in FP32.
"""
from __future__ import annotations

import numpy as np

FIELD_SEED = 20261202


def node_coords(n: int, d: float, lo: float) -> np.ndarray:
    """Positions of Fortran nodes -1 .. n+2 (array index 0 .. n+3)."""
    return lo + (np.arange(n + 4, dtype=np.float64) - 2.0) * d


def _modes(rng, nmodes, kmax, ndim):
    k = rng.integers(1, kmax + 1, size=(nmodes, ndim)) * rng.choice([-1, 1], size=(nmodes, ndim))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    amp = rng.uniform(0.5, 1.0, size=nmodes)
    return k.astype(np.float64), phase, amp


def fields_2d(nx: int, ny: int, seed: int = FIELD_SEED, lx: float = 1.0, ly: float = 1.0,
              dtype=np.float32) -> np.ndarray:
    """(1, ny+4, nx+4, 8) node fields of the 2D snapshot."""
    rng = np.random.default_rng(seed)
    x = node_coords(nx, lx / nx, 0.0)[None, :]
    y = node_coords(ny, ly / ny, 0.0)[:, None]
    tx, ty = 2.0 * np.pi / lx, 2.0 * np.pi / ly
    # vector potential perturbation: A = sum a cos(kx x + ky y + phi)
    ka, pa, aa = _modes(rng, 8, 4, 2)
    dAdx = np.zeros((y.shape[0], x.shape[1]))
    dAdy = np.zeros_like(dAdx)
    for (kx, ky), ph, a in zip(ka, pa, aa):
        s = -np.sin(tx * kx * x + ty * ky * y + ph)
        dAdx += a * tx * kx * s
        dAdy += a * ty * ky * s
    rms = np.sqrt(np.mean(dAdx[2:-2, 2:-2] ** 2 + dAdy[2:-2, 2:-2] ** 2))
    scale = 0.01 / rms
    bx = np.tanh((y - 0.25) / 0.02) - np.tanh((y - 0.75) / 0.02) - 1.0 + scale * dAdy
    by = -scale * dAdx + 0.0 * y
    bz = np.full_like(bx, 0.1)
    # velocity: 16 modes with random in-plane directions
    kv, pv, av = _modes(rng, 16, 4, 2)
    ang = rng.uniform(0.0, 2.0 * np.pi, size=16)
    vx = np.zeros_like(bx)
    vy = np.zeros_like(bx)
    for (kx, ky), ph, a, th in zip(kv, pv, av, ang):
        s = a * np.sin(tx * kx * x + ty * ky * y + ph)
        vx += np.cos(th) * s
        vy += np.sin(th) * s
    vrms = np.sqrt(np.mean(vx[2:-2, 2:-2] ** 2 + vy[2:-2, 2:-2] ** 2))
    vx *= 0.1 / vrms
    vy *= 0.1 / vrms
    out = np.zeros((1, ny + 4, nx + 4, 8), dtype=np.float64)
    out[0, ..., 0] = vx
    out[0, ..., 1] = vy
    out[0, ..., 2] = 0.0
    out[0, ..., 3] = 1.0
    out[0, ..., 4] = bx
    out[0, ..., 5] = by
    out[0, ..., 6] = bz
    out[0, ..., 7] = np.sqrt(bx * bx + by * by + bz * bz)
    return out.astype(dtype)


def config_2d(n: int, **over) -> dict:
    """sp_config keys of the 2D snapshot on an n x n grid over [0,1]^2 with the
    pusher parameters of the 512^2 / 2048^2 benchmark settings."""
    c = dict(
        ndim=2, nx=n, ny=n, nz=1, dx=1.0 / n, dy=1.0 / n, dz=1.0,
        xmin=0.0, ymin=0.0, zmin=0.0, xmax=1.0, ymax=1.0, zmax=1.0, lx=1.0, ly=1.0, lz=1.0,
        pbcx=0, pbcy=1, pbcz=0, momentum_dependency=1, mag_dependency=0, charge=-1,
        p0=1.0, pmin=1e-2, pmax=1e4, b0=1.0, gamma_turb=5.0 / 3.0, kpara0=1e-3, kret=0.01,
        dt_min_rel=1e-7, dt_max_rel=1e-3, drift1=4e7, drift2=2e8, check_drift_2d=0,
        acc_region_flag=0, time_interp=0,
    )
    c.update(over)
    return c


def interval_2d(n: int, nsteps: int, kpara0: float = 1e-3) -> float:
    """MHD interval dtf that holds `nsteps` diffusion-limited steps of a
    particle at p0: dt = (0.5 dx)^2 / (2 kpara0)."""
    dx = 1.0 / n
    return nsteps * (0.5 * dx) ** 2 / (2.0 * kpara0)


def dist_config_2d(n: int, map_bins: int = 128, npbins1: int = 16) -> dict:
    """Global f(p): 256 log bins on [1e-2, 1e4]; local set 1: a map_bins^2
    spatial map with npbins1 momentum bins on the same range."""
    r = n // map_bins
    return dict(
        npp_global=256, nmu_global=1, local_dist=1,
        local=[dict(enabled=1, npbins=npbins1, nmu=1, rx=r, ry=r, rz=1, pmin=1e-2, pmax=1e4),
               None, None, None],
    )


# ---------------------------------------------------------------------------
# 3D turbulence snapshot (512^3 configuration)
#   B = B0 x + dB, dB solenoidal with E(k) ~ k^-5/3 on shells 2 <= |k| <= 64,
#   rms |dB| = 0.5 B0;  V = compressive (curl-free) turbulence of the same
#   spectrum with rms |V| = 0.05, plus a smooth step of Vx from U1 to U1/3
#   across a central slab (divV < 0 there: a compression of ratio 3).  The
#   turbulence is periodic on the unit cube; the Vx step is evaluated at every
#   node, ghosts included (the box is open along x for particles).
SLAB_U1 = 0.15
SLAB_RATIO = 3.0
SLAB_WIDTH = 0.02


def _turbulence(n, rng, kmin, kmax, solenoidal, workers):
    """Three real components on an n^3 periodic grid (index order z, y, x)
    whose Fourier amplitudes follow |a_k|^2 ~ k^-11/3 (E(k) ~ k^-5/3) on
    kmin <= |k| <= kmax, projected on the plane normal to k (solenoidal) or
    on k (compressive).  Returned unscaled, FP64."""
    from scipy import fft as sfft

    kk = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    # only the band holds modes: draw them sparse, then scatter
    lim = int(kmax)
    ax = np.concatenate([np.arange(0, lim + 1), np.arange(-lim, 0)])
    kz, ky, kx = np.meshgrid(ax, ax, np.arange(0, lim + 1), indexing="ij")
    kabs = np.sqrt(kx * kx + ky * ky + kz * kz)
    sel = (kabs >= kmin) & (kabs <= kmax)
    kx, ky, kz, kabs = kx[sel], ky[sel], kz[sel], kabs[sel]
    m = kabs.size
    xi = rng.standard_normal((3, m)) + 1j * rng.standard_normal((3, m))
    khat = np.stack([kx, ky, kz]) / kabs
    kdotxi = (khat * xi).sum(axis=0)
    if solenoidal:
        vec = xi - khat * kdotxi
    else:
        vec = khat * kdotxi
    vec *= kabs ** (-11.0 / 6.0)
    iz = np.where(kz >= 0, kz, kz + n)
    iy = np.where(ky >= 0, ky, ky + n)
    out = []
    for c in range(3):
        spec = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
        spec[iz, iy, kx] = vec[c]
        out.append(sfft.irfftn(spec, s=(n, n, n), workers=workers))
        del spec
    del kk, kr
    return out


def _ghosted(a, n):
    """Periodic (n+4)^3 copy of an n^3 array (node i <-> grid index i - 2)."""
    idx = (np.arange(n + 4) - 2) % n
    return a[np.ix_(idx, idx, idx)]


def fields_3d(n: int, seed: int = FIELD_SEED, kmin: float = 2.0, kmax: float = 64.0,
              db_rms: float = 0.5, v_rms: float = 0.05, workers: int = -1) -> np.ndarray:
    """(n+4, n+4, n+4, 8) FP32 node fields of the 3D snapshot on [0,1]^3."""
    rng = np.random.default_rng(seed + 3)
    kmax = min(kmax, n // 2 - 1)
    out = np.empty((n + 4, n + 4, n + 4, 8), dtype=np.float32)
    db = _turbulence(n, rng, kmin, kmax, True, workers)
    s = db_rms / np.sqrt(sum(np.mean(c * c) for c in db))
    btot2 = np.zeros((n + 4,) * 3, dtype=np.float64)
    for c in range(3):
        g = _ghosted(db[c], n) * s
        if c == 0:
            g += 1.0
        out[..., 4 + c] = g
        btot2 += g * g
        db[c] = None
    out[..., 7] = np.sqrt(btot2)
    del btot2
    v = _turbulence(n, rng, kmin, kmax, False, workers)
    s = v_rms / np.sqrt(sum(np.mean(c * c) for c in v))
    x = node_coords(n, 1.0 / n, 0.0)
    step = SLAB_U1 - 0.5 * SLAB_U1 * (1.0 - 1.0 / SLAB_RATIO) * (1.0 + np.tanh((x - 0.5) / SLAB_WIDTH))
    for c in range(3):
        g = _ghosted(v[c], n) * s
        if c == 0:
            g += step[None, None, :]
        out[..., c] = g
        v[c] = None
    out[..., 3] = 1.0
    return out


def config_3d(n: int, **over) -> dict:
    """sp_config keys of the 3D snapshot on an n^3 grid over [0,1]^3: open
    along x (the compression slab's normal), periodic along y and z.  The
    particles are sorted once per interval, before the mover (with the
    8^3-node record bricks this measured faster on MI355X for 512^3 and 1e8
    particles than re-sorting within the interval every 100 or 200 steps;
    results do not depend on it)."""
    c = dict(
        ndim=3, nx=n, ny=n, nz=n, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n,
        xmin=0.0, ymin=0.0, zmin=0.0, xmax=1.0, ymax=1.0, zmax=1.0, lx=1.0, ly=1.0, lz=1.0,
        pbcx=1, pbcy=0, pbcz=0, momentum_dependency=1, mag_dependency=0, charge=-1,
        p0=1.0, pmin=1e-2, pmax=1e4, b0=1.0, gamma_turb=5.0 / 3.0, kpara0=1e-3, kret=0.01,
        dt_min_rel=1e-7, dt_max_rel=1e-3, drift1=4e7, drift2=2e8, check_drift_2d=0,
        acc_region_flag=0, time_interp=0, sort_steps=0,
    )
    c.update(over)
    return c


def interval_3d(n: int, nsteps: int, kpara0: float = 1e-3) -> float:
    """dtf holding `nsteps` diffusion-limited steps at p0 on an n^3 grid."""
    return interval_2d(n, nsteps, kpara0)


def dist_config_3d(n: int, map_bins: int = 64, npbins1: int = 16) -> dict:
    """Global f(p) as in 2D; local set 1: a map_bins^3 spatial map."""
    r = max(1, n // map_bins)
    return dict(
        npp_global=256, nmu_global=1, local_dist=1,
        local=[dict(enabled=1, npbins=npbins1, nmu=1, rx=r, ry=r, rz=r, pmin=1e-2, pmax=1e4),
               None, None, None],
    )


# ---------------------------------------------------------------------------
# Batched parameter sweep over the 512^2 2D snapshot
#   kpara0 log-uniform in [1e-4, 1e-1], kret log-uniform in [1e-3, 1], the
#   momentum exponent 3 - gamma_turb uniform in [1, 2], injection in one of
#   four regions, particle counts log-uniform in [1e5, 1e7] per variant.
SWEEP_SEED = 20261203
SWEEP_REGIONS = (
    (0.0, 0.20, 0.0, 1.0, 0.30, 0.0),   # band around the current sheet at y = 0.25
    (0.0, 0.70, 0.0, 1.0, 0.80, 0.0),   # band around the current sheet at y = 0.75
    (0.0, 0.45, 0.0, 1.0, 0.55, 0.0),   # band between the sheets
    (0.0, 0.00, 0.0, 1.0, 1.00, 0.0),   # the whole box
)


def sweep_variants(nvar: int = 64, nmin: float = 1e5, nmax: float = 1e7,
                   seed: int = SWEEP_SEED) -> dict:
    """Parameters of a batched sweep: arrays kpara0, kret, pindex, gamma_turb,
    region (0..3), nptl, and boxes (nvar, 6)."""
    rng = np.random.default_rng(seed)
    kpara0 = 10.0 ** rng.uniform(-4.0, -1.0, nvar)
    kret = 10.0 ** rng.uniform(-3.0, 0.0, nvar)
    pindex = rng.uniform(1.0, 2.0, nvar)
    region = rng.integers(0, len(SWEEP_REGIONS), nvar)
    nptl = np.round(10.0 ** rng.uniform(np.log10(nmin), np.log10(nmax), nvar)).astype(np.int64)
    boxes = np.array([SWEEP_REGIONS[r] for r in region], dtype=np.float64)
    return dict(kpara0=kpara0, kret=kret, pindex=pindex, gamma_turb=3.0 - pindex, region=region,
                nptl=nptl, boxes=boxes)


def sweep_interval(n: int, kpara_max: float, nsteps: int = 1000) -> float:
    """dtf holding `nsteps` diffusion-limited steps at p0 for the most
    diffusive variant (the others take dt_max = dt_max_rel dtf steps)."""
    return interval_2d(n, nsteps, kpara_max)
