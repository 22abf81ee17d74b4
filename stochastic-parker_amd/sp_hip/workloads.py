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
