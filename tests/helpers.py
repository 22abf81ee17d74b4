"""Shared builders for the tests: configurations, analytic field snapshots and
particle states."""
from __future__ import annotations

import numpy as np

import oracle
from sp_hip import workloads
from sp_hip.pusher import CONFIG_DEFAULTS


def cfg2d(n=32, **over):
    c = dict(CONFIG_DEFAULTS)
    c.update(workloads.config_2d(n))
    c.update(over)
    return c


def cfg3d(n=16, **over):
    c = dict(CONFIG_DEFAULTS)
    c.update(dict(ndim=3, nx=n, ny=n, nz=n, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n,
                  xmin=0.0, ymin=0.0, zmin=0.0, xmax=1.0, ymax=1.0, zmax=1.0,
                  lx=1.0, ly=1.0, lz=1.0, pbcx=0, pbcy=0, pbcz=0))
    c.update(over)
    return c


def grid_shape(c):
    nzg = c["nz"] + 4 if c["ndim"] == 3 else c["nz"]
    return (nzg, c["ny"] + 4, c["nx"] + 4)


def coords(c):
    """Node positions (z, y, x) broadcastable over grid_shape(c)."""
    x = workloads.node_coords(c["nx"], c["dx"], c["xmin"])
    y = workloads.node_coords(c["ny"], c["dy"], c["ymin"])
    if c["ndim"] == 3:
        z = workloads.node_coords(c["nz"], c["dz"], c["zmin"])
    else:
        z = np.array([c["zmin"]])
    return z[:, None, None], y[None, :, None], x[None, None, :]


def nodes_from_funcs(c, funcs):
    """Evaluate 8 callables f(x, y, z) on every node -> (nzg, nyg, nxg, 8) f32."""
    z, y, x = coords(c)
    out = np.zeros(grid_shape(c) + (8,), dtype=np.float64)
    for k, f in enumerate(funcs):
        out[..., k] = np.broadcast_to(f(x, y, z), grid_shape(c))
    return out.astype(np.float32)


def smooth_nodes_3d(c, seed=7, amp_b=0.3, amp_v=0.05):
    """A smooth periodic 3D snapshot with gradients in every component."""
    rng = np.random.default_rng(seed)
    tp = 2.0 * np.pi
    ph = rng.uniform(0, tp, size=12)

    def bx(x, y, z):
        return 1.0 + amp_b * np.sin(tp * y + ph[0]) * np.cos(tp * z + ph[1])

    def by(x, y, z):
        return amp_b * np.sin(tp * z + ph[2]) * np.cos(tp * x + ph[3])

    def bz(x, y, z):
        return 0.2 + amp_b * np.sin(tp * x + ph[4]) * np.cos(tp * y + ph[5])

    def vx(x, y, z):
        return amp_v * np.sin(tp * x + ph[6]) + 0.0 * y

    def vy(x, y, z):
        return amp_v * np.sin(tp * y + ph[7]) * np.cos(tp * z + ph[8])

    def vz(x, y, z):
        return amp_v * np.cos(tp * z + ph[9]) * np.sin(tp * x + ph[10])

    def rho(x, y, z):
        return 1.0 + 0.0 * x

    def btot(x, y, z):
        return np.sqrt(bx(x, y, z) ** 2 + by(x, y, z) ** 2 + bz(x, y, z) ** 2)

    return nodes_from_funcs(c, [vx, vy, vz, rho, bx, by, bz, btot])


def nodes_for(c, seed=7):
    if c["ndim"] == 2:
        return workloads.fields_2d(c["nx"], c["ny"], seed=seed)
    return smooth_nodes_3d(c, seed=seed)


def random_states(c, n, seed=3, slack=True):
    """Random particle rows inside the (slack) box with p in [0.3, 30]."""
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=oracle.PARTICLE_DTYPE)
    pad = 0.49 if slack else 0.0
    a["x"] = rng.uniform(c["xmin"] - pad * c["dx"], c["xmax"] + pad * c["dx"], n)
    a["y"] = rng.uniform(c["ymin"] - pad * c["dy"], c["ymax"] + pad * c["dy"], n)
    if c["ndim"] == 3:
        a["z"] = rng.uniform(c["zmin"] - pad * c["dz"], c["zmax"] + pad * c["dz"], n)
    else:
        a["z"] = rng.uniform(c["zmin"], c["zmax"], n)
    a["p"] = np.exp(rng.uniform(np.log(0.3), np.log(30.0), n))
    a["t"] = rng.uniform(0.0, 0.01, n)
    a["dt"] = rng.uniform(1e-6, 1e-4, n)
    a["weight"] = 1.0
    a["count_flag"] = 1
    a["tag_injected"] = np.arange(n)
    a["tag_splitted"] = 1
    return a


def escaped_particles(c, n, seed):
    """Particles just outside a random face each (count_flag = -face), their
    other coordinates spread 5 % beyond the domain so that the local face maps'
    range checks matter; p and mu partly outside the binned ranges."""
    rng = np.random.default_rng(seed)
    nd = c["ndim"]
    a = np.zeros(n, dtype=oracle.PARTICLE_DTYPE)
    faces = rng.integers(1, 2 * nd + 1, n)
    lo = np.array([c["xmin"], c["ymin"], c["zmin"]])
    hi = np.array([c["xmax"], c["ymax"], c["zmax"]])
    pos = rng.uniform(lo - 0.05, hi + 0.05, size=(n, 3))
    if nd == 2:
        pos[:, 2] = rng.uniform(c["zmin"], c["zmax"], n)
    axis, side = (faces - 1) // 2, (faces - 1) % 2
    depth = rng.uniform(1e-6, 0.01, n)
    pos[np.arange(n), axis] = np.where(side == 0, lo[axis] - depth, hi[axis] + depth)
    a["x"], a["y"], a["z"] = pos.T
    a["p"] = np.exp(rng.uniform(np.log(5e-3), np.log(2e4), n))
    a["mu"] = rng.uniform(-1.05, 1.05, n)
    a["weight"] = 0.5 ** rng.integers(0, 4, n)
    a["count_flag"] = -faces
    a["tag_injected"] = np.arange(n)
    a["tag_splitted"] = 1
    return a


ESC_DIST = dict(npp_global=64, nmu_global=2, local_dist=1,
                local=[dict(enabled=1, npbins=16, nmu=1, rx=4, ry=4, rz=4, pmin=1e-2, pmax=1e4),
                       None, None,
                       dict(enabled=1, npbins=4, nmu=2, rx=8, ry=2, rz=3, pmin=0.1, pmax=100.0)])


def face_layout(c, dcfg, k):
    """(axis name, C shape) of local set k's face arrays, back to back."""
    L = dcfg["local"][k]
    nr = [-(-c["nx"] // L["rx"]), -(-c["ny"] // L["ry"]), -(-c["nz"] // L["rz"])]
    tail = (L["npbins"], L["nmu"])
    out = [("x", (2, nr[2], nr[1]) + tail), ("y", (2, nr[2], nr[0]) + tail)]
    if c["ndim"] == 3:
        out.append(("z", (2, nr[1], nr[0]) + tail))
    return out, nr
