"""Host-side mirror of the reference's pusher modules over libsp_hip.so.

``ParkerPusher`` exposes the procedures the reference driver calls
(stochastic-mhd.f90:397-566) with the reference's names and argument meaning:

=====================================  ==============================================
reference procedure                    method
=====================================  ==============================================
read_field_data_parallel               upload_fields
calc_fields_gradients                  calc_fields_gradients
copy_fields                            copy_fields
inject_particles_spatial_uniform       inject_particles_spatial_uniform
particle_mover                         particle_mover
split_particle                         split_particle
init_particle_distributions            init_particle_distributions
calc_particle_distributions            calc_particle_distributions
quick_check / get_pmax_global          quick_check
=====================================  ==============================================

Failures raise ``SpError`` with the library's message; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (PARTICLE_FIELDS, SpConfig, SpDistConfig, SpMoverStats, SpParticles,
                   SpTuneResult, SpVariant, check, dptr)

CONFIG_DEFAULTS = dict(
    ndim=2, nx=64, ny=64, nz=1, dx=1.0 / 64, dy=1.0 / 64, dz=1.0,
    xmin=0.0, ymin=0.0, zmin=0.0, xmax=1.0, ymax=1.0, zmax=1.0, lx=1.0, ly=1.0, lz=1.0,
    pbcx=0, pbcy=0, pbcz=0, momentum_dependency=1, mag_dependency=0, charge=-1,
    p0=1.0, pmin=1e-2, pmax=1e4, b0=1.0, gamma_turb=5.0 / 3.0, kpara0=1e-3, kret=0.01,
    dt_min_rel=1e-7, dt_max_rel=1e-3, drift1=4e7, drift2=2e8, check_drift_2d=0,
    acc_region_flag=0, acc_region=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), time_interp=0,
    precision=_abi.SP_PREC_F64, seed=20261202, stream=0, sort_every=1, sort_steps=0, reserved0=0,
    nptl_max=1000, dpp_wave=0, dpp_shear=0, weak_scattering=0, nlgc=0, tau0=0.0,
    kperp_kpara=0.01,
)


def make_config(**kw) -> SpConfig:
    d = dict(CONFIG_DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError(f"unknown config keys: {sorted(unknown)}")
    d.update(kw)
    if isinstance(d["precision"], str):
        d["precision"] = {"f64": _abi.SP_PREC_F64, "f32": _abi.SP_PREC_F32,
                          "f32h": _abi.SP_PREC_F32H}[d["precision"]]
    c = SpConfig()
    for k, v in d.items():
        if k == "acc_region":
            for i in range(6):
                c.acc_region[i] = float(v[i])
        else:
            setattr(c, k, v)
    return c


def empty_particles(n: int) -> dict:
    return {name: np.zeros(n, dtype=dt) for name, dt in PARTICLE_FIELDS}


def _soa_struct(arrs: dict) -> SpParticles:
    s = SpParticles()
    for name, dt in PARTICLE_FIELDS:
        a = arrs.get(name)
        if a is not None:
            ct = np.ctypeslib.as_ctypes_type(dt)
            setattr(s, name, a.ctypes.data_as(C.POINTER(ct)))
    return s


class ParkerPusher:
    def __init__(self, **config):
        self.lib = _abi.load()
        self.cfg = make_config(**config)
        ctx = C.c_void_p()
        check(self.lib, "sp_ctx_create", self.lib.sp_ctx_create(C.byref(self.cfg), C.byref(ctx)))
        self.ctx = ctx
        nxg = self.cfg.nx + 4
        nyg = self.cfg.ny + 4
        nzg = self.cfg.nz + 4 if self.cfg.ndim == 3 else self.cfg.nz
        self.shape = (nzg, nyg, nxg)
        self._dist_dims = None
        self._nvar = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.sp_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        check(self.lib, name, getattr(self.lib, name)(self.ctx, *args))

    # -- fields -----------------------------------------------------------
    def upload_fields(self, nodes8: np.ndarray, slot: int = 0):
        """read_field_data_parallel: nodes8 is (nzg, nyg, nxg, 8) float32."""
        a = np.ascontiguousarray(nodes8, dtype=np.float32)
        if a.shape != self.shape + (8,):
            raise ValueError(f"nodes8 must have shape {self.shape + (8,)}, got {a.shape}")
        self._call("sp_fields_upload", slot, a.ctypes.data_as(C.POINTER(C.c_float)), a.size)

    def calc_fields_gradients(self, slot: int = 0, return_farray: bool = False):
        out = None
        ptr = None
        if return_farray:
            out = np.zeros(self.shape + (32,), dtype=np.float32)
            ptr = out.ctypes.data_as(C.POINTER(C.c_float))
        self._call("sp_fields_gradients", slot, ptr)
        return out

    def copy_fields(self):
        self._call("sp_fields_copy")

    # -- particles --------------------------------------------------------
    def set_particles(self, arrs: dict):
        n = len(arrs["x"])
        full = {name: np.ascontiguousarray(arrs[name], dtype=dt) for name, dt in PARTICLE_FIELDS}
        s = _soa_struct(full)
        self._call("sp_particles_upload", C.byref(s), n)

    def counts(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        t = C.c_int32()
        self._call("sp_particles_counts", C.byref(a), C.byref(b), C.byref(c), C.byref(t))
        return dict(nptl_current=a.value, nptl_split=b.value, nptl_escaped=c.value, tag_max=t.value)

    def get_particles(self) -> dict:
        n = self.counts()["nptl_current"]
        arrs = empty_particles(n)
        s = _soa_struct(arrs)
        got = C.c_int64()
        self._call("sp_particles_download", C.byref(s), n, C.byref(got))
        return arrs

    def inject_particles_spatial_uniform(self, nptl, dt, dist_flag, particle_v0, t0, dtf,
                                         part_box, power_index=7.0):
        box = np.ascontiguousarray(part_box, dtype=np.float64)
        self._call("sp_inject_particles_spatial_uniform", int(nptl), float(dt), int(dist_flag),
                   float(particle_v0), float(t0), float(dtf), dptr(box), float(power_index))

    def ncells_above_threshold(self, quantity, threshold, part_box) -> int:
        """get_ncells_large_{jz,absj,divv,rho}; quantity is one of 'jz', 'absj',
        'divv', 'rho'."""
        box = np.ascontiguousarray(part_box, dtype=np.float64)
        n = C.c_int64(0)
        self._call("sp_ncells_above_threshold", _abi.SP_INJECT[quantity], float(threshold),
                   dptr(box), C.byref(n))
        return n.value

    def inject_particles_threshold(self, quantity, nptl, dt, dist_flag, particle_v0, t0, dtf,
                                   inject_same_nptl, threshold, ncells_norm, part_box,
                                   power_index=7.0) -> int:
        """inject_particles_at_large_{jz,absj,divv,rho}; returns the number
        injected."""
        box = np.ascontiguousarray(part_box, dtype=np.float64)
        n = C.c_int64(0)
        self._call("sp_inject_particles_threshold", _abi.SP_INJECT[quantity], int(nptl), float(dt),
                   int(dist_flag), float(particle_v0), float(t0), float(dtf),
                   int(bool(inject_same_nptl)), float(threshold), int(ncells_norm), dptr(box),
                   float(power_index), C.byref(n))
        return n.value

    def particle_mover(self, t0, dtf, num_fine_steps=1, nsteps_interval=10) -> dict:
        st = SpMoverStats()
        self._call("sp_particle_mover", float(t0), float(dtf), int(num_fine_steps),
                   int(nsteps_interval), C.byref(st))
        return {k: getattr(st, k) for k, _ in SpMoverStats._fields_}

    def autotune_mover(self, t0, dtf, num_fine_steps=1, nsteps_interval=10) -> dict:
        """Times mover launch settings over [t0, t0 + dtf] and keeps the
        fastest; the particles are left as they were."""
        r = SpTuneResult()
        self._call("sp_autotune_mover", float(t0), float(dtf), int(num_fine_steps),
                   int(nsteps_interval), C.byref(r))
        return {k: getattr(r, k) for k, _ in SpTuneResult._fields_}

    def split_particle(self, split_ratio, pmin_split):
        self._call("sp_split_particle", float(split_ratio), float(pmin_split))

    def sort_particles(self):
        self._call("sp_sort_particles")

    def step_once(self, state: np.ndarray, rands: np.ndarray, fixed: np.ndarray, t0: float,
                  dtf: float, noalloc: bool = False):
        """One step per row (sp_step_once); noalloc uses sp_step_once_noalloc,
        which stages the rows in the library's fixed per-GPU array."""
        st = np.ascontiguousarray(state, dtype=np.float64).copy()
        r = np.ascontiguousarray(rands, dtype=np.float64)
        fx = np.ascontiguousarray(fixed, dtype=np.int8)
        n = st.shape[0]
        d = np.zeros((n, 4), dtype=np.float64)
        self._call("sp_step_once_noalloc" if noalloc else "sp_step_once", n, float(t0), float(dtf),
                   dptr(st), dptr(r), fx.ctypes.data_as(C.POINTER(C.c_int8)), dptr(d))
        return st, d

    # -- diagnostics ------------------------------------------------------
    def init_particle_distributions(self, npp_global=256, nmu_global=1, local_dist=0,
                                    local=(), escaped_dist=0):
        dc = SpDistConfig()
        dc.npp_global = npp_global
        dc.nmu_global = nmu_global
        dc.local_dist = int(local_dist)
        dc.escaped_dist = int(escaped_dist)
        for k, L in enumerate(local):
            if L is None:
                continue
            for key in ("enabled", "npbins", "nmu", "rx", "ry", "rz", "pmin", "pmax"):
                setattr(dc.local[k], key, L[key])
        self._call("sp_distributions_init", C.byref(dc))
        self._npp = npp_global * nmu_global
        self._npp_p = npp_global
        dims = []
        for k in range(4):
            d = (C.c_int64 * 5)()
            self._call("sp_local_dist_dims", k + 1, d)
            dims.append(tuple(d))
        self._dist_dims = dims

    def calc_particle_distributions(self):
        """Returns (fglobal, [flocal1..4]); flocal arrays are Fortran-ordered
        (nmu, npbins, nrx, nry, nrz) reshaped to C order (nrz, nry, nrx, npbins, nmu)."""
        fg = np.zeros(self._npp, dtype=np.float64)
        outs = []
        ptrs = []
        for d in self._dist_dims:
            size = int(np.prod(d)) if d[0] else 0
            if size:
                a = np.zeros(size, dtype=np.float64)
                outs.append(a)
                ptrs.append(dptr(a))
            else:
                outs.append(None)
                ptrs.append(None)
        self._call("sp_distributions", dptr(fg), *ptrs)
        shaped = []
        for a, d in zip(outs, self._dist_dims):
            shaped.append(None if a is None else a.reshape(d[4], d[3], d[2], d[1], d[0]))
        return fg, shaped

    def calc_escaped_distributions(self):
        """Distributions of the particles removed with an escape flag since the
        previous call (the bins are cleared).  Returns (fescaped, sets):
        fescaped has C shape (2 ndim, npp_global, nmu_global), faces in the
        order low x, high x, low y, high y[, low z, high z]; sets[k] is None or
        a dict of face arrays 'x', 'y'[, 'z'], each of C shape
        (2, n_c, n_b, npbins, nmu) with (b, c) the other two axes."""
        nd = self.cfg.ndim
        fe = np.zeros(2 * nd * self._npp, dtype=np.float64)
        outs, ptrs = [], []
        for d in self._dist_dims:
            if d[0]:
                nmu, npb, nrx, nry, nrz = d
                per = nmu * npb * 2
                size = per * (nry * nrz + nrx * nrz + (nrx * nry if nd > 2 else 0))
                a = np.zeros(size, dtype=np.float64)
                outs.append(a)
                ptrs.append(dptr(a))
            else:
                outs.append(None)
                ptrs.append(None)
        self._call("sp_escaped_distributions", dptr(fe), *ptrs)
        sets = []
        for a, d in zip(outs, self._dist_dims):
            if a is None:
                sets.append(None)
                continue
            nmu, npb, nrx, nry, nrz = d
            shapes = {"x": (2, nrz, nry, npb, nmu), "y": (2, nrz, nrx, npb, nmu)}
            if nd > 2:
                shapes["z"] = (2, nry, nrx, npb, nmu)
            faces, off = {}, 0
            for axis, shp in shapes.items():
                m = int(np.prod(shp))
                faces[axis] = a[off:off + m].reshape(shp)
                off += m
            sets.append(faces)
        return fe.reshape(2 * nd, self._npp_p, -1), sets

    # -- batched parameter variants (no reference counterpart) -----------
    def set_variants(self, variants):
        """variants: sequence of (kpara0, kret, gamma_turb); [] returns to the
        configuration's values."""
        n = len(variants)
        arr = (SpVariant * max(n, 1))()
        for i, (k0, kr, gt) in enumerate(variants):
            arr[i].kpara0, arr[i].kret, arr[i].gamma_turb, arr[i].reserved0 = k0, kr, gt, 0.0
        self._call("sp_set_variants", n, arr)
        self._nvar = n

    def inject_variants(self, nptl, dt, dist_flag, particle_v0, t0, dtf, boxes, power_index=7.0):
        """nptl[v] particles of variant v uniformly in boxes[v] (6 values each)."""
        counts = np.ascontiguousarray(nptl, dtype=np.int64)
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        if len(b) != len(counts):
            raise ValueError("one box per variant")
        self._call("sp_inject_variants", len(counts), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                   float(dt), int(dist_flag), float(particle_v0), float(t0), float(dtf), dptr(b),
                   float(power_index))

    def variant_distributions(self, nvar=None):
        """(f, counts): f[v] is variant v's f(p) over the npp_global bins."""
        nvar = self._nvar if nvar is None else nvar
        f = np.zeros((nvar, self._npp_p), dtype=np.float64)
        cnt = np.zeros(nvar, dtype=np.int64)
        self._call("sp_variant_distributions", nvar, dptr(f), cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        return f, cnt

    def calibration_copy(self, nbytes: int) -> float:
        ms = C.c_double()
        self._call("sp_calibration_copy", int(nbytes), C.byref(ms))
        return ms.value

    def quick_check(self) -> dict:
        out = np.zeros(9, dtype=np.float64)
        self._call("sp_quick_check", dptr(out))
        keys = ("nptl_current", "nptl_split", "ntot", "leak", "leak_negp", "pdt_min", "pdt_max",
                "pdt_avg", "pmax")
        return dict(zip(keys, out.tolist()))
