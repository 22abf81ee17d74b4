"""ctypes wrapper of the CPU oracle (oracle/libsp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsp_oracle.so")

# orc_particle (sp_oracle.h), C layout
PARTICLE_DTYPE = np.dtype([
    ("split_times", np.int8), ("count_flag", np.int8),
    ("origin", np.int32), ("nsteps_tracked", np.int32), ("nsteps_pushed", np.int32),
    ("tag_injected", np.int32), ("tag_splitted", np.int32), ("rng_step", np.uint32),
    ("x", np.float64), ("y", np.float64), ("z", np.float64), ("p", np.float64),
    ("v", np.float64), ("mu", np.float64), ("weight", np.float64), ("t", np.float64),
    ("dt", np.float64), ("padding", np.float64),
], align=True)


class OrcParams(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("pbcx", C.c_int32), ("pbcy", C.c_int32), ("pbcz", C.c_int32),
        ("p0", C.c_double), ("pmin", C.c_double), ("pmax", C.c_double),
        ("momentum_dependency", C.c_int32), ("mag_dependency", C.c_int32),
        ("gamma_turb", C.c_double), ("pindex", C.c_double),
        ("kpara0", C.c_double), ("kret", C.c_double),
        ("dt_min_rel", C.c_double), ("dt_max_rel", C.c_double),
        ("drift1", C.c_double), ("drift2", C.c_double),
        ("pcharge", C.c_int32), ("check_drift_2d", C.c_int32),
        ("acc_region_flag", C.c_int32), ("acc_region", C.c_double * 6),
        ("time_interp", C.c_int32),
        ("seed", C.c_uint64), ("stream", C.c_uint32),
        ("dpp_wave", C.c_int32), ("dpp_shear", C.c_int32), ("weak_scattering", C.c_int32),
        ("tau0", C.c_double), ("nlgc", C.c_int32), ("kperp_kpara", C.c_double),
    ]


class OrcKappa(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "knorm_para", "knorm_perp", "kpara", "kperp", "skpara", "skperp", "skpara_perp",
        "kxx", "kyy", "kzz", "kxy", "kxz", "kyz",
        "dkxx_dx", "dkyy_dy", "dkzz_dz", "dkxy_dx", "dkxy_dy", "dkxz_dx", "dkxz_dz",
        "dkyz_dy", "dkyz_dz")]


class OrcMoverAcc(C.Structure):
    _fields_ = [("leak", C.c_double), ("leak_negp", C.c_double), ("steps", C.c_int64),
                ("nptl_escaped", C.c_int64)]


class OrcDistConfig(C.Structure):
    _fields_ = [("npp_global", C.c_int32), ("nmu_global", C.c_int32), ("local_dist", C.c_int32),
                ("local", (C.c_int32 * 6) * 4), ("local_p", (C.c_double * 2) * 4)]


class OrcMT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_I64 = C.POINTER(C.c_int64)
_PP = C.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.POINTER(OrcParams)
    sig = {
        "orc_philox4x32_10": (None, [_U32, _U32, _U32]),
        "orc_u01": (C.c_double, [C.c_uint32]),
        "orc_mt_init_genrand": (None, [C.POINTER(OrcMT), C.c_uint32]),
        "orc_mt_init_by_array": (None, [C.POINTER(OrcMT), _U32, C.c_int]),
        "orc_mt_genrand_int32": (C.c_uint32, [C.POINTER(OrcMT)]),
        "orc_field_nodes": (C.c_int64, [P]),
        "orc_calc_fields_gradients": (None, [P, _F]),
        "orc_get_interp_paramters": (None, [P, C.c_double, C.c_double, C.c_double,
                                            C.POINTER(C.c_int), _D]),
        "orc_interp_fields": (None, [P, _F, _F, C.POINTER(C.c_int), _D, C.c_double, _D]),
        "orc_kappa_standard": (None, [P, C.c_double, _D, C.POINTER(OrcKappa)]),
        "orc_kappa_nlgc": (None, [P, C.c_double, C.c_double, _D, C.POINTER(OrcKappa)]),
        "orc_push_particle_2d": (None, [P, _D, _PP, _D, C.POINTER(OrcKappa), C.c_int, _D, _D]),
        "orc_push_particle_3d": (None, [P, _D, _PP, _D, C.POINTER(OrcKappa), C.c_int, _D, _D]),
        "orc_step_once": (None, [P, _F, _F, C.c_double, C.c_double, _PP, C.c_int, _D, _D]),
        "orc_particle_mover": (None, [P, _F, _F, _PP, _I64, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.POINTER(OrcMoverAcc)]),
        "orc_remove_particles": (None, [_PP, _I64, _I64]),
        "orc_inject_particles_spatial_uniform": (None, [P, _PP, _I64, C.c_int64, C.c_int64,
                                                        C.c_double, C.c_int, C.c_double,
                                                        C.c_double, C.c_double, _D, C.c_double,
                                                        C.POINTER(C.c_int32)]),
        "orc_split_particle": (None, [P, _PP, _I64, C.c_int64, C.c_double, C.c_double, _I64]),
        "orc_ncells_above_threshold": (C.c_int64, [P, _F, C.c_int, C.c_double, _D]),
        "orc_inject_particles_threshold": (C.c_int64, [P, _F, _PP, _I64, C.c_int64, C.c_int,
                                                       C.c_int64, C.c_double, C.c_int, C.c_double,
                                                       C.c_double, C.c_double, C.c_int, C.c_double,
                                                       C.c_int64, _D, C.c_double,
                                                       C.POINTER(C.c_int32)]),
        "orc_local_dist_size": (C.c_int64, [P, C.POINTER(OrcDistConfig), C.c_int]),
        "orc_calc_particle_distributions": (None, [P, C.POINTER(OrcDistConfig), _PP, C.c_int64,
                                                   _D, C.POINTER(_D)]),
        "orc_escaped_dist_size": (C.c_int64, [P, C.POINTER(OrcDistConfig), C.c_int]),
        "orc_calc_escaped_distributions": (None, [P, C.POINTER(OrcDistConfig), _PP, C.c_int64,
                                                  _D, C.POINTER(_D)]),
        "orc_quick_check": (None, [_PP, C.c_int64, C.c_int64, C.c_double, C.c_double, _D]),
        "orc_sizeof_particle": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    assert lib.orc_sizeof_particle() == PARTICLE_DTYPE.itemsize, "orc_particle layout mismatch"
    _lib = lib
    return lib


def params(cfg: dict, seed: int = 20261202, stream: int = 0) -> OrcParams:
    """OrcParams from an sp_config-style dict (same key names)."""
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        if name in ("pindex", "pcharge", "acc_region", "seed", "stream", "dpp_wave", "dpp_shear",
                    "weak_scattering", "tau0", "nlgc", "kperp_kpara"):
            continue
        setattr(p, name, cfg[name])
    p.pindex = 3.0 - cfg["gamma_turb"]
    p.pcharge = cfg["charge"]
    ar = cfg.get("acc_region", (0.0, 1.0, 0.0, 1.0, 0.0, 1.0))
    for i in range(6):
        p.acc_region[i] = ar[i]
    p.dpp_wave = cfg.get("dpp_wave", 0)
    p.dpp_shear = cfg.get("dpp_shear", 0)
    p.weak_scattering = cfg.get("weak_scattering", 0)
    p.tau0 = cfg.get("tau0", 0.0)
    p.nlgc = cfg.get("nlgc", 0)
    p.kperp_kpara = cfg.get("kperp_kpara", 0.0)
    p.seed = cfg.get("seed", seed)
    p.stream = cfg.get("stream", stream)
    return p


def fptr(a):
    return None if a is None else a.ctypes.data_as(_F)


def dptr(a):
    return a.ctypes.data_as(_D)


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def farray_from_nodes8(nodes8: np.ndarray) -> np.ndarray:
    """(nzg, nyg, nxg, 8) -> zeroed (nzg, nyg, nxg, 32) farray with fields 1..8."""
    fa = np.zeros(nodes8.shape[:-1] + (32,), dtype=np.float32)
    fa[..., :8] = nodes8
    return fa


def gradients(prm: OrcParams, nodes8: np.ndarray) -> np.ndarray:
    lib = load()
    fa = farray_from_nodes8(nodes8)
    lib.orc_calc_fields_gradients(C.byref(prm), fptr(fa))
    return fa


def particles_from_soa(arrs: dict) -> np.ndarray:
    n = len(arrs["x"])
    a = np.zeros(n, dtype=PARTICLE_DTYPE)
    for name in PARTICLE_DTYPE.names:
        if name in arrs:
            a[name] = arrs[name]
    return a


def soa_from_particles(a: np.ndarray) -> dict:
    return {name: np.ascontiguousarray(a[name]) for name in PARTICLE_DTYPE.names if name != "padding"}


def mover(prm, fa1, ptls, t0, dtf, num_fine_steps=1, nsteps_interval=10, rng_mode=0,
          nthreads=1, fa2=None, return_removed=False):
    """Runs orc_particle_mover in place on a copy; returns (particles, acc), and
    with return_removed also the removed particles (remove_particles swaps
    them to the tail of the array, where they are read back)."""
    lib = load()
    buf = np.ascontiguousarray(ptls.copy())
    n = C.c_int64(len(buf))
    acc = OrcMoverAcc()
    lib.orc_particle_mover(C.byref(prm), fptr(fa1), fptr(fa2 if fa2 is not None else fa1),
                           buf.ctypes.data_as(_PP), C.byref(n), t0, dtf, num_fine_steps,
                           nsteps_interval, rng_mode, nthreads, C.byref(acc))
    if return_removed:
        return buf[: n.value].copy(), acc, buf[n.value:].copy()
    return buf[: n.value].copy(), acc


def step_once(prm, fa1, ptl_row, t0, dtf, fixed, rands, fa2=None):
    lib = load()
    buf = np.ascontiguousarray(ptl_row.copy().reshape(1))
    r = np.ascontiguousarray(rands, dtype=np.float64)
    d = np.zeros(4)
    lib.orc_step_once(C.byref(prm), fptr(fa1), fptr(fa2 if fa2 is not None else fa1), t0, dtf,
                      buf.ctypes.data_as(_PP), int(fixed), dptr(r), dptr(d))
    return buf[0], d


def inject(prm, n_existing_arr, nptl_max, nptl, dt, dist_flag, v0, t0, dtf, box, power_index,
           tag_max=0):
    lib = load()
    buf = np.zeros(nptl_max, dtype=PARTICLE_DTYPE)
    k = len(n_existing_arr)
    buf[:k] = n_existing_arr
    n = C.c_int64(k)
    tm = C.c_int32(tag_max)
    b = np.ascontiguousarray(box, dtype=np.float64)
    lib.orc_inject_particles_spatial_uniform(C.byref(prm), buf.ctypes.data_as(_PP), C.byref(n),
                                             nptl_max, nptl, dt, dist_flag, v0, t0, dtf, dptr(b),
                                             power_index, C.byref(tm))
    return buf[: n.value].copy(), tm.value


THRESHOLD = {"jz": 1, "absj": 2, "divv": 4, "rho": 5}


def ncells_above_threshold(prm, fa1, quantity, threshold, box):
    lib = load()
    b = np.ascontiguousarray(box, dtype=np.float64)
    return lib.orc_ncells_above_threshold(C.byref(prm), fptr(fa1), THRESHOLD[quantity], threshold,
                                          dptr(b))


def inject_threshold(prm, fa1, existing, nptl_max, quantity, nptl, dt, dist_flag, v0, t0, dtf,
                     same, threshold, ncells_norm, box, power_index, tag_max=0):
    """Returns (particles, tag_max, number injected)."""
    lib = load()
    buf = np.zeros(nptl_max, dtype=PARTICLE_DTYPE)
    k = len(existing)
    buf[:k] = existing
    n = C.c_int64(k)
    tm = C.c_int32(tag_max)
    b = np.ascontiguousarray(box, dtype=np.float64)
    ninj = lib.orc_inject_particles_threshold(C.byref(prm), fptr(fa1), buf.ctypes.data_as(_PP),
                                              C.byref(n), nptl_max, THRESHOLD[quantity], nptl, dt,
                                              dist_flag, v0, t0, dtf, int(same), threshold,
                                              ncells_norm, dptr(b), power_index, C.byref(tm))
    return buf[: n.value].copy(), tm.value, ninj


def split(prm, ptls, nptl_max, split_ratio, pmin_split):
    lib = load()
    buf = np.zeros(nptl_max, dtype=PARTICLE_DTYPE)
    buf[: len(ptls)] = ptls
    n = C.c_int64(len(ptls))
    ns = C.c_int64(0)
    lib.orc_split_particle(C.byref(prm), buf.ctypes.data_as(_PP), C.byref(n), nptl_max,
                           split_ratio, pmin_split, C.byref(ns))
    return buf[: n.value].copy(), ns.value


def remove(ptls):
    lib = load()
    buf = np.ascontiguousarray(ptls.copy())
    n = C.c_int64(len(buf))
    ne = C.c_int64(0)
    lib.orc_remove_particles(buf.ctypes.data_as(_PP), C.byref(n), C.byref(ne))
    return buf[: n.value].copy(), ne.value


def dist_config(d: dict) -> OrcDistConfig:
    dc = OrcDistConfig()
    dc.npp_global = d["npp_global"]
    dc.nmu_global = d["nmu_global"]
    dc.local_dist = d.get("local_dist", 0)
    for k, L in enumerate(d.get("local", [])):
        if L is None:
            continue
        vals = [L["enabled"], L["npbins"], L["nmu"], L["rx"], L["ry"], L["rz"]]
        for i in range(6):
            dc.local[k][i] = vals[i]
        dc.local_p[k][0] = L["pmin"]
        dc.local_p[k][1] = L["pmax"]
    return dc


def distributions(prm, dcfg: dict, ptls):
    lib = load()
    dc = dist_config(dcfg)
    fg = np.zeros(dc.npp_global * dc.nmu_global)
    locs = []
    arr = (_D * 4)()
    for k in range(4):
        size = lib.orc_local_dist_size(C.byref(prm), C.byref(dc), k)
        a = np.zeros(max(size, 1))
        locs.append(a if size else None)
        arr[k] = dptr(a)
    buf = np.ascontiguousarray(ptls)
    lib.orc_calc_particle_distributions(C.byref(prm), C.byref(dc), buf.ctypes.data_as(_PP),
                                        len(buf), dptr(fg), arr)
    return fg, locs


def escaped_distributions(prm, dcfg: dict, escaped):
    """calc_escaped_distributions over the given removed particles (those with
    count_flag >= 0 are skipped).  Returns (fescaped, [set arrays or None])."""
    lib = load()
    dc = dist_config(dcfg)
    fe = np.zeros(lib.orc_escaped_dist_size(C.byref(prm), C.byref(dc), -1))
    locs = []
    arr = (_D * 4)()
    for k in range(4):
        size = lib.orc_escaped_dist_size(C.byref(prm), C.byref(dc), k)
        a = np.zeros(max(size, 1))
        locs.append(a if size else None)
        arr[k] = dptr(a)
    buf = np.ascontiguousarray(escaped)
    lib.orc_calc_escaped_distributions(C.byref(prm), C.byref(dc), buf.ctypes.data_as(_PP),
                                       len(buf), dptr(fe), arr)
    return fe, locs


def quick_check(ptls, nptl_split, leak, leak_negp):
    lib = load()
    out = np.zeros(9)
    buf = np.ascontiguousarray(ptls)
    lib.orc_quick_check(buf.ctypes.data_as(_PP), len(buf), nptl_split, leak, leak_negp, dptr(out))
    keys = ("nptl_current", "nptl_split", "ntot", "leak", "leak_negp", "pdt_min", "pdt_max",
            "pdt_avg", "pmax")
    return dict(zip(keys, out.tolist()))


def mt_outputs(key, count):
    lib = load()
    s = OrcMT()
    k = (C.c_uint32 * len(key))(*key)
    lib.orc_mt_init_by_array(C.byref(s), k, len(key))
    return [lib.orc_mt_genrand_int32(C.byref(s)) for _ in range(count)]
