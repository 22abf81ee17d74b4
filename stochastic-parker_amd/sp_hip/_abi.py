"""ctypes mirror of include/sp_hip.h and the loader of libsp_hip.so.

The library is built in-tree (stochastic-parker_amd/lib/libsp_hip.so) by
``make -C stochastic-parker_amd``; loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsp_hip.so")

SP_OK = 0
SP_PREC_F64 = 0
SP_PREC_F32 = 1
SP_PREC_F32H = 2
SP_INJECT = {"jz": 1, "absj": 2, "db2": 3, "divv": 4, "rho": 5}
COUNT_FLAG_INBOX = 1
COUNT_FLAG_OTHERS = 0

ERRORS = {
    -1: "SP_ERR_ARG",
    -2: "SP_ERR_HIP",
    -3: "SP_ERR_STATE",
    -4: "SP_ERR_CAPACITY",
    -5: "SP_ERR_UNSUPPORTED",
    -6: "SP_ERR_NOMEM",
}


class SpConfig(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("pbcx", C.c_int32), ("pbcy", C.c_int32), ("pbcz", C.c_int32),
        ("momentum_dependency", C.c_int32), ("mag_dependency", C.c_int32),
        ("charge", C.c_int32),
        ("p0", C.c_double), ("pmin", C.c_double), ("pmax", C.c_double), ("b0", C.c_double),
        ("gamma_turb", C.c_double),
        ("kpara0", C.c_double), ("kret", C.c_double),
        ("dt_min_rel", C.c_double), ("dt_max_rel", C.c_double),
        ("drift1", C.c_double), ("drift2", C.c_double),
        ("check_drift_2d", C.c_int32), ("acc_region_flag", C.c_int32),
        ("acc_region", C.c_double * 6),
        ("time_interp", C.c_int32), ("precision", C.c_int32),
        ("seed", C.c_uint64), ("stream", C.c_uint32), ("sort_every", C.c_int32),
        ("sort_steps", C.c_int32), ("reserved0", C.c_int32),
        ("nptl_max", C.c_int64),
        ("dpp_wave", C.c_int32), ("dpp_shear", C.c_int32), ("weak_scattering", C.c_int32),
        ("nlgc", C.c_int32), ("tau0", C.c_double), ("kperp_kpara", C.c_double),
    ]


class SpParticles(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in
                ("x", "y", "z", "p", "v", "mu", "weight", "t", "dt")] + [
        ("split_times", C.POINTER(C.c_int8)), ("count_flag", C.POINTER(C.c_int8)),
        ("origin", C.POINTER(C.c_int32)), ("nsteps_tracked", C.POINTER(C.c_int32)),
        ("nsteps_pushed", C.POINTER(C.c_int32)), ("tag_injected", C.POINTER(C.c_int32)),
        ("tag_splitted", C.POINTER(C.c_int32)), ("rng_step", C.POINTER(C.c_uint32)),
    ]


class SpMoverStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("nptl_current", C.c_int64), ("nptl_escaped", C.c_int64),
                ("leak", C.c_double), ("leak_negp", C.c_double),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double)]


class SpLocalDist(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("npbins", C.c_int32), ("nmu", C.c_int32),
                ("rx", C.c_int32), ("ry", C.c_int32), ("rz", C.c_int32),
                ("pmin", C.c_double), ("pmax", C.c_double)]


class SpDistConfig(C.Structure):
    _fields_ = [("npp_global", C.c_int32), ("nmu_global", C.c_int32),
                ("local_dist", C.c_int32), ("escaped_dist", C.c_int32),
                ("local", SpLocalDist * 4)]


class SpVariant(C.Structure):
    _fields_ = [("kpara0", C.c_double), ("kret", C.c_double), ("gamma_turb", C.c_double),
                ("reserved0", C.c_double)]


class SpTuneResult(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int32), ("sort_steps", C.c_int32), ("ntrials", C.c_int32),
                ("particles_per_lane", C.c_int32), ("steps_per_s", C.c_double),
                ("default_steps_per_s", C.c_double)]


class SpMlpConfig(C.Structure):
    _fields_ = [("nin", C.c_int32), ("nhidden", C.c_int32), ("nlayers", C.c_int32),
                ("nout", C.c_int32), ("batch_max", C.c_int32), ("reserved0", C.c_int32),
                ("seed", C.c_uint64), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("weight_decay", C.c_double)]


# Particle record members: name -> numpy dtype (order of sp_particles)
PARTICLE_FIELDS = [
    ("x", np.float64), ("y", np.float64), ("z", np.float64), ("p", np.float64),
    ("v", np.float64), ("mu", np.float64), ("weight", np.float64), ("t", np.float64),
    ("dt", np.float64), ("split_times", np.int8), ("count_flag", np.int8),
    ("origin", np.int32), ("nsteps_tracked", np.int32), ("nsteps_pushed", np.int32),
    ("tag_injected", np.int32), ("tag_splitted", np.int32), ("rng_step", np.uint32),
]

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)

# name -> (argtypes)
EXPORTS = {
    "sp_ctx_create": [C.POINTER(SpConfig), C.POINTER(C.c_void_p)],
    "sp_ctx_destroy": [_P],
    "sp_last_error": [],
    "sp_abi_version": [],
    "sp_fields_upload": [_P, C.c_int32, C.POINTER(C.c_float), C.c_int64],
    "sp_fields_gradients": [_P, C.c_int32, C.POINTER(C.c_float)],
    "sp_fields_copy": [_P],
    "sp_particles_upload": [_P, C.POINTER(SpParticles), C.c_int64],
    "sp_particles_download": [_P, C.POINTER(SpParticles), C.c_int64, C.POINTER(C.c_int64)],
    "sp_particles_counts": [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                            C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
    "sp_inject_particles_spatial_uniform": [_P, C.c_int64, C.c_double, C.c_int32, C.c_double,
                                            C.c_double, C.c_double, _D, C.c_double],
    "sp_particle_mover": [_P, C.c_double, C.c_double, C.c_int32, C.c_int32,
                          C.POINTER(SpMoverStats)],
    "sp_split_particle": [_P, C.c_double, C.c_double],
    "sp_sort_particles": [_P],
    "sp_step_once": [_P, C.c_int64, C.c_double, C.c_double, _D, _D, C.POINTER(C.c_int8), _D],
    "sp_step_once_noalloc": [_P, C.c_int64, C.c_double, C.c_double, _D, _D,
                             C.POINTER(C.c_int8), _D],
    "sp_distributions_init": [_P, C.POINTER(SpDistConfig)],
    "sp_distributions": [_P, _D, _D, _D, _D, _D],
    "sp_local_dist_dims": [_P, C.c_int32, C.POINTER(C.c_int64)],
    "sp_escaped_distributions": [_P, _D, _D, _D, _D, _D],
    "sp_ncells_above_threshold": [_P, C.c_int32, C.c_double, _D, C.POINTER(C.c_int64)],
    "sp_inject_particles_threshold": [_P, C.c_int32, C.c_int64, C.c_double, C.c_int32, C.c_double,
                                      C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int64,
                                      _D, C.c_double, C.POINTER(C.c_int64)],
    "sp_quick_check": [_P, _D],
    "sp_calibration_copy": [_P, C.c_int64, _D],
    "sp_autotune_mover": [_P, C.c_double, C.c_double, C.c_int32, C.c_int32,
                          C.POINTER(SpTuneResult)],
    "sp_set_variants": [_P, C.c_int32, C.POINTER(SpVariant)],
    "sp_inject_variants": [_P, C.c_int32, C.POINTER(C.c_int64), C.c_double, C.c_int32,
                           C.c_double, C.c_double, C.c_double, _D, C.c_double],
    "sp_variant_distributions": [_P, C.c_int32, _D, C.POINTER(C.c_int64)],
    "sp_mlp_create": [C.POINTER(SpMlpConfig), C.POINTER(C.c_void_p)],
    "sp_mlp_destroy": [_P],
    "sp_mlp_num_params": [_P, C.POINTER(C.c_int64)],
    "sp_mlp_get_params": [_P, _F],
    "sp_mlp_set_params": [_P, _F],
    "sp_mlp_forward": [_P, _F, C.c_int64, _F],
    "sp_mlp_last_forward_ms": [_P, _D],
    "sp_mlp_gradients": [_P, _F, _F, C.c_int64, _F, _D],
    "sp_mlp_train_step": [_P, _F, _F, C.c_int64, C.c_double, _D],
    "sp_mlp_ensemble_spectra": [C.POINTER(C.c_void_p), C.c_int32, _F, C.c_int64, C.c_double, _F,
                                _F],
    "sp_mlp_ensemble_spectra_device": [C.POINTER(C.c_void_p), C.c_int32, C.c_void_p, C.c_int64,
                                       C.c_double, C.c_void_p, C.c_void_p],
    "sp_mlp_set_data": [_P, _F, _F, C.c_int64],
    "sp_mlp_train": [_P, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_uint64, _F],
}

_lib = None


def load(path: str | None = None):
    """Load libsp_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"libsp_hip.so not found at {path}: build it with `make -C stochastic-parker_amd` "
            "(or __graft_entry__.build()); there is no fallback path")
    lib = C.CDLL(path)
    for name, args in EXPORTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "sp_last_error" else C.c_int
    _lib = lib
    return lib


class SpError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(lib, fn: str, rc: int):
    if rc != SP_OK:
        raise SpError(fn, rc, lib.sp_last_error().decode(errors="replace"))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)
