"""The C-ABI library loads and exports every entry point the header declares;
host-side argument validation (no GPU work) behaves as documented."""
import ctypes as C
import os
import re

import pytest

from sp_hip import _abi
from sp_hip.pusher import make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sp_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "sp_particle_mover" in names and "sp_ctx_create" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes mirror covers exactly the declared set
    assert sorted(_abi.EXPORTS) == declared_functions()


def test_abi_version():
    assert _abi.load().sp_abi_version() == 4


def test_struct_sizes_match_header_layout():
    assert C.sizeof(_abi.SpConfig) == 352
    assert C.sizeof(_abi.SpVariant) == 32
    assert C.sizeof(_abi.SpTuneResult) == 32
    assert C.sizeof(_abi.SpMoverStats) == 56
    assert C.sizeof(_abi.SpLocalDist) == 40
    assert C.sizeof(_abi.SpDistConfig) == 176
    assert C.sizeof(_abi.SpParticles) == 17 * 8


@pytest.mark.parametrize("bad,msg", [
    (dict(ndim=1), "ndim"),
    (dict(nx=0), "nx"),
    (dict(nz=3), "nz == 1"),
    (dict(dx=0.0), "dx"),
    (dict(pmax=1e-3), "pmin < pmax"),
    (dict(kpara0=-1.0), "kpara0"),
    (dict(kret=2.0), "kret"),
    (dict(dt_min_rel=1e-2, dt_max_rel=1e-3), "dt_min_rel"),
    (dict(charge=0), "charge"),
    (dict(drift1=0.0, drift2=0.0), "drift"),
    (dict(pbcx=2), "pbc"),
    (dict(precision=7), "precision"),
    (dict(nptl_max=0), "nptl_max"),
])
def test_invalid_configs_are_rejected_before_any_device_work(bad, msg):
    lib = _abi.load()
    cfg = make_config(**bad)
    ctx = C.c_void_p()
    rc = lib.sp_ctx_create(C.byref(cfg), C.byref(ctx))
    assert rc in (-1, -5)
    assert msg in lib.sp_last_error().decode()
    assert not ctx.value


def test_null_arguments():
    lib = _abi.load()
    assert lib.sp_ctx_create(None, None) == -1
    assert lib.sp_ctx_destroy(None) == 0
    assert lib.sp_particle_mover(None, 0.0, 1.0, 1, 1, None) == -1
    assert b"NULL" in lib.sp_last_error()
