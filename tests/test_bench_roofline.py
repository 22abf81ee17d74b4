"""bench.py's roofline: every committed counter run of a mover configuration
is found by the configuration it was taken on, and at that run's own rate the
binding roof (VALU issue or texture-data busy cycles) is a fraction in (0, 1],
the larger of the two; algorithmic gather bytes count both frames of a
time-interpolated step."""
import importlib.util
import json
import os
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def entries():
    with open(os.path.join(ROOT, "profiles", "mover_counters.json")) as f:
        return sorted(json.load(f).items())


@pytest.mark.parametrize("key,e", entries(), ids=[k for k, _ in entries()])
def test_counter_runs_give_a_fraction_at_their_own_rate(bench, key, e):
    kind = key.split("_")[0]
    wl = types.SimpleNamespace(kind=kind, ndim=3 if kind == "3d" else 2, grid=e["grid"],
                               nsteps=e["nsteps"], ti=e["time_interp"])
    assert key == f"{kind}_{'ti_' if e['time_interp'] else ''}{e['precision']}"
    rate = e["bench_value"]
    r = bench.roofline(wl, e["nptl"], e["precision"], rate, 1000.0, 1)
    assert r["bound"] in ("valu_issue", "texture_data")
    valu = e["valu_instr_per_step"] * rate / bench.VALU_LANE_INSTR_PEAK
    td = e["td_cycles_per_step"] * rate / bench.TD_CYCLES_PEAK
    assert r["frac"] == pytest.approx(max(valu, td), rel=1e-12)
    assert 0.0 < r["frac"] <= 1.0
    assert r["hbm"]["measured_bytes_per_step"] == e["hbm_bytes_per_step"]
    frames = 2 if e["time_interp"] else 1
    assert r["hbm"]["algorithmic_bytes_per_step"] == frames * bench.bytes_per_step(wl.ndim, e["precision"])


def test_other_configurations_fall_back_to_the_hbm_view(bench):
    # another particle count than the counter run's
    wl = types.SimpleNamespace(kind="2d", ndim=2, grid=2048, nsteps=2000, ti=False)
    r = bench.roofline(wl, 12345, "f32", 1e10, 1000.0, 1)
    assert r["bound"] == "hbm" and r["traffic"] is None and "note" in r
    assert r["frac"] == pytest.approx(224 * 1e10 / 1e9 / bench.HBM_PEAK_GBS)
