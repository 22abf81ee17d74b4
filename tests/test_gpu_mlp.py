"""The surrogate network's BF16 MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py, which rounds to BF16 at the same points)."""
import numpy as np
import pytest

import mlp_oracle as mo
from sp_hip.surrogate import Mlp

pytestmark = pytest.mark.gpu

SHAPE = (6, 256, 4, 64)


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, SHAPE[0])).astype(np.float32)
    t = (np.sin(x @ rng.standard_normal((SHAPE[0], SHAPE[3]))) * 0.5).astype(np.float32)
    return x, t


def _net(**kw):
    return Mlp(*SHAPE, **kw)


def test_forward_matches_oracle():
    x, _ = _data(300)
    with _net(seed=3) as net:
        y = net.forward(x)
        p = net.get_params()
    yr, _ = mo.forward(p, x, SHAPE, bf=True)
    err = np.abs(y - yr)
    assert err.max() < 2e-3 * np.abs(yr).max()
    assert np.median(err) < 1e-4 * np.abs(yr).max()


def test_forward_splits_rows_beyond_batch_max():
    x, _ = _data(1000, seed=1)
    with _net(seed=4, batch_max=128) as a, _net(seed=4, batch_max=4096) as b:
        np.testing.assert_array_equal(a.forward(x), b.forward(x))


def test_gradients_match_oracle():
    x, t = _data(500, seed=2)
    with _net(seed=5) as net:
        p = net.get_params()
        loss, g = net.gradients(x, t)
    lr, gr = mo.loss_and_grads(p, x, t, SHAPE, bf=True)
    assert loss == pytest.approx(lr, rel=1e-4)
    layers = mo.unpack(np.arange(len(p)), *SHAPE)
    for w, b in layers:
        for idx in (w.ravel(), b):
            num = np.linalg.norm(g[idx] - gr[idx])
            assert num < 1e-3 * np.linalg.norm(gr[idx]) + 1e-9
    # and the rounded network's gradient is the exact one to BF16 precision
    _, g_exact = mo.loss_and_grads(p, x, t, SHAPE, bf=False)
    assert np.dot(g, g_exact) / (np.linalg.norm(g) * np.linalg.norm(g_exact)) > 0.999


def test_adamw_step_matches_oracle():
    x, t = _data(256, seed=3)
    with _net(seed=6, weight_decay=0.01) as net:
        p0 = net.get_params().astype(np.float64)
        _, g = net.gradients(x, t)
        net.train_step(x, t, 1e-3)
        p1 = net.get_params()
    ref, _, _ = mo.adamw(p0, g.astype(np.float64), 0.0, 0.0, 1, 1e-3, 0.9, 0.999, 1e-8, 0.01)
    np.testing.assert_allclose(p1, ref, rtol=1e-6, atol=1e-8)


def test_training_fits_and_is_deterministic():
    x, t = _data(20000, seed=4)
    finals = []
    for _ in range(2):
        with _net(seed=8) as net:
            net.set_data(x, t)
            losses = net.train(400, batch=2048, lr0=2e-3, lr_min=1e-4, seed=11)
            finals.append(net.get_params())
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    np.testing.assert_array_equal(finals[0], finals[1])


def test_ensemble_spectra_match_members():
    from sp_hip.surrogate import FLOOR, Surrogate, spectra_from_log
    x, _ = _data(3000, seed=5)
    with Surrogate(members=2, seed=3, batch_max=1024) as s:
        spec, spread = s.predict(x)
        ys = [spectra_from_log(n.forward(x)) for n in s.nets]
    mean = (ys[0] + ys[1]) / 2
    np.testing.assert_allclose(spec, mean, rtol=1e-5, atol=1e-7 + 1e-5 * FLOOR)
    ref = 0.5 * (np.abs(ys[0] - mean).sum(1) + np.abs(ys[1] - mean).sum(1))
    np.testing.assert_allclose(spread, ref, rtol=1e-4, atol=1e-6)


def test_training_graph_replays_match_eager(monkeypatch):
    """sp_mlp_train records one step as a HIP graph and replays it; every
    chunk of replays, with other work on the stream between chunks, equals
    the same steps launched one by one."""
    x, t = _data(8000, seed=6)
    with _net(seed=9, batch_max=1024) as g, _net(seed=9, batch_max=1024) as e:
        g.set_data(x, t)
        e.set_data(x, t)
        for _ in range(3):
            lg = g.train(40, batch=1024, lr0=1e-3, lr_min=1e-4, seed=3)
            monkeypatch.setenv("SP_MLP_EAGER", "1")
            le = e.train(40, batch=1024, lr0=1e-3, lr_min=1e-4, seed=3)
            monkeypatch.delenv("SP_MLP_EAGER")
            g.forward(x[:300])
            e.forward(x[:300])
            np.testing.assert_array_equal(lg, le)
            np.testing.assert_array_equal(g.get_params(), e.get_params())
        assert lg[-1] < lg[0]


@pytest.mark.parametrize("shape", [SHAPE, (5, 128, 2, 32), (6, 256, 1, 128)])
def test_fused_inference_equals_layer_gemms(monkeypatch, shape):
    """The one-kernel inference path gives the layer-by-layer GEMM path's
    outputs bit for bit (ragged row counts, several chunks)."""
    rng = np.random.default_rng(7)
    x = rng.uniform(-1.0, 1.0, (1000, shape[0])).astype(np.float32)
    with Mlp(*shape, seed=12, batch_max=384) as net:
        y_fused = net.forward(x)
        fwd_ms = net.last_forward_ms()
        monkeypatch.setenv("SP_MLP_UNFUSED", "1")
        y_gemm = net.forward(x)
    assert fwd_ms > 0
    np.testing.assert_array_equal(y_fused, y_gemm)


def test_ensemble_device_arrays_equal_host_arrays(monkeypatch):
    from sp_hip.device import DeviceArray
    from sp_hip.surrogate import Surrogate
    x, _ = _data(2500, seed=8)
    with Surrogate(members=2, seed=4, batch_max=1024) as s:
        s.c0, s.c1 = 0.01, 2.0
        spec_h, est_h = s.predict(x)
        with DeviceArray.from_host(x) as xd, DeviceArray((len(x), SHAPE[3])) as sd, \
                DeviceArray(len(x)) as rd:
            s.predict_device(xd, sd, rd)
            spec_d, spread_d = sd.to_host(), rd.to_host()
        monkeypatch.setenv("SP_MLP_UNFUSED", "1")
        spec_g, est_g = s.predict(x)
    np.testing.assert_array_equal(spec_d, spec_h)
    np.testing.assert_array_equal(s.estimate(spread_d), est_h)
    np.testing.assert_array_equal(spec_g, spec_h)
    np.testing.assert_array_equal(est_g, est_h)


def test_fused_inference_follows_parameter_updates(monkeypatch):
    """The fragment-ordered weights of the fused path are rewritten after
    every kind of parameter update (AdamW step, training loop, set_params)."""
    x, t = _data(600, seed=9)
    with _net(seed=13, batch_max=512) as net:
        ys = [net.forward(x)]
        net.train_step(x[:256], t[:256], 1e-2)
        ys.append(net.forward(x))
        net.set_data(x, t)
        net.train(5, batch=256, lr0=1e-2, lr_min=1e-3, seed=2)
        ys.append(net.forward(x))
        net.set_params(net.get_params() * 0.5)
        ys.append(net.forward(x))
        monkeypatch.setenv("SP_MLP_UNFUSED", "1")
        y_gemm = net.forward(x)
    for a, b in zip(ys, ys[1:]):
        assert not np.array_equal(a, b)
    np.testing.assert_array_equal(ys[-1], y_gemm)
