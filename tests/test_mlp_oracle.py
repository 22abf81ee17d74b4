"""The surrogate network's CPU oracle, pinned against independent
implementations: torch's bfloat16 conversion, torch autograd and torch's
AdamW (all on the CPU), and finite differences."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo

SHAPE = (3, 5, 2, 4)   # nin, nhidden, nlayers, nout


def _params(seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.8, 0.8, mo.num_params(*SHAPE))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 10000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)
    # ties go to the even mantissa
    assert mo.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -8)) == np.float32(1.0 + 2 ** -6)


def _torch_net(params):
    layers = mo.unpack(np.asarray(params, np.float64), *SHAPE)
    ws = [(torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in layers]
    return ws


def test_gradients_match_torch_autograd():
    p = _params()
    rng = np.random.default_rng(2)
    x = rng.standard_normal((17, SHAPE[0]))
    t = rng.standard_normal((17, SHAPE[3]))
    loss, g = mo.loss_and_grads(p, x, t, SHAPE, bf=False)
    ws = _torch_net(p)
    h = torch.tensor(x)
    for i, (w, b) in enumerate(ws):
        h = h @ w + b
        if i + 1 < len(ws):
            h = torch.nn.functional.silu(h)
    tl = torch.mean((h - torch.tensor(t)) ** 2)
    tl.backward()
    tg = np.concatenate([np.concatenate([w.grad.numpy().ravel(), b.grad.numpy()]) for w, b in ws])
    assert loss == pytest.approx(tl.item(), rel=1e-12)
    np.testing.assert_allclose(g, tg, rtol=1e-10, atol=1e-13)


def test_gradients_match_finite_differences():
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((5, SHAPE[0]))
    t = rng.standard_normal((5, SHAPE[3]))
    _, g = mo.loss_and_grads(p, x, t, SHAPE, bf=False)
    h = 1e-6
    for i in rng.choice(len(p), 25, replace=False):
        q = p.copy()
        q[i] += h
        lp, _ = mo.loss_and_grads(q, x, t, SHAPE, bf=False)
        q[i] -= 2 * h
        lm, _ = mo.loss_and_grads(q, x, t, SHAPE, bf=False)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-9)


def test_adamw_matches_torch():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.AdamW([tp], lr=3e-3, betas=(0.9, 0.99), eps=1e-8, weight_decay=0.01)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.standard_normal(50)
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = mo.adamw(p, g, m, v, step, 3e-3, 0.9, 0.99, 1e-8, 0.01)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_bf16_mode_is_close_to_exact():
    # the rounded network stays within BF16 precision of the exact one
    p = _params(6)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((33, SHAPE[0]))
    t = rng.standard_normal((33, SHAPE[3]))
    l0, g0 = mo.loss_and_grads(p, x, t, SHAPE, bf=False)
    l1, g1 = mo.loss_and_grads(p, x, t, SHAPE, bf=True)
    assert l1 == pytest.approx(l0, rel=0.05)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)
