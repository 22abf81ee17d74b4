"""CPU restatement of the f(p) surrogate network (test infrastructure only).

The library's network (libsp_hip.so, sp_mlp_*) has no counterpart in the
reference; this module is the checker of its forward pass, its gradients and
its AdamW update.  It follows the same parameter layout (for each linear layer
W_l as K_l x N_l row-major, then b_l) and, with bf16=True, rounds to BF16
(round-to-nearest-even) at the same points as the GPU kernels:

  X_0 = bf16(x); W~ = bf16(W)
  hidden layer l:  Z_l = X_l W~_l + b_l (FP32 parameters, FP64 sums here)
                   X_{l+1} = bf16(silu(Z_l))
  output:          Y = X_L W~_L + b_L
  loss = mean((Y - T)^2);  G_L = bf16(2 (Y - T) / count)
  dW_l = X_l^T G_l;  db_l = column sums of G_l
  G_{l-1} = bf16((G_l W~_l^T) * silu'(Z_{l-1}))

With bf16=False every rounding is skipped, which makes the gradients exact
derivatives of the loss (checked against finite differences)."""
from __future__ import annotations

import numpy as np


def bf16(a) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even); returns FP32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def layer_sizes(nin, nhidden, nlayers, nout):
    k = [nin] + [nhidden] * nlayers
    n = [nhidden] * nlayers + [nout]
    return k, n


def unpack(params, nin, nhidden, nlayers, nout):
    """[(W_l, b_l)] views of the flat parameter vector."""
    ks, ns = layer_sizes(nin, nhidden, nlayers, nout)
    out, o = [], 0
    for k, n in zip(ks, ns):
        w = params[o:o + k * n].reshape(k, n)
        o += k * n
        b = params[o:o + n]
        o += n
        out.append((w, b))
    assert o == len(params)
    return out


def num_params(nin, nhidden, nlayers, nout):
    ks, ns = layer_sizes(nin, nhidden, nlayers, nout)
    return sum(k * n + n for k, n in zip(ks, ns))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(params, x, shape, bf=True):
    """Returns (Y, cache) in FP64 arithmetic from the rounded inputs."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if bf else np.float64), *shape)
    xs, zs = [r(x).astype(np.float64)], []
    for i, (w, b) in enumerate(layers):
        z = xs[-1] @ r(w).astype(np.float64) + b.astype(np.float64)
        if i + 1 < len(layers):
            z32 = z.astype(np.float32) if bf else z
            zs.append(z32)
            s = sigmoid(np.asarray(z32, dtype=np.float64))
            xs.append(r((z32 * s).astype(np.float32) if bf else z32 * s).astype(np.float64))
        else:
            y = z
    return y, (layers, xs, zs)


def loss_and_grads(params, x, t, shape, bf=True):
    """(loss, flat gradient) of the mean squared error."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, (layers, xs, zs) = forward(params, x, shape, bf)
    e = y - np.asarray(t, dtype=np.float64)
    count = e.size
    loss = float(np.mean(e * e))
    g = 2.0 * e / count
    g = r(g.astype(np.float32) if bf else g).astype(np.float64)
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        grads.append((xs[i].T @ g, g.sum(axis=0)))
        if i > 0:
            z = np.asarray(zs[i - 1], dtype=np.float64)
            s = sigmoid(z)
            d = (g @ r(w).astype(np.float64).T) * (s * (1.0 + z * (1.0 - s)))
            g = r(d.astype(np.float32) if bf else d).astype(np.float64)
    flat = []
    for dw, db in reversed(grads):
        flat.append(dw.ravel())
        flat.append(db.ravel())
    return loss, np.concatenate(flat)


def adamw(p, g, m, v, step, lr, beta1, beta2, eps, wd):
    """One AdamW step (decoupled decay on every parameter, bias correction)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    c1 = 1.0 - beta1 ** step
    c2 = 1.0 - beta2 ** step
    upd = (m / c1) / (np.sqrt(v / c2) + eps)
    return p - lr * (upd + wd * p), m, v
