"""The f(p) surrogate: a dense SiLU network on the GPU (libsp_hip.so, sp_mlp_*)
trained on spectra that the batched pusher computes, with the pusher itself as
the fallback for queries the network cannot answer well.

Query features (the parameters of one pusher run and a time):
  log10 kpara0, log10 kret, momentum exponent 3 - gamma_turb, the injection
  region as two bits, and the time since injection; each scaled to [-1, 1].
Output: log10 of f(p), the fraction of particles in each of nbins log-spaced
momentum bins (floored at FLOOR)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import SpMlpConfig, check

FLOOR = 1e-5


class Mlp:
    """A network context: nin -> nhidden (x nlayers, SiLU) -> nout."""

    def __init__(self, nin=6, nhidden=256, nlayers=4, nout=64, batch_max=65536, seed=1,
                 beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4):
        self.lib = _abi.load()
        c = SpMlpConfig(nin=nin, nhidden=nhidden, nlayers=nlayers, nout=nout,
                        batch_max=batch_max, reserved0=0, seed=seed, beta1=beta1, beta2=beta2,
                        eps=eps, weight_decay=weight_decay)
        h = C.c_void_p()
        check(self.lib, "sp_mlp_create", self.lib.sp_mlp_create(C.byref(c), C.byref(h)))
        self.h = h
        self.shape = (nin, nhidden, nlayers, nout)
        n = C.c_int64()
        self._call("sp_mlp_num_params", C.byref(n))
        self.nparams = n.value

    def _call(self, name, *args):
        check(self.lib, name, getattr(self.lib, name)(self.h, *args))

    def close(self):
        if self.h:
            self.lib.sp_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        self._call("sp_mlp_get_params", self._f(out))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        self._call("sp_mlp_set_params", self._f(a))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.shape[0])
        y = np.zeros((len(x), self.shape[3]), dtype=np.float32)
        self._call("sp_mlp_forward", self._f(x), len(x), self._f(y))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.shape[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.shape[3])
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double()
        self._call("sp_mlp_gradients", self._f(x), self._f(t), len(x), self._f(g), C.byref(loss))
        return loss.value, g

    def train_step(self, x, t, lr) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.shape[0])
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.shape[3])
        loss = C.c_double()
        self._call("sp_mlp_train_step", self._f(x), self._f(t), len(x), float(lr), C.byref(loss))
        return loss.value

    def set_data(self, x, t):
        self._x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.shape[0])
        self._t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.shape[3])
        self._call("sp_mlp_set_data", self._f(self._x), self._f(self._t), len(self._x))

    def train(self, nsteps, batch=4096, lr0=2e-3, lr_min=2e-5, seed=7) -> np.ndarray:
        losses = np.zeros(max(nsteps, 1), dtype=np.float32)
        self._call("sp_mlp_train", int(nsteps), int(batch), float(lr0), float(lr_min), int(seed),
                   self._f(losses))
        return losses[:nsteps]
