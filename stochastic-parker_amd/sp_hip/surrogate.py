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

    def last_forward_ms(self) -> float:
        """Device time of the layer kernels of the last forward call."""
        ms = C.c_double()
        self._call("sp_mlp_last_forward_ms", C.byref(ms))
        return ms.value

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


# ---------------------------------------------------------------------------
# Query features
LOG_KPARA0 = (-4.0, -1.0)
LOG_KRET = (-3.0, 0.0)
PINDEX = (1.0, 2.0)


def _scale(v, lo, hi):
    return 2.0 * (np.asarray(v, dtype=np.float64) - lo) / (hi - lo) - 1.0


def features(kpara0, kret, pindex, region, t, tmax):
    """(n, 6) network inputs; region is 0..3 (two bits), t in (0, tmax]."""
    region = np.asarray(region, dtype=np.int64)
    cols = [_scale(np.log10(kpara0), *LOG_KPARA0), _scale(np.log10(kret), *LOG_KRET),
            _scale(pindex, *PINDEX), 2.0 * (region & 1) - 1.0, 2.0 * ((region >> 1) & 1) - 1.0,
            _scale(t, 0.0, tmax)]
    return np.stack(np.broadcast_arrays(*cols), axis=1).astype(np.float32)


def in_domain(kpara0, kret, pindex, region):
    lk, lr = np.log10(kpara0), np.log10(kret)
    region = np.asarray(region)
    return ((lk >= LOG_KPARA0[0]) & (lk <= LOG_KPARA0[1]) & (lr >= LOG_KRET[0]) & (lr <= LOG_KRET[1]) &
            (np.asarray(pindex) >= PINDEX[0]) & (np.asarray(pindex) <= PINDEX[1]) &
            (region >= 0) & (region <= 3))


def spectra_from_log(logf):
    """Spectra from network outputs: 10^y - FLOOR, clipped at 0."""
    return np.clip(10.0 ** np.asarray(logf, dtype=np.float64) - FLOOR, 0.0, None)


def l1(a, b):
    """Per-row L1 distance of spectra (fractions of the injected particles)."""
    return np.abs(np.asarray(a) - np.asarray(b)).sum(axis=-1)


class SweepSDE:
    """The plain method: spectra of parameter sets from the batched pusher.

    Each query (kpara0, kret, pindex, region) injects `nptl` particles at p0 in
    its region of the 2D snapshot and records f(p) after every interval of
    `dt` for `ntimes` intervals: the fraction of the injected particles in each
    of `nbins` log bins over prange."""

    def __init__(self, grid=512, nptl=2048, dt=0.05, ntimes=8, nbins=64, prange=(0.5, 2.0),
                 precision="f32", max_particles=4e7, nodes=None):
        from . import workloads
        from .pusher import ParkerPusher
        self._pusher = ParkerPusher
        self.grid, self.nptl, self.dt, self.ntimes, self.nbins = grid, nptl, dt, ntimes, nbins
        self.cfg = workloads.config_2d(grid, pmin=prange[0], pmax=prange[1], dt_max_rel=1e-2,
                                       precision=precision)
        self.nodes = nodes if nodes is not None else workloads.fields_2d(grid, grid)
        self.boxes = np.array(workloads.SWEEP_REGIONS, dtype=np.float64)
        self.max_particles = int(max_particles)
        self.tmax = dt * ntimes
        self.times = dt * np.arange(1, ntimes + 1)
        self.steps = 0

    def run(self, kpara0, kret, pindex, region, nptl=None):
        """(nq, ntimes, nbins) spectra; chunks of queries share one launch."""
        nptl = int(nptl or self.nptl)
        kpara0, kret, pindex, region = (np.atleast_1d(np.asarray(a)) for a in (kpara0, kret, pindex, region))
        nq = len(kpara0)
        out = np.zeros((nq, self.ntimes, self.nbins))
        per = max(1, self.max_particles // nptl)
        for q0 in range(0, nq, per):
            q = slice(q0, min(nq, q0 + per))
            m = q.stop - q.start
            with self._pusher(**dict(self.cfg, nptl_max=m * nptl)) as pz:
                pz.upload_fields(self.nodes)
                pz.calc_fields_gradients()
                pz.init_particle_distributions(npp_global=self.nbins, nmu_global=1)
                pz.set_variants(list(zip(kpara0[q], kret[q], 3.0 - pindex[q])))
                pz.inject_variants(np.full(m, nptl), 1e-7, 1, 5.0, 0.0, 0.0, self.boxes[region[q]])
                for k in range(self.ntimes):
                    st = pz.particle_mover(k * self.dt, self.dt, 1, 10)
                    self.steps += st["steps"]
                    f, _ = pz.variant_distributions(m)
                    out[q, k] = f / nptl
        return out


def sample_queries(n, seed):
    """Parameter sets drawn like the batched sweep: log-uniform kpara0 and
    kret, uniform momentum exponent, one of four regions."""
    rng = np.random.default_rng(seed)
    return dict(kpara0=10.0 ** rng.uniform(*LOG_KPARA0, n), kret=10.0 ** rng.uniform(*LOG_KRET, n),
                pindex=rng.uniform(*PINDEX, n), region=rng.integers(0, 4, n))


def training_set(sde: SweepSDE, runs: dict):
    """(X, Y, F): features of every (run, time), targets log10(f + FLOOR), and
    the spectra themselves."""
    f = sde.run(runs["kpara0"], runs["kret"], runs["pindex"], runs["region"])
    nq, nt, nb = f.shape
    rep = lambda a: np.repeat(np.asarray(a), nt)
    t = np.tile(sde.times, nq)
    x = features(rep(runs["kpara0"]), rep(runs["kret"]), rep(runs["pindex"]), rep(runs["region"]),
                 t, sde.tmax)
    y = np.log10(f.reshape(nq * nt, nb) + FLOOR).astype(np.float32)
    return x, y, f


class Surrogate:
    """An ensemble of networks on log f(p); the spread between members is the
    per-query error estimate that decides the fallback to the pusher."""

    def __init__(self, members=2, nhidden=256, nlayers=4, nbins=64, seed=1, batch_max=65536):
        self.nets = [Mlp(6, nhidden, nlayers, nbins, batch_max=batch_max, seed=seed + 101 * i)
                     for i in range(members)]
        self.c0, self.c1 = 0.0, 1.0

    def close(self):
        for n in self.nets:
            n.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def fit(self, x, y, nsteps=3000, batch=4096, lr0=2e-3, lr_min=2e-5, seed=7):
        losses = []
        for i, n in enumerate(self.nets):
            n.set_data(x, y)
            losses.append(n.train(nsteps, batch, lr0, lr_min, seed + 17 * i))
        return losses

    def predict(self, x):
        """(spectra, error estimate) per row: the members' mean spectrum and
        the calibrated error estimate c0 + c1 * (L1 spread between members),
        both computed on the GPU."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = len(x)
        spec = np.zeros((n, self.nets[0].shape[3]), dtype=np.float32)
        spread = np.zeros(n, dtype=np.float32)
        hs = (C.c_void_p * len(self.nets))(*[m.h for m in self.nets])
        lib = self.nets[0].lib
        f = Mlp._f
        check(lib, "sp_mlp_ensemble_spectra",
              lib.sp_mlp_ensemble_spectra(hs, len(self.nets), f(x), n, FLOOR, f(spec), f(spread)))
        return spec, self.estimate(spread)

    def predict_device(self, x, spec, spread, rows=None):
        """predict() on rows held in device memory: x (rows, nin) float32,
        spec (rows, nout) and spread (rows,) float32 are device arrays
        (sp_hip.device.DeviceArray) that receive the members' mean spectrum
        and their L1 spread; the error estimate is estimate(spread)."""
        nin, nout = self.nets[0].shape[0], self.nets[0].shape[3]
        n = x.shape[0] if rows is None else int(rows)
        if (len(x.shape) != 2 or x.shape[1] != nin or x.shape[0] < n or spec.shape[0] < n or
                spec.shape[1:] != (nout,) or spread.shape[0] < n or
                any(a.dtype != np.float32 for a in (x, spec, spread))):
            raise ValueError("device arrays of the wrong shape or dtype")
        hs = (C.c_void_p * len(self.nets))(*[m.h for m in self.nets])
        lib = self.nets[0].lib
        check(lib, "sp_mlp_ensemble_spectra_device",
              lib.sp_mlp_ensemble_spectra_device(hs, len(self.nets), x.ptr, n, FLOOR, spec.ptr,
                                                 spread.ptr))

    def estimate(self, spread):
        """Calibrated error estimate c0 + c1 * spread."""
        return self.c0 + self.c1 * np.asarray(spread)

    def calibrate(self, x, f_true, quantile=0.95):
        """Fits the error estimate on held-out rows: err ~ c0 + c1 spread by
        least squares (c1 >= 0), then raises c0 until the estimate covers the
        true L1 error in `quantile` of the rows.  Members share part of their
        error, which the spread alone cannot see; c0 carries it.  On the
        512^2 sweep at tolerance 0.05 (max over a query's eight times), 0.95
        keeps 52 % of held-out queries on the network with 99.3 % of them
        within the tolerance (0.90: 80 % and 96.6 %; 0.98: 2 %)."""
        self.c0, self.c1 = 0.0, 1.0
        f, spread = self.predict(x)
        err = l1(f, f_true)
        a = np.stack([np.ones_like(spread), spread], axis=1).astype(np.float64)
        c0, c1 = np.linalg.lstsq(a, err, rcond=None)[0]
        if c1 < 0:
            c0, c1 = float(np.mean(err)), 0.0
        resid = err - (c0 + c1 * spread)
        self.c0, self.c1 = float(c0 + np.quantile(resid, quantile)), float(c1)
        return self.c0, self.c1


class SpectrumService:
    """Answers spectrum queries with the surrogate; a query outside the
    trained domain or whose error estimate exceeds `tol` is answered by the
    pusher (the same SweepSDE that made the training data)."""

    def __init__(self, surrogate: Surrogate, sde: SweepSDE, tol=0.05):
        self.s, self.sde, self.tol = surrogate, sde, tol

    def answer(self, kpara0, kret, pindex, region):
        """(spectra (nq, ntimes, nbins), used_sde (nq,) bool)."""
        kpara0, kret, pindex, region = (np.atleast_1d(np.asarray(a)) for a in (kpara0, kret, pindex, region))
        nq, nt = len(kpara0), self.sde.ntimes
        out = np.zeros((nq, nt, self.sde.nbins))
        ok = in_domain(kpara0, kret, pindex, region)
        use_sde = ~ok
        idx = np.nonzero(ok)[0]
        if len(idx):
            rep = lambda a: np.repeat(a[idx], nt)
            x = features(rep(kpara0), rep(kret), rep(pindex), rep(region), np.tile(self.sde.times, len(idx)),
                         self.sde.tmax)
            f, est = self.s.predict(x)
            f = f.reshape(len(idx), nt, -1)
            est = est.reshape(len(idx), nt).max(axis=1)
            out[idx] = f
            use_sde[idx[est > self.tol]] = True
        j = np.nonzero(use_sde)[0]
        if len(j):
            out[j] = self.sde.run(kpara0[j], kret[j], pindex[j], region[j])
        return out, use_sde
