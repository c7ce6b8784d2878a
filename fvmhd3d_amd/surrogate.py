"""Host side of the learned HLLD surrogate (include/fvm_surrogate.h).

No counterpart in the reference: the surrogate stands in for
riemannSolver_HLLD (computeFlux.cpp:178-425) inside computeFlux
(computeFlux.cpp:6-148) and answers the same batch call as
``Context.compute_flux``; faces its gate rejects are answered by the FP64 HLLD.

Everything computes in the HIP library: featurisation, training (Adam over
bf16 MFMA GEMMs), inference and the fallback.  This module holds the handle,
the training-data harvest from real face states of the reference-order
pipeline, and parameter files (.npz of float32 arrays).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import gpu as G

N_IN, N_H, N_OUT = 16, 128, 8
NPARAM = 36232
SHAPES = (("W0", (N_IN, N_H)), ("b0", (N_H,)), ("W1", (N_H, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)),
          ("b2", (N_H,)), ("W3", (N_H, N_OUT)), ("b3", (N_OUT,)))
DEFAULT_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "hlld_surrogate.npz")


class Gate(ctypes.Structure):
    _fields_ = [("max_log_rho_jump", ctypes.c_double), ("max_log_p_jump", ctypes.c_double),
                ("max_vn", ctypes.c_double), ("max_log10_err", ctypes.c_double)]


def flatten(P):
    return np.concatenate([np.asarray(P[k], np.float32).ravel() for k, _ in SHAPES])


def unflatten(v):
    P, o = {}, 0
    for k, shp in SHAPES:
        n = int(np.prod(shp))
        P[k] = np.asarray(v[o:o + n], np.float64).reshape(shp)
        o += n
    assert o == NPARAM
    return P


def init_params(seed=1):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    P = {}
    for k, shp in SHAPES:
        if k.startswith("W"):
            lim = np.sqrt(6.0 / shp[0])
            P[k] = rng.uniform(-lim, lim, shp)
        else:
            P[k] = np.zeros(shp)
    return P


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class Surrogate:
    """One fvm_sg handle (device, stream, parameters, training set)."""

    def __init__(self, gamma=5.0 / 3.0, device=0, params=None,
                 gate=(0.5, 1.0, 3.0, -3.0)):
        self.L = G.lib()
        h = ctypes.c_void_p()
        rc = self.L.fvm_sg_create(ctypes.byref(h), device, float(gamma))
        if rc != G.FVM_OK:
            raise G.FvmError(rc, "fvm_sg_create failed")
        self.h = h
        self.set_params(init_params() if params is None else params)
        self.set_gate(*gate)

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != G.FVM_OK:
            raise G.FvmError(rc, self.L.fvm_sg_last_error(self.h).decode())

    # parameters and gate
    def set_params(self, P):
        v = flatten(P) if isinstance(P, dict) else np.ascontiguousarray(P, np.float32)
        self._chk(self.L.fvm_sg_set_params(self.h, _p(v), NPARAM))

    def params(self):
        v = np.empty(NPARAM, np.float32)
        self._chk(self.L.fvm_sg_get_params(self.h, _p(v), NPARAM))
        return unflatten(v)

    def set_gate(self, max_log_rho_jump, max_log_p_jump, max_vn, max_log10_err):
        self.gate = Gate(max_log_rho_jump, max_log_p_jump, max_vn, max_log10_err)
        self._chk(self.L.fvm_sg_set_gate(self.h, ctypes.byref(self.gate)))

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        P = self.params()
        np.savez(path, **{k: np.asarray(v, np.float32) for k, v in P.items()},
                 gate=np.array([self.gate.max_log_rho_jump, self.gate.max_log_p_jump, self.gate.max_vn,
                                self.gate.max_log10_err]),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=DEFAULT_PARAMS, gamma=5.0 / 3.0, device=0):
        with np.load(path, allow_pickle=False) as z:
            P = {k: z[k] for k, _ in SHAPES}
            gate = tuple(float(v) for v in z["gate"])
        return cls(gamma=gamma, device=device, params=P, gate=gate)

    # training set
    def add_faces(self, wij, Wi, Wj, jbnd, normal):
        a = [np.ascontiguousarray(x, dt) for x, dt in ((wij, np.float64), (Wi, np.float64), (Wj, np.float64),
                                                        (jbnd, np.int32), (normal, np.float64))]
        kept = ctypes.c_int()
        self._chk(self.L.fvm_sg_add_faces(self.h, len(a[1]), *[_p(x) for x in a], ctypes.byref(kept)))
        return kept.value

    def clear_samples(self):
        self._chk(self.L.fvm_sg_clear_samples(self.h))

    def num_samples(self):
        n = ctypes.c_longlong()
        self._chk(self.L.fvm_sg_num_samples(self.h, ctypes.byref(n)))
        return n.value

    def samples(self, first=0, n=None):
        n = self.num_samples() - first if n is None else n
        x, t = np.empty((n, N_IN), np.float32), np.empty((n, N_OUT), np.float32)
        self._chk(self.L.fvm_sg_get_samples(self.h, int(first), int(n), _p(x), _p(t)))
        return x, t

    def set_loss_weights(self, w):
        w = np.ascontiguousarray(w, np.float64)
        self._chk(self.L.fvm_sg_set_loss_weights(self.h, _p(w), 7))
        self.loss_w = w

    def standardize(self, max_rows=1_000_000):
        """Weight each flux output by the inverse variance of its target over the
        training set (mean weight 1), so every output is learned to a similar
        relative precision."""
        n = min(self.num_samples(), max_rows)
        _, t = self.samples(0, n)
        var = np.var(t[:, :7].astype(np.float64), axis=0) + 1e-12
        w = 1.0 / var
        self.set_loss_weights(w / w.mean())
        return self.loss_w

    def train(self, steps, batch=8192, lr=1e-3, lam=0.05, seed=7):
        hist = np.empty(max(steps, 1), np.float32)
        self._chk(self.L.fvm_sg_train(self.h, int(steps), int(batch), float(lr), float(lam), int(seed), _p(hist)))
        return hist[:steps]

    def gradients(self, x, t, lam=0.05):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n = len(x)
        y, g = np.empty((n, N_OUT), np.float32), np.empty(NPARAM, np.float32)
        loss = ctypes.c_double()
        self._chk(self.L.fvm_sg_gradients(self.h, n, _p(x), _p(t), float(lam), _p(y), _p(g), ctypes.byref(loss)))
        return y, unflatten(g), loss.value

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((len(x), N_OUT), np.float32)
        self._chk(self.L.fvm_sg_forward(self.h, len(x), _p(x), _p(y)))
        return y

    # the computeFlux batch
    def compute_flux(self, wij, Wi, Wj, jbnd, normal):
        a = [np.ascontiguousarray(x, dt) for x, dt in ((wij, np.float64), (Wi, np.float64), (Wj, np.float64),
                                                        (jbnd, np.int32), (normal, np.float64))]
        n = len(a[1])
        psi, Bn, flux = np.empty(n), np.empty(n), np.empty((n, 10))
        nfb = ctypes.c_int()
        self._chk(self.L.fvm_sg_compute_flux(self.h, n, *[_p(x) for x in a], _p(psi), _p(Bn), _p(flux),
                                             ctypes.byref(nfb)))
        return psi, Bn, flux, nfb.value

    def time_flux(self, wij, Wi, Wj, jbnd, normal, iters=10):
        """Average ms per launch: (surrogate + fallback, FP64 HLLD, FP32 HLLD), and the fallback count."""
        a = [np.ascontiguousarray(x, dt) for x, dt in ((wij, np.float64), (Wi, np.float64), (Wj, np.float64),
                                                        (jbnd, np.int32), (normal, np.float64))]
        ms = np.empty(3)
        nfb = ctypes.c_int()
        self._chk(self.L.fvm_sg_time_flux(self.h, len(a[1]), *[_p(x) for x in a], int(iters), _p(ms),
                                          ctypes.byref(nfb)))
        return ms, nfb.value


def rung_dt(dtmax, dt):
    """Scheduler::get_dt(get_rung(dt)) with min_rung = 0 (Scheduler.h:94-111)."""
    e = lambda v: (np.frombuffer(np.float64(v).tobytes(), np.uint64)[0] >> np.uint64(52)) & np.uint64(0xfff)
    rung = max(int(e(dtmax)) - int(e(dt)), 0)
    return dtmax / 2.0 ** 31 * 2.0 ** (31 - rung)


class FaceStates:
    """The computeFlux inputs of the active faces of state W on `mesh`, as
    computeFluidUpdate forms them after computeTimestep, computePredictor,
    computeReconstruction and slopeLimiter of one global step; read in
    chunks from one prepared reference-order context."""

    def __init__(self, mesh, W, limiter, t_global=0.0, dtmax=1.0):
        from .state import global_step_state
        st = global_step_state(mesh, W, dt=0.0, t_global=t_global)
        self.ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=limiter)
        c = self.ctx
        c.upload_state(st)
        self.nfa = len(c.get_active_faces())
        dt = rung_dt(dtmax, c.compute_timestep())
        c.upload("tend", np.full(st.np, t_global + dt))
        c.compute_predictor(dtmax, 0)
        c.compute_reconstruction()
        c.slope_limiter()

    def read(self, first=0, count=None):
        count = self.nfa - first if count is None else min(count, self.nfa - first)
        return self.ctx.face_states(first, count)

    def close(self):
        self.ctx.close()


def face_states(mesh, W, limiter, t_global=0.0, dtmax=1.0, first=0, count=None):
    """All (or a range of) the active faces' computeFlux inputs; see FaceStates."""
    fs = FaceStates(mesh, W, limiter, t_global, dtmax)
    try:
        return fs.read(first, count)
    finally:
        fs.close()


def evolve(mesh, W, limiter, nsteps, precision=G.FUSED_FP64, dtmax=1.0):
    """W after `nsteps` global steps of the device-resident path."""
    from .state import global_step_state
    st = global_step_state(mesh, W, dt=0.0)
    ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=limiter)
    try:
        ctx.upload_state(st)
        ctx.fused_setup(precision)
        t = ctx.fused_run(nsteps, dtmax) if nsteps else 0.0
        return ctx.download("w"), t
    finally:
        ctx.close()
