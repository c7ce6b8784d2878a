"""ctypes driver of the CPU oracle (oracle/fvm_oracle.c).

TEST INFRASTRUCTURE ONLY -- the checker.  Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
product package.  Two builds: ``mode="f64"`` (all-FP64 reference) and
``mode="shipped"`` (the reference's float storage).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p

# per-cell fields, in o_state order: (name, width, kind) kind: f double, s storage real, i int
_CELL_FIELDS = [
    ("w", 10, "f"), ("x", 10, "s"), ("y", 10, "s"), ("z", 10, "s"), ("t", 10, "s"), ("pos", 3, "f"),
    ("vel", 3, "s"), ("J", 3, "s"), ("tend", 1, "f"), ("bnd", 1, "i"), ("is_ngb", 1, "i"),
    ("is_active", 1, "i"), ("mm_lo", 10, "s"), ("mm_hi", 10, "s"), ("lim", 10, "s"), ("dU", 10, "s"),
    ("dgp", 3, "s"), ("ddivB", 1, "s"), ("U", 10, "f"), ("mvol", 1, "f"), ("tbeg", 1, "f"),
    ("mvel", 3, "f"), ("acc0", 3, "f"), ("acc1", 3, "f"), ("dt_new", 1, "f"), ("divB", 1, "f"),
]


class OState(ctypes.Structure):
    _fields_ = ([("np", ctypes.c_int), ("nactive", ctypes.c_int), ("nimport_loc", ctypes.c_int),
                 ("t_global", ctypes.c_double), ("gamma_gas", ctypes.c_double), ("courant_no", ctypes.c_double),
                 ("L", ctypes.c_double * 3), ("limiter_mode", ctypes.c_int), ("pad0", ctypes.c_int)]
                + [(name, _P) for name, _, _ in _CELL_FIELDS]
                + [("nface", ctypes.c_int), ("nface_act", ctypes.c_int), ("faces", _P), ("cell_off", _P),
                   ("cell_face", _P), ("cvol", _P), ("face_act", _P)]
                + [(n, ctypes.c_longlong) for n in
                   ("n_pos_viol", "n_closure_viol", "n_ethm_failed", "n_mass_failed", "n_dsl_viol")])


def build():
    """Compile both oracle libraries (gcc, strict IEEE)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def _lib(mode, path=None):
    """The oracle library of ``mode``; ``path`` loads another build of the same
    source instead (the tests' planted-mistake builds)."""
    key = path or mode
    if key not in _libs:
        if path is None:
            path = os.path.join(HERE, f"liboracle_{mode}.so")
            if not os.path.exists(path):
                build()
        L = ctypes.CDLL(path)
        for fn in ("o_compute_reconstruction", "o_slope_limiter", "o_compute_fluid_update",
                   "o_compute_timestep", "o_compute_timestep_limiter"):
            getattr(L, fn).argtypes = [ctypes.POINTER(OState)]
            getattr(L, fn).restype = None
        L.o_compute_primitives.argtypes = [ctypes.POINTER(OState), ctypes.c_int]
        L.o_compute_primitives.restype = None
        L.o_compute_predictor.argtypes = [ctypes.POINTER(OState), ctypes.c_double, ctypes.c_int]
        L.o_compute_predictor.restype = None
        L.o_get_active_faces.argtypes = [ctypes.POINTER(OState), _P]
        L.o_get_active_faces.restype = ctypes.c_int
        L.o_compute_flux_batch.argtypes = [ctypes.POINTER(OState), ctypes.c_int] + [_P] * 8
        L.o_compute_flux_batch.restype = None
        L.o_step_global.argtypes = [ctypes.POINTER(OState), ctypes.c_double]
        L.o_step_global.restype = ctypes.c_double
        L.o_rung_dt.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.o_rung_dt.restype = ctypes.c_double
        if hasattr(L, "o_step_global_par"):
            L.o_step_global_par.argtypes = [ctypes.POINTER(OState), ctypes.c_double]
            L.o_step_global_par.restype = ctypes.c_double
            L.o_threads.restype = ctypes.c_int
        L.o_limiter_phi.argtypes = [ctypes.c_int, ctypes.c_double]
        L.o_limiter_phi.restype = ctypes.c_double
        L.o_sizeof_state.restype = ctypes.c_int
        assert L.o_sizeof_state() == ctypes.sizeof(OState), "o_state layout mismatch"
        _libs[key] = L
    return _libs[key]


class Oracle:
    """Runs the reference functions on a fvmhd3d_amd.state.CellState in place."""

    def __init__(self, mode="f64", limiter=0, path=None):
        """mode: "f64" (all-FP64), "shipped" (float storage) or "f64_omp" (all-FP64
        built with OpenMP: step_global_par runs on every thread it is given)."""
        assert mode in ("f64", "shipped", "f64_omp")
        self.mode = mode
        self.L = _lib(mode, path)
        self.sdtype = np.float32 if mode == "shipped" else np.float64
        self.limiter = limiter
        self.diag = dict(positivity=0, closure=0, ethm_failed=0, mass_failed=0, dsl=0)

    def _pack(self, st):
        keep = {}
        o = OState()
        o.np, o.nactive, o.nimport_loc = st.np, st.nactive, st.nimport_loc
        o.t_global, o.gamma_gas, o.courant_no = st.t_global, st.gamma, st.courant
        for d in range(3):
            o.L[d] = st.box[d]
        o.limiter_mode = self.limiter
        for name, width, kind in _CELL_FIELDS:
            src = getattr(st, name)
            dt = np.int32 if kind == "i" else (self.sdtype if kind == "s" else np.float64)
            a = np.ascontiguousarray(src, dtype=dt)
            keep[name] = a
            setattr(o, name, a.ctypes.data)
        faces = np.ascontiguousarray(st.faces)
        keep["faces"] = faces
        o.faces = faces.ctypes.data
        o.nface = len(faces)
        for name in ("cell_off", "cell_face"):
            a = np.ascontiguousarray(getattr(st, name), dtype=np.int32)
            keep[name] = a
            setattr(o, name, a.ctypes.data if a.size else None)
        cv = np.ascontiguousarray(st.cvol, dtype=np.float64)
        keep["cvol"] = cv
        o.cvol = cv.ctypes.data
        fa = st.face_act if st.face_act is not None else np.zeros(0, np.int32)
        fa = np.ascontiguousarray(fa, dtype=np.int32)
        keep["face_act"] = fa
        o.face_act = fa.ctypes.data if fa.size else None
        o.nface_act = len(fa)
        return o, keep

    def _unpack(self, st, o, keep):
        for name, width, kind in _CELL_FIELDS:
            a = keep[name]
            setattr(st, name, a.astype(np.int32 if kind == "i" else np.float64))
        self.diag["positivity"] += o.n_pos_viol
        self.diag["closure"] += o.n_closure_viol
        self.diag["ethm_failed"] += o.n_ethm_failed
        self.diag["mass_failed"] += o.n_mass_failed
        self.diag["dsl"] += o.n_dsl_viol

    def _run(self, st, fn, *args):
        o, keep = self._pack(st)
        fn(ctypes.byref(o), *args)
        self._unpack(st, o, keep)

    def get_active_faces(self, st):
        o, keep = self._pack(st)
        out = np.empty(max(1, len(st.faces)), dtype=np.int32)
        n = self.L.o_get_active_faces(ctypes.byref(o), out.ctypes.data)
        st.face_act = out[:n].copy()
        return st.face_act

    def compute_reconstruction(self, st):
        self._run(st, self.L.o_compute_reconstruction)

    def slope_limiter(self, st):
        self._run(st, self.L.o_slope_limiter)

    def compute_fluid_update(self, st):
        self._run(st, self.L.o_compute_fluid_update)

    def compute_primitives(self, st, predictor_flag=False):
        self._run(st, self.L.o_compute_primitives, int(bool(predictor_flag)))

    def compute_timestep(self, st):
        self._run(st, self.L.o_compute_timestep)
        return float(np.min(st.dt_new[:st.nactive])) if st.nactive else float("inf")

    def compute_timestep_limiter(self, st):
        self._run(st, self.L.o_compute_timestep_limiter)

    def compute_predictor(self, st, dtmax, min_rung=0):
        self._run(st, self.L.o_compute_predictor, float(dtmax), int(min_rung))

    def step_global(self, st, dtmax, nsteps=1):
        """nsteps of the global-timestep driver; returns the list of dt."""
        o, keep = self._pack(st)
        dts = []
        for _ in range(nsteps):
            dts.append(self.L.o_step_global(ctypes.byref(o), float(dtmax)))
        st.t_global = o.t_global
        self._unpack(st, o, keep)
        return dts

    def phi(self, mode, x):
        """The limiter function Phi of ``mode`` at x."""
        return self.L.o_limiter_phi(int(mode), float(x))

    def step_global_par(self, st, dtmax, nsteps=1):
        """nsteps of the global-timestep driver with the threaded passes (bit-identical
        to step_global); returns the list of dt."""
        o, keep = self._pack(st)
        dts = []
        for _ in range(nsteps):
            dts.append(self.L.o_step_global_par(ctypes.byref(o), float(dtmax)))
        st.t_global = o.t_global
        self._unpack(st, o, keep)
        return dts

    def threads(self):
        return self.L.o_threads()

    def rung_dt(self, dtmax, min_rung, dt):
        return self.L.o_rung_dt(dtmax, min_rung, dt)

    def compute_flux(self, gamma, wij, Wi, Wj, jbnd, normal):
        n = len(Wi)
        o = OState()
        o.gamma_gas = gamma
        arrs = [np.ascontiguousarray(a, dtype=dt) for a, dt in
                ((wij, np.float64), (Wi, np.float64), (Wj, np.float64), (jbnd, np.int32), (normal, np.float64))]
        psi, Bn, flux = np.empty(n), np.empty(n), np.empty((n, 10))
        self.L.o_compute_flux_batch(ctypes.byref(o), n, *[a.ctypes.data for a in arrs],
                                    psi.ctypes.data, Bn.ctypes.data, flux.ctypes.data)
        return psi, Bn, flux
