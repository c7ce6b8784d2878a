"""ctypes binding of the C ABI in include/fvm_gpu.h (libfvm_gpu.so).

The binding fails loudly when the HIP library has not been built; there is no
CPU fallback anywhere on this path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvm_gpu.so")

FVM_OK, FVM_EARG, FVM_EHIP, FVM_EPOS, FVM_ESTATE = 0, -1, -2, -3, -4
LIMITER_SHIPPED, LIMITER_MINMOD, LIMITER_SMOOTH = 0, 1, 2
STORAGE_F64, STORAGE_SHIPPED = 0, 1
NO_BOUNDARY, REFLECTING, OUTFLOW, DIOD, INFLOW = 0, 1, 2, 3, 4

# per-cell fields: name -> (id, width, kind) ; kind: 'f' double, 's' float-as-shipped, 'i' int
FIELDS = {
    "w": (0, 10, "f"), "x": (1, 10, "s"), "y": (2, 10, "s"), "z": (3, 10, "s"), "t": (4, 10, "s"),
    "pos": (5, 3, "f"), "vel": (6, 3, "s"), "J": (7, 3, "s"), "tend": (8, 1, "f"), "bnd": (9, 1, "i"),
    "is_ngb": (10, 1, "i"), "is_active": (11, 1, "i"), "mm_lo": (12, 10, "s"), "mm_hi": (13, 10, "s"),
    "lim": (14, 10, "s"), "dU": (15, 10, "s"), "dgp": (16, 3, "s"), "ddivB": (17, 1, "s"),
    "U": (18, 10, "f"), "mvol": (19, 1, "f"), "tbeg": (20, 1, "f"), "mvel": (21, 3, "f"),
    "acc0": (22, 3, "f"), "acc1": (23, 3, "f"), "dt_new": (24, 1, "f"), "divB": (25, 1, "f"),
}
SUBSTEP_RUNG = len(FIELDS)  # rung per cell in fvm_substep_upload / _download
NDIAG = 8
FLUX_NOALLOC_MAX = 65536  # FVM_FLUX_NOALLOC_MAX of include/fvm_gpu.h
# reference-order kernels timed by fvm_kernel_timing, in FVM_KT_* order
KT_NAMES = ["k_reconstruction", "k_slope_limiter", "k_face_flux", "k_scatter", "k_primitives", "k_timestep",
            "k_predictor_face", "k_predictor_cell"]
DIAG_NAMES = ["positivity", "closure", "ethm_failed", "mass_failed", "dsl", "sg_fallback"]


class FvmParams(ctypes.Structure):
    _fields_ = [("gamma_gas", ctypes.c_double), ("courant_no", ctypes.c_double),
                ("domain_size", ctypes.c_double * 3), ("limiter_mode", ctypes.c_int),
                ("storage", ctypes.c_int)]


class FvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fvm status {code}: {msg}")
        self.code = code


_lib = None

# exported entry points and their argument types (also used by the symbol test)
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
SIGNATURES = {
    "fvm_create": [ctypes.POINTER(ctypes.c_void_p), _I, ctypes.POINTER(FvmParams)],
    "fvm_destroy": [_P],
    "fvm_set_params": [_P, ctypes.POINTER(FvmParams)],
    "fvm_set_time": [_P, _D],
    "fvm_upload_mesh": [_P, _I, _I, _I, _I, _P, _P, _P, _P],
    "fvm_upload_field": [_P, _I, _P, _I],
    "fvm_download_field": [_P, _I, _P, _I],
    "fvm_get_active_faces": [_P, ctypes.POINTER(ctypes.c_int)],
    "fvm_set_active_faces": [_P, _I, _P],
    "fvm_download_active_faces": [_P, _P, _I],
    "fvm_compute_reconstruction": [_P],
    "fvm_slope_limiter": [_P],
    "fvm_compute_fluid_update": [_P],
    "fvm_compute_primitives": [_P, _I],
    "fvm_compute_timestep": [_P, ctypes.POINTER(ctypes.c_double)],
    "fvm_compute_timestep_limiter": [_P],
    "fvm_compute_predictor": [_P, _D, _I],
    "fvm_compute_flux": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "fvm_compute_flux_noalloc": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "fvm_get_diagnostics": [_P, ctypes.POINTER(ctypes.c_longlong), _I],
    "fvm_reset_diagnostics": [_P],
    "fvm_synchronize": [_P],
    "fvm_fused_setup": [_P, _I, _I],
    "fvm_fused_run": [_P, _I, _D, ctypes.POINTER(ctypes.c_double)],
    "fvm_fused_dt_history": [_P, _P, _I],
    "fvm_fused_sync": [_P],
    "fvm_fused_timing": [_P, _I],
    "fvm_fused_kernel_ms": [_P, _P, _P, _I],
    "fvm_kernel_timing": [_P, _I],
    "fvm_set_face_precision": [_P, _I],
    "fvm_kernel_ms": [_P, _P, _P, _I],
    "fvm_measure_copy_bandwidth": [_P, ctypes.c_longlong, _I, ctypes.POINTER(ctypes.c_double)],
    "fvm_measure_runtime_copy_bandwidth": [_P, ctypes.c_longlong, _I, ctypes.POINTER(ctypes.c_double)],
    "fvm_get_face_states": [_P, _I, _I, _P, _P, _P, _P, _P],
    "fvm_energy": [_P, _P],
    "fvm_dump_binary": [_P, ctypes.c_char_p, _P, _P, _P],
    # learned HLLD surrogate (include/fvm_surrogate.h)
    "fvm_sg_create": [ctypes.POINTER(ctypes.c_void_p), _I, _D],
    "fvm_sg_destroy": [_P],
    "fvm_upload_faces": [_P, _I, _I, _I, _I, _P, _P],
    "fvm_download_cell_faces": [_P, _P, _P],
    "fvm_substep_setup": [_P, _I, _I, _P, _P, _P, _P],
    "fvm_timestep_limiter_sweep": [_I, _P, _P, _P, _P],
    "fvm_substep_upload": [_P, _I, _P, _I],
    "fvm_substep_download": [_P, _I, _P, _I],
    "fvm_substep_init_rungs": [_P, _D, _I, _P],
    "fvm_substep_run": [_P, ctypes.c_uint, _D, _D, _I, _I, _P, _P],
    "fvm_sg_set_params": [_P, _P, _I],
    "fvm_sg_get_params": [_P, _P, _I],
    "fvm_sg_set_gate": [_P, _P],
    "fvm_sg_set_loss_weights": [_P, _P, _I],
    "fvm_sg_add_faces": [_P, _I, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_int)],
    "fvm_sg_clear_samples": [_P],
    "fvm_sg_num_samples": [_P, ctypes.POINTER(ctypes.c_longlong)],
    "fvm_sg_get_samples": [_P, ctypes.c_longlong, _I, _P, _P],
    "fvm_sg_train": [_P, _I, _I, _D, _D, ctypes.c_ulonglong, _P],
    "fvm_sg_gradients": [_P, _I, _P, _P, _D, _P, _P, ctypes.POINTER(ctypes.c_double)],
    "fvm_sg_forward": [_P, _I, _P, _P],
    "fvm_sg_compute_flux": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_int)],
    "fvm_fused_set_surrogate": [_P, _P],
    "fvm_sg_compute_flux_device": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_int)],
    "fvm_sg_time_flux": [_P, _I, _P, _P, _P, _P, _P, _I, _P, ctypes.POINTER(ctypes.c_int)],
}
FUSED_FP64, FUSED_FP32 = 0, 1
ENERGY_NAMES = ("MASS", "ETHM", "EPOT", "ETOT", "EKINX", "EKINY", "EKINZ", "EKIN", "EMAGX", "EMAGY", "EMAGZ", "EMAG",
                "MOMX", "MOMY", "MOMZ", "MOM", "ANMX", "ANMY", "ANMZ", "ANM", "VOLUME")


class SnapshotInfo(ctypes.Structure):
    _fields_ = [("chare_index", ctypes.c_int), ("num_elements", ctypes.c_int), ("global_n", ctypes.c_int),
                ("iteration", ctypes.c_int), ("tsysU", ctypes.c_ulonglong), ("dt_global", ctypes.c_double),
                ("rmin", ctypes.c_double * 3), ("rmax", ctypes.c_double * 3)]


def lib():
    """Load libfvm_gpu.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP library must be built "
                               "(__graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        for name in ("fvm_last_error", "fvm_sg_last_error"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_char_p
        L.fvm_sg_destroy.restype = None
        _lib = L
    return _lib


def timestep_limiter_sweep(cell_off, nbr, active, dti):
    """fvm_timestep_limiter_sweep on host arrays (no GPU): returns the swept dti."""
    cell_off = np.ascontiguousarray(cell_off, dtype=np.int32)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    active = np.ascontiguousarray(active, dtype=np.uint8)
    out = np.array(dti, dtype=np.float64)
    rc = lib().fvm_timestep_limiter_sweep(len(cell_off) - 1, _ptr(cell_off), _ptr(nbr), _ptr(active), _ptr(out))
    if rc != FVM_OK:
        raise FvmError(rc, "fvm_timestep_limiter_sweep: bad arguments")
    return out


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class Context:
    """One fvm_ctx: the device-side counterpart of a System chare's arrays."""

    def __init__(self, gamma=5.0 / 3.0, courant=0.4, box=(1.0, 1.0, 1.0), limiter=LIMITER_SHIPPED,
                 storage=STORAGE_F64, device=0):
        self.L = lib()
        self.params = FvmParams(gamma, courant, (ctypes.c_double * 3)(*box), limiter, storage)
        h = ctypes.c_void_p()
        rc = self.L.fvm_create(ctypes.byref(h), device, ctypes.byref(self.params))
        if rc != FVM_OK:
            raise FvmError(rc, "fvm_create failed")
        self.h = h
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != FVM_OK:
            raise FvmError(rc, self.L.fvm_last_error(self.h).decode())
        return rc

    def set_time(self, t):
        self._chk(self.L.fvm_set_time(self.h, float(t)))

    def upload_mesh(self, np_, nactive, nimport_loc, faces, cell_off, cell_face, cvol):
        faces = np.ascontiguousarray(faces)
        cell_off = np.ascontiguousarray(cell_off, dtype=np.int32)
        cell_face = np.ascontiguousarray(cell_face, dtype=np.int32)
        cvol = np.ascontiguousarray(cvol, dtype=np.float64)
        self._chk(self.L.fvm_upload_mesh(self.h, np_, nactive, nimport_loc, len(faces), _ptr(faces),
                                         _ptr(cell_off), _ptr(cell_face), _ptr(cvol)))
        self.np = np_

    def upload_faces(self, np_, nactive, nimport_loc, faces, cvol):
        """Mesh upload with the cell lists built on the device (fvm_upload_faces)."""
        faces = np.ascontiguousarray(faces)
        cvol = np.ascontiguousarray(cvol, dtype=np.float64)
        self._chk(self.L.fvm_upload_faces(self.h, np_, nactive, nimport_loc, len(faces), _ptr(faces), _ptr(cvol)))
        self.np = np_
        self.nactive = nactive

    def cell_faces(self):
        """The uploaded mesh's cell lists (offsets, face indices)."""
        off = np.empty(self.nactive + 1, dtype=np.int32)
        self._chk(self.L.fvm_download_cell_faces(self.h, _ptr(off), None))
        cf = np.empty(max(int(off[-1]), 1), dtype=np.int32)
        self._chk(self.L.fvm_download_cell_faces(self.h, None, _ptr(cf)))
        return off, cf[:int(off[-1])]

    def upload(self, name, arr):
        fid, width, kind = FIELDS[name]
        dt = np.int32 if kind == "i" else np.float64
        a = np.ascontiguousarray(arr, dtype=dt).reshape(self.np, -1) if width > 1 else \
            np.ascontiguousarray(arr, dtype=dt).reshape(self.np)
        self._chk(self.L.fvm_upload_field(self.h, fid, _ptr(a), self.np))

    def download(self, name):
        fid, width, kind = FIELDS[name]
        dt = np.int32 if kind == "i" else np.float64
        a = np.empty((self.np, width) if width > 1 else (self.np,), dtype=dt)
        self._chk(self.L.fvm_download_field(self.h, fid, _ptr(a), self.np))
        return a

    def get_active_faces(self):
        n = ctypes.c_int()
        self._chk(self.L.fvm_get_active_faces(self.h, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.int32)
        self._chk(self.L.fvm_download_active_faces(self.h, _ptr(out), n.value))
        return out

    def set_active_faces(self, fa):
        fa = np.ascontiguousarray(fa, dtype=np.int32)
        self._chk(self.L.fvm_set_active_faces(self.h, len(fa), _ptr(fa)))

    def compute_reconstruction(self):
        self._chk(self.L.fvm_compute_reconstruction(self.h))

    def slope_limiter(self):
        self._chk(self.L.fvm_slope_limiter(self.h))

    def compute_fluid_update(self, raise_on_positivity=True):
        rc = self.L.fvm_compute_fluid_update(self.h)
        if rc == FVM_EPOS and not raise_on_positivity:
            return rc
        return self._chk(rc)

    def compute_primitives(self, predictor_flag=False):
        self._chk(self.L.fvm_compute_primitives(self.h, int(bool(predictor_flag))))

    def compute_timestep(self):
        d = ctypes.c_double()
        self._chk(self.L.fvm_compute_timestep(self.h, ctypes.byref(d)))
        return d.value

    def compute_timestep_limiter(self):
        self._chk(self.L.fvm_compute_timestep_limiter(self.h))

    def compute_predictor(self, dtmax, min_rung=0):
        self._chk(self.L.fvm_compute_predictor(self.h, float(dtmax), int(min_rung)))

    def compute_flux(self, wij, Wi, Wj, jbnd, normal, noalloc=False):
        """fvm_compute_flux (noalloc: fvm_compute_flux_noalloc, staged through the
        library's fixed per-GPU array, at most FLUX_NOALLOC_MAX faces)."""
        n = len(Wi)
        wij = np.ascontiguousarray(wij, dtype=np.float64)
        Wi = np.ascontiguousarray(Wi, dtype=np.float64)
        Wj = np.ascontiguousarray(Wj, dtype=np.float64)
        jbnd = np.ascontiguousarray(jbnd, dtype=np.int32)
        normal = np.ascontiguousarray(normal, dtype=np.float64)
        psi = np.empty(n)
        Bn = np.empty(n)
        flux = np.empty((n, 10))
        fn = self.L.fvm_compute_flux_noalloc if noalloc else self.L.fvm_compute_flux
        self._chk(fn(self.h, n, _ptr(wij), _ptr(Wi), _ptr(Wj), _ptr(jbnd), _ptr(normal), _ptr(psi), _ptr(Bn), _ptr(flux)))
        return psi, Bn, flux

    def face_states(self, first, count):
        """computeFlux inputs of active faces [first, first+count): (wij, Wi, Wj, jbnd, normal)."""
        wij, Wi, Wj = np.empty((count, 3)), np.empty((count, 10)), np.empty((count, 10))
        jbnd, normal = np.empty(count, np.int32), np.empty((count, 3))
        self._chk(self.L.fvm_get_face_states(self.h, int(first), int(count), _ptr(wij), _ptr(Wi), _ptr(Wj),
                                             _ptr(jbnd), _ptr(normal)))
        return wij, Wi, Wj, jbnd, normal

    def energy(self):
        """Energy totals (Energy.h:27-79) as a dict in Energy's order."""
        out = np.empty(len(ENERGY_NAMES))
        self._chk(self.L.fvm_energy(self.h, _ptr(out)))
        return dict(zip(ENERGY_NAMES, out))

    def dump_binary(self, path, dt_global, iteration=0, chare_index=0, num_elements=1, global_n=None, tsysU=0,
                    rmin=(0.0, 0.0, 0.0), rmax=(1.0, 1.0, 1.0), idx=None, dt_cell=None):
        """System::dump_binary (IO.cpp:7-99) of the active cells into `path`."""
        info = SnapshotInfo(chare_index, num_elements, self.np if global_n is None else global_n, iteration, tsysU,
                            dt_global, (ctypes.c_double * 3)(*rmin), (ctypes.c_double * 3)(*rmax))
        idx = None if idx is None else np.ascontiguousarray(idx, np.int32)
        dt_cell = None if dt_cell is None else np.ascontiguousarray(dt_cell, np.float64)
        self._chk(self.L.fvm_dump_binary(self.h, os.fsencode(path), ctypes.byref(info), _ptr(idx), _ptr(dt_cell)))

    def diagnostics(self):
        out = (ctypes.c_longlong * NDIAG)()
        self._chk(self.L.fvm_get_diagnostics(self.h, out, NDIAG))
        return {k: int(out[i]) for i, k in enumerate(DIAG_NAMES)}

    def reset_diagnostics(self):
        self._chk(self.L.fvm_reset_diagnostics(self.h))

    # device-resident stepping
    def fused_setup(self, precision=FUSED_FP32, moving=False):
        self._chk(self.L.fvm_fused_setup(self.h, int(precision), int(bool(moving))))

    def fused_set_surrogate(self, sg):
        """The learned HLLD surrogate (a surrogate.Surrogate, or None) in the
        flux step of the FP32 face path (fvm_fused_set_surrogate)."""
        self._chk(self.L.fvm_fused_set_surrogate(self.h, sg.h if sg is not None else None))
        self._sg = sg  # kept alive while set

    def fused_run(self, nsteps, dtmax):
        t = ctypes.c_double()
        self._chk(self.L.fvm_fused_run(self.h, int(nsteps), float(dtmax), ctypes.byref(t)))
        return t.value

    def fused_dt_history(self, n):
        out = np.empty(n)
        self._chk(self.L.fvm_fused_dt_history(self.h, _ptr(out), n))
        return out

    def fused_sync(self):
        self._chk(self.L.fvm_fused_sync(self.h))

    def fused_timing(self, enable=True):
        self._chk(self.L.fvm_fused_timing(self.h, int(bool(enable))))

    def fused_kernel_ms(self):
        ms = np.zeros(4)
        nl = np.zeros(4, dtype=np.int64)
        self._chk(self.L.fvm_fused_kernel_ms(self.h, _ptr(ms), _ptr(nl), 4))
        return ms, nl

    def set_face_precision(self, bits):
        """Face arithmetic of the fluid update and predictor calls: 64 (the
        reference's) or 32 (FP32 face states and HLLD, FP64 sums)."""
        self._chk(self.L.fvm_set_face_precision(self.h, int(bits)))

    def kernel_timing(self, enable=True):
        """HIP-event timing of the reference-order kernels (fvm_kernel_timing);
        enabling it resets the totals."""
        self._chk(self.L.fvm_kernel_timing(self.h, int(bool(enable))))

    def kernel_ms(self):
        """{kernel name: (total ms, launches)} since timing was enabled."""
        n = len(KT_NAMES)
        ms = np.zeros(n)
        nl = np.zeros(n, dtype=np.int64)
        self._chk(self.L.fvm_kernel_ms(self.h, _ptr(ms), _ptr(nl), n))
        return {k: (float(ms[i]), int(nl[i])) for i, k in enumerate(KT_NAMES)}

    def copy_bandwidth(self, nbytes=1 << 30, iters=20):
        g = ctypes.c_double()
        self._chk(self.L.fvm_measure_copy_bandwidth(self.h, int(nbytes), int(iters), ctypes.byref(g)))
        return g.value

    def runtime_copy_bandwidth(self, nbytes=1 << 30, iters=20):
        """The HIP runtime's own device-to-device copy, timed the same way."""
        g = ctypes.c_double()
        self._chk(self.L.fvm_measure_runtime_copy_bandwidth(self.h, int(nbytes), int(iters), ctypes.byref(g)))
        return g.value

    # individual timesteps with the state resident on the device
    def substep_setup(self, faces, cell_off, cell_face, volume):
        faces = np.ascontiguousarray(faces)
        cell_off = np.ascontiguousarray(cell_off, dtype=np.int32)
        cell_face = np.ascontiguousarray(cell_face, dtype=np.int32)
        volume = np.ascontiguousarray(volume, dtype=np.float64)
        self.nsub = len(volume)
        self._chk(self.L.fvm_substep_setup(self.h, self.nsub, len(faces), _ptr(faces), _ptr(cell_off),
                                           _ptr(cell_face), _ptr(volume)))

    def substep_upload(self, name, arr):
        if name == "rung":
            a = np.ascontiguousarray(arr, dtype=np.int32).reshape(self.nsub)
            fid = SUBSTEP_RUNG
        else:
            fid, width, kind = FIELDS[name]
            a = np.ascontiguousarray(arr, dtype=np.int32 if kind == "i" else np.float64)
        self._chk(self.L.fvm_substep_upload(self.h, fid, _ptr(a), self.nsub))

    def substep_download(self, name):
        if name == "rung":
            a = np.empty(self.nsub, dtype=np.int32)
            fid = SUBSTEP_RUNG
        else:
            fid, width, kind = FIELDS[name]
            a = np.empty((self.nsub, width) if width > 1 else (self.nsub,), dtype=np.int32 if kind == "i" else np.float64)
        self._chk(self.L.fvm_substep_download(self.h, fid, _ptr(a), self.nsub))
        return a

    def substep_init_rungs(self, dtmax, nrung):
        hist = np.zeros(32, dtype=np.int32)
        self._chk(self.L.fvm_substep_init_rungs(self.h, float(dtmax), int(nrung), _ptr(hist)))
        return hist

    def substep_run(self, rung_mask, t_global, dtmax, min_rung, nrung):
        counts = np.zeros(5, dtype=np.int32)
        hist = np.zeros(32, dtype=np.int32)
        self._chk(self.L.fvm_substep_run(self.h, int(rung_mask), float(t_global), float(dtmax), int(min_rung),
                                         int(nrung), _ptr(counts), _ptr(hist)))
        return counts, hist

    # convenience: whole state in / out
    def upload_state(self, st):
        self.upload_mesh(st.np, st.nactive, st.nimport_loc, st.faces, st.cell_off, st.cell_face, st.cvol)
        for name in FIELDS:
            self.upload(name, getattr(st, name))
        self.set_time(st.t_global)

    def download_state(self, st, names=None):
        for name in (names or FIELDS):
            setattr(st, name, self.download(name).reshape(getattr(st, name).shape))
        return st
