"""Shared fixtures-in-code for the parity tests: small seeded meshes and states."""
from __future__ import annotations

import numpy as np

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state

_CACHE = {}


def lattice_mesh(n=12, jitter=0.1, seed=123):
    key = ("lat", n, jitter, seed)
    if key not in _CACHE:
        pos = M.jittered_lattice(n, jitter=jitter, seed=seed)
        _CACHE[key] = M.build_periodic(pos)
    return _CACHE[key]


def workload_mesh(generators, n, seed=123):
    """The mesh of a workload (mesh.workload_mesh), built once per test process:
    the 256^3 mesh takes a minute or two to build and is shared by the tests
    that run on it."""
    key = ("workload", generators, n, seed)
    if key not in _CACHE:
        _CACHE[key] = M.workload_mesh(generators, n, seed=seed)
    return _CACHE[key]


def random_mesh(n=2000, seed=7):
    key = ("rand", n, seed)
    if key not in _CACHE:
        pos = M.uniform_random(n, seed=seed)
        _CACHE[key] = M.build_periodic(pos)
    return _CACHE[key]


def alfven_state(n=12, dt=2e-3, moving=False, seed=123):
    m = lattice_mesh(n, seed=seed)
    W = M.ic_alfven(m.pos)
    vel = None
    if moving:
        # generator velocity = fluid velocity plus a small seeded perturbation,
        # so the face-velocity correction (computeFluidUpdate.cpp:194-207) is exercised
        rng = np.random.default_rng(seed)
        vel = W[:, 1:4] + 0.01 * rng.standard_normal((m.ncell, 3))
    return global_step_state(m, W, dt=dt, vel=vel)


def perturb_gradients(st, scale=0.05, seed=5):
    """Non-trivial per-cell time derivatives Wt (computePredictor output)."""
    rng = np.random.default_rng(seed)
    st.t = scale * rng.standard_normal(st.t.shape) * np.maximum(np.abs(st.w), 1e-3)
    return st


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if scale is None:
        scale = np.maximum(np.abs(b), 1e-300)
    return np.max(np.abs(a - b) / scale) if a.size else 0.0


def max_ulp_f64(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return int(np.max(np.abs(ia - ib))) if a.size else 0
