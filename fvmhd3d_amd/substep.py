"""Individual timesteps: the host side around the path.

The reference keeps the rung bookkeeping on the host (Scheduler.h, Iterate.cpp);
the path functions run on the active subset of each substep in the reference's
index layout (active cells, then their inactive neighbours; localMesh.cpp:216-322).
This module is that host side for a static periodic mesh:

* ``RungClock``   power-of-two rungs on an integer tick clock with the
                  Scheduler's semantics (Scheduler.h:4-206): dt(rung) =
                  dtmax/2^rung, get_rung from exponent bits, pull of every rung
                  whose boundary is reached, rung changes limited to rungs that
                  are synchronised now (min_rung).
* ``substep_layout`` the per-substep reference layout and mesh views.
* ``Substepper``  one substep of Iterate.cpp:92-202 (GetActiveList,
                  ComputeFluidUpdate .. computeFluidUpdate) against a backend:
                  either the HIP library (``GpuBackend``) or any object with the
                  same methods (the oracle, in the tests).
"""
from __future__ import annotations

import math
import struct

import numpy as np

from .state import CellState, to_conservative

RUNGMAX = 32


def exp_of(dt: float) -> int:
    return (struct.unpack("<Q", struct.pack("<d", dt))[0] >> 52) & 0xFFF


class RungClock:
    """Power-of-two rungs on an integer tick clock (Scheduler.h)."""

    def __init__(self, dtmax: float):
        self.dtmax = dtmax
        self.dt_tick = dtmax / (1 << (RUNGMAX - 1))
        self.tsysU = 0
        self.min_rung = 0

    def get_rung(self, dt: float, floor_rung: int | None = None) -> int:
        r = max(exp_of(self.dtmax) - exp_of(dt), self.min_rung if floor_rung is None else floor_rung)
        assert r < RUNGMAX
        return r

    def get_dt(self, rung: int) -> float:
        return self.dt_tick * float(1 << (RUNGMAX - 1 - rung))

    def rungs_of(self, dt: np.ndarray, nrung: int) -> np.ndarray:
        """get_rung over an array, clamped to [min_rung, nrung - 1] as the
        driver clamps (Iterate.cpp:146-168); the exponent is read from the bits."""
        e = (np.ascontiguousarray(dt, dtype=np.float64).view(np.uint64) >> np.uint64(52)) & np.uint64(0xFFF)
        r = np.maximum(exp_of(self.dtmax) - e.astype(np.int64), self.min_rung)
        assert np.all(r < RUNGMAX)
        return np.maximum(np.minimum(r, nrung - 1), self.min_rung)

    def dts_of(self, rung: np.ndarray) -> np.ndarray:
        """get_dt over an array: dt_tick * 2^(RUNGMAX-1-rung), exact scalings."""
        return np.ldexp(self.dt_tick, (RUNGMAX - 1 - np.asarray(rung, dtype=np.int64)).astype(np.int32))

    def tsys(self) -> float:
        return self.dt_tick * self.tsysU

    def advance(self, max_rung: int) -> tuple[int, float]:
        """Advance to the next synchronisation time of the finest occupied rung;
        return (bitmask of the active rungs, bit r = rung r; dt_global)."""
        assert max_rung != RUNGMAX - 1
        dtU = 1 << (RUNGMAX - 1 - max_rung)
        tnextU = (self.tsysU // dtU + 1) * dtU
        flags = self.tsysU ^ tnextU
        mask = 0
        for i in range(RUNGMAX):
            if flags & (1 << i):
                mask |= 1 << (RUNGMAX - 1 - i)
                self.min_rung = RUNGMAX - 1 - i
        prev = self.tsysU
        self.tsysU = tnextU
        return mask, self.dt_tick * (self.tsysU - prev)

    def pull(self, rung: np.ndarray) -> tuple[np.ndarray, float]:
        """Advance to the next synchronisation time; return (active mask, dt_global)."""
        mask, dt = self.advance(int(rung.max()))
        return ((mask >> np.asarray(rung, dtype=np.int64)) & 1).astype(bool), dt


def substep_layout(mesh, active: np.ndarray, bnd: np.ndarray, order: np.ndarray | None = None):
    """Reference layout of one substep.

    Returns (perm, nactive, nimport_loc, faces, cell_off, cell_face, cvol, is_ngb)
    where perm maps local index -> mesh cell: the active cells first, in
    ``order`` (the active list; mesh order when None), then the inactive cells
    that share a face with an active cell, in mesh order.  Faces are the faces
    of active cells, oriented so that s1 is active (the creating cell,
    localMesh.cpp:497-591) with the centroid in s1's frame."""
    A = np.nonzero(active)[0] if order is None else np.asarray(order, dtype=np.int64)
    n = mesh.ncell
    starts, ends = mesh.cell_off[A], mesh.cell_off[A + 1]
    lens = ends - starts
    idx = np.repeat(starts, lens) + (np.arange(lens.sum()) - np.repeat(np.cumsum(lens) - lens, lens))
    # faces of active cells, ascending, and the inactive cells they touch
    # (boolean marks: linear in the mesh size, no sort)
    fmark = np.zeros(mesh.nface, dtype=bool)
    fmark[mesh.cell_face[idx]] = True
    fl = np.nonzero(fmark)[0]
    # the face records as raw rows (a structured gather is several times slower)
    rows = mesh.faces.view(np.uint8).reshape(mesh.nface, mesh.faces.dtype.itemsize)
    faces = rows[fl].view(mesh.faces.dtype).reshape(len(fl))
    s1 = faces["s1"].astype(np.int64)
    s2 = faces["s2"].astype(np.int64)
    cmark = np.zeros(n, dtype=bool)
    cmark[s1] = True
    cmark[s2] = True
    inact = np.nonzero(cmark & ~active)[0]
    perm = np.concatenate([A, inact])
    g2l = np.full(n, -1, dtype=np.int64)
    g2l[perm] = np.arange(len(perm))
    # orient: s1 must be active; swap when only s2 is
    swap = ~active[s1] & active[s2]
    if np.any(swap):
        sw = np.nonzero(swap)[0]
        box = np.asarray(mesh.box)
        p_new = mesh.pos[s2[sw]]
        d = faces["centroid"][sw] - p_new
        d -= np.round(d / box) * box
        faces["centroid"][sw] = p_new + d
        faces["n"][sw] = -faces["n"][sw]
    faces["s1"] = g2l[np.where(swap, s2, s1)]
    faces["s2"] = g2l[np.where(swap, s1, s2)]
    # cell lists of the active cells, local face ids in ascending mesh-face order
    f2l = np.full(mesh.nface, -1, dtype=np.int64)
    f2l[fl] = np.arange(len(fl))
    cell_face = f2l[mesh.cell_face[idx]].astype(np.int32)
    cell_off = np.zeros(len(A) + 1, dtype=np.int32)
    cell_off[1:] = np.cumsum(lens)
    cvol = mesh.volume[A].copy()
    is_ngb = np.zeros(len(perm), dtype=np.int32)
    is_ngb[len(A):] = 1
    return perm, len(A), len(inact), faces, cell_off, cell_face, cvol, is_ngb


class Substepper:
    """Individual-timestep driver for a static periodic mesh (global arrays in mesh order).

    State kept per mesh cell, as the reference keeps it in ptcl_list / mesh_pnts /
    Wrec_list / U_list / dU_list: W, U, gradients, Wt, min/max, dU (with
    gradPsi, divB), rung, tbeg, tend, dt_new."""

    def __init__(self, mesh, W, gamma=5.0 / 3.0, courant=0.4, dtmax=1.0 / 64, limiter=0, nrung=6):
        n = mesh.ncell
        self.mesh = mesh
        # the Scheduler's bucket order (Scheduler.h push_particle / pull_active_list):
        # each cell's push stamp; a pull lists the finest rung first, each rung's
        # cells in push order, and the active cells are pushed again, in that
        # order, at the end of their step (Iterate.cpp:146-168)
        self.seq = np.arange(n, dtype=np.int64)
        self.next_seq = n
        self.gamma, self.courant, self.limiter, self.nrung = gamma, courant, limiter, nrung
        self.clock = RungClock(dtmax)
        self.W = np.array(W, dtype=np.float64)
        self.U = to_conservative(self.W, mesh.volume)
        z10 = lambda: np.zeros((n, 10))
        self.x, self.y, self.z, self.t = z10(), z10(), z10(), z10()
        self.mm_lo, self.mm_hi = z10(), z10()
        self.J = np.zeros((n, 3))
        self.dU, self.dgp, self.ddivB = z10(), np.zeros((n, 3)), np.zeros(n)
        self.mvol = mesh.volume.copy()
        self.bnd = np.zeros(n, dtype=np.int32)
        self.rung = np.zeros(n, dtype=np.int64)
        self.tbeg = np.zeros(n)
        self.tend = np.zeros(n)
        self.dt_new = np.zeros(n)
        self.t_global = 0.0
        self.history = []

    # --- initial rungs from the CFL of every cell (System.cpp:255-300), but
    # per cell instead of the global minimum, clamped to nrung rungs
    def init_rungs(self, backend):
        st, perm = self._state(np.ones(self.mesh.ncell, dtype=bool))
        backend.load(st)
        backend.compute_timestep(st)
        backend.fetch(st, ["dt_new"])
        dtn = np.empty(self.mesh.ncell)
        dtn[perm] = st.dt_new
        self.rung[:] = self.clock.rungs_of(dtn, self.nrung)
        self.tend = self.tbeg + self.clock.dts_of(self.rung)
        # every cell pushed in index order (System.cpp:300-315)
        self.seq[:] = np.arange(self.mesh.ncell)
        self.next_seq = self.mesh.ncell

    def active_list(self, active):
        """The Scheduler's pull order of the active cells: finest rung first,
        each rung in push order."""
        A = np.nonzero(active)[0]
        return A[np.lexsort((self.seq[A], -self.rung[A]))]

    def _state(self, active, order=None):
        perm, nact, nimp, faces, coff, cface, cvol, is_ngb = substep_layout(self.mesh, active, self.bnd, order)
        st = CellState(np=len(perm), nactive=nact, nimport_loc=nimp, faces=faces, cell_off=coff,
                       cell_face=cface, cvol=cvol, t_global=self.t_global, gamma=self.gamma,
                       courant=self.courant, box=tuple(float(b) for b in self.mesh.box))
        from .gpu import FIELDS
        for name, (_, width, kind) in FIELDS.items():
            st.arrays[name] = np.zeros(len(perm), dtype=np.int32) if kind == "i" else \
                np.zeros((len(perm), width) if width > 1 else len(perm))
        st.w = self.W[perm].copy()
        st.U = self.U[perm].copy()
        for name in ("x", "y", "z", "t", "mm_lo", "mm_hi", "J", "dU", "dgp", "ddivB", "mvol", "bnd", "tbeg",
                     "tend", "dt_new"):
            setattr(st, name, getattr(self, name)[perm].copy())
        st.pos = self.mesh.pos[perm].copy()
        st.is_ngb = is_ngb
        st.is_active = np.zeros(len(perm), dtype=np.int32)
        st.is_active[:nact] = 1
        st.lim = np.ones((len(perm), 10))
        # local inactive cells: their dU is a fresh buffer that is added back
        # to the owner afterwards (localMesh.cpp:266, computeFluidUpdate.cpp:320-345)
        st.dU[nact:] = 0.0
        st.dgp[nact:] = 0.0
        st.ddivB[nact:] = 0.0
        return st, perm

    def substep(self, backend):
        """One substep; returns the number of active cells and active faces."""
        active, _ = self.clock.pull(self.rung)
        self.t_global = self.clock.tsys()
        assert np.all(self.tend[active] == self.t_global)
        order = self.active_list(active)
        self.last_active = order
        st, perm = self._state(active, order)
        nact = st.nactive
        # predict inactive local cells to the current time (Iterate.cpp:99-106), on the copy
        dtp = self.t_global - st.tbeg[nact:]
        st.w[nact:] = st.w[nact:] + st.t[nact:] * dtp[:, None]
        backend.load(st)
        nfa = backend.get_active_faces(st)
        backend.compute_primitives(st)
        backend.compute_timestep(st)
        backend.compute_timestep_limiter(st)
        backend.compute_predictor(st, self.clock.dtmax, self.clock.min_rung)
        backend.fetch(st, ["dt_new"])
        # new rungs of the active cells (Iterate.cpp:146-168)
        r = self.clock.rungs_of(st.dt_new[:nact], self.nrung)
        self.rung[perm[:nact]] = r
        st.tbeg[:nact] = self.t_global
        st.tend[:nact] = self.t_global + self.clock.dts_of(r)
        backend.put(st, ["tbeg", "tend"])
        backend.compute_reconstruction(st)
        backend.slope_limiter(st)
        backend.compute_fluid_update(st)
        backend.fetch(st, None)
        # write back: active cells keep everything; inactive neighbours add their dU
        A = perm[:nact]
        for name in ("W", "U", "x", "y", "z", "t", "mm_lo", "mm_hi", "J", "mvol", "tbeg", "tend", "dt_new",
                     "dU", "dgp", "ddivB"):
            src = getattr(st, "w" if name == "W" else name)
            getattr(self, name)[A] = src[:nact]
        I = perm[nact:]
        self.dU[I] += st.dU[nact:]
        self.dgp[I] += st.dgp[nact:]
        self.ddivB[I] += st.ddivB[nact:]
        # pushed again in the order of the active list
        self.seq[A] = self.next_seq + np.arange(nact)
        self.next_seq += nact
        self.history.append((self.t_global, nact, nfa))
        return nact, nfa


class GpuBackend:
    """The HIP library behind the Substepper (one upload/download per substep)."""

    def __init__(self, gamma, courant, box, limiter=0):
        from . import gpu as G
        self.ctx = G.Context(gamma=gamma, courant=courant, box=box, limiter=limiter)
        self.kernel_s = 0.0

    def load(self, st):
        self.ctx.upload_state(st)

    def put(self, st, names):
        for n in names:
            self.ctx.upload(n, getattr(st, n))

    def fetch(self, st, names):
        self.ctx.download_state(st, names)

    def get_active_faces(self, st):
        st.face_act = self.ctx.get_active_faces()
        return len(st.face_act)

    def _timed(self, fn, *a):
        import time
        t = time.perf_counter()
        r = fn(*a)
        self.kernel_s += time.perf_counter() - t
        return r

    def compute_primitives(self, st):
        self._timed(self.ctx.compute_primitives, False)

    def compute_timestep(self, st):
        self._timed(self.ctx.compute_timestep)

    def compute_timestep_limiter(self, st):
        self._timed(self.ctx.compute_timestep_limiter)

    def compute_predictor(self, st, dtmax, min_rung):
        self._timed(self.ctx.compute_predictor, dtmax, min_rung)

    def compute_reconstruction(self, st):
        self._timed(self.ctx.compute_reconstruction)

    def slope_limiter(self, st):
        self._timed(self.ctx.slope_limiter)

    def compute_fluid_update(self, st):
        self._timed(self.ctx.compute_fluid_update, False)


class DeviceSubstepper:
    """The Substepper with the per-cell state resident on the device.

    The same substeps (Iterate.cpp:92-202) as ``Substepper`` with ``GpuBackend``,
    but the layout, the gathers, the new rungs and the write-back run on the
    device (``fvm_substep_run``); the host keeps only the rung clock, fed by
    the device's rung histogram.  The mesh must be periodic (no boundary cells).
    face_bits = 32 runs the face states and HLLD of the fluid update and the
    predictor in FP32 (fvm_set_face_precision); 64 is the reference's FP64."""

    def __init__(self, mesh, W, gamma=5.0 / 3.0, courant=0.4, dtmax=1.0 / 64, limiter=0, nrung=6, face_bits=64):
        from . import gpu as G
        self.mesh = mesh
        self.gamma, self.courant, self.limiter, self.nrung = gamma, courant, limiter, nrung
        self.clock = RungClock(dtmax)
        self.ctx = G.Context(gamma=gamma, courant=courant, box=tuple(float(b) for b in mesh.box), limiter=limiter)
        self.ctx.set_face_precision(face_bits)
        self.ctx.substep_setup(mesh.faces, mesh.cell_off, mesh.cell_face, mesh.volume)
        W = np.array(W, dtype=np.float64)
        self.ctx.substep_upload("w", W)
        self.ctx.substep_upload("U", to_conservative(W, mesh.volume))
        self.ctx.substep_upload("pos", mesh.pos)
        self.ctx.substep_upload("mvol", mesh.volume)
        self.hist = np.zeros(RUNGMAX, dtype=np.int32)
        self.t_global = 0.0
        self.history = []
        self.local_counts = []  # per substep: active cells, active faces, inactive neighbours, local faces

    def init_rungs(self, backend=None):
        self.hist = self.ctx.substep_init_rungs(self.clock.dtmax, self.nrung)

    def max_rung(self) -> int:
        return int(np.nonzero(self.hist)[0].max())

    def substep(self, backend=None):
        """One substep; returns the number of active cells and active faces."""
        mask, _ = self.clock.advance(self.max_rung())
        self.t_global = self.clock.tsys()
        counts, self.hist = self.ctx.substep_run(mask, self.t_global, self.clock.dtmax, self.clock.min_rung, self.nrung)
        nact, nfa = int(counts[0]), int(counts[1])
        self.history.append((self.t_global, nact, nfa))
        self.local_counts.append(tuple(int(x) for x in counts[:4]))
        return nact, nfa

    def state(self, name):
        """A global per-cell array in mesh order ("W", "U", "rung", field names)."""
        return self.ctx.substep_download({"W": "w"}.get(name, name))


def ic_disk(pos, gamma=5.0 / 3.0, H=0.1, beta=400.0, noise=0.01, seed=7, rsoft=0.05, r_out=0.4):
    """Keplerian-like disk in the periodic unit cube, centred at (0.5, 0.5, 0.5):
    rho ~ exp(-z^2/2H^2) (floor 1e-2), v_phi ~ (r^2+rsoft^2)^(-1/4) tapered to zero
    between 0.75 r_out and r_out so the periodic images stay at rest, vertical B at
    plasma beta 400 (mid-plane), 1% seeded velocity noise."""
    from .mesh import DENS, VELX, VELY, VELZ, ETHM, BZ, ENTR, NFLUID, entropy_from_ethm
    x, y, z = pos[:, 0] - 0.5, pos[:, 1] - 0.5, pos[:, 2] - 0.5
    r = np.sqrt(x * x + y * y)
    rho = np.maximum(np.exp(-z * z / (2 * H * H)), 1e-2)
    p = 0.1 * rho
    s = np.clip((r_out - r) / (0.25 * r_out), 0.0, 1.0)
    taper = s * s * (3.0 - 2.0 * s)
    vphi = 0.3 * (r * r + rsoft * rsoft) ** -0.25 * taper
    rng = np.random.default_rng(seed)
    W = np.zeros((len(pos), NFLUID))
    W[:, DENS] = rho
    W[:, VELX] = -vphi * y / np.maximum(r, 1e-12)
    W[:, VELY] = vphi * x / np.maximum(r, 1e-12)
    W[:, VELX:VELZ + 1] += noise * (0.3 / math.sqrt(rsoft)) * rng.standard_normal((len(pos), 3))
    W[:, ETHM] = p / (gamma - 1.0)
    W[:, BZ] = math.sqrt(2.0 * 0.1 / beta)
    W[:, ENTR] = entropy_from_ethm(W, gamma)
    return W


def ic_thin_disk(pos, gamma=5.0 / 3.0, H=0.1, beta=400.0, noise=0.01, seed=7, rsoft=0.05, r_out=0.25,
                 v0=0.3, cs_out=0.05, cs_max=0.9):
    """Keplerian-like thin disk in the periodic unit cube, centred at (0.5, 0.5, 0.5),
    whose timesteps span the six rungs of the individual-timestep workload:
      rho = exp(-z^2/2H^2) (floor 1e-2);
      v_phi = v0 (r^2+rsoft^2)^(-1/4) (~ r^-1/2 outside the softening), tapered to
        zero between 0.75 r_out and r_out so the periodic images stay at rest;
      sound speed of a thin disk of scale height H, c_s = H Omega with
        Omega = v_K / sqrt(r^2+rsoft^2), capped at cs_max, tapered like v_phi and
        floored at cs_out outside the disk; p = rho c_s^2 / gamma;
      vertical B = sqrt(2 p_mid / beta), beta = 400 against the local mid-plane
        pressure (B_z depends on x, y only, so div B = 0);
      seeded velocity noise of 1 % of the local v_K.
    The signal speed falls from ~2.2 at the centre to ~0.05 outside the disk, so the
    CFL steps span a factor ~32 with a sharp lower end (no cell below the finest
    rung): on 128^3 jittered lattice sites with dtmax = 1/64, 81 % of the cells sit
    on the coarsest rung and each of the first 16 substeps updates 1.8-19 % of the
    faces."""
    from .mesh import DENS, VELX, VELY, VELZ, ETHM, BZ, ENTR, NFLUID, entropy_from_ethm
    x, y, z = pos[:, 0] - 0.5, pos[:, 1] - 0.5, pos[:, 2] - 0.5
    r = np.sqrt(x * x + y * y)
    rho = np.maximum(np.exp(-z * z / (2 * H * H)), 1e-2)
    s = np.clip((r_out - r) / (0.25 * r_out), 0.0, 1.0)
    taper = s * s * (3.0 - 2.0 * s)
    rs = np.sqrt(r * r + rsoft * rsoft)
    vk = v0 / np.sqrt(rs)
    vphi = vk * taper
    cs2 = np.maximum(np.minimum((H * vk / rs) ** 2, cs_max * cs_max) * taper, cs_out * cs_out)
    rng = np.random.default_rng(seed)
    W = np.zeros((len(pos), NFLUID))
    W[:, DENS] = rho
    W[:, VELX] = -vphi * y / np.maximum(r, 1e-12)
    W[:, VELY] = vphi * x / np.maximum(r, 1e-12)
    W[:, VELX:VELZ + 1] += noise * vk[:, None] * rng.standard_normal((len(pos), 3))
    W[:, ETHM] = rho * cs2 / gamma / (gamma - 1.0)
    W[:, BZ] = np.sqrt(2.0 * cs2 / gamma / beta)
    W[:, ENTR] = entropy_from_ethm(W, gamma)
    return W
