"""Global steps on a moving mesh: the mesh side of the boundary, per step.

The reference moves its generators every step and rebuilds the Voronoi mesh
(moveParticles.cpp, localMesh_build, localMesh.cpp:333-614) before the fluid
update.  Here the host builds each step's mesh at the moved generators
(``mesh.build_periodic``), and a backend runs one global step of the path on
it.  With the HIP library behind it (``GpuStep``) the per-step mesh goes to the
device as the reference's face list only: ``fvm_upload_faces`` builds the
cell lists (cell_list[i].ngb, localMesh.cpp:586-589) and the SoA face records
on the device, every step.

Between steps the conserved state U of every cell carries over (a cell is its
generator); W is U over the new mesh's volume (Fluid::to_primitive,
Fluid.h:61-98).  The generator velocity is computePvel's (computePvel.cpp:6-54):
the fluid velocity plus a pull towards the cell's centroid.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, build_periodic
from .state import global_step_state, to_primitive


def compute_pvel(mesh: Mesh, W: np.ndarray, gamma: float) -> np.ndarray:
    """System::computePvel (computePvel.cpp:6-54) for every cell of a periodic
    mesh without boundary cells: v_mesh = v + f (c - r), with c - r the
    offset of the cell centroid from its generator (wrapped periodically), d
    its length, R = (3V/4pi)^(1/3), and f = 0 below d = 0.9 eta R, rising
    linearly to 1 at 1.1 eta R (eta = 0.25), times ki/tau with tau =
    d / sqrt(c_f^2 + |v|^2), c_f^2 = (gamma p + |B|^2)/rho."""
    box = np.asarray(mesh.box, dtype=np.float64)
    vel = W[:, 1:4].copy()
    B2 = np.sum(W[:, 5:8] ** 2, axis=1)
    pres = (gamma - 1.0) * W[:, 4] if gamma > 1.0 else W[:, 4]
    cs2 = (gamma * pres + B2) / W[:, 0]
    vabs = np.sqrt(cs2 + np.sum(vel ** 2, axis=1))
    cen = mesh.centroid - mesh.pos
    cen -= np.round(cen / box) * box
    d = np.sqrt(np.sum(cen ** 2, axis=1))
    eta, ki, f1, f2 = 0.25, 1.0, 0.9, 1.1
    R = np.power(mesh.volume * (3.0 / (4.0 * np.pi)), 1.0 / 3.0)
    fac = d / (eta * R)
    f = np.where(fac < f1, 0.0, np.where(fac < f2, (d - f1 * eta * R) / ((f2 - f1) * eta * R), 1.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        tau = d / vabs
        corr = np.where(d > 0.0, f * ki / tau, 0.0)
    return vel + cen * corr[:, None]


class MovingMesh:
    """Generators, conserved state and time of a moving-mesh run; ``step`` runs
    one global step on the mesh at the current generators and moves them."""

    def __init__(self, pos, W, gamma=5.0 / 3.0, courant=0.4, box=(1.0, 1.0, 1.0)):
        from .state import to_conservative
        self.box = np.asarray(box, dtype=np.float64)
        self.pos = np.mod(np.asarray(pos, dtype=np.float64), self.box)
        self.gamma, self.courant = gamma, courant
        self.mesh = build_periodic(self.pos, self.box)
        self.U = to_conservative(np.asarray(W, dtype=np.float64), self.mesh.volume)
        self.t = 0.0
        self.dts = []

    def state(self):
        """The global-step state on the current mesh: W from U and the mesh's
        volumes, the generator velocity from computePvel."""
        W = to_primitive(self.U, self.mesh.volume)
        vel = compute_pvel(self.mesh, W, self.gamma)
        return global_step_state(self.mesh, W, dt=0.0, gamma=self.gamma, courant=self.courant,
                                 t_global=self.t, vel=vel)

    def step(self, backend, dtmax, move_with=None):
        """One global step of ``backend`` (``step(state, dtmax) -> (dt, U)``) on the
        current mesh; then the generators move by their velocity times dt and
        the mesh is rebuilt.  ``move_with`` (another MovingMesh's last state and
        dt) moves them with that run's velocities instead, so that two backends
        can be compared on the same sequence of meshes."""
        st = self.state()
        dt, U = backend.step(st, dtmax)
        self.U = U
        self.t += dt
        self.dts.append(dt)
        vel, dtm = (st.vel, dt) if move_with is None else move_with
        self.last = (st.vel, dt)
        self.pos = np.mod(self.pos + vel * dtm, self.box)
        self.mesh = build_periodic(self.pos, self.box)
        return dt


class GpuStep:
    """One global step through the reference-order entry points of the HIP
    library (the calls of Iterate.cpp:92-202 for every cell active), with the
    step's mesh uploaded as its face list (fvm_upload_faces)."""

    def __init__(self, gamma=5.0 / 3.0, courant=0.4, box=(1.0, 1.0, 1.0), limiter=0):
        from . import gpu as G
        self.G = G
        self.ctx = G.Context(gamma=gamma, courant=courant, box=tuple(float(b) for b in box), limiter=limiter)

    def step(self, st, dtmax):
        from .surrogate import rung_dt
        c = self.ctx
        c.upload_faces(st.np, st.nactive, st.nimport_loc, st.faces, st.cvol)
        for name in self.G.FIELDS:
            c.upload(name, getattr(st, name))
        c.set_time(st.t_global)
        c.get_active_faces()
        c.compute_timestep()
        c.compute_predictor(dtmax, 0)
        dt = rung_dt(dtmax, float(np.min(c.download("dt_new"))))
        t0 = st.t_global
        c.upload("tbeg", np.full(st.np, t0))
        c.upload("tend", np.full(st.np, t0 + dt))
        c.compute_reconstruction()
        c.slope_limiter()
        c.compute_fluid_update()
        c.set_time(t0 + dt)
        c.compute_primitives(False)
        return dt, c.download("U").reshape(st.np, 10)

    def close(self):
        self.ctx.close()
