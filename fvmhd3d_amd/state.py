"""Host-side cell/face state of one step, in the reference's index layout.

``CellState`` holds, as numpy arrays, what the reference keeps behind the
System chare's pointer tables (fvmhd3d.h:343-406) for one step: the mesh
(face_list, cell_list[i].ngb, cell_list[i].Volume) and per-cell records
(Fluid_rec, MeshPoint, U_act, dU_act, limiters).  Arrays are arrays of the
reference's records: Fluid -> (n, 10) float64, vec3 -> (n, 3) float64.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

from .mesh import DENS, ETHM, ENTR, BX, BZ, VELX, VELZ, NFLUID, FACE_DTYPE, Mesh

MASS, MOMX, MOMY, MOMZ, ENER, WBX, WBY, WBZ, MPSI, MENTR = range(10)


def to_conservative(W, V):
    """Fluid::to_conservative (Fluid.h:100-135), vectorised."""
    V = np.asarray(V, dtype=np.float64)
    M = W[:, DENS] * V
    U = np.empty_like(W)
    U[:, MASS] = W[:, DENS] * V
    U[:, MOMX:MOMZ + 1] = W[:, VELX:VELZ + 1] * M[:, None]
    U[:, WBX:WBZ + 1] = W[:, BX:BZ + 1] * V[:, None]
    U[:, MPSI] = W[:, 8] * M
    U[:, MENTR] = W[:, ENTR] * M
    B2 = W[:, BX] ** 2 + W[:, 6] ** 2 + W[:, BZ] ** 2
    v2 = W[:, VELX] ** 2 + W[:, 2] ** 2 + W[:, VELZ] ** 2
    U[:, ENER] = W[:, ETHM] * V + 0.5 * B2 * V + 0.5 * v2 * M
    return U


def to_primitive(U, V):
    """Fluid::to_primitive (Fluid.h:61-98), vectorised."""
    V = np.asarray(V, dtype=np.float64)
    iM = 1.0 / U[:, MASS]
    iV = 1.0 / V
    W = np.empty_like(U)
    W[:, DENS] = U[:, MASS] * iV
    W[:, VELX:VELZ + 1] = U[:, MOMX:MOMZ + 1] * iM[:, None]
    W[:, BX:BZ + 1] = U[:, WBX:WBZ + 1] * iV[:, None]
    W[:, 8] = U[:, MPSI] * iM
    W[:, ENTR] = U[:, MENTR] * iM
    W[:, ETHM] = U[:, ENER] * iV - 0.5 * (W[:, BX] ** 2 + W[:, 6] ** 2 + W[:, BZ] ** 2) \
        - 0.5 * (W[:, VELX] ** 2 + W[:, 2] ** 2 + W[:, VELZ] ** 2) * W[:, DENS]
    return W


@dataclass
class CellState:
    np: int
    nactive: int
    nimport_loc: int
    faces: np.ndarray
    cell_off: np.ndarray
    cell_face: np.ndarray
    cvol: np.ndarray
    t_global: float = 0.0
    gamma: float = 5.0 / 3.0
    courant: float = 0.4
    box: tuple = (1.0, 1.0, 1.0)
    face_act: np.ndarray | None = None
    arrays: dict = field(default_factory=dict)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            arrays[name] = value
        else:
            object.__setattr__(self, name, value)

    def copy(self):
        return copy.deepcopy(self)


def _zeros(n, w):
    return np.zeros((n, w) if w > 1 else (n,), dtype=np.float64)


def global_step_state(mesh: Mesh, W: np.ndarray, dt: float, gamma=5.0 / 3.0, courant=0.4,
                      t_global: float = 0.0, vel: np.ndarray | None = None) -> CellState:
    """All cells active on one periodic domain, one global step of size ``dt``
    starting at ``t_global`` (tbeg = t_global - dt of the previous step is not
    needed: dU is zero).  ``vel`` is the mesh-generator velocity (w), zero for a
    static mesh."""
    n = mesh.ncell
    st = CellState(np=n, nactive=n, nimport_loc=0, faces=mesh.faces, cell_off=mesh.cell_off,
                   cell_face=mesh.cell_face, cvol=mesh.volume.copy(), t_global=t_global, gamma=gamma,
                   courant=courant, box=tuple(float(b) for b in mesh.box))
    from .gpu import FIELDS
    for name, (_, width, kind) in FIELDS.items():
        st.arrays[name] = np.zeros(n, dtype=np.int32) if kind == "i" else _zeros(n, width)
    st.w = np.array(W, dtype=np.float64)
    st.pos = mesh.pos.copy()
    st.mvol = mesh.volume.copy()
    st.U = to_conservative(st.w, st.mvol)
    st.tend = np.full(n, t_global + dt)
    st.tbeg = np.full(n, t_global)
    st.lim = np.ones((n, 10))
    st.is_active = np.ones(n, dtype=np.int32)
    if vel is not None:
        st.vel = np.array(vel, dtype=np.float64)
        st.mvel = np.array(vel, dtype=np.float64)
    return st
