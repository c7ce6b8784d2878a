"""Synthetic meshes and initial conditions for the Voronoi-face pipeline.

The reference builds its mesh with CGAL inside ``System::localMesh_build``
(localMesh.cpp:333-614) and draws generator positions with its own Rand48
(fvmhd3d.h:27-49).  Mesh construction is not part of the accelerated path; this
module wraps a host builder (``csrc/voronoi.cpp``) that produces the same
``Face`` / cell records (fvmhd3d.h:169-192, 262-284) for periodic boxes, plus the
initial states of the workloads the pipeline is measured on.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# fvmhd3d::Face as laid out by GCC: int s1, s2; double centroid[3]; double n[3] (56 B)
FACE_DTYPE = np.dtype([("s1", "<i4"), ("s2", "<i4"), ("centroid", "<f8", (3,)), ("n", "<f8", (3,))])
assert FACE_DTYPE.itemsize == 56

# primitive variable order (Fluid.h:22-33)
DENS, VELX, VELY, VELZ, ETHM, BX, BY, BZ, PSI, ENTR = range(10)
NFLUID = 10

_mesh_lib = None


def _lib():
    global _mesh_lib
    if _mesh_lib is None:
        path = os.path.join(_HERE, "libfvm_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(path)
        L.fvm_mesh_build_periodic.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.fvm_mesh_build_periodic.restype = ctypes.c_int
        L.fvm_mesh_nface.argtypes = [ctypes.c_void_p]
        L.fvm_mesh_nface.restype = ctypes.c_longlong
        L.fvm_mesh_nref.argtypes = [ctypes.c_void_p]
        L.fvm_mesh_nref.restype = ctypes.c_longlong
        L.fvm_mesh_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5
        L.fvm_mesh_copy.restype = ctypes.c_int
        L.fvm_mesh_free.argtypes = [ctypes.c_void_p]
        L.fvm_mesh_free.restype = None
        L.fvm_mesh_error.argtypes = []
        L.fvm_mesh_error.restype = ctypes.c_char_p
        L.fvm_rand48_drand.argtypes = [ctypes.c_longlong, ctypes.c_longlong, ctypes.c_void_p]
        L.fvm_rand48_drand.restype = ctypes.c_int
        _mesh_lib = L
    return _mesh_lib


def rand48(seed: int, n: int) -> np.ndarray:
    """``n`` successive ``Rand48::drand()`` values after ``srand(seed)`` (fvmhd3d.h:27-49)."""
    out = np.empty(n, dtype=np.float64)
    if _lib().fvm_rand48_drand(seed, n, out.ctypes.data) != 0:
        raise ValueError("rand48: bad arguments")
    return out


@dataclass
class Mesh:
    """One step's mesh, laid out like the reference hands it to the fluid kernels.

    faces      -- structured ``FACE_DTYPE`` array (the reference ``face_list``)
    cell_off   -- CSR offsets (n+1) of each cell's face list (``cell_list[i].ngb``)
    cell_face  -- face indices, ascending within each cell
    volume     -- cell volumes (``cell_list[i].Volume``)
    centroid   -- cell centroids (``cell_list[i].centroid``)
    pos        -- generator positions
    box        -- periodic box size (``global_domain_size``)
    """

    faces: np.ndarray
    cell_off: np.ndarray
    cell_face: np.ndarray
    volume: np.ndarray
    centroid: np.ndarray
    pos: np.ndarray
    box: np.ndarray

    @property
    def ncell(self) -> int:
        return len(self.volume)

    @property
    def nface(self) -> int:
        return len(self.faces)


def build_periodic(pos: np.ndarray, box=(1.0, 1.0, 1.0), nthreads: int = 0) -> Mesh:
    """Periodic Voronoi mesh of the sites ``pos`` (n x 3, inside [0, box))."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    box = np.ascontiguousarray(np.asarray(box, dtype=np.float64))
    n = pos.shape[0]
    L = _lib()
    handle = ctypes.c_void_p()
    rc = L.fvm_mesh_build_periodic(n, pos.ctypes.data, box.ctypes.data, int(nthreads), ctypes.byref(handle))
    if rc != 0:
        raise RuntimeError(L.fvm_mesh_error().decode())
    try:
        nf = L.fvm_mesh_nface(handle)
        nr = L.fvm_mesh_nref(handle)
        faces = np.empty(nf, dtype=FACE_DTYPE)
        cell_off = np.empty(n + 1, dtype=np.int32)
        cell_face = np.empty(nr, dtype=np.int32)
        volume = np.empty(n, dtype=np.float64)
        centroid = np.empty((n, 3), dtype=np.float64)
        L.fvm_mesh_copy(handle, faces.ctypes.data, cell_off.ctypes.data, cell_face.ctypes.data,
                        volume.ctypes.data, centroid.ctypes.data)
    finally:
        L.fvm_mesh_free(handle)
    return Mesh(faces, cell_off, cell_face, volume, centroid, pos, box)


_MESH_FIELDS = ("faces", "cell_off", "cell_face", "volume", "centroid", "pos", "box")


def workload_mesh(generators: str, n: int, seed: int = 123) -> Mesh:
    """The periodic unit-cube mesh of a workload: ``generators`` is "lattice" (n^3
    jittered lattice sites) or "random" (n^3 uniform sites), Rand48-seeded, sorted
    along the Morton curve.  When the environment variable FVM_MESH_CACHE names a
    directory, the mesh is stored there as .npy files on first use and loaded from
    there afterwards (a 256^3 mesh takes minutes to build and ~9 GB to store)."""
    cache = os.environ.get("FVM_MESH_CACHE")
    tag = os.path.join(cache, f"{generators}_{n}_{seed}") if cache else None
    if tag and os.path.exists(os.path.join(tag, "done")):
        return Mesh(*(np.load(os.path.join(tag, f + ".npy"), allow_pickle=False) for f in _MESH_FIELDS))
    if generators == "lattice":
        pos = jittered_lattice(n, seed=seed)
    elif generators == "random":
        pos = uniform_random(n ** 3, seed=seed)
    else:
        raise ValueError(f"unknown generators {generators!r}")
    m = build_periodic(pos[morton_order(pos, (1.0, 1.0, 1.0))])
    if tag:
        try:
            os.makedirs(tag, exist_ok=True)
            for f in _MESH_FIELDS:
                np.save(os.path.join(tag, f + ".npy"), getattr(m, f), allow_pickle=False)
            open(os.path.join(tag, "done"), "w").close()
        except OSError:  # a full disk only loses the cache, not the mesh
            import shutil
            shutil.rmtree(tag, ignore_errors=True)
    return m


# ----------------------------------------------------------------------------
# generator positions
# ----------------------------------------------------------------------------

def jittered_lattice(nx: int, ny: int | None = None, nz: int | None = None, jitter: float = 0.1,
                     box=(1.0, 1.0, 1.0), seed: int = 123) -> np.ndarray:
    """Cell-centred lattice with uniform jitter of +-jitter*dx per axis, Rand48-seeded."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    box = np.asarray(box, dtype=np.float64)
    dx = box / np.array([nx, ny, nz], dtype=np.float64)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([(i.ravel() + 0.5) * dx[0], (j.ravel() + 0.5) * dx[1], (k.ravel() + 0.5) * dx[2]], axis=1)
    r = rand48(seed, base.size).reshape(-1, 3)
    pos = base + (2.0 * r - 1.0) * jitter * dx
    return np.mod(pos, box)


def uniform_random(n: int, box=(1.0, 1.0, 1.0), seed: int = 123) -> np.ndarray:
    """n sites drawn uniformly with Rand48 as ``Problem_generate_geometry`` does (Problem.cpp:104-113)."""
    box = np.asarray(box, dtype=np.float64)
    r = rand48(seed, 3 * n).reshape(n, 3)
    return r * box


def morton_order(pos: np.ndarray, box) -> np.ndarray:
    """Permutation sorting sites along a Morton (Z-order) curve, 21 bits per axis."""
    box = np.asarray(box, dtype=np.float64)
    q = np.clip((pos / box * (1 << 21)).astype(np.int64), 0, (1 << 21) - 1).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v

    key = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


# ----------------------------------------------------------------------------
# initial primitive states (Fluid order, Fluid.h:22-33)
# ----------------------------------------------------------------------------

def entropy_from_ethm(W: np.ndarray, gamma: float) -> np.ndarray:
    """``Problem_compute_entropy_from_ethm`` (Problem.cpp:239-243)."""
    return (gamma - 1.0) * W[:, ETHM] / np.power(W[:, DENS], gamma)


def ic_alfven(pos: np.ndarray, gamma: float = 5.0 / 3.0, amp: float = 0.1) -> np.ndarray:
    """Circularly polarised Alfven wave, B0 = 1 along (1,1,1)/sqrt3, one wavelength
    per unit length along each axis (k = 2pi(1,1,1)), rho = 1, p = 0.1.

    The transverse pattern follows the reference's alfven.cpp:182-187
    (v_perp = B_perp = amp (sin phi, cos phi) in the plane normal to B0)."""
    e0 = np.array([1.0, 1.0, 1.0]) / math.sqrt(3.0)
    e1 = np.array([1.0, -1.0, 0.0]) / math.sqrt(2.0)
    e2 = np.array([1.0, 1.0, -2.0]) / math.sqrt(6.0)
    phi = 2.0 * math.pi * (pos[:, 0] + pos[:, 1] + pos[:, 2])
    perp = amp * (np.sin(phi)[:, None] * e1 + np.cos(phi)[:, None] * e2)
    W = np.zeros((pos.shape[0], NFLUID))
    W[:, DENS] = 1.0
    W[:, ETHM] = 0.1 / (gamma - 1.0)
    W[:, VELX:VELZ + 1] = perp
    W[:, BX:BZ + 1] = e0[None, :] + perp
    W[:, ENTR] = entropy_from_ethm(W, gamma)
    return W


def ic_blast(pos: np.ndarray, gamma: float = 5.0 / 3.0, centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """MHD blast: rho = 1, p = 100 inside r < 0.1 of the centre, 0.1 outside,
    B = (10/sqrt2)(1, 1, 0), v = 0."""
    r = np.linalg.norm(pos - np.asarray(centre)[None, :], axis=1)
    p = np.where(r < 0.1, 100.0, 0.1)
    W = np.zeros((pos.shape[0], NFLUID))
    W[:, DENS] = 1.0
    W[:, ETHM] = p / (gamma - 1.0)
    W[:, BX] = 10.0 / math.sqrt(2.0)
    W[:, BY] = 10.0 / math.sqrt(2.0)
    W[:, ENTR] = entropy_from_ethm(W, gamma)
    return W


def ic_orszag_tang(pos: np.ndarray, gamma: float = 5.0 / 3.0) -> np.ndarray:
    """3-D Orszag-Tang-type vortex (after Problem.cpp:118-191, with v_z = 0.1 sin 2pi z):
    rho = 25/(36pi), p = 5/(12pi), v = (-sin2piy, sin2pix, 0.1 sin2piz),
    B = (-sin2piy, sin4pix, 0)/sqrt(4pi)."""
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    W = np.zeros((pos.shape[0], NFLUID))
    W[:, DENS] = 25.0 / (36.0 * math.pi)
    W[:, ETHM] = (5.0 / (12.0 * math.pi)) / (gamma - 1.0)
    W[:, VELX] = -np.sin(2.0 * math.pi * y)
    W[:, VELY] = np.sin(2.0 * math.pi * x)
    W[:, VELZ] = 0.1 * np.sin(2.0 * math.pi * z)
    W[:, BX] = -b0 * np.sin(2.0 * math.pi * y)
    W[:, BY] = b0 * np.sin(4.0 * math.pi * x)
    W[:, ENTR] = entropy_from_ethm(W, gamma)
    return W
