"""Mesh upload with the cell lists built on the device (fvm_upload_faces).

localMesh.cpp:586-589 lists every new face under its s1 and, when s2 is
active, under its s2, so each active cell's list is its faces in ascending
index.  The host mesh builder (fvmhd3d_amd/mesh.py) and the substep layout
(fvmhd3d_amd/substep.py) produce their lists independently of the device
code; the device-built lists must equal theirs exactly."""
import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.substep import substep_layout
from tests.helpers import random_mesh


@pytest.mark.gpu
def test_device_cell_lists_equal_host_lists_full_mesh():
    from fvmhd3d_amd import gpu as G
    pos = M.uniform_random(12 ** 3, seed=4)
    m = M.build_periodic(pos[M.morton_order(pos, (1, 1, 1))])
    ctx = G.Context()
    ctx.upload_faces(m.ncell, m.ncell, 0, m.faces, m.volume)
    off, cf = ctx.cell_faces()
    assert np.array_equal(off, m.cell_off) and np.array_equal(cf, m.cell_face)


@pytest.mark.gpu
@pytest.mark.parametrize("frac", [0.05, 0.3, 0.8])
def test_device_cell_lists_equal_substep_layout(frac):
    from fvmhd3d_amd import gpu as G
    m = random_mesh(1500)
    active = np.random.default_rng(int(frac * 100)).random(m.ncell) < frac
    perm, nact, nimp, faces, coff, cface, cvol, _ = substep_layout(m, active, np.zeros(m.ncell, np.int32))
    ctx = G.Context()
    ctx.upload_faces(len(perm), nact, nimp, faces, cvol)
    off, cf = ctx.cell_faces()
    assert np.array_equal(off, coff) and np.array_equal(cf, cface)
    # the same mesh through fvm_upload_mesh gives the same path results
    ref = G.Context()
    ref.upload_mesh(len(perm), nact, nimp, faces, coff, cface, cvol)
    assert ctx.get_active_faces().tolist() == ref.get_active_faces().tolist()


@pytest.mark.gpu
def test_upload_faces_rejects_inactive_creator():
    """A face whose s1 is not active has no creating cell: refused on the host."""
    from fvmhd3d_amd import gpu as G
    m = random_mesh(300)
    active = np.zeros(m.ncell, bool)
    active[:50] = True
    perm, nact, nimp, faces, coff, cface, cvol, _ = substep_layout(m, active, np.zeros(m.ncell, np.int32))
    k = int(np.nonzero(faces["s2"] >= nact)[0][0])
    bad = faces.copy()
    bad["s1"][k], bad["s2"][k] = faces["s2"][k], faces["s1"][k]
    ctx = G.Context()
    with pytest.raises(G.FvmError):
        ctx.upload_faces(len(perm), nact, nimp, bad, cvol)
