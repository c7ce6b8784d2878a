"""Calls out of order and incomplete cell lists through the C ABI (GPU).

The fused stepping state belongs to the mesh it was set up on: any later mesh
upload (fvm_upload_mesh, fvm_upload_faces, a substep layout) invalidates it,
and fvm_fused_run / fvm_fused_sync then refuse with FVM_ESTATE instead of
running on buffers sized for another mesh.

get_active_faces (Iterate.cpp:298-324) takes a direct path when every face is
listed once under each of its active cells, and a list-position minimum
otherwise.  The completeness check is per (cell, face) pair, so a list that
has the right number of entries per face but names one face twice under the
same cell takes the general path; both must give the oracle's list."""
import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state
from oracle.oracle import Oracle
from tests.helpers import lattice_mesh, random_mesh


@pytest.mark.gpu
@pytest.mark.parametrize("upload", ["mesh", "faces"])
def test_fused_state_invalidated_by_mesh_upload(upload):
    from fvmhd3d_amd import gpu as G
    small = lattice_mesh(6)
    big = lattice_mesh(8)
    ctx = G.Context()
    ctx.upload_state(global_step_state(small, M.ic_alfven(small.pos), dt=0.0))
    ctx.fused_setup(G.FUSED_FP64)
    ctx.fused_run(1, 1.0)
    if upload == "mesh":
        ctx.upload_mesh(big.ncell, big.ncell, 0, big.faces, big.cell_off, big.cell_face, big.volume)
    else:
        ctx.upload_faces(big.ncell, big.ncell, 0, big.faces, big.volume)
    for call in (lambda: ctx.fused_run(1, 1.0), ctx.fused_sync):
        with pytest.raises(G.FvmError) as e:
            call()
        assert e.value.code == G.FVM_ESTATE
    # a new setup on the new mesh works
    ctx.upload_state(global_step_state(big, M.ic_alfven(big.pos), dt=0.0))
    ctx.fused_setup(G.FUSED_FP64)
    ctx.fused_run(1, 1.0)
    ctx.close()


def _drop_face(m, cell):
    """Cell lists with the first face of ``cell`` removed from its list."""
    off, cf = m.cell_off.copy(), m.cell_face.copy()
    k = off[cell]
    cf = np.delete(cf, k)
    off[cell + 1:] -= 1
    return off, cf


def _double_face(m):
    """Per-face counts kept, pairs broken: a face of cell a listed twice under
    its s1 and no longer under its s2."""
    f = 0
    a, b = int(m.faces[f]["s1"]), int(m.faces[f]["s2"])
    lists = [list(m.cell_face[m.cell_off[i]:m.cell_off[i + 1]]) for i in range(m.ncell)]
    lists[b].remove(f)
    lists[a].insert(lists[a].index(f), f)
    off = np.zeros(m.ncell + 1, np.int32)
    off[1:] = np.cumsum([len(x) for x in lists])
    return off, np.array([x for y in lists for x in y], np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["complete", "dropped", "doubled"])
def test_active_faces_with_incomplete_lists_match_oracle(case):
    from fvmhd3d_amd import gpu as G
    m = random_mesh(1500)
    if case == "complete":
        off, cf = m.cell_off, m.cell_face
    elif case == "dropped":
        off, cf = _drop_face(m, 17)
    else:
        off, cf = _double_face(m)
    st = global_step_state(m, M.ic_alfven(m.pos), dt=1e-3)
    st.cell_off, st.cell_face = off, cf
    ref = Oracle("f64").get_active_faces(st.copy())
    ctx = G.Context()
    ctx.upload_state(st)
    got = ctx.get_active_faces()
    ctx.close()
    assert np.array_equal(got, ref)
