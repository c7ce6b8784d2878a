"""The threaded global step of the oracle (the CPU baseline's code) against
the serial one: every cell adds its faces in the serial loop's order, so the
states must be bit-identical (CPU only)."""
import os

import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state
from oracle.oracle import Oracle
from tests.helpers import lattice_mesh, random_mesh


def _bits(st):
    return {k: np.ascontiguousarray(v).view(np.uint8).copy() for k, v in st.arrays.items()}


@pytest.mark.parametrize("case", ["blast_smooth", "ot_minmod", "alfven_shipped"])
def test_parallel_global_step_bit_identical(case):
    m = random_mesh(1500) if case.startswith("blast") else lattice_mesh(10)
    ic = {"blast": M.ic_blast, "ot": M.ic_orszag_tang, "alfven": M.ic_alfven}[case.split("_")[0]]
    lim = {"smooth": 2, "minmod": 1, "shipped": 0}[case.split("_")[1]]
    st = global_step_state(m, ic(m.pos), dt=0.0)
    a, b = st.copy(), st.copy()
    So, Po = Oracle("f64", limiter=lim), Oracle("f64_omp", limiter=lim)
    So.get_active_faces(a)
    Po.get_active_faces(b)
    assert Po.threads() >= 1
    da = So.step_global(a, 1.0, 3)
    db = Po.step_global_par(b, 1.0, 3)
    assert da == db
    ba, bb = _bits(a), _bits(b)
    for k in ba:
        assert np.array_equal(ba[k], bb[k]), k
    assert So.diag == Po.diag


def test_parallel_step_falls_back_with_incomplete_lists():
    m = random_mesh(1500)
    st = global_step_state(m, M.ic_blast(m.pos), dt=0.0)
    Oracle("f64").get_active_faces(st)
    # drop one face from one cell's list: the threaded passes would miss a side
    off, cf = m.cell_off.copy(), m.cell_face.copy()
    cf = np.delete(cf, off[3])
    off[4:] -= 1
    st.cell_off, st.cell_face = off, cf
    a, b = st.copy(), st.copy()
    So, Po = Oracle("f64", limiter=2), Oracle("f64_omp", limiter=2)
    So.step_global(a, 1.0, 1)
    Po.step_global_par(b, 1.0, 1)
    for k, v in _bits(a).items():
        assert np.array_equal(v, _bits(b)[k]), k
