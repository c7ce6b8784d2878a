"""Pins of three oracle functions that conservation and the uniform-state
checks leave open, each shown to catch planted mistakes (CPU only).

  computeTimestepLimiter (computeTimestep.cpp:64-110)
      a single ascending sweep over a chain has the closed form
          a_k = min(F_k, 2 d_{k+1}),   F_k = min_{j<=k} 2^(k-j) d_j
      (d = initial values: the propagation runs forward along the whole chain,
      backward by one hop only); a 4-cell case worked by hand with an inactive
      neighbour (initial value tend - tbeg) and a cell outside the local set
      (left alone); the identity when no face has a ratio above 2; the output
      never exceeds the input.
  Phi of the smooth limiter (MPI/compute_reconstruction.cpp:5-15)
      phi(1) = 3/4 and phi(2) = 1 exactly; phi(x) <= min(1, x) on [0, 2] and
      phi >= 1 beyond, so the limited face values stay within the neighbour
      bounds as with min(1, x).
  space-time extrapolation and half step of computeFluidUpdate
  (computeFluidUpdate.cpp:129-169)
      for W(x, s) = W0 + G x + Wt s with each cell's state taken at
      s = -(tend_k - t), both extrapolated face states of a face equal
      W0 + G c_f - Wt dt/2, dt = min(dtI, dtJ), whatever the two cells' tend;
      so dU_i = -sum_f dt A_f F(W*_f) with F the textbook flux of equal states
      (tests/textbook.py).  Unequal tends make the asymmetric dti/dtj matter.
"""
import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.gpu import FIELDS
from fvmhd3d_amd.state import CellState, global_step_state
from oracle.oracle import Oracle
from tests import textbook as TB
from tests.helpers import random_mesh
from tests.mutants import build_mutant, fails

GAMMA = 5.0 / 3.0


# ------------------------------------------------------------ graphs ------
def graph_state(n, nactive, edges, dt_new):
    """A state on an abstract cell graph: face f = edges[f] = (s1, s2); each
    active cell lists its faces in ascending face index (localMesh.cpp:586-589).
    Geometry is irrelevant to the timestep limiter and left at unit values."""
    faces = np.zeros(len(edges), dtype=M.FACE_DTYPE)
    for f, (a, b) in enumerate(edges):
        faces[f]["s1"], faces[f]["s2"] = a, b
        faces[f]["n"] = (1.0, 0.0, 0.0)
    lists = [[] for _ in range(nactive)]
    for f, (a, b) in enumerate(edges):
        for c in (a, b):
            if c < nactive:
                lists[c].append(f)
    off = np.zeros(nactive + 1, np.int32)
    off[1:] = np.cumsum([len(x) for x in lists])
    cf = np.array([f for x in lists for f in x], np.int32)
    st = CellState(np=n, nactive=nactive, nimport_loc=n - nactive, faces=faces, cell_off=off,
                   cell_face=cf, cvol=np.ones(nactive))
    for name, (_, width, kind) in FIELDS.items():
        st.arrays[name] = (np.zeros(n, np.int32) if kind == "i" else
                           np.zeros((n, width) if width > 1 else n))
    st.dt_new = np.array(dt_new, dtype=np.float64)
    st.is_active = (np.arange(n) < nactive).astype(np.int32)
    st.is_ngb = np.zeros(n, np.int32)
    return st


def chain_closed_form(d):
    n = len(d)
    F = np.empty(n)
    F[0] = d[0]
    for k in range(1, n):
        F[k] = min(d[k], 2.0 * F[k - 1])
    out = F.copy()
    out[:-1] = np.minimum(F[:-1], 2.0 * d[1:])
    return out


def check_chain(path):
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(3, 40))
        d = 10.0 ** rng.uniform(-4, 0, n)
        st = graph_state(n, n, [(k, k + 1) for k in range(n - 1)], d)
        Oracle("f64", path=path).compute_timestep_limiter(st)
        assert np.array_equal(st.dt_new, chain_closed_form(d)), trial


def check_hand_case(path):
    # cells 0,1,2 active; 3 an inactive neighbour (initial value tend - tbeg = 0.5;
    # its dt_new 999 must not be used); 4 outside the local set (left alone).
    # faces f0=(0,1) f1=(1,2) f2=(2,3) f3=(0,2); lists 0:[f0,f3] 1:[f0,f1] 2:[f1,f2,f3]
    # visit 0: f0 a1=min(6,20)=6 a0=min(10,12)=10; f3 a2=min(40,20)=20 a0=10
    # visit 1: f0 no change; f1 a2=min(20,12)=12 a1=min(6,24)=6
    # visit 2: f1 no change; f2 a3=min(.5,24)=.5 a2=min(12,1)=1; f3 a0=min(10,2)=2 a2=1
    st = graph_state(5, 3, [(0, 1), (1, 2), (2, 3), (0, 2)], [10.0, 6.0, 40.0, 999.0, 123.0])
    st.is_ngb[3] = 1
    st.tbeg[3], st.tend[3] = 0.25, 0.75
    Oracle("f64", path=path).compute_timestep_limiter(st)
    assert np.array_equal(st.dt_new, [2.0, 6.0, 1.0, 0.5, 123.0]), st.dt_new


def check_mesh_invariants(path):
    m = random_mesh(1500)
    rng = np.random.default_rng(12)
    st = global_step_state(m, M.ic_blast(m.pos), dt=1e-3)
    # identity: every face within ratio 2
    st.dt_new = 1e-3 * (1.0 + 0.999 * rng.random(m.ncell))
    d0 = st.dt_new.copy()
    Oracle("f64", path=path).compute_timestep_limiter(st)
    assert np.array_equal(st.dt_new, d0)
    # spread of four decades: never above the input, some cells lowered, and
    # every face of the first visited cell within ratio 2 of its neighbours
    st.dt_new = 10.0 ** rng.uniform(-6, -2, m.ncell)
    d0 = st.dt_new.copy()
    Oracle("f64", path=path).compute_timestep_limiter(st)
    assert np.all(st.dt_new <= d0) and np.any(st.dt_new < d0)
    for k in range(m.cell_off[0], m.cell_off[1]):
        f = m.faces[m.cell_face[k]]
        j = f["s2"] if f["s1"] == 0 else f["s1"]
        assert st.dt_new[0] <= 2.0 * st.dt_new[j]


TSL_CHECKS = (check_chain, check_hand_case, check_mesh_invariants)


def test_timestep_limiter_pins():
    for check in TSL_CHECKS:
        check(None)


TSL_MUTANTS = {
    "tsl_scale4": [("const real scale_fac = 2.0;", "const real scale_fac = 4.0;")],
    "tsl_no_backward": [("dti_list[i] = smin(dti_list[i], scale_fac * dti_list[j]);", ";")],
    "tsl_descending": [("for (int i = 0; i < nactive; i++)\n      for (int k",
                        "for (int i = nactive - 1; i >= 0; i--)\n      for (int k")],
    "tsl_ngb_from_dt_new": [("else dti_list[i] = s->tend[i] - s->tbeg[i];", "else dti_list[i] = s->dt_new[i];")],
    "tsl_writes_active_only": [("if (i < s->nactive || s->is_ngb[i]) s->dt_new[i]", "if (i < s->nactive) s->dt_new[i]")],
}


@pytest.mark.parametrize("tag", sorted(TSL_MUTANTS))
def test_timestep_limiter_pins_catch_mutant(tag):
    path = build_mutant(tag, "o_compute_timestep_limiter", TSL_MUTANTS[tag])
    assert any(fails(check, path) for check in TSL_CHECKS), tag


# ---------------------------------------------------- smooth limiter ------
def check_smooth_phi(path):
    O = Oracle("f64", path=path)
    assert O.phi(2, 1.0) == 0.75
    assert O.phi(2, 2.0) == 1.0
    x = np.linspace(0.0, 6.0, 6001)
    phi = np.array([O.phi(2, v) for v in x])
    lo = x <= 2.0
    assert np.all(phi[lo] <= np.minimum(1.0, x[lo]) + 1e-16)
    assert np.all(phi[~lo] >= 1.0)
    # the other modes
    assert O.phi(0, 0.3) == 1.0 and O.phi(1, 0.3) == 0.3 and O.phi(1, 3.0) == 1.0


def check_limited_values_in_bounds(path, lim=2):
    """Every face value after the slope limiter lies inside the cell's
    neighbour bounds (min(1, phi) <= min(1, x) makes this hold for the smooth
    mode as for min(1, x))."""
    m = random_mesh(1500)
    rng = np.random.default_rng(7)
    W = M.ic_blast(m.pos)
    W[:, 0] *= 1.0 + 0.5 * rng.random(m.ncell)
    st = global_step_state(m, W, dt=1e-4)
    O = Oracle("f64", limiter=lim, path=path)
    O.get_active_faces(st)
    O.compute_reconstruction(st)
    O.slope_limiter(st)
    assert O.diag["closure"] == 0
    assert np.all((st.lim >= 0) & (st.lim <= 1)) and np.any(st.lim < 1)
    f = st.faces[st.face_act]
    for side in (0, 1):
        c = f["s1"] if side == 0 else f["s2"]
        dri = f["centroid"] - st.pos[f["s1"]]
        nrm = f["n"] / np.linalg.norm(f["n"], axis=1)[:, None]
        dr = dri if side == 0 else dri - nrm * 2 * np.sum(dri * nrm, 1)[:, None]
        dW = st.x[c] * dr[:, [0]] + st.y[c] * dr[:, [1]] + st.z[c] * dr[:, [2]]
        val = st.w[c] + dW
        tol = 1e-9 * (1 + np.abs(st.w[c]))
        assert np.all(val <= np.maximum(st.mm_hi[c], st.w[c]) + tol)
        assert np.all(val >= np.minimum(st.mm_lo[c], st.w[c]) - tol)


def test_smooth_limiter_pins():
    check_smooth_phi(None)
    check_limited_values_in_bounds(None, 2)
    check_limited_values_in_bounds(None, 1)


PHI_MUTANTS = {
    # numerator without the factor 2: phi(1) = 1/2
    "phi_num": [("(x * x + 2.0 * x + fac * eps1) / (x * x + x + 2.0 + fac * eps1)",
                 "(x * x + x + fac * eps1) / (x * x + x + 2.0 + fac * eps1)")],
    # denominator without x: phi(x) > x below 1, face values overshoot the bounds
    "phi_den": [("(x * x + 2.0 * x + fac * eps1) / (x * x + x + 2.0 + fac * eps1)",
                 "(x * x + 2.0 * x + fac * eps1) / (x * x + 2.0 + fac * eps1)")],
}


@pytest.mark.parametrize("tag", sorted(PHI_MUTANTS))
def test_smooth_limiter_pins_catch_mutant(tag):
    path = build_mutant(tag, "Phi", PHI_MUTANTS[tag])
    assert fails(check_smooth_phi, path), tag
    if tag == "phi_den":
        assert fails(check_limited_values_in_bounds, path), tag


# ------------------------------------ extrapolation and half step ------
def extrapolation_case():
    m = random_mesh(1500)
    rng = np.random.default_rng(21)
    W0 = np.array([1.0, 0.1, -0.2, 0.05, 1.0, 0.5, 0.3, -0.2, 0.01, 1.0])
    G = 0.1 * rng.standard_normal((10, 3))
    Wt = rng.standard_normal(10) * np.array([0.5, 1, 1, 1, 0.5, 1, 1, 1, 0.2, 0.5])
    t0 = 0.125
    dT = 1e-2 * rng.uniform(0.5, 2.0, m.ncell)  # unequal steps: tend_k - t
    W = W0[None, :] + m.pos @ G.T - Wt[None, :] * dT[:, None]
    st = global_step_state(m, W, dt=1e-2, t_global=t0)
    st.tend = t0 + dT
    st.x = np.tile(G[:, 0], (m.ncell, 1))
    st.y = np.tile(G[:, 1], (m.ncell, 1))
    st.z = np.tile(G[:, 2], (m.ncell, 1))
    st.t = np.tile(Wt, (m.ncell, 1))
    # faces that do not cross the periodic boundary (the linear field is not periodic)
    f = m.faces
    d = m.pos[f["s2"]] - m.pos[f["s1"]]
    inner = np.all(np.abs(d) < 0.5, axis=1) & np.all(np.abs(f["centroid"] - m.pos[f["s1"]]) < 0.5, axis=1)
    st.face_act = np.nonzero(inner)[0].astype(np.int32)
    # expected: both sides at W0 + G c_f - Wt dt/2, dt = min(dtI, dtJ)
    fa = f[st.face_act]
    s1, s2 = fa["s1"], fa["s2"]
    dt = np.minimum(dT[s1], dT[s2])
    area = np.linalg.norm(fa["n"], axis=1)
    nrm = fa["n"] / area[:, None]
    Ws = W0[None, :] + fa["centroid"] @ G.T - Wt[None, :] * (0.5 * dt)[:, None]
    psi, Bn, flux = TB.compute_flux_textbook(GAMMA, np.zeros_like(nrm), Ws, Ws, nrm)
    ta = dt * area
    dU = np.zeros((m.ncell, 10))
    dgp = np.zeros((m.ncell, 3))
    ddivB = np.zeros(m.ncell)
    np.add.at(dU, s1, -flux * ta[:, None])
    np.add.at(dU, s2, flux * ta[:, None])
    np.add.at(dgp, s1, nrm * (psi * ta)[:, None])
    np.add.at(dgp, s2, -nrm * (psi * ta)[:, None])
    np.add.at(ddivB, s1, Bn * ta)
    np.add.at(ddivB, s2, -Bn * ta)
    return st, dict(dU=dU, dgp=dgp, ddivB=ddivB)


def check_extrapolation(path):
    st, ref = extrapolation_case()
    O = Oracle("f64", path=path)
    O.compute_fluid_update(st)
    assert O.diag["positivity"] == 0 and O.diag["dsl"] == 0
    for name in ("dU", "dgp"):
        got, exp = getattr(st, name), ref[name]
        scale = np.max(np.abs(exp), axis=0) + 1e-300
        assert np.max(np.abs(got - exp) / scale) < 1e-12, name
    assert np.max(np.abs(st.ddivB - ref["ddivB"])) < 1e-12 * np.max(np.abs(ref["ddivB"]))


def test_extrapolation_and_half_step_pin():
    check_extrapolation(None)


FU_MUTANTS = {
    "fu_swap_dti_dtj": [("const real dti = dtI - 0.5 * dt;", "const real dti = dtJ - 0.5 * dt;"),
                        ("const real dtj = dtJ - 0.5 * dt;", "const real dtj = dtI - 0.5 * dt;")],
    "fu_drop_t_dti": [("(real)s->t[ik] * dti;", "0.0 * dti;")],
    "fu_drj_sign": [("dri[0] - normal[0] * dsl", "dri[0] + normal[0] * dsl"),
                    ("dri[1] - normal[1] * dsl", "dri[1] + normal[1] * dsl"),
                    ("dri[2] - normal[2] * dsl", "dri[2] + normal[2] * dsl")],
    "fu_full_step": [("const real dti = dtI - 0.5 * dt;", "const real dti = dtI - 1.0 * dt;"),
                     ("const real dtj = dtJ - 0.5 * dt;", "const real dtj = dtJ - 1.0 * dt;")],
    "fu_dt_max": [("const real dt = smin(dtI, dtJ);", "const real dt = smax(dtI, dtJ);")],
}


@pytest.mark.parametrize("tag", sorted(FU_MUTANTS))
def test_extrapolation_pin_catches_mutant(tag):
    path = build_mutant(tag, "fluid_face", FU_MUTANTS[tag])
    assert fails(check_extrapolation, path), tag


# ------------------------------------------- substep visiting order ------
def test_limiter_depends_on_the_active_order():
    """The reference's active list is the Scheduler's (Scheduler.h
    pull_active_list: finest rung bucket first, each bucket in push order);
    this build's substep layouts list active cells in ascending mesh order.
    The sequential sweep (computeTimestep.cpp:64-110) is order dependent, so
    the two orders can give different dt_new: on the chain 0-1-2 with
    d = (100, 100, 1), ascending order reaches cell 0 with 100 (one backward
    hop), while the finest-first order (2, 1, 0) carries the small step to
    cell 0 as 4."""
    d = [100.0, 100.0, 1.0]
    st = graph_state(3, 3, [(0, 1), (1, 2)], d)
    Oracle("f64").compute_timestep_limiter(st)
    assert np.array_equal(st.dt_new, [100.0, 2.0, 1.0])
    # the same cells in the finest-first order: local 0 = cell 2, 1 = cell 1, 2 = cell 0
    st = graph_state(3, 3, [(0, 1), (1, 2)], d[::-1])
    Oracle("f64").compute_timestep_limiter(st)
    assert np.array_equal(st.dt_new[::-1], [4.0, 2.0, 1.0])
