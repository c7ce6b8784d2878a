"""Device-resident stepping vs the oracle's global-timestep driver.

The FP64 fused path evaluates each face with the reference's FP64 operations
in the reference's summation order, so after several steps it must agree with
the all-FP64 oracle to <= 1e-12 relative (per component, scaled by the
component's largest magnitude).  The FP32 face path is measured against the
FP64 oracle by the relative L1 error of density and total energy, <= 1e-4.
"""
import numpy as np
import pytest

from fvmhd3d_amd import gpu as G
from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _case(kind, n):
    if kind == "alfven":
        pos = M.jittered_lattice(n)
        lim = G.LIMITER_SHIPPED
    elif kind == "ot":
        pos = M.jittered_lattice(n)
        lim = G.LIMITER_SMOOTH
    elif kind == "blast_minmod":
        pos = M.uniform_random(n ** 3, seed=123)
        lim = G.LIMITER_MINMOD
        kind = "blast"
    else:
        pos = M.uniform_random(n ** 3, seed=123)
        lim = G.LIMITER_SMOOTH
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    m = M.build_periodic(pos)
    W = {"alfven": M.ic_alfven, "ot": M.ic_orszag_tang, "blast": M.ic_blast}[kind](m.pos)
    return m, global_step_state(m, W, dt=0.0), lim


def _run_gpu(st, lim, precision, nsteps, dtmax):
    ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=lim)
    ctx.upload_state(st)
    ctx.fused_setup(precision)
    t = ctx.fused_run(nsteps, dtmax)
    W = ctx.download("w")
    U = ctx.download("U")
    dts = ctx.fused_dt_history(nsteps)
    dg = ctx.diagnostics()
    ctx.close()
    return W, U, dts, t, dg


@pytest.mark.parametrize("kind,n,nsteps", [("alfven", 12, 5), ("ot", 12, 5), ("blast", 14, 4), ("blast_minmod", 12, 4),
                                            ("blast", 32, 2)])
def test_fused_fp64_matches_oracle(kind, n, nsteps):
    m, st, lim = _case(kind, n)
    ref = st.copy()
    o = Oracle("f64", limiter=lim)
    o.get_active_faces(ref)
    dts_o = o.step_global(ref, 1.0, nsteps)
    W, U, dts, t, dg = _run_gpu(st, lim, G.FUSED_FP64, nsteps, 1.0)
    assert np.array_equal(np.asarray(dts_o), dts), (dts_o, dts)
    for name, a, b in (("W", W, ref.w), ("U", U, ref.U)):
        scale = np.max(np.abs(b), axis=0) + 1e-300
        err = np.max(np.abs(a - b) / scale)
        print(f"{kind} {name}: max rel err {err:.3e}")
        assert err <= 1e-12, (name, err)
    assert dg["positivity"] == o.diag["positivity"]


@pytest.mark.parametrize("kind,n,nsteps", [("alfven", 16, 10), ("blast", 16, 10), ("blast_minmod", 16, 10)])
def test_fused_fp32_error_vs_fp64(kind, n, nsteps):
    m, st, lim = _case(kind, n)
    W64, U64, dts64, _, _ = _run_gpu(st.copy(), lim, G.FUSED_FP64, nsteps, 1.0)
    W32, U32, dts32, _, dg = _run_gpu(st.copy(), lim, G.FUSED_FP32, nsteps, 1.0)
    assert np.array_equal(dts32, dts64)
    l1 = lambda a, b: np.sum(np.abs(a - b)) / np.sum(np.abs(b))
    e_rho = l1(W32[:, 0], W64[:, 0])
    e_E = l1(U32[:, 4], U64[:, 4])
    e_p = l1(W32[:, 4], W64[:, 4])
    print(f"{kind}: L1 rel err rho {e_rho:.2e} E {e_E:.2e} ethm {e_p:.2e}")
    assert e_rho <= 1e-4 and e_E <= 1e-4, (e_rho, e_E)
    assert dg["positivity"] == 0


def test_fused_conservation_and_determinism():
    m, st, lim = _case("ot", 14)
    W1, U1, _, _, _ = _run_gpu(st.copy(), lim, G.FUSED_FP32, 3, 1.0)
    W2, U2, _, _, _ = _run_gpu(st.copy(), lim, G.FUSED_FP32, 3, 1.0)
    assert np.array_equal(W1, W2) and np.array_equal(U1, U2)  # same bits every run
    tot0 = st.U.sum(0)
    tot1 = U1.sum(0)
    mag = np.abs(st.U).sum(0)
    # mass, momentum, energy, magnetic flux are conserved by the scatter; the
    # Powell/Dedner source terms act on B and psi only
    for k in (0, 1, 2, 3, 4):
        assert abs(tot1[k] - tot0[k]) <= 1e-12 * mag[k] + 1e-15, (k, tot1[k] - tot0[k])


@pytest.mark.parametrize("n", [12, 13])
def test_fused_fp32_addressing_variants_agree(monkeypatch, n):
    """The FP32 flux kernel addresses the neighbour records with 32-bit offsets
    up to 2^24 cells and with 64-bit addresses beyond; FVM_FUSED_ADDR64 (read
    at every launch) forces the latter.  Both must give the same bits: with
    12^3 = 1728 cells (27 whole wavefronts) and 13^3 = 2197 (the last
    wavefront partly past the last cell, whose block read of its own records
    runs past the array: zeros in the 32-bit form, the first record in the
    64-bit form, neither used)."""
    m, st, lim = _case("blast", n)
    monkeypatch.delenv("FVM_FUSED_ADDR64", raising=False)
    a = _run_gpu(st.copy(), lim, G.FUSED_FP32, 4, 1.0)
    monkeypatch.setenv("FVM_FUSED_ADDR64", "1")
    b = _run_gpu(st.copy(), lim, G.FUSED_FP32, 4, 1.0)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_fused_moving_generators():
    """Generator velocities set, geometry static: the face velocity with
    Springel's correction (computeFluidUpdate.cpp:194-207) and the CFL's
    relative speeds (computeTimestep.cpp:37-54) use them.  The FP64 fused path
    matches the oracle's global step; the FP32 face path stays within 1e-4 L1
    of the FP64 one."""
    from tests.helpers import alfven_state
    st = alfven_state(12, dt=0.0, moving=True)
    lim, nsteps = G.LIMITER_SHIPPED, 4
    ref = st.copy()
    o = Oracle("f64", limiter=lim)
    o.get_active_faces(ref)
    dts_o = o.step_global(ref, 1.0, nsteps)
    res = {}
    for prec in (G.FUSED_FP64, G.FUSED_FP32):
        ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=lim)
        ctx.upload_state(st.copy())
        ctx.fused_setup(prec, moving=True)
        ctx.fused_run(nsteps, 1.0)
        res[prec] = (ctx.download("w"), ctx.download("U"), ctx.fused_dt_history(nsteps))
        ctx.close()
    W64, U64, d64 = res[G.FUSED_FP64]
    assert np.array_equal(np.asarray(dts_o), d64)
    for name, a, b in (("W", W64, ref.w), ("U", U64, ref.U)):
        scale = np.max(np.abs(b), axis=0) + 1e-300
        err = np.max(np.abs(a - b) / scale)
        print(f"moving FP64 {name}: max rel err {err:.3e}")
        assert err <= 1e-12, (name, err)
    W32, U32, d32 = res[G.FUSED_FP32]
    l1 = lambda a, b: np.sum(np.abs(a - b)) / np.sum(np.abs(b))
    print(f"moving FP32: L1 rho {l1(W32[:, 0], W64[:, 0]):.2e} E {l1(U32[:, 4], U64[:, 4]):.2e}")
    assert l1(W32[:, 0], W64[:, 0]) <= 1e-4 and l1(U32[:, 4], U64[:, 4]) <= 1e-4
