"""HIP reference-order path vs the CPU oracle, through the C ABI.

Bar: the all-FP64 path matches the all-FP64 oracle to <= 1e-12 relative (per
component, scaled by the component's largest magnitude); the as-shipped float
storage matches the as-shipped oracle to one float rounding (<= 1e-6).
"""
import numpy as np
import pytest

from fvmhd3d_amd import gpu as G
from oracle.oracle import Oracle
from tests.helpers import alfven_state, perturb_gradients, rel_err, max_ulp_f64

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "shipped": 1e-6}


def _cmp(name, a, b, tol, report):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, name
    assert np.array_equal(np.isfinite(a), np.isfinite(b)), f"{name}: non-finite mismatch"
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    for k in range(a.shape[1]):
        scale = max(np.max(np.abs(b[:, k])), 1e-300)
        e = np.max(np.abs(a[:, k] - b[:, k])) / scale
        report.append((name, k, e, max_ulp_f64(a[:, k], b[:, k])))
        assert e <= tol, f"{name}[{k}]: rel err {e:.3e} > {tol:.0e}"


def _ctx(st, storage, limiter=G.LIMITER_SHIPPED):
    ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=limiter,
                    storage=G.STORAGE_F64 if storage == "f64" else G.STORAGE_SHIPPED)
    ctx.upload_state(st)
    return ctx


FIELDS_OUT = {
    "reconstruction": ["x", "y", "z", "t", "J", "mm_lo", "mm_hi", "w"],
    "limiter": ["lim", "x", "y", "z", "t"],
    "fluid_update": ["dU", "dgp", "ddivB"],
    "primitives": ["w", "U", "dU", "t", "mvol", "divB"],
    "timestep": ["dt_new"],
    "predictor": ["t", "dt_new"],
}


def _run_pipeline(storage, limiter, st, with_predictor=True):
    report = []
    tol = TOL[storage]
    o = Oracle(storage, limiter=limiter)
    ref = st.copy()
    ctx = _ctx(st, storage, limiter)

    fa_o = o.get_active_faces(ref)
    fa_g = ctx.get_active_faces()
    assert np.array_equal(fa_o, fa_g), "get_active_faces order differs"

    dt_o = o.compute_timestep(ref)
    dt_g = ctx.compute_timestep()
    assert abs(dt_o - dt_g) <= 1e-12 * dt_o
    for n in FIELDS_OUT["timestep"]:
        _cmp("timestep." + n, ctx.download(n)[:st.nactive], getattr(ref, n)[:st.nactive], tol, report)

    if with_predictor:
        dtmax = 1.0 / 64
        o.compute_predictor(ref, dtmax, 0)
        ctx.compute_predictor(dtmax, 0)
        for n in FIELDS_OUT["predictor"]:
            _cmp("predictor." + n, ctx.download(n), getattr(ref, n), tol, report)

    for stage, fo, fg in (("reconstruction", o.compute_reconstruction, ctx.compute_reconstruction),
                          ("limiter", o.slope_limiter, ctx.slope_limiter),
                          ("fluid_update", o.compute_fluid_update, ctx.compute_fluid_update)):
        fo(ref)
        fg()
        for n in FIELDS_OUT[stage]:
            _cmp(f"{stage}.{n}", ctx.download(n), getattr(ref, n), tol, report)

    t1 = float(st.tend[0])
    ref.t_global = t1
    ctx.set_time(t1)
    o.compute_primitives(ref, False)
    ctx.compute_primitives(False)
    for n in FIELDS_OUT["primitives"]:
        _cmp("primitives." + n, ctx.download(n), getattr(ref, n), tol, report)
    dg = ctx.diagnostics()
    for k in ("positivity", "closure", "ethm_failed", "mass_failed", "dsl"):
        assert dg[k] == o.diag[k], f"diagnostic {k}: gpu {dg[k]} oracle {o.diag[k]}"
    ctx.close()
    return report


@pytest.mark.parametrize("storage", ["f64", "shipped"])
def test_pipeline_parity_alfven(storage):
    st = perturb_gradients(alfven_state(12, moving=True))
    report = _run_pipeline(storage, G.LIMITER_SHIPPED, st)
    worst = max(report, key=lambda r: r[2])
    print(f"[{storage}] worst rel err {worst[2]:.3e} at {worst[0]}[{worst[1]}]; "
          f"max ulp {max(r[3] for r in report)}")


@pytest.mark.parametrize("limiter", [G.LIMITER_MINMOD, G.LIMITER_SMOOTH])
def test_pipeline_parity_limiters(limiter):
    st = perturb_gradients(alfven_state(10, moving=True), scale=0.2)
    _run_pipeline("f64", limiter, st)


def test_pipeline_parity_boundaries():
    st = perturb_gradients(alfven_state(10, moving=True))
    rng = np.random.default_rng(11)
    # a few boundary cells of every kind (reflecting, outflow, diod, inflow)
    idx = rng.choice(st.np, size=40, replace=False)
    st.bnd[idx] = np.tile(np.array([1, 2, 3, 4], dtype=np.int32), 10)
    _run_pipeline("f64", G.LIMITER_SHIPPED, st)


def test_flux_batch_parity():
    rng = np.random.default_rng(3)
    n = 4096
    Wi = np.zeros((n, 10))
    Wj = np.zeros((n, 10))
    for W in (Wi, Wj):
        W[:, 0] = 10.0 ** rng.uniform(-6, 1, n)
        W[:, 1:4] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-2, 1, (n, 1))
        W[:, 4] = 10.0 ** rng.uniform(-4, 2, n)
        W[:, 5:8] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, 1.5, (n, 1))
        W[:, 8] = 0.1 * rng.standard_normal(n)
        W[:, 9] = rng.uniform(0.1, 2.0, n)
    Wj[:64] = Wi[:64]  # equal states
    Wi[64:128, 5] = Wj[64:128, 5] = 0.0  # zero normal field in some faces
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    nrm[128:160] = [0.0, 0.0, 1.0]  # ds == 0 branch of the rotation
    nrm[160:192] = [0.0, 0.0, -1.0]
    wij = 0.3 * rng.standard_normal((n, 3))
    jbnd = rng.integers(0, 5, n).astype(np.int32)
    ctx = G.Context()
    g = ctx.compute_flux(wij, Wi, Wj, jbnd, nrm)
    o = Oracle("f64").compute_flux(5.0 / 3.0, wij, Wi, Wj, jbnd, nrm)
    report = []
    for name, a, b in zip(("psi", "Bn", "flux"), g, o):
        _cmp("flux." + name, a, b, 1e-12, report)
    print("flux batch max ulp", max(r[3] for r in report))
    ctx.close()
