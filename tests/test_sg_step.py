"""The learned HLLD surrogate inside the step of the FP32 face path
(fvm_fused_set_surrogate), through the C ABI on the GPU.

  gate closed     every face goes to the FP32 HLLD: the surrogate step (faces
                  once, flux records, per-cell FP64 sums in face order) must give
                  the FP32 face path's dt history and positivity count, and its
                  state to FP32 rounding (max relative difference 1e-5 per
                  component after 6 steps; a wrong face, side or sign shows at
                  order one), and count every face as a fallback.  The two
                  kernels share the source of the face states and the HLLD but
                  not their instruction selection: the compiler contracts
                  multiplies and adds into FMAs per kernel, so bits may differ.
  conservation    with the trained network answering faces: each face's flux
                  is shared by its two cells, so total mass, momentum and energy
                  change only by FP64 roundoff; two runs give the same bits.
  gain            the target configuration: the Orszag-Tang vortex on
                  256^3 = 16,777,216 jittered generators, 100 steps, the
                  surrogate step against the plain FP64 path (FP64 HLLD, the
                  reference's arithmetic, which matches the oracle to 1e-12)
                  from the same state.  It must be faster per step, and its
                  density and total energy must stay within relative L1 1e-4
                  of the FP64 run."""
import time

import numpy as np
import pytest

from fvmhd3d_amd import gpu as G
from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state
from tests.helpers import workload_mesh

pytestmark = pytest.mark.gpu


def _state(gen, n, ic):
    m = workload_mesh(gen, n, seed=123)
    return m, global_step_state(m, ic(m.pos), dt=0.0)


def _run(st, precision, sg, nsteps):
    c = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=G.LIMITER_SMOOTH)
    c.upload_state(st)
    c.fused_setup(precision)
    if sg is not None:
        c.fused_set_surrogate(sg)
    c.reset_diagnostics()
    c.L.fvm_synchronize(c.h)
    t = time.perf_counter()
    c.fused_run(nsteps, 1.0)
    c.L.fvm_synchronize(c.h)
    ms = (time.perf_counter() - t) * 1e3 / nsteps
    out = dict(W=c.download("w"), U=c.download("U"), dts=c.fused_dt_history(nsteps), diag=c.diagnostics(), ms=ms)
    c.close()
    return out


def test_gate_closed_is_the_fp32_path():
    from fvmhd3d_amd.surrogate import Surrogate
    m, st = _state("random", 16, M.ic_blast)
    nsteps = 6
    ref = _run(st.copy(), G.FUSED_FP32, None, nsteps)
    sg = Surrogate.load()
    sg.set_gate(0.5, 1.0, 3.0, -1e30)
    got = _run(st.copy(), G.FUSED_FP32, sg, nsteps)
    sg.close()
    print("diag", got["diag"], ref["diag"])
    assert np.array_equal(got["dts"], ref["dts"])
    for f in ("W", "U"):
        scale = np.max(np.abs(ref[f]), axis=0) + 1e-300
        err = float(np.max(np.abs(got[f] - ref[f]) / scale))
        print(f"{f}: max relative difference {err:.2e}")
        assert err <= 1e-5, (f, err)
    assert got["diag"]["sg_fallback"] == m.nface * nsteps
    assert got["diag"]["positivity"] == ref["diag"]["positivity"]


def test_trained_surrogate_conserves_and_repeats():
    from fvmhd3d_amd.surrogate import Surrogate
    m, st = _state("lattice", 24, M.ic_orszag_tang)
    sg = Surrogate.load()
    a = _run(st.copy(), G.FUSED_FP32, sg, 10)
    b = _run(st.copy(), G.FUSED_FP32, sg, 10)
    sg.close()
    frac = a["diag"]["sg_fallback"] / (10.0 * m.nface)
    print(f"fallback fraction {frac:.4f}")
    assert frac < 0.5
    assert np.array_equal(a["U"].view(np.uint64), b["U"].view(np.uint64))
    U0 = st.U
    for k, name in ((0, "mass"), (1, "momx"), (2, "momy"), (3, "momz"), (4, "energy")):
        scale = np.sum(np.abs(U0[:, k])) + np.sum(np.abs(a["U"][:, k]))
        drift = abs(np.sum(a["U"][:, k]) - np.sum(U0[:, k])) / scale
        print(f"{name}: relative change of the total {drift:.2e}")
        assert drift < 1e-12, (name, drift)


@pytest.mark.timeout(900)  # the 256^3 mesh, unless an earlier test built it
def test_surrogate_step_gain_and_accuracy():
    from fvmhd3d_amd.surrogate import Surrogate
    m, st = _state("lattice", 256, M.ic_orszag_tang)
    nsteps = 100
    sg = Surrogate.load()
    _run(st.copy(), G.FUSED_FP32, sg, 2)  # first launches (kernel attributes, code load)
    r64 = _run(st.copy(), G.FUSED_FP64, None, nsteps)
    rsg = _run(st.copy(), G.FUSED_FP32, sg, nsteps)
    sg.close()
    l1 = lambda x, y: float(np.sum(np.abs(x - y)) / np.sum(np.abs(y)))
    l1_rho = l1(rsg["W"][:, 0], r64["W"][:, 0])
    l1_E = l1(rsg["U"][:, 4], r64["U"][:, 4])
    frac = rsg["diag"]["sg_fallback"] / (nsteps * m.nface)
    print(f"{m.ncell} cells, {m.nface} faces, {nsteps} steps: surrogate {rsg['ms']:.3f} ms/step, "
          f"plain FP64 {r64['ms']:.3f} ms/step ({r64['ms'] / rsg['ms']:.2f}x); L1 rho {l1_rho:.2e}, "
          f"E {l1_E:.2e}; fallback fraction {frac:.4f}")
    assert l1_rho <= 1e-4 and l1_E <= 1e-4, (l1_rho, l1_E)
    assert rsg["ms"] < r64["ms"], (rsg["ms"], r64["ms"])
