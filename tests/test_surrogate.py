"""Learned HLLD surrogate.

CPU (oracle pins, oracle/surrogate_ref.py):
  * features/targets/denormalise round-trip to the oracle's computeFlux output;
  * the two symmetries the featurisation divides out (Galilean boost, MHD
    scaling) are exact symmetries of the oracle's HLLD: boosted and rescaled
    faces give identical features AND identical normalised targets;
  * backward pass vs central finite differences; Adam's first step; bf16 rounding.
GPU (through the C ABI, against the oracle restatement):
  * training GEMMs: forward outputs, loss and every gradient tensor;
  * the fused MFMA inference kernel, ragged batch sizes;
  * training samples (device featurisation + FP64 HLLD targets) on harvested
    face states of the reference-order pipeline;
  * fallback answers are the FP64 computeFlux bit for bit; training lowers the
    loss; the gate keeps the accepted faces within a stated accuracy; timing
    against the FP64/FP32 HLLD on the same faces is reported.
"""
import numpy as np
import pytest

from oracle import surrogate_ref as R
from oracle.oracle import Oracle

GAMMA = 5.0 / 3.0


def _random_faces(n, seed=3, jump=0.3):
    rng = np.random.default_rng(seed)
    normal = rng.normal(size=(n, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    normal[:8] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, 0, -1], [0, -1, 0], [0.6, 0.8, 0], [0, 0.6, 0.8]]
    Wi = np.zeros((n, 10))
    Wi[:, 0] = np.exp(rng.normal(0, 1, n))
    Wi[:, 1:4] = rng.normal(0, 1, (n, 3))
    Wi[:, 4] = np.exp(rng.normal(0, 1, n))
    Wi[:, 5:8] = rng.normal(0, 1, (n, 3))
    Wi[:, 8] = rng.normal(0, 0.1, n)
    Wi[:, 9] = np.exp(rng.normal(0, 1, n))
    Wj = Wi.copy()
    Wj[:, [0, 4, 9]] *= np.exp(rng.normal(0, jump, (n, 3)))
    Wj[:, [1, 2, 3, 5, 6, 7, 8]] += rng.normal(0, jump, (n, 7))
    wij = rng.normal(0, 0.5, (n, 3))
    return wij, Wi, Wj, np.zeros(n, np.int32), normal


def _scale(a):
    return np.max(np.abs(a), axis=0) + 1e-300


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3, 0.0])
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == 1.0 + 2 ** -6                 # tie -> even (up)
    assert r[3] == 1.0 and r[4] == -2.5 and r[6] == 0.0
    assert abs(r[5] - 3.0e-3) <= 2 ** -8 * 3.0e-3   # half an ulp of an 8-bit significand


def test_features_targets_roundtrip_oracle_flux():
    o = Oracle("f64")
    wij, Wi, Wj, jb, nrm = _random_faces(3000)
    psi, Bn, F = o.compute_flux(GAMMA, wij, Wi, Wj, jb, nrm)
    x, ctx = R.featurize(GAMMA, wij, Wi, Wj, nrm)
    y = R.targets(F, ctx)
    psi2, Bn2, F2 = R.denormalize(y, ctx)
    assert np.max(np.abs(psi2 - psi) / _scale(psi)) < 1e-13
    assert np.max(np.abs(Bn2 - Bn) / _scale(Bn)) < 1e-13
    err = np.max(np.abs(F2 - F) / _scale(F), axis=0)
    assert np.all(err < 1e-12), err


def test_central_flux_is_hlld_for_equal_states():
    """With W_L = W_R the HLLD flux is the physical flux, so the residual targets
    vanish and the normalised central flux equals the oracle's HLLD."""
    o = Oracle("f64")
    wij, Wi, _, jb, nrm = _random_faces(1000, seed=21)
    Wi[:, 8] = 0.0                          # psi = 0: B_n is the sides' normal field
    F = o.compute_flux(GAMMA, wij, Wi, Wi, jb, nrm)[2]
    x, ctx = R.featurize(GAMMA, wij, Wi, Wi, nrm)
    y = R.targets(F, ctx)
    assert np.all(ctx["J"] == R.J_EPS)
    assert np.max(np.abs(y * R.J_EPS)) < 1e-12 * np.max(np.abs(ctx["Fc"]))


def test_targets_depend_on_features_only():
    """Boost by an arbitrary velocity u (face velocity by its normal part) and
    rescale rho->a rho, v->b v, p->a b^2 p, B->sqrt(a) b B, psi->sqrt(a) b^2 psi
    (the Dedner psi carries B times a speed): features
    and normalised HLLD targets are unchanged.  A sign or index slip in the
    boost or the scales breaks this."""
    o = Oracle("f64")
    wij, Wi, Wj, jb, nrm = _random_faces(2000, seed=11)
    rng = np.random.default_rng(4)
    u = rng.normal(0, 2, (len(Wi), 3))
    a, b = 7.3, 0.37
    def transform(W):
        V = W.copy()
        V[:, 1:4] = b * (W[:, 1:4] + u)
        V[:, 0] = a * W[:, 0]
        V[:, 4] = a * b * b * W[:, 4]
        V[:, 5:8] = np.sqrt(a) * b * W[:, 5:8]
        V[:, 8] = np.sqrt(a) * b * b * W[:, 8]
        return V
    un = np.sum(u * nrm, axis=1, keepdims=True) * nrm
    wij2 = b * (wij + un)
    x1, c1 = R.featurize(GAMMA, wij, Wi, Wj, nrm)
    x2, c2 = R.featurize(GAMMA, wij2, transform(Wi), transform(Wj), nrm)
    y1 = R.targets(o.compute_flux(GAMMA, wij, Wi, Wj, jb, nrm)[2], c1)
    y2 = R.targets(o.compute_flux(GAMMA, wij2, transform(Wi), transform(Wj), jb, nrm)[2], c2)
    assert np.max(np.abs(x1 - x2)) < 1e-11
    assert np.max(np.abs(y1 - y2)) < 1e-10


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = R.init_params(3)
    for k in P:
        if k.startswith("b"):
            P[k] = rng.normal(0, 0.1, P[k].shape)
    x = rng.normal(0, 1, (16, R.N_IN))
    t = rng.normal(0, 1, (16, 7))
    lam = 0.3
    y, cache = R.forward(P, x, round_bf16=False)
    c = np.log10(((y[:, :7] - t) ** 2).sum(1) + R.CONF_EPS)   # the held-fixed confidence target

    def loss_fixed(Q):
        yq, _ = R.forward(Q, x, round_bf16=False)
        d = yq[:, :7] - t
        return float(((d * d).sum(1) + lam * (yq[:, 7] - c) ** 2).sum() / len(x))

    loss, dy = R.loss_and_dy(y, t, lam)
    assert abs(loss - loss_fixed(P)) < 1e-12 * abs(loss)
    G = R.backward(P, cache, dy, round_bf16=False)
    h = 1e-6
    for k in P:
        flat = P[k].ravel()
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            Qp = {kk: vv.copy() for kk, vv in P.items()}
            Qm = {kk: vv.copy() for kk, vv in P.items()}
            Qp[k].ravel()[idx] += h
            Qm[k].ravel()[idx] -= h
            fd = (loss_fixed(Qp) - loss_fixed(Qm)) / (2 * h)
            g = G[k].ravel()[idx]
            assert abs(fd - g) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, g)


def test_adam_first_step_moves_by_lr():
    P = {"w": np.array([1.0, -2.0, 0.5])}
    G = {"w": np.array([3.0, -1e-3, 0.0])}
    M = {"w": np.zeros(3)}
    V = {"w": np.zeros(3)}
    R.adam_step(P, G, M, V, 1, 0.01)
    assert np.allclose(P["w"], [1.0 - 0.01, -2.0 + 0.01 * 1e-3 / (1e-3 + 1e-8), 0.5], rtol=0, atol=1e-9)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _rel_fro(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (np.linalg.norm(b) + 1e-300))


def _harvest(n=14, steps=3):
    from fvmhd3d_amd import mesh as M
    from fvmhd3d_amd import surrogate as S
    pos = M.uniform_random(n ** 3, seed=9)
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    m = M.build_periodic(pos)
    W0 = M.ic_blast(m.pos)
    W, t = S.evolve(m, W0, 2, steps)
    return S.face_states(m, W, 2, t_global=t)


@pytest.mark.gpu
def test_gpu_samples_and_training_are_deterministic():
    """Samples are kept in face order and minibatches come from a counter hash,
    so two surrogates fed the same faces hold the same training set and train
    to the same parameters bit for bit."""
    from fvmhd3d_amd.surrogate import Surrogate, flatten
    faces = _harvest(n=12, steps=2)
    out = []
    for _ in range(2):
        sg = Surrogate(gate=(0.5, 1.0, 3.0, 1e30))
        sg.add_faces(*faces)
        x, t = sg.samples()
        sg.standardize()
        sg.train(40, batch=2048, lr=2e-3)
        out.append((x, t, flatten(sg.params())))
        sg.close()
    assert out[0][0].shape[0] > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_gpu_gradients_match_oracle():
    from fvmhd3d_amd.surrogate import Surrogate
    rng = np.random.default_rng(8)
    P = R.init_params(5)
    for k in P:
        if k.startswith("b"):
            P[k] = rng.normal(0, 0.1, P[k].shape)
    sg = Surrogate(params=P)
    P = {k: v.astype(np.float32).astype(np.float64) for k, v in sg.params().items()}
    n = 1000
    x = rng.normal(0, 1, (n, R.N_IN)).astype(np.float32)
    t = np.zeros((n, 8), np.float32)
    t[:, :7] = rng.normal(0, 1, (n, 7))
    y, G, loss = sg.gradients(x, t, lam=0.05)
    yo, cache = R.forward(P, x.astype(np.float64))
    lo, dy = R.loss_and_dy(yo, t[:, :7].astype(np.float64), 0.05)
    Go = R.backward(P, cache, dy)
    print("y", _rel_fro(y, yo), "loss", abs(loss - lo) / lo)
    assert _rel_fro(y, yo) < 2e-3
    assert abs(loss - lo) < 2e-3 * lo
    for k in Go:
        e = _rel_fro(G[k], Go[k])
        print(k, e)
        assert e < 1e-2, (k, e)
    sg.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 63, 65, 1000, 5000])
def test_gpu_inference_kernel_matches_oracle(n):
    from fvmhd3d_amd.surrogate import Surrogate
    rng = np.random.default_rng(n)
    P = R.init_params(7)
    for k in P:
        if k.startswith("b"):
            P[k] = rng.normal(0, 0.1, P[k].shape)
    sg = Surrogate(params=P)
    P = {k: v.astype(np.float64) for k, v in sg.params().items()}
    x = rng.normal(0, 1, (n, R.N_IN)).astype(np.float32)
    x[:, 15] = 0.0
    y = sg.forward(x)
    yo, _ = R.forward(P, x.astype(np.float64))
    e = np.max(np.abs(y - yo)) / (np.max(np.abs(yo)) + 1e-30)
    print("max rel err", e)
    assert e < 2e-2
    assert _rel_fro(y, yo) < 5e-3
    sg.close()


@pytest.mark.gpu
def test_gpu_samples_match_oracle_featurisation():
    from fvmhd3d_amd.surrogate import Surrogate
    wij, Wi, Wj, jb, nrm = _harvest()
    sg = Surrogate(gate=(50.0, 50.0, 1e6, 1e6))
    kept = sg.add_faces(wij, Wi, Wj, jb, nrm)
    assert kept == len(Wi)            # periodic, positive: every face is a sample
    x, t = sg.samples()
    o = Oracle("f64")
    xo, ctx = R.featurize(GAMMA, wij, Wi, Wj, nrm)
    yo = R.targets(o.compute_flux(GAMMA, wij, Wi, Wj, jb, nrm)[2], ctx)
    # the device appends in wave order: match rows by nearest features
    from scipy.spatial import cKDTree
    d, k = cKDTree(xo).query(x.astype(np.float64))
    assert len(np.unique(k)) == len(k)
    # features are formed in FP32 (they are rounded to bf16 for the network); targets in FP64
    ex = np.max(np.abs(x - xo[k]) / (np.abs(xo[k]) + 1e-2))
    et = np.max(np.abs(t[:, :7] - yo[k]) / (_scale(yo)))
    print("features", ex, "targets", et)
    assert ex < 1e-5 and et < 1e-6
    sg.close()


@pytest.mark.gpu
def test_gpu_surrogate_flux_matches_restatement():
    """Every face answered by the network: the device's featurisation, network and
    flux assembly against oracle featurize -> forward (bf16) -> denormalize."""
    from fvmhd3d_amd.surrogate import Surrogate
    wij, Wi, Wj, jb, nrm = _harvest()
    rng = np.random.default_rng(12)
    P = R.init_params(9)
    for k in P:
        P[k] = P[k] * 0.3 + (rng.normal(0, 0.05, P[k].shape) if k.startswith("b") else 0.0)
    sg = Surrogate(params=P, gate=(50.0, 50.0, 1e6, 1e30))
    P = {k: v.astype(np.float64) for k, v in sg.params().items()}
    psi, Bn, F, nfb = sg.compute_flux(wij, Wi, Wj, jb, nrm)
    sg.close()
    assert nfb == 0
    x, ctx = R.featurize(GAMMA, wij, Wi, Wj, nrm)
    y, _ = R.forward(P, x)
    psi_o, Bn_o, F_o = R.denormalize(y[:, :7], ctx)
    assert np.max(np.abs(psi - psi_o) / _scale(psi_o)) < 1e-13
    assert np.max(np.abs(Bn - Bn_o) / _scale(Bn_o)) < 1e-13
    err = np.abs(F[:, :8] - F_o[:, :8]) / _scale(F_o[:, :8])
    print("max", err.max(0), "median", np.median(err, 0))
    assert np.all(np.median(err, 0) < 1e-3) and np.all(err.max(0) < 3e-2)


@pytest.mark.gpu
def test_gpu_fallback_is_fp64_compute_flux():
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.surrogate import Surrogate
    wij, Wi, Wj, jb, nrm = _harvest()
    jb[::97] = G.REFLECTING              # boundary faces always fall back
    ctx = G.Context()
    psi, Bn, F = ctx.compute_flux(wij, Wi, Wj, jb, nrm)
    ctx.close()
    sg = Surrogate(gate=(0.5, 1.0, 3.0, -1e30))    # the confidence gate rejects everything
    psi2, Bn2, F2, nfb = sg.compute_flux(wij, Wi, Wj, jb, nrm)
    assert nfb == len(Wi)
    assert np.array_equal(psi, psi2) and np.array_equal(Bn, Bn2) and np.array_equal(F, F2)
    sg.set_gate(50.0, 50.0, 1e6, 1e30)            # the network answers every interior face
    psi3, Bn3, F3, nfb = sg.compute_flux(wij, Wi, Wj, jb, nrm)
    assert nfb == int(np.sum(jb != 0))
    inter = jb == 0
    assert np.array_equal(psi3[inter], psi[inter]) and np.array_equal(Bn3[inter], Bn[inter])
    assert np.array_equal(F3[~inter], F[~inter])
    sg.close()


@pytest.mark.gpu
def test_gpu_training_lowers_loss_and_gates_accuracy():
    """Train on harvested blast faces, evaluate on a held-out harvest: the loss
    falls by 10x and the faces the gate accepts are within 2e-2 of the FP64 HLLD
    (L1 over the accepted faces' mass, momentum, energy and induction fluxes,
    relative to the FP64 L1); the rest are answered by the FP64 HLLD.  The
    timing against the FP64 and FP32 HLLD on the same faces is printed only: a
    batch of faces staged through HBM is bound by its bytes for all three.  What
    the network gains over the plain FP64 method is asserted where it is used,
    in the step (tests/test_sg_step.py::test_surrogate_step_gain_and_accuracy:
    faster per step, density and energy within 1e-4 after 100 steps)."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.surrogate import Surrogate
    train = _harvest(n=20, steps=4)
    test = _harvest(n=24, steps=6)
    sg = Surrogate(gate=(0.5, 1.0, 3.0, 1e30))
    sg.add_faces(*train)
    sg.standardize()
    hist = sg.train(1500, batch=4096, lr=2e-3)
    print("loss", hist[:3], hist[-3:])
    assert np.mean(hist[-50:]) < 0.1 * np.mean(hist[:10])
    sg.set_gate(0.5, 1.0, 3.0, -3.0)
    ctx = G.Context()
    psi, Bn, F = ctx.compute_flux(*test)
    ctx.close()
    psi2, Bn2, F2, nfb = sg.compute_flux(*test)
    ans = np.any(F2 != F, axis=1)            # faces the network answered; the rest are FP64 bit for bit
    l1 = np.sum(np.abs(F2[ans, :8] - F[ans, :8])) / np.sum(np.abs(F[ans, :8]))
    print(f"fallback {nfb}/{len(F)}, answered {int(ans.sum())}, accepted-face L1 rel {l1:.3e}")
    assert int(ans.sum()) + nfb >= len(F) - 1 and l1 < 2e-2
    big = tuple(np.concatenate([a] * max(1, 400000 // len(a))) for a in test)
    ms, _ = sg.time_flux(*big, iters=5)
    print(f"{len(big[1])} faces: surrogate+fallback {ms[0]:.3f} ms, FP64 HLLD {ms[1]:.3f} ms, FP32 HLLD {ms[2]:.3f} ms")
    sg.close()
