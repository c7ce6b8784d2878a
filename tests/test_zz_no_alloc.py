"""fvm_compute_flux_noalloc: the batched computeFlux (computeFlux.cpp:6-148)
staged through the library's fixed per-GPU array instead of allocated device
memory.  Through the C ABI, against fvm_compute_flux on the same inputs: bit
for bit (integer-exact bar for the same kernel), on two calls in a row, the
second with fewer faces, so a stale tail of the shared array cannot pass
unnoticed; a batch above the stated maximum is refused on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _faces(n, seed):
    rng = np.random.default_rng(seed)
    Wi = np.zeros((n, 10))
    Wj = np.zeros((n, 10))
    for W in (Wi, Wj):
        W[:, 0] = 10.0 ** rng.uniform(-6, 1, n)
        W[:, 1:4] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-2, 1, (n, 1))
        W[:, 4] = 10.0 ** rng.uniform(-4, 2, n)
        W[:, 5:8] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, 1.5, (n, 1))
        W[:, 8] = 0.1 * rng.standard_normal(n)
        W[:, 9] = rng.uniform(0.1, 2.0, n)
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    wij = 0.3 * rng.standard_normal((n, 3))
    jbnd = rng.integers(0, 5, n).astype(np.int32)
    return wij, Wi, Wj, jbnd, nrm


def test_noalloc_flux_matches_allocating_entry_point():
    from fvmhd3d_amd import gpu as G
    ctx = G.Context()
    for n, seed in ((G.FLUX_NOALLOC_MAX, 11), (3000, 12)):
        faces = _faces(n, seed)
        print(f"noalloc flux: {n} faces, seed {seed}", flush=True)
        want = ctx.compute_flux(*faces)
        got = ctx.compute_flux(*faces, noalloc=True)
        for name, a, b in zip(("psi", "Bn", "flux"), got, want):
            assert a.shape == b.shape, name
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), name
    with pytest.raises(G.FvmError) as e:
        ctx.compute_flux(*_faces(G.FLUX_NOALLOC_MAX + 1, 13), noalloc=True)
    assert e.value.code == G.FVM_EARG
    assert "FVM_FLUX_NOALLOC_MAX" in str(e.value)
    # the context still works after the refusal
    faces = _faces(64, 14)
    want = ctx.compute_flux(*faces)
    got = ctx.compute_flux(*faces, noalloc=True)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    ctx.close()
