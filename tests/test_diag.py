"""Energy totals (Energy.h:27-79) and the dump_binary snapshot (IO.cpp:7-99).

Pinned by closed forms on uniform states, by exact conservation of mass,
momentum and total energy over periodic device-resident steps (the flux form
conserves them; no source term touches them with zero gravity), and by reading
a snapshot back with the record layout of IO.cpp:29-95.
"""
import numpy as np
import pytest

from fvmhd3d_amd import gpu as G
from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.state import global_step_state

pytestmark = pytest.mark.gpu


def _mesh(n=10, seed=3):
    pos = M.uniform_random(n ** 3, seed=seed)
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    return M.build_periodic(pos)


def _uniform(m, rho=2.0, v=(0.1, -0.2, 0.3), ethm=0.5, B=(0.3, 0.4, 0.5)):
    W = np.zeros((m.ncell, 10))
    W[:, 0] = rho
    W[:, 1:4] = v
    W[:, 4] = ethm
    W[:, 5:8] = B
    W[:, 9] = M.entropy_from_ethm(W, 5.0 / 3.0)
    return W


def test_energy_closed_form_uniform_state():
    m = _mesh()
    rho, v, ethm, B = 2.0, np.array([0.1, -0.2, 0.3]), 0.5, np.array([0.3, 0.4, 0.5])
    st = global_step_state(m, _uniform(m), dt=0.0)
    ctx = G.Context()
    ctx.upload_state(st)
    e = ctx.energy()
    ctx.close()
    Vbox = 1.0
    assert abs(e["VOLUME"] - Vbox) < 1e-12
    assert abs(e["MASS"] - rho * Vbox) < 1e-12
    assert abs(e["ETHM"] - ethm * Vbox) < 1e-12
    for k, ax in enumerate("XYZ"):
        assert abs(e["EKIN" + ax] - 0.5 * rho * v[k] ** 2 * Vbox) < 1e-12
        assert abs(e["EMAG" + ax] - 0.5 * B[k] ** 2 * Vbox) < 1e-12
        assert abs(e["MOM" + ax] - rho * v[k] * Vbox) < 1e-12
    assert abs(e["ETOT"] - (ethm + 0.5 * rho * v @ v + 0.5 * B @ B) * Vbox) < 1e-12
    assert e["EPOT"] == 0.0
    # angular terms reduce to first moments of the volume: sum V x, sum V y, sum V z
    mx, my, mz = (np.sum(m.volume * m.pos[:, d]) for d in range(3))
    assert abs(e["ANMX"] - rho * (v[1] * mz - v[2] * my)) < 1e-12
    assert abs(e["ANMY"] - rho * (v[2] * mx - v[0] * mz)) < 1e-12
    assert abs(e["ANMZ"] - rho * (v[0] * my - v[1] * mz)) < 1e-12   # the reference's pairing


@pytest.mark.parametrize("precision", [G.FUSED_FP64, G.FUSED_FP32])
def test_energy_conserved_over_device_steps(precision):
    pos = M.jittered_lattice(12, seed=5)
    m = M.build_periodic(pos[M.morton_order(pos, (1, 1, 1))])
    W = M.ic_alfven(m.pos)
    st = global_step_state(m, W, dt=0.0)
    ctx = G.Context()
    ctx.upload_state(st)
    e0 = ctx.energy()
    ctx.fused_setup(precision)
    ctx.fused_run(8, 1.0)
    e1 = ctx.energy()
    e1b = ctx.energy()
    ctx.close()
    assert e1 == e1b                                   # deterministic reduction
    for k in ("MASS", "ETOT"):
        assert abs(e1[k] - e0[k]) <= 1e-13 * abs(e0[k]), (k, e0[k], e1[k])
    scale = np.sqrt(2 * e0["MASS"] * e0["ETOT"])
    for k in ("MOMX", "MOMY", "MOMZ"):
        assert abs(e1[k] - e0[k]) <= 1e-13 * scale, (k, e0[k], e1[k])
    assert e1["EKIN"] != e0["EKIN"]                    # the state did evolve


def test_dump_binary_layout(tmp_path):
    m = _mesh(8, seed=9)
    W = M.ic_alfven(m.pos)
    st = global_step_state(m, W, dt=0.0, t_global=0.25)
    rng = np.random.default_rng(2)
    st.vel = rng.normal(size=(m.ncell, 3))
    st.divB = rng.normal(size=m.ncell)
    ctx = G.Context()
    ctx.upload_state(st)
    idx = np.arange(m.ncell, dtype=np.int32) * 3 + 7
    dtc = rng.uniform(1e-3, 2e-3, m.ncell)
    p = tmp_path / "snap_part000004"
    ctx.dump_binary(str(p), dt_global=1.5e-3, iteration=17, chare_index=4, num_elements=8, global_n=999,
                    tsysU=(5 << 32) + 9, rmin=(0, 0, 0), rmax=(1, 1, 1), idx=idx, dt_cell=dtc)
    ctx.close()
    raw = np.fromfile(p, dtype=np.uint32)
    n = m.ncell
    assert raw.size == 22 + 28 * n
    hi, hf = raw[:22].view(np.int32), raw[:22].view(np.float32)
    assert hi[0] == 80 and hi[21] == 80
    assert list(hi[1:9]) == [4, 8, 8, 9, 5, 999, n, 3]
    assert hf[9] == np.float32(0.25) and hf[10] == np.float32(1.5e-3) and hi[11] == 17
    assert hf[12] == np.float32(0.4) and hf[13] == np.float32(5.0 / 3.0) and hi[14] == -1
    assert list(hf[15:21]) == [0, 0, 0, 1, 1, 1]
    rec = raw[22:].reshape(n, 28)
    ri, rf = rec.view(np.int32), rec.view(np.float32)
    assert np.all(ri[:, 0] == 104) and np.all(ri[:, 27] == 104)
    assert np.array_equal(ri[:, 1], idx)
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32)
    assert np.array_equal(rf[:, 2:5], f32(m.pos))
    assert np.array_equal(rf[:, 5:8], f32(st.vel))
    Wd = st.w
    for col, ref in ((8, Wd[:, 0]), (9, Wd[:, 4]), (13, Wd[:, 1]), (14, Wd[:, 2]), (15, Wd[:, 3]), (16, Wd[:, 5]),
                     (17, Wd[:, 6]), (18, Wd[:, 7]), (21, Wd[:, 8]), (23, Wd[:, 9])):
        assert np.allclose(rf[:, col], f32(ref), rtol=2e-7, atol=1e-12), col
    assert np.allclose(rf[:, 10], f32(Wd[:, 4] * (2.0 / 3.0)), rtol=2e-7)
    assert np.all(rf[:, 11] == -1.0) and np.all(ri[:, 12] == 0)
    assert np.array_equal(rf[:, 19], f32(dtc))
    assert np.array_equal(rf[:, 20], f32(m.volume))
    L = (m.volume * 3.0 / 4.0 / np.pi) ** (1.0 / 3.0)
    assert np.allclose(rf[:, 22], f32(L * st.divB), rtol=1e-6, atol=1e-12)
    assert np.all(rf[:, 24:27] == -1.0)
