"""Global steps on a moving mesh: the mesh is rebuilt at the moved generators
every step and reaches the device as the reference's face list only
(fvm_upload_faces builds the cell lists and the SoA records there), then one
global step runs through the reference-order entry points.  The oracle runs
the same steps on the same meshes (o_step_global).  The FP64 path must stay
within 1e-12 of it (each component's largest magnitude) over every step.

Also, on the CPU: the generator velocity of computePvel (computePvel.cpp:6-54)
in its closed-form limits."""
import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.moving import MovingMesh, compute_pvel
from oracle.oracle import Oracle


class OracleStep:
    def __init__(self, limiter):
        self.o = Oracle("f64", limiter=limiter)

    def step(self, st, dtmax):
        self.o.get_active_faces(st)
        dt = self.o.step_global(st, dtmax, 1)[0]
        return dt, st.U.copy()


def _drifting_alfven(n):
    pos = M.jittered_lattice(n, seed=3)
    W = M.ic_alfven(pos)
    W[:, 1:4] += np.array([0.5, 0.3, 0.2])  # a uniform drift: every generator moves
    return pos, W


@pytest.mark.gpu
def test_moving_mesh_steps_match_oracle():
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.moving import GpuStep
    pos, W = _drifting_alfven(10)
    a, b = MovingMesh(pos, W), MovingMesh(pos, W)
    o, g = OracleStep(G.LIMITER_SHIPPED), GpuStep(limiter=G.LIMITER_SHIPPED)
    vol0 = a.mesh.volume.copy()
    for k in range(4):
        nface = a.mesh.nface
        dt_a = a.step(o, 1.0)
        dt_b = b.step(g, 1.0, move_with=a.last)
        assert dt_a == dt_b, (k, dt_a, dt_b)
        scale = np.max(np.abs(a.U), axis=0) + 1e-300
        err = float(np.max(np.abs(b.U - a.U) / scale))
        print(f"step {k}: dt {dt_a:.3e}, {nface} faces, max rel err of U {err:.2e}")
        assert err <= 1e-12, (k, err)
        assert np.array_equal(a.pos, b.pos)
    g.close()
    # the generators moved and the mesh changed under them
    assert np.max(np.abs(a.mesh.volume - vol0)) > 1e-3 * np.max(vol0)


def test_pvel_limits():
    """computePvel: v where the centroid sits on the generator or within 0.9
    eta R of it; v + (c - r) sqrt(c_f^2 + |v|^2) / d beyond 1.1 eta R (f = 1,
    tau = d / vabs); half of that at 1.0 eta R (the linear ramp)."""
    m = M.build_periodic(M.jittered_lattice(6, seed=1))
    n = m.ncell
    W = np.zeros((n, 10))
    W[:, 0] = 1.0
    W[:, 1:4] = np.array([0.3, -0.2, 0.1])
    W[:, 4] = 0.6
    W[:, 5:8] = np.array([0.2, 0.1, -0.3])
    gamma = 5.0 / 3.0
    vabs = np.sqrt((gamma * (gamma - 1.0) * 0.6 + 0.14) / 1.0 + 0.14)
    R = np.power(m.volume * 3.0 / (4.0 * np.pi), 1.0 / 3.0)
    u = np.array([0.6, 0.0, 0.8])
    for fac, f in ((0.0, 0.0), (0.5, 0.0), (0.89, 0.0), (1.0, 0.5), (1.2, 1.0), (2.0, 1.0)):
        mm = M.Mesh(m.faces, m.cell_off, m.cell_face, m.volume, m.pos + fac * 0.25 * R[:, None] * u, m.pos, m.box)
        v = compute_pvel(mm, W, gamma)
        want = W[:, 1:4] + f * vabs * u[None, :]
        assert np.allclose(v, want, rtol=1e-12, atol=1e-12), (fac, np.max(np.abs(v - want)))
