"""Individual timesteps: the host rung clock and per-substep layout (CPU), and
the device path against the oracle over several substeps (GPU).

The same host driver (fvmhd3d_amd/substep.py: the rung clock of Scheduler.h and
the substep sequence of Iterate.cpp:92-202) runs once with the oracle and once
with the HIP library behind it; each substep's path calls see identical inputs,
so the states must agree to the FP64 bar (1e-12, relative to each component's
largest magnitude)."""
import numpy as np
import pytest

from fvmhd3d_amd import mesh as M
from fvmhd3d_amd.substep import RungClock, Substepper, ic_disk, substep_layout
from tests.helpers import random_mesh, workload_mesh
from tests.substep_oracle import OracleBackend


def test_rung_clock_matches_scheduler_rules():
    c = RungClock(1.0 / 16)
    rung = np.array([0, 1, 2, 2, 3])
    seen = []
    for _ in range(8):
        act, dt = c.pull(rung)
        seen.append((c.tsys(), tuple(np.nonzero(act)[0])))
    # the finest rung (3) is active every 1/128, rung r every (1/16)/2^r
    assert seen[0] == (1.0 / 128, (4,))
    assert seen[1] == (2.0 / 128, (2, 3, 4))
    assert seen[3] == (4.0 / 128, (1, 2, 3, 4))
    assert seen[7] == (8.0 / 128, (0, 1, 2, 3, 4))
    assert c.get_rung(1.0 / 16) == c.min_rung  # rungs below the synchronised ones are not allowed
    assert c.get_dt(3) == 1.0 / 128


def test_substep_layout_is_reference_layout():
    m = random_mesh(1500)
    rng = np.random.default_rng(3)
    active = rng.random(m.ncell) < 0.2
    perm, nact, nimp, faces, coff, cface, cvol, is_ngb = substep_layout(m, active, np.zeros(m.ncell, np.int32))
    assert np.all(active[perm[:nact]]) and not np.any(active[perm[nact:]])
    assert np.all(faces["s1"] < nact)  # every face is created by an active cell
    # every active cell lists exactly its mesh faces; centroid sits in s1's frame
    assert coff[-1] == sum(m.cell_off[i + 1] - m.cell_off[i] for i in perm[:nact])
    dri = faces["centroid"] - m.pos[perm[faces["s1"]]]
    assert np.max(np.abs(dri)) < 0.25
    assert np.all(np.sum(dri * faces["n"], 1) > 0)


class ReferenceScheduler:
    """Scheduler.h's rung lists, written out: push_particle appends a cell to
    its rung's list; pull_active_list takes every list whose rung ends its step
    now, finest rung first (bit i of tsys ^ tnext is rung 31 - i), each list in
    push order, and clears it."""

    def __init__(self, dtmax):
        self.lists = [[] for _ in range(32)]
        self.dt_tick = dtmax / (1 << 31)
        self.tsysU = 0

    def push(self, cell, rung):
        self.lists[rung].append(int(cell))

    def pull(self):
        max_rung = max(r for r in range(32) if self.lists[r])
        dtU = 1 << (31 - max_rung)
        tnext = (self.tsysU // dtU + 1) * dtU
        flags = self.tsysU ^ tnext
        out = []
        for i in range(32):
            if flags & (1 << i):
                out += self.lists[31 - i]
                self.lists[31 - i] = []
        self.tsysU = tnext
        return out


def test_active_list_follows_the_scheduler():
    """The host driver's active list, substep by substep, equals the pull of
    Scheduler.h's lists fed with the same pushes: every cell in index order at
    the start (System.cpp:300-315), then each substep's active cells in the
    order of its active list with their new rungs (Iterate.cpp:146-168)."""
    pos = M.uniform_random(10 ** 3, seed=9)
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    m = M.build_periodic(pos)
    S = Substepper(m, ic_disk(m.pos), dtmax=0.25, limiter=2)
    B = OracleBackend(limiter=2)
    S.init_rungs(B)
    ref = ReferenceScheduler(S.clock.dtmax)
    for i in range(m.ncell):
        ref.push(i, S.rung[i])
    orders_differ = False
    for _ in range(12):
        want = ref.pull()
        S.substep(B)
        got = [int(c) for c in S.last_active]
        assert got == want
        orders_differ = orders_differ or got != sorted(got)
        for c in got:
            ref.push(c, S.rung[c])
    assert len(set(int(r) for r in S.rung)) >= 3
    assert orders_differ  # the pull order is not the mesh order on this run


def test_substepper_oracle_runs_conservatively():
    pos = M.uniform_random(12 ** 3, seed=5)
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    m = M.build_periodic(pos)
    S = Substepper(m, ic_disk(m.pos), dtmax=0.25, limiter=2)
    B = OracleBackend(limiter=2)
    S.init_rungs(B)
    M0 = np.sum(S.U[:, 0])
    for _ in range(8):
        S.substep(B)
    assert B.o.diag["positivity"] == 0
    assert len(set(int(r) for r in S.rung)) >= 2
    # conserved mass: applied U plus pending dU
    assert abs(np.sum(S.U[:, 0] + S.dU[:, 0]) - M0) < 1e-12 * M0


def wt_bound(S):
    """Per-cell, per-component bound on |Wt_gpu - Wt_oracle| after a run.

    Wt is the predictor's finite difference (W1 - W)/dtp with the probe
    dtp = get_dt(get_rung(dt_new))/128 (computePredictor.cpp:271).  W1 passes
    through exp (psi damping) and pow (entropy) in computePrimitives, where the
    device's results may differ from glibc's by an ulp; that reaches Wt as
    ulp(|W1|)/dtp.  The probe's rung is the cell's stored rung unless the
    driver clamped it to the finest rung, in which case it is the unclamped
    exp(dtmax) - exp(dt_new).  |W1| is bounded by |W| + 2|Wt| (tend - tbeg).
    Bound: 1e-12 of the component's largest |Wt| (the FP64 bar) plus 4 ulp of
    |W1| over dtp."""
    from fvmhd3d_amd.substep import exp_of
    r0 = np.array([exp_of(S.clock.dtmax) - exp_of(d) if d > 0 else 0 for d in S.dt_new])
    dtp = S.clock.dts_of(np.maximum(S.rung, r0)) / 128.0
    w1 = np.abs(S.W) + 2.0 * np.abs(S.t) * (S.tend - S.tbeg)[:, None]
    return 1e-12 * (np.max(np.abs(S.t), axis=0) + 1e-300) + 4.0 * 2.0 ** -52 * w1 / dtp[:, None]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["disk_random_14", "thin_disk_lattice_16"])
def test_substeps_gpu_match_oracle(case):
    """The host driver with the oracle and with the library behind it, 10
    substeps.  The thin disk is the individual-timestep workload's initial
    state on 16^3 lattice sites, with dtmax scaled with the cell size (1/64 at
    128^3, 1/8 here)."""
    from fvmhd3d_amd.substep import GpuBackend, ic_thin_disk
    if case == "disk_random_14":
        pos = M.uniform_random(14 ** 3, seed=5)
        pos = pos[M.morton_order(pos, (1, 1, 1))]
        m = M.build_periodic(pos)
        W, dtmax = ic_disk(m.pos), 0.25
    else:
        m = workload_mesh("lattice", 16)
        W, dtmax = ic_thin_disk(m.pos), 1.0 / 8
    So = Substepper(m, W, dtmax=dtmax, limiter=2)
    Sg = Substepper(m, W, dtmax=dtmax, limiter=2)
    Bo = OracleBackend(limiter=2)
    Bg = GpuBackend(Sg.gamma, Sg.courant, tuple(m.box), limiter=2)
    So.init_rungs(Bo)
    Sg.init_rungs(Bg)
    assert np.array_equal(So.rung, Sg.rung)
    for k in range(10):
        a = So.substep(Bo)
        b = Sg.substep(Bg)
        assert a == b, (k, a, b)
        assert np.array_equal(So.rung, Sg.rung)
    for name in ("W", "U", "dU", "x"):
        ref = getattr(So, name)
        got = getattr(Sg, name)
        scale = np.max(np.abs(ref), axis=0) + 1e-300
        err = np.max(np.abs(got - ref) / scale)
        print(f"{name}: max rel err {err:.2e}")
        assert err <= 1e-12, (name, err)
    ratio = np.max(np.abs(Sg.t - So.t) / wt_bound(So))
    print(f"t: max err / derived bound {ratio:.3f}")
    assert ratio <= 1.0, ratio
    fr = [h[2] / m.nface for h in Sg.history]
    print("active face fractions", np.round(fr, 3))


def test_rung_arrays_match_scalar_scheduler():
    """RungClock.rungs_of / dts_of (the driver's array form) against get_rung /
    get_dt cell by cell (Scheduler.h:94-111), with the driver's clamp to
    [min_rung, nrung - 1]; dts are compared bit for bit."""
    c = RungClock(1.0 / 64)
    rng = np.random.default_rng(3)
    dt = np.exp(rng.uniform(np.log(1e-7), np.log(1.0 / 64), 20000))
    dt[:4] = [1.0 / 64, 1.0 / 128, np.nextafter(1.0 / 128, 0.0), 3e-7]
    for min_rung in (0, 2):
        c.min_rung = min_rung
        r_ref = np.array([max(min(c.get_rung(x), 5), min_rung) for x in dt])
        d_ref = np.array([c.get_dt(int(r)) for r in r_ref])
        r = c.rungs_of(dt, 6)
        assert np.array_equal(r, r_ref)
        assert np.array_equal(c.dts_of(r).view(np.uint64), d_ref.view(np.uint64))


@pytest.mark.gpu
def test_device_substeps_match_host_driver():
    """fvm_substep_run (layout, gathers, rungs and write-back on the device)
    against the host driver with the same library behind it: every step of the
    device substep is a copy, a compaction or the host's own FP64 expression,
    so the states, the rungs and the per-substep counts agree bit for bit."""
    from fvmhd3d_amd.substep import DeviceSubstepper, GpuBackend
    pos = M.uniform_random(14 ** 3, seed=5)
    pos = pos[M.morton_order(pos, (1, 1, 1))]
    m = M.build_periodic(pos)
    W = ic_disk(m.pos)
    Sh = Substepper(m, W, dtmax=0.25, limiter=2)
    Bh = GpuBackend(Sh.gamma, Sh.courant, tuple(m.box), limiter=2)
    Sd = DeviceSubstepper(m, W, dtmax=0.25, limiter=2)
    Sh.init_rungs(Bh)
    Sd.init_rungs()
    assert np.array_equal(Sh.rung, Sd.state("rung"))
    assert np.array_equal(Sh.tend, Sd.state("tend"))
    for k in range(12):
        a = Sh.substep(Bh)
        b = Sd.substep()
        print(f"substep {k}: host {a} device {b}", flush=True)
        assert a == b, (k, a, b)
        assert np.array_equal(Sh.rung, Sd.state("rung")), k
    assert Sh.history == Sd.history
    names = {"W": "W", "U": "U", "x": "x", "y": "y", "z": "z", "t": "t", "mm_lo": "mm_lo", "mm_hi": "mm_hi",
             "J": "J", "dU": "dU", "dgp": "dgp", "ddivB": "ddivB", "mvol": "mvol", "tbeg": "tbeg", "tend": "tend",
             "dt_new": "dt_new"}
    for host, dev in names.items():
        ref = np.asarray(getattr(Sh, host), dtype=np.float64)
        got = Sd.state(dev).reshape(ref.shape)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), host
    # some substeps were partial (the layout with inactive neighbours was exercised)
    assert any(0 < h[1] < m.ncell for h in Sd.history)


@pytest.mark.parametrize("spread,frac", [(0.05, 0.3), (0.35, 0.5), (1.0, 0.2), (2.0, 1.0), (1.5, 0.05)])
def test_sparse_timestep_limiter_sweep_matches_oracle(spread, frac):
    """The sweep that fvm_substep_run uses (visiting only the cells a change can
    reach; fvm_timestep_limiter_sweep runs it on host arrays) against the
    oracle's full sequential sweep (computeTimestep.cpp:64-110) on the
    reference layout of a substep: dt_new of every layout cell, bit for bit.
    Spreads from almost no hot faces to most faces hot, and partial and full
    active sets."""
    from fvmhd3d_amd import gpu as G
    m = random_mesh(1500)
    rng = np.random.default_rng(int(spread * 100) + int(frac * 10))
    active = rng.random(m.ncell) < frac
    active[0] = True
    S = Substepper(m, ic_disk(m.pos), dtmax=1.0 / 64)
    dt = 1e-3 * np.exp(spread * rng.standard_normal(m.ncell))
    S.dt_new[:] = dt
    S.tbeg[:] = rng.uniform(0.0, 1e-3, m.ncell)
    S.tend[:] = S.tbeg + 1e-3 * np.exp(spread * rng.standard_normal(m.ncell))
    st, perm = S._state(active)
    nact = st.nactive
    OracleBackend().compute_timestep_limiter(st)
    # global form: neighbour of every face-end, initial values as in the layout
    coff = m.cell_off
    i_of = np.repeat(np.arange(m.ncell), np.diff(coff))
    f = m.cell_face
    nbr = np.where(m.faces["s1"][f] == i_of, m.faces["s2"][f], m.faces["s1"][f])
    dti = np.where(active, S.dt_new, S.tend - S.tbeg)
    got = G.timestep_limiter_sweep(coff, nbr, active, dti)
    assert np.array_equal(st.dt_new.view(np.uint64), got[perm].view(np.uint64))
    changed = np.sum(got[perm] != dti[perm])
    print(f"spread {spread} active {nact}: {changed} of {len(perm)} layout cells changed")
    if spread >= 1.0:
        assert changed > 0


def test_sparse_sweep_rejects_bad_adjacency():
    from fvmhd3d_amd import gpu as G
    off = np.array([0, 1, 2], dtype=np.int32)
    with pytest.raises(G.FvmError):
        G.timestep_limiter_sweep(off, np.array([1, 2], dtype=np.int32), np.ones(2, np.uint8), np.ones(2))
    with pytest.raises(G.FvmError):
        G.timestep_limiter_sweep(np.array([0, 2, 1], dtype=np.int32), np.array([1, 0], dtype=np.int32),
                                 np.ones(2, np.uint8), np.ones(2))
    # two cells, both active: 1 -> 2 dt[0] caps dt[1]
    got = G.timestep_limiter_sweep(off, np.array([1, 0], dtype=np.int32), np.ones(2, np.uint8), np.array([1.0, 5.0]))
    assert got.tolist() == [1.0, 2.0]


@pytest.mark.gpu
def test_device_substep_rejects_inconsistent_calls():
    """A rung mask that activates cells whose step does not end now is refused
    (the host driver asserts the same), leaving the state usable; the
    shipped-precision storage is refused at setup."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.substep import DeviceSubstepper
    pos = M.uniform_random(8 ** 3, seed=2)
    m = M.build_periodic(pos[M.morton_order(pos, (1, 1, 1))])
    D = DeviceSubstepper(m, ic_disk(m.pos), dtmax=0.25)
    D.init_rungs()
    W0 = D.state("W")
    with pytest.raises(G.FvmError):
        D.ctx.substep_run(0xFFFFFFFF, 0.0, 0.25, 0, 6)
    assert np.array_equal(W0, D.state("W"))
    nact, nfa = D.substep()
    assert nact > 0 and nfa > 0
    ctx = G.Context(storage=G.STORAGE_SHIPPED)
    with pytest.raises(G.FvmError):
        ctx.substep_setup(m.faces, m.cell_off, m.cell_face, m.volume)


@pytest.mark.parametrize("frac", [1.0, 0.4])
def test_sparse_sweep_long_cascades_match_oracle(frac):
    """A few cells with a tiny dt in an otherwise uniform field: each change
    makes the next face exceed ratio 2, so the sweep's changes run in long
    chains through the mesh.  Compared with the oracle's full sweep bit for bit."""
    from fvmhd3d_amd import gpu as G
    m = random_mesh(1500)
    rng = np.random.default_rng(11)
    active = rng.random(m.ncell) < frac
    active[[0, 700, 1499]] = True
    S = Substepper(m, ic_disk(m.pos), dtmax=1.0 / 64)
    S.dt_new[:] = 1e-2 * (1.0 + 0.1 * rng.random(m.ncell))
    S.dt_new[[0, 700, 1499]] = [1e-7, 3e-8, 2e-6]
    S.tbeg[:] = 0.0
    S.tend[:] = 1e-2
    st, perm = S._state(active)
    OracleBackend().compute_timestep_limiter(st)
    coff = m.cell_off
    i_of = np.repeat(np.arange(m.ncell), np.diff(coff))
    f = m.cell_face
    nbr = np.where(m.faces["s1"][f] == i_of, m.faces["s2"][f], m.faces["s1"][f])
    dti = np.where(active, S.dt_new, S.tend - S.tbeg)
    got = G.timestep_limiter_sweep(coff, nbr, active, dti)
    assert np.array_equal(st.dt_new.view(np.uint64), got[perm].view(np.uint64))
    changed = np.sum(got[perm] != dti[perm])
    print(f"active {st.nactive}: {changed} of {len(perm)} layout cells changed")
    assert changed > 50


@pytest.mark.gpu
def test_device_substeps_face32_against_oracle():
    """The substeps with the FP32 face arithmetic (fvm_set_face_precision 32: face
    states and HLLD of the fluid update and the predictor in FP32; geometry, dt,
    dti, the per-cell sums, pressure recovery and the CFL step in FP64) against
    the FP64 oracle behind the host driver: the thin disk on 16^3 lattice sites,
    dtmax 1/8, 10 substeps.  The substeps (active cells and faces) are the
    oracle's; the state is within FP32 rounding of it (largest difference 1e-5
    of each column's largest magnitude, relative L1 of density and total energy
    1e-6); mass, applied plus pending, is conserved to FP64 roundoff.  Any face
    precision but 32 or 64 is refused."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.substep import DeviceSubstepper, ic_thin_disk
    m = workload_mesh("lattice", 16)
    W, dtmax = ic_thin_disk(m.pos), 1.0 / 8
    So = Substepper(m, W, dtmax=dtmax, limiter=2)
    Bo = OracleBackend(limiter=2)
    So.init_rungs(Bo)
    D = DeviceSubstepper(m, W, dtmax=dtmax, limiter=2, face_bits=32)
    with pytest.raises(G.FvmError):
        D.ctx.set_face_precision(16)
    D.init_rungs()
    M0 = float(np.sum(D.state("U").reshape(m.ncell, 10)[:, 0]))
    for k in range(10):
        a = So.substep(Bo)
        b = D.substep()
        assert a == b, (k, a, b)
    for name in ("W", "U"):
        ref = np.asarray(getattr(So, name))
        got = D.state(name).reshape(ref.shape)
        scale = np.max(np.abs(ref), axis=0) + 1e-300
        err = float(np.max(np.abs(got - ref) / scale))
        print(f"{name}: largest difference / column scale {err:.2e}")
        assert err <= 1e-5, (name, err)
    l1 = lambda a, b: float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))
    Wd, Ud = D.state("W").reshape(m.ncell, 10), D.state("U").reshape(m.ncell, 10)
    l1_rho, l1_E = l1(Wd[:, 0], So.W[:, 0]), l1(Ud[:, 4], So.U[:, 4])
    print(f"relative L1: rho {l1_rho:.2e}, E {l1_E:.2e}")
    assert l1_rho <= 1e-6 and l1_E <= 1e-6, (l1_rho, l1_E)
    mass = float(np.sum(Ud[:, 0] + D.state("dU").reshape(m.ncell, 10)[:, 0]))
    assert abs(mass - M0) <= 1e-12 * M0, (mass, M0)
    D.ctx.close()


@pytest.mark.gpu
def test_device_substeps_conserve_mass_minmod():
    """Device substeps with the minmod limiter over two rung cycles: every substep's
    count matches the host driver's, and mass (applied U plus the neighbours' pending
    dU) is conserved to FP64 roundoff."""
    from fvmhd3d_amd.substep import DeviceSubstepper, GpuBackend
    pos = M.uniform_random(12 ** 3, seed=9)
    m = M.build_periodic(pos[M.morton_order(pos, (1, 1, 1))])
    W = ic_disk(m.pos)
    D = DeviceSubstepper(m, W, dtmax=0.25, limiter=1)
    D.init_rungs()
    M0 = float(np.sum(D.state("U")[:, 0]))
    Sh = Substepper(m, W, dtmax=0.25, limiter=1)
    Bh = GpuBackend(Sh.gamma, Sh.courant, tuple(m.box), limiter=1)
    Sh.init_rungs(Bh)
    for k in range(24):
        assert D.substep() == Sh.substep(Bh), k
    U, dU = D.state("U"), D.state("dU")
    assert abs(np.sum(U[:, 0] + dU[:, 0]) - M0) < 1e-12 * M0
    assert np.array_equal(D.state("W").view(np.uint64), Sh.W.view(np.uint64))
    assert len({h[1] for h in D.history}) > 2  # several distinct active sets


def test_rung_clock_advance_matches_pull():
    """advance(max_rung) (the device driver's form: a bitmask of active rungs)
    against pull(rung) (the host driver's per-cell mask), over a sequence of
    substeps with changing rung sets."""
    rng = np.random.default_rng(5)
    a, b = RungClock(1.0 / 64), RungClock(1.0 / 64)
    rung = rng.integers(0, 5, 400)
    for _ in range(40):
        mask, dt_a = a.advance(int(rung.max()))
        act, dt_b = b.pull(rung)
        assert dt_a == dt_b and a.tsysU == b.tsysU and a.min_rung == b.min_rung
        assert np.array_equal(((mask >> rung) & 1).astype(bool), act)
        # active cells may move to any rung at or above min_rung
        rung[act] = rng.integers(a.min_rung, 6, int(act.sum()))


@pytest.mark.gpu
@pytest.mark.timeout(900)  # builds the 128^3 lattice mesh unless an earlier test did
def test_thin_disk_substeps_in_activity_band():
    """The individual-timestep workload: the thin disk (ic_thin_disk) on 128^3 =
    2,097,152 jittered lattice sites, 6 rungs, dtmax = 1/64.  Each of the first 16
    substeps updates 1-30 % of the faces; the device driver, run with the
    reference-order kernels timed (fvm_kernel_timing), equals the host driver bit
    for bit over them; and every slope-limiter launch was timed."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.substep import DeviceSubstepper, GpuBackend, ic_thin_disk
    m = workload_mesh("lattice", 128)
    W = ic_thin_disk(m.pos)
    D = DeviceSubstepper(m, W, dtmax=1.0 / 64, limiter=G.LIMITER_SMOOTH)
    D.init_rungs()
    D.ctx.kernel_timing(True)
    for _ in range(16):
        D.substep()
    frac = [h[2] / m.nface for h in D.history]
    print("active faces per substep:", " ".join(f"{100 * f:.2f}%" for f in frac))
    assert all(0.01 <= f <= 0.30 for f in frac), frac
    kms = D.ctx.kernel_ms()
    print("kernel ms:", {k: round(v[0], 3) for k, v in kms.items()})
    assert kms["k_slope_limiter"][1] == 16 and kms["k_slope_limiter"][0] > 0.0
    S = Substepper(m, W, dtmax=1.0 / 64, limiter=G.LIMITER_SMOOTH)
    B = GpuBackend(S.gamma, S.courant, tuple(float(b) for b in m.box), limiter=G.LIMITER_SMOOTH)
    S.init_rungs(B)
    for _ in range(16):
        S.substep(B)
    assert S.history == D.history
    for f in ("W", "U", "dU", "tend", "dt_new"):
        ref = np.asarray(getattr(S, f), dtype=np.float64)
        got = D.state(f).reshape(ref.shape)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), f
    D.ctx.close()
