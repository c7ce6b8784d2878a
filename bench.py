"""Benchmark: Voronoi face-flux updates/s on one MI355X.

Workload (default): the 3D Orszag-Tang-type vortex on 256^3 = 16,777,216
generators on a jittered periodic lattice (Rand48, seed 123, +-0.1 dx),
rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2pi y, sin 2pi x, 0.1 sin 2pi z),
B = (-sin 2pi y, sin 4pi x, 0)/sqrt(4 pi), gamma = 5/3, Courant 0.4, smooth
limiter, static mesh.  One step = one pass of the path over every face: CFL +
predictor + reconstruction, limiter, flux + FP64 apply, on the FP32 face path
(FP64 flux sums, pressure recovery and CFL).  Outside the timed region the
same initial state is run for the configuration's 100 steps on the FP32 face
path and on the plain FP64 path, and the relative L1 difference of every
conserved component, the dt history and the diagnostic counts are compared.

Prints ONE JSON line.  Everything else goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "Voronoi face-flux updates/s (1 GPU) at ≤1e-4 rel L1 vs FP64; % HBM roofline"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, n, ic, limiter, steps in the configuration)
    "mhd_blast_128^3": ("random", 128, "blast", 2, 20),
    "orszag_tang_256^3": ("lattice", 256, "ot", 2, 100),
    "alfven_32^3": ("lattice", 32, "alfven", 0, 1),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from fvmhd3d_amd import mesh as M
    from fvmhd3d_amd.state import global_step_state
    gen, n, ic, lim, _ = WORKLOADS[name]
    t = time.perf_counter()
    m = M.workload_mesh(gen, n, seed=123)
    W = {"blast": M.ic_blast, "ot": M.ic_orszag_tang, "alfven": M.ic_alfven}[ic](m.pos)
    st = global_step_state(m, W, dt=0.0)
    log(f"[bench] mesh {name}: {m.ncell} cells, {m.nface} faces, built in {time.perf_counter() - t:.1f} s")
    return m, st, lim


def k2_bytes(ncell, nfe, prec_bytes):
    """Algorithmic HBM bytes of one flux+apply launch: every array it must touch, once.
    per cell: the gather record and Wt (FP64: W 10, Wt 10, gradients 30; FP32: the
    half-step state 10 and gradients 30, Wt 10 beside them), FP64 U, volume, CSR offset;
    writes FP64 U and the next step's W (10).  The FP64 path also reads and writes the
    stored FP64 primitives (the FP32 path re-derives them from U).
    per face-end: one 32-byte face-end record (neighbour, area vector, centroid offset)."""
    p = prec_bytes
    per_cell = 50 * p + 80 + 8 + 4 + 80 + 10 * p + (160 if p == 8 else 0)
    per_fe = 32 if p == 4 else 80
    return ncell * per_cell + nfe * per_fe


def k1_bytes(ncell, nfe, prec_bytes, limiter):
    """CFL + predictor + reconstruction launch: per cell read FP64 U, volume, offset and W;
    write Wt + gradients (40) and dt_new (the FP32 path also the half-step state W + Wt dti
    at the head of its record, 10 more); per face-end the same 32-byte record stream.
    With an opt-in limiter the FP32 path also limits here (a second pass over the 12-byte
    own-side offsets of each face-end); the FP64 path writes min/max (20) for its separate
    limiter launch."""
    p = prec_bytes
    per_cell = 80 + 8 + 4 + 10 * p + 40 * p + 8 + (10 * p if p == 4 else 0) + (20 * p if limiter and p == 8 else 0)
    per_fe = (32 + (12 if limiter else 0)) if p == 4 else 80
    return ncell * per_cell + nfe * per_fe


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(st, lim):
    """The oracle's global step on the host cores (OpenMP over every thread the
    process is given: OMP_NUM_THREADS, else the CPUs it may run on), one step of the
    same mesh and state: CFL, predictor, reconstruction, limiter, flux scatter and
    primitives.  The threaded passes are bit-identical to the serial oracle."""
    from oracle.oracle import Oracle
    o = Oracle("f64_omp", limiter=lim)
    s = st.copy()
    o.get_active_faces(s)
    t = time.perf_counter()
    o.step_global_par(s, 1.0, 1)
    dt = time.perf_counter() - t
    threads = o.threads()
    return {"value": len(s.face_act) / dt, "unit": "face-flux updates/s", "cores": threads, "kind": "port",
            "sample": f"1 global step of the same mesh and state ({st.np} cells, {len(s.face_act)} faces), "
                      f"oracle/fvm_oracle.c FP64 o_step_global_par on {threads} OpenMP threads "
                      f"(host: {len(os.sched_getaffinity(0))} usable CPUs, {_cpu_model()}), {dt:.1f} s"}


def substep_cpu_baseline(m, W0, dtmax, nsub, budget_s=15.0):
    """The oracle (CPU restatement, one core) behind the host-layout driver, on the
    same disk and substep cycle from its start, for about budget_s seconds."""
    from oracle.oracle import Oracle
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.substep import Substepper

    class OracleCalls:
        def __init__(self):
            self.o = Oracle("f64", limiter=G.LIMITER_SMOOTH)

        def load(self, st):
            pass

        def put(self, st, names):
            pass

        def fetch(self, st, names):
            pass

        def get_active_faces(self, st):
            return len(self.o.get_active_faces(st))

        def compute_primitives(self, st):
            self.o.compute_primitives(st, False)

        def compute_timestep(self, st):
            self.o.compute_timestep(st)

        def compute_timestep_limiter(self, st):
            self.o.compute_timestep_limiter(st)

        def compute_predictor(self, st, dtmax, min_rung):
            self.o.compute_predictor(st, dtmax, min_rung)

        def compute_reconstruction(self, st):
            self.o.compute_reconstruction(st)

        def slope_limiter(self, st):
            self.o.slope_limiter(st)

        def compute_fluid_update(self, st):
            self.o.compute_fluid_update(st)

    S = Substepper(m, W0, dtmax=dtmax, limiter=G.LIMITER_SMOOTH)
    B = OracleCalls()
    S.init_rungs(B)
    t0 = time.perf_counter()
    k = 0
    while k < nsub and time.perf_counter() - t0 < budget_s:
        S.substep(B)
        k += 1
        log(f"[substeps] oracle {k}/{nsub}: {S.history[-1][2]} active faces, {time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    faces = sum(h[2] for h in S.history)
    return {"value": faces / wall, "unit": "face-flux updates/s", "cores": 1, "kind": "port",
            "sample": f"the first {k} of the {nsub} substeps ({faces} active faces), host-layout driver over "
                      f"oracle/fvm_oracle.c FP64, {wall:.1f} s"}


# Algorithmic bytes of one k_slope_limiter<double> launch (slopeLimiter.cpp:27-237
# + computeFluidUpdateI, computeFluidUpdate.cpp:25-75, smooth limiter, FP64
# storage), each array it must touch counted once.  Per local cell: W, the two
# neighbour extremes and the gradients x, y, z and Wt read (6 x 80 B), the
# gradients and Wt written back (320 B) and the limiter written (80 B), tend,
# position, boundary flag, neighbour flag, volume and face-end offset (52 B):
# 1012 B.  Per active face: both cell indices, the face normal and centroid
# (56 B), its entry of the active-face list (4 B) and its two face-end entries
# (8 B): 68 B.
SLOPE_LIMITER_B_PER_CELL = 1012
SLOPE_LIMITER_B_PER_FACE = 68


def substep_roofline(m, W0, dtmax, nsub, history):
    """The same substeps again, from the same state, with HIP-event timing of the
    reference-order kernels (fvm_kernel_timing waits for each timed launch, so this
    run is not the timed one): the roofline of the substeps' dominant kernel,
    k_slope_limiter, as algorithmic bytes over its summed launch time."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.substep import DeviceSubstepper
    D = DeviceSubstepper(m, W0, dtmax=dtmax, limiter=G.LIMITER_SMOOTH)
    D.init_rungs()
    D.ctx.kernel_timing(True)
    for _ in range(nsub):
        D.substep()
    kms = D.ctx.kernel_ms()
    D.ctx.kernel_timing(False)
    assert D.history == history, "the timed repeat took other substeps"
    ms, launches = kms["k_slope_limiter"]
    nbytes = sum(SLOPE_LIMITER_B_PER_CELL * (c[0] + c[2]) + SLOPE_LIMITER_B_PER_FACE * c[1] for c in D.local_counts)
    D.ctx.close()
    achieved = nbytes / (ms * 1e-3) / 1e9
    total = sum(v[0] for v in kms.values())
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4), "traffic": None,
            "kernel": "k_slope_limiter<double> (slopeLimiter + computeFluidUpdateI on the active subset)",
            "bytes": int(nbytes), "ms": round(ms, 3), "launches": launches,
            "kernel_ms": {k: round(v[0], 3) for k, v in kms.items()}, "share_of_timed_kernels": round(ms / total, 3)}


def substep_bench(nsub, n=128, dtmax=1.0 / 64, host_driver=True, cpu=True):
    """Individual timesteps on the thin disk (substep.ic_thin_disk: rho ~ exp(-z^2/2H^2),
    Keplerian v_phi, c_s = H Omega, B_z at beta = 400, 1 % velocity noise) on n^3 =
    2,097,152 jittered lattice sites, 6 rungs.  dtmax = 1/64: the initial CFL steps
    then span the six rungs with none below the finest, 81 % of the cells sit on the
    coarsest rung, and each of the first 16 substeps updates 1.8-19 % of the faces.
    The measured driver keeps the per-cell state on the device (fvm_substep_run:
    layout, gathers, rungs and write-back on the GPU; the host keeps the rung clock
    and the sequential timestep-limiter sweep).  Reports active faces per second of
    wall time.  With host_driver, the host-side driver (per-substep layout and
    full-state transfers) runs the same substeps after it, for its time and as a
    check: its per-substep counts and final state must equal the device driver's
    bit for bit."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd import mesh as M
    from fvmhd3d_amd.substep import DeviceSubstepper, GpuBackend, Substepper, ic_thin_disk
    t = time.perf_counter()
    m = M.workload_mesh("lattice", n, seed=123)
    log(f"[substeps] mesh {m.ncell} cells, {m.nface} faces in {time.perf_counter() - t:.0f} s")
    W0 = ic_thin_disk(m.pos)
    D = DeviceSubstepper(m, W0, dtmax=dtmax, limiter=G.LIMITER_SMOOTH)
    D.init_rungs()
    D.ctx.reset_diagnostics()
    t0 = time.perf_counter()
    for k in range(nsub):
        D.substep()
        log(f"[substeps] device driver {k + 1}/{nsub}: {D.history[-1][1]} active cells, {D.history[-1][2]} active faces")
    wall = time.perf_counter() - t0
    faces = [h[2] for h in D.history]
    out = {"metric": "individual-timestep active-face updates/s (device-resident substeps)",
           "value": sum(faces) / wall, "unit": "face-flux updates/s", "substeps": nsub, "wall_s": round(wall, 3),
           "active_face_fraction": [round(f / m.nface, 4) for f in faces],
           "rungs": int(D.nrung), "dtmax": dtmax, "diagnostics": D.ctx.diagnostics(),
           "workload": f"thin disk, {m.ncell} jittered lattice generators"}
    # the same substeps with the FP32 face arithmetic (fvm_set_face_precision 32):
    # its rate, and its state against the FP64 run's
    D32 = DeviceSubstepper(m, W0, dtmax=dtmax, limiter=G.LIMITER_SMOOTH, face_bits=32)
    D32.init_rungs()
    D32.ctx.reset_diagnostics()
    t0 = time.perf_counter()
    for k in range(nsub):
        D32.substep()
    wall32 = time.perf_counter() - t0
    f32 = [h[2] for h in D32.history]
    l1 = lambda a, b: float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))
    U64, U32 = D.state("U").reshape(m.ncell, 10), D32.state("U").reshape(m.ncell, 10)
    W64, W32 = D.state("W").reshape(m.ncell, 10), D32.state("W").reshape(m.ncell, 10)
    scale = np.max(np.abs(W64), axis=0) + 1e-300
    out["face32"] = {"value": sum(f32) / wall32, "wall_s": round(wall32, 3),
                     "same_substeps_as_fp64": D32.history == D.history,
                     "l1_rho_vs_fp64": l1(W32[:, 0], W64[:, 0]), "l1_E_vs_fp64": l1(U32[:, 4], U64[:, 4]),
                     "max_rel_W_vs_fp64": float(np.max(np.abs(W32 - W64) / scale)),
                     "diagnostics": D32.ctx.diagnostics()}
    D32.ctx.close()
    log(f"[substeps] FP32 faces: {out['face32']}")
    out["roofline"] = substep_roofline(m, W0, dtmax, nsub, D.history)
    if host_driver:
        S = Substepper(m, W0, dtmax=dtmax, limiter=G.LIMITER_SMOOTH)
        be = GpuBackend(S.gamma, S.courant, tuple(float(b) for b in m.box), limiter=G.LIMITER_SMOOTH)
        S.init_rungs(be)
        be.kernel_s = 0.0
        t0 = time.perf_counter()
        for k in range(nsub):
            S.substep(be)
            log(f"[substeps] host driver {k + 1}/{nsub}: {S.history[-1][2]} active faces")
        hwall = time.perf_counter() - t0
        same = S.history == D.history and all(
            np.array_equal(np.asarray(getattr(S, f)).view(np.uint64), D.state(f).reshape(np.shape(getattr(S, f))).view(np.uint64))
            for f in ("W", "U", "dU", "tend", "dt_new"))
        out["host_driver"] = {"value": sum(faces) / hwall, "wall_s": round(hwall, 3),
                              "library_call_s": round(be.kernel_s, 3), "host_s": round(hwall - be.kernel_s, 3),
                              "bit_identical_to_device_driver": bool(same)}
    if cpu:
        out["cpu_baseline"] = substep_cpu_baseline(m, W0, dtmax, nsub)
    return out


def surrogate_bench(workload, nsteps=0):
    """The learned HLLD surrogate in the step against the plain FP64 path.

    From the same initial state: the plain FP64 path (the reference's
    arithmetic, FP64 HLLD), the FP32 face path (FP32 HLLD from both sides of
    every face) and the FP32 face path with the surrogate in its flux step
    (every face once, network where the gate accepts it, FP32 HLLD where it
    does not), each for the configuration's step count.  Reports face-flux
    updates/s of each (wall time of the whole run, synchronised), the relative
    L1 of density and total energy of the surrogate run against the FP64 run,
    the fraction of faces the gate sent to the HLLD, and the per-kernel times of
    the surrogate step.  Parameters: fvmhd3d_amd/data/hlld_surrogate.npz."""
    from fvmhd3d_amd import gpu as G
    from fvmhd3d_amd.surrogate import Surrogate
    m, st, lim = build_workload(workload)
    nsteps = nsteps or WORKLOADS[workload][4]
    sg = Surrogate.load()

    def run(precision, use_sg):
        c = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=lim)
        c.upload_state(st)
        c.fused_setup(precision)
        if use_sg:
            c.fused_set_surrogate(sg)
        c.reset_diagnostics()
        c.fused_timing(True)
        c.L.fvm_synchronize(c.h)
        t = time.perf_counter()
        c.fused_run(nsteps, 1.0)
        c.L.fvm_synchronize(c.h)
        wall = time.perf_counter() - t
        ms, nl = c.fused_kernel_ms()
        out = dict(W=c.download("w"), U=c.download("U"), dts=c.fused_dt_history(nsteps), diag=c.diagnostics(),
                   ms=wall * 1e3 / nsteps, kms=[round(ms[k] / max(nl[k], 1), 4) for k in range(4)])
        c.close()
        log(f"[surrogate] {'FP64' if precision == G.FUSED_FP64 else 'FP32'}{' + surrogate' if use_sg else ''}: "
            f"{out['ms']:.2f} ms/step, kernels {out['kms']}, diag {out['diag']}")
        return out

    r64 = run(G.FUSED_FP64, False)
    r32 = run(G.FUSED_FP32, False)
    rsg = run(G.FUSED_FP32, True)
    sg.close()
    l1 = lambda x, y: float(np.sum(np.abs(x - y)) / max(np.sum(np.abs(y)), 1e-300))
    fps = lambda r: m.nface / (r["ms"] * 1e-3)
    out = {"metric": "Voronoi face-flux updates/s (1 GPU): learned HLLD surrogate + HLLD fallback in the step "
                     "vs the plain FP64 path", "unit": "face-flux updates/s", "value": fps(rsg),
           "workload": workload, "steps": nsteps, "cells": m.ncell, "faces": m.nface,
           "ms_per_step": {"surrogate_fp32": rsg["ms"], "plain_fp64": r64["ms"], "fp32_hlld": r32["ms"]},
           "faces_per_s": {"surrogate_fp32": fps(rsg), "plain_fp64": fps(r64), "fp32_hlld": fps(r32)},
           "speedup_vs_plain_fp64": r64["ms"] / rsg["ms"],
           "l1_rel_vs_fp64": {"rho": l1(rsg["W"][:, 0], r64["W"][:, 0]), "E": l1(rsg["U"][:, 4], r64["U"][:, 4])},
           "l1_rel_fp32_hlld_vs_fp64": {"rho": l1(r32["W"][:, 0], r64["W"][:, 0]), "E": l1(r32["U"][:, 4], r64["U"][:, 4])},
           "fallback_fraction": rsg["diag"]["sg_fallback"] / float(m.nface * nsteps),
           "same_dt_every_step": bool(np.array_equal(rsg["dts"], r64["dts"])),
           "surrogate_step_kernel_ms": {"cfl_predict_recon_limit": rsg["kms"][0], "dt_check": rsg["kms"][1],
                                        "faces_plus_cells": rsg["kms"][2]},
           "diagnostics": {"surrogate": rsg["diag"], "fp64": r64["diag"]}}
    return out


def verify_fp32_vs_fp64(fresh, nsteps):
    """The FP32 face path and the plain FP64 path from the same initial state for
    nsteps global steps: relative L1 difference of every conserved component and of
    rho, whether both chose the same dt every step, and both paths' diagnostic
    counts (face states where the reference would assert, entropy and mass
    fallbacks).  Outside the timed region."""
    from fvmhd3d_amd import gpu as G
    res = {}
    for prec, name in ((G.FUSED_FP32, "fp32"), (G.FUSED_FP64, "fp64")):
        c = fresh(prec)
        c.reset_diagnostics()
        c.L.fvm_synchronize(c.h)
        t = time.perf_counter()
        c.fused_run(nsteps, 1.0)
        c.L.fvm_synchronize(c.h)
        wall = time.perf_counter() - t
        res[name] = dict(W=c.download("w"), U=c.download("U"), dts=c.fused_dt_history(nsteps),
                         diag=c.diagnostics(), ms=wall * 1e3 / nsteps)
        c.close()
        log(f"[bench] {name} path: {nsteps} steps, {res[name]['ms']:.2f} ms/step, diag {res[name]['diag']}")
    a, b = res["fp32"], res["fp64"]
    l1 = lambda x, y: float(np.sum(np.abs(x - y)) / max(np.sum(np.abs(y)), 1e-300))
    names = ["mass", "momx", "momy", "momz", "energy", "Bx", "By", "Bz", "psi_mass", "entropy_mass"]
    per = {n: l1(a["U"][:, k], b["U"][:, k]) for k, n in enumerate(names)}
    out = {"fp64_path_ms_per_step": round(b["ms"], 4),
           "l1_rel_vs_fp64": {"rho": l1(a["W"][:, 0], b["W"][:, 0]), "E": per["energy"],
                              "ethm": l1(a["W"][:, 4], b["W"][:, 4]), "steps": nsteps,
                              "conserved": per},
           "same_dt_every_step": bool(np.array_equal(np.asarray(a["dts"]), np.asarray(b["dts"]))),
           "diagnostics": {"fp32": a["diag"], "fp64": b["diag"], "equal": a["diag"] == b["diag"]}}
    log(f"[bench] FP32 vs FP64 after {nsteps} steps: {out['l1_rel_vs_fp64']}; same dt "
        f"{out['same_dt_every_step']}; diagnostics equal {out['diagnostics']['equal']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="orszag_tang_256^3", choices=sorted(WORKLOADS))
    ap.add_argument("--verify-steps", type=int, default=0,
                    help="steps of the FP32-vs-FP64 comparison (default: the configuration's step count)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--copy-only", action="store_true", help="only the stream-copy probe (counter calibration)")
    ap.add_argument("--surrogate", action="store_true",
                    help="the learned HLLD surrogate in the step against the plain FP64 path, the "
                         "configuration's step count (--verify-steps overrides it)")
    ap.add_argument("--substeps", type=int, default=0,
                    help="individual timesteps: this many substeps of the 2M-generator thin disk (6 rungs)")
    ap.add_argument("--no-host-driver", action="store_true",
                    help="with --substeps: skip the host-layout driver run (comparison and bit-identity check)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "sharding the path over several GPUs is a design, not an implementation", file=sys.stderr)
        sys.exit(2)

    from fvmhd3d_amd import gpu as G
    if args.surrogate:
        print(json.dumps(surrogate_bench(args.workload, args.verify_steps)), flush=True)
        return
    if args.substeps:
        print(json.dumps(substep_bench(args.substeps, host_driver=not args.no_host_driver,
                                       cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.copy_only:
        c = G.Context()
        gb = c.copy_bandwidth(1 << 30, 10)
        c.close()
        print(json.dumps({"stream_copy_GBps": gb, "bytes_per_launch": 2 * (1 << 30)}), flush=True)
        return
    probe = G.Context()
    copy_gbps = probe.copy_bandwidth(1 << 30, 10)
    memcpy_gbps = probe.runtime_copy_bandwidth(1 << 30, 10)
    probe.close()
    log(f"[bench] stream copy {copy_gbps:.0f} GB/s; the runtime's device-to-device copy {memcpy_gbps:.0f} GB/s")
    m, st, lim = build_workload(args.workload)
    nfe = int(m.cell_off[-1])

    def fresh(precision):
        ctx = G.Context(gamma=st.gamma, courant=st.courant, box=st.box, limiter=lim)
        ctx.upload_state(st)
        ctx.fused_setup(precision)
        return ctx

    ctx = fresh(G.FUSED_FP32)
    ctx.fused_run(args.warmup, 1.0)
    ctx.reset_diagnostics()
    ctx.fused_timing(True)
    ctx.L.fvm_synchronize(ctx.h)
    t0 = time.perf_counter()
    ctx.fused_run(args.steps, 1.0)
    ctx.L.fvm_synchronize(ctx.h)
    t1 = time.perf_counter()
    ms, nl = ctx.fused_kernel_ms()
    diag = ctx.diagnostics()
    W32 = ctx.download("w")
    U32 = ctx.download("U")
    dts = ctx.fused_dt_history(args.steps)
    ctx.close()
    elapsed = t1 - t0
    value = m.nface * args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed * 1e3:.1f} ms; kernel ms/launch: "
        f"predict+recon+limit {ms[0] / max(nl[0], 1):.3f} dt check {ms[1] / max(nl[1], 1):.3f} "
        f"flux {ms[2] / max(nl[2], 1):.3f} advance {ms[3] / max(nl[3], 1):.4f}; dt {dts[0]:.3e}; diag {diag}")

    k2_ms = ms[2] / max(nl[2], 1)
    k2_b = k2_bytes(m.ncell, nfe, 4)
    achieved = k2_b / (k2_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_flux_apply_{args.workload.replace('^3', '')}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_flux_apply_coop<float> (flux + FP64 apply)", "bytes_per_launch": k2_b,
                "ms_per_launch": round(k2_ms, 4)}
    k1_ms = ms[0] / max(nl[0], 1)
    k1_b = k1_bytes(m.ncell, nfe, 4, lim != 0)
    # with an opt-in limiter, the second launch of k_cfl_predict_recon checks the dt its
    # limiters took (the last step's) against the true one and redoes the step if they differ
    kc_ms = ms[1] / max(nl[1], 1)
    extra = {"kernels": {
        "k_cfl_predict_recon": {"ms": round(k1_ms, 4), "GB/s": round(k1_b / (k1_ms * 1e-3) / 1e9, 1)},
        "k_cfl_predict_recon_check": {"ms": round(kc_ms, 4), "redone": "first step of a run, and when the rung changes"},
        "k_flux_apply": {"ms": round(k2_ms, 4), "GB/s": round(achieved, 1)}},
        "step_GB_alg": round((k1_b + k2_b) / 1e9, 3), "stream_copy_GBps": round(copy_gbps, 1),
        "runtime_d2d_copy_GBps": round(memcpy_gbps, 1)}
    # the whole step against the same peak: every kernel's algorithmic bytes over the step time
    step_gbs = (k1_b + k2_b) / (elapsed / args.steps) / 1e9
    extra["pipeline_roofline"] = {"achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    if not args.no_verify:
        extra.update(verify_fp32_vs_fp64(fresh, args.verify_steps or WORKLOADS[args.workload][4]))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(st, lim)
        log(f"[bench] cpu baseline: {cpu}")

    gen, n, ic, _, cfg_steps = WORKLOADS[args.workload]
    out = {
        "metric": METRIC, "value": value, "unit": "face-flux updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 face arithmetic, fp64 sums/pressure/CFL",
        "data": "synthetic (Rand48-seeded generators, analytic initial state)",
        "config": {"workload": args.workload, "cells": m.ncell, "faces": m.nface,
                   "generators": gen, "limiter": ["shipped", "minmod", "smooth"][lim], "mesh": "static"},
        "roofline": roofline, "cpu_baseline": cpu, "diagnostics_timed": diag, **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
