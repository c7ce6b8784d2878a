"""CPU restatement of the learned HLLD surrogate's arithmetic, numpy FP64.

TEST INFRASTRUCTURE ONLY -- the checker.  Imported by tests/ (and by nothing
in the product package).  The surrogate has no counterpart in the reference;
what it stands in for is ``riemannSolver_HLLD`` (computeFlux.cpp:178-425)
inside ``computeFlux`` (computeFlux.cpp:6-148).  The face frame, the Dedner
normal field and the psi/entropy upwinding follow computeFlux.cpp exactly; the
MLP sees the HLLD inputs of one face after two exact symmetries of the HLLD
flux have been divided out:

* a Galilean boost by V = w_n n + u_t (the face speed along the normal and the
  mean tangential velocity of the two sides); the moving-frame flux G' and the
  face flux G are related by
      G_m = G'_m,   G_mom = G'_mom + V G'_m,
      G_E = G'_E + V.G'_mom + |V|^2/2 G'_m,   G_Bt = G'_Bt - B_n V_t,
      G_Bn = -w_n B_n;
* the MHD scaling rho -> a rho, v -> b v, p -> a b^2 p, B -> sqrt(a) b B,
  psi -> sqrt(a) b^2 psi (scales rho_s = (rho_L + rho_R)/2,
  c_s^2 = (gamma pbar + |B|^2bar)/rho_s).

Both are pinned in tests/test_surrogate.py against the oracle's HLLD.

Targets are residuals: with the normalised moving-frame states of both sides,
their HLL flux Fb (``hll_flux``) and the size J of the normalised jump, the
network learns (G' - Fb)/J (G' the normalised moving-frame HLLD flux): what
the HLLD's intermediate states add to the two-wave flux, per unit jump, so its
error in the flux scales with the jump.  The features are the mean state of
the two sides, the unit jump direction d/J, J, the log density and pressure
ratios and the larger fast speed: a smooth-flow jump is a small difference of
O(1) values that bf16 cannot hold, while its direction and size survive the
rounding.

Network: 16 -> 128 -> 128 -> 128 -> 8, ReLU, weights W[in][out]; every GEMM
operand is rounded to bfloat16 (round-to-nearest-even) and products are
accumulated in higher precision, as the device kernels do (FP32 there, FP64
here).  Outputs 0..6 are the normalised moving-frame fluxes (mass, mom n,
mom t1, mom t2, energy, B t1, B t2); output 7 is a confidence head trained
to predict log10 of the squared error of outputs 0..6.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 16, 128, 8
J_EPS = 1e-6
LAYERS = ((N_IN, N_H), (N_H, N_H), (N_H, N_H), (N_H, N_OUT))
CONF_EPS = 1e-10


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def frame(normal):
    """Rows of the face-frame rotation A (computeFlux.cpp:17-40); A @ v = (v_n, v_t1, v_t2)."""
    ex, ey, ez = normal[:, 0], normal[:, 1], normal[:, 2]
    ds = np.sqrt(ex * ex + ey * ey)
    with np.errstate(divide="ignore"):
        ids = np.where(ds != 0.0, 1.0 / np.where(ds != 0.0, ds, 1.0), 0.0)
    cosph = np.where(ids == 0.0, 1.0, ex * ids)
    sinph = np.where(ids == 0.0, 0.0, ey * ids)
    costh, sinth = ez, ds
    A = np.empty((len(normal), 3, 3))
    A[:, 0] = np.stack([cosph * sinth, sinth * sinph, costh], 1)
    A[:, 1] = np.stack([-sinph, cosph, np.zeros_like(ex)], 1)
    A[:, 2] = np.stack([-costh * cosph, -costh * sinph, sinth], 1)
    return A


def pressure(gamma, ethm):
    return (gamma - 1.0) * ethm if gamma > 1.0 else ethm


def featurize(gamma, wij, Wi, Wj, normal):
    """HLLD inputs of each face -> (x[n,16], ctx) with ctx holding what the
    inverse map needs.  Interior faces only (no boundary ghost rules)."""
    A = frame(normal)
    rot = lambda v: np.einsum("nij,nj->ni", A, v)
    vL, vR = rot(Wi[:, 1:4]), rot(Wj[:, 1:4])
    BL, BR = rot(Wi[:, 5:8]), rot(Wj[:, 5:8])
    rL, rR = Wi[:, 0], Wj[:, 0]
    pL, pR = pressure(gamma, Wi[:, 4]), pressure(gamma, Wj[:, 4])
    psiL, psiR = Wi[:, 8], Wj[:, 8]
    wn = np.einsum("ni,ni->n", A[:, 0], wij)
    # Dedner normal field and psi (computeFlux.cpp:101-107)
    chL = np.sqrt((gamma * pL + (BL * BL).sum(1)) / rL)
    chR = np.sqrt((gamma * pR + (BR * BR).sum(1)) / rR)
    Bn = (BL[:, 0] * chL + BR[:, 0] * chR + (psiL - psiR)) / (chL + chR)
    psi = (psiL * chR + psiR * chL + chR * chL * (BL[:, 0] - BR[:, 0])) / (chL + chR)
    ut = 0.5 * (vL[:, 1:] + vR[:, 1:])
    rs = 0.5 * (rL + rR)
    cs = np.sqrt((gamma * 0.5 * (pL + pR) + 0.5 * ((BL * BL).sum(1) + (BR * BR).sum(1))) / rs)
    Ps = rs * cs * cs
    Bs = np.sqrt(Ps)
    vnL, vnR = (vL[:, 0] - wn) / cs, (vR[:, 0] - wn) / cs
    # normalised moving-frame states: rho, v_n, v_t1, v_t2, p, e, B_n, B_t1, B_t2
    eL = Wi[:, 4] / Ps if gamma > 1.0 else pL / Ps
    eR = Wj[:, 4] / Ps if gamma > 1.0 else pR / Ps
    sideL = (rL / rs, vnL, (vL[:, 1] - ut[:, 0]) / cs, (vL[:, 2] - ut[:, 1]) / cs, pL / Ps, eL,
             Bn / Bs, BL[:, 1] / Bs, BL[:, 2] / Bs)
    sideR = (rR / rs, vnR, (vR[:, 1] - ut[:, 0]) / cs, (vR[:, 2] - ut[:, 1]) / cs, pR / Ps, eR,
             Bn / Bs, BR[:, 1] / Bs, BR[:, 2] / Bs)
    Fc = hll_flux(gamma, sideL, sideR)
    dl = np.stack([sideR[k] - sideL[k] for k in (0, 1, 2, 3, 4, 7, 8)], 1)
    J = np.sqrt((dl * dl).sum(1)) + J_EPS
    x = np.zeros((len(Wi), N_IN))
    x[:, 0] = 0.5 * (vnL + vnR)
    x[:, 1] = 0.5 * (pL + pR) / Ps
    x[:, 2] = Bn / Bs
    x[:, 3] = 0.5 * (BL[:, 1] + BR[:, 1]) / Bs
    x[:, 4] = 0.5 * (BL[:, 2] + BR[:, 2]) / Bs
    x[:, 5:12] = dl / J[:, None]
    x[:, 12] = J
    x[:, 13] = np.log(rR / rL)
    x[:, 14] = np.log(pR / pL)
    x[:, 15] = fast_speed_max(gamma, sideL, sideR)
    ctx = dict(A=A, wn=wn, ut=ut, rs=rs, cs=cs, Ps=Ps, Bs=Bs, Bn=Bn, psi=psi, psiL=psiL, psiR=psiR,
               entrL=Wi[:, 9], entrR=Wj[:, 9], Fc=Fc, J=J)
    return x, ctx


def fast_speed_max(gamma, sL, sR):
    """max of the two sides' fast speeds sqrt((gamma p + |B|^2)/rho), normalised units."""
    cf2 = lambda q: (gamma * q[4] + q[6] * q[6] + q[7] * q[7] + q[8] * q[8]) / q[0]
    return np.sqrt(np.maximum(cf2(sL), cf2(sR)))


def hll_flux(gamma, sL, sR):
    """HLL flux (Harten, Lax & van Leer 1983) of two normalised moving-frame
    states at the face, which is at rest in that frame, with the wave speeds
    riemannSolver_HLLD starts from (computeFlux.cpp:218-229):
    S_L = min(v_nL, v_nR) - c_f,max, S_R = max(v_nL, v_nR) + c_f,max.
    Components: mass, mom n, mom t1, mom t2, energy, B t1, B t2."""
    FL, FR = physical_flux(*sL), physical_flux(*sR)
    cfm = fast_speed_max(gamma, sL, sR)
    SL = np.minimum(sL[1], sR[1]) - cfm
    SR = np.maximum(sL[1], sR[1]) + cfm

    def cons(q):
        E = q[5] + 0.5 * q[0] * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) + 0.5 * (q[6] ** 2 + q[7] ** 2 + q[8] ** 2)
        return np.stack([q[0], q[0] * q[1], q[0] * q[2], q[0] * q[3], E, q[7], q[8]], 1)

    UL, UR = cons(sL), cons(sR)
    mid = (SR[:, None] * FL - SL[:, None] * FR + (SL * SR)[:, None] * (UR - UL)) / (SR - SL)[:, None]
    return np.where((SL >= 0.0)[:, None], FL, np.where((SR <= 0.0)[:, None], FR, mid))


def physical_flux(r, u, v1, v2, p, e, bn, b1, b2):
    """Ideal-MHD flux through a face at rest with normal n, normalised units
    (mass, mom n, mom t1, mom t2, energy, B t1, B t2); e is the thermal energy."""
    B2 = bn * bn + b1 * b1 + b2 * b2
    PT = p + 0.5 * B2
    E = e + 0.5 * r * (u * u + v1 * v1 + v2 * v2) + 0.5 * B2
    vB = u * bn + v1 * b1 + v2 * b2
    return np.stack([r * u, r * u * u + PT - bn * bn, r * u * v1 - bn * b1, r * u * v2 - bn * b2,
                     (E + PT) * u - bn * vB, u * b1 - bn * v1, u * b2 - bn * v2], 1)


def targets(flux_lab, ctx):
    """Face flux (lab axes, computeFlux output order) -> normalised moving-frame targets y[n,7]."""
    A = ctx["A"]
    rot = lambda v: np.einsum("nij,nj->ni", A, v)
    Gm = flux_lab[:, 0]
    Gmom = rot(flux_lab[:, 1:4])
    GE = flux_lab[:, 4]
    GB = rot(flux_lab[:, 5:8])
    V = np.concatenate([ctx["wn"][:, None], ctx["ut"]], 1)
    Gpm = Gm
    Gpmom = Gmom - V * Gpm[:, None]
    GpE = GE - (V * Gpmom).sum(1) - 0.5 * (V * V).sum(1) * Gpm
    GpBt = GB[:, 1:] + ctx["Bn"][:, None] * V[:, 1:]
    rs, cs, Ps, Bs = ctx["rs"], ctx["cs"], ctx["Ps"], ctx["Bs"]
    g = np.empty((len(Gm), 7))
    g[:, 0] = Gpm / (rs * cs)
    g[:, 1:4] = Gpmom / Ps[:, None]
    g[:, 4] = GpE / (Ps * cs)
    g[:, 5:7] = GpBt / (Bs * cs)[:, None]
    return (g - ctx["Fc"]) / ctx["J"][:, None]


def denormalize(y, ctx):
    """Network outputs 0..6 -> (psi, Bn, flux[n,10]) exactly as computeFlux assembles them."""
    rs, cs, Ps, Bs = ctx["rs"], ctx["cs"], ctx["Ps"], ctx["Bs"]
    V = np.concatenate([ctx["wn"][:, None], ctx["ut"]], 1)
    g = ctx["Fc"] + ctx["J"][:, None] * y
    Gpm = g[:, 0] * rs * cs
    Gpmom = g[:, 1:4] * Ps[:, None]
    GpE = g[:, 4] * Ps * cs
    GpBt = g[:, 5:7] * (Bs * cs)[:, None]
    Gmom = Gpmom + V * Gpm[:, None]
    GE = GpE + (V * Gpmom).sum(1) + 0.5 * (V * V).sum(1) * Gpm
    GB = np.empty((len(y), 3))
    GB[:, 0] = -ctx["wn"] * ctx["Bn"]
    GB[:, 1:] = GpBt - ctx["Bn"][:, None] * V[:, 1:]
    AT = np.transpose(ctx["A"], (0, 2, 1))
    back = lambda v: np.einsum("nij,nj->ni", AT, v)
    flux = np.zeros((len(y), 10))
    flux[:, 0] = Gpm
    flux[:, 1:4] = back(Gmom)
    flux[:, 4] = GE
    flux[:, 5:8] = back(GB)
    flux[:, 8] = Gpm * np.where(Gpm > 0.0, ctx["psiL"], ctx["psiR"])
    flux[:, 9] = Gpm * np.where(Gpm > 0.0, ctx["entrL"], ctx["entrR"])
    return ctx["psi"], ctx["Bn"], flux


# --------------------------------------------------------------------------
# network
# --------------------------------------------------------------------------
def init_params(seed=1):
    """He-uniform weights, zero biases (the device trainer takes any start)."""
    rng = np.random.default_rng(seed)
    P = {}
    for l, (i, o) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / i)
        P[f"W{l}"] = rng.uniform(-lim, lim, (i, o))
        P[f"b{l}"] = np.zeros(o)
    return P


def forward(P, x, round_bf16=True):
    """Returns (y, cache).  cache[l] = (input of layer l as used, pre-activation)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    h = x
    cache = []
    for l in range(len(LAYERS)):
        hin = r(h)
        z = hin @ r(P[f"W{l}"]) + P[f"b{l}"]
        cache.append((hin, z))
        h = np.maximum(z, 0.0) if l < len(LAYERS) - 1 else z
    return h, cache


def loss_and_dy(y, t, lam, w=None):
    """Mean over the batch of e2 + lam (y[7] - log10(e2 + eps))^2 with
    e2 = sum_k w_k (y[k] - t[k])^2 over outputs 0..6, the confidence target held
    fixed.  Returns (loss, dL/dy)."""
    B = len(y)
    w = np.ones(7) if w is None else np.asarray(w, np.float64)
    d = y[:, :7] - t
    e2 = (w * d * d).sum(1)
    c = np.log10(e2 + CONF_EPS)
    dc = y[:, 7] - c
    loss = float((e2 + lam * dc * dc).sum() / B)
    dy = np.zeros_like(y)
    dy[:, :7] = 2.0 * w * d / B
    dy[:, 7] = 2.0 * lam * dc / B
    return loss, dy


def backward(P, cache, dy, round_bf16=True):
    """Gradients of the loss wrt every W and b, given dL/dy."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    G = {}
    g = dy
    for l in reversed(range(len(LAYERS))):
        hin, z = cache[l]
        if l < len(LAYERS) - 1:
            g = g * (z > 0.0)
        G[f"W{l}"] = r(hin).T @ r(g)
        G[f"b{l}"] = g.sum(0)
        if l > 0:
            g = r(g) @ r(P[f"W{l}"]).T
    return G


def adam_step(P, G, M, V, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected), in place."""
    for k in P:
        M[k] = b1 * M[k] + (1.0 - b1) * G[k]
        V[k] = b2 * V[k] + (1.0 - b2) * G[k] * G[k]
        mh = M[k] / (1.0 - b1 ** step)
        vh = V[k] / (1.0 - b2 ** step)
        P[k] = P[k] - lr * mh / (np.sqrt(vh) + eps)
