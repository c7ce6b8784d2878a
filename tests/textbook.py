"""Independent textbook formulations used to pin the oracle (tests only).

The HLLD solver here is written from Miyoshi & Kusano, J. Comput. Phys. 208
(2005) 315-344, in its standard jump-condition form
    F*_a = F_a + S_a (U*_a - U_a),   F**_a = F*_a + S*_a (U**_a - U*_a),
sampled at the face speed w as F(U_K) - w U_K of the fan region K containing
w.  The reference (computeFlux.cpp:178-425) evaluates the same fan in an
a*U** + b*U* + c*U combination, so agreement pins its algebra.  Wave speeds
follow the reference's estimate S_L = min(u_L,u_R) - c, S_R = max(u_L,u_R) + c
with c^2 = max((gamma p + B^2)/rho) (computeFlux.cpp:218-229), which is a
parameter of the method, not of the algebra.
"""
import numpy as np


def hlld_moving(gamma, Bx, w, rl, pl, el, ul, vl, wl, byl, bzl, rr, pr, er, ur, vr, wr, byr, bzr):
    """Vectorised over faces. Returns the 8 face-frame fluxes
    (mass, momx, momy, momz, energy, Bx, By, Bz) of the interface moving at w."""
    sgn = np.sign(Bx)

    def state(r, p, e, u, v, ww, by, bz):
        B2 = Bx * Bx + by * by + bz * bz
        E = e + 0.5 * (r * (u * u + v * v + ww * ww) + B2)
        pT = p + 0.5 * B2
        U = np.stack([r, r * u, r * v, r * ww, E, by, bz])
        vB = u * Bx + v * by + ww * bz
        F = np.stack([r * u, r * u * u + pT - Bx * Bx, r * u * v - Bx * by, r * u * ww - Bx * bz,
                      (E + pT) * u - Bx * vB, by * u - Bx * v, bz * u - Bx * ww])
        return U, F, E, pT, B2

    UL, FL, EL, pTL, B2L = state(rl, pl, el, ul, vl, wl, byl, bzl)
    UR, FR, ER, pTR, B2R = state(rr, pr, er, ur, vr, wr, byr, bzr)
    cf = np.sqrt(np.maximum((gamma * pl + B2L) / rl, (gamma * pr + B2R) / rr))
    SL = np.minimum(ul, ur) - cf
    SR = np.maximum(ul, ur) + cf
    den = (SR - ur) * rr - (SL - ul) * rl
    SM = ((SR - ur) * rr * ur - (SL - ul) * rl * ul - pTR + pTL) / den
    pTs = ((SR - ur) * rr * pTL - (SL - ul) * rl * pTR + rl * rr * (SR - ur) * (SL - ul) * (ur - ul)) / den

    def star(r, u, v, ww, by, bz, E, pT, S):
        rs = r * (S - u) / (S - SM)
        d = r * (S - u) * (S - SM) - Bx * Bx
        vs = v - Bx * by * (SM - u) / d
        wsv = ww - Bx * bz * (SM - u) / d
        bys = by * (r * (S - u) ** 2 - Bx * Bx) / d
        bzs = bz * (r * (S - u) ** 2 - Bx * Bx) / d
        vB = u * Bx + v * by + ww * bz
        vBs = SM * Bx + vs * bys + wsv * bzs
        Es = ((S - u) * E - pT * u + pTs * SM + Bx * (vB - vBs)) / (S - SM)
        Us = np.stack([rs, rs * SM, rs * vs, rs * wsv, Es, bys, bzs])
        return Us, rs, vs, wsv, bys, bzs, Es, vBs

    UsL, rsL, vsL, wsL, bysL, bzsL, EsL, vBsL = star(rl, ul, vl, wl, byl, bzl, EL, pTL, SL)
    UsR, rsR, vsR, wsR, bysR, bzsR, EsR, vBsR = star(rr, ur, vr, wr, byr, bzr, ER, pTR, SR)
    sL, sR = np.sqrt(rsL), np.sqrt(rsR)
    SsL = SM - np.abs(Bx) / sL
    SsR = SM + np.abs(Bx) / sR
    vss = (sL * vsL + sR * vsR + (bysR - bysL) * sgn) / (sL + sR)
    wss = (sL * wsL + sR * wsR + (bzsR - bzsL) * sgn) / (sL + sR)
    byss = (sL * bysR + sR * bysL + sL * sR * (vsR - vsL) * sgn) / (sL + sR)
    bzss = (sL * bzsR + sR * bzsL + sL * sR * (wsR - wsL) * sgn) / (sL + sR)
    vBss = SM * Bx + vss * byss + wss * bzss
    EssL = EsL - sL * (vBsL - vBss) * sgn
    EssR = EsR + sR * (vBsR - vBss) * sgn
    UssL = np.stack([rsL, rsL * SM, rsL * vss, rsL * wss, EssL, byss, bzss])
    UssR = np.stack([rsR, rsR * SM, rsR * vss, rsR * wss, EssR, byss, bzss])
    FsL = FL + SL * (UsL - UL)
    FssL = FsL + SsL * (UssL - UsL)
    FsR = FR + SR * (UsR - UR)
    FssR = FsR + SsR * (UssR - UsR)

    out = np.where(w <= SL, FL - w * UL,
          np.where(w <= SsL, FsL - w * UsL,
          np.where(w <= SM, FssL - w * UssL,
          np.where(w <= SsR, FssR - w * UssR,
          np.where(w <= SR, FsR - w * UsR, FR - w * UR)))))
    # (mass, momx, momy, momz, E, Bx, By, Bz)
    return np.stack([out[0], out[1], out[2], out[3], out[4], -w * Bx, out[5], out[6]])


def face_frame(n):
    """An orthonormal frame (n, t1, t2) per face, built by Gram-Schmidt
    (independent of the reference's spherical-angle rotation)."""
    a = np.where(np.abs(n[:, [0]]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t1 = a - np.sum(a * n, axis=1)[:, None] * n
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 = np.cross(n, t1)
    return n, t1, t2


def compute_flux_textbook(gamma, wij, Wi, Wj, normal):
    """computeFlux for interior faces (jbnd = 0) from the textbook pieces:
    Gram-Schmidt face frame, Dedner-type face values, textbook HLLD, upwinded
    passive scalars.  Returns (psi, Bn, flux[n,10]) in the lab frame."""
    n, t1, t2 = face_frame(normal)

    def rot(W):
        v = W[:, 1:4]
        B = W[:, 5:8]
        return (np.sum(v * n, 1), np.sum(v * t1, 1), np.sum(v * t2, 1),
                np.sum(B * n, 1), np.sum(B * t1, 1), np.sum(B * t2, 1))

    ul, vl, wl, bxl, byl, bzl = rot(Wi)
    ur, vr, wr, bxr, byr, bzr = rot(Wj)
    rl, el, psil, sl = Wi[:, 0], Wi[:, 4], Wi[:, 8], Wi[:, 9]
    rr, er, psir, sr = Wj[:, 0], Wj[:, 4], Wj[:, 8], Wj[:, 9]
    pl = (gamma - 1.0) * el if gamma > 1.0 else el
    pr = (gamma - 1.0) * er if gamma > 1.0 else er
    chl = np.sqrt((gamma * pl + bxl ** 2 + byl ** 2 + bzl ** 2) / rl)
    chr_ = np.sqrt((gamma * pr + bxr ** 2 + byr ** 2 + bzr ** 2) / rr)
    Bn = (bxl * chl + bxr * chr_ + (psil - psir)) / (chl + chr_)
    psi = (psil * chr_ + psir * chl + chr_ * chl * (bxl - bxr)) / (chl + chr_)
    wn = np.sum(wij * n, 1)
    F = hlld_moving(gamma, Bn, wn, rl, pl, el, ul, vl, wl, byl, bzl, rr, pr, er, ur, vr, wr, byr, bzr)
    flux = np.zeros((len(rl), 10))
    flux[:, 0] = F[0]
    flux[:, 4] = F[4]
    flux[:, 1:4] = F[1][:, None] * n + F[2][:, None] * t1 + F[3][:, None] * t2
    flux[:, 5:8] = F[5][:, None] * n + F[6][:, None] * t1 + F[7][:, None] * t2
    flux[:, 8] = F[0] * np.where(F[0] > 0.0, psil, psir)
    flux[:, 9] = F[0] * np.where(F[0] > 0.0, sl, sr)
    return psi, Bn, flux
