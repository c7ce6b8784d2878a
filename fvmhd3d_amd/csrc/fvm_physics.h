// Device-side physics of the Voronoi-face pipeline, templated on the
// arithmetic type T (double for the FP64 path, float for the FP32 face path).
//
// Every expression keeps the reference's evaluation order so that, built with
// -ffp-contract=off, the FP64 instantiation reproduces the reference's double
// arithmetic operation for operation.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace fvm {

enum { NF = 10 };
enum { DENS = 0, VELX, VELY, VELZ, ETHM, BX, BY, BZ, PSI, ENTR };
enum { MASS = 0, MOMX, MOMY, MOMZ, ENER, WBX, WBY, WBZ, MPSI, MENTR };
enum { NO_BOUNDARY = 0, REFLECTING = 1, OUTFLOW = 2, DIOD = 3, INFLOW = 4 };
enum { LIM_SHIPPED = 0, LIM_MINMOD = 1, LIM_SMOOTH = 2 };

constexpr double REF_HUGE = 1e30;  // Boundary.h:5-7
constexpr double REF_PI = 3.14159265358979323846;

template <typename T> __device__ __forceinline__ T sqr(T x) { return x * x; }
// std::min / std::max argument conventions
template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

// Division and square root of the face arithmetic: IEEE double for the FP64
// path (operation for operation the reference's); for the FP32 face path the
// hardware reciprocal and square root (v_rcp_f32 / v_sqrt_f32, about 1 ulp).
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Problem_compute_pressure (Problem.cpp:234-237)
template <typename T> __device__ __forceinline__ T pressure(T gamma, T ethm) {
  return gamma > T(1.0) ? (gamma - T(1.0)) * ethm : ethm;
}

// Fluid::to_primitive (Fluid.h:61-98), total-energy form
__device__ __forceinline__ void to_primitive(const double *U, double volume, double *W) {
  const double iM = 1.0 / U[MASS];
  const double iV = 1.0 / volume;
  W[DENS] = U[MASS] * iV;
  W[VELX] = U[MOMX] * iM;
  W[VELY] = U[MOMY] * iM;
  W[VELZ] = U[MOMZ] * iM;
  W[BX] = U[WBX] * iV;
  W[BY] = U[WBY] * iV;
  W[BZ] = U[WBZ] * iV;
  W[PSI] = U[MPSI] * iM;
  W[ENTR] = U[MENTR] * iM;
  W[ETHM] = U[ENER] * iV - 0.5 * (sqr(W[BX]) + sqr(W[BY]) + sqr(W[BZ])) -
            0.5 * (sqr(W[VELX]) + sqr(W[VELY]) + sqr(W[VELZ])) * W[DENS];
}

// Fluid::to_conservative (Fluid.h:100-135)
__device__ __forceinline__ void to_conservative(const double *W, double volume, double *U) {
  const double V = volume;
  const double M = W[DENS] * volume;
  U[MASS] = W[DENS] * V;
  U[MOMX] = W[VELX] * M;
  U[MOMY] = W[VELY] * M;
  U[MOMZ] = W[VELZ] * M;
  U[WBX] = W[BX] * V;
  U[WBY] = W[BY] * V;
  U[WBZ] = W[BZ] * V;
  U[MPSI] = W[PSI] * M;
  U[MENTR] = W[ENTR] * M;
  U[ENER] = W[ETHM] * V + 0.5 * (sqr(W[BX]) + sqr(W[BY]) + sqr(W[BZ])) * V +
            0.5 * (sqr(W[VELX]) + sqr(W[VELY]) + sqr(W[VELZ])) * M;
}

// Phi of the limiter modes (computeReconstruction.cpp:5-17, slopeLimiter.cpp:5-18,
// MPI/compute_reconstruction.cpp:5-15)
__device__ __forceinline__ double phi(int mode, double x) {
  if (mode == LIM_SHIPPED) return 1.0;
  if (mode == LIM_MINMOD) return smin(1.0, x);
  const double fac = 0.2, eps1 = 1.0e-16;
  return (x * x + 2.0 * x + fac * eps1) / (x * x + x + 2.0 + fac * eps1);
}
// computeReconstruction.cpp:19-40
__device__ __forceinline__ double limiter_rec(int mode, double dWi, double mn, double mx) {
  if (dabs(dWi) <= 1.0e-16 * smax(dabs(mx), dabs(mn))) return 1.0;
  const double invWi = 1.0 / dWi;
  if (dWi < 0.0) return phi(mode, mn * invWi);
  return phi(mode, mx * invWi);
}
// slopeLimiter.cpp:20-25
__device__ __forceinline__ double limiter_sl(int mode, double dWi, double mn, double mx) {
  if (dabs(dWi) <= 1.0e-16 * smax(dabs(mx), dabs(mn))) return 1.0;
  if (dWi < 0.0) return phi(mode, mn / dWi);
  return phi(mode, mx / dWi);
}

// limiters in the face precision T: the exact double versions above for the
// FP64 path, hardware reciprocal for the FP32 face path
template <int MODE, typename T> __device__ __forceinline__ T phi_c(T x) {
  if constexpr (MODE == LIM_MINMOD) {
    return smin(T(1.0), x);
  } else if constexpr (MODE == LIM_SMOOTH) {
    const T fac = T(0.2), eps1 = T(1.0e-16);
    return fdiv(x * x + T(2.0) * x + fac * eps1, x * x + x + T(2.0) + fac * eps1);
  } else {
    return T(1.0);
  }
}
// The mode is a compile-time constant.  FP32: both outcomes are formed and the
// flat-difference test selects (no divergent branch per variable); the ratio
// of a flat difference may be inf/NaN and is discarded by the select.
template <int MODE, typename T> __device__ __forceinline__ T limiter_rec_c(T dWi, T mn, T mx) {
  if constexpr (sizeof(T) == 8) {
    return limiter_rec(MODE, dWi, mn, mx);
  } else {
    const bool flat = dabs(dWi) <= T(1.0e-16) * smax(dabs(mx), dabs(mn));
    const T r = phi_c<MODE, T>((dWi < T(0.0) ? mn : mx) * frcp(dWi));
    return flat ? T(1.0) : r;
  }
}
template <int MODE, typename T> __device__ __forceinline__ T limiter_sl_c(T dWi, T mn, T mx) {
  if constexpr (sizeof(T) == 8) {
    return limiter_sl(MODE, dWi, mn, mx);
  } else {
    const bool flat = dabs(dWi) <= T(1.0e-16) * smax(dabs(mx), dabs(mn));
    const T r = phi_c<MODE, T>(fdiv(dWi < T(0.0) ? mn : mx, dWi));
    return flat ? T(1.0) : r;
  }
}

// riemannSolver_HLLD (computeFlux.cpp:178-425): HLLD sampled at the face speed w,
// flux of the moving interface in the a*U** + b*U* + c*U form.
template <typename T>
__device__ __forceinline__ void riemann_hlld(T gamma_gas, T *flux, const T Bx, const T w,
                                             const T dens_L, const T pres_L, const T ethm_L,
                                             const T velx_L, const T vely_L, const T velz_L,
                                             const T By_L, const T Bz_L, const T dens_R,
                                             const T pres_R, const T ethm_R, const T velx_R,
                                             const T vely_R, const T velz_R, const T By_R,
                                             const T Bz_R) {
  const T c0 = T(0.0), c1 = T(1.0);
  const T signBx = Bx == c0 ? c0 : (Bx > c0 ? +c1 : -c1);
  const T momx_L = dens_L * velx_L;
  const T momx_R = dens_R * velx_R;
  const T chalf = T(0.5);
  const T B2_L = sqr(Bx) + sqr(By_L) + sqr(Bz_L);
  const T B2_R = sqr(Bx) + sqr(By_R) + sqr(Bz_R);
  const T gpl = gamma_gas * pres_L;
  const T gpr = gamma_gas * pres_R;
  const T gpbl = gpl + B2_L;
  const T gpbr = gpr + B2_R;
  const T cfl2 = fdiv(gpbl, dens_L);
  const T cfr2 = fdiv(gpbr, dens_R);
  const T cfmax = fsqrt(smax(cfl2, cfr2));
  const T S_L = smin(velx_L, velx_R) - cfmax;
  const T S_R = smax(velx_L, velx_R) + cfmax;
  const T pT_L = pres_L + chalf * B2_L;
  const T pT_R = pres_R + chalf * B2_R;
  const T iSM = frcp((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);
  const T S_M = iSM * ((S_R - velx_R) * momx_R - (S_L - velx_L) * momx_L - pT_R + pT_L);
  const T ipTs = frcp((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);
  const T pT_s = ipTs * ((S_R - velx_R) * dens_R * pT_L - (S_L - velx_L) * dens_L * pT_R +
                         dens_L * dens_R * (S_R - velx_R) * (S_L - velx_L) * (velx_R - velx_L));
  const T B2x = Bx * Bx;
  const T iSLmSM = frcp(S_L - S_M);
  const T iSRmSM = frcp(S_R - S_M);
  const T dens_L_s = dens_L * (S_L - velx_L) * iSLmSM;
  const T dens_R_s = dens_R * (S_R - velx_R) * iSRmSM;
  const T divL = dens_L * (S_L - velx_L) * (S_L - S_M) - B2x;
  const T divR = dens_R * (S_R - velx_R) * (S_R - S_M) - B2x;
  const T idivL = (divL != c0) ? frcp(divL) : c0;
  const T idivR = (divR != c0) ? frcp(divR) : c0;
  const T vely_L_s = vely_L - Bx * By_L * (S_M - velx_L) * idivL;
  const T velz_L_s = velz_L - Bx * Bz_L * (S_M - velx_L) * idivL;
  const T By_L_s = By_L * (dens_L * sqr(S_L - velx_L) - B2x) * idivL;
  const T Bz_L_s = Bz_L * (dens_L * sqr(S_L - velx_L) - B2x) * idivL;
  const T vely_R_s = vely_R - Bx * By_R * (S_M - velx_R) * idivR;
  const T velz_R_s = velz_R - Bx * Bz_R * (S_M - velx_R) * idivR;
  const T By_R_s = By_R * (dens_R * sqr(S_R - velx_R) - B2x) * idivR;
  const T Bz_R_s = Bz_R * (dens_R * sqr(S_R - velx_R) - B2x) * idivR;
  T sDens_L_s, sDens_R_s, S_L_s, S_R_s;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: one reciprocal square root gives both sqrt(rho*) and Bx / sqrt(rho*)
    const T rL = __builtin_amdgcn_rsqf(dens_L_s), rR = __builtin_amdgcn_rsqf(dens_R_s);
    sDens_L_s = dens_L_s * rL;
    sDens_R_s = dens_R_s * rR;
    S_L_s = S_M - dabs(Bx * rL);
    S_R_s = S_M + dabs(Bx * rR);
  } else {
    sDens_L_s = fsqrt(dens_L_s);
    sDens_R_s = fsqrt(dens_R_s);
    S_L_s = S_M - dabs(fdiv(Bx, sDens_L_s));
    S_R_s = S_M + dabs(fdiv(Bx, sDens_R_s));
  }
  const T idsqroot = frcp(sDens_L_s + sDens_R_s);
  const T vely_ss =
      idsqroot * (sDens_L_s * vely_L_s + sDens_R_s * vely_R_s + (By_R_s - By_L_s) * signBx);
  const T velz_ss =
      idsqroot * (sDens_L_s * velz_L_s + sDens_R_s * velz_R_s + (Bz_R_s - Bz_L_s) * signBx);
  const T By_ss = idsqroot * (sDens_L_s * By_R_s + sDens_R_s * By_L_s +
                              sDens_L_s * sDens_R_s * (vely_R_s - vely_L_s) * signBx);
  const T Bz_ss = idsqroot * (sDens_L_s * Bz_R_s + sDens_R_s * Bz_L_s +
                              sDens_L_s * sDens_R_s * (velz_R_s - velz_L_s) * signBx);
  const bool w_lt_SL = w <= S_L;
  const bool w_le_SLs = w <= S_L_s;
  const bool w_le_SM = w <= S_M;
  const bool w_le_SRs = w <= S_R_s;
  const bool w_le_SR = w <= S_R;
  const T cnull = c0;
  const T a = (w_le_SLs ? cnull : (w_le_SM ? S_L_s - w : (w_le_SRs ? S_R_s - w : cnull)));
  const T b =
      (w_lt_SL ? cnull
               : (w_le_SLs ? +S_L - w
                           : (w_le_SM ? -S_L_s + S_L
                                      : (w_le_SRs ? -S_R_s + S_R : (w_le_SR ? +S_R - w : cnull)))));
  const T c =
      (w_lt_SL ? -w : (w_le_SLs ? -S_L : (w_le_SM ? -S_L : (w_le_SRs ? -S_R : (w_le_SR ? -S_R : -w)))));
  // The flux takes the states of one side of the contact (L for w <= S_M, R
  // otherwise), so that side's terms are formed alone: its inputs are
  // selected first.  Each term is the expression the reference evaluates for
  // that side (momx_L = dens_L velx_L, etot_L_s, Fetot_L, ...), operation for
  // operation, so the value is the one the reference selects.
  const T dens = w_le_SM ? dens_L : dens_R;
  const T velx = w_le_SM ? velx_L : velx_R;
  const T vely = w_le_SM ? vely_L : vely_R;
  const T velz = w_le_SM ? velz_L : velz_R;
  const T by = w_le_SM ? By_L : By_R;
  const T bz = w_le_SM ? Bz_L : Bz_R;
  const T ethm = w_le_SM ? ethm_L : ethm_R;
  const T B2 = w_le_SM ? B2_L : B2_R;
  const T pT = w_le_SM ? pT_L : pT_R;
  const T S_X = w_le_SM ? S_L : S_R;
  const T iSXmSM = w_le_SM ? iSLmSM : iSRmSM;
  const T dens_s = w_le_SM ? dens_L_s : dens_R_s;
  const T vely_s = w_le_SM ? vely_L_s : vely_R_s;
  const T velz_s = w_le_SM ? velz_L_s : velz_R_s;
  const T by_s = w_le_SM ? By_L_s : By_R_s;
  const T bz_s = w_le_SM ? Bz_L_s : Bz_R_s;
  const T sDens_s = w_le_SM ? sDens_L_s : sDens_R_s;
  const T velx_s = S_M, velx_ss = S_M;
  const T momx = w_le_SM ? momx_L : momx_R;
  const T momy = dens * vely;
  const T momz = dens * velz;
  const T v2 = sqr(velx) + sqr(vely) + sqr(velz);
  const T etot = ethm + chalf * (dens * v2 + B2);
  const T vB = velx * Bx + vely * by + velz * bz;
  const T vB_s = velx_s * Bx + vely_s * by_s + velz_s * bz_s;
  const T etot_s = ((S_X - velx) * etot - pT * velx + pT_s * S_M + Bx * (vB - vB_s)) * iSXmSM;
  const T vB_ss = velx_ss * Bx + vely_ss * By_ss + velz_ss * Bz_ss;
  // etot_L_ss = etot_L_s - d, etot_R_ss = etot_R_s + d
  const T dE = sDens_s * (vB_s - vB_ss) * signBx;
  const T etot_ss = w_le_SM ? etot_s - dE : etot_s + dE;
  const T fdens = dens * velx;
  const T fmomx = momx * velx + pT - B2x;
  const T fmomy = momy * velx - Bx * by;
  const T fmomz = momz * velx - Bx * bz;
  const T fetot = etot * velx + pT * velx - Bx * vB;
  const T fby = by * velx - Bx * vely;
  const T fbz = bz * velx - Bx * velz;
  const T dens_ss = dens_s;
  const T momx_s = dens_s * velx_s;
  const T momx_ss = dens_ss * velx_ss;
  const T momy_s = dens_s * vely_s;
  const T momy_ss = dens_ss * vely_ss;
  const T momz_s = dens_s * velz_s;
  const T momz_ss = dens_ss * velz_ss;
  const T by_ss = By_ss;
  const T bz_ss = Bz_ss;

  flux[MASS] = fdens + a * dens_ss + b * dens_s + c * dens;
  flux[ENER] = fetot + a * etot_ss + b * etot_s + c * etot;
  flux[MOMX] = fmomx + a * momx_ss + b * momx_s + c * momx;
  flux[MOMY] = fmomy + a * momy_ss + b * momy_s + c * momy;
  flux[MOMZ] = fmomz + a * momz_ss + b * momz_s + c * momz;
  flux[WBY] = fby + a * by_ss + b * by_s + c * by;
  flux[WBZ] = fbz + a * bz_ss + b * bz_s + c * bz;
  flux[WBX] = -w * Bx;
}

// System::computeFlux (computeFlux.cpp:6-148): rotate into the face frame,
// boundary ghost states, Dedner face values Bn/psi, HLLD, rotate back, upwind
// psi and entropy fluxes by the sign of the mass flux.
template <typename T>
__device__ __forceinline__ void compute_flux(const T gamma_gas, const T *wij, const T *Wi,
                                             const T *Wj, const int jbnd, const T *normal,
                                             T &psi_ij, T &Bn_ij, T *flux) {
  const T c0 = T(0.0), c1 = T(1.0);
  const T ex = normal[0];
  const T ey = normal[1];
  const T ez = normal[2];
  T ds, ids;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: sin(theta) and its reciprocal from one reciprocal square root; a
    // normal within 1e-15 of the z axis takes the axis frame (no subnormal input)
    const T s2 = ex * ex + ey * ey;
    const bool axis = !(s2 > T(1.0e-30));
    const T r = __builtin_amdgcn_rsqf(axis ? c1 : s2);
    ds = axis ? c0 : s2 * r;
    ids = axis ? c0 : r;
  } else {
    ds = fsqrt(ex * ex + ey * ey);
    ids = (ds != c0) ? frcp(ds) : c0;
  }
  const T cosph = (ids == c0) ? c1 : ex * ids;
  const T sinph = (ids == c0) ? c0 : ey * ids;
  const T costh = ez;
  const T sinth = ds;
  const T Axx = cosph * sinth;
  const T Axy = sinth * sinph;
  const T Axz = costh;
  const T Ayx = -sinph;
  const T Ayy = cosph;
  const T Ayz = c0;
  const T Azx = -costh * cosph;
  const T Azy = -costh * sinph;
  const T Azz = sinth;
  T dens_L = Wi[DENS];
  T ethm_L = Wi[ETHM];
  T velx_L = Axx * Wi[VELX] + Axy * Wi[VELY] + Axz * Wi[VELZ];
  T vely_L, By_L, vely_R, By_R;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: the second frame axis has no z component (Ayz = 0), so its term is
    // left out rather than multiplied by zero
    vely_L = Ayx * Wi[VELX] + Ayy * Wi[VELY];
    By_L = Ayx * Wi[BX] + Ayy * Wi[BY];
    vely_R = Ayx * Wj[VELX] + Ayy * Wj[VELY];
    By_R = Ayx * Wj[BX] + Ayy * Wj[BY];
  } else {
    vely_L = Ayx * Wi[VELX] + Ayy * Wi[VELY] + Ayz * Wi[VELZ];
    By_L = Ayx * Wi[BX] + Ayy * Wi[BY] + Ayz * Wi[BZ];
    vely_R = Ayx * Wj[VELX] + Ayy * Wj[VELY] + Ayz * Wj[VELZ];
    By_R = Ayx * Wj[BX] + Ayy * Wj[BY] + Ayz * Wj[BZ];
  }
  T velz_L = Azx * Wi[VELX] + Azy * Wi[VELY] + Azz * Wi[VELZ];
  T Bx_L = Axx * Wi[BX] + Axy * Wi[BY] + Axz * Wi[BZ];
  T Bz_L = Azx * Wi[BX] + Azy * Wi[BY] + Azz * Wi[BZ];
  T psiL = Wi[PSI];
  T entrL = Wi[ENTR];
  T dens_R = Wj[DENS];
  T ethm_R = Wj[ETHM];
  T velx_R = Axx * Wj[VELX] + Axy * Wj[VELY] + Axz * Wj[VELZ];
  T velz_R = Azx * Wj[VELX] + Azy * Wj[VELY] + Azz * Wj[VELZ];
  T Bx_R = Axx * Wj[BX] + Axy * Wj[BY] + Axz * Wj[BZ];
  T Bz_R = Azx * Wj[BX] + Azy * Wj[BY] + Azz * Wj[BZ];
  T psiR = Wj[PSI];
  T entrR = Wj[ENTR];

  T pres_L = pressure(gamma_gas, Wi[ETHM]);
  T pres_R = pressure(gamma_gas, Wj[ETHM]);

  const T wn_ij = Axx * wij[0] + Axy * wij[1] + Axz * wij[2];
  switch (jbnd) {
    case REFLECTING:
      velx_R = -velx_R;
      break;
    case DIOD:
      velx_R = smax(velx_L - c1 * wn_ij, c0) + c1 * wn_ij;
      vely_R = vely_L;
      velz_R = velz_L;
      dens_R = dens_L;
      pres_R = pres_L;
      Bx_R = Bx_L;
      By_R = By_L;
      Bz_R = Bz_L;
      psiR = psiL;
      entrR = entrL;
      break;
    case OUTFLOW:
      dens_R = dens_L;
      pres_R = pres_L;
      Bx_R = Bx_L;
      By_R = By_L;
      Bz_R = Bz_L;
      psiR = psiL;
      entrR = entrL;
      break;
    default:
      break;
  }

  const T cfl2 = fdiv(gamma_gas * pres_L + sqr(Bx_L) + sqr(By_L) + sqr(Bz_L), dens_L);
  const T cfr2 = fdiv(gamma_gas * pres_R + sqr(Bx_R) + sqr(By_R) + sqr(Bz_R), dens_R);
  T chL, chR;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: the reference's 0 * (v_n - w)^2 term is exactly 0 for finite states
    chL = fsqrt(cfl2);
    chR = fsqrt(cfr2);
  } else {
    chL = fsqrt(cfl2 + c0 * sqr(velx_L - wn_ij));
    chR = fsqrt(cfr2 + c0 * sqr(velx_R - wn_ij));
  }

  Bn_ij = fdiv(Bx_L * chL + Bx_R * chR + (psiL - psiR), chL + chR);
  psi_ij = fdiv(psiL * chR + psiR * chL + chR * chL * (Bx_L - Bx_R), chL + chR);

  T F[NF];
  riemann_hlld<T>(gamma_gas, F, Bn_ij, wn_ij, dens_L, pres_L, ethm_L, velx_L, vely_L, velz_L, By_L,
                  Bz_L, dens_R, pres_R, ethm_R, velx_R, vely_R, velz_R, By_R, Bz_R);

  const T iAxx = cosph * sinth;
  const T iAxy = -sinph;
  const T iAxz = -costh * cosph;
  const T iAyx = sinth * sinph;
  const T iAyy = cosph;
  const T iAyz = -costh * sinph;
  const T iAzx = costh;
  const T iAzy = c0;
  const T iAzz = sinth;
  flux[MASS] = F[MASS];
  flux[ENER] = F[ENER];
  flux[MOMX] = iAxx * F[MOMX] + iAxy * F[MOMY] + iAxz * F[MOMZ];
  flux[MOMY] = iAyx * F[MOMX] + iAyy * F[MOMY] + iAyz * F[MOMZ];
  flux[WBX] = iAxx * F[WBX] + iAxy * F[WBY] + iAxz * F[WBZ];
  flux[WBY] = iAyx * F[WBX] + iAyy * F[WBY] + iAyz * F[WBZ];
  if constexpr (std::is_same<T, float>::value) {
    // FP32: iAzy = 0, as above
    flux[MOMZ] = iAzx * F[MOMX] + iAzz * F[MOMZ];
    flux[WBZ] = iAzx * F[WBX] + iAzz * F[WBZ];
  } else {
    flux[MOMZ] = iAzx * F[MOMX] + iAzy * F[MOMY] + iAzz * F[MOMZ];
    flux[WBZ] = iAzx * F[WBX] + iAzy * F[WBY] + iAzz * F[WBZ];
  }
  flux[MPSI] = F[MASS] * (F[MASS] > c0 ? psiL : psiR);
  flux[MENTR] = F[MASS] * (F[MASS] > c0 ? entrL : entrR);
}

}  // namespace fvm
