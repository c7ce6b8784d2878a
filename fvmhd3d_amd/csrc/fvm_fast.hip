// Device-resident stepping of the Voronoi-face pipeline ("fused" path):
// static mesh, every cell active, one global timestep (all cells on the
// smallest rung).  Per step three launches:
//
//   K1 k_cfl_predict_recon  one thread per cell, one pass over its face-ends:
//        computeTimestep (CFL, FP64), computePredictor (first-order HLLD sweep
//        + FP64 cell update -> Wt), computeReconstruction (gradients, min/max)
//   K1b k_limit             opt-in limiters only (slopeLimiter + apply)
//   K2 k_flux_apply         one thread per cell, one pass over its face-ends:
//        space-time extrapolation + HLLD (computeFluidUpdate), FP64 sums of
//        flux*dt*A in the cell's face order, then computePrimitives (FP64
//        pressure recovery, Dedner/Powell terms, fallbacks)
//   K3 k_advance            t += dt, reset the CFL minimum
//
// Each face is evaluated by both of its cells in the same canonical
// orientation (L = s1, R = s2, normal L->R), so both see bit-identical fluxes:
// the scatter is conservative, deterministic and needs no atomics.  Face
// arithmetic runs in T = double (the plain FP64 path, operation-for-operation
// the reference's) or T = float (FP32 face path); accumulation into conserved
// variables, pressure recovery and the CFL reduction are FP64 in both.
//
// Face-ends are stored per cell in CSR order (the cell's face list order),
// structure of arrays, so the streams of a cell block are read sequentially;
// cells are expected in a space-filling-curve order so that neighbour gathers
// hit L2.  Blocks are mapped XCD-aware: the 8 XCDs each sweep a contiguous
// eighth of the cells, keeping neighbour reuse inside one XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"
#include "fvm_physics.h"

using namespace fvm;

namespace fvm {

constexpr int FB = 256;  // threads per block of the fused kernels

__device__ __forceinline__ int xcd_cell(int nb) {
  // gridDim.x is padded to a multiple of 8; block b runs on XCD b % 8
  const int per = (int)gridDim.x >> 3;
  const int tile = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
  (void)nb;
  return tile * FB + (int)threadIdx.x;
}

template <typename T> struct FusedArgs {
  int n;                    // cells (all active)
  const int *off;           // [n+1] face-end offsets
  const int *nbr;           // [E] neighbour j (>= 0: this cell is L = s1; ~j: this cell is R)
  const T *fn;              // [3][E] face.n (area vector, s1 -> s2)
  const T *drL;             // [3][E] centroid - pos[s1]
  const T *cen;             // [3][E] wrapped centroid - pos[this cell] (FP64 only)
  int E;
  const double *cvol;       // [n]
  const double *mvel;       // [3][n] generator velocity (moving mesh only)
  const T *Wg;              // [10][n] gathered primitive state (this step)
  T *Wg_next;               // [10][n]
  double *W64;              // [10][n] FP64 primitive state (own cell only)
  double *U;                // [10][n] FP64 conserved state (own cell only)
  T *grad;                  // [30][n] x[10], y[10], z[10]
  T *Wt;                    // [10][n]
  T *mm;                    // [20][n] min/max (opt-in limiters)
  double *dt_new;           // [n]
  unsigned long long *dtmin_bits;
  double *tdev;             // [0] t_global, [1] dt of the last step
  unsigned long long *diag;
  double gamma, courant, dtmax;
  int limiter;
};

#define SF(p, k, i, n) (p)[(size_t)(k) * (size_t)(n) + (size_t)(i)]
#define FLD(p, k, i, n) SF(p, k, i, n)

// geometry of one face-end: canonical (L->R) normal, area, drL, and this cell's
// reconstruction centroid/normal
template <typename T> struct FE {
  int j;
  bool isL;
  T area, nrm[3], drL[3], drR[3], cen[3], ncen[3];
};

template <typename T>
__device__ __forceinline__ void load_fe(const FusedArgs<T> &A, int e, FE<T> &g) {
  const int v = A.nbr[e];
  g.isL = v >= 0;
  g.j = g.isL ? v : ~v;
  const T fx = SF(A.fn, 0, e, A.E), fy = SF(A.fn, 1, e, A.E), fz = SF(A.fn, 2, e, A.E);
  g.area = dsqrt(fx * fx + fy * fy + fz * fz);
  const T ia = T(1.0) / g.area;
  g.nrm[0] = fx * ia;
  g.nrm[1] = fy * ia;
  g.nrm[2] = fz * ia;
  for (int d = 0; d < 3; d++) g.drL[d] = SF(A.drL, d, e, A.E);
  const T dsh = g.drL[0] * g.nrm[0] + g.drL[1] * g.nrm[1] + g.drL[2] * g.nrm[2];
  const T dsl = T(2.0) * dsh;
  for (int d = 0; d < 3; d++) g.drR[d] = g.drL[d] - g.nrm[d] * dsl;
  if constexpr (std::is_same<T, double>::value) {
    // the reference's reconstruction geometry (computeReconstruction.cpp:97-123)
    for (int d = 0; d < 3; d++) g.cen[d] = SF(A.cen, d, e, A.E);
    const T cn = g.cen[0] * fx + g.cen[1] * fy + g.cen[2] * fz;
    const T sc = (cn < 0.0) ? (-1.0 / g.area) : (1.0 / g.area);
    g.ncen[0] = fx * sc;
    g.ncen[1] = fy * sc;
    g.ncen[2] = fz * sc;
  } else {
    for (int d = 0; d < 3; d++) {
      g.cen[d] = g.isL ? g.drL[d] : g.drR[d];
      g.ncen[d] = g.isL ? g.nrm[d] : -g.nrm[d];
    }
  }
}

// face velocity with Springel's correction (computeFluidUpdate.cpp:194-207), canonical orientation
template <typename T, bool MOVING>
__device__ __forceinline__ void face_velocity(const FusedArgs<T> &A, int L, int R, const FE<T> &g, T *wij,
                                              T &dV) {
  if constexpr (!MOVING) {
    wij[0] = wij[1] = wij[2] = T(0.0);
    dV = T(0.0);
  } else {
    T vi[3], vj[3];
    for (int d = 0; d < 3; d++) {
      vi[d] = (T)SF(A.mvel, d, L, A.n);
      vj[d] = (T)SF(A.mvel, d, R, A.n);
    }
    const T dsh = g.drL[0] * g.nrm[0] + g.drL[1] * g.nrm[1] + g.drL[2] * g.nrm[2];
    const T dsl = T(2.0) * dsh;
    T fij[3], rij[3];
    for (int d = 0; d < 3; d++) {
      fij[d] = g.drL[d] - g.nrm[d] * dsh;
      rij[d] = g.nrm[d] * dsl;
    }
    const T fdv = (vj[0] - vi[0]) * fij[0] + (vj[1] - vi[1]) * fij[1] + (vj[2] - vi[2]) * fij[2];
    const T r = T(1.0) / sqr(dsl);
    for (int d = 0; d < 3; d++) wij[d] = (vj[d] + vi[d]) * T(0.5) - rij[d] * fdv * r;
    dV = g.area * (rij[0] * wij[0] + rij[1] * wij[1] + rij[2] * wij[2]) / dsl;
  }
}

__device__ __forceinline__ int exp_of_d(double dt) {
  return (int)(((unsigned long long)__double_as_longlong(dt) >> 52) & 0xfff);
}
// Scheduler::get_dt(get_rung(dt)) with min_rung = 0 (Scheduler.h:94-111)
__device__ __forceinline__ double rung_dt0(double dtmax, double dt) {
  int rung = exp_of_d(dtmax) - exp_of_d(dt);
  if (rung < 0) rung = 0;
  const double dt_tick = dtmax / (double)(1ULL << 31);
  return dt_tick * (double)(1ULL << (31 - rung));
}

__device__ __forceinline__ void wave_count(unsigned long long *ctr, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m && __lane_id() == (__ffsll((long long)m) - 1)) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

__device__ __forceinline__ double ethm_from_entropy_d(double g, const double *W) {
  return W[ENTR] * pow(W[DENS], g) / (g - 1.0);
}
__device__ __forceinline__ double entropy_from_ethm_d(double g, const double *W) {
  return (g - 1.0) * W[ETHM] / pow(W[DENS], g);
}

// ---------------------------------------------------------------------------
// K1: computeTimestep + computePredictor + computeReconstruction
// ---------------------------------------------------------------------------
template <typename T, bool MOVING>
__global__ void __launch_bounds__(FB) k_cfl_predict_recon(FusedArgs<T> A) {
  const int c = xcd_cell(A.n);
  const bool live = c < A.n;
  double dtn = REF_HUGE;
  if (live) {
    const int n = A.n;
    const double g64 = A.gamma;
    const T g = (T)A.gamma;
    const double V = A.cvol[c];
    double U0[NF], W0[NF];
    for (int k = 0; k < NF; k++) U0[k] = SF(A.U, k, c, n);
    to_primitive(U0, V, W0);
    T Wc[NF];
    for (int k = 0; k < NF; k++) Wc[k] = SF(A.Wg, k, c, n);

    // CFL (computeTimestep.cpp:13-31)
    const double h = pow(V / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
    double pv[3] = {0.0, 0.0, 0.0};
    if constexpr (MOVING)
      for (int d = 0; d < 3; d++) pv[d] = SF(A.mvel, d, c, n);
    const double Pi = pressure(g64, W0[ETHM]) * g64;
    const double B2i = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
    const double csi = dsqrt((Pi + B2i) / W0[DENS]);
    const double dvr[3] = {W0[VELX] - pv[0], W0[VELY] - pv[1], W0[VELZ] - pv[2]};
    const double vrel = dsqrt(dvr[0] * dvr[0] + dvr[1] * dvr[1] + dvr[2] * dvr[2]);
    double dti = h / (csi + vrel + 1.0 / REF_HUGE);
    double arg_max = -1.0;  // min_j h/(cs + sqrt(a_j)) = h/(cs + sqrt(max_j a_j)), monotone roundings

    T gx[NF], gy[NF], gz[NF], dmn[NF], dmx[NF], pU[NF], pgp[3], pdv;
    double pdV = 0.0;
    for (int k = 0; k < NF; k++) gx[k] = gy[k] = gz[k] = dmn[k] = dmx[k] = pU[k] = T(0.0);
    pgp[0] = pgp[1] = pgp[2] = pdv = T(0.0);

    const int eb = A.off[c], ee = A.off[c + 1];
    for (int e = eb; e < ee; e++) {
      FE<T> fe;
      load_fe(A, e, fe);
      const int j = fe.j;
      T Wj[NF];
      for (int k = 0; k < NF; k++) Wj[k] = SF(A.Wg, k, j, n);

      // --- reconstruction sums (computeReconstruction.cpp:108-149)
      {
        const T dsh = fe.cen[0] * fe.ncen[0] + fe.cen[1] * fe.ncen[1] + fe.cen[2] * fe.ncen[2];
        const T ids = (dsh != T(0.0)) ? T(0.5) / dsh : T(0.0);
        const T idsA = fe.area * ids;
        T drh[3], fij[3];
        for (int d = 0; d < 3; d++) {
          drh[d] = fe.ncen[d] * dsh;
          fij[d] = fe.cen[d] - drh[d];
        }
        for (int k = 0; k < NF; k++) {
          const T sum = Wj[k] + Wc[k];
          const T diff = Wj[k] - Wc[k];
          dmn[k] = smin(dmn[k], diff);
          dmx[k] = smax(dmx[k], diff);
          gx[k] += (sum * drh[0] + diff * fij[0]) * idsA;
          gy[k] += (sum * drh[1] + diff * fij[1]) * idsA;
          gz[k] += (sum * drh[2] + diff * fij[2]) * idsA;
        }
      }
      // --- CFL neighbour term (computeTimestep.cpp:37-54)
      {
        double dwi[3] = {0.0, 0.0, 0.0};
        if constexpr (MOVING)
          for (int d = 0; d < 3; d++) dwi[d] = (double)(T)SF(A.mvel, d, j, n) - pv[d];
        const double dvi[3] = {(double)Wj[VELX] - W0[VELX], (double)Wj[VELY] - W0[VELY],
                               (double)Wj[VELZ] - W0[VELZ]};
        const double a = (dwi[0] * dwi[0] + dwi[1] * dwi[1] + dwi[2] * dwi[2]) +
                         1.0 * (dvi[0] * dvi[0] + dvi[1] * dvi[1] + dvi[2] * dvi[2]) + 1.0 / REF_HUGE;
        arg_max = (arg_max < a) ? a : arg_max;
      }
      // --- predictor face flux, first order (computePredictor.cpp:44-136)
      {
        const int L = fe.isL ? c : j, R = fe.isL ? j : c;
        const T *WL = fe.isL ? Wc : Wj;
        const T *WR = fe.isL ? Wj : Wc;
        T wij[3], dV;
        face_velocity<T, MOVING>(A, L, R, fe, wij, dV);
        T psi, Bn, flux[NF];
        compute_flux<T>(g, wij, WL, WR, NO_BOUNDARY, fe.nrm, psi, Bn, flux);
        if (fe.isL) {
          for (int k = 0; k < NF; k++) pU[k] = pU[k] - flux[k] * fe.area;
          for (int d = 0; d < 3; d++) pgp[d] = pgp[d] + fe.nrm[d] * (psi * fe.area);
          pdv = pdv + Bn * fe.area;
          pdV += (double)dV;
        } else {
          pdV -= (double)dV;
          for (int k = 0; k < NF; k++) pU[k] = pU[k] + flux[k] * fe.area;
          for (int d = 0; d < 3; d++) pgp[d] = pgp[d] - fe.nrm[d] * (psi * fe.area);
          pdv = pdv - Bn * fe.area;
        }
      }
    }
    if (arg_max >= 0.0) dti = smin(dti, h / (csi + dsqrt(arg_max)));
    dti = smin(dti, REF_HUGE);
    dtn = dti * A.courant;

    // --- predictor cell update (computePredictor.cpp:141-379), FP64
    {
      while (U0[MASS] + (double)pU[MASS] * dtn <= 0.0) dtn *= 0.5;
      const double dt = rung_dt0(A.dtmax, dtn) / 128;
      const double divB = (double)pdv;
      double Uc[NF];
      for (int k = 0; k < NF; k++) Uc[k] = U0[k] + (double)pU[k] * dt;
      Uc[WBX] -= W0[VELX] * divB * dt;
      Uc[WBY] -= W0[VELY] * divB * dt;
      Uc[WBZ] -= W0[VELZ] * divB * dt;
      const double pres = pressure(g64, W0[ETHM]);
      const double B2 = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
      const double dcs2 = g64 * pres + B2;
      const double ch_sig = dsqrt(dcs2 / W0[DENS]);
      Uc[WBX] -= (double)pgp[0] * dt;
      Uc[WBY] -= (double)pgp[1] * dt;
      Uc[WBZ] -= (double)pgp[2] * dt;
      Uc[MPSI] -= dcs2 * divB * dt;
      const double cr = 0.5;
      const double l_min = pow(V / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
      Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
      double W1[NF];
      to_primitive(Uc, V * exp(pdV * dt / V), W1);
      const bool ef = W1[ETHM] <= 0.0;
      if (ef) W1[ETHM] = ethm_from_entropy_d(g64, W1);
      else W1[ENTR] = entropy_from_ethm_d(g64, W1);
      wave_count(&A.diag[FVM_DIAG_ETHM_FAILED], ef);
      const double idt = 1.0 / dt;
      T tv[NF];
      for (int k = 0; k < NF; k++) tv[k] = (T)((W1[k] - W0[k]) * idt);

      // --- reconstruction finish (computeReconstruction.cpp:162-239)
      const double invV = 1.0 / V;
      for (int k = 0; k < NF; k++) {
        gx[k] = (T)((double)gx[k] * invV);
        gy[k] = (T)((double)gy[k] * invV);
        gz[k] = (T)((double)gz[k] * invV);
      }
      if (A.limiter != LIM_SHIPPED) {
        double lm[NF];
        for (int k = 0; k < NF; k++) lm[k] = 1.0;
        for (int e = eb; e < ee; e++) {
          FE<T> fe;
          load_fe(A, e, fe);
          for (int k = 0; k < NF; k++) {
            const double dWx = (double)(gx[k] * fe.cen[0] + gy[k] * fe.cen[1] + gz[k] * fe.cen[2]);
            lm[k] = smin(lm[k], limiter_rec(A.limiter, dWx, (double)dmn[k], (double)dmx[k]));
          }
        }
        for (int k = 0; k < NF; k++) {
          SF(A.mm, k, c, n) = (T)((double)dmn[k] + (double)Wc[k]);
          SF(A.mm, NF + k, c, n) = (T)((double)dmx[k] + (double)Wc[k]);
          tv[k] = (T)((double)tv[k] * lm[k]);
          gx[k] = (T)((double)gx[k] * lm[k]);
          gy[k] = (T)((double)gy[k] * lm[k]);
          gz[k] = (T)((double)gz[k] * lm[k]);
        }
      }
      for (int k = 0; k < NF; k++) {
        SF(A.grad, k, c, n) = gx[k];
        SF(A.grad, NF + k, c, n) = gy[k];
        SF(A.grad, 2 * NF + k, c, n) = gz[k];
        SF(A.Wt, k, c, n) = tv[k];
      }
    }
    A.dt_new[c] = dtn;
  }
  // global minimum of dt_new (dt > 0: IEEE bit patterns order like the values)
  double v = dtn;
  for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o));
  if (__lane_id() == 0 && v < REF_HUGE && v >= 0.0)
    atomicMin(A.dtmin_bits, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// K1b (opt-in limiters): slopeLimiter compute part + computeFluidUpdateI for
// the global step, per cell over its own face-ends (own data only)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(FB) k_limit(FusedArgs<T> A) {
  const int c = xcd_cell(A.n);
  if (c >= A.n) return;
  const int n = A.n;
  const double t0 = A.tdev[0];
  const double dt_glob = rung_dt0(A.dtmax, __longlong_as_double((long long)*A.dtmin_bits));
  const double tend = t0 + dt_glob;
  const double dtI = tend - t0;
  const double dth = 0.5 * smin(dtI, dtI);
  double wc[NF], lo[NF], hi[NF], lm[NF];
  for (int k = 0; k < NF; k++) {
    wc[k] = (double)SF(A.Wg, k, c, n);
    lo[k] = (double)SF(A.mm, k, c, n);
    hi[k] = (double)SF(A.mm, NF + k, c, n);
    lm[k] = 1.0;
  }
  for (int e = A.off[c]; e < A.off[c + 1]; e++) {
    FE<T> fe;
    load_fe(A, e, fe);
    const T *dr = fe.isL ? fe.drL : fe.drR;
    for (int k = 0; k < NF; k++) {
      const double dWt = (double)SF(A.Wt, k, c, n) * (dtI - dth);
      const double dWx = (double)SF(A.grad, k, c, n) * (double)dr[0] + (double)SF(A.grad, NF + k, c, n) * (double)dr[1] +
                         (double)SF(A.grad, 2 * NF + k, c, n) * (double)dr[2];
      lm[k] = smin(lm[k], limiter_sl(A.limiter, dWx + dWt, smin(lo[k] - wc[k], 0.0), smax(hi[k] - wc[k], 0.0)));
    }
  }
  for (int k = 0; k < NF; k++) {
    SF(A.grad, k, c, n) = (T)((double)SF(A.grad, k, c, n) * lm[k]);
    SF(A.grad, NF + k, c, n) = (T)((double)SF(A.grad, NF + k, c, n) * lm[k]);
    SF(A.grad, 2 * NF + k, c, n) = (T)((double)SF(A.grad, 2 * NF + k, c, n) * lm[k]);
    SF(A.Wt, k, c, n) = (T)((double)SF(A.Wt, k, c, n) * lm[k]);
  }
}

// ---------------------------------------------------------------------------
// K2: computeFluidUpdate (extrapolation + HLLD + FP64 scatter sums) and
// computePrimitives (FP64) for the own cell
// ---------------------------------------------------------------------------
template <typename T, bool MOVING>
__global__ void __launch_bounds__(FB) k_flux_apply(FusedArgs<T> A) {
  const int c = xcd_cell(A.n);
  if (c >= A.n) return;
  const int n = A.n;
  const T g = (T)A.gamma;
  const double g64 = A.gamma;
  const double t0 = A.tdev[0];
  const double dt_glob = rung_dt0(A.dtmax, __longlong_as_double((long long)*A.dtmin_bits));
  const double tend = t0 + dt_glob;
  // every cell has tend = t0 + dt: dtI = dtJ, dt = min(dtI, dtJ), dti = dtI - dt/2
  const double dtI = tend - t0;
  const double dtf = smin(dtI, dtI);
  const T dti = (T)(dtI - 0.5 * dtf);

  T Wc[NF], gc[3 * NF], tc[NF];
  for (int k = 0; k < NF; k++) {
    Wc[k] = SF(A.Wg, k, c, n);
    tc[k] = SF(A.Wt, k, c, n);
  }
  for (int k = 0; k < 3 * NF; k++) gc[k] = SF(A.grad, k, c, n);

  double dU[NF], gp[3], dv = 0.0;
  for (int k = 0; k < NF; k++) dU[k] = 0.0;
  gp[0] = gp[1] = gp[2] = 0.0;
  bool bad = false;

  for (int e = A.off[c]; e < A.off[c + 1]; e++) {
    FE<T> fe;
    load_fe(A, e, fe);
    const int j = fe.j;
    T Wj[NF], gj[3 * NF], tj[NF];
    for (int k = 0; k < NF; k++) {
      Wj[k] = SF(A.Wg, k, j, n);
      tj[k] = SF(A.Wt, k, j, n);
    }
    for (int k = 0; k < 3 * NF; k++) gj[k] = SF(A.grad, k, j, n);
    const T *W0L = fe.isL ? Wc : Wj, *W0R = fe.isL ? Wj : Wc;
    const T *gL = fe.isL ? gc : gj, *gR = fe.isL ? gj : gc;
    const T *tL = fe.isL ? tc : tj, *tR = fe.isL ? tj : tc;
    T WL[NF], WR[NF];
    for (int k = 0; k < NF; k++) {
      WL[k] = W0L[k] + (gL[k] * fe.drL[0] + gL[NF + k] * fe.drL[1] + gL[2 * NF + k] * fe.drL[2] + tL[k] * dti);
      WR[k] = W0R[k] + (gR[k] * fe.drR[0] + gR[NF + k] * fe.drR[1] + gR[2 * NF + k] * fe.drR[2] + tR[k] * dti);
    }
    bad = bad || !(WL[DENS] > T(0.0)) || !(WL[ETHM] > T(0.0)) || !(WR[DENS] > T(0.0)) || !(WR[ETHM] > T(0.0));
    T wij[3], dV;
    face_velocity<T, MOVING>(A, fe.isL ? c : j, fe.isL ? j : c, fe, wij, dV);
    T psi, Bn, flux[NF];
    compute_flux<T>(g, wij, WL, WR, NO_BOUNDARY, fe.nrm, psi, Bn, flux);
    const double tarea = dtf * (double)fe.area;
    if (fe.isL) {
      for (int k = 0; k < NF; k++) dU[k] = dU[k] - (double)flux[k] * tarea;
      for (int d = 0; d < 3; d++) gp[d] = gp[d] + (double)fe.nrm[d] * ((double)psi * tarea);
      dv = dv + (double)Bn * tarea;
    } else {
      for (int k = 0; k < NF; k++) dU[k] = dU[k] + (double)flux[k] * tarea;
      for (int d = 0; d < 3; d++) gp[d] = gp[d] - (double)fe.nrm[d] * ((double)psi * tarea);
      dv = dv - (double)Bn * tarea;
    }
  }
  wave_count(&A.diag[FVM_DIAG_POSITIVITY], bad);

  // computePrimitives at t1 = t0 + dt (computeFluidUpdate.cpp:367-505)
  const double t1 = t0 + dt_glob;
  const double dt = t1 - t0;
  const double V = A.cvol[c];
  double Ui[NF], wi[NF], Uc[NF], W1[NF];
  for (int k = 0; k < NF; k++) {
    Ui[k] = SF(A.U, k, c, n);
    wi[k] = SF(A.W64, k, c, n);
    Uc[k] = Ui[k] + (0.0 + dU[k]);
  }
  const double divB = 0.0 + dv;
  const bool failed = Uc[MASS] <= 0.0;
  bool ef = false;
  if (!failed) {
    const double dth = 0.5 * dt;
    double Wh[NF];
    for (int k = 0; k < NF; k++) Wh[k] = wi[k] + (double)tc[k] * dth;
    Uc[WBX] -= Wh[VELX] * divB;
    Uc[WBY] -= Wh[VELY] * divB;
    Uc[WBZ] -= Wh[VELZ] * divB;
    const double pres = pressure(g64, Wh[ETHM]);
    const double B2 = Wh[BX] * Wh[BX] + Wh[BY] * Wh[BY] + Wh[BZ] * Wh[BZ];
    const double dcs2 = g64 * pres + B2;
    const double ch_sig = dsqrt(dcs2 / Wh[DENS]);
    Uc[WBX] -= 0.0 + gp[0];
    Uc[WBY] -= 0.0 + gp[1];
    Uc[WBZ] -= 0.0 + gp[2];
    Uc[MPSI] -= dcs2 * divB;
    // gravity terms with acc0 = acc1 = 0 (computeFluidUpdate.cpp:433-447)
    {
      const double M0 = Ui[MASS], M1 = Uc[MASS];
      double dmom[3];
      for (int d = 0; d < 3; d++) dmom[d] = ((0.0 * M0 + 0.0 * M1) * 0.5) * dt;
      Uc[MOMX] += dmom[0];
      Uc[MOMY] += dmom[1];
      Uc[MOMZ] += dmom[2];
      Uc[ENER] += ((Ui[MOMX] * 0.0 + Ui[MOMY] * 0.0 + Ui[MOMZ] * 0.0) + (Uc[MOMX] * 0.0 + Uc[MOMY] * 0.0 + Uc[MOMZ] * 0.0)) *
                  dt * 0.5;
    }
    const double cr = 0.5;
    const double l_min = pow(V / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
    Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
    to_primitive(Uc, V, W1);
    ef = W1[ETHM] <= 0.0;
    if (ef) W1[ETHM] = ethm_from_entropy_d(g64, W1);
    else W1[ENTR] = entropy_from_ethm_d(g64, W1);
  } else {
    double Ut[NF];
    for (int k = 0; k < NF; k++) W1[k] = wi[k] + (double)tc[k] * dt;
    to_conservative(W1, V, Ut);
    to_primitive(Ut, V, W1);
    if (W1[ETHM] <= 0.0) W1[ETHM] = ethm_from_entropy_d(g64, W1);
    else W1[ENTR] = entropy_from_ethm_d(g64, W1);
  }
  wave_count(&A.diag[FVM_DIAG_ETHM_FAILED], ef);
  wave_count(&A.diag[FVM_DIAG_MASS_FAILED], failed);
  to_conservative(W1, V, Ui);
  for (int k = 0; k < NF; k++) {
    SF(A.W64, k, c, n) = W1[k];
    SF(A.U, k, c, n) = Ui[k];
    SF(A.Wg_next, k, c, n) = (T)W1[k];
  }
}

__global__ void k_advance(double *tdev, unsigned long long *dtmin_bits, double dtmax, double *hist, int step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double dt = rung_dt0(dtmax, __longlong_as_double((long long)*dtmin_bits));
    const double t0 = tdev[0];
    tdev[0] = t0 + dt;
    tdev[1] = dt;
    if (hist) hist[step] = dt;
    *dtmin_bits = 0x7FF0000000000000ULL;
  }
}

// face-ends from the uploaded mesh: cell c's list cell_face[off[c]..off[c+1]);
// drL = centroid - pos[s1] and the reconstruction centroid wrap(centroid - pos[c])
// are formed in FP64 exactly as the reference forms them
template <typename T>
__global__ void k_build_faceends(int n, const int *coff, const int *cface, const int *s1, const int *s2,
                                 const double *fc, const double *fnv, int nface, const double *pos, double Lx,
                                 double Ly, double Lz, int E, int *nbr, T *fn, T *drL, T *cen) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double L[3] = {Lx, Ly, Lz};
  for (int e = coff[c]; e < coff[c + 1]; e++) {
    const int f = cface[e];
    const int a = s1[f], b = s2[f];
    const bool isL = a == c;
    nbr[e] = isL ? b : ~a;
    for (int d = 0; d < 3; d++) {
      const double cf = FLD(fc, d, f, nface);
      SF(fn, d, e, E) = (T)FLD(fnv, d, f, nface);
      SF(drL, d, e, E) = (T)(cf - FLD(pos, d, a, n));
      if (cen) {
        double v = cf - FLD(pos, d, c, n);
        if (v > 0.5 * L[d]) v -= L[d];
        else if (v < -0.5 * L[d]) v += L[d];
        SF(cen, d, e, E) = (T)v;
      }
    }
  }
}

template <typename S, typename D> __global__ void k_convert(const S *src, D *dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (D)src[i];
}

}  // namespace fvm

// ===========================================================================
// host side
// ===========================================================================
using fvm::FB;

static unsigned fused_grid(int n) {
  unsigned nb = (unsigned)((n + FB - 1) / FB);
  return (nb + 7u) & ~7u;  // multiple of 8 for the XCD mapping
}

void fvm_fast_release(fvm_ctx *c) {
  FusedState &z = c->fz;
  for (DBuf *b : {&z.nbr, &z.fn, &z.drL, &z.cen, &z.Wg[0], &z.Wg[1], &z.grad, &z.Wt, &z.mm, &z.dt_new, &z.dtmin,
                  &z.tdev, &z.hist})
    fvm_release(*b);
  for (auto &e : z.ev)
    if (e) (void)hipEventDestroy(e);
  z.ev.clear();
  z.ready = false;
}

template <typename T> static fvm::FusedArgs<T> fused_args(fvm_ctx *c, double dtmax) {
  FusedState &z = c->fz;
  fvm::FusedArgs<T> A;
  A.n = c->np;
  A.off = (const int *)c->m_coff.p;
  A.nbr = (const int *)z.nbr.p;
  A.fn = (const T *)z.fn.p;
  A.drL = (const T *)z.drL.p;
  A.cen = (const T *)z.cen.p;
  A.E = z.E;
  A.cvol = (const double *)c->m_cvol.p;
  A.mvel = (const double *)c->field[FVM_FIELD_MESH_VEL].p;
  A.Wg = (const T *)z.Wg[z.cur].p;
  A.Wg_next = (T *)z.Wg[z.cur ^ 1].p;
  A.W64 = (double *)c->field[FVM_FIELD_W].p;
  A.U = (double *)c->field[FVM_FIELD_U].p;
  A.grad = (T *)z.grad.p;
  A.Wt = (T *)z.Wt.p;
  A.mm = (T *)z.mm.p;
  A.dt_new = (double *)z.dt_new.p;
  A.dtmin_bits = (unsigned long long *)z.dtmin.p;
  A.tdev = (double *)z.tdev.p;
  A.diag = (unsigned long long *)c->diag.p;
  A.gamma = c->prm.gamma_gas;
  A.courant = c->prm.courant_no;
  A.dtmax = dtmax;
  A.limiter = c->prm.limiter_mode;
  return A;
}

template <typename T> static int fused_setup_t(fvm_ctx *c) {
  FusedState &z = c->fz;
  const int n = c->np, E = c->nref;
  z.E = E;
  int rc;
  const size_t es = sizeof(T);
  if ((rc = fvm_ensure(c, z.nbr, sizeof(int) * (size_t)E)) || (rc = fvm_ensure(c, z.fn, es * 3 * (size_t)E)) ||
      (rc = fvm_ensure(c, z.drL, es * 3 * (size_t)E)) ||
      (rc = fvm_ensure(c, z.cen, std::is_same<T, double>::value ? es * 3 * (size_t)E : 8)) ||
      (rc = fvm_ensure(c, z.Wg[0], es * 10 * (size_t)n)) || (rc = fvm_ensure(c, z.Wg[1], es * 10 * (size_t)n)) ||
      (rc = fvm_ensure(c, z.grad, es * 30 * (size_t)n)) || (rc = fvm_ensure(c, z.Wt, es * 10 * (size_t)n)) ||
      (rc = fvm_ensure(c, z.mm, es * 20 * (size_t)n)) || (rc = fvm_ensure(c, z.dt_new, sizeof(double) * (size_t)n)) ||
      (rc = fvm_ensure(c, z.dtmin, sizeof(unsigned long long))) || (rc = fvm_ensure(c, z.tdev, sizeof(double) * 2)))
    return rc;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(fvm::k_build_faceends<T>, dim3(nb), dim3(256), 0, c->stream, n, (const int *)c->m_coff.p,
                     (const int *)c->m_cface.p, (const int *)c->m_s1.p, (const int *)c->m_s2.p, (const double *)c->m_c.p,
                     (const double *)c->m_n.p, c->nface, (const double *)c->field[FVM_FIELD_POS].p,
                     c->prm.domain_size[0], c->prm.domain_size[1], c->prm.domain_size[2], E, (int *)z.nbr.p, (T *)z.fn.p,
                     (T *)z.drL.p, std::is_same<T, double>::value ? (T *)z.cen.p : (T *)nullptr);
  FVM_HIPCHK(c, hipGetLastError());
  const size_t nw = 10 * (size_t)n;
  hipLaunchKernelGGL((fvm::k_convert<double, T>), dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, c->stream,
                     (const double *)c->field[FVM_FIELD_W].p, (T *)z.Wg[0].p, nw);
  FVM_HIPCHK(c, hipGetLastError());
  const unsigned long long inf = 0x7FF0000000000000ULL;
  const double t2[2] = {c->t_global, 0.0};
  FVM_HIPCHK(c, hipMemcpyAsync(z.dtmin.p, &inf, sizeof inf, hipMemcpyHostToDevice, c->stream));
  FVM_HIPCHK(c, hipMemcpyAsync(z.tdev.p, t2, sizeof t2, hipMemcpyHostToDevice, c->stream));
  FVM_HIPCHK(c, hipMemsetAsync(z.grad.p, 0, es * 30 * (size_t)n, c->stream));
  FVM_HIPCHK(c, hipMemsetAsync(z.Wt.p, 0, es * 10 * (size_t)n, c->stream));
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  z.cur = 0;
  return FVM_OK;
}

int fvm_fused_setup(fvm_ctx *c, int precision, int moving_mesh) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fvm_fail(c, FVM_ESTATE, "fvm_fused_setup before fvm_upload_mesh");
  if (precision != FVM_FUSED_FP64 && precision != FVM_FUSED_FP32) return fvm_fail(c, FVM_EARG, "bad precision");
  if (c->nactive != c->np || c->nimport_loc != 0)
    return fvm_fail(c, FVM_EARG, "fused stepping needs every cell active on one domain (nactive == np)");
  (void)hipSetDevice(c->device);
  try {
    std::vector<int> bnd(c->np);
    FVM_HIPCHK(c, hipMemcpy(bnd.data(), c->field[FVM_FIELD_BND].p, sizeof(int) * (size_t)c->np, hipMemcpyDeviceToHost));
    for (int b : bnd)
      if (b != FVM_NO_BOUNDARY) return fvm_fail(c, FVM_EARG, "fused stepping supports periodic meshes without boundary cells");
  } catch (...) {
    return fvm_fail(c, FVM_EARG, "host allocation failed");
  }
  fvm_fast_release(c);
  c->fz.precision = precision;
  c->fz.moving = moving_mesh ? 1 : 0;
  const int rc = precision == FVM_FUSED_FP64 ? fused_setup_t<double>(c) : fused_setup_t<float>(c);
  if (rc) return rc;
  c->fz.ready = true;
  return FVM_OK;
}

template <typename T, bool MOVING> static int fused_run_t(fvm_ctx *c, int nsteps, double dtmax) {
  FusedState &z = c->fz;
  const unsigned grid = fused_grid(c->np);
  const bool lim = c->prm.limiter_mode != FVM_LIMITER_SHIPPED;
  if (z.timing)
    while ((int)z.ev.size() < 5 * nsteps) {
      hipEvent_t e;
      FVM_HIPCHK(c, hipEventCreate(&e));
      z.ev.push_back(e);
    }
  for (int s = 0; s < nsteps; s++) {
    fvm::FusedArgs<T> A = fused_args<T>(c, dtmax);
    hipEvent_t *ev = z.timing ? &z.ev[5 * s] : nullptr;
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[0], c->stream));
    hipLaunchKernelGGL((fvm::k_cfl_predict_recon<T, MOVING>), dim3(grid), dim3(FB), 0, c->stream, A);
    FVM_HIPCHK(c, hipGetLastError());
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[1], c->stream));
    if (lim) {
      hipLaunchKernelGGL((fvm::k_limit<T>), dim3(grid), dim3(FB), 0, c->stream, A);
      FVM_HIPCHK(c, hipGetLastError());
    }
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[2], c->stream));
    hipLaunchKernelGGL((fvm::k_flux_apply<T, MOVING>), dim3(grid), dim3(FB), 0, c->stream, A);
    FVM_HIPCHK(c, hipGetLastError());
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[3], c->stream));
    hipLaunchKernelGGL(fvm::k_advance, dim3(1), dim3(64), 0, c->stream, (double *)z.tdev.p,
                       (unsigned long long *)z.dtmin.p, dtmax, (double *)z.hist.p, s);
    FVM_HIPCHK(c, hipGetLastError());
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[4], c->stream));
    z.cur ^= 1;
  }
  if (z.timing && nsteps > 0) {
    FVM_HIPCHK(c, hipEventSynchronize(z.ev[5 * nsteps - 1]));
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < 4; k++) {
        float ms = 0.0f;
        FVM_HIPCHK(c, hipEventElapsedTime(&ms, z.ev[5 * s + k], z.ev[5 * s + k + 1]));
        z.ms[k] += ms;
      }
    for (int k = 0; k < 4; k++) z.launches[k] += (k == 1 && !lim) ? 0 : nsteps;
  }
  return FVM_OK;
}

int fvm_fused_run(fvm_ctx *c, int nsteps, double dtmax, double *t_out) {
  if (!c) return FVM_EARG;
  if (!c->fz.ready) return fvm_fail(c, FVM_ESTATE, "fvm_fused_run before fvm_fused_setup");
  if (nsteps < 0 || !(dtmax > 0.0)) return fvm_fail(c, FVM_EARG, "bad nsteps/dtmax");
  (void)hipSetDevice(c->device);
  FusedState &z = c->fz;
  int rc = fvm_ensure(c, z.hist, sizeof(double) * (size_t)std::max(nsteps, 1));
  if (rc) return rc;
  const bool f64 = z.precision == FVM_FUSED_FP64;
  if (f64) rc = z.moving ? fused_run_t<double, true>(c, nsteps, dtmax) : fused_run_t<double, false>(c, nsteps, dtmax);
  else rc = z.moving ? fused_run_t<float, true>(c, nsteps, dtmax) : fused_run_t<float, false>(c, nsteps, dtmax);
  if (rc) return rc;
  double t2[2];
  FVM_HIPCHK(c, hipMemcpyAsync(t2, z.tdev.p, sizeof t2, hipMemcpyDeviceToHost, c->stream));
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  c->t_global = t2[0];
  if (t_out) *t_out = t2[0];
  return FVM_OK;
}

int fvm_fused_dt_history(fvm_ctx *c, double *dts, int n) {
  if (!c || !dts || n < 0) return FVM_EARG;
  if (!c->fz.ready || (size_t)n * sizeof(double) > c->fz.hist.bytes) return fvm_fail(c, FVM_EARG, "no such history");
  (void)hipSetDevice(c->device);
  FVM_HIPCHK(c, hipMemcpy(dts, c->fz.hist.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  return FVM_OK;
}

// gradients and Wt of the last step into the per-cell fields
int fvm_fused_sync(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->fz.ready) return fvm_fail(c, FVM_ESTATE, "no fused state");
  (void)hipSetDevice(c->device);
  FusedState &z = c->fz;
  const size_t n = (size_t)c->np;
  const bool shipped = c->prm.storage == FVM_STORAGE_SHIPPED;
  const int dstf[4] = {FVM_FIELD_GRAD_X, FVM_FIELD_GRAD_Y, FVM_FIELD_GRAD_Z, FVM_FIELD_DWDT};
  for (int q = 0; q < 4; q++) {
    const size_t cnt = 10 * n;
    const unsigned nb = (unsigned)((cnt + 255) / 256);
    const char *src = (q < 3) ? (const char *)z.grad.p + (size_t)q * cnt * (z.precision == FVM_FUSED_FP64 ? 8 : 4)
                              : (const char *)z.Wt.p;
    void *dst = c->field[dstf[q]].p;
    if (z.precision == FVM_FUSED_FP64) {
      if (shipped) hipLaunchKernelGGL((fvm::k_convert<double, float>), dim3(nb), dim3(256), 0, c->stream, (const double *)src, (float *)dst, cnt);
      else hipLaunchKernelGGL((fvm::k_convert<double, double>), dim3(nb), dim3(256), 0, c->stream, (const double *)src, (double *)dst, cnt);
    } else {
      if (shipped) hipLaunchKernelGGL((fvm::k_convert<float, float>), dim3(nb), dim3(256), 0, c->stream, (const float *)src, (float *)dst, cnt);
      else hipLaunchKernelGGL((fvm::k_convert<float, double>), dim3(nb), dim3(256), 0, c->stream, (const float *)src, (double *)dst, cnt);
    }
    FVM_HIPCHK(c, hipGetLastError());
  }
  FVM_HIPCHK(c, hipMemcpyAsync(c->field[FVM_FIELD_DT_NEW].p, z.dt_new.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_fused_timing(fvm_ctx *c, int enable) {
  if (!c) return FVM_EARG;
  c->fz.timing = enable != 0;
  for (int k = 0; k < 4; k++) {
    c->fz.ms[k] = 0.0;
    c->fz.launches[k] = 0;
  }
  return FVM_OK;
}

int fvm_fused_kernel_ms(fvm_ctx *c, double *ms, long long *launches, int n) {
  if (!c || !ms || n < 0) return FVM_EARG;
  for (int k = 0; k < n && k < 4; k++) {
    ms[k] = c->fz.ms[k];
    if (launches) launches[k] = c->fz.launches[k];
  }
  return FVM_OK;
}
