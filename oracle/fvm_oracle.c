/*
 * oracle/fvm_oracle.c -- CPU restatement of fvmhd3d's per-step Voronoi-face
 * pipeline.  TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
 * path is compared against (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg).  The product library never links, loads or calls it.
 *
 * It restates, in the same arithmetic order, the reference functions
 *   System::computeReconstruction   computeReconstruction.cpp:5-264
 *   System::slopeLimiter (compute)  slopeLimiter.cpp:5-237
 *   System::computeFluidUpdateI     computeFluidUpdate.cpp:25-75
 *   System::computeFluidUpdate      computeFluidUpdate.cpp:88-247
 *   System::computeFlux             computeFlux.cpp:6-148
 *   System::riemannSolver_HLLD      computeFlux.cpp:178-425
 *   System::computePrimitives       computeFluidUpdate.cpp:347-542
 *   System::computeTimestep         computeTimestep.cpp:6-62
 *   System::computeTimestepLimiter  computeTimestep.cpp:64-110
 *   System::computePredictor        computePredictor.cpp:11-391
 *   System::get_active_faces        Iterate.cpp:298-324
 *   Fluid::to_primitive / to_conservative   Fluid.h:61-136
 *   Problem_compute_pressure, entropy<->ethm, enforce_limiter   Problem.cpp:234-254
 *
 * It is compiled twice (oracle/Makefile):
 *   SREAL=double  the all-FP64 oracle (the reference with every float field
 *                 widened to double);
 *   SREAL=float   the reference "as shipped": Fluid_flt/fvec3 storage of
 *                 gradients, time derivatives, min/max, limiters, mesh
 *                 velocity and the flux accumulator FluidD (Fluid.h:140,
 *                 154-206), with the same rounding points.
 * Build with -O2 -ffp-contract=off and no fast-math: every expression below
 * is evaluated exactly as written (IEEE double, left to right).
 *
 * Where the reference asserts, the oracle counts (o_state.n_* diagnostics).
 *
 * Parity status of each function, and the checks that pin it, are listed in
 * tests/test_oracle.py; the pins are closed forms, invariants and an
 * independent textbook HLLD (Miyoshi & Kusano 2005) -- the reference itself
 * cannot be built here (Charm++/CGAL absent).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef SREAL
#define SREAL double
#endif

typedef double real;
typedef SREAL sreal;

#define NF 10
#define REF_HUGE 1e30 /* Boundary.h:5-7 (glibc's math.h does not define HUGE) */
#define REF_PI 3.14159265358979323846

enum { DENS = 0, VELX, VELY, VELZ, ETHM, BX, BY, BZ, PSI, ENTR };
enum { MASS = 0, MOMX, MOMY, MOMZ, ENER, WBX, WBY, WBZ, MPSI, MENTR };
enum { NO_BOUNDARY = 0, REFLECTING = 1, OUTFLOW = 2, DIOD = 3, INFLOW = 4 };
enum { LIM_SHIPPED = 0, LIM_MINMOD = 1, LIM_SMOOTH = 2 };

#define API __attribute__((visibility("default")))
/* diagnostic counters may be bumped from threads of the parallel global step */
#define DIAG_INC(x) \
  do {                \
    _Pragma("omp atomic") (x)++; \
  } while (0)

typedef struct {
  int s1, s2;
  double centroid[3];
  double n[3];
} o_face; /* fvmhd3d.h:262-284 */

typedef struct {
  int np, nactive, nimport_loc;
  double t_global, gamma_gas, courant_no;
  double L[3]; /* global_domain_size */
  int limiter_mode;
  int pad0;
  /* per cell, index order of ptcl_act: [0,nactive) active, then local inactive, then ghosts */
  double *w;        /* [np][10] Wrec.w */
  sreal *x, *y, *z; /* [np][10] Wrec.x, y, z */
  sreal *t;         /* [np][10] Wrec.t */
  double *pos;      /* [np][3]  Wrec.pos */
  sreal *vel;       /* [np][3]  Wrec.vel */
  sreal *J;         /* [np][3]  Wrec.J */
  double *tend;     /* [np]     Wrec.tend */
  int *bnd;         /* [np]     Wrec.bnd == mesh_act->boundary */
  int *is_ngb;      /* [np]     ptcl_act[i]->is_ngb() */
  int *is_active;   /* [np]     ptcl_act[i]->is_active() */
  sreal *mm_lo, *mm_hi; /* [np][10] Wrec_minmax */
  sreal *lim;       /* [np][10] slopeLimiter_act */
  sreal *dU;        /* [np][10] dU_act->U */
  sreal *dgp;       /* [np][3]  dU_act->gradPsi */
  sreal *ddivB;     /* [np]     dU_act->divB */
  double *U;        /* [np][10] U_act */
  double *mvol;     /* [np]     mesh_act->Volume */
  double *tbeg;     /* [np]     mesh_act->tbeg */
  double *mvel;     /* [np][3]  mesh_act->vel */
  double *acc0, *acc1; /* [np][3] mesh_act->acc0/acc1 */
  double *dt_new;   /* [np]     mesh_act->dt_new */
  double *divBi;    /* [np]     divBi output of computePrimitives */
  /* mesh */
  int nface;
  int nface_act;
  const o_face *faces;
  const int *cell_off, *cell_face; /* cell_list[i].ngb for i < nactive */
  const double *cvol;              /* cell_list[i].Volume */
  int *face_act;                   /* face_active_list (face indices) */
  /* diagnostics: how often a reference assert would have fired */
  long long n_pos_viol, n_closure_viol, n_ethm_failed, n_mass_failed, n_dsl_viol;
} o_state;

static inline real sqr(real x) { return x * x; }
/* std::min / std::max argument conventions */
static inline real smin(real a, real b) { return (b < a) ? b : a; }
static inline real smax(real a, real b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------ */
/* EOS hooks, Problem.cpp:234-254                                       */
/* ------------------------------------------------------------------ */
static inline real P_pressure(const o_state *s, const real *W) {
  return s->gamma_gas > 1.0 ? (s->gamma_gas - 1.0) * W[ETHM] : W[ETHM];
}
static inline real P_entropy_from_ethm(const o_state *s, const real *W) {
  return (s->gamma_gas - 1.0) * W[ETHM] / pow(W[DENS], s->gamma_gas);
}
static inline real P_ethm_from_entropy(const o_state *s, const real *W) {
  return W[ENTR] * pow(W[DENS], s->gamma_gas) / (s->gamma_gas - 1.0);
}

/* Fluid::to_primitive, Fluid.h:61-98 (total-energy form, __ENER_UB__ undefined) */
static void to_primitive(const real *U, real volume, real *W) {
  const real iM = 1.0 / U[MASS];
  const real iV = 1.0 / volume;
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

/* Fluid::to_conservative, Fluid.h:100-135 */
static void to_conservative(const real *W, real volume, real *U) {
  const real V = volume;
  const real M = W[DENS] * volume;
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

/* ------------------------------------------------------------------ */
/* limiters                                                             */
/* ------------------------------------------------------------------ */
/* Phi: computeReconstruction.cpp:5-17 / slopeLimiter.cpp:5-18 return 1 as
 * shipped; the opt-in modes are the dead branch below it (min(1,x)) and the
 * MPI variant's smooth limiter (MPI/compute_reconstruction.cpp:5-15). */
static inline real Phi(int mode, real x) {
  if (mode == LIM_SHIPPED) return 1.0;
  if (mode == LIM_MINMOD) return smin(1.0, x);
  {
    const real fac = 0.2, eps1 = 1.0e-16;
    return (x * x + 2.0 * x + fac * eps1) / (x * x + x + 2.0 + fac * eps1);
  }
}

/* Phi itself, for the closed-form checks of each mode */
API double o_limiter_phi(int mode, double x) { return Phi(mode, x); }

/* limiter of computeReconstruction.cpp:19-40 (multiplies by 1/dWi) */
static inline real limiter_rec(int mode, real dWi, real dWj_min, real dWj_max) {
  if (fabs(dWi) <= 1.0e-16 * smax(fabs(dWj_max), fabs(dWj_min))) return 1.0;
  {
    const real invWi = 1.0 / dWi;
    if (dWi < 0.0) return Phi(mode, dWj_min * invWi);
    else return Phi(mode, dWj_max * invWi);
  }
}

/* limiter of slopeLimiter.cpp:20-25 (divides by dWi) */
static inline real limiter_sl(int mode, real dWi, real dWj_min, real dWj_max) {
  if (fabs(dWi) <= 1.0e-16 * smax(fabs(dWj_max), fabs(dWj_min))) return 1.0;
  else if (dWi < 0.0) return Phi(mode, dWj_min / dWi);
  else return Phi(mode, dWj_max / dWi);
}

/* ------------------------------------------------------------------ */
/* HLLD, computeFlux.cpp:178-425                                        */
/* ------------------------------------------------------------------ */
static void riemann_hlld(real gamma_gas, real *flux, const real Bx, const real w,
                         const real dens_L, const real pres_L, const real ethm_L,
                         const real velx_L, const real vely_L, const real velz_L,
                         const real By_L, const real Bz_L, const real dens_R,
                         const real pres_R, const real ethm_R, const real velx_R,
                         const real vely_R, const real velz_R, const real By_R,
                         const real Bz_R) {
  const real signBx = Bx == 0.0 ? 0.0 : (Bx > 0.0 ? +1.0 : -1.0);

  const real momx_L = dens_L * velx_L;
  const real momy_L = dens_L * vely_L;
  const real momz_L = dens_L * velz_L;

  const real momx_R = dens_R * velx_R;
  const real momy_R = dens_R * vely_R;
  const real momz_R = dens_R * velz_R;
  const real chalf = 0.5;

  const real B2_L = sqr(Bx) + sqr(By_L) + sqr(Bz_L);
  const real v2_L = sqr(velx_L) + sqr(vely_L) + sqr(velz_L);
  const real etot_L = ethm_L + chalf * (dens_L * v2_L + B2_L);
  const real B2_R = sqr(Bx) + sqr(By_R) + sqr(Bz_R);
  const real v2_R = sqr(velx_R) + sqr(vely_R) + sqr(velz_R);
  const real etot_R = ethm_R + chalf * (dens_R * v2_R + B2_R);
  const real gpl = gamma_gas * pres_L;
  const real gpr = gamma_gas * pres_R;
  const real gpbl = gpl + B2_L;
  const real gpbr = gpr + B2_R;

  const real cfl2 = gpbl / dens_L;
  const real cfr2 = gpbr / dens_R;
  const real cfmax = sqrt(smax(cfl2, cfr2));

  const real S_L = smin(velx_L, velx_R) - cfmax;
  const real S_R = smax(velx_L, velx_R) + cfmax;

  const real pT_L = pres_L + chalf * B2_L;
  const real pT_R = pres_R + chalf * B2_R;
  const real iSM = 1.0 / ((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);
  const real S_M = iSM * ((S_R - velx_R) * momx_R - (S_L - velx_L) * momx_L - pT_R + pT_L);

  const real ipTs = 1.0 / ((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);
  const real pT_s = ipTs * ((S_R - velx_R) * dens_R * pT_L - (S_L - velx_L) * dens_L * pT_R +
                            dens_L * dens_R * (S_R - velx_R) * (S_L - velx_L) * (velx_R - velx_L));
  const real velx_L_s = S_M;
  const real velx_L_ss = S_M;
  const real velx_R_s = S_M;
  const real velx_R_ss = S_M;
  const real B2x = Bx * Bx;
  const real iSLmSM = 1.0 / (S_L - S_M);
  const real iSRmSM = 1.0 / (S_R - S_M);
  const real dens_L_s = dens_L * (S_L - velx_L) * iSLmSM;
  const real dens_R_s = dens_R * (S_R - velx_R) * iSRmSM;
  const real divL = dens_L * (S_L - velx_L) * (S_L - S_M) - B2x;
  const real divR = dens_R * (S_R - velx_R) * (S_R - S_M) - B2x;
  const real idivL = (divL != 0.0) ? 1.0 / divL : 0.0;
  const real idivR = (divR != 0.0) ? 1.0 / divR : 0.0;
  const real vely_L_s = vely_L - Bx * By_L * (S_M - velx_L) * idivL;
  const real velz_L_s = velz_L - Bx * Bz_L * (S_M - velx_L) * idivL;
  const real By_L_s = By_L * (dens_L * sqr(S_L - velx_L) - B2x) * idivL;
  const real Bz_L_s = Bz_L * (dens_L * sqr(S_L - velx_L) - B2x) * idivL;
  const real vely_R_s = vely_R - Bx * By_R * (S_M - velx_R) * idivR;
  const real velz_R_s = velz_R - Bx * Bz_R * (S_M - velx_R) * idivR;
  const real By_R_s = By_R * (dens_R * sqr(S_R - velx_R) - B2x) * idivR;
  const real Bz_R_s = Bz_R * (dens_R * sqr(S_R - velx_R) - B2x) * idivR;
  const real vB_L = velx_L * Bx + vely_L * By_L + velz_L * Bz_L;
  const real vB_L_s = velx_L_s * Bx + vely_L_s * By_L_s + velz_L_s * Bz_L_s;
  const real etot_L_s =
      ((S_L - velx_L) * etot_L - pT_L * velx_L + pT_s * S_M + Bx * (vB_L - vB_L_s)) * iSLmSM;
  const real vB_R = velx_R * Bx + vely_R * By_R + velz_R * Bz_R;
  const real vB_R_s = velx_R_s * Bx + vely_R_s * By_R_s + velz_R_s * Bz_R_s;
  const real etot_R_s =
      ((S_R - velx_R) * etot_R - pT_R * velx_R + pT_s * S_M + Bx * (vB_R - vB_R_s)) * iSRmSM;
  const real dens_L_ss = dens_L_s;
  const real dens_R_ss = dens_R_s;
  const real sDens_L_s = sqrt(dens_L_s);
  const real sDens_R_s = sqrt(dens_R_s);
  const real S_L_s = S_M - fabs(Bx / sDens_L_s);
  const real S_R_s = S_M + fabs(Bx / sDens_R_s);
  const real idsqroot = 1.0 / (sDens_L_s + sDens_R_s);
  const real vely_ss =
      idsqroot * (sDens_L_s * vely_L_s + sDens_R_s * vely_R_s + (By_R_s - By_L_s) * signBx);
  const real velz_ss =
      idsqroot * (sDens_L_s * velz_L_s + sDens_R_s * velz_R_s + (Bz_R_s - Bz_L_s) * signBx);
  const real By_ss = idsqroot * (sDens_L_s * By_R_s + sDens_R_s * By_L_s +
                                 sDens_L_s * sDens_R_s * (vely_R_s - vely_L_s) * signBx);
  const real Bz_ss = idsqroot * (sDens_L_s * Bz_R_s + sDens_R_s * Bz_L_s +
                                 sDens_L_s * sDens_R_s * (velz_R_s - velz_L_s) * signBx);
  const real vely_L_ss = vely_ss, velz_L_ss = velz_ss, By_L_ss = By_ss, Bz_L_ss = Bz_ss;
  const real vely_R_ss = vely_ss, velz_R_ss = velz_ss, By_R_ss = By_ss, Bz_R_ss = Bz_ss;

  const real vB_L_ss = velx_L_ss * Bx + vely_L_ss * By_L_ss + velz_L_ss * Bz_L_ss;
  const real etot_L_ss = etot_L_s - sDens_L_s * (vB_L_s - vB_L_ss) * signBx;
  const real vB_R_ss = velx_R_ss * Bx + vely_R_ss * By_R_ss + velz_R_ss * Bz_R_ss;
  const real etot_R_ss = etot_R_s + sDens_R_s * (vB_R_s - vB_R_ss) * signBx;
  const real Fdens_L = dens_L * velx_L;
  const real Fmomx_L = momx_L * velx_L + pT_L - B2x;
  const real Fmomy_L = momy_L * velx_L - Bx * By_L;
  const real Fmomz_L = momz_L * velx_L - Bx * Bz_L;
  const real Fetot_L = etot_L * velx_L + pT_L * velx_L - Bx * vB_L;
  const real Fdens_R = dens_R * velx_R;
  const real Fmomx_R = momx_R * velx_R + pT_R - B2x;
  const real Fmomy_R = momy_R * velx_R - Bx * By_R;
  const real Fmomz_R = momz_R * velx_R - Bx * Bz_R;
  const real Fetot_R = etot_R * velx_R + pT_R * velx_R - Bx * vB_R;
  const real momx_L_s = dens_L_s * velx_L_s;
  const real momy_L_s = dens_L_s * vely_L_s;
  const real momz_L_s = dens_L_s * velz_L_s;
  const real momx_L_ss = dens_L_ss * velx_L_ss;
  const real momy_L_ss = dens_L_ss * vely_L_ss;
  const real momz_L_ss = dens_L_ss * velz_L_ss;
  const real momx_R_s = dens_R_s * velx_R_s;
  const real momy_R_s = dens_R_s * vely_R_s;
  const real momz_R_s = dens_R_s * velz_R_s;
  const real momx_R_ss = dens_R_ss * velx_R_ss;
  const real momy_R_ss = dens_R_ss * vely_R_ss;
  const real momz_R_ss = dens_R_ss * velz_R_ss;
  const real Fby_L = By_L * velx_L - Bx * vely_L;
  const real Fbz_L = Bz_L * velx_L - Bx * velz_L;
  const real Fby_R = By_R * velx_R - Bx * vely_R;
  const real Fbz_R = Bz_R * velx_R - Bx * velz_R;
  const int w_lt_SL = w <= S_L;
  const int w_le_SLs = w <= S_L_s;
  const int w_le_SM = w <= S_M;
  const int w_le_SRs = w <= S_R_s;
  const int w_le_SR = w <= S_R;
  const real fdens = (w_le_SM) ? Fdens_L : Fdens_R;
  const real fetot = (w_le_SM) ? Fetot_L : Fetot_R;
  const real fmomx = (w_le_SM) ? Fmomx_L : Fmomx_R;
  const real fmomy = (w_le_SM) ? Fmomy_L : Fmomy_R;
  const real fmomz = (w_le_SM) ? Fmomz_L : Fmomz_R;
  const real fby = (w_le_SM) ? Fby_L : Fby_R;
  const real fbz = (w_le_SM) ? Fbz_L : Fbz_R;
  const real cnull = 0.0;
  const real a = (w_le_SLs ? cnull : (w_le_SM ? S_L_s - w : (w_le_SRs ? S_R_s - w : cnull)));
  const real b =
      (w_lt_SL ? cnull
               : (w_le_SLs ? +S_L - w
                           : (w_le_SM ? -S_L_s + S_L
                                      : (w_le_SRs ? -S_R_s + S_R : (w_le_SR ? +S_R - w : cnull)))));
  const real c =
      (w_lt_SL ? -w : (w_le_SLs ? -S_L : (w_le_SM ? -S_L : (w_le_SRs ? -S_R : (w_le_SR ? -S_R : -w)))));
  const real dens = (w_le_SM ? dens_L : dens_R);
  const real dens_s = (w_le_SM ? dens_L_s : dens_R_s);
  const real dens_ss = (w_le_SM ? dens_L_ss : dens_R_ss);
  const real etot = (w_le_SM ? etot_L : etot_R);
  const real etot_s = (w_le_SM ? etot_L_s : etot_R_s);
  const real etot_ss = (w_le_SM ? etot_L_ss : etot_R_ss);
  const real momx = (w_le_SM ? momx_L : momx_R);
  const real momx_s = (w_le_SM ? momx_L_s : momx_R_s);
  const real momx_ss = (w_le_SM ? momx_L_ss : momx_R_ss);
  const real momy = (w_le_SM ? momy_L : momy_R);
  const real momy_s = (w_le_SM ? momy_L_s : momy_R_s);
  const real momy_ss = (w_le_SM ? momy_L_ss : momy_R_ss);
  const real momz = (w_le_SM ? momz_L : momz_R);
  const real momz_s = (w_le_SM ? momz_L_s : momz_R_s);
  const real momz_ss = (w_le_SM ? momz_L_ss : momz_R_ss);
  const real by = (w_le_SM ? By_L : By_R);
  const real by_s = (w_le_SM ? By_L_s : By_R_s);
  const real by_ss = (w_le_SM ? By_L_ss : By_R_ss);
  const real bz = (w_le_SM ? Bz_L : Bz_R);
  const real bz_s = (w_le_SM ? Bz_L_s : Bz_R_s);
  const real bz_ss = (w_le_SM ? Bz_L_ss : Bz_R_ss);

  flux[MASS] = fdens + a * dens_ss + b * dens_s + c * dens;
  flux[ENER] = fetot + a * etot_ss + b * etot_s + c * etot;
  flux[MOMX] = fmomx + a * momx_ss + b * momx_s + c * momx;
  flux[MOMY] = fmomy + a * momy_ss + b * momy_s + c * momy;
  flux[MOMZ] = fmomz + a * momz_ss + b * momz_s + c * momz;
  flux[WBY] = fby + a * by_ss + b * by_s + c * by;
  flux[WBZ] = fbz + a * bz_ss + b * bz_s + c * bz;
  flux[WBX] = -w * Bx;
}

/* computeFlux.cpp:6-148 */
static void compute_flux(const o_state *s, const real *wij, const real *Wi, const real *Wj,
                         const int jbnd, const real *normal, real *psi_ij, real *Bn_ij,
                         real *flux) {
  const real gamma_gas = s->gamma_gas;
  const real ex = normal[0];
  const real ey = normal[1];
  const real ez = normal[2];
  const real ds = sqrt(ex * ex + ey * ey);
  const real ids = (ds != 0.0) ? 1.0 / ds : 0.0;
  const real cosph = (ids == 0.0) ? 1.0 : ex * ids;
  const real sinph = (ids == 0.0) ? 0.0 : ey * ids;
  const real costh = ez;
  const real sinth = ds;
  const real Axx = cosph * sinth;
  const real Axy = sinth * sinph;
  const real Axz = costh;
  const real Ayx = -sinph;
  const real Ayy = cosph;
  const real Ayz = 0.0;
  const real Azx = -costh * cosph;
  const real Azy = -costh * sinph;
  const real Azz = sinth;
  real dens_L = Wi[DENS];
  real ethm_L = Wi[ETHM];
  real velx_L = Axx * Wi[VELX] + Axy * Wi[VELY] + Axz * Wi[VELZ];
  real vely_L = Ayx * Wi[VELX] + Ayy * Wi[VELY] + Ayz * Wi[VELZ];
  real velz_L = Azx * Wi[VELX] + Azy * Wi[VELY] + Azz * Wi[VELZ];
  real Bx_L = Axx * Wi[BX] + Axy * Wi[BY] + Axz * Wi[BZ];
  real By_L = Ayx * Wi[BX] + Ayy * Wi[BY] + Ayz * Wi[BZ];
  real Bz_L = Azx * Wi[BX] + Azy * Wi[BY] + Azz * Wi[BZ];
  real psiL = Wi[PSI];
  real entrL = Wi[ENTR];
  real dens_R = Wj[DENS];
  real ethm_R = Wj[ETHM];
  real velx_R = Axx * Wj[VELX] + Axy * Wj[VELY] + Axz * Wj[VELZ];
  real vely_R = Ayx * Wj[VELX] + Ayy * Wj[VELY] + Ayz * Wj[VELZ];
  real velz_R = Azx * Wj[VELX] + Azy * Wj[VELY] + Azz * Wj[VELZ];
  real Bx_R = Axx * Wj[BX] + Axy * Wj[BY] + Axz * Wj[BZ];
  real By_R = Ayx * Wj[BX] + Ayy * Wj[BY] + Ayz * Wj[BZ];
  real Bz_R = Azx * Wj[BX] + Azy * Wj[BY] + Azz * Wj[BZ];
  real psiR = Wj[PSI];
  real entrR = Wj[ENTR];

  real pres_L = P_pressure(s, Wi);
  real pres_R = P_pressure(s, Wj);

  const real wn_ij = Axx * wij[0] + Axy * wij[1] + Axz * wij[2];
  switch (jbnd) {
    case REFLECTING:
      velx_R = -velx_R;
      break;
    case DIOD:
      velx_R = smax(velx_L - 1.0 * wn_ij, 0.0) + 1.0 * wn_ij;
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
    case INFLOW:
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

  const real cfl2 = (gamma_gas * pres_L + sqr(Bx_L) + sqr(By_L) + sqr(Bz_L)) / dens_L;
  const real cfr2 = (gamma_gas * pres_R + sqr(Bx_R) + sqr(By_R) + sqr(Bz_R)) / dens_R;
  const real chL = sqrt(cfl2 + 0.0 * sqr(velx_L - wn_ij));
  const real chR = sqrt(cfr2 + 0.0 * sqr(velx_R - wn_ij));

  *Bn_ij = (Bx_L * chL + Bx_R * chR + (psiL - psiR)) / (chL + chR);
  *psi_ij = (psiL * chR + psiR * chL + chR * chL * (Bx_L - Bx_R)) / (chL + chR);

  real F[NF];
  F[MPSI] = F[MENTR] = 0.0;
  riemann_hlld(gamma_gas, F, *Bn_ij, wn_ij, dens_L, pres_L, ethm_L, velx_L, vely_L, velz_L, By_L,
               Bz_L, dens_R, pres_R, ethm_R, velx_R, vely_R, velz_R, By_R, Bz_R);

  {
    const real iAxx = cosph * sinth;
    const real iAxy = -sinph;
    const real iAxz = -costh * cosph;
    const real iAyx = sinth * sinph;
    const real iAyy = cosph;
    const real iAyz = -costh * sinph;
    const real iAzx = costh;
    const real iAzy = 0.0;
    const real iAzz = sinth;
    flux[MASS] = F[MASS];
    flux[ENER] = F[ENER];
    flux[MOMX] = iAxx * F[MOMX] + iAxy * F[MOMY] + iAxz * F[MOMZ];
    flux[MOMY] = iAyx * F[MOMX] + iAyy * F[MOMY] + iAyz * F[MOMZ];
    flux[MOMZ] = iAzx * F[MOMX] + iAzy * F[MOMY] + iAzz * F[MOMZ];
    flux[WBX] = iAxx * F[WBX] + iAxy * F[WBY] + iAxz * F[WBZ];
    flux[WBY] = iAyx * F[WBX] + iAyy * F[WBY] + iAyz * F[WBZ];
    flux[WBZ] = iAzx * F[WBX] + iAzy * F[WBY] + iAzz * F[WBZ];
    flux[MPSI] = F[MASS] * (F[MASS] > 0.0 ? psiL : psiR);
    flux[MENTR] = F[MASS] * (F[MASS] > 0.0 ? entrL : entrR);
  }
}

/* ------------------------------------------------------------------ */
/* get_active_faces, Iterate.cpp:298-324                                */
/* ------------------------------------------------------------------ */
API int o_get_active_faces(o_state *s, int *face_act_out) {
  /* the sign-flip marks of the reference, on a private copy of s1 */
  int *mark = (int *)calloc((size_t)s->nface + 1, sizeof(int));
  int n = 0;
  for (int i = 0; i < s->nactive; i++) {
    if (s->bnd[i] != NO_BOUNDARY) continue;
    for (int k = s->cell_off[i]; k < s->cell_off[i + 1]; k++) {
      const int f = s->cell_face[k];
      if (mark[f]) continue;
      face_act_out[n++] = f;
      mark[f] = 1;
    }
  }
  free(mark);
  return n;
}

/* ------------------------------------------------------------------ */
/* computeReconstruction.cpp:42-264                                     */
/* ------------------------------------------------------------------ */
/* per-face scratch of one reconstruction (grown on demand) */
typedef struct {
  real *centroid_list, *dtij_list;
  int cap;
} o_rec_buf;

static void rec_buf_init(o_rec_buf *b) {
  b->cap = 64;
  b->centroid_list = (real *)malloc(sizeof(real) * 3 * b->cap);
  b->dtij_list = (real *)malloc(sizeof(real) * b->cap);
}

static void rec_buf_free(o_rec_buf *b) {
  free(b->centroid_list);
  free(b->dtij_list);
}

/* the body of the cell loop of computeReconstruction.cpp:57-262 for cell i */
static void reconstruction_cell(o_state *s, int i, o_rec_buf *buf) {
  const double *L = s->L;
  {
    real Wi[NF];
    for (int k = 0; k < NF; k++) Wi[k] = s->w[i * NF + k];
    const real *ipos = s->pos + 3 * i;

    if (s->bnd[i] != NO_BOUNDARY) {
      for (int k = 0; k < NF; k++) {
        s->w[i * NF + k] = Wi[k];
        s->x[i * NF + k] = 0.0;
        s->y[i * NF + k] = 0.0;
        s->z[i * NF + k] = 0.0;
        s->t[i * NF + k] = 0.0;
      }
      return;
    }

    real Wx[NF], Wy[NF], Wz[NF];
    sreal dWj_min[NF], dWj_max[NF];
    for (int k = 0; k < NF; k++) Wx[k] = Wy[k] = Wz[k] = 0.0, dWj_min[k] = dWj_max[k] = 0.0;

    const int fb = s->cell_off[i], nface = s->cell_off[i + 1] - fb;
    if (nface > buf->cap) {
      buf->cap = 2 * nface;
      buf->centroid_list = (real *)realloc(buf->centroid_list, sizeof(real) * 3 * buf->cap);
      buf->dtij_list = (real *)realloc(buf->dtij_list, sizeof(real) * buf->cap);
    }
    real *centroid_list = buf->centroid_list, *dtij_list = buf->dtij_list;
    int zero_flag = 0;
    const real dtI = s->tend[i] - s->t_global;

    int nlist = 0;
    for (int iface = 0; iface < nface; iface++) {
      const o_face *face = &s->faces[s->cell_face[fb + iface]];
      real dri[3];
      for (int d = 0; d < 3; d++) {
        dri[d] = face->centroid[d] - ipos[d];
        if (dri[d] > 0.5 * L[d]) dri[d] -= L[d];
        else if (dri[d] < -0.5 * L[d]) dri[d] += L[d];
      }
      const real *centroid = dri;
      centroid_list[3 * iface + 0] = centroid[0];
      centroid_list[3 * iface + 1] = centroid[1];
      centroid_list[3 * iface + 2] = centroid[2];

      const real *fn = face->n;
      const real area = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
      const real cn = centroid[0] * fn[0] + centroid[1] * fn[1] + centroid[2] * fn[2];
      const real sc = (cn < 0.0) ? (-1.0 / area) : (1.0 / area);
      const real normal[3] = {fn[0] * sc, fn[1] * sc, fn[2] * sc};
      const real dsh = centroid[0] * normal[0] + centroid[1] * normal[1] + centroid[2] * normal[2];

      const real ids = (dsh != 0.0) ? 0.5 / dsh : 0.0;
      const real idsA = area * ids;

      const real drh[3] = {normal[0] * dsh, normal[1] * dsh, normal[2] * dsh};
      const real fij[3] = {centroid[0] - drh[0], centroid[1] - drh[1], centroid[2] - drh[2]};

      const int j = (i == face->s1) ? face->s2 : face->s1;

      if (s->bnd[j] != NO_BOUNDARY && s->bnd[j] != INFLOW) {
        zero_flag = 1;
        break;
      }

      const real dtJ = s->tend[j] - s->t_global;
      const real dth = 0.5 * smin(dtI, dtJ);
      dtij_list[nlist++] = dtI - dth;
      const real *Wj = s->w + (size_t)j * NF;
      for (int k = 0; k < NF; k++) {
        const real sum = Wj[k] + Wi[k];
        const real diff = Wj[k] - Wi[k];
        dWj_min[k] = (sreal)smin((real)dWj_min[k], diff);
        dWj_max[k] = (sreal)smax((real)dWj_max[k], diff);
        Wx[k] += (sum * drh[0] + diff * fij[0]) * idsA;
        Wy[k] += (sum * drh[1] + diff * fij[1]) * idsA;
        Wz[k] += (sum * drh[2] + diff * fij[2]) * idsA;
      }
    }

    if (zero_flag) {
      for (int k = 0; k < NF; k++) {
        s->w[i * NF + k] = Wi[k];
        s->x[i * NF + k] = 0.0;
        s->y[i * NF + k] = 0.0;
        s->z[i * NF + k] = 0.0;
        s->t[i * NF + k] = 0.0;
      }
      return;
    }

    {
      const real invV = 1.0 / s->cvol[i];
      for (int k = 0; k < NF; k++) {
        Wx[k] *= invV;
        Wy[k] *= invV;
        Wz[k] *= invV;
      }
    }

    /* J = curl B (computeReconstruction.cpp:171-177) */
    s->J[3 * i + 0] = (sreal)(Wy[BZ] - Wz[BY]);
    s->J[3 * i + 1] = (sreal)(Wz[BX] - Wx[BZ]);
    s->J[3 * i + 2] = (sreal)(Wx[BY] - Wy[BX]);

    real limiter_min[NF];
    for (int k = 0; k < NF; k++) limiter_min[k] = 1.0;
    for (int iface = 0; iface < nface; iface++)
      for (int k = 0; k < NF; k++) {
        const real *c = centroid_list + 3 * iface;
        const real dWx = Wx[k] * c[0] + Wy[k] * c[1] + Wz[k] * c[2];
        const real dWt = 0.0 * (real)s->t[i * NF + k] * dtij_list[iface];
        const real dWi = dWx + dWt;
        limiter_min[k] =
            smin(limiter_min[k], limiter_rec(s->limiter_mode, dWi, dWj_min[k], dWj_max[k]));
      }

    for (int k = 0; k < NF; k++) {
      s->mm_lo[i * NF + k] = dWj_min[k];
      s->mm_hi[i * NF + k] = dWj_max[k];
    }

    {
      const real f0 = 1.0; /* Problem_enforce_limiter, Problem.cpp:251-254 */
      for (int k = 0; k < NF; k++) {
        s->mm_lo[i * NF + k] = (sreal)((real)s->mm_lo[i * NF + k] + Wi[k]);
        s->mm_hi[i * NF + k] = (sreal)((real)s->mm_hi[i * NF + k] + Wi[k]);
        const real tau = f0 * limiter_min[k];
        Wx[k] *= tau;
        Wy[k] *= tau;
        Wz[k] *= tau;
        s->t[i * NF + k] = (sreal)((real)s->t[i * NF + k] * tau);
        s->w[i * NF + k] = Wi[k];
        s->x[i * NF + k] = (sreal)Wx[k];
        s->y[i * NF + k] = (sreal)Wy[k];
        s->z[i * NF + k] = (sreal)Wz[k];
      }
    }
  }
}

API void o_compute_reconstruction(o_state *s) {
  o_rec_buf buf;
  rec_buf_init(&buf);
  for (int i = 0; i < s->nactive; i++) reconstruction_cell(s, i, &buf);
  rec_buf_free(&buf);
}

/* ------------------------------------------------------------------ */
/* slopeLimiter.cpp:27-237 (compute) and computeFluidUpdateI            */
/* (computeFluidUpdate.cpp:25-75, apply)                                */
/* ------------------------------------------------------------------ */
/* one active face of slopeLimiter.cpp:40-216: the limiter minimum and the
 * area-closure sums of the sides in `sides` (bit 0: s1, bit 1: s2); the
 * dsl check is counted with the s1 side */
static void slope_face(o_state *s, const o_face *face, int sides, real *area_test) {
  const int i = face->s1;
  const int j = face->s2;
  const real dtI = s->tend[i] - s->t_global;
  const real dtJ = s->tend[j] - s->t_global;
  const real dth = 0.5 * smin(dtI, dtJ);

  const real *ip = s->pos + 3 * i;
  const real dri[3] = {face->centroid[0] - ip[0], face->centroid[1] - ip[1],
                       face->centroid[2] - ip[2]};
  const real *fn = face->n;
  const real area = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const real ia = 1.0 / area;
  const real normal[3] = {fn[0] * ia, fn[1] * ia, fn[2] * ia};
  const real dsh = dri[0] * normal[0] + dri[1] * normal[1] + dri[2] * normal[2];
  const real dsl = 2.0 * dsh;
  if (!(dsl > 0.0) && (sides & 1)) DIAG_INC(s->n_dsl_viol);
  const real drj[3] = {dri[0] - normal[0] * dsl, dri[1] - normal[1] * dsl,
                       dri[2] - normal[2] * dsl};

  for (int d = 0; d < 3; d++) {
    const real an = normal[d] * area;
    if (sides & 1) area_test[3 * i + d] = area_test[3 * i + d] + an;
    if (sides & 2) area_test[3 * j + d] = area_test[3 * j + d] - an;
  }

  if (s->bnd[i] != NO_BOUNDARY || s->bnd[j] != NO_BOUNDARY) {
    for (int k = 0; k < NF; k++) {
      if (sides & 1) s->lim[i * NF + k] = 0;
      if (sides & 2) s->lim[j * NF + k] = 0;
    }
    return;
  }

  for (int k = 0; k < NF; k++) {
    const real dWit = (real)s->t[i * NF + k] * (dtI - dth);
    const real dWjt = (real)s->t[j * NF + k] * (dtJ - dth);
    if (sides & 1) {
      const real dWix = (real)s->x[i * NF + k] * dri[0] + (real)s->y[i * NF + k] * dri[1] +
                        (real)s->z[i * NF + k] * dri[2];
      const real dWi = dWix + dWit;
      s->lim[i * NF + k] = (sreal)smin(
          (real)s->lim[i * NF + k],
          limiter_sl(s->limiter_mode, dWi, smin((real)s->mm_lo[i * NF + k] - s->w[i * NF + k], 0.0),
                     smax((real)s->mm_hi[i * NF + k] - s->w[i * NF + k], 0.0)));
    }
    if (sides & 2) {
      const real dWjx = (real)s->x[j * NF + k] * drj[0] + (real)s->y[j * NF + k] * drj[1] +
                        (real)s->z[j * NF + k] * drj[2];
      const real dWj = dWjx + dWjt;
      s->lim[j * NF + k] = (sreal)smin(
          (real)s->lim[j * NF + k],
          limiter_sl(s->limiter_mode, dWj, smin((real)s->mm_lo[j * NF + k] - s->w[j * NF + k], 0.0),
                     smax((real)s->mm_hi[j * NF + k] - s->w[j * NF + k], 0.0)));
    }
  }
}

/* slopeLimiter.cpp:218-237: area closure of active cell i */
static void closure_cell(o_state *s, int i, const real *area_test) {
  if (s->bnd[i] != NO_BOUNDARY) return;
  const real *a = area_test + 3 * i;
  const real aabs = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (!(aabs < 1.0e-10 * pow(s->cvol[i], 2.0 / 3.0))) DIAG_INC(s->n_closure_viol);
}

/* computeFluidUpdateI (computeFluidUpdate.cpp:25-75) for cell i */
static void limiter_apply_cell(o_state *s, int i) {
  if (!(i < s->nactive || s->is_ngb[i])) return;
  const real f0 = 1.0;
  for (int k = 0; k < NF; k++) {
    const real f = f0 * (real)s->lim[i * NF + k];
    s->x[i * NF + k] = (sreal)((real)s->x[i * NF + k] * f);
    s->y[i * NF + k] = (sreal)((real)s->y[i * NF + k] * f);
    s->z[i * NF + k] = (sreal)((real)s->z[i * NF + k] * f);
    s->t[i * NF + k] = (sreal)((real)s->t[i * NF + k] * f);
  }
}

API void o_slope_limiter(o_state *s) {
  const int np = s->np;
  real *area_test = (real *)calloc((size_t)np * 3, sizeof(real));
  /* slopeLimiter_all / import limiters are 1 after localMesh_import (localMesh.cpp:57, 267) */
  for (int i = 0; i < np * NF; i++) s->lim[i] = 1.0;
  for (int fa = 0; fa < s->nface_act; fa++) slope_face(s, &s->faces[s->face_act[fa]], 3, area_test);
  for (int i = 0; i < s->nactive; i++) closure_cell(s, i, area_test);
  free(area_test);
  for (int i = 0; i < np; i++) limiter_apply_cell(s, i);
}

/* ------------------------------------------------------------------ */
/* computeFluidUpdate.cpp:88-247                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  sreal U[NF];
  sreal gradPsi[3];
  sreal divB;
} o_fluidD;

/* one active face of computeFluidUpdate.cpp:101-234: after the boundary swap
 * the face adds -ft, +dPsi, +bt to cell *pi and +ft, -dPsi, -bt to cell *pj,
 * with ft = flux*dt*A, dPsi = n psi dt A, bt = Bn dt A */
static void fluid_face(o_state *s, const o_face *face, int *pi, int *pj, real *ft, real *dPsi, real *bt) {
  int i = face->s1;
  int j = face->s2;

  const real *fn = face->n;
  const real area = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const real ia = 1.0 / area;
  real normal[3] = {fn[0] * ia, fn[1] * ia, fn[2] * ia};

  const real *ip = s->pos + 3 * i;
  real dri[3] = {face->centroid[0] - ip[0], face->centroid[1] - ip[1],
                 face->centroid[2] - ip[2]};

  if (s->bnd[i] != NO_BOUNDARY) {
    const int tmp = i;
    i = j;
    j = tmp;
    for (int d = 0; d < 3; d++) {
      normal[d] = normal[d] * -1.0;
      dri[d] = dri[d] * -1.0;
    }
  }

  const real dtI = s->tend[i] - s->t_global;
  const real dtJ = s->tend[j] - s->t_global;
  const real dt = smin(dtI, dtJ);
  const real dti = dtI - 0.5 * dt;
  const real dtj = dtJ - 0.5 * dt;

  const real dsh = dri[0] * normal[0] + dri[1] * normal[1] + dri[2] * normal[2];
  const real dsl = 2.0 * dsh;
  if (!(dsl > 0.0)) DIAG_INC(s->n_dsl_viol);

  const real drj[3] = {dri[0] - normal[0] * dsl, dri[1] - normal[1] * dsl,
                       dri[2] - normal[2] * dsl};

  real Wi[NF], Wj[NF];
  for (int k = 0; k < NF; k++) {
    Wi[k] = s->w[(size_t)i * NF + k];
    Wj[k] = s->w[(size_t)j * NF + k];
  }
  const int jbnd = s->bnd[j];
  if (jbnd == NO_BOUNDARY) {
    for (int k = 0; k < NF; k++) {
      const size_t ik = (size_t)i * NF + k, jk = (size_t)j * NF + k;
      Wi[k] += (real)s->x[ik] * dri[0] + (real)s->y[ik] * dri[1] + (real)s->z[ik] * dri[2] +
               (real)s->t[ik] * dti;
      Wj[k] += (real)s->x[jk] * drj[0] + (real)s->y[jk] * drj[1] + (real)s->z[jk] * drj[2] +
               (real)s->t[jk] * dtj;
    }
  }
  if (!(Wi[DENS] > 0.0) || !(Wi[ETHM] > 0.0) || !(Wj[DENS] > 0.0) || !(Wj[ETHM] > 0.0))
    DIAG_INC(s->n_pos_viol);

  const real vi[3] = {(real)s->vel[3 * i], (real)s->vel[3 * i + 1], (real)s->vel[3 * i + 2]};
  const real vj[3] = {(real)s->vel[3 * j], (real)s->vel[3 * j + 1], (real)s->vel[3 * j + 2]};
  const real vij[3] = {vj[0] + vi[0], vj[1] + vi[1], vj[2] + vi[2]};
  const real dvij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
  const real fij[3] = {dri[0] - normal[0] * dsh, dri[1] - normal[1] * dsh,
                       dri[2] - normal[2] * dsh};
  const real fdv = dvij[0] * fij[0] + dvij[1] * fij[1] + dvij[2] * fij[2];
  const real rji[3] = {normal[0] * dsl, normal[1] * dsl, normal[2] * dsl};
  const real r = 1.0 / sqr(dsl);
  const real dwij[3] = {rji[0] * fdv * r, rji[1] * fdv * r, rji[2] * fdv * r};
  const real wij[3] = {vij[0] * 0.5 - dwij[0], vij[1] * 0.5 - dwij[1], vij[2] * 0.5 - dwij[2]};

  real psi, Bn, flux[NF];
  compute_flux(s, wij, Wi, Wj, jbnd, normal, &psi, &Bn, flux);

  const real tarea = dt * area;
  dPsi[0] = normal[0] * (psi * tarea);
  dPsi[1] = normal[1] * (psi * tarea);
  dPsi[2] = normal[2] * (psi * tarea);
  for (int k = 0; k < NF; k++) ft[k] = flux[k] * tarea;
  *bt = Bn * tarea;
  *pi = i;
  *pj = j;
}

/* the FluidD accumulation of cell c for one face (computeFluidUpdate.cpp:221-234) */
static inline void fluid_acc(o_fluidD *acc, int minus, const real *ft, const real *dPsi, real bt) {
  if (minus) {
    for (int k = 0; k < NF; k++) acc->U[k] = (sreal)((real)acc->U[k] - ft[k]);
    for (int d = 0; d < 3; d++) acc->gradPsi[d] = acc->gradPsi[d] + (sreal)dPsi[d];
    acc->divB = (sreal)((real)acc->divB + bt);
  } else {
    for (int k = 0; k < NF; k++) acc->U[k] = (sreal)((real)acc->U[k] + ft[k]);
    for (int d = 0; d < 3; d++) acc->gradPsi[d] = acc->gradPsi[d] - (sreal)dPsi[d];
    acc->divB = (sreal)((real)acc->divB - bt);
  }
}

/* computeFluidUpdate.cpp:237-245: dU_act[i] += FluidD of cell i */
static inline void fluid_store(o_state *s, int i, const o_fluidD *acc) {
  s->ddivB[i] = s->ddivB[i] + acc->divB;
  for (int d = 0; d < 3; d++) s->dgp[3 * i + d] = s->dgp[3 * i + d] + acc->gradPsi[d];
  for (int k = 0; k < NF; k++) s->dU[i * NF + k] = s->dU[i * NF + k] + acc->U[k];
}

API void o_compute_fluid_update(o_state *s) {
  const int np = s->np;
  o_fluidD *acc = (o_fluidD *)calloc((size_t)np, sizeof(o_fluidD));
  for (int fa = 0; fa < s->nface_act; fa++) {
    int i, j;
    real ft[NF], dPsi[3], bt;
    fluid_face(s, &s->faces[s->face_act[fa]], &i, &j, ft, dPsi, &bt);
    fluid_acc(&acc[i], 1, ft, dPsi, bt);
    fluid_acc(&acc[j], 0, ft, dPsi, bt);
  }
  for (int i = 0; i < np; i++) fluid_store(s, i, &acc[i]);
  free(acc);
}

/* batched computeFlux for per-face parity */
API void o_compute_flux_batch(o_state *s, int n, const double *wij, const double *Wi,
                              const double *Wj, const int *jbnd, const double *normal,
                              double *psi, double *Bn, double *flux) {
  for (int f = 0; f < n; f++)
    compute_flux(s, wij + 3 * f, Wi + NF * f, Wj + NF * f, jbnd[f], normal + 3 * f, psi + f,
                 Bn + f, flux + NF * f);
}

/* ------------------------------------------------------------------ */
/* computePrimitives, computeFluidUpdate.cpp:347-542                    */
/* ------------------------------------------------------------------ */
/* the cell loop of computeFluidUpdate.cpp:356-502 for active cell i; returns
 * 1 when the mass is not positive (the cell is left for the fallback) */
static int primitives_cell(o_state *s, int i, int predictor_flag) {
  const real gamma_gas = s->gamma_gas;
  real *Ui = s->U + (size_t)i * NF;
  real *wi = s->w + (size_t)i * NF;
  if (s->bnd[i] != NO_BOUNDARY) {
    to_primitive(Ui, s->mvol[i], wi);
    for (int k = 0; k < NF; k++)
      s->x[i * NF + k] = s->y[i * NF + k] = s->z[i * NF + k] = s->t[i * NF + k] = 0.0;
    s->mvol[i] = s->cvol[i];
    to_conservative(wi, s->mvol[i], Ui);
    return 0;
  }

  real Uc[NF];
  for (int k = 0; k < NF; k++) Uc[k] = Ui[k] + (real)s->dU[i * NF + k];

  const real divB = (real)s->ddivB[i];
  const real dPsi[3] = {(real)s->dgp[3 * i], (real)s->dgp[3 * i + 1], (real)s->dgp[3 * i + 2]};

  const real dt = s->t_global - s->tbeg[i];
  s->divBi[i] = dt > 0.0 ? divB / dt / s->cvol[i] : divB;

  if (Uc[MASS] <= 0.0) {
    DIAG_INC(s->n_mass_failed);
    return 1;
  }

  {
    const real dth = 0.5 * dt;
    real Wh[NF];
    for (int k = 0; k < NF; k++) Wh[k] = wi[k] + (real)s->t[i * NF + k] * dth;

    Uc[WBX] -= Wh[VELX] * divB;
    Uc[WBY] -= Wh[VELY] * divB;
    Uc[WBZ] -= Wh[VELZ] * divB;

    /* DEDNER (computeFluidUpdate.cpp:3, 416-430) */
    const real pres = P_pressure(s, Wh);
    const real B2 = Wh[BX] * Wh[BX] + Wh[BY] * Wh[BY] + Wh[BZ] * Wh[BZ];
    const real dcs2 = gamma_gas * pres + B2;
    const real ch_sig = sqrt(dcs2 / Wh[DENS]);
    Uc[WBX] -= dPsi[0];
    Uc[WBY] -= dPsi[1];
    Uc[WBZ] -= dPsi[2];
    Uc[MPSI] -= dcs2 * divB;

    if (!predictor_flag) {
      const real *a0 = s->acc0 + 3 * i, *a1 = s->acc1 + 3 * i;
      const real M0 = Ui[MASS], M1 = Uc[MASS];
      real dmom[3];
      for (int d = 0; d < 3; d++) dmom[d] = ((a0[d] * M0 + a1[d] * M1) * 0.5) * dt;
      Uc[MOMX] += dmom[0];
      Uc[MOMY] += dmom[1];
      Uc[MOMZ] += dmom[2];
      const real mom0[3] = {Ui[MOMX], Ui[MOMY], Ui[MOMZ]};
      const real mom1[3] = {Uc[MOMX], Uc[MOMY], Uc[MOMZ]};
      Uc[ENER] += ((mom0[0] * a0[0] + mom0[1] * a0[1] + mom0[2] * a0[2]) +
                   (mom1[0] * a1[0] + mom1[1] * a1[1] + mom1[2] * a1[2])) *
                  dt * 0.5;
    }

    {
      const double cr = 0.5;
      const real l_min = pow(s->cvol[i] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
      Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
    }
  }

  real W1[NF];
  to_primitive(Uc, s->cvol[i], W1);
  /* Problem_compute_ethm_update returns W[ETHM] (Problem.cpp:228-232) */
  if (W1[ETHM] <= 0.0) {
    W1[ETHM] = P_ethm_from_entropy(s, W1);
    DIAG_INC(s->n_ethm_failed);
  } else {
    W1[ENTR] = P_entropy_from_ethm(s, W1);
  }

  if (predictor_flag) {
    if (dt > 0.0)
      for (int k = 0; k < NF; k++) s->t[i * NF + k] = (sreal)((W1[k] - wi[k]) / dt);
    else
      for (int k = 0; k < NF; k++) s->t[i * NF + k] = 0.0;
  }

  if (!predictor_flag) {
    s->mvol[i] = s->cvol[i];
    for (int k = 0; k < NF; k++) wi[k] = W1[k];
    to_conservative(W1, s->cvol[i], Ui);
    for (int k = 0; k < NF; k++) s->dU[i * NF + k] = 0.0;
    s->dgp[3 * i] = s->dgp[3 * i + 1] = s->dgp[3 * i + 2] = 0.0;
    s->ddivB[i] = 0.0;
  }
  return 0;
}

/* the failed-mass fallback of computeFluidUpdate.cpp:504-540 for cell i */
static void primitives_failed_cell(o_state *s, int i) {
  real *Ui = s->U + (size_t)i * NF;
  real *wi = s->w + (size_t)i * NF;
  const real dt = s->t_global - s->tbeg[i];
  real W1[NF], Ut[NF];
  for (int k = 0; k < NF; k++) W1[k] = wi[k] + (real)s->t[i * NF + k] * dt;
  to_conservative(W1, s->mvol[i], Ut);
  to_primitive(Ut, s->cvol[i], W1);
  s->mvol[i] = s->cvol[i];
  if (W1[ETHM] <= 0.0) {
    W1[ETHM] = P_ethm_from_entropy(s, W1);
  } else {
    W1[ENTR] = P_entropy_from_ethm(s, W1);
  }
  for (int k = 0; k < NF; k++) wi[k] = W1[k];
  to_conservative(W1, s->cvol[i], Ui);
  for (int k = 0; k < NF; k++) s->dU[i * NF + k] = 0.0;
  s->dgp[3 * i] = s->dgp[3 * i + 1] = s->dgp[3 * i + 2] = 0.0;
  s->ddivB[i] = 0.0;
}

API void o_compute_primitives(o_state *s, int predictor_flag) {
  int *failed = (int *)malloc(sizeof(int) * (size_t)(s->nactive + 1));
  int nfailed = 0;
  for (int i = 0; i < s->nactive; i++)
    if (primitives_cell(s, i, predictor_flag)) failed[nfailed++] = i;
  for (int f = 0; f < nfailed; f++) primitives_failed_cell(s, failed[f]);
  free(failed);
}

/* ------------------------------------------------------------------ */
/* computeTimestep.cpp:6-62                                             */
/* ------------------------------------------------------------------ */
/* the cell loop of computeTimestep.cpp:13-60 for active cell i */
static void timestep_cell(o_state *s, int i) {
  const real gamma_gas = s->gamma_gas;
  const real Volume = s->mvol[i];
  real Wi[NF];
  to_primitive(s->U + (size_t)i * NF, Volume, Wi);
  const real h = pow(Volume / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
  const real *pv = s->mvel + 3 * i;

  real csi = 0.0, vrel = 0.0;
  const int isb = s->bnd[i] != NO_BOUNDARY;
  if (!isb) {
    const real Pi = P_pressure(s, Wi) * gamma_gas;
    const real B2i = Wi[BX] * Wi[BX] + Wi[BY] * Wi[BY] + Wi[BZ] * Wi[BZ];
    csi = sqrt((Pi + B2i) / Wi[DENS]);
    const real dv[3] = {Wi[VELX] - pv[0], Wi[VELY] - pv[1], Wi[VELZ] - pv[2]};
    vrel = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  }

  real dti = h / (csi + vrel + 1.0 / REF_HUGE);
  const real fac = isb ? 0.0 : 1.0;
  const real vi[3] = {Wi[VELX], Wi[VELY], Wi[VELZ]};
  const real wi[3] = {pv[0], pv[1], pv[2]};
  for (int k = s->cell_off[i]; k < s->cell_off[i + 1]; k++) {
    const o_face *face = &s->faces[s->cell_face[k]];
    const int j = (i == face->s1) ? face->s2 : face->s1;
    const real *wj_ = s->w + (size_t)j * NF;
    const real vj[3] = {wj_[VELX], wj_[VELY], wj_[VELZ]};
    const real dvi[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
    const real wj[3] = {(real)s->vel[3 * j], (real)s->vel[3 * j + 1], (real)s->vel[3 * j + 2]};
    const real dwi[3] = {wj[0] - wi[0], wj[1] - wi[1], wj[2] - wi[2]};
    const real dvmax =
        sqrt((dwi[0] * dwi[0] + dwi[1] * dwi[1] + dwi[2] * dwi[2]) +
             fac * (dvi[0] * dvi[0] + dvi[1] * dvi[1] + dvi[2] * dvi[2]) + 1.0 / REF_HUGE);
    dti = smin(dti, h / (csi + dvmax));
  }
  /* Problem_extra_timestep_criterion returns HUGE (Problem.cpp:223-226) */
  dti = smin(dti, REF_HUGE);
  s->dt_new[i] = dti * s->courant_no;
}

API void o_compute_timestep(o_state *s) {
  for (int i = 0; i < s->nactive; i++) timestep_cell(s, i);
}

/* computeTimestep.cpp:64-110 */
API void o_compute_timestep_limiter(o_state *s) {
  const int np = s->np, nactive = s->nactive;
  real *dti_list = (real *)malloc(sizeof(real) * (size_t)np);
  for (int i = 0; i < np; i++) dti_list[i] = -REF_HUGE;
  for (int i = 0; i < nactive; i++) dti_list[i] = s->dt_new[i];
  for (int i = nactive; i < np; i++)
    if (s->is_ngb[i]) {
      if (s->is_active[i]) dti_list[i] = s->dt_new[i];
      else dti_list[i] = s->tend[i] - s->tbeg[i];
    }
  {
    const real scale_fac = 2.0;
    for (int i = 0; i < nactive; i++)
      for (int k = s->cell_off[i]; k < s->cell_off[i + 1]; k++) {
        const o_face *face = &s->faces[s->cell_face[k]];
        const int j = (i == face->s1) ? face->s2 : face->s1;
        dti_list[j] = smin(dti_list[j], scale_fac * dti_list[i]);
        dti_list[i] = smin(dti_list[i], scale_fac * dti_list[j]);
      }
  }
  for (int i = 0; i < np; i++)
    if (i < s->nactive || s->is_ngb[i]) s->dt_new[i] = dti_list[i];
  free(dti_list);
}

/* ------------------------------------------------------------------ */
/* computePredictor.cpp:11-391                                          */
/* The probe step scheduler.get_dt(scheduler.get_rung(dt_new))/128     */
/* (computePredictor.cpp:271) uses the power-of-two rungs of           */
/* Scheduler.h:37-42, 94-111 with the scheduler's dtmax and min_rung.  */
/* ------------------------------------------------------------------ */
static int exp_of(double dt) {
  unsigned long long u;
  memcpy(&u, &dt, sizeof u);
  return (int)((u >> 52) & 0xfff);
}
API double o_rung_dt(double dtmax, int min_rung, double dt) {
  int rung = exp_of(dtmax) - exp_of(dt);
  if (rung < min_rung) rung = min_rung;
  {
    const double dt_tick = dtmax / (double)(1ULL << 31);
    return dt_tick * (double)(1ULL << (31 - rung));
  }
}

/* one active face of computePredictor.cpp:47-140: after the boundary swap,
 * fl = flux*A, gp = n psi A, bA = Bn A and the face's volume rate dV, added
 * to cell *pi with - (and to *pj with +) */
static void predictor_face(o_state *s, const o_face *face, int *pi, int *pj, int *pjbnd, real *fl, real *gp,
                           real *bA, real *pdV) {
  int i = face->s1;
  int j = face->s2;
  const real *fn = face->n;
  const real area = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const real ia = 1.0 / area;
  real normal[3] = {fn[0] * ia, fn[1] * ia, fn[2] * ia};
  const real *ip = s->pos + 3 * i;
  real dri[3] = {face->centroid[0] - ip[0], face->centroid[1] - ip[1],
                 face->centroid[2] - ip[2]};
  if (s->bnd[i] != NO_BOUNDARY) {
    const int tmp = i;
    i = j;
    j = tmp;
    for (int d = 0; d < 3; d++) {
      normal[d] = normal[d] * -1.0;
      dri[d] = dri[d] * -1.0;
    }
  }
  const real dsh = dri[0] * normal[0] + dri[1] * normal[1] + dri[2] * normal[2];
  const real dsl = 2.0 * dsh;
  if (!(dsl > 0.0)) DIAG_INC(s->n_dsl_viol);

  const int jbnd = s->bnd[j];
  const real *Wi = s->w + (size_t)i * NF;
  const real *Wj = (jbnd == NO_BOUNDARY || jbnd == INFLOW) ? s->w + (size_t)j * NF : Wi;

  const real vi[3] = {(real)s->vel[3 * i], (real)s->vel[3 * i + 1], (real)s->vel[3 * i + 2]};
  const real vj[3] = {(real)s->vel[3 * j], (real)s->vel[3 * j + 1], (real)s->vel[3 * j + 2]};
  const real vij[3] = {vj[0] + vi[0], vj[1] + vi[1], vj[2] + vi[2]};
  const real dvij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
  const real fij[3] = {dri[0] - normal[0] * dsh, dri[1] - normal[1] * dsh,
                       dri[2] - normal[2] * dsh};
  const real fdv = dvij[0] * fij[0] + dvij[1] * fij[1] + dvij[2] * fij[2];
  const real rij[3] = {normal[0] * dsl, normal[1] * dsl, normal[2] * dsl};
  const real r = 1.0 / sqr(dsl);
  const real dwij[3] = {rij[0] * fdv * r, rij[1] * fdv * r, rij[2] * fdv * r};
  const real wij[3] = {vij[0] * 0.5 - dwij[0], vij[1] * 0.5 - dwij[1], vij[2] * 0.5 - dwij[2]};

  real psi, Bn, flux[NF];
  compute_flux(s, wij, Wi, Wj, jbnd, normal, &psi, &Bn, flux);

  for (int k = 0; k < NF; k++) fl[k] = flux[k] * area;
  for (int d = 0; d < 3; d++) gp[d] = normal[d] * (psi * area);
  *bA = Bn * area;
  *pdV = area * (rij[0] * wij[0] + rij[1] * wij[1] + rij[2] * wij[2]) / dsl;
  *pi = i;
  *pj = j;
  *pjbnd = jbnd;
}

/* the dUdt/dVdt accumulation of one face into cell c (computePredictor.cpp:120-140) */
static inline void predictor_acc(o_fluidD *d, real *dv, int minus, const real *fl, const real *gp, real bA, real dV) {
  if (minus) {
    for (int k = 0; k < NF; k++) d->U[k] = (sreal)((real)d->U[k] - fl[k]);
    for (int q = 0; q < 3; q++) d->gradPsi[q] = d->gradPsi[q] + (sreal)gp[q];
    d->divB = (sreal)((real)d->divB + bA);
    *dv += dV;
  } else {
    *dv -= dV;
    for (int k = 0; k < NF; k++) d->U[k] = (sreal)((real)d->U[k] + fl[k]);
    for (int q = 0; q < 3; q++) d->gradPsi[q] = d->gradPsi[q] - (sreal)gp[q];
    d->divB = (sreal)((real)d->divB - bA);
  }
}

/* computePredictor.cpp:142-389 for active cell i, from its face sums */
static void predictor_cell(o_state *s, int i, double dtmax, int min_rung, const o_fluidD *dUdti, real dVdti) {
  const real gamma_gas = s->gamma_gas;
  if (s->bnd[i] != NO_BOUNDARY) return;
  const sreal *dotUi = dUdti->U;
  const real divB = (real)dUdti->divB; /* a float as shipped (computePredictor.cpp:147) */
  const real dPsi[3] = {(real)dUdti->gradPsi[0], (real)dUdti->gradPsi[1],
                        (real)dUdti->gradPsi[2]};
  const real *U0 = s->U + (size_t)i * NF;
  real W0[NF];
  to_primitive(U0, s->cvol[i], W0);

  /* the dt-halving loop of computePredictor.cpp:156-164 on mass */
  while (U0[MASS] + (real)dotUi[MASS] * s->dt_new[i] <= 0.0) s->dt_new[i] *= 0.5;

  const real dt = o_rung_dt(dtmax, min_rung, s->dt_new[i]) / 128;
  real Uc[NF];
  for (int k = 0; k < NF; k++) Uc[k] = U0[k] + (real)dotUi[k] * dt;

  Uc[WBX] -= W0[VELX] * divB * dt;
  Uc[WBY] -= W0[VELY] * divB * dt;
  Uc[WBZ] -= W0[VELZ] * divB * dt;
  {
    const real pres = P_pressure(s, W0);
    const real B2 = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
    const real dcs2 = gamma_gas * pres + B2;
    const real ch_sig = sqrt(dcs2 / W0[DENS]);
    Uc[WBX] -= dPsi[0] * dt;
    Uc[WBY] -= dPsi[1] * dt;
    Uc[WBZ] -= dPsi[2] * dt;
    Uc[MPSI] -= dcs2 * divB * dt;
    const double cr = 0.5;
    const real l_min = pow(s->cvol[i] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
    Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
  }

  real W1[NF];
  to_primitive(Uc, s->cvol[i] * exp(dVdti * dt / s->cvol[i]), W1);
  if (W1[ETHM] <= 0.0) {
    W1[ETHM] = P_ethm_from_entropy(s, W1);
    DIAG_INC(s->n_ethm_failed);
  } else {
    W1[ENTR] = P_entropy_from_ethm(s, W1);
  }
  {
    const real idt = 1.0 / dt;
    for (int k = 0; k < NF; k++) s->t[i * NF + k] = (sreal)((W1[k] - W0[k]) * idt);
  }
}

API void o_compute_predictor(o_state *s, double dtmax, int min_rung) {
  const int nactive = s->nactive;
  o_fluidD *dUdt = (o_fluidD *)calloc((size_t)nactive + 1, sizeof(o_fluidD));
  real *dVdt = (real *)calloc((size_t)nactive + 1, sizeof(real));
  for (int fa = 0; fa < s->nface_act; fa++) {
    int i, j, jbnd;
    real fl[NF], gp[3], bA, dV;
    predictor_face(s, &s->faces[s->face_act[fa]], &i, &j, &jbnd, fl, gp, &bA, &dV);
    /* dUdt lives only for [0, nactive) (computePredictor.cpp:40-41) */
    if (i < nactive) predictor_acc(&dUdt[i], &dVdt[i], 1, fl, gp, bA, dV);
    if (j >= nactive || jbnd != NO_BOUNDARY) continue;
    predictor_acc(&dUdt[j], &dVdt[j], 0, fl, gp, bA, dV);
  }
  for (int i = 0; i < nactive; i++) predictor_cell(s, i, dtmax, min_rung, &dUdt[i], dVdt[i]);
  free(dUdt);
  free(dVdt);
}

API int o_sizeof_sreal(void) { return (int)sizeof(sreal); }
API int o_sizeof_state(void) { return (int)sizeof(o_state); }

/* ------------------------------------------------------------------ */
/* Global-timestep driver: one step of Iterate.cpp:92-202 with every    */
/* cell on the smallest rung.  CFL (computeTimestep) -> predictor       */
/* (probe from each cell's own dt_new) -> dt = rung of the global       */
/* minimum dt_new (Scheduler.h:94-111, System.cpp:258-266) -> tend ->   */
/* reconstruction -> limiter -> flux scatter -> computePrimitives at    */
/* t + dt.  face_act must hold the active face list.  Returns dt.      */
/* ------------------------------------------------------------------ */
API double o_step_global(o_state *s, double dtmax) {
  o_compute_timestep(s);
  o_compute_predictor(s, dtmax, 0);
  double dtmin = REF_HUGE;
  for (int i = 0; i < s->nactive; i++) dtmin = smin(dtmin, s->dt_new[i]);
  {
    const double dt = o_rung_dt(dtmax, 0, dtmin);
    const double t0 = s->t_global;
    for (int i = 0; i < s->np; i++) {
      s->tbeg[i] = t0;
      s->tend[i] = t0 + dt;
    }
    o_compute_reconstruction(s);
    o_slope_limiter(s);
    o_compute_fluid_update(s);
    s->t_global = t0 + dt;
    o_compute_primitives(s, 0);
    return dt;
  }
}

/* ------------------------------------------------------------------ */
/* The same global step on every host core (OpenMP; the CPU baseline).  */
/* The reference runs one chare per core (fvmhd3d.cpp:56-73); here the  */
/* cell loops run in parallel as they are, and each face loop runs in   */
/* two passes: the per-face work in parallel into one record per face,  */
/* then each cell adds its faces' records in ascending face_act order,  */
/* the order of the serial loops, so every cell's sums and therefore    */
/* the results are bit-identical to o_step_global.  Needs every cell    */
/* active and no boundary cells (the global step); otherwise it runs    */
/* o_step_global.  Built single-threaded it runs the same passes.       */
/* ------------------------------------------------------------------ */
typedef struct {
  int i, j, jbnd, pad;
  real v[NF + 3 + 2]; /* flux term, psi term, Bn term, dV */
} o_face_rec;

API int o_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

API double o_step_global_par(o_state *s, double dtmax) {
  const int n = s->np, nfa = s->nface_act;
  int ok = s->nactive == n;
  for (int i = 0; i < n && ok; i++) ok = s->bnd[i] == NO_BOUNDARY;
  if (!ok) return o_step_global(s, dtmax);

  /* each cell's active faces by ascending face_act position */
  int *posf = (int *)malloc(sizeof(int) * ((size_t)s->nface + 1));
  int *cpos = (int *)malloc(sizeof(int) * ((size_t)s->cell_off[n] + 1));
  int *ccnt = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  o_face_rec *rec = (o_face_rec *)malloc(sizeof(o_face_rec) * ((size_t)nfa + 1));
  int *failed = (int *)calloc((size_t)n + 1, sizeof(int));
#pragma omp parallel for schedule(static)
  for (int f = 0; f < s->nface; f++) posf[f] = -1;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < nfa; a++) posf[s->face_act[a]] = a;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    int *p = cpos + s->cell_off[c], m = 0;
    for (int k = s->cell_off[c]; k < s->cell_off[c + 1]; k++) {
      const int a = posf[s->cell_face[k]];
      if (a < 0) continue;
      int q = m++;
      while (q > 0 && p[q - 1] > a) {
        p[q] = p[q - 1];
        q--;
      }
      p[q] = a;
    }
    ccnt[c] = m;
  }
  /* every active face must be in the lists of both its cells (else the serial step) */
  {
    long long listed = 0;
#pragma omp parallel for schedule(static) reduction(+ : listed)
    for (int c = 0; c < n; c++) {
      const int *p = cpos + s->cell_off[c];
      for (int q = 0; q < ccnt[c]; q++) {
        const o_face *face = &s->faces[s->face_act[p[q]]];
        listed += (face->s1 == c) + (face->s2 == c) * 2 + (q > 0 && p[q] == p[q - 1]) * 4;
      }
    }
    if (listed != 3LL * nfa) {
      free(posf);
      free(cpos);
      free(ccnt);
      free(rec);
      free(failed);
      return o_step_global(s, dtmax);
    }
  }

  /* computeTimestep */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) timestep_cell(s, i);

  /* computePredictor: faces, then each cell's sums and update */
#pragma omp parallel for schedule(static)
  for (int a = 0; a < nfa; a++) {
    o_face_rec *r = &rec[a];
    predictor_face(s, &s->faces[s->face_act[a]], &r->i, &r->j, &r->jbnd, r->v, r->v + NF, r->v + NF + 3,
                   r->v + NF + 4);
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    o_fluidD d;
    real dv = 0.0;
    memset(&d, 0, sizeof d);
    const int *p = cpos + s->cell_off[c];
    for (int q = 0; q < ccnt[c]; q++) {
      const o_face_rec *r = &rec[p[q]];
      if (r->i == c) predictor_acc(&d, &dv, 1, r->v, r->v + NF, r->v[NF + 3], r->v[NF + 4]);
      else if (r->jbnd == NO_BOUNDARY) predictor_acc(&d, &dv, 0, r->v, r->v + NF, r->v[NF + 3], r->v[NF + 4]);
    }
    predictor_cell(s, c, dtmax, 0, &d, dv);
  }

  double dtmin = REF_HUGE;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
  for (int i = 0; i < n; i++) dtmin = smin(dtmin, s->dt_new[i]);
  const double dt = o_rung_dt(dtmax, 0, dtmin);
  const double t0 = s->t_global;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    s->tbeg[i] = t0;
    s->tend[i] = t0 + dt;
  }

  /* computeReconstruction */
#pragma omp parallel
  {
    o_rec_buf buf;
    rec_buf_init(&buf);
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) reconstruction_cell(s, i, &buf);
    rec_buf_free(&buf);
  }

  /* slopeLimiter + computeFluidUpdateI: each cell takes its own side of its faces */
  {
    real *area_test = (real *)calloc((size_t)n * 3, sizeof(real));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
      for (int k = 0; k < NF; k++) s->lim[c * NF + k] = 1.0;
      const int *p = cpos + s->cell_off[c];
      for (int q = 0; q < ccnt[c]; q++) {
        const o_face *face = &s->faces[s->face_act[p[q]]];
        slope_face(s, face, face->s1 == c ? 1 : 2, area_test);
      }
      closure_cell(s, c, area_test);
      limiter_apply_cell(s, c);
    }
    free(area_test);
  }

  /* computeFluidUpdate: faces, then each cell's FluidD */
#pragma omp parallel for schedule(static)
  for (int a = 0; a < nfa; a++) {
    o_face_rec *r = &rec[a];
    fluid_face(s, &s->faces[s->face_act[a]], &r->i, &r->j, r->v, r->v + NF, r->v + NF + 3);
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    o_fluidD acc;
    memset(&acc, 0, sizeof acc);
    const int *p = cpos + s->cell_off[c];
    for (int q = 0; q < ccnt[c]; q++) {
      const o_face_rec *r = &rec[p[q]];
      fluid_acc(&acc, r->i == c, r->v, r->v + NF, r->v[NF + 3]);
    }
    fluid_store(s, c, &acc);
  }

  /* computePrimitives at t + dt */
  s->t_global = t0 + dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) failed[i] = primitives_cell(s, i, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    if (failed[i]) primitives_failed_cell(s, i);

  free(posf);
  free(cpos);
  free(ccnt);
  free(rec);
  free(failed);
  return dt;
}
