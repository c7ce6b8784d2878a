// Device-resident stepping of the Voronoi-face pipeline ("fused" path):
// static mesh, every cell active, one global timestep (all cells on the
// smallest rung).  Per step three launches:
//
//   K1 k_cfl_predict_recon  one thread per cell, one pass over its face-ends:
//        computeTimestep (CFL, FP64), computePredictor (first-order HLLD sweep
//        + FP64 cell update -> Wt), computeReconstruction (gradients, min/max;
//        the FP64 path also applies the reconstruction limiter here)
//   K1b k_limit             opt-in limiters only: FP64 slopeLimiter + apply;
//        FP32 both limiters from one pass over 12-byte own-side offsets
//   K2 k_flux_apply (FP64) / k_flux_apply_coop (FP32, neighbour records
//        gathered cooperatively into LDS two face-end steps ahead), one
//        thread per cell, one pass over its face-ends: space-time
//        extrapolation + HLLD (computeFluidUpdate), FP64 sums of flux*dt*A in
//        the cell's face order, then computePrimitives (FP64 pressure recovery,
//        Dedner/Powell terms, fallbacks)
//   K3 k_advance            t += dt, reset the CFL minimum
//
// Each face is evaluated by both of its cells in the same canonical
// orientation (L = s1, R = s2, normal L->R), so both see bit-identical fluxes:
// the scatter is conservative, deterministic and needs no atomics.  Face
// arithmetic runs in T = double (the plain FP64 path, operation-for-operation
// the reference's) or T = float (FP32 face path); accumulation into conserved
// variables, pressure recovery and the CFL reduction are FP64 in both.
//
// Layout (SELL-C-sigma, C = 64 = one wavefront, sigma = 256 cells): the mesh's
// cells, expected in a space-filling-curve order, are renumbered inside each
// window of 256 consecutive cells by decreasing face count, so the 64 cells of
// a wavefront have nearly equal face counts and the per-face-end loop wastes
// few lanes.  Face-end k of the 64 cells of a wavefront is stored as 64
// consecutive records, so every record load of a wavefront is one contiguous
// 2 KiB (FP32) read.  Windows are spatially compact, so neighbour gathers
// still hit L2.  The state lives in internal order on the device while the
// path runs and is scattered back into the mesh-order fields after each run.
// Blocks are mapped XCD-aware: the 8 XCDs each sweep a contiguous eighth of
// the cells, keeping neighbour reuse inside one XCD's L2.
//
// This header is compiled once per face precision (fvm_fused_f32.hip,
// fvm_fused_f64.hip): the FP32 face path with fused multiply-add contraction,
// the FP64 path without, operation-for-operation the reference's.  Everything
// here has internal linkage, so the two translation units never share a
// symbol.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"
#include "fvm_physics.h"

namespace fvm {
namespace {

constexpr int FB = 256;  // threads per block of the fused kernels

__device__ __forceinline__ int xcd_cell(int nb) {
  // gridDim.x is padded to a multiple of 8; block b runs on XCD b % 8
  const int per = (int)gridDim.x >> 3;
  const int tile = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
  (void)nb;
  return tile * FB + (int)threadIdx.x;
}

// Per-cell gather record: the state a neighbour needs, contiguous, so one
// gather is a few 16-byte vector loads.
//   FP64: [0,10) W   [10,20) Wt   [20,50) gradients x[10] y[10] z[10]   (stride 52)
//   FP32: [0,10) W + Wt dti, the state at the half step that every face state of
//         the step starts from   [10,40) gradients   (stride 40: ten 16-byte
//         words, 160 bytes); Wt itself is kept per cell beside it (FusedArgs::tv)
// The FP64 path keeps W and Wt apart, because the reference forms a face state
// as W + (grad . dr + Wt dti) and the order is part of its arithmetic.
constexpr int REC = 52;
enum { R_W = 0, R_T = 10, R_G = 20 };
constexpr int REC32 = 40;
constexpr int R32_G = 10;
template <typename T> constexpr int rec_len() { return sizeof(T) == 4 ? REC32 : REC; }
// A second, compact copy of each cell's W, 12 values apart (48 bytes in FP32):
// the predictor kernel gathers only the neighbours' W, and from this array a
// cache line holds the W of more than two cells instead of part of one record.
constexpr int RWC = 12;

// Per face-end record (CSR order of the cell's face list): the neighbour
// (j, or ~j when this cell is s2) and the face geometry in canonical
// orientation s1 -> s2; FP64 also carries the reconstruction centroid of the
// own side as the reference forms it.
template <typename T> struct FaceEndRec;
// The unit normal and the area are formed once, in FP64, when the records are
// built (area = |face.n|, n = face.n / area as the reference forms them): the
// kernels skip a square root and a reciprocal per face-end, and the FP32 path
// never squares an area-vector component (faces of 1e-20 would underflow).
template <> struct alignas(16) FaceEndRec<float> {
  int nbr;
  float nrm[3];
  float drL[3];
  float area;
};
// FP64: area carries the sign of cen . face.n (the orientation of the
// reconstruction normal, computeReconstruction.cpp:113-123)
template <> struct alignas(16) FaceEndRec<double> {
  int nbr, pad;
  double nrm[3];
  double drL[3];
  double cen[3];
  double area;
};

template <typename T> struct FusedArgs {
  int n;                    // cells (all active), in the internal (window-sorted) order
  const int *deg;           // [n] face count of each cell
  const int *gbase;         // [n/64] first slot of each 64-cell group in fe
  const FaceEndRec<T> *fe;  // SELL: face-end k of cell c at gbase[c/64] + 64*k + c%64
  const float *fdr;         // FP32: own-side centroid offset of each face-end slot, [slot][3]
  const int *fnb;           // FP32: the neighbour field of each face-end slot (fe[slot].nbr), compact
  const double *cvol;       // [n]
  const double *lmin;       // [n] pow(V / (4 pi / 3), 1 / 3), formed once at setup
  const double *mvel;       // [3][n] generator velocity (moving mesh only)
  const T *rec;             // [n][rec_len<T>()] records of this step
  T *rec_w;                 // records written by this launch (K1: rec itself; K2: next step's)
  T *rec_next;
  const T *wcur;            // [n][RWC] compact W of this step (predictor gathers)
  T *wnext;                 // [n][RWC] compact W written for the next step
  T *tv;                    // FP32: [n][NF] Wt of this step (the predictor kernel writes it)
  double *W64;              // [10][n] FP64 primitive state (own cell only)
  double *U;                // [10][n] FP64 conserved state (own cell only)
  T *mm;                    // [20][n] min/max (opt-in limiters)
  double *dt_new;           // [n]
  unsigned long long *dtmin_bits;
  double *tdev;             // [0] t_global, [1] dt of the last step
  unsigned long long *diag;
  double gamma, courant, dtmax;
  int write_w64;            // FP32 path: store FP64 W (last step of a run)
  int k1_pass;              // FP32 path: 0 first pass of K1, 1 its check-and-redo pass
};

#define SF(p, k, i, n) (p)[(size_t)(k) * (size_t)(n) + (size_t)(i)]
#define FLD(p, k, i, n) SF(p, k, i, n)

// vector loads/stores of record slices (offsets are multiples of 16 bytes)
template <int N> __device__ __forceinline__ void ldv(const float *__restrict__ p, float *o) {
  int k = 0;
#pragma unroll
  for (; k + 4 <= N; k += 4) {
    const float4 v = *reinterpret_cast<const float4 *>(p + k);
    o[k] = v.x; o[k + 1] = v.y; o[k + 2] = v.z; o[k + 3] = v.w;
  }
#pragma unroll
  for (; k < N; k += 2) {
    const float2 v = *reinterpret_cast<const float2 *>(p + k);
    o[k] = v.x; o[k + 1] = v.y;
  }
}
template <int N> __device__ __forceinline__ void ldv(const double *__restrict__ p, double *o) {
#pragma unroll
  for (int k = 0; k < N; k += 2) {
    const double2 v = *reinterpret_cast<const double2 *>(p + k);
    o[k] = v.x; o[k + 1] = v.y;
  }
}
template <int N> __device__ __forceinline__ void stv(float *__restrict__ p, const float *o) {
  int k = 0;
#pragma unroll
  for (; k + 4 <= N; k += 4) *reinterpret_cast<float4 *>(p + k) = make_float4(o[k], o[k + 1], o[k + 2], o[k + 3]);
#pragma unroll
  for (; k < N; k += 2) *reinterpret_cast<float2 *>(p + k) = make_float2(o[k], o[k + 1]);
}
template <int N> __device__ __forceinline__ void stv(double *__restrict__ p, const double *o) {
#pragma unroll
  for (int k = 0; k < N; k += 2) *reinterpret_cast<double2 *>(p + k) = make_double2(o[k], o[k + 1]);
}

// geometry of one face-end: canonical (s1->s2) normal, area, drL, drR, and
// this cell's reconstruction centroid and outward normal
template <typename T> struct FE {
  int j;
  bool isL;
  T area, nrm[3], drL[3], drR[3], cen[3], ncen[3];
};

// streamed once per launch (own-cell state): non-temporal, so they do not
// evict the gathered neighbour records from L2
template <typename T> __device__ __forceinline__ T ntl(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void nts(T v, T *p) { __builtin_nontemporal_store(v, p); }

template <typename T>
__device__ __forceinline__ void decode_fe(const FaceEndRec<T> &r, FE<T> &g) {
  g.isL = r.nbr >= 0;
  g.j = g.isL ? r.nbr : ~r.nbr;
  for (int d = 0; d < 3; d++) {
    g.nrm[d] = r.nrm[d];
    g.drL[d] = r.drL[d];
  }
  const T dsh = g.drL[0] * g.nrm[0] + g.drL[1] * g.nrm[1] + g.drL[2] * g.nrm[2];
  const T dsl = T(2.0) * dsh;
  for (int d = 0; d < 3; d++) g.drR[d] = g.drL[d] - g.nrm[d] * dsl;
  if constexpr (std::is_same<T, double>::value) {
    // the reference's reconstruction geometry (computeReconstruction.cpp:97-123)
    g.area = dabs(r.area);
    for (int d = 0; d < 3; d++) {
      g.cen[d] = r.cen[d];
      g.ncen[d] = r.area < 0.0 ? -g.nrm[d] : g.nrm[d];
    }
  } else {
    g.area = r.area;
    for (int d = 0; d < 3; d++) {
      g.cen[d] = g.isL ? g.drL[d] : g.drR[d];
      g.ncen[d] = g.isL ? g.nrm[d] : -g.nrm[d];
    }
  }
}

template <typename T>
__device__ __forceinline__ void load_fe(const FusedArgs<T> &A, int e, FE<T> &g) {
  decode_fe(A.fe[e], g);
}
__device__ __forceinline__ int fe_nbr(int nbr) { return nbr >= 0 ? nbr : ~nbr; }

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
    const T r = frcp(sqr(dsl));
    for (int d = 0; d < 3; d++) wij[d] = (vj[d] + vi[d]) * T(0.5) - rij[d] * fdv * r;
    dV = fdiv(g.area * (rij[0] * wij[0] + rij[1] * wij[1] + rij[2] * wij[2]), dsl);
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

// adds the sum over the wavefront's active lanes of v (0 <= v < 2^16), from
// ballots only, so lanes that have left the kernel contribute nothing
__device__ __forceinline__ void wave_add(unsigned long long *ctr, int v) {
  if (__ballot(v != 0) == 0) return;
  unsigned long long tot = 0;
  for (int b = 0; b < 16; b++) tot += (unsigned long long)__popcll(__ballot((v >> b) & 1)) << b;
  if (__lane_id() == __ffsll((long long)__ballot(1)) - 1) atomicAdd(ctr, tot);
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
// FP32 predictor kernel: a lane's reconstruction accumulators (gradient sums
// x, y, z; neighbour min/max differences) live in its LDS slice between face-end
// steps, so they do not hold registers through the HLLD and the kernel fits
// three wavefronts per SIMD.  13 float4 per lane at a 208-byte stride: the 16
// lanes of one LDS cycle hit distinct banks; 53 KiB per block, three blocks in
// the 160 KiB of a CU.
constexpr int K1ACC = 13;
__device__ __forceinline__ float4 *k1_acc_slice() {
  __shared__ float4 k1acc[FB * K1ACC];
  return k1acc + threadIdx.x * K1ACC;
}
template <int NQ> __device__ __forceinline__ void acc_ld(const float4 *p, float *x) {
#pragma unroll
  for (int i = 0; i < NQ; i++) {
    const float4 v = p[i];
    x[4 * i] = v.x; x[4 * i + 1] = v.y; x[4 * i + 2] = v.z; x[4 * i + 3] = v.w;
  }
}
template <int NQ> __device__ __forceinline__ void acc_st(float4 *p, const float *x) {
#pragma unroll
  for (int i = 0; i < NQ; i++) p[i] = make_float4(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
}
// FP32 limiters from one pass.  phi (minmod min(1, x); smooth
// (x^2 + 2x + e)/(x^2 + x + 2 + e)) rises on [0, X1] to phi(X1) = 1 and stays
// >= 1 beyond (X1 = 1 minmod, 2 smooth), so a limiter min(1, min_f phi(x_f))
// with x_f = bound / d_f >= 0 equals min(1, phi(min_f x_f)), and min_f x_f
// = min(mx / max_f d_f, mn / min_f d_f).  Faces with |d_f| <= tau = 1e-16
// max(|mn|, |mx|) count as 1 in the reference; they never set the minimum
// while a face beyond tau exists, and a side whose extreme is within tau is
// ignored, as the reference ignores it.  The reconstruction limiter takes d_f
// = grad . cen_f (computeReconstruction.cpp:214-239); the slope limiter takes
// the reconstruction-limited lr (grad . dr_f + t dtt) (slopeLimiter.cpp:20-77);
// in the FP32 records cen_f = dr_f.
template <int LIM> __device__ __forceinline__ float lim_from_ratio(float x) {
  constexpr float X1 = LIM == LIM_MINMOD ? 1.0f : 2.0f;
  return x >= X1 ? 1.0f : phi_c<LIM, float>(x);
}
// min_f bound / (s d_f) over the faces beyond the flat threshold
__device__ __forceinline__ float min_ratio(float mn, float mx, float dmin, float dmax, float s) {
  const float inf = __builtin_huge_valf();
  const float tau = 1.0e-16f * smax(dabs(mx), dabs(mn));
  const float hi = dmax * s, lo = dmin * s;
  const float a = hi > tau ? (mx > 0.0f ? fdiv(mx, hi) : 0.0f) : inf;
  const float b = lo < -tau ? (mn < 0.0f ? fdiv(mn, lo) : 0.0f) : inf;
  return smin(a, b);
}

struct Dr3 {
  float x, y, z;
};

// the slope limiter's time offset for a step of length dt from t0: every cell
// has tend = t0 + dt, so dtI = dt, dth = dt / 2 and dtt = dtI - dth
__device__ __forceinline__ float limiter_dtt(double t0, double dt_glob) {
  const double tend = t0 + dt_glob;
  const double dtI = tend - t0;
  const double dth = 0.5 * smin(dtI, dtI);
  return (float)(dtI - dth);
}

// both FP32 limiters of one cell from one pass over its own-side offsets: mn /
// mx the raw neighbour differences, tv and gg (x, y, z blocks) Wt and the
// gradients, limited in place
template <int LIM>
__device__ __forceinline__ void limit_cell32(const FusedArgs<float> &A, int e0, int deg, const float *mn,
                                             const float *mx, float dtt, float *tv, float *gg) {
  // extremes of d_f = grad . dr_f over the faces; the slope limiter's
  // y_f = d_f + t dtt differs by a face-independent term, and rounding is
  // monotone, so max_f fl(d_f + t dtt) = fl(max_f d_f + t dtt)
  float dmin[NF], dmax[NF];
  for (int k = 0; k < NF; k++) {
    dmin[k] = __builtin_huge_valf();
    dmax[k] = -__builtin_huge_valf();
  }
  // the own-side offsets only (12 bytes per face-end, not the 32-byte record)
  const Dr3 *fdr = reinterpret_cast<const Dr3 *>(A.fdr);
  Dr3 rn;
  if (deg > 0) rn = fdr[e0];
  for (int q = 0, e = e0; q < deg; q++, e += 64) {
    const Dr3 r = rn;
    if (q + 1 < deg) rn = fdr[e + 64];
    const float dr[3] = {r.x, r.y, r.z};
    for (int k = 0; k < NF; k++) {
      const float d = gg[k] * dr[0] + gg[NF + k] * dr[1] + gg[2 * NF + k] * dr[2];
      dmin[k] = smin(dmin[k], d);
      dmax[k] = smax(dmax[k], d);
    }
  }
  for (int k = 0; k < NF; k++) {
    // the extremes start at 0 (a cell without faces gives 0, 0)
    const float dlo = smin(0.0f, dmin[k]), dhi = smax(0.0f, dmax[k]);
    const float ylo = smin(0.0f, __builtin_fmaf(tv[k], dtt, dmin[k]));
    const float yhi = smax(0.0f, __builtin_fmaf(tv[k], dtt, dmax[k]));
    const float lr = lim_from_ratio<LIM>(min_ratio(mn[k], mx[k], dlo, dhi, 1.0f));
    const float lm = lr * lim_from_ratio<LIM>(min_ratio(mn[k], mx[k], ylo, yhi, lr));
    tv[k] *= lm;
    gg[k] *= lm;
    gg[NF + k] *= lm;
    gg[2 * NF + k] *= lm;
  }
}

// the step's dt and the extrapolation time of the face states: every cell has
// tend = t0 + dt, so dtI = dtJ, dt = min(dtI, dtJ), dti = dtI - dt/2
struct StepTimes {
  double t0, dt_glob, dtf, dti;
};
__device__ __forceinline__ StepTimes step_times_dt(double t0, double dt_glob) {
  StepTimes s;
  s.t0 = t0;
  s.dt_glob = dt_glob;
  const double tend = s.t0 + s.dt_glob;
  const double dtI = tend - s.t0;
  s.dtf = smin(dtI, dtI);
  s.dti = dtI - 0.5 * s.dtf;
  return s;
}
template <typename T> __device__ __forceinline__ StepTimes step_times(const FusedArgs<T> &A) {
  return step_times_dt(A.tdev[0], rung_dt0(A.dtmax, __longlong_as_double((long long)*A.dtmin_bits)));
}
// the half-step state W + Wt dti of one variable (FP32 path)
__device__ __forceinline__ float half_step32(float W, float t, float dti) { return __builtin_fmaf(t, dti, W); }

template <typename T> struct K1Waves {
  static constexpr int value = std::is_same<T, float>::value ? 3 : 1;
};

template <typename T, bool MOVING, int LIM>
__global__ void __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(K1Waves<T>::value)))
k_cfl_predict_recon(FusedArgs<T> A) {
  constexpr bool LIMIT = LIM != LIM_SHIPPED;
  constexpr bool FP32 = std::is_same<T, float>::value;
  constexpr bool LIMIT32 = LIMIT && FP32;
  // FP32 path: the opt-in limiters and the half-step states of the records are
  // formed here with the step's dt taken to be the last step's (dt is the
  // global minimum this kernel is computing).  A second launch (k1_pass 1)
  // compares it with the true dt: it returns at once when they agree (every
  // step of a run but the first, while the rung holds) and otherwise redoes
  // every cell with the true dt, so the result never depends on the guess
  double dt_step = 0.0;
  if constexpr (FP32) {
    dt_step = A.tdev[1];
    if (A.k1_pass) {
      const double dt_true = rung_dt0(A.dtmax, __longlong_as_double((long long)*A.dtmin_bits));
      if (dt_true == dt_step) return;  // uniform over the grid
      dt_step = dt_true;
    }
  }
  const int c = xcd_cell(A.n);
  const bool live = c < A.n;
  double dtn = REF_HUGE;
  if (live) {
    const int n = A.n;
    const double g64 = A.gamma;
    const T g = (T)A.gamma;
    const double V = ntl(A.cvol + c);
    T Wc[NF];
    ldv<NF>(A.wcur + (size_t)c * RWC, Wc);

    // CFL (computeTimestep.cpp:13-31), FP64 from U
    double pv[3] = {0.0, 0.0, 0.0};
    if constexpr (MOVING)
      for (int d = 0; d < 3; d++) pv[d] = SF(A.mvel, d, c, n);
    double v0[3], csi = 0.0, h = 0.0, dti = 0.0;
    // the own-cell terms, from the FP64 primitives
    auto cfl_cell = [&](const double *W0) {
      h = A.lmin[c];  // pow(V / (4 pi / 3), 1 / 3)
      const double Pi = pressure(g64, W0[ETHM]) * g64;
      const double B2i = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
      csi = dsqrt((Pi + B2i) / W0[DENS]);
      const double dvr[3] = {W0[VELX] - pv[0], W0[VELY] - pv[1], W0[VELZ] - pv[2]};
      const double vrel = dsqrt(dvr[0] * dvr[0] + dvr[1] * dvr[1] + dvr[2] * dvr[2]);
      dti = h / (csi + vrel + 1.0 / REF_HUGE);
    };
    if constexpr (std::is_same<T, float>::value) {
      // FP32 face path: the neighbour speed differences take the own velocity
      // from the FP32 record, as the neighbours' come from theirs; U is read
      // once, after the face loop, for the own-cell terms and the predictor
      for (int d = 0; d < 3; d++) v0[d] = (double)Wc[VELX + d];
    } else {
      double U0[NF], W0[NF];
      for (int k = 0; k < NF; k++) U0[k] = ntl(&SF(A.U, k, c, n));
      to_primitive(U0, V, W0);
      cfl_cell(W0);
      v0[0] = W0[VELX];
      v0[1] = W0[VELY];
      v0[2] = W0[VELZ];
    }
    double arg_max = -1.0;  // min_j h/(cs + sqrt(a_j)) = h/(cs + sqrt(max_j a_j)), monotone roundings

    T gx[NF], gy[NF], gz[NF], dmn[NF], dmx[NF], pU[NF], pgp[3], pdv;
    double pdV = 0.0;
    for (int k = 0; k < NF; k++) gx[k] = gy[k] = gz[k] = dmn[k] = dmx[k] = pU[k] = T(0.0);
    pgp[0] = pgp[1] = pgp[2] = pdv = T(0.0);
    T gw[3] = {T(0.0), T(0.0), T(0.0)};  // FP32 reconstruction: geometric weight of the own state
    // FP32: accumulators x[0,30) gradient sums, x[30,50) min/max, in LDS
    constexpr int K1NQ = LIMIT ? K1ACC : 8;
    if constexpr (std::is_same<T, float>::value) {
      float z[4 * K1ACC];
      for (int k = 0; k < 4 * K1ACC; k++) z[k] = 0.0f;
      acc_st<K1NQ>(k1_acc_slice(), z);
    }

    const int deg = A.deg[c];
    const int e0 = A.gbase[c >> 6] + (c & 63);
    // the next face-end record is loaded one iteration ahead, so the neighbour
    // gather of an iteration issues without waiting for its own record
    FaceEndRec<T> rn;
    if (deg > 0) rn = A.fe[e0];
    for (int q = 0, e = e0; q < deg; q++, e += 64) {
      const FaceEndRec<T> rc0 = rn;
      if (q + 1 < deg) rn = A.fe[e + 64];
      const int j = fe_nbr(rc0.nbr);
      T Wj[NF];
      ldv<NF>(A.wcur + (size_t)j * RWC, Wj);
      FE<T> fe;
      decode_fe(rc0, fe);

      // --- reconstruction sums (computeReconstruction.cpp:108-149)
      {
        const T dsh = fe.cen[0] * fe.ncen[0] + fe.cen[1] * fe.ncen[1] + fe.cen[2] * fe.ncen[2];
        const T ids = (dsh != T(0.0)) ? fdiv(T(0.5), dsh) : T(0.0);
        const T idsA = fe.area * ids;
        if constexpr (std::is_same<T, float>::value) {
          // FP32: (Wj + Wc) drh + (Wj - Wc) fij = Wj cen + Wc (2 drh - cen); the
          // Wc part is summed as geometry once per face and applied per cell
          T a[3];
          for (int d = 0; d < 3; d++) {
            a[d] = idsA * fe.cen[d];
            gw[d] += idsA * (T(2.0) * fe.ncen[d] * dsh - fe.cen[d]);
          }
          // accumulators from and back to the lane's LDS slice; the barrier
          // keeps the compiler from carrying them in registers across steps
          __asm__ __volatile__("" ::: "memory");
          float4 *acc = k1_acc_slice();
          float x[4 * K1ACC];
          acc_ld<K1NQ>(acc, x);
          for (int k = 0; k < NF; k++) {
            if constexpr (LIMIT) {
              const T diff = Wj[k] - Wc[k];
              x[3 * NF + k] = smin(x[3 * NF + k], diff);
              x[4 * NF + k] = smax(x[4 * NF + k], diff);
            }
            x[k] += a[0] * Wj[k];
            x[NF + k] += a[1] * Wj[k];
            x[2 * NF + k] += a[2] * Wj[k];
          }
          acc_st<K1NQ>(acc, x);
        } else {
          T drh[3], fij[3];
          for (int d = 0; d < 3; d++) {
            drh[d] = fe.ncen[d] * dsh;
            fij[d] = fe.cen[d] - drh[d];
          }
          for (int k = 0; k < NF; k++) {
            const T sum = Wj[k] + Wc[k];
            const T diff = Wj[k] - Wc[k];
            if constexpr (LIMIT) {
              dmn[k] = smin(dmn[k], diff);
              dmx[k] = smax(dmx[k], diff);
            }
            gx[k] += (sum * drh[0] + diff * fij[0]) * idsA;
            gy[k] += (sum * drh[1] + diff * fij[1]) * idsA;
            gz[k] += (sum * drh[2] + diff * fij[2]) * idsA;
          }
        }
      }
      // --- CFL neighbour term (computeTimestep.cpp:37-54)
      {
        double dwi[3] = {0.0, 0.0, 0.0};
        if constexpr (MOVING)
          for (int d = 0; d < 3; d++) dwi[d] = (double)(T)SF(A.mvel, d, j, n) - pv[d];
        const double dvi[3] = {(double)Wj[VELX] - v0[0], (double)Wj[VELY] - v0[1], (double)Wj[VELZ] - v0[2]};
        const double a = (dwi[0] * dwi[0] + dwi[1] * dwi[1] + dwi[2] * dwi[2]) +
                         1.0 * (dvi[0] * dvi[0] + dvi[1] * dvi[1] + dvi[2] * dvi[2]) + 1.0 / REF_HUGE;
        arg_max = (arg_max < a) ? a : arg_max;
      }
      // --- predictor face flux, first order (computePredictor.cpp:44-136)
      {
        const int L = fe.isL ? c : j, R = fe.isL ? j : c;
        T WL[NF], WR[NF];
        for (int k = 0; k < NF; k++) {
          WL[k] = fe.isL ? Wc[k] : Wj[k];
          WR[k] = fe.isL ? Wj[k] : Wc[k];
        }
        T wij[3], dV;
        face_velocity<T, MOVING>(A, L, R, fe, wij, dV);
        T psi, Bn, flux[NF];
        compute_flux<T>(g, wij, WL, WR, NO_BOUNDARY, fe.nrm, psi, Bn, flux);
        // signed area: the L cell loses what flows L -> R (negation is exact)
        const T sa = fe.isL ? -fe.area : fe.area;
        for (int k = 0; k < NF; k++) pU[k] = pU[k] + flux[k] * sa;
        for (int d = 0; d < 3; d++) pgp[d] = pgp[d] - fe.nrm[d] * (psi * sa);
        pdv = pdv - Bn * sa;
        pdV += fe.isL ? (double)dV : -(double)dV;
      }
    }
    // --- predictor cell update (computePredictor.cpp:141-379), FP64
    T tv[NF];
    {
      double U0[NF], W0[NF];
      for (int k = 0; k < NF; k++) U0[k] = ntl(&SF(A.U, k, c, n));
      to_primitive(U0, V, W0);
      if constexpr (std::is_same<T, float>::value) cfl_cell(W0);
      if constexpr (std::is_same<T, float>::value) {
      __asm__ __volatile__("" ::: "memory");
      float x[4 * K1ACC];
      acc_ld<K1NQ>(k1_acc_slice(), x);
      for (int k = 0; k < NF; k++) {
        gx[k] = x[k];
        gy[k] = x[NF + k];
        gz[k] = x[2 * NF + k];
        if constexpr (LIMIT) {
          dmn[k] = x[3 * NF + k];
          dmx[k] = x[4 * NF + k];
        }
      }
    }
    if (arg_max >= 0.0) dti = smin(dti, h / (csi + dsqrt(arg_max)));
      dti = smin(dti, REF_HUGE);
      dtn = dti * A.courant;
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
      const double l_min = A.lmin[c];
      Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
      double W1[NF];
      // exp(0) = 1 exactly: a static mesh skips the exponential
      to_primitive(Uc, pdV == 0.0 ? V : V * exp(pdV * dt / V), W1);
      const bool ef = W1[ETHM] <= 0.0;
      if (ef) W1[ETHM] = ethm_from_entropy_d(g64, W1);
      else W1[ENTR] = entropy_from_ethm_d(g64, W1);
      if (!A.k1_pass) wave_count(&A.diag[FVM_DIAG_ETHM_FAILED], ef);  // a redo counts nothing twice
      const double idt = 1.0 / dt;
      for (int k = 0; k < NF; k++) tv[k] = (T)((W1[k] - W0[k]) * idt);
    }

    // --- reconstruction finish (computeReconstruction.cpp:162-239)
    {
      const double invV = 1.0 / V;
      if constexpr (std::is_same<T, float>::value)
        for (int k = 0; k < NF; k++) {
          gx[k] += gw[0] * Wc[k];
          gy[k] += gw[1] * Wc[k];
          gz[k] += gw[2] * Wc[k];
        }
      for (int k = 0; k < NF; k++) {
        gx[k] = (T)((double)gx[k] * invV);
        gy[k] = (T)((double)gy[k] * invV);
        gz[k] = (T)((double)gz[k] * invV);
      }
      if constexpr (LIMIT32) {
        // FP32: the reconstruction and slope limiters in one pass over the
        // own-side offsets, from the raw differences and the unlimited gradients
        T gg[3 * NF];
        for (int k = 0; k < NF; k++) {
          gg[k] = gx[k];
          gg[NF + k] = gy[k];
          gg[2 * NF + k] = gz[k];
        }
        limit_cell32<LIM>(A, e0, deg, dmn, dmx, limiter_dtt(A.tdev[0], dt_step), tv, gg);
        for (int k = 0; k < NF; k++) {
          gx[k] = gg[k];
          gy[k] = gg[NF + k];
          gz[k] = gg[2 * NF + k];
        }
      } else if constexpr (LIMIT) {
        T lm[NF];
        for (int k = 0; k < NF; k++) lm[k] = T(1.0);
        for (int q = 0, e = e0; q < deg; q++, e += 64) {
          FE<T> fe;
          load_fe(A, e, fe);
          for (int k = 0; k < NF; k++) {
            const T dWx = gx[k] * fe.cen[0] + gy[k] * fe.cen[1] + gz[k] * fe.cen[2];
            lm[k] = smin(lm[k], limiter_rec_c<LIM, T>(dWx, dmn[k], dmx[k]));
          }
        }
        for (int k = 0; k < NF; k++) {
          SF(A.mm, k, c, n) = dmn[k] + Wc[k];
          SF(A.mm, NF + k, c, n) = dmx[k] + Wc[k];
          tv[k] = tv[k] * lm[k];
          gx[k] = gx[k] * lm[k];
          gy[k] = gy[k] * lm[k];
          gz[k] = gz[k] * lm[k];
        }
      }
      if constexpr (FP32) {
        // the record starts from the half-step state W + Wt dti of the step,
        // formed as the flux kernel formed it; Wt goes to its own array
        const float dti = (float)step_times_dt(A.tdev[0], dt_step).dti;
        T wh[NF];
        for (int k = 0; k < NF; k++) wh[k] = half_step32(Wc[k], tv[k], dti);
        T *r = A.rec_w + (size_t)c * REC32;
        stv<NF>(r, wh);
        stv<NF>(r + R32_G, gx);
        stv<NF>(r + R32_G + NF, gy);
        stv<NF>(r + R32_G + 2 * NF, gz);
        stv<NF>(A.tv + (size_t)c * NF, tv);
      } else {
        T *wc = A.rec_w + (size_t)c * REC;
        stv<NF>(wc + R_T, tv);
        stv<NF>(wc + R_G, gx);
        stv<NF>(wc + R_G + NF, gy);
        stv<NF>(wc + R_G + 2 * NF, gz);
      }
    }
    A.dt_new[c] = dtn;
  }
  // global minimum of dt_new (dt > 0: IEEE bit patterns order like the values),
  // one atomic per block
  double v = dtn;
  for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o));
  __shared__ double wmin[FB / 64];
  if (__lane_id() == 0) wmin[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < FB / 64; w++) v = smin(v, wmin[w]);
    if (v < REF_HUGE && v >= 0.0) atomicMin(A.dtmin_bits, (unsigned long long)__double_as_longlong(v));
  }
}

// ---------------------------------------------------------------------------
// K1b (opt-in limiters, FP64 path): slopeLimiter compute part +
// computeFluidUpdateI for the global step, per cell over its own face-ends
// (own data only)
// ---------------------------------------------------------------------------
template <typename T, int LIM>
__global__ void __launch_bounds__(FB) k_limit(FusedArgs<T> A) {
  static_assert(std::is_same<T, double>::value, "the FP32 path limits inside k_cfl_predict_recon");
  const int c = xcd_cell(A.n);
  if (c >= A.n) return;
  const int n = A.n;
  const double t0 = A.tdev[0];
  const double dt_glob = rung_dt0(A.dtmax, __longlong_as_double((long long)*A.dtmin_bits));
  const double tend = t0 + dt_glob;
  const double dtI = tend - t0;
  const double dth = 0.5 * smin(dtI, dtI);
  T *rc = A.rec_w + (size_t)c * REC;
  T W[NF], tv[NF], gg[3 * NF];
  ldv<NF>(rc + R_W, W);
  ldv<NF>(rc + R_T, tv);
  ldv<3 * NF>(rc + R_G, gg);
  T lo[NF], hi[NF];
  for (int k = 0; k < NF; k++) {
    lo[k] = SF(A.mm, k, c, n) - W[k];
    hi[k] = SF(A.mm, NF + k, c, n) - W[k];
    lo[k] = smin(lo[k], T(0.0));
    hi[k] = smax(hi[k], T(0.0));
  }
  T lm[NF];
  for (int k = 0; k < NF; k++) lm[k] = T(1.0);
  const T dtt = (T)(dtI - dth);
  const int deg = A.deg[c];
  const int e0 = A.gbase[c >> 6] + (c & 63);
  FaceEndRec<T> rn;
  if (deg > 0) rn = A.fe[e0];
  for (int q = 0, e = e0; q < deg; q++, e += 64) {
    const FaceEndRec<T> rc0 = rn;
    if (q + 1 < deg) rn = A.fe[e + 64];
    FE<T> fe;
    decode_fe(rc0, fe);
    const T *dr = fe.isL ? fe.drL : fe.drR;
    for (int k = 0; k < NF; k++) {
      const T dW = gg[k] * dr[0] + gg[NF + k] * dr[1] + gg[2 * NF + k] * dr[2] + tv[k] * dtt;
      lm[k] = smin(lm[k], limiter_sl_c<LIM, T>(dW, lo[k], hi[k]));
    }
  }
  for (int k = 0; k < NF; k++) {
    tv[k] *= lm[k];
    gg[k] *= lm[k];
    gg[NF + k] *= lm[k];
    gg[2 * NF + k] *= lm[k];
  }
  stv<NF>(rc + R_T, tv);
  stv<3 * NF>(rc + R_G, gg);
}

// ---------------------------------------------------------------------------
// K2: computeFluidUpdate (extrapolation + HLLD + FP64 scatter sums) and
// computePrimitives (FP64) for the own cell
// ---------------------------------------------------------------------------
// one face-end, given both extrapolated face states (Wo own side, Wn
// neighbour side): HLLD in canonical orientation, FP64 sums of flux * dt * A
// into the own cell (computeFluidUpdate.cpp:209-307)
template <typename T, bool MOVING>
__device__ __forceinline__ void face_update(const FusedArgs<T> &A, int c, const FE<T> &fe, const T *Wn, const T *Wo,
                                            double dtf, double *dU, double *gp, double &dv, int &nbad) {
  T WL[NF], WR[NF];
  for (int k = 0; k < NF; k++) {
    WL[k] = fe.isL ? Wo[k] : Wn[k];
    WR[k] = fe.isL ? Wn[k] : Wo[k];
  }
  // a face state where the reference asserts (computeFluidUpdate.cpp:179-189),
  // counted once per face: by its L cell
  nbad += (fe.isL && (!(WL[DENS] > T(0.0)) || !(WL[ETHM] > T(0.0)) || !(WR[DENS] > T(0.0)) || !(WR[ETHM] > T(0.0))))
              ? 1 : 0;
  T wij[3], dV;
  face_velocity<T, MOVING>(A, fe.isL ? c : fe.j, fe.isL ? fe.j : c, fe, wij, dV);
  T psi, Bn, flux[NF];
  compute_flux<T>((T)A.gamma, wij, WL, WR, NO_BOUNDARY, fe.nrm, psi, Bn, flux);
  // the L cell loses what flows L -> R: one signed area instead of two
  // branches (negation is exact, so the sums are those of the reference's
  // dU - F dt A and dU + F dt A)
  const double tarea = dtf * (double)fe.area;
  const double sa = fe.isL ? -tarea : tarea;
  for (int k = 0; k < NF; k++) dU[k] = dU[k] + (double)flux[k] * sa;
  for (int d = 0; d < 3; d++) gp[d] = gp[d] - (double)fe.nrm[d] * ((double)psi * sa);
  dv = dv - (double)Bn * sa;
}

// computePrimitives at t1 = t0 + dt (computeFluidUpdate.cpp:367-505) for the
// own cell, from the summed face terms
template <typename T>
__device__ __forceinline__ void cell_update(const FusedArgs<T> &A, int c, double t0, double dt_glob, const double *dU,
                                            const double *gp, double dv, const T *tc) {
  const int n = A.n;
  const double g64 = A.gamma;
  const double t1 = t0 + dt_glob;
  const double dt = t1 - t0;
  const double V = ntl(A.cvol + c);
  double Ui[NF], wi[NF], Uc[NF], W1[NF];
  for (int k = 0; k < NF; k++) Ui[k] = ntl(&SF(A.U, k, c, n));
  if constexpr (sizeof(T) == 8) {
    // the FP64 path keeps the reference's stored primitives
    for (int k = 0; k < NF; k++) wi[k] = ntl(&SF(A.W64, k, c, n));
  } else {
    // the FP32 face path re-derives them from U (they differ by rounding only)
    to_primitive(Ui, V, wi);
  }
  for (int k = 0; k < NF; k++) Uc[k] = Ui[k] + (0.0 + dU[k]);
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
    const double l_min = A.lmin[c];
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
  T Wn[NF];
  for (int k = 0; k < NF; k++) {
    if (sizeof(T) == 8 || A.write_w64) nts(W1[k], &SF(A.W64, k, c, n));
    nts(Ui[k], &SF(A.U, k, c, n));
    Wn[k] = (T)W1[k];
  }
  if constexpr (sizeof(T) == 8) stv<NF>(A.rec_next + (size_t)c * REC + R_W, Wn);  // FP32: K1 writes the record
  stv<NF>(A.wnext + (size_t)c * RWC, Wn);
}

// generic K2: each lane gathers its own neighbour records
template <typename T, bool MOVING>
__global__ void __launch_bounds__(FB) k_flux_apply(FusedArgs<T> A) {
  const int c = xcd_cell(A.n);
  if (c >= A.n) return;
  const StepTimes st = step_times(A);
  const T dti = (T)st.dti;

  const T *rc = A.rec + (size_t)c * REC;
  T Wc[NF], tc[NF], gc[3 * NF];
  ldv<NF>(rc + R_W, Wc);
  ldv<NF>(rc + R_T, tc);
  ldv<3 * NF>(rc + R_G, gc);

  double dU[NF], gp[3], dv = 0.0;
  for (int k = 0; k < NF; k++) dU[k] = 0.0;
  gp[0] = gp[1] = gp[2] = 0.0;
  int nbad = 0;

  const int deg = A.deg[c];
  const int e0 = A.gbase[c >> 6] + (c & 63);
  FaceEndRec<T> rn;
  if (deg > 0) rn = A.fe[e0];
  for (int q = 0, e = e0; q < deg; q++, e += 64) {
    const FaceEndRec<T> rc0 = rn;
    if (q + 1 < deg) rn = A.fe[e + 64];
    const int j = fe_nbr(rc0.nbr);
    // neighbour face state, extrapolated as soon as it is loaded
    T Wn[NF];
    FE<T> fe;
    {
      const T *rj = A.rec + (size_t)j * REC;
      T Wj[NF], tj[NF], gj[3 * NF];
      ldv<NF>(rj + R_W, Wj);
      ldv<NF>(rj + R_T, tj);
      ldv<3 * NF>(rj + R_G, gj);
      decode_fe(rc0, fe);
      const T *dr = fe.isL ? fe.drR : fe.drL;
      for (int k = 0; k < NF; k++) Wn[k] = Wj[k] + (gj[k] * dr[0] + gj[NF + k] * dr[1] + gj[2 * NF + k] * dr[2] + tj[k] * dti);
    }
    T Wo[NF];
    {
      const T *dr = fe.isL ? fe.drL : fe.drR;
      for (int k = 0; k < NF; k++) Wo[k] = Wc[k] + (gc[k] * dr[0] + gc[NF + k] * dr[1] + gc[2 * NF + k] * dr[2] + tc[k] * dti);
    }
    face_update<T, MOVING>(A, c, fe, Wn, Wo, st.dtf, dU, gp, dv, nbad);
  }
  wave_add(&A.diag[FVM_DIAG_POSITIVITY], nbad);
  cell_update<T>(A, c, st.t0, st.dt_glob, dU, gp, dv, tc);
}

// FP32 K2 with a cooperative, asynchronous neighbour gather.  At face-end
// step q the wavefront needs 64 neighbour records of 160 bytes (10 16-byte
// slots).  Lane-per-record loads touch 64 records per instruction and saturate
// the vector-memory address unit; here the 64 records of a step are fetched as
// one stream of 640 slots, 64 per load: lane l of load k fetches slot
// (64k + l) % 10 of record (64k + l) / 10, so each load instruction reads 6.4
// whole records contiguously, and the loads go straight to LDS
// (global_load_lds_dwordx4: lane l of load k lands at byte 1024k + 16l of the
// buffer, i.e. record r at 160r, packed).  Each wavefront has two buffers: the
// gather of step q+2 is issued as soon as every lane has read its step-q
// record, so a gather has two HLLD steps to arrive.  Reading its record, lane
// l addresses slot s as its base 160l plus the instruction's constant 16s;
// lanes l and l + 8 of a 16-lane LDS cycle then share banks (a two-way
// conflict on 10 reads per step, small beside the step's HLLD).
constexpr int GQ = REC32 / 4;  // float4 of an FP32 record (all of it)
constexpr int GBUF = 64 * GQ;  // float4 of one step's buffer (10 KiB)

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// s_waitcnt immediates (gfx9 encoding: vmcnt [3:0]+[15:14], expcnt [6:4], lgkmcnt [11:8])
constexpr int WAIT_ALL = 0;           // vmcnt(0) expcnt(0) lgkmcnt(0)
constexpr int WAIT_LGKM = 0xC07F;     // lgkmcnt(0) only
constexpr int WAIT_VM0 = 0x0F70;      // vmcnt(0) only
constexpr int WAIT_VMGQ = 0x0F70 | GQ;  // vmcnt(GQ): all but the last GQ vector-memory operations
static_assert(GQ < 16, "vmcnt immediate");

// gather of one face-end step: lane l asks for the record of neighbour jsel
// (>= 0; lanes without a face-end ask for record 0, so that every lane issues
// the same loads).  buf is the wavefront's buffer of the step (uniform).
// Position 64k + l = 60k + t (t = 4k + l < 100) is slot t % 10 of record
// 6k + t / 10.
// Two forms.  Up to 2^24 cells (records under 2.7 GB) they are read as a
// buffer (rs: base and byte size) at a 32-bit byte offset, one 24-bit
// multiply-add per load; a load past the size reads zeros.  Larger sets use
// 64-bit global addresses.  The buffer form also keeps the compiler's wait
// counting exact: it treats a pending global (FLAT-encoded) load into LDS as
// possibly out of order and waits for every load before using any register
// loaded after it, which would also wait for the gather of the step after
// next.
struct GatherSrc {
  __amdgpu_buffer_rsrc_t rs;
  const float *rec;
};
template <bool OFF32>
__device__ __forceinline__ void gather_async(const GatherSrc &G, int jsel, float4 *buf, int lane) {
  // every lane's neighbour index first (the permutes overlap one another),
  // then the loads
  int jr[GQ];
#pragma unroll
  for (int k = 0; k < GQ; k++) {
    // record 6k + t / 10 belongs to that lane: byte address 4 (t / 10) plus
    // the constant 24 k, which the instruction takes as its offset
    const unsigned t = (unsigned)(4 * k + lane);
    jr[k] = __builtin_amdgcn_ds_bpermute((int)(4u * (t / 10u)) + 24 * k, jsel);
  }
#pragma unroll
  for (int k = 0; k < GQ; k++) {
    const unsigned t = (unsigned)(4 * k + lane);
    const unsigned slot = t - 10u * (t / 10u);
    if constexpr (OFF32) {
      const unsigned off = __umul24((unsigned)jr[k], (unsigned)(REC32 * sizeof(float))) + 16u * slot;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(G.rs, (lds_void_t *)(buf + 64 * k), 16, (int)off, 0, 0, 0);
    } else {
      const float *src = G.rec + 4 * slot + (size_t)(unsigned)jr[k] * REC32;
      __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)(buf + 64 * k), 16, 0, 0);
    }
  }
}

typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const f4v_t lds_cf4_t;

// face state of one side: (W + t dti) + grad . dr with explicit FMAs.  A cell
// forms its own side and its neighbour's with the same operations, so the two
// cells of a face build bit-identical states and fluxes.
__device__ __forceinline__ float face_state32(float Wh, float g0, float g1, float g2, const float *dr) {
  return __builtin_fmaf(g2, dr[2], __builtin_fmaf(g1, dr[1], __builtin_fmaf(g0, dr[0], Wh)));
}

// a face-end record as the two 16-byte vectors it is loaded as: a record
// carried across loop iterations in this form stays in the registers its
// load writes
struct FeRaw {
  f4v_t a, b;
};
__device__ __forceinline__ FeRaw fe_raw(const FaceEndRec<float> *fe, int slot) {
  const f4v_t *p = reinterpret_cast<const f4v_t *>(fe + slot);
  return FeRaw{p[0], p[1]};
}
__device__ __forceinline__ FaceEndRec<float> fe_rec(const FeRaw &r) {
  FaceEndRec<float> x;
  x.nbr = __float_as_int(r.a.x);
  x.nrm[0] = r.a.y; x.nrm[1] = r.a.z; x.nrm[2] = r.a.w;
  x.drL[0] = r.b.x; x.drL[1] = r.b.y; x.drL[2] = r.b.z;
  x.area = r.b.w;
  return x;
}

template <bool MOVING, bool OFF32>
__global__ void __launch_bounds__(FB) k_flux_apply_coop(FusedArgs<float> A) {
  using T = float;
  __shared__ __attribute__((aligned(256))) float4 sbuf[FB / 64][2][GBUF];
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  float4 *buf0 = sbuf[wv][0], *buf1 = sbuf[wv][1];
  const int lane = (int)__lane_id();
  // LDS byte address of this lane's record in the first buffer
  const unsigned myb = (unsigned)(size_t)(lds_cf4_t *)(buf0 + GQ * lane);
  const int c = xcd_cell(A.n);
  const bool live = c < A.n;
  // a wavefront is 64 consecutive cells of one 64-cell group; lanes past the
  // last cell keep loading for the others but own no face-ends
  const int deg = live ? A.deg[c] : 0;
  int maxdeg = deg;
  for (int o = 32; o > 0; o >>= 1) maxdeg = max(maxdeg, __shfl_xor(maxdeg, o));
  maxdeg = __builtin_amdgcn_readfirstlane(maxdeg);
  if (maxdeg == 0 && !live) return;  // whole wavefront past the last cell (uniform)

  // gathers of steps 0 and 1; every later one is issued by the loop, and the
  // wave waits for all of them after it, so no load into LDS is in flight
  // when the wave exits.  A load whose value a later step needs is issued before the
  // gather that precedes that step and lands in the register that consumes
  // it: no copy of a load in flight (which would wait for it, and for every
  // gather issued after it) is ever made.  rcur holds the face-end record of
  // the step, jn the neighbour of the step after next.  Lanes past their
  // degree load slot 0 (always present), so that every load is unconditional.
  GatherSrc G;
  G.rec = A.rec;
  if constexpr (OFF32)  // raw buffer: stride 0, size in bytes, the gfx9 format word
    G.rs = __builtin_amdgcn_make_buffer_rsrc((void *)A.rec, (short)0, (int)((unsigned)A.n * (unsigned)(REC32 * sizeof(float))),
                                             0x00020000);
  const int e0 = live ? A.gbase[c >> 6] + (c & 63) : 0;
  FeRaw rcur = fe_raw(A.fe, deg > 0 ? e0 : 0);
  // the neighbours of later steps come from the compact neighbour array: 256
  // contiguous bytes per wavefront and step, against 2 KiB of records
  const int j1 = fe_nbr(A.fnb[deg > 1 ? e0 + 64 : 0]);
  int jn = A.fnb[deg > 2 ? e0 + 128 : 0];  // raw, decoded where it is used
  // own state at the half step, extrapolated per face-end along drL / drR.
  // The wavefront's 64 cells are consecutive, so their records are one
  // contiguous 10 KiB block: read through LDS with 10 loads of 1 KiB (lane l of
  // load k: byte 1024k + 16l of the block), then each lane takes its record
  // from 160 l.  Past the last cell the buffer form reads zeros and the global
  // form re-reads the first record.
  T Wc[NF], gc[3 * NF];
  {
    const int cw = c - lane;
#pragma unroll
    for (int k = 0; k < GQ; k++) {
      const size_t off = (size_t)cw * (REC32 * sizeof(float)) + 1024u * k + 16u * (unsigned)lane;
      if constexpr (OFF32) {  // below 2^24 cells the offset fits 32 bits
        __builtin_amdgcn_raw_ptr_buffer_load_lds(G.rs, (lds_void_t *)(buf1 + 64 * k), 16, (int)(unsigned)off, 0, 0, 0);
      } else {
        const bool in = off < (size_t)A.n * (REC32 * sizeof(float));
        const float *src = reinterpret_cast<const float *>(reinterpret_cast<const char *>(A.rec) + (in ? off : 0));
        __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)(buf1 + 64 * k), 16, 0, 0);
      }
    }
  }
  gather_async<OFF32>(G, deg > 0 ? fe_nbr(fe_rec(rcur).nbr) : 0, buf0, lane);
  __builtin_amdgcn_s_waitcnt(WAIT_VMGQ);  // the own records are in LDS (step 0's gather may not be)
  wave_lds_sync();
  {
    const unsigned ob = myb + (unsigned)(GBUF * sizeof(float4));
    T x[4 * GQ];
#pragma unroll
    for (int k = 0; k < GQ; k++) {
      const f4v_t w = *(lds_cf4_t *)(size_t)(ob + 16u * (unsigned)k);
      x[4 * k] = w.x; x[4 * k + 1] = w.y; x[4 * k + 2] = w.z; x[4 * k + 3] = w.w;
    }
    for (int k = 0; k < NF; k++) Wc[k] = x[k];
    for (int k = 0; k < 3 * NF; k++) gc[k] = x[R32_G + k];
  }
  __builtin_amdgcn_s_waitcnt(WAIT_LGKM);  // every lane's reads of the second buffer are done
  wave_lds_sync();
  gather_async<OFF32>(G, deg > 1 ? j1 : 0, buf1, lane);

  const StepTimes st = step_times(A);
  double dU[NF], gp[3], dv = 0.0;
  for (int k = 0; k < NF; k++) dU[k] = 0.0;
  gp[0] = gp[1] = gp[2] = 0.0;
  int nbad = 0;

  // Every step issues a gather, so that the loads after any load of the loop
  // are the same on every path: the last two fetch record 0 into buffers that
  // are read no more.
  for (int q = 0, e = e0; q < maxdeg; q++, e += 64) {
    // step q's records are in LDS: all but the last GQ loads (step q+1's
    // gather, issued after step q's) have completed.  From this wait until its
    // next gather is issued the wavefront runs at raised priority, so that the
    // other wavefront's face arithmetic does not delay its loads.
    __builtin_amdgcn_s_setprio(2);
    __builtin_amdgcn_s_waitcnt(WAIT_VMGQ);
    wave_lds_sync();
    const bool act = q < deg;
    const unsigned rb = myb + (unsigned)(q & 1) * (unsigned)(GBUF * sizeof(float4));
    FE<T> fe;
    decode_fe(fe_rec(rcur), fe);
    T Wn[NF];
    if (act) {
      T x[4 * GQ];
#pragma unroll
      for (int k = 0; k < GQ; k++) {
        const f4v_t w = *(lds_cf4_t *)(size_t)(rb + 16u * (unsigned)k);
        x[4 * k] = w.x; x[4 * k + 1] = w.y; x[4 * k + 2] = w.z; x[4 * k + 3] = w.w;
      }
      const T *dr = fe.isL ? fe.drR : fe.drL;
      for (int k = 0; k < NF; k++)
        Wn[k] = face_state32(x[k], x[R32_G + k], x[R32_G + NF + k], x[R32_G + 2 * NF + k], dr);
    }
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM);  // every lane's reads of the buffer are done
    wave_lds_sync();
    rcur = fe_raw(A.fe, q + 1 < deg ? e + 64 : 0);
    const int jnext = A.fnb[q + 3 < deg ? e + 192 : 0];
    gather_async<OFF32>(G, q + 2 < deg ? fe_nbr(jn) : 0, (q & 1) ? buf1 : buf0, lane);
    __builtin_amdgcn_s_setprio(0);
    jn = jnext;
    if (act) {
      T Wo[NF];
      const T *dr = fe.isL ? fe.drL : fe.drR;
      for (int k = 0; k < NF; k++) Wo[k] = face_state32(Wc[k], gc[k], gc[NF + k], gc[2 * NF + k], dr);
      face_update<T, MOVING>(A, c, fe, Wn, Wo, st.dtf, dU, gp, dv, nbad);
    }
  }
  __builtin_amdgcn_s_waitcnt(WAIT_VM0);  // no load into LDS in flight past here
  wave_add(&A.diag[FVM_DIAG_POSITIVITY], nbad);
  if (live) {
    T tc[NF];  // Wt of the step, for the Powell and Dedner terms
    ldv<NF>(A.tv + (size_t)c * NF, tc);
    cell_update<T>(A, c, st.t0, st.dt_glob, dU, gp, dv, tc);
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

// face-end records from the uploaded mesh, in the internal cell order: internal
// cell i is mesh cell o = perm[i], its list cell_face[off[o]..off[o+1]) in the
// mesh's order; neighbours are stored as internal indices (inv).  drL =
// centroid - pos[s1] and the reconstruction centroid wrap(centroid - pos[o])
// are formed in FP64 exactly as the reference forms them.
template <typename T>
__global__ void k_build_faceends(int n, const int *perm, const int *inv, const int *coff, const int *cface,
                                 const int *s1, const int *s2, const double *fc, const double *fnv, int nface,
                                 const double *pos, double Lx, double Ly, double Lz, const int *gbase,
                                 FaceEndRec<T> *fe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  const double L[3] = {Lx, Ly, Lz};
  for (int e = coff[o]; e < coff[o + 1]; e++) {
    const int slot = gbase[i >> 6] + 64 * (e - coff[o]) + (i & 63);
    const int f = cface[e];
    const int a = s1[f], b = s2[f];
    const bool isL = a == o;
    FaceEndRec<T> r;
    r.nbr = isL ? inv[b] : ~inv[a];
    const double fx = FLD(fnv, 0, f, nface), fy = FLD(fnv, 1, f, nface), fz = FLD(fnv, 2, f, nface);
    const double area = dsqrt(fx * fx + fy * fy + fz * fz);
    const double ia = 1.0 / area;
    const double fn[3] = {fx, fy, fz};
    double cn = 0.0;
    for (int d = 0; d < 3; d++) {
      const double cf = FLD(fc, d, f, nface);
      r.nrm[d] = (T)(fn[d] * ia);
      r.drL[d] = (T)(cf - FLD(pos, d, a, n));
      if constexpr (std::is_same<T, double>::value) {
        double v = cf - FLD(pos, d, o, n);
        if (v > 0.5 * L[d]) v -= L[d];
        else if (v < -0.5 * L[d]) v += L[d];
        r.cen[d] = v;
      }
    }
    if constexpr (std::is_same<T, double>::value) {
      r.pad = 0;
      cn = r.cen[0] * fx + r.cen[1] * fy + r.cen[2] * fz;
      r.area = cn < 0.0 ? -area : area;
    } else {
      r.area = (float)area;
    }
    fe[slot] = r;
  }
}

// FP32: own-side centroid offset of every face-end slot (drL or drR as the
// kernels decode them) and its neighbour field; unused slots of short cells
// keep their offsets unwritten and neighbour 0
__global__ void k_build_dr(int n, const int *deg, const int *gbase, const FaceEndRec<float> *fe, float *fdr, int *fnb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < deg[i]; q++) {
    const int slot = gbase[i >> 6] + 64 * q + (i & 63);
    fnb[slot] = fe[slot].nbr;
    FE<float> g;
    decode_fe(fe[slot], g);
    const float *dr = g.isL ? g.drL : g.drR;
    for (int d = 0; d < 3; d++) fdr[3 * (size_t)slot + d] = dr[d];
  }
}

// the cell length scale of the CFL step and the Dedner damping,
// pow(V / (4 pi / 3), 1 / 3) (computeTimestep.cpp:17, computePredictor.cpp:319,
// computeFluidUpdate.cpp:451): the volume is fixed on this path, so it is
// formed once
__global__ void k_cell_lmin(const double *cvol, double *lmin, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lmin[i] = pow(cvol[i] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
}

// SoA fields between the mesh order and the internal order
__global__ void k_gather_soa(const double *src, double *dst, const int *perm, int n, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  for (int k = 0; k < nk; k++) SF(dst, k, i, n) = SF(src, k, o, n);
}
__global__ void k_scatter_soa(const double *src, double *dst, const int *perm, int n, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  for (int k = 0; k < nk; k++) SF(dst, k, o, n) = SF(src, k, i, n);
}

// W field (SoA FP64) -> W slice of the records; records -> fields
template <typename T> __global__ void k_rec_from_W(const double *W, T *rec, T *wc, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  T *r = rec + (size_t)c * rec_len<T>();
  T *w = wc + (size_t)c * RWC;
  for (int k = 0; k < rec_len<T>(); k++) r[k] = T(0.0);
  for (int k = 0; k < RWC; k++) w[k] = T(0.0);
  for (int k = 0; k < NF; k++) r[k] = w[k] = (T)SF(W, k, c, n);
}
// ten values per cell, `stride` apart, at offset `off` -> a mesh-order field
template <typename T, typename S>
__global__ void k_rec_to_field(const T *rec, int stride, int off, S *dst, const int *perm, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int o = perm[c];
  for (int k = 0; k < NF; k++) SF(dst, k, o, n) = (S)rec[(size_t)c * stride + off + k];
}

// ===========================================================================
// host side
// ===========================================================================
constexpr int SIGMA = 256;  // cells per sorting window (4 wavefronts)

unsigned fused_grid(int n) {
  unsigned nb = (unsigned)((n + FB - 1) / FB);
  return (nb + 7u) & ~7u;  // multiple of 8 for the XCD mapping
}

template <typename T> FusedArgs<T> fused_args(fvm_ctx *c, double dtmax) {
  FusedState &z = c->fz;
  FusedArgs<T> A;
  A.n = c->np;
  A.deg = (const int *)z.deg.p;
  A.gbase = (const int *)z.gbase.p;
  A.fe = (const FaceEndRec<T> *)z.fe.p;
  A.fdr = (const float *)z.fdr.p;
  A.fnb = (const int *)z.fnb.p;
  A.cvol = (const double *)z.cvol.p;
  A.lmin = (const double *)z.lmin.p;
  A.mvel = (const double *)z.mvel.p;
  A.rec = (const T *)z.rec[z.cur].p;
  A.rec_w = (T *)z.rec[z.cur].p;
  A.rec_next = (T *)z.rec[z.cur ^ 1].p;
  A.wcur = (const T *)z.wc[z.cur].p;
  A.wnext = (T *)z.wc[z.cur ^ 1].p;
  A.tv = (T *)z.tv.p;
  A.W64 = (double *)z.W64.p;
  A.U = (double *)z.U.p;
  A.mm = (T *)z.mm.p;
  A.dt_new = (double *)z.dt_new.p;
  A.dtmin_bits = (unsigned long long *)z.dtmin.p;
  A.tdev = (double *)z.tdev.p;
  A.diag = (unsigned long long *)c->diag.p;
  A.gamma = c->prm.gamma_gas;
  A.courant = c->prm.courant_no;
  A.dtmax = dtmax;
  A.write_w64 = 0;
  A.k1_pass = 0;
  return A;
}

// internal order: inside each window of SIGMA cells, decreasing face count
// (ties keep the mesh order); perm[i] = mesh cell of internal cell i
void window_sort(const std::vector<int> &off, int n, std::vector<int> &perm, std::vector<int> &inv) {
  perm.resize(n);
  inv.resize(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int w = 0; w < n; w += SIGMA) {
    const int e = std::min(n, w + SIGMA);
    std::stable_sort(perm.begin() + w, perm.begin() + e,
                     [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
  }
  for (int i = 0; i < n; i++) inv[perm[i]] = i;
}

template <typename T> int fused_setup_t(fvm_ctx *c) {
  FusedState &z = c->fz;
  const int n = c->np;
  z.E = c->nref;
  int rc;
  const size_t es = sizeof(T);
  std::vector<int> off(n + 1), perm, inv;
  FVM_HIPCHK(c, hipMemcpy(off.data(), c->m_coff.p, sizeof(int) * (size_t)(n + 1), hipMemcpyDeviceToHost));
  window_sort(off, n, perm, inv);
  std::vector<int> deg(n);
  for (int i = 0; i < n; i++) deg[i] = off[perm[i] + 1] - off[perm[i]];
  // SELL slots: each 64-cell group gets 64 x (largest face count in the group) records
  const int ng = (n + 63) / 64;
  std::vector<int> gb(ng + 1, 0);
  long long acc = 0;
  for (int g = 0; g < ng; g++) {
    int mx = 0;
    for (int i = g * 64; i < std::min(n, g * 64 + 64); i++) mx = std::max(mx, deg[i]);
    acc += 64LL * mx;
    if (acc > 0x7fffffffLL) return fvm_fail(c, FVM_EARG, "fused path: face-end slots exceed 2^31");
    gb[g + 1] = (int)acc;
  }
  z.slots = (size_t)gb[ng];
  const size_t ni = sizeof(int) * (size_t)n, nd = sizeof(double) * (size_t)n;
  if ((rc = fvm_ensure(c, z.gbase, sizeof(int) * (size_t)(ng + 1))) || (rc = fvm_ensure(c, z.perm, ni)) ||
      (rc = fvm_ensure(c, z.inv, ni)) || (rc = fvm_ensure(c, z.deg, ni)) ||
      (rc = fvm_ensure(c, z.fe, sizeof(FaceEndRec<T>) * z.slots)) || (rc = fvm_ensure(c, z.U, NF * nd)) ||
      (rc = fvm_ensure(c, z.W64, NF * nd)) || (rc = fvm_ensure(c, z.cvol, nd)) ||
      (rc = fvm_ensure(c, z.mvel, 3 * nd)) || (rc = fvm_ensure(c, z.rec[0], es * rec_len<T>() * (size_t)n)) ||
      (rc = fvm_ensure(c, z.rec[1], es * rec_len<T>() * (size_t)n)) ||
      (rc = fvm_ensure(c, z.tv, es * NF * (size_t)(es == 4 ? n : 1)))  /* FP32 path only */ || (rc = fvm_ensure(c, z.mm, es * 20 * (size_t)(es == 8 ? n : 1)))  /* FP64 path only */ ||
      (rc = fvm_ensure(c, z.wc[0], es * RWC * (size_t)n)) || (rc = fvm_ensure(c, z.wc[1], es * RWC * (size_t)n)) ||
      (rc = fvm_ensure(c, z.dt_new, nd)) || (rc = fvm_ensure(c, z.dtmin, sizeof(unsigned long long))) ||
      (rc = fvm_ensure(c, z.tdev, sizeof(double) * 2)))
    return rc;
  FVM_HIPCHK(c, hipMemcpy(z.gbase.p, gb.data(), sizeof(int) * (size_t)(ng + 1), hipMemcpyHostToDevice));
  FVM_HIPCHK(c, hipMemcpy(z.perm.p, perm.data(), ni, hipMemcpyHostToDevice));
  FVM_HIPCHK(c, hipMemcpy(z.inv.p, inv.data(), ni, hipMemcpyHostToDevice));
  FVM_HIPCHK(c, hipMemcpy(z.deg.p, deg.data(), ni, hipMemcpyHostToDevice));
  const unsigned nb = (unsigned)((n + 255) / 256);
  const int *pp = (const int *)z.perm.p;
  hipLaunchKernelGGL(k_build_faceends<T>, dim3(nb), dim3(256), 0, c->stream, n, pp, (const int *)z.inv.p,
                     (const int *)c->m_coff.p, (const int *)c->m_cface.p, (const int *)c->m_s1.p,
                     (const int *)c->m_s2.p, (const double *)c->m_c.p, (const double *)c->m_n.p, c->nface,
                     (const double *)c->field[FVM_FIELD_POS].p, c->prm.domain_size[0], c->prm.domain_size[1],
                     c->prm.domain_size[2], (const int *)z.gbase.p, (FaceEndRec<T> *)z.fe.p);
  FVM_HIPCHK(c, hipGetLastError());
  if constexpr (std::is_same<T, float>::value) {
    if ((rc = fvm_ensure(c, z.fdr, 3 * sizeof(float) * z.slots)) || (rc = fvm_ensure(c, z.fnb, sizeof(int) * z.slots)))
      return rc;
    FVM_HIPCHK(c, hipMemsetAsync(z.fnb.p, 0, sizeof(int) * z.slots, c->stream));
    hipLaunchKernelGGL(k_build_dr, dim3(nb), dim3(256), 0, c->stream, n, (const int *)z.deg.p,
                       (const int *)z.gbase.p, (const FaceEndRec<float> *)z.fe.p, (float *)z.fdr.p, (int *)z.fnb.p);
    FVM_HIPCHK(c, hipGetLastError());
  }
  struct {
    const void *src;
    DBuf *dst;
    int nk;
  } g[4] = {{c->field[FVM_FIELD_U].p, &z.U, NF},
            {c->field[FVM_FIELD_W].p, &z.W64, NF},
            {c->m_cvol.p, &z.cvol, 1},
            {c->field[FVM_FIELD_MESH_VEL].p, &z.mvel, 3}};
  for (auto &q : g) {
    hipLaunchKernelGGL(k_gather_soa, dim3(nb), dim3(256), 0, c->stream, (const double *)q.src, (double *)q.dst->p, pp,
                       n, q.nk);
    FVM_HIPCHK(c, hipGetLastError());
  }
  if ((rc = fvm_ensure(c, z.lmin, nd))) return rc;
  hipLaunchKernelGGL(k_cell_lmin, dim3(nb), dim3(256), 0, c->stream, (const double *)z.cvol.p, (double *)z.lmin.p, n);
  FVM_HIPCHK(c, hipGetLastError());
  for (int q = 0; q < 2; q++) {
    hipLaunchKernelGGL(k_rec_from_W<T>, dim3(nb), dim3(256), 0, c->stream, (const double *)z.W64.p, (T *)z.rec[q].p,
                       (T *)z.wc[q].p, n);
    FVM_HIPCHK(c, hipGetLastError());
  }
  const unsigned long long inf = 0x7FF0000000000000ULL;
  const double t2[2] = {c->t_global, 0.0};
  FVM_HIPCHK(c, hipMemcpyAsync(z.dtmin.p, &inf, sizeof inf, hipMemcpyHostToDevice, c->stream));
  FVM_HIPCHK(c, hipMemcpyAsync(z.tdev.p, t2, sizeof t2, hipMemcpyHostToDevice, c->stream));
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  z.cur = 0;
  return FVM_OK;
}

// the surrogate flux step of the FP32 path (fvm_sg_step.h, FP32 translation unit only)
int sg_step_launch(fvm_ctx *c, const FusedArgs<float> &A);

template <typename T, bool MOVING, int LIM> int launch_step(fvm_ctx *c, const FusedArgs<T> &A, hipEvent_t *ev) {
  const unsigned grid = fused_grid(c->np);
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  hipLaunchKernelGGL((k_cfl_predict_recon<T, MOVING, LIM>), dim3(grid), dim3(FB), 0, c->stream, A);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if constexpr (std::is_same<T, float>::value) {
    FusedArgs<T> A2 = A;  // the check-and-redo pass of the step's dt
    A2.k1_pass = 1;
    hipLaunchKernelGGL((k_cfl_predict_recon<T, MOVING, LIM>), dim3(grid), dim3(FB), 0, c->stream, A2);
  } else if constexpr (LIM != LIM_SHIPPED) {
    hipLaunchKernelGGL((k_limit<T, LIM>), dim3(grid), dim3(FB), 0, c->stream, A);
  }
  if (ev) (void)hipEventRecord(ev[2], c->stream);
  if constexpr (std::is_same<T, float>::value) {
    // FVM_FUSED_ADDR64 set: always the 64-bit addressing (it serves meshes past
    // 2^24 cells; the switch lets small meshes exercise it)
    const bool addr64 = std::getenv("FVM_FUSED_ADDR64") != nullptr;
    if (c->fz.sg) {
      if constexpr (!MOVING) {
        const int rc = sg_step_launch(c, A);
        if (rc) return rc;
      }
    } else if (!addr64 && c->np <= (1 << 24))  // 24-bit record indices, < 2.7 GB of records
      hipLaunchKernelGGL((k_flux_apply_coop<MOVING, true>), dim3(grid), dim3(FB), 0, c->stream, A);
    else
      hipLaunchKernelGGL((k_flux_apply_coop<MOVING, false>), dim3(grid), dim3(FB), 0, c->stream, A);
  }
  else
    hipLaunchKernelGGL((k_flux_apply<T, MOVING>), dim3(grid), dim3(FB), 0, c->stream, A);
  if (ev) (void)hipEventRecord(ev[3], c->stream);
  return FVM_OK;
}

template <typename T, bool MOVING> int fused_run_t(fvm_ctx *c, int nsteps, double dtmax) {
  FusedState &z = c->fz;
  const int lim = c->prm.limiter_mode;
  if (z.timing)
    while ((int)z.ev.size() < 5 * nsteps) {
      hipEvent_t e;
      FVM_HIPCHK(c, hipEventCreate(&e));
      z.ev.push_back(e);
    }
  for (int s = 0; s < nsteps; s++) {
    FusedArgs<T> A = fused_args<T>(c, dtmax);
    A.write_w64 = (s == nsteps - 1) ? 1 : 0;
    hipEvent_t *ev = z.timing ? &z.ev[5 * s] : nullptr;
    int rc;
    if (lim == LIM_MINMOD) rc = launch_step<T, MOVING, LIM_MINMOD>(c, A, ev);
    else if (lim == LIM_SMOOTH) rc = launch_step<T, MOVING, LIM_SMOOTH>(c, A, ev);
    else rc = launch_step<T, MOVING, LIM_SHIPPED>(c, A, ev);
    if (rc) return rc;
    FVM_HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, (double *)z.tdev.p, (unsigned long long *)z.dtmin.p,
                       dtmax, (double *)z.hist.p, s);
    FVM_HIPCHK(c, hipGetLastError());
    if (ev) FVM_HIPCHK(c, hipEventRecord(ev[4], c->stream));
    z.cur ^= 1;
  }
  if (nsteps > 0) {
    // the run's result in the mesh-order fields
    const unsigned nb = (unsigned)((c->np + 255) / 256);
    const int *pp = (const int *)z.perm.p;
    hipLaunchKernelGGL(k_scatter_soa, dim3(nb), dim3(256), 0, c->stream, (const double *)z.U.p,
                       (double *)c->field[FVM_FIELD_U].p, pp, c->np, NF);
    hipLaunchKernelGGL(k_scatter_soa, dim3(nb), dim3(256), 0, c->stream, (const double *)z.W64.p,
                       (double *)c->field[FVM_FIELD_W].p, pp, c->np, NF);
    FVM_HIPCHK(c, hipGetLastError());
  }
  if (z.timing && nsteps > 0) {
    FVM_HIPCHK(c, hipEventSynchronize(z.ev[5 * nsteps - 1]));
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < 4; k++) {
        float ms = 0.0f;
        FVM_HIPCHK(c, hipEventElapsedTime(&ms, z.ev[5 * s + k], z.ev[5 * s + k + 1]));
        z.ms[k] += ms;
      }
    for (int k = 0; k < 4; k++) z.launches[k] += (k == 1 && lim == LIM_SHIPPED) ? 0 : nsteps;
  }
  return FVM_OK;
}

// gradients, Wt and dt_new of the last step into the mesh-order fields
template <typename T, typename S> int fused_sync_t(fvm_ctx *c) {
  FusedState &z = c->fz;
  const int n = c->np;
  const unsigned nb = (unsigned)((n + 255) / 256);
  const int *pp = (const int *)z.perm.p;
  // the last completed step's gradients sit in the records it read (z.cur ^ 1)
  const T *rec = (const T *)z.rec[z.cur ^ 1].p;
  const int dstf[4] = {FVM_FIELD_GRAD_X, FVM_FIELD_GRAD_Y, FVM_FIELD_GRAD_Z, FVM_FIELD_DWDT};
  constexpr int G = sizeof(T) == 4 ? R32_G : R_G;
  const int off[4] = {G, G + NF, G + 2 * NF, R_T};
  for (int q = 0; q < 4; q++) {
    // FP32: Wt sits in its own array
    const bool own = sizeof(T) == 4 && q == 3;
    hipLaunchKernelGGL((k_rec_to_field<T, S>), dim3(nb), dim3(256), 0, c->stream, own ? (const T *)z.tv.p : rec,
                       own ? NF : rec_len<T>(), own ? 0 : off[q], (S *)c->field[dstf[q]].p, pp, n);
    FVM_HIPCHK(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_scatter_soa, dim3(nb), dim3(256), 0, c->stream, (const double *)z.dt_new.p,
                     (double *)c->field[FVM_FIELD_DT_NEW].p, pp, n, 1);
  FVM_HIPCHK(c, hipGetLastError());
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

}  // namespace
}  // namespace fvm
