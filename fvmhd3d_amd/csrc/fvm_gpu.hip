// HIP implementation of the fvm_gpu.h C ABI: the reference-order ("parity")
// kernels of the Voronoi-face pipeline and the context that owns their device
// buffers.
//
// Device layout: every per-cell field is structure-of-arrays, component k of
// cell i at [k*np + i]; face records likewise [k*nface_act + a].  Fields the
// reference stores as float (Fluid_flt, fvec3, FluidD) use the storage type S
// (float as shipped, double for the all-FP64 build); everything else is double.
//
// Sums that the reference accumulates face by face (flux scatter, area
// closure, predictor) run per cell over a CSR list of its active face-ends
// sorted by face_active_list position, i.e. in the reference's own order, so
// the results do not depend on scheduling (no atomics on data).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"
#include "fvm_physics.h"

using namespace fvm;

namespace fvm {

// ---------------------------------------------------------------------------
// field table
// ---------------------------------------------------------------------------
enum { K_DOUBLE = 0, K_STORE = 1, K_INT = 2 };
struct FieldInfo {
  int width, kind;
};
static const FieldInfo kFields[FVM_NFIELDS] = {
    {10, K_DOUBLE},  // W
    {10, K_STORE},   // GRAD_X
    {10, K_STORE},   // GRAD_Y
    {10, K_STORE},   // GRAD_Z
    {10, K_STORE},   // DWDT
    {3, K_DOUBLE},   // POS
    {3, K_STORE},    // VEL
    {3, K_STORE},    // J
    {1, K_DOUBLE},   // TEND
    {1, K_INT},      // BND
    {1, K_INT},      // IS_NGB
    {1, K_INT},      // IS_ACTIVE
    {10, K_STORE},   // MINMAX_LO
    {10, K_STORE},   // MINMAX_HI
    {10, K_STORE},   // LIMITER
    {10, K_STORE},   // DU
    {3, K_STORE},    // DGRADPSI
    {1, K_STORE},    // DDIVB
    {10, K_DOUBLE},  // U
    {1, K_DOUBLE},   // MESH_VOLUME
    {1, K_DOUBLE},   // TBEG
    {3, K_DOUBLE},   // MESH_VEL
    {3, K_DOUBLE},   // ACC0
    {3, K_DOUBLE},   // ACC1
    {1, K_DOUBLE},   // DT_NEW
    {1, K_DOUBLE},   // DIVB
};

template <typename S> struct Cells {
  int np, nactive;
  double *w;
  S *x, *y, *z, *t;
  double *pos;
  S *vel, *J;
  double *tend;
  int *bnd, *is_ngb, *is_active;
  S *mm_lo, *mm_hi, *lim, *dU, *dgp, *ddivB;
  double *U, *mvol, *tbeg, *mvel, *acc0, *acc1, *dt_new, *divB;
};

struct MeshView {
  int nface, nfa;
  const int *s1, *s2;
  const double *c, *n;  // [3][nface]
  const double *cn;     // [nface][6]: centroid, n (the same values per face)
  const int *coff, *cface;
  const double *cvol;
  const int *fact;        // face_active_list
  const int *soff, *sent; // per-cell active face-ends, entry = a<<1 | side
};

struct Phys {
  double gamma, courant, t_global;
  double L[3];
  int limiter;
};

#define FLD(p, k, i, n) (p)[(size_t)(k) * (size_t)(n) + (size_t)(i)]
// per-face records of the flux and predictor sweeps, one record of `stride`
// doubles per active face: the cell-side loops read a face's whole record from
// one or two cache lines
#define REC(p, k, a, stride) (p)[(size_t)(a) * (size_t)(stride) + (size_t)(k)]
// per-face geometry record: centroid[3] then area vector n[3]
#define GEO_C(M, d, f) (M).cn[(size_t)(f) * 6 + (size_t)(d)]
#define GEO_N(M, d, f) (M).cn[(size_t)(f) * 6 + 3 + (size_t)(d)]

__device__ __forceinline__ int gid() { return blockIdx.x * blockDim.x + threadIdx.x; }
// XCD-aware index of the cell and face sweeps: block b runs on XCD b mod 8
// and the 8 XCDs each take a contiguous eighth of the range, so neighbour
// gathers of nearby (Morton-ordered) cells share one XCD's L2.  gridDim.x is
// a multiple of 8 (LAUNCH_X); every thread still computes exactly one item.
__device__ __forceinline__ int xgid() {
  const int per = (int)gridDim.x >> 3;
  const int tile = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
  return tile * (int)blockDim.x + (int)threadIdx.x;
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
template <typename S>
__global__ void k_aos_to_soa(const double *__restrict__ src, S *__restrict__ dst, int n, int width) {
  const int i = gid();
  if (i >= n) return;
  for (int k = 0; k < width; k++) FLD(dst, k, i, n) = (S)src[(size_t)i * width + k];
}
template <typename S>
__global__ void k_soa_to_aos(const S *__restrict__ src, double *__restrict__ dst, int n, int width) {
  const int i = gid();
  if (i >= n) return;
  for (int k = 0; k < width; k++) dst[(size_t)i * width + k] = (double)FLD(src, k, i, n);
}
__global__ void k_faces_to_soa(const fvm_face *__restrict__ f, int n, int *s1, int *s2, double *c,
                               double *nn, double *cn) {
  const int i = gid();
  if (i >= n) return;
  const fvm_face F = f[i];
  s1[i] = F.s1;
  s2[i] = F.s2;
  for (int d = 0; d < 3; d++) {
    FLD(c, d, i, n) = F.centroid[d];
    FLD(nn, d, i, n) = F.n[d];
    cn[(size_t)i * 6 + d] = F.centroid[d];
    cn[(size_t)i * 6 + 3 + d] = F.n[d];
  }
}

// ---------------------------------------------------------------------------
// System::computeReconstruction (computeReconstruction.cpp:42-264).  Two forms
// of the cell loops here (reconstruction, slope limiter, flux scatter,
// predictor cell update), with the same values bit for bit:
//   k_<name>    one thread per cell, all variables in registers;
//   k_<name>_v  one lane per (cell, variable): NF lanes per cell, CPW cells per
//               wavefront (the last 64 - NF * CPW lanes idle).  Every lane of a
//               cell repeats the face geometry and forms its own variable's
//               terms in the reference's face order and operation order.
// The per-variable form fills the chip on the small active sets of individual
// timesteps, where one thread per cell leaves most SIMDs idle; on large sets
// it repeats more geometry and touches more cache lines, and one thread per
// cell is faster (ONE_THREAD_MIN_CELLS below).
// ---------------------------------------------------------------------------
constexpr int CPW = 64 / NF;  // cells per wavefront of the per-variable kernels
// the cell and variable of this lane; false for the idle lanes
__device__ __forceinline__ bool cell_var(int &cell, int &k) {
  const int t = xgid();
  const int sub = (t & 63) / NF;
  k = (t & 63) - NF * sub;
  cell = (t >> 6) * CPW + sub;
  return sub < CPW;
}
// threads of a per-variable launch over n cells
__host__ __device__ constexpr int cell_var_threads(int n) { return (n + CPW - 1) / CPW * 64; }
// cell count from which the thread-per-cell forms are used
constexpr int ONE_THREAD_MIN_CELLS = 131072;

template <typename S>
__global__ void __launch_bounds__(256) k_reconstruction(Cells<S> C, MeshView M, Phys P) {
  const int i = xgid();
  if (i >= C.nactive) return;
  const int np = C.np;
  double Wi[NF];
  for (int k = 0; k < NF; k++) Wi[k] = FLD(C.w, k, i, np);
  const double ipos[3] = {FLD(C.pos, 0, i, np), FLD(C.pos, 1, i, np), FLD(C.pos, 2, i, np)};

  if (C.bnd[i] != NO_BOUNDARY) {
    for (int k = 0; k < NF; k++) {
      FLD(C.x, k, i, np) = (S)0.0;
      FLD(C.y, k, i, np) = (S)0.0;
      FLD(C.z, k, i, np) = (S)0.0;
      FLD(C.t, k, i, np) = (S)0.0;
    }
    return;
  }

  double Wx[NF], Wy[NF], Wz[NF];
  S dmin[NF], dmax[NF];
  for (int k = 0; k < NF; k++) {
    Wx[k] = Wy[k] = Wz[k] = 0.0;
    dmin[k] = dmax[k] = (S)0.0;
  }
  const int fb = M.coff[i], fe = M.coff[i + 1];
  bool zero_flag = false;
  const double dtI = C.tend[i] - P.t_global;

  auto face_dri = [&](int f, double *dri) {
    for (int d = 0; d < 3; d++) {
      double v = GEO_C(M, d, f) - ipos[d];
      if (v > 0.5 * P.L[d]) v -= P.L[d];
      else if (v < -0.5 * P.L[d]) v += P.L[d];
      dri[d] = v;
    }
  };

  for (int e = fb; e < fe; e++) {
    const int f = M.cface[e];
    double cen[3];
    face_dri(f, cen);
    const double fn[3] = {GEO_N(M, 0, f), GEO_N(M, 1, f), GEO_N(M, 2, f)};
    const double area = dsqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
    const double cn = cen[0] * fn[0] + cen[1] * fn[1] + cen[2] * fn[2];
    const double sc = (cn < 0.0) ? (-1.0 / area) : (1.0 / area);
    const double normal[3] = {fn[0] * sc, fn[1] * sc, fn[2] * sc};
    const double dsh = cen[0] * normal[0] + cen[1] * normal[1] + cen[2] * normal[2];
    const double ids = (dsh != 0.0) ? 0.5 / dsh : 0.0;
    const double idsA = area * ids;
    const double drh[3] = {normal[0] * dsh, normal[1] * dsh, normal[2] * dsh};
    const double fij[3] = {cen[0] - drh[0], cen[1] - drh[1], cen[2] - drh[2]};
    const int s1 = M.s1[f];
    const int j = (i == s1) ? M.s2[f] : s1;
    const int bj = C.bnd[j];
    if (bj != NO_BOUNDARY && bj != INFLOW) {
      zero_flag = true;
      break;
    }
    for (int k = 0; k < NF; k++) {
      const double wj = FLD(C.w, k, j, np);
      const double sum = wj + Wi[k];
      const double diff = wj - Wi[k];
      dmin[k] = (S)smin((double)dmin[k], diff);
      dmax[k] = (S)smax((double)dmax[k], diff);
      Wx[k] += (sum * drh[0] + diff * fij[0]) * idsA;
      Wy[k] += (sum * drh[1] + diff * fij[1]) * idsA;
      Wz[k] += (sum * drh[2] + diff * fij[2]) * idsA;
    }
  }

  if (zero_flag) {
    for (int k = 0; k < NF; k++) {
      FLD(C.x, k, i, np) = (S)0.0;
      FLD(C.y, k, i, np) = (S)0.0;
      FLD(C.z, k, i, np) = (S)0.0;
      FLD(C.t, k, i, np) = (S)0.0;
    }
    return;
  }

  const double invV = 1.0 / M.cvol[i];
  for (int k = 0; k < NF; k++) {
    Wx[k] *= invV;
    Wy[k] *= invV;
    Wz[k] *= invV;
  }
  FLD(C.J, 0, i, np) = (S)(Wy[BZ] - Wz[BY]);
  FLD(C.J, 1, i, np) = (S)(Wz[BX] - Wx[BZ]);
  FLD(C.J, 2, i, np) = (S)(Wx[BY] - Wy[BX]);

  double limiter_min[NF];
  for (int k = 0; k < NF; k++) limiter_min[k] = 1.0;
  if (P.limiter != LIM_SHIPPED) {
    // the shipped Phi() returns 1 for every argument, so the face loop is
    // only evaluated for the opt-in limiters
    for (int e = fb; e < fe; e++) {
      const int f = M.cface[e];
      double cen[3];
      face_dri(f, cen);
      const int s1 = M.s1[f];
      const int j = (i == s1) ? M.s2[f] : s1;
      const double dtJ = C.tend[j] - P.t_global;
      const double dtij = dtI - 0.5 * smin(dtI, dtJ);
      for (int k = 0; k < NF; k++) {
        const double dWx = Wx[k] * cen[0] + Wy[k] * cen[1] + Wz[k] * cen[2];
        const double dWt = 0.0 * (double)FLD(C.t, k, i, np) * dtij;
        const double dWi = dWx + dWt;
        limiter_min[k] = smin(limiter_min[k], limiter_rec(P.limiter, dWi, (double)dmin[k], (double)dmax[k]));
      }
    }
  }
  for (int k = 0; k < NF; k++) {
    FLD(C.mm_lo, k, i, np) = (S)((double)dmin[k] + Wi[k]);
    FLD(C.mm_hi, k, i, np) = (S)((double)dmax[k] + Wi[k]);
    const double tau = 1.0 * limiter_min[k];
    FLD(C.t, k, i, np) = (S)((double)FLD(C.t, k, i, np) * tau);
    FLD(C.x, k, i, np) = (S)(Wx[k] * tau);
    FLD(C.y, k, i, np) = (S)(Wy[k] * tau);
    FLD(C.z, k, i, np) = (S)(Wz[k] * tau);
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_reconstruction_v(Cells<S> C, MeshView M, Phys P) {
  int i, k;
  if (!cell_var(i, k) || i >= C.nactive) return;
  const int np = C.np;
  const int base = (int)__lane_id() - k;  // the lane of this cell's variable 0
  const double Wi = FLD(C.w, k, i, np);
  const double ipos[3] = {FLD(C.pos, 0, i, np), FLD(C.pos, 1, i, np), FLD(C.pos, 2, i, np)};

  if (C.bnd[i] != NO_BOUNDARY) {
    FLD(C.x, k, i, np) = (S)0.0;
    FLD(C.y, k, i, np) = (S)0.0;
    FLD(C.z, k, i, np) = (S)0.0;
    FLD(C.t, k, i, np) = (S)0.0;
    return;
  }

  double Wx = 0.0, Wy = 0.0, Wz = 0.0;
  S dmin = (S)0.0, dmax = (S)0.0;
  const int fb = M.coff[i], fe = M.coff[i + 1];
  bool zero_flag = false;
  const double dtI = C.tend[i] - P.t_global;

  auto face_dri = [&](int f, double *dri) {
    for (int d = 0; d < 3; d++) {
      double v = GEO_C(M, d, f) - ipos[d];
      if (v > 0.5 * P.L[d]) v -= P.L[d];
      else if (v < -0.5 * P.L[d]) v += P.L[d];
      dri[d] = v;
    }
  };

  for (int e = fb; e < fe; e++) {
    const int f = M.cface[e];
    double cen[3];
    face_dri(f, cen);
    const double fn[3] = {GEO_N(M, 0, f), GEO_N(M, 1, f), GEO_N(M, 2, f)};
    const double area = dsqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
    const double cn = cen[0] * fn[0] + cen[1] * fn[1] + cen[2] * fn[2];
    const double sc = (cn < 0.0) ? (-1.0 / area) : (1.0 / area);
    const double normal[3] = {fn[0] * sc, fn[1] * sc, fn[2] * sc};
    const double dsh = cen[0] * normal[0] + cen[1] * normal[1] + cen[2] * normal[2];
    const double ids = (dsh != 0.0) ? 0.5 / dsh : 0.0;
    const double idsA = area * ids;
    const double drh[3] = {normal[0] * dsh, normal[1] * dsh, normal[2] * dsh};
    const double fij[3] = {cen[0] - drh[0], cen[1] - drh[1], cen[2] - drh[2]};
    const int s1 = M.s1[f];
    const int j = (i == s1) ? M.s2[f] : s1;
    const int bj = C.bnd[j];
    if (bj != NO_BOUNDARY && bj != INFLOW) {
      zero_flag = true;
      break;
    }
    const double wj = FLD(C.w, k, j, np);
    const double sum = wj + Wi;
    const double diff = wj - Wi;
    dmin = (S)smin((double)dmin, diff);
    dmax = (S)smax((double)dmax, diff);
    Wx += (sum * drh[0] + diff * fij[0]) * idsA;
    Wy += (sum * drh[1] + diff * fij[1]) * idsA;
    Wz += (sum * drh[2] + diff * fij[2]) * idsA;
  }

  if (zero_flag) {
    FLD(C.x, k, i, np) = (S)0.0;
    FLD(C.y, k, i, np) = (S)0.0;
    FLD(C.z, k, i, np) = (S)0.0;
    FLD(C.t, k, i, np) = (S)0.0;
    return;
  }

  const double invV = 1.0 / M.cvol[i];
  Wx *= invV;
  Wy *= invV;
  Wz *= invV;
  // J = curl B from the B lanes of this cell
  const double yBZ = __shfl(Wy, base + BZ), zBY = __shfl(Wz, base + BY);
  const double zBX = __shfl(Wz, base + BX), xBZ = __shfl(Wx, base + BZ);
  const double xBY = __shfl(Wx, base + BY), yBX = __shfl(Wy, base + BX);
  if (k < 3) FLD(C.J, k, i, np) = (S)(k == 0 ? yBZ - zBY : k == 1 ? zBX - xBZ : xBY - yBX);

  double limiter_min = 1.0;
  if (P.limiter != LIM_SHIPPED) {
    // the shipped Phi() returns 1 for every argument, so the face loop is
    // only evaluated for the opt-in limiters
    for (int e = fb; e < fe; e++) {
      const int f = M.cface[e];
      double cen[3];
      face_dri(f, cen);
      const int s1 = M.s1[f];
      const int j = (i == s1) ? M.s2[f] : s1;
      const double dtJ = C.tend[j] - P.t_global;
      const double dtij = dtI - 0.5 * smin(dtI, dtJ);
      const double dWx = Wx * cen[0] + Wy * cen[1] + Wz * cen[2];
      const double dWt = 0.0 * (double)FLD(C.t, k, i, np) * dtij;
      const double dWi = dWx + dWt;
      limiter_min = smin(limiter_min, limiter_rec(P.limiter, dWi, (double)dmin, (double)dmax));
    }
  }
  FLD(C.mm_lo, k, i, np) = (S)((double)dmin + Wi);
  FLD(C.mm_hi, k, i, np) = (S)((double)dmax + Wi);
  const double tau = 1.0 * limiter_min;
  FLD(C.t, k, i, np) = (S)((double)FLD(C.t, k, i, np) * tau);
  FLD(C.x, k, i, np) = (S)(Wx * tau);
  FLD(C.y, k, i, np) = (S)(Wy * tau);
  FLD(C.z, k, i, np) = (S)(Wz * tau);
}

// ---------------------------------------------------------------------------
// slopeLimiter compute part (slopeLimiter.cpp:27-237) per cell over its active
// face-ends, then computeFluidUpdateI (computeFluidUpdate.cpp:25-75)
// ---------------------------------------------------------------------------
struct FaceGeo {
  double area, normal[3], dri[3], drj[3], dsh, dsl;
};
__device__ __forceinline__ void face_geo(const MeshView &M, const double *pos_i, int f, int np,
                                         FaceGeo &g) {
  (void)np;
  const double fn[3] = {GEO_N(M, 0, f), GEO_N(M, 1, f), GEO_N(M, 2, f)};
  g.area = dsqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const double ia = 1.0 / g.area;
  for (int d = 0; d < 3; d++) {
    g.normal[d] = fn[d] * ia;
    g.dri[d] = GEO_C(M, d, f) - pos_i[d];
  }
  g.dsh = g.dri[0] * g.normal[0] + g.dri[1] * g.normal[1] + g.dri[2] * g.normal[2];
  g.dsl = 2.0 * g.dsh;
  for (int d = 0; d < 3; d++) g.drj[d] = g.dri[d] - g.normal[d] * g.dsl;
}

template <typename S>
__global__ void __launch_bounds__(256) k_slope_limiter(Cells<S> C, MeshView M, Phys P, unsigned long long *diag) {
  const int c = xgid();
  if (c >= C.np) return;
  const int np = C.np;
  S lim[NF];
  for (int k = 0; k < NF; k++) lim[k] = (S)1.0;
  double at[3] = {0.0, 0.0, 0.0};
  double wc[NF], mlo[NF], mhi[NF];
  for (int k = 0; k < NF; k++) {
    wc[k] = FLD(C.w, k, c, np);
    mlo[k] = (double)FLD(C.mm_lo, k, c, np);
    mhi[k] = (double)FLD(C.mm_hi, k, c, np);
  }
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const int f = M.fact[a];
    const int i = M.s1[f], j = M.s2[f];
    const double dtI = C.tend[i] - P.t_global;
    const double dtJ = C.tend[j] - P.t_global;
    const double dth = 0.5 * smin(dtI, dtJ);
    const double pos_i[3] = {FLD(C.pos, 0, i, np), FLD(C.pos, 1, i, np), FLD(C.pos, 2, i, np)};
    FaceGeo g;
    face_geo(M, pos_i, f, np, g);
    for (int d = 0; d < 3; d++) {
      const double an = g.normal[d] * g.area;
      at[d] = side == 0 ? at[d] + an : at[d] - an;
    }
    if (C.bnd[i] != NO_BOUNDARY || C.bnd[j] != NO_BOUNDARY) {
      for (int k = 0; k < NF; k++) lim[k] = (S)0.0;
      continue;
    }
    if (P.limiter == LIM_SHIPPED) continue;  // Phi() == 1: min(lim, 1) leaves lim unchanged
    const double *dr = side == 0 ? g.dri : g.drj;
    const double dtt = side == 0 ? dtI - dth : dtJ - dth;
    for (int k = 0; k < NF; k++) {
      const double dWt = (double)FLD(C.t, k, c, np) * dtt;
      const double dWx = (double)FLD(C.x, k, c, np) * dr[0] + (double)FLD(C.y, k, c, np) * dr[1] +
                         (double)FLD(C.z, k, c, np) * dr[2];
      const double dW = dWx + dWt;
      lim[k] = (S)smin((double)lim[k],
                       limiter_sl(P.limiter, dW, smin(mlo[k] - wc[k], 0.0), smax(mhi[k] - wc[k], 0.0)));
    }
  }
  if (c < C.nactive && C.bnd[c] == NO_BOUNDARY) {
    const double aabs = dsqrt(at[0] * at[0] + at[1] * at[1] + at[2] * at[2]);
    if (!(aabs < 1.0e-10 * pow(M.cvol[c], 2.0 / 3.0))) atomicAdd(&diag[FVM_DIAG_CLOSURE], 1ULL);
  }
  for (int k = 0; k < NF; k++) FLD(C.lim, k, c, np) = lim[k];
  if (c < C.nactive || C.is_ngb[c]) {
    for (int k = 0; k < NF; k++) {
      const double f = 1.0 * (double)lim[k];
      FLD(C.x, k, c, np) = (S)((double)FLD(C.x, k, c, np) * f);
      FLD(C.y, k, c, np) = (S)((double)FLD(C.y, k, c, np) * f);
      FLD(C.z, k, c, np) = (S)((double)FLD(C.z, k, c, np) * f);
      FLD(C.t, k, c, np) = (S)((double)FLD(C.t, k, c, np) * f);
    }
  }
}

// one lane per (cell, variable): every lane repeats the face-end geometry and
// applies the limiter to its own variable
template <typename S>
__global__ void __launch_bounds__(256) k_slope_limiter_v(Cells<S> C, MeshView M, Phys P, unsigned long long *diag) {
  int c, k;
  if (!cell_var(c, k) || c >= C.np) return;
  const int np = C.np;
  S lim = (S)1.0;
  double at[3] = {0.0, 0.0, 0.0};
  const double wc = FLD(C.w, k, c, np);
  const double mlo = (double)FLD(C.mm_lo, k, c, np);
  const double mhi = (double)FLD(C.mm_hi, k, c, np);
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const int f = M.fact[a];
    const int i = M.s1[f], j = M.s2[f];
    const double dtI = C.tend[i] - P.t_global;
    const double dtJ = C.tend[j] - P.t_global;
    const double dth = 0.5 * smin(dtI, dtJ);
    const double pos_i[3] = {FLD(C.pos, 0, i, np), FLD(C.pos, 1, i, np), FLD(C.pos, 2, i, np)};
    FaceGeo g;
    face_geo(M, pos_i, f, np, g);
    for (int d = 0; d < 3; d++) {
      const double an = g.normal[d] * g.area;
      at[d] = side == 0 ? at[d] + an : at[d] - an;
    }
    if (C.bnd[i] != NO_BOUNDARY || C.bnd[j] != NO_BOUNDARY) {
      lim = (S)0.0;
      continue;
    }
    if (P.limiter == LIM_SHIPPED) continue;  // Phi() == 1: min(lim, 1) leaves lim unchanged
    const double *dr = side == 0 ? g.dri : g.drj;
    const double dtt = side == 0 ? dtI - dth : dtJ - dth;
    const double dWt = (double)FLD(C.t, k, c, np) * dtt;
    const double dWx = (double)FLD(C.x, k, c, np) * dr[0] + (double)FLD(C.y, k, c, np) * dr[1] +
                       (double)FLD(C.z, k, c, np) * dr[2];
    const double dW = dWx + dWt;
    lim = (S)smin((double)lim, limiter_sl(P.limiter, dW, smin(mlo - wc, 0.0), smax(mhi - wc, 0.0)));
  }
  if (k == 0 && c < C.nactive && C.bnd[c] == NO_BOUNDARY) {
    const double aabs = dsqrt(at[0] * at[0] + at[1] * at[1] + at[2] * at[2]);
    if (!(aabs < 1.0e-10 * pow(M.cvol[c], 2.0 / 3.0))) atomicAdd(&diag[FVM_DIAG_CLOSURE], 1ULL);
  }
  FLD(C.lim, k, c, np) = lim;
  if (c < C.nactive || C.is_ngb[c]) {
    const double f = 1.0 * (double)lim;
    FLD(C.x, k, c, np) = (S)((double)FLD(C.x, k, c, np) * f);
    FLD(C.y, k, c, np) = (S)((double)FLD(C.y, k, c, np) * f);
    FLD(C.z, k, c, np) = (S)((double)FLD(C.z, k, c, np) * f);
    FLD(C.t, k, c, np) = (S)((double)FLD(C.t, k, c, np) * f);
  }
}

// ---------------------------------------------------------------------------
// System::computeFluidUpdate face part (computeFluidUpdate.cpp:99-223): one
// thread per active face writes the face record
//   rec[0..9] = flux*dt*A, rec[10..12] = n*(psi*dt*A), rec[13] = Bn*dt*A
// and swp[a] = 1 when i and j were swapped (s1 on a boundary).
// ---------------------------------------------------------------------------
// 14 entries per face record (NREC_P = 15 adds the predictor's dV)
constexpr int NREC_F = NF + 4;

// computeFluidUpdate.cpp:99-209: the inputs of computeFlux for active face a
// (boundary swap, space-time extrapolation of both sides, face velocity)
// F: the face arithmetic (double: the reference's; float: the FP32 face
// states, formed from FP64 geometry, dt and dti as the fused FP32 path forms
// them)
template <typename F> struct FaceIn {
  int i, j, jbnd;
  bool sw;
  double dt, area, normal[3], wij[3];
  F Wi[NF], Wj[NF];
};

template <typename S, typename F>
__device__ __forceinline__ void face_inputs(const Cells<S> &C, const MeshView &M, const Phys &P, int a,
                                            unsigned long long *diag, FaceIn<F> &q) {
  const int np = C.np;
  const int f = M.fact[a];
  int i = M.s1[f], j = M.s2[f];
  const double fn[3] = {GEO_N(M, 0, f), GEO_N(M, 1, f), GEO_N(M, 2, f)};
  const double area = dsqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const double ia = 1.0 / area;
  double normal[3] = {fn[0] * ia, fn[1] * ia, fn[2] * ia};
  double dri[3];
  for (int d = 0; d < 3; d++) dri[d] = GEO_C(M, d, f) - FLD(C.pos, d, i, np);
  const bool sw = C.bnd[i] != NO_BOUNDARY;
  if (sw) {
    const int tmp = i;
    i = j;
    j = tmp;
    for (int d = 0; d < 3; d++) {
      normal[d] = normal[d] * -1.0;
      dri[d] = dri[d] * -1.0;
    }
  }
  const double dtI = C.tend[i] - P.t_global;
  const double dtJ = C.tend[j] - P.t_global;
  const double dt = smin(dtI, dtJ);
  const double dti = dtI - 0.5 * dt;
  const double dtj = dtJ - 0.5 * dt;
  const double dsh = dri[0] * normal[0] + dri[1] * normal[1] + dri[2] * normal[2];
  const double dsl = 2.0 * dsh;
  if (diag && !(dsl > 0.0)) atomicAdd(&diag[FVM_DIAG_DSL], 1ULL);
  const double drj[3] = {dri[0] - normal[0] * dsl, dri[1] - normal[1] * dsl, dri[2] - normal[2] * dsl};

  F *Wi = q.Wi, *Wj = q.Wj;
  for (int k = 0; k < NF; k++) {
    Wi[k] = (F)FLD(C.w, k, i, np);
    Wj[k] = (F)FLD(C.w, k, j, np);
  }
  const int jbnd = C.bnd[j];
  if (jbnd == NO_BOUNDARY) {
    if constexpr (std::is_same<F, double>::value) {
      for (int k = 0; k < NF; k++) {
        Wi[k] += (double)FLD(C.x, k, i, np) * dri[0] + (double)FLD(C.y, k, i, np) * dri[1] +
                 (double)FLD(C.z, k, i, np) * dri[2] + (double)FLD(C.t, k, i, np) * dti;
        Wj[k] += (double)FLD(C.x, k, j, np) * drj[0] + (double)FLD(C.y, k, j, np) * drj[1] +
                 (double)FLD(C.z, k, j, np) * drj[2] + (double)FLD(C.t, k, j, np) * dtj;
      }
    } else {
      // FP32: the half-step state W + Wt dti, then the gradient terms, in FMAs
      const float ri[3] = {(float)dri[0], (float)dri[1], (float)dri[2]};
      const float rj[3] = {(float)drj[0], (float)drj[1], (float)drj[2]};
      const float ti = (float)dti, tj = (float)dtj;
      for (int k = 0; k < NF; k++) {
        Wi[k] = __builtin_fmaf((float)FLD(C.z, k, i, np), ri[2],
                               __builtin_fmaf((float)FLD(C.y, k, i, np), ri[1],
                                              __builtin_fmaf((float)FLD(C.x, k, i, np), ri[0],
                                                             __builtin_fmaf((float)FLD(C.t, k, i, np), ti, Wi[k]))));
        Wj[k] = __builtin_fmaf((float)FLD(C.z, k, j, np), rj[2],
                               __builtin_fmaf((float)FLD(C.y, k, j, np), rj[1],
                                              __builtin_fmaf((float)FLD(C.x, k, j, np), rj[0],
                                                             __builtin_fmaf((float)FLD(C.t, k, j, np), tj, Wj[k]))));
      }
    }
  }
  if (diag && (!(Wi[DENS] > F(0.0)) || !(Wi[ETHM] > F(0.0)) || !(Wj[DENS] > F(0.0)) || !(Wj[ETHM] > F(0.0))))
    atomicAdd(&diag[FVM_DIAG_POSITIVITY], 1ULL);

  const double vi[3] = {(double)FLD(C.vel, 0, i, np), (double)FLD(C.vel, 1, i, np), (double)FLD(C.vel, 2, i, np)};
  const double vj[3] = {(double)FLD(C.vel, 0, j, np), (double)FLD(C.vel, 1, j, np), (double)FLD(C.vel, 2, j, np)};
  const double vij[3] = {vj[0] + vi[0], vj[1] + vi[1], vj[2] + vi[2]};
  const double dvij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
  const double fij[3] = {dri[0] - normal[0] * dsh, dri[1] - normal[1] * dsh, dri[2] - normal[2] * dsh};
  const double fdv = dvij[0] * fij[0] + dvij[1] * fij[1] + dvij[2] * fij[2];
  const double rji[3] = {normal[0] * dsl, normal[1] * dsl, normal[2] * dsl};
  const double r = 1.0 / sqr(dsl);
  q.wij[0] = vij[0] * 0.5 - rji[0] * fdv * r;
  q.wij[1] = vij[1] * 0.5 - rji[1] * fdv * r;
  q.wij[2] = vij[2] * 0.5 - rji[2] * fdv * r;
  for (int d = 0; d < 3; d++) q.normal[d] = normal[d];
  q.i = i;
  q.j = j;
  q.jbnd = jbnd;
  q.sw = sw;
  q.dt = dt;
  q.area = area;
}

// F = float: the FP32 face arithmetic (face states and HLLD); the records
// times dt and area, and every sum over them, stay FP64
template <typename S, typename F>
__global__ void __launch_bounds__(256) k_face_flux(Cells<S> C, MeshView M, Phys P, double *rec, int *swp,
                                   unsigned long long *diag) {
  const int a = xgid();
  if (a >= M.nfa) return;
  FaceIn<F> q;
  face_inputs<S, F>(C, M, P, a, diag, q);
  F psi, Bn, flux[NF];
  F wij[3], nrm[3];
  for (int d = 0; d < 3; d++) {
    wij[d] = (F)q.wij[d];
    nrm[d] = (F)q.normal[d];
  }
  compute_flux<F>((F)P.gamma, wij, q.Wi, q.Wj, q.jbnd, nrm, psi, Bn, flux);

  const double tarea = q.dt * q.area;
  for (int k = 0; k < NF; k++) REC(rec, k, a, NREC_F) = (double)flux[k] * tarea;
  for (int d = 0; d < 3; d++) REC(rec, NF + d, a, NREC_F) = q.normal[d] * ((double)psi * tarea);
  REC(rec, NF + 3, a, NREC_F) = (double)Bn * tarea;
  swp[a] = q.sw ? 1 : 0;
}

// the computeFlux inputs of active faces [first, first + count), in the
// batch layout of fvm_compute_flux
template <typename S>
__global__ void __launch_bounds__(256) k_face_states(Cells<S> C, MeshView M, Phys P, int first, int count, double *wij,
                                                     double *Wi, double *Wj, int *jbnd, double *normal) {
  const int q0 = gid();
  if (q0 >= count) return;
  FaceIn<double> q;
  face_inputs<S, double>(C, M, P, first + q0, nullptr, q);
  for (int d = 0; d < 3; d++) {
    wij[3 * (size_t)q0 + d] = q.wij[d];
    normal[3 * (size_t)q0 + d] = q.normal[d];
  }
  for (int k = 0; k < NF; k++) {
    Wi[(size_t)NF * q0 + k] = q.Wi[k];
    Wj[(size_t)NF * q0 + k] = q.Wj[k];
  }
  jbnd[q0] = q.jbnd;
}

// computeFluidUpdate.cpp:225-245: per cell, in face_active_list order
template <typename S>
__global__ void __launch_bounds__(256) k_scatter(Cells<S> C, MeshView M, const double *rec, const int *swp) {
  const int c = xgid();
  if (c >= C.np) return;
  const int np = C.np;
  S U[NF], gp[3], dv;
  for (int k = 0; k < NF; k++) U[k] = (S)0.0;
  gp[0] = gp[1] = gp[2] = (S)0.0;
  dv = (S)0.0;
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const bool is_i = (side == 0) != (swp[a] != 0);
    if (is_i) {
      for (int k = 0; k < NF; k++) U[k] = (S)((double)U[k] - REC(rec, k, a, NREC_F));
      for (int d = 0; d < 3; d++) gp[d] = gp[d] + (S)REC(rec, NF + d, a, NREC_F);
      dv = (S)((double)dv + REC(rec, NF + 3, a, NREC_F));
    } else {
      for (int k = 0; k < NF; k++) U[k] = (S)((double)U[k] + REC(rec, k, a, NREC_F));
      for (int d = 0; d < 3; d++) gp[d] = gp[d] - (S)REC(rec, NF + d, a, NREC_F);
      dv = (S)((double)dv - REC(rec, NF + 3, a, NREC_F));
    }
  }
  FLD(C.ddivB, 0, c, np) = FLD(C.ddivB, 0, c, np) + dv;
  for (int d = 0; d < 3; d++) FLD(C.dgp, d, c, np) = FLD(C.dgp, d, c, np) + gp[d];
  for (int k = 0; k < NF; k++) FLD(C.dU, k, c, np) = FLD(C.dU, k, c, np) + U[k];
}

// one lane per (cell, variable); lanes 0-2 also sum the gradPsi terms and
// lane 3 the div B term
template <typename S>
__global__ void __launch_bounds__(256) k_scatter_v(Cells<S> C, MeshView M, const double *rec, const int *swp) {
  int c, k;
  if (!cell_var(c, k) || c >= C.np) return;
  const int np = C.np;
  S U = (S)0.0, gx = (S)0.0;  // gx: gradPsi[k] (k < 3) or div B (k == 3)
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const bool is_i = (side == 0) != (swp[a] != 0);
    const double r = REC(rec, k, a, NREC_F);
    const double x = k < 4 ? REC(rec, NF + k, a, NREC_F) : 0.0;
    if (is_i) {
      U = (S)((double)U - r);
      if (k < 3) gx = gx + (S)x;
      else if (k == 3) gx = (S)((double)gx + x);
    } else {
      U = (S)((double)U + r);
      if (k < 3) gx = gx - (S)x;
      else if (k == 3) gx = (S)((double)gx - x);
    }
  }
  if (k == 3) FLD(C.ddivB, 0, c, np) = FLD(C.ddivB, 0, c, np) + gx;
  if (k < 3) FLD(C.dgp, k, c, np) = FLD(C.dgp, k, c, np) + gx;
  FLD(C.dU, k, c, np) = FLD(C.dU, k, c, np) + U;
}

// ---------------------------------------------------------------------------
// System::computePrimitives (computeFluidUpdate.cpp:347-542), one thread per
// active cell; the failed-mass fallback (:508-541) touches only its own cell
// and runs in the same thread.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ethm_from_entropy(double g, const double *W) {
  return W[ENTR] * pow(W[DENS], g) / (g - 1.0);
}
__device__ __forceinline__ double entropy_from_ethm(double g, const double *W) {
  return (g - 1.0) * W[ETHM] / pow(W[DENS], g);
}

template <typename S>
__global__ void __launch_bounds__(256) k_primitives(Cells<S> C, MeshView M, Phys P, int predictor_flag,
                                    unsigned long long *diag) {
  const int i = xgid();
  if (i >= C.nactive) return;
  const int np = C.np;
  const double g = P.gamma;
  double Ui[NF], wi[NF];
  for (int k = 0; k < NF; k++) {
    Ui[k] = FLD(C.U, k, i, np);
    wi[k] = FLD(C.w, k, i, np);
  }
  if (C.bnd[i] != NO_BOUNDARY) {
    to_primitive(Ui, C.mvol[i], wi);
    for (int k = 0; k < NF; k++) {
      FLD(C.x, k, i, np) = (S)0.0;
      FLD(C.y, k, i, np) = (S)0.0;
      FLD(C.z, k, i, np) = (S)0.0;
      FLD(C.t, k, i, np) = (S)0.0;
    }
    C.mvol[i] = M.cvol[i];
    to_conservative(wi, C.mvol[i], Ui);
    for (int k = 0; k < NF; k++) {
      FLD(C.w, k, i, np) = wi[k];
      FLD(C.U, k, i, np) = Ui[k];
    }
    return;
  }
  double Uc[NF];
  for (int k = 0; k < NF; k++) Uc[k] = Ui[k] + (double)FLD(C.dU, k, i, np);
  const double divB = (double)FLD(C.ddivB, 0, i, np);
  const double dPsi[3] = {(double)FLD(C.dgp, 0, i, np), (double)FLD(C.dgp, 1, i, np),
                          (double)FLD(C.dgp, 2, i, np)};
  const double dt = P.t_global - C.tbeg[i];
  C.divB[i] = dt > 0.0 ? divB / dt / M.cvol[i] : divB;

  double W1[NF];
  const bool failed = Uc[MASS] <= 0.0;
  if (!failed) {
    const double dth = 0.5 * dt;
    double Wh[NF];
    for (int k = 0; k < NF; k++) Wh[k] = wi[k] + (double)FLD(C.t, k, i, np) * dth;
    Uc[WBX] -= Wh[VELX] * divB;
    Uc[WBY] -= Wh[VELY] * divB;
    Uc[WBZ] -= Wh[VELZ] * divB;
    const double pres = pressure(g, Wh[ETHM]);
    const double B2 = Wh[BX] * Wh[BX] + Wh[BY] * Wh[BY] + Wh[BZ] * Wh[BZ];
    const double dcs2 = g * pres + B2;
    const double ch_sig = dsqrt(dcs2 / Wh[DENS]);
    Uc[WBX] -= dPsi[0];
    Uc[WBY] -= dPsi[1];
    Uc[WBZ] -= dPsi[2];
    Uc[MPSI] -= dcs2 * divB;
    if (!predictor_flag) {
      double a0[3], a1[3];
      for (int d = 0; d < 3; d++) {
        a0[d] = FLD(C.acc0, d, i, np);
        a1[d] = FLD(C.acc1, d, i, np);
      }
      const double M0 = Ui[MASS], M1 = Uc[MASS];
      double dmom[3];
      for (int d = 0; d < 3; d++) dmom[d] = ((a0[d] * M0 + a1[d] * M1) * 0.5) * dt;
      Uc[MOMX] += dmom[0];
      Uc[MOMY] += dmom[1];
      Uc[MOMZ] += dmom[2];
      const double mom0[3] = {Ui[MOMX], Ui[MOMY], Ui[MOMZ]};
      const double mom1[3] = {Uc[MOMX], Uc[MOMY], Uc[MOMZ]};
      Uc[ENER] += ((mom0[0] * a0[0] + mom0[1] * a0[1] + mom0[2] * a0[2]) +
                   (mom1[0] * a1[0] + mom1[1] * a1[1] + mom1[2] * a1[2])) *
                  dt * 0.5;
    }
    const double cr = 0.5;
    const double l_min = pow(M.cvol[i] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
    Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
    to_primitive(Uc, M.cvol[i], W1);
    if (W1[ETHM] <= 0.0) {
      W1[ETHM] = ethm_from_entropy(g, W1);
      atomicAdd(&diag[FVM_DIAG_ETHM_FAILED], 1ULL);
    } else {
      W1[ENTR] = entropy_from_ethm(g, W1);
    }
    if (predictor_flag) {
      if (dt > 0.0)
        for (int k = 0; k < NF; k++) FLD(C.t, k, i, np) = (S)((W1[k] - wi[k]) / dt);
      else
        for (int k = 0; k < NF; k++) FLD(C.t, k, i, np) = (S)0.0;
      return;
    }
  } else {
    atomicAdd(&diag[FVM_DIAG_MASS_FAILED], 1ULL);
    double Ut[NF];
    for (int k = 0; k < NF; k++) W1[k] = wi[k] + (double)FLD(C.t, k, i, np) * dt;
    to_conservative(W1, C.mvol[i], Ut);
    to_primitive(Ut, M.cvol[i], W1);
    if (W1[ETHM] <= 0.0) W1[ETHM] = ethm_from_entropy(g, W1);
    else W1[ENTR] = entropy_from_ethm(g, W1);
  }
  C.mvol[i] = M.cvol[i];
  to_conservative(W1, M.cvol[i], Ui);
  for (int k = 0; k < NF; k++) {
    FLD(C.w, k, i, np) = W1[k];
    FLD(C.U, k, i, np) = Ui[k];
    FLD(C.dU, k, i, np) = (S)0.0;
  }
  for (int d = 0; d < 3; d++) FLD(C.dgp, d, i, np) = (S)0.0;
  FLD(C.ddivB, 0, i, np) = (S)0.0;
}

// ---------------------------------------------------------------------------
// System::computeTimestep (computeTimestep.cpp:6-62)
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_timestep(Cells<S> C, MeshView M, Phys P, unsigned long long *dtmin_bits) {
  const int i = xgid();
  double dti_out = REF_HUGE;
  if (i < C.nactive) {
    const int np = C.np;
    const double g = P.gamma;
    const double Volume = C.mvol[i];
    double Ui[NF], Wi[NF];
    for (int k = 0; k < NF; k++) Ui[k] = FLD(C.U, k, i, np);
    to_primitive(Ui, Volume, Wi);
    const double h = pow(Volume / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
    const double pv[3] = {FLD(C.mvel, 0, i, np), FLD(C.mvel, 1, i, np), FLD(C.mvel, 2, i, np)};
    double csi = 0.0, vrel = 0.0;
    const bool isb = C.bnd[i] != NO_BOUNDARY;
    if (!isb) {
      const double Pi = pressure(g, Wi[ETHM]) * g;
      const double B2i = Wi[BX] * Wi[BX] + Wi[BY] * Wi[BY] + Wi[BZ] * Wi[BZ];
      csi = dsqrt((Pi + B2i) / Wi[DENS]);
      const double dv[3] = {Wi[VELX] - pv[0], Wi[VELY] - pv[1], Wi[VELZ] - pv[2]};
      vrel = dsqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    }
    double dti = h / (csi + vrel + 1.0 / REF_HUGE);
    const double fac = isb ? 0.0 : 1.0;
    for (int e = M.coff[i]; e < M.coff[i + 1]; e++) {
      const int f = M.cface[e];
      const int s1 = M.s1[f];
      const int j = (i == s1) ? M.s2[f] : s1;
      const double dvi[3] = {FLD(C.w, VELX, j, np) - Wi[VELX], FLD(C.w, VELY, j, np) - Wi[VELY],
                             FLD(C.w, VELZ, j, np) - Wi[VELZ]};
      const double dwi[3] = {(double)FLD(C.vel, 0, j, np) - pv[0], (double)FLD(C.vel, 1, j, np) - pv[1],
                             (double)FLD(C.vel, 2, j, np) - pv[2]};
      const double dvmax = dsqrt((dwi[0] * dwi[0] + dwi[1] * dwi[1] + dwi[2] * dwi[2]) +
                                 fac * (dvi[0] * dvi[0] + dvi[1] * dvi[1] + dvi[2] * dvi[2]) +
                                 1.0 / REF_HUGE);
      dti = smin(dti, h / (csi + dvmax));
    }
    dti = smin(dti, REF_HUGE);
    C.dt_new[i] = dti * P.courant;
    dti_out = C.dt_new[i];
  }
  // min over active cells; dt > 0 so the IEEE bit patterns order like the values
  double v = dti_out;
  for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0 && i - (int)(threadIdx.x & 63) < C.nactive) {
    if (v >= 0.0) atomicMin(dtmin_bits, (unsigned long long)__double_as_longlong(v));
  }
}

// ---------------------------------------------------------------------------
// System::computePredictor (computePredictor.cpp:11-391)
// face part: record rec[0..9] = flux*A, [10..12] = n*(psi*A), [13] = Bn*A, [14] = dV
// ---------------------------------------------------------------------------
constexpr int NREC_P = 15;

// F = float: the FP32 HLLD of the first-order states; the records stay FP64
template <typename S, typename F>
__global__ void __launch_bounds__(256) k_predictor_face(Cells<S> C, MeshView M, Phys P, double *rec, int *swp,
                                        unsigned long long *diag) {
  const int a = xgid();
  if (a >= M.nfa) return;
  const int np = C.np;
  const int f = M.fact[a];
  int i = M.s1[f], j = M.s2[f];
  const double fn[3] = {GEO_N(M, 0, f), GEO_N(M, 1, f), GEO_N(M, 2, f)};
  const double area = dsqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const double ia = 1.0 / area;
  double normal[3] = {fn[0] * ia, fn[1] * ia, fn[2] * ia};
  double dri[3];
  for (int d = 0; d < 3; d++) dri[d] = GEO_C(M, d, f) - FLD(C.pos, d, i, np);
  const bool sw = C.bnd[i] != NO_BOUNDARY;
  if (sw) {
    const int tmp = i;
    i = j;
    j = tmp;
    for (int d = 0; d < 3; d++) {
      normal[d] = normal[d] * -1.0;
      dri[d] = dri[d] * -1.0;
    }
  }
  const double dsh = dri[0] * normal[0] + dri[1] * normal[1] + dri[2] * normal[2];
  const double dsl = 2.0 * dsh;
  if (!(dsl > 0.0)) atomicAdd(&diag[FVM_DIAG_DSL], 1ULL);
  const int jbnd = C.bnd[j];
  const int jw = (jbnd == NO_BOUNDARY || jbnd == INFLOW) ? j : i;
  double Wi[NF], Wj[NF];
  for (int k = 0; k < NF; k++) {
    Wi[k] = FLD(C.w, k, i, np);
    Wj[k] = FLD(C.w, k, jw, np);
  }
  const double vi[3] = {(double)FLD(C.vel, 0, i, np), (double)FLD(C.vel, 1, i, np), (double)FLD(C.vel, 2, i, np)};
  const double vj[3] = {(double)FLD(C.vel, 0, j, np), (double)FLD(C.vel, 1, j, np), (double)FLD(C.vel, 2, j, np)};
  const double vij[3] = {vj[0] + vi[0], vj[1] + vi[1], vj[2] + vi[2]};
  const double dvij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
  const double fij[3] = {dri[0] - normal[0] * dsh, dri[1] - normal[1] * dsh, dri[2] - normal[2] * dsh};
  const double fdv = dvij[0] * fij[0] + dvij[1] * fij[1] + dvij[2] * fij[2];
  const double rij[3] = {normal[0] * dsl, normal[1] * dsl, normal[2] * dsl};
  const double r = 1.0 / sqr(dsl);
  const double wij[3] = {vij[0] * 0.5 - rij[0] * fdv * r, vij[1] * 0.5 - rij[1] * fdv * r,
                         vij[2] * 0.5 - rij[2] * fdv * r};
  F psi, Bn, flux[NF];
  {
    F wf[3], nf[3], Wif[NF], Wjf[NF];
    for (int d = 0; d < 3; d++) {
      wf[d] = (F)wij[d];
      nf[d] = (F)normal[d];
    }
    for (int k = 0; k < NF; k++) {
      Wif[k] = (F)Wi[k];
      Wjf[k] = (F)Wj[k];
    }
    compute_flux<F>((F)P.gamma, wf, Wif, Wjf, jbnd, nf, psi, Bn, flux);
  }
  for (int k = 0; k < NF; k++) REC(rec, k, a, NREC_P) = (double)flux[k] * area;
  for (int d = 0; d < 3; d++) REC(rec, NF + d, a, NREC_P) = normal[d] * ((double)psi * area);
  REC(rec, NF + 3, a, NREC_P) = (double)Bn * area;
  REC(rec, NF + 4, a, NREC_P) = area * (rij[0] * wij[0] + rij[1] * wij[1] + rij[2] * wij[2]) / dsl;
  swp[a] = sw ? 1 : 0;
}

__device__ __forceinline__ int exp_of(double dt) {
  return (int)(((unsigned long long)__double_as_longlong(dt) >> 52) & 0xfff);
}
// Scheduler::get_dt(get_rung(dt)) (Scheduler.h:94-111)
__device__ __forceinline__ double rung_dt(double dtmax, int min_rung, double dt) {
  int rung = exp_of(dtmax) - exp_of(dt);
  if (rung < min_rung) rung = min_rung;
  const double dt_tick = dtmax / (double)(1ULL << 31);
  return dt_tick * (double)(1ULL << (31 - rung));
}

template <typename S>
__global__ void __launch_bounds__(256) k_predictor_cell(Cells<S> C, MeshView M, Phys P, const double *rec,
                                        const int *swp, double dtmax, int min_rung,
                                        unsigned long long *diag) {
  const int c = xgid();
  if (c >= C.nactive) return;
  const int np = C.np, nactive = C.nactive;
  S dU[NF], gp[3], dv;
  double dV = 0.0;
  for (int k = 0; k < NF; k++) dU[k] = (S)0.0;
  gp[0] = gp[1] = gp[2] = (S)0.0;
  dv = (S)0.0;
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const int f = M.fact[a];
    const bool sw = swp[a] != 0;
    const bool is_i = (side == 0) != sw;
    if (is_i) {
      for (int k = 0; k < NF; k++) dU[k] = (S)((double)dU[k] - REC(rec, k, a, NREC_P));
      for (int d = 0; d < 3; d++) gp[d] = gp[d] + (S)REC(rec, NF + d, a, NREC_P);
      dv = (S)((double)dv + REC(rec, NF + 3, a, NREC_P));
      dV += REC(rec, NF + 4, a, NREC_P);
    } else {
      // j receives only when active and not a boundary (computePredictor.cpp:127-128)
      const int jeff = sw ? M.s1[f] : M.s2[f];
      (void)jeff;
      if (c >= nactive || C.bnd[c] != NO_BOUNDARY) continue;
      dV -= REC(rec, NF + 4, a, NREC_P);
      for (int k = 0; k < NF; k++) dU[k] = (S)((double)dU[k] + REC(rec, k, a, NREC_P));
      for (int d = 0; d < 3; d++) gp[d] = gp[d] - (S)REC(rec, NF + d, a, NREC_P);
      dv = (S)((double)dv - REC(rec, NF + 3, a, NREC_P));
    }
  }
  if (C.bnd[c] != NO_BOUNDARY) return;
  const double g = P.gamma;
  const double divB = (double)dv;
  const double dPsi[3] = {(double)gp[0], (double)gp[1], (double)gp[2]};
  double U0[NF], W0[NF];
  for (int k = 0; k < NF; k++) U0[k] = FLD(C.U, k, c, np);
  to_primitive(U0, M.cvol[c], W0);
  double dtn = C.dt_new[c];
  while (U0[MASS] + (double)dU[MASS] * dtn <= 0.0) dtn *= 0.5;
  C.dt_new[c] = dtn;
  const double dt = rung_dt(dtmax, min_rung, dtn) / 128;
  double Uc[NF];
  for (int k = 0; k < NF; k++) Uc[k] = U0[k] + (double)dU[k] * dt;
  Uc[WBX] -= W0[VELX] * divB * dt;
  Uc[WBY] -= W0[VELY] * divB * dt;
  Uc[WBZ] -= W0[VELZ] * divB * dt;
  const double pres = pressure(g, W0[ETHM]);
  const double B2 = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
  const double dcs2 = g * pres + B2;
  const double ch_sig = dsqrt(dcs2 / W0[DENS]);
  Uc[WBX] -= dPsi[0] * dt;
  Uc[WBY] -= dPsi[1] * dt;
  Uc[WBZ] -= dPsi[2] * dt;
  Uc[MPSI] -= dcs2 * divB * dt;
  const double cr = 0.5;
  const double l_min = pow(M.cvol[c] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
  Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
  double W1[NF];
  to_primitive(Uc, M.cvol[c] * exp(dV * dt / M.cvol[c]), W1);
  if (W1[ETHM] <= 0.0) {
    W1[ETHM] = ethm_from_entropy(g, W1);
    atomicAdd(&diag[FVM_DIAG_ETHM_FAILED], 1ULL);
  } else {
    W1[ENTR] = entropy_from_ethm(g, W1);
  }
  const double idt = 1.0 / dt;
  for (int k = 0; k < NF; k++) FLD(C.t, k, c, np) = (S)((W1[k] - W0[k]) * idt);
}

// The face sums take one lane per (cell, variable) (lanes 0-2 also gradPsi,
// lane 3 div B, lane 4 the volume change); the cell update then runs in every
// lane of the cell on the sums gathered from its lanes, and lane k stores Wt[k].
template <typename S>
__global__ void __launch_bounds__(256) k_predictor_cell_v(Cells<S> C, MeshView M, Phys P, const double *rec,
                                        const int *swp, double dtmax, int min_rung,
                                        unsigned long long *diag) {
  int c, k;
  if (!cell_var(c, k) || c >= C.nactive) return;
  const int np = C.np;
  const int base = (int)__lane_id() - k;
  S dUk = (S)0.0, gx = (S)0.0;  // gx: gradPsi[k] (k < 3) or div B (k == 3)
  double dVk = 0.0;             // lane 4: the volume change
  for (int e = M.soff[c]; e < M.soff[c + 1]; e++) {
    const int ent = M.sent[e];
    const int a = ent >> 1, side = ent & 1;
    const bool sw = swp[a] != 0;
    const bool is_i = (side == 0) != sw;
    const double r = REC(rec, k, a, NREC_P);
    const double x = k < 5 ? REC(rec, NF + k, a, NREC_P) : 0.0;
    if (is_i) {
      dUk = (S)((double)dUk - r);
      if (k < 3) gx = gx + (S)x;
      else if (k == 3) gx = (S)((double)gx + x);
      else if (k == 4) dVk += x;
    } else {
      // j receives only when not a boundary (computePredictor.cpp:127-128);
      // every cell of this kernel is active
      if (C.bnd[c] != NO_BOUNDARY) continue;
      if (k == 4) dVk -= x;
      dUk = (S)((double)dUk + r);
      if (k < 3) gx = gx - (S)x;
      else if (k == 3) gx = (S)((double)gx - x);
    }
  }
  if (C.bnd[c] != NO_BOUNDARY) return;
  // every lane of the cell gathers the sums
  double dU[NF];
  for (int m = 0; m < NF; m++) dU[m] = (double)__shfl((double)dUk, base + m);
  const double dPsi[3] = {(double)__shfl((double)gx, base), (double)__shfl((double)gx, base + 1),
                          (double)__shfl((double)gx, base + 2)};
  const double divB = __shfl((double)gx, base + 3);
  const double dV = __shfl(dVk, base + 4);
  const double g = P.gamma;
  double U0[NF], W0[NF];
  for (int m = 0; m < NF; m++) U0[m] = FLD(C.U, m, c, np);
  to_primitive(U0, M.cvol[c], W0);
  double dtn = C.dt_new[c];
  while (U0[MASS] + dU[MASS] * dtn <= 0.0) dtn *= 0.5;
  if (k == 0) C.dt_new[c] = dtn;
  const double dt = rung_dt(dtmax, min_rung, dtn) / 128;
  double Uc[NF];
  for (int m = 0; m < NF; m++) Uc[m] = U0[m] + dU[m] * dt;
  Uc[WBX] -= W0[VELX] * divB * dt;
  Uc[WBY] -= W0[VELY] * divB * dt;
  Uc[WBZ] -= W0[VELZ] * divB * dt;
  const double pres = pressure(g, W0[ETHM]);
  const double B2 = W0[BX] * W0[BX] + W0[BY] * W0[BY] + W0[BZ] * W0[BZ];
  const double dcs2 = g * pres + B2;
  const double ch_sig = dsqrt(dcs2 / W0[DENS]);
  Uc[WBX] -= dPsi[0] * dt;
  Uc[WBY] -= dPsi[1] * dt;
  Uc[WBZ] -= dPsi[2] * dt;
  Uc[MPSI] -= dcs2 * divB * dt;
  const double cr = 0.5;
  const double l_min = pow(M.cvol[c] / (4.0 * REF_PI / 3.0), 1.0 / 3.0);
  Uc[MPSI] *= exp(-cr * ch_sig / l_min * dt);
  double W1[NF];
  to_primitive(Uc, M.cvol[c] * exp(dV * dt / M.cvol[c]), W1);
  if (W1[ETHM] <= 0.0) {
    W1[ETHM] = ethm_from_entropy(g, W1);
    if (k == 0) atomicAdd(&diag[FVM_DIAG_ETHM_FAILED], 1ULL);
  } else {
    W1[ENTR] = entropy_from_ethm(g, W1);
  }
  const double idt = 1.0 / dt;
  double out = 0.0;
  for (int m = 0; m < NF; m++)
    if (m == k) out = (W1[m] - W0[m]) * idt;
  FLD(C.t, k, c, np) = (S)out;
}

// ---------------------------------------------------------------------------
// batched System::computeFlux (FP64)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_flux_batch(int n, double gamma, const double *wij, const double *Wi, const double *Wj,
                             const int *jbnd, const double *normal, double *psi, double *Bn, double *flux) {
  const int f = gid();
  if (f >= n) return;
  double w3[3], n3[3], L[NF], R[NF], F[NF];
  for (int d = 0; d < 3; d++) {
    w3[d] = wij[3 * (size_t)f + d];
    n3[d] = normal[3 * (size_t)f + d];
  }
  for (int k = 0; k < NF; k++) {
    L[k] = Wi[(size_t)NF * f + k];
    R[k] = Wj[(size_t)NF * f + k];
  }
  double p, b;
  compute_flux<double>(gamma, w3, L, R, jbnd[f], n3, p, b, F);
  psi[f] = p;
  Bn[f] = b;
  for (int k = 0; k < NF; k++) flux[(size_t)NF * f + k] = F[k];
}

// ---------------------------------------------------------------------------
// get_active_faces (Iterate.cpp:298-324) and the per-cell face-end lists
// ---------------------------------------------------------------------------
__global__ void k_fill_int(int *p, int v, int n) {
  const int i = gid();
  if (i < n) p[i] = v;
}
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) {
  if (gid() == 0) *p = v;
}

// Cell lists from the face list as localMesh.cpp:586-589 builds them: every
// face is listed by its s1 and, when s2 is active, by its s2, in the order the
// faces were created (ascending index).  Count, scan, fill, then sort each
// cell's short list.
__global__ void k_csr_count(const int *__restrict__ s1, const int *__restrict__ s2, int nface, int nactive,
                            int *__restrict__ cnt) {
  const int f = gid();
  if (f >= nface) return;
  atomicAdd(&cnt[s1[f]], 1);
  if (s2[f] < nactive) atomicAdd(&cnt[s2[f]], 1);
}
__global__ void k_csr_fill(const int *__restrict__ s1, const int *__restrict__ s2, int nface, int nactive,
                           int *__restrict__ cursor, int *__restrict__ cface) {
  const int f = gid();
  if (f >= nface) return;
  cface[atomicAdd(&cursor[s1[f]], 1)] = f;
  if (s2[f] < nactive) cface[atomicAdd(&cursor[s2[f]], 1)] = f;
}
__global__ void k_csr_sort(const int *__restrict__ coff, int nactive, int *__restrict__ cface) {
  const int i = gid();
  if (i >= nactive) return;
  const int b = coff[i], e = coff[i + 1];
  for (int k = b + 1; k < e; k++) {
    const int v = cface[k];
    int m = k - 1;
    while (m >= b && cface[m] > v) {
      cface[m + 1] = cface[m];
      m--;
    }
    cface[m + 1] = v;
  }
}
// first position of each face in the concatenated face lists of active,
// non-boundary cells
__global__ void k_face_firstpos(MeshView M, const int *bnd, int nactive, int *firstpos) {
  const int i = gid();
  if (i >= nactive || bnd[i] != NO_BOUNDARY) return;
  for (int e = M.coff[i]; e < M.coff[i + 1]; e++) atomicMin(&firstpos[M.cface[e]], e);
}
__global__ void k_face_firstflag(MeshView M, const int *bnd, int nactive, const int *firstpos, int *flag) {
  const int i = gid();
  if (i >= nactive) return;
  const bool ok = bnd[i] == NO_BOUNDARY;
  for (int e = M.coff[i]; e < M.coff[i + 1]; e++) flag[e] = (ok && firstpos[M.cface[e]] == e) ? 1 : 0;
}
// The same flags when every active cell lists all its faces (each face under
// s1 when s1 is active and under s2 when s2 is active): a face first appears
// in the lower-indexed of its two cells that are active and not on a
// boundary, since the lists are concatenated in cell order.
__global__ void k_face_firstflag_direct(MeshView M, const int *bnd, int nactive, int *flag) {
  const int i = gid();
  if (i >= nactive) return;
  const bool ok = bnd[i] == NO_BOUNDARY;
  for (int e = M.coff[i]; e < M.coff[i + 1]; e++) {
    const int f = M.cface[e];
    const int j = (M.s1[f] == i) ? M.s2[f] : M.s1[f];
    const bool earlier = j < i && j < nactive && bnd[j] == NO_BOUNDARY;
    flag[e] = (ok && !earlier) ? 1 : 0;
  }
}
__global__ void k_face_compact(MeshView M, const int *flag, const int *pos, int nref, int *fact) {
  const int e = gid();
  if (e < nref && flag[e]) fact[pos[e]] = M.cface[e];
}
// Consecutive active faces mostly share the cell whose list introduced them,
// so lanes with the same target add once per wavefront: the lowest such lane
// adds the lane count, and each lane takes base + its rank among them.  Every
// lane of the wavefront runs the loop (inactive ones with target -1).
__device__ __forceinline__ int wave_add(int *ctr, int t, bool want_pos) {
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long left = __ballot(t >= 0);
  int pos = -1;
  while (left) {
    const int lead = __ffsll((long long)left) - 1;
    const int tv = __shfl(t, lead);
    const unsigned long long same = __ballot(t == tv);
    int base = 0;
    if (lane == lead) base = atomicAdd(&ctr[tv], __popcll(same));
    if (want_pos) {
      base = __shfl(base, lead);
      if (t == tv) pos = base + __popcll(same & ((1ull << lane) - 1ull));
    }
    left &= ~same;
  }
  return pos;
}
__global__ void k_end_count(MeshView M, int *cnt) {
  const int a = gid();
  int t1 = -1, t2 = -1;
  if (a < M.nfa) {
    const int f = M.fact[a];
    t1 = M.s1[f];
    t2 = M.s2[f];
  }
  wave_add(cnt, t1, false);
  wave_add(cnt, t2, false);
}
__global__ void k_end_fill(MeshView M, int *cursor, int *sent) {
  const int a = gid();
  int t1 = -1, t2 = -1;
  if (a < M.nfa) {
    const int f = M.fact[a];
    t1 = M.s1[f];
    t2 = M.s2[f];
  }
  const int p1 = wave_add(cursor, t1, true);
  const int p2 = wave_add(cursor, t2, true);
  if (a < M.nfa) {
    sent[p1] = (a << 1) | 0;
    sent[p2] = (a << 1) | 1;
  }
}
// order each cell's face-ends by face_active_list position (insertion sort of
// short segments) so that every sum runs in the reference's face order
// Each cell's face-end entries (a << 1 | side, all distinct) in ascending
// order.  Up to END_SORT_REG entries are sorted in registers by a fully
// unrolled odd-even transposition network (static indices, padded with
// INT_MAX); longer lists are insertion-sorted in place.  Distinct keys have
// one sorted order, so both give the same list.
constexpr int END_SORT_REG = 48;
__global__ void k_end_sort(const int *soff, int *sent, int n) {
  const int c = gid();
  if (c >= n) return;
  const int b = soff[c], e = soff[c + 1];
  const int deg = e - b;
  if (deg <= 1) return;
  if (deg <= END_SORT_REG) {
    int v[END_SORT_REG];
#pragma unroll
    for (int k = 0; k < END_SORT_REG; k++) v[k] = k < deg ? sent[b + k] : INT_MAX;
#pragma unroll
    for (int pass = 0; pass < END_SORT_REG; pass++) {
#pragma unroll
      for (int k = pass & 1; k + 1 < END_SORT_REG; k += 2) {
        const int lo = min(v[k], v[k + 1]), hi = max(v[k], v[k + 1]);
        v[k] = lo;
        v[k + 1] = hi;
      }
    }
#pragma unroll
    for (int k = 0; k < END_SORT_REG; k++)
      if (k < deg) sent[b + k] = v[k];
    return;
  }
  for (int p = b + 1; p < e; p++) {
    const int v = sent[p];
    int q = p - 1;
    while (q >= b && sent[q] > v) {
      sent[q + 1] = sent[q];
      q--;
    }
    sent[q + 1] = v;
  }
}

}  // namespace fvm

// ===========================================================================
// staging memory of fvm_compute_flux_noalloc
// ===========================================================================
// A file-scope device array: every GPU the library runs on holds its own copy
// (at its own address) until the process ends.  In host code the name
// `scratch` is only a host-side stand-in; the device address of each GPU's
// copy comes from hipGetSymbolAddress, called once per device by fvm_create
// with that device current, and only that address is ever passed to kernels
// or copies.
static __device__ __align__(256) unsigned char scratch[FVM_FLUX_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device, or null
}  // namespace

// device address of this device's copy of `scratch` (the caller has made the
// device current); looked up once per device
static int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return FVM_EARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return FVM_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FVM_EHIP;
  g_scratch_ptr[device] = p;
  return FVM_OK;
}

static void *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ===========================================================================
// context
// ===========================================================================
int fvm_fail(fvm_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}
static int fail(fvm_ctx *c, int code, const std::string &msg) { return fvm_fail(c, code, msg); }
#define HIPCHK(c, call) FVM_HIPCHK(c, call)

// the end of an entry point: wait for its work unless a substep defers it
static int finish(fvm_ctx *c) {
  if (!c->defer_sync) HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_ensure(fvm_ctx *c, DBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return FVM_OK;
  if (b.p) HIPCHK(c, hipFree(b.p));
  b.p = nullptr;
  b.bytes = 0;
  if (bytes == 0) bytes = 8;
  HIPCHK(c, hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return FVM_OK;
}
static int ensure(fvm_ctx *c, DBuf &b, size_t bytes) { return fvm_ensure(c, b, bytes); }
void fvm_release(DBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}
static void release(DBuf &b) { fvm_release(b); }

static size_t store_size(const fvm_ctx *c) { return c->prm.storage == FVM_STORAGE_SHIPPED ? sizeof(float) : sizeof(double); }
static size_t field_bytes(const fvm_ctx *c, int f, int np) {
  const auto &fi = fvm::kFields[f];
  const size_t es = fi.kind == fvm::K_INT ? sizeof(int) : (fi.kind == fvm::K_STORE ? store_size(c) : sizeof(double));
  return es * (size_t)fi.width * (size_t)np;
}

template <typename S> static fvm::Cells<S> cells(fvm_ctx *c) {
  fvm::Cells<S> C;
  C.np = c->np;
  C.nactive = c->nactive;
  auto D = [&](int f) { return (double *)c->field[f].p; };
  auto St = [&](int f) { return (S *)c->field[f].p; };
  auto I = [&](int f) { return (int *)c->field[f].p; };
  C.w = D(FVM_FIELD_W);
  C.x = St(FVM_FIELD_GRAD_X);
  C.y = St(FVM_FIELD_GRAD_Y);
  C.z = St(FVM_FIELD_GRAD_Z);
  C.t = St(FVM_FIELD_DWDT);
  C.pos = D(FVM_FIELD_POS);
  C.vel = St(FVM_FIELD_VEL);
  C.J = St(FVM_FIELD_J);
  C.tend = D(FVM_FIELD_TEND);
  C.bnd = I(FVM_FIELD_BND);
  C.is_ngb = I(FVM_FIELD_IS_NGB);
  C.is_active = I(FVM_FIELD_IS_ACTIVE);
  C.mm_lo = St(FVM_FIELD_MINMAX_LO);
  C.mm_hi = St(FVM_FIELD_MINMAX_HI);
  C.lim = St(FVM_FIELD_LIMITER);
  C.dU = St(FVM_FIELD_DU);
  C.dgp = St(FVM_FIELD_DGRADPSI);
  C.ddivB = St(FVM_FIELD_DDIVB);
  C.U = D(FVM_FIELD_U);
  C.mvol = D(FVM_FIELD_MESH_VOLUME);
  C.tbeg = D(FVM_FIELD_TBEG);
  C.mvel = D(FVM_FIELD_MESH_VEL);
  C.acc0 = D(FVM_FIELD_ACC0);
  C.acc1 = D(FVM_FIELD_ACC1);
  C.dt_new = D(FVM_FIELD_DT_NEW);
  C.divB = D(FVM_FIELD_DIVB);
  return C;
}

static fvm::MeshView meshview(fvm_ctx *c) {
  fvm::MeshView M;
  M.nface = c->nface;
  M.nfa = c->nfa < 0 ? 0 : c->nfa;
  M.s1 = (const int *)c->m_s1.p;
  M.s2 = (const int *)c->m_s2.p;
  M.c = (const double *)c->m_c.p;
  M.n = (const double *)c->m_n.p;
  M.cn = (const double *)c->m_cn.p;
  M.coff = (const int *)c->m_coff.p;
  M.cface = (const int *)c->m_cface.p;
  M.cvol = (const double *)c->m_cvol.p;
  M.fact = (const int *)c->m_fact.p;
  M.soff = (const int *)c->m_soff.p;
  M.sent = (const int *)c->m_sent.p;
  return M;
}

static fvm::Phys phys(const fvm_ctx *c) {
  fvm::Phys P;
  P.gamma = c->prm.gamma_gas;
  P.courant = c->prm.courant_no;
  P.t_global = c->t_global;
  for (int d = 0; d < 3; d++) P.L[d] = c->prm.domain_size[d];
  P.limiter = c->prm.limiter_mode;
  return P;
}

static inline unsigned nblk(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

#define LAUNCH(c, kern, n, ...)                                                        \
  do {                                                                                 \
    if ((n) > 0) {                                                                     \
      hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(256), 0, (c)->stream, __VA_ARGS__); \
      HIPCHK(c, hipGetLastError());                                                    \
    }                                                                                  \
  } while (0)
// the same with the grid padded to a multiple of 8 blocks (kernels indexed by xgid)
#define LAUNCH_X(c, kern, n, ...)                                                        \
  do {                                                                                 \
    if ((n) > 0) {                                                                     \
      hipLaunchKernelGGL(kern, dim3((nblk(n) + 7u) & ~7u), dim3(256), 0, (c)->stream, __VA_ARGS__); \
      HIPCHK(c, hipGetLastError());                                                    \
    }                                                                                  \
  } while (0)

// LAUNCH_X timed as kernel `id` when fvm_kernel_timing is on: events around the
// launch on the context's stream, waited for, the elapsed time added up
#define LAUNCH_T(c, id, kern, n, ...)                                                   \
  do {                                                                                  \
    if ((n) > 0 && (c)->kt.on) HIPCHK(c, hipEventRecord((c)->kt.ev[0], (c)->stream));   \
    LAUNCH_X(c, kern, n, __VA_ARGS__);                                                  \
    if ((n) > 0 && (c)->kt.on) {                                                        \
      HIPCHK(c, hipEventRecord((c)->kt.ev[1], (c)->stream));                            \
      HIPCHK(c, hipEventSynchronize((c)->kt.ev[1]));                                    \
      float ms_ = 0.0f;                                                                 \
      HIPCHK(c, hipEventElapsedTime(&ms_, (c)->kt.ev[0], (c)->kt.ev[1]));               \
      (c)->kt.ms[id] += ms_;                                                            \
      (c)->kt.launches[id] += 1;                                                        \
    }                                                                                   \
  } while (0)

static int check_params(const fvm_params *p, std::string &why) {
  if (!p) { why = "params is NULL"; return FVM_EARG; }
  if (!(p->gamma_gas > 0.0)) { why = "gamma_gas must be > 0"; return FVM_EARG; }
  if (!(p->courant_no > 0.0)) { why = "courant_no must be > 0"; return FVM_EARG; }
  for (int d = 0; d < 3; d++)
    if (!(p->domain_size[d] > 0.0)) { why = "domain_size must be > 0"; return FVM_EARG; }
  if (p->limiter_mode < 0 || p->limiter_mode > 2) { why = "limiter_mode out of range"; return FVM_EARG; }
  if (p->storage != FVM_STORAGE_F64 && p->storage != FVM_STORAGE_SHIPPED) { why = "storage out of range"; return FVM_EARG; }
  return FVM_OK;
}

// Entry points: C linkage comes from the declarations in fvm_gpu.h.

int fvm_create(fvm_ctx **out, int device, const fvm_params *params) {
  if (!out) return FVM_EARG;
  *out = nullptr;
  std::string why;
  if (check_params(params, why) != FVM_OK) return FVM_EARG;
  fvm_ctx *c = nullptr;
  try {
    c = new fvm_ctx;
  } catch (...) {
    return FVM_EARG;
  }
  c->device = device;
  c->prm = *params;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return FVM_EHIP;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return FVM_EHIP;
  }
  if (ensure(c, c->diag, sizeof(unsigned long long) * FVM_NDIAG) != FVM_OK ||
      ensure(c, c->scalars, sizeof(unsigned long long) * 8) != FVM_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FVM_EHIP;
  }
  (void)hipMemsetAsync(c->diag.p, 0, c->diag.bytes, c->stream);
  if (noalloc_scratch_init(device) != FVM_OK) {
    (void)fvm_destroy(c);
    return FVM_EHIP;
  }
  *out = c;
  return FVM_OK;
}

int fvm_destroy(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &b : c->field) release(b);
  for (DBuf *b : {&c->m_s1, &c->m_s2, &c->m_c, &c->m_n, &c->m_cn, &c->m_coff, &c->m_cface, &c->m_cvol, &c->m_fact,
                  &c->m_soff, &c->m_sent, &c->rec, &c->swp, &c->tmp0, &c->tmp1, &c->tmp2, &c->stage,
                  &c->diag, &c->scalars})
    release(*b);
  fvm::scan_release(c->scan);
  fvm_fast_release(c);
  fvm_substep_release(c);
  for (hipEvent_t &e : c->kt.ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FVM_OK;
}

int fvm_kernel_timing(fvm_ctx *c, int enable) {
  if (!c) return FVM_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  for (hipEvent_t &e : c->kt.ev)
    if (!e) HIPCHK(c, hipEventCreate(&e));
  c->kt.on = enable != 0;
  for (int k = 0; k < FVM_NKT; k++) {
    c->kt.ms[k] = 0.0;
    c->kt.launches[k] = 0;
  }
  return FVM_OK;
}

int fvm_kernel_ms(fvm_ctx *c, double *ms, long long *launches, int n) {
  if (!c) return FVM_EARG;
  if (n < 0 || n > FVM_NKT || (n > 0 && (!ms || !launches))) return fvm_fail(c, FVM_EARG, "fvm_kernel_ms: bad n or null output");
  for (int k = 0; k < n; k++) {
    ms[k] = c->kt.ms[k];
    launches[k] = c->kt.launches[k];
  }
  return FVM_OK;
}

const char *fvm_last_error(const fvm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fvm_set_params(fvm_ctx *c, const fvm_params *p) {
  if (!c) return FVM_EARG;
  std::string why;
  if (check_params(p, why) != FVM_OK) return fail(c, FVM_EARG, why);
  if (c->have_mesh && p->storage != c->prm.storage)
    return fail(c, FVM_ESTATE, "storage mode cannot change after fvm_upload_mesh");
  c->prm = *p;
  return FVM_OK;
}

int fvm_set_time(fvm_ctx *c, double t) {
  if (!c) return FVM_EARG;
  c->t_global = t;
  return FVM_OK;
}

namespace fvm {
// Complete cell lists: every face appears exactly once in the list of its s1
// when s1 is active and exactly once in the list of its s2 when s2 is active
// (localMesh.cpp:586-589).  Checked per (cell, face) pair; the lists must
// already be known to hold only faces their cell owns.
bool lists_complete(int nactive, int nface, const fvm_face *faces, const int *cell_face_off, const int *cell_face) {
  try {
    std::vector<unsigned char> under1(nface, 0), under2(nface, 0);
    for (int i = 0; i < nactive; i++)
      for (int e = cell_face_off[i]; e < cell_face_off[i + 1]; e++) {
        const int f = cell_face[e];
        unsigned char &u = faces[f].s1 == i ? under1[f] : under2[f];
        if (u) return false;  // listed twice under the same cell
        u = 1;
      }
    for (int f = 0; f < nface; f++)
      if (under1[f] != (faces[f].s1 < nactive) || under2[f] != (faces[f].s2 < nactive)) return false;
  } catch (...) {
    return false;
  }
  return true;
}
}  // namespace fvm

int fvm_upload_mesh(fvm_ctx *c, int np, int nactive, int nimport_loc, int nface, const fvm_face *faces,
                    const int *cell_face_off, const int *cell_face, const double *cell_volume) {
  if (!c) return FVM_EARG;
  if (np <= 0 || nactive < 0 || nactive > np || nimport_loc < 0 || nactive + nimport_loc > np || nface < 0)
    return fail(c, FVM_EARG, "fvm_upload_mesh: inconsistent sizes");
  if ((nface > 0 && !faces) || !cell_face_off || (nactive > 0 && !cell_volume))
    return fail(c, FVM_EARG, "fvm_upload_mesh: NULL array");
  // host-side validation of every index the kernels will follow
  const int nref = cell_face_off[nactive];
  if (cell_face_off[0] != 0 || nref < 0 || (nref > 0 && !cell_face))
    return fail(c, FVM_EARG, "fvm_upload_mesh: bad cell_face_off");
  for (int i = 0; i < nactive; i++)
    if (cell_face_off[i + 1] < cell_face_off[i]) return fail(c, FVM_EARG, "fvm_upload_mesh: cell_face_off not monotone");
  for (int e = 0; e < nref; e++)
    if (cell_face[e] < 0 || cell_face[e] >= nface) return fail(c, FVM_EARG, "fvm_upload_mesh: cell_face index out of range");
  for (int f = 0; f < nface; f++)
    if (faces[f].s1 < 0 || faces[f].s1 >= np || faces[f].s2 < 0 || faces[f].s2 >= np || faces[f].s1 == faces[f].s2)
      return fail(c, FVM_EARG, "fvm_upload_mesh: face cell index out of range");
  for (int i = 0; i < nactive; i++)
    for (int e = cell_face_off[i]; e < cell_face_off[i + 1]; e++) {
      const fvm_face &F = faces[cell_face[e]];
      if (F.s1 != i && F.s2 != i) return fail(c, FVM_EARG, "fvm_upload_mesh: cell lists a face it does not own");
    }
  bool complete = fvm::lists_complete(nactive, nface, faces, cell_face_off, cell_face);
  (void)hipSetDevice(c->device);
  fvm_fast_release(c);   // the fused state belongs to the previous mesh
  c->have_mesh = false;  // until the new mesh is complete
  c->lists_complete = complete;
  const bool resize = np != c->np;
  c->np = np;
  c->nactive = nactive;
  c->nimport_loc = nimport_loc;
  c->nface = nface;
  c->nref = nref;
  c->nfa = -1;
  for (int f = 0; f < FVM_NFIELDS; f++) {
    const size_t bytes = field_bytes(c, f, np);
    if (resize || c->field[f].bytes < bytes) {
      int rc = ensure(c, c->field[f], bytes);
      if (rc) return rc;
      HIPCHK(c, hipMemsetAsync(c->field[f].p, 0, bytes, c->stream));
    }
  }
  int rc;
  if ((rc = ensure(c, c->m_s1, sizeof(int) * (size_t)nface)) || (rc = ensure(c, c->m_s2, sizeof(int) * (size_t)nface)) ||
      (rc = ensure(c, c->m_c, sizeof(double) * 3 * (size_t)nface)) || (rc = ensure(c, c->m_n, sizeof(double) * 3 * (size_t)nface)) ||
      (rc = ensure(c, c->m_cn, sizeof(double) * 6 * (size_t)nface + 8)) ||
      (rc = ensure(c, c->m_coff, sizeof(int) * (size_t)(nactive + 1))) || (rc = ensure(c, c->m_cface, sizeof(int) * (size_t)nref)) ||
      (rc = ensure(c, c->m_cvol, sizeof(double) * (size_t)nactive)) || (rc = ensure(c, c->stage, sizeof(fvm_face) * (size_t)nface)))
    return rc;
  if (nface > 0) {
    HIPCHK(c, hipMemcpyAsync(c->stage.p, faces, sizeof(fvm_face) * (size_t)nface, hipMemcpyHostToDevice, c->stream));
    LAUNCH(c, fvm::k_faces_to_soa, nface, (const fvm_face *)c->stage.p, nface, (int *)c->m_s1.p, (int *)c->m_s2.p,
           (double *)c->m_c.p, (double *)c->m_n.p, (double *)c->m_cn.p);
  }
  HIPCHK(c, hipMemcpyAsync(c->m_coff.p, cell_face_off, sizeof(int) * (size_t)(nactive + 1), hipMemcpyHostToDevice, c->stream));
  if (nref > 0) HIPCHK(c, hipMemcpyAsync(c->m_cface.p, cell_face, sizeof(int) * (size_t)nref, hipMemcpyHostToDevice, c->stream));
  if (nactive > 0) HIPCHK(c, hipMemcpyAsync(c->m_cvol.p, cell_volume, sizeof(double) * (size_t)nactive, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_mesh = true;
  return FVM_OK;
}

int fvm_upload_faces(fvm_ctx *c, int np, int nactive, int nimport_loc, int nface, const fvm_face *faces,
                     const double *cell_volume) {
  if (!c) return FVM_EARG;
  if (np <= 0 || nactive < 0 || nactive > np || nimport_loc < 0 || nactive + nimport_loc > np || nface < 0)
    return fail(c, FVM_EARG, "fvm_upload_faces: inconsistent sizes");
  if ((nface > 0 && !faces) || (nactive > 0 && !cell_volume)) return fail(c, FVM_EARG, "fvm_upload_faces: NULL array");
  // every face is created by an active cell (its s1); s2 may be any local cell
  for (int f = 0; f < nface; f++)
    if (faces[f].s1 < 0 || faces[f].s1 >= nactive || faces[f].s2 < 0 || faces[f].s2 >= np || faces[f].s1 == faces[f].s2)
      return fail(c, FVM_EARG, "fvm_upload_faces: face cell index out of range (s1 must be active)");
  (void)hipSetDevice(c->device);
  fvm_fast_release(c);   // the fused state belongs to the previous mesh
  c->have_mesh = false;  // until the new mesh is complete
  c->lists_complete = true;  // built below, every face under each active cell
  const bool resize = np != c->np;
  for (int f = 0; f < FVM_NFIELDS; f++) {
    const size_t bytes = field_bytes(c, f, np);
    if (resize || c->field[f].bytes < bytes) {
      int rc = ensure(c, c->field[f], bytes);
      if (rc) return rc;
      HIPCHK(c, hipMemsetAsync(c->field[f].p, 0, bytes, c->stream));
    }
  }
  int rc;
  if ((rc = ensure(c, c->m_s1, sizeof(int) * (size_t)nface)) || (rc = ensure(c, c->m_s2, sizeof(int) * (size_t)nface)) ||
      (rc = ensure(c, c->m_c, sizeof(double) * 3 * (size_t)nface)) || (rc = ensure(c, c->m_n, sizeof(double) * 3 * (size_t)nface)) ||
      (rc = ensure(c, c->m_cn, sizeof(double) * 6 * (size_t)nface + 8)) ||
      (rc = ensure(c, c->m_coff, sizeof(int) * (size_t)(nactive + 1))) || (rc = ensure(c, c->m_cface, sizeof(int) * 2 * (size_t)nface + 4)) ||
      (rc = ensure(c, c->m_cvol, sizeof(double) * (size_t)nactive + 8)) || (rc = ensure(c, c->stage, sizeof(fvm_face) * (size_t)nface + 8)) ||
      (rc = ensure(c, c->tmp0, sizeof(int) * (size_t)(nactive + 1))) || (rc = ensure(c, c->tmp1, sizeof(int) * (size_t)(nactive + 1))))
    return rc;
  int *cnt = (int *)c->tmp0.p, *cursor = (int *)c->tmp1.p;
  LAUNCH(c, fvm::k_fill_int, nactive + 1, cnt, 0, nactive + 1);
  if (nface > 0) {
    HIPCHK(c, hipMemcpyAsync(c->stage.p, faces, sizeof(fvm_face) * (size_t)nface, hipMemcpyHostToDevice, c->stream));
    LAUNCH(c, fvm::k_faces_to_soa, nface, (const fvm_face *)c->stage.p, nface, (int *)c->m_s1.p, (int *)c->m_s2.p,
           (double *)c->m_c.p, (double *)c->m_n.p, (double *)c->m_cn.p);
    LAUNCH(c, fvm::k_csr_count, nface, (const int *)c->m_s1.p, (const int *)c->m_s2.p, nface, nactive, cnt);
  }
  if (fvm::exclusive_scan(c->scan, cnt, (int *)c->m_coff.p, nactive + 1, c->stream))
    return fail(c, FVM_EHIP, "fvm_upload_faces: scan failed");
  int nref = 0;
  HIPCHK(c, hipMemcpyAsync(&nref, (int *)c->m_coff.p + nactive, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(cursor, c->m_coff.p, sizeof(int) * (size_t)(nactive + 1), hipMemcpyDeviceToDevice, c->stream));
  if (nface > 0) {
    LAUNCH(c, fvm::k_csr_fill, nface, (const int *)c->m_s1.p, (const int *)c->m_s2.p, nface, nactive, cursor,
           (int *)c->m_cface.p);
    LAUNCH(c, fvm::k_csr_sort, nactive, (const int *)c->m_coff.p, nactive, (int *)c->m_cface.p);
  }
  if (nactive > 0) HIPCHK(c, hipMemcpyAsync(c->m_cvol.p, cell_volume, sizeof(double) * (size_t)nactive, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->np = np;
  c->nactive = nactive;
  c->nimport_loc = nimport_loc;
  c->nface = nface;
  c->nref = nref;
  c->nfa = -1;
  c->have_mesh = true;
  return FVM_OK;
}

int fvm_download_cell_faces(fvm_ctx *c, int *cell_face_off, int *cell_face) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "fvm_download_cell_faces before a mesh upload");
  if (!cell_face_off && !cell_face) return fail(c, FVM_EARG, "fvm_download_cell_faces: NULL arrays");
  (void)hipSetDevice(c->device);
  if (cell_face_off)
    HIPCHK(c, hipMemcpy(cell_face_off, c->m_coff.p, sizeof(int) * (size_t)(c->nactive + 1), hipMemcpyDeviceToHost));
  if (cell_face && c->nref > 0)
    HIPCHK(c, hipMemcpy(cell_face, c->m_cface.p, sizeof(int) * (size_t)c->nref, hipMemcpyDeviceToHost));
  return FVM_OK;
}

int fvm_upload_field(fvm_ctx *c, int field, const void *host, int np) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "fvm_upload_field before fvm_upload_mesh");
  if (field < 0 || field >= FVM_NFIELDS) return fail(c, FVM_EARG, "fvm_upload_field: unknown field");
  if (np != c->np || !host) return fail(c, FVM_EARG, "fvm_upload_field: np mismatch or NULL");
  (void)hipSetDevice(c->device);
  const auto &fi = fvm::kFields[field];
  if (fi.kind == fvm::K_INT) {
    HIPCHK(c, hipMemcpyAsync(c->field[field].p, host, sizeof(int) * (size_t)np, hipMemcpyHostToDevice, c->stream));
  } else {
    const size_t bytes = sizeof(double) * (size_t)fi.width * (size_t)np;
    int rc = ensure(c, c->stage, bytes);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->stage.p, host, bytes, hipMemcpyHostToDevice, c->stream));
    if (fi.kind == fvm::K_STORE && c->prm.storage == FVM_STORAGE_SHIPPED)
      LAUNCH(c, fvm::k_aos_to_soa<float>, np, (const double *)c->stage.p, (float *)c->field[field].p, np, fi.width);
    else
      LAUNCH(c, fvm::k_aos_to_soa<double>, np, (const double *)c->stage.p, (double *)c->field[field].p, np, fi.width);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_download_field(fvm_ctx *c, int field, void *host, int np) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "fvm_download_field before fvm_upload_mesh");
  if (field < 0 || field >= FVM_NFIELDS) return fail(c, FVM_EARG, "fvm_download_field: unknown field");
  if (np != c->np || !host) return fail(c, FVM_EARG, "fvm_download_field: np mismatch or NULL");
  (void)hipSetDevice(c->device);
  const auto &fi = fvm::kFields[field];
  if (fi.kind == fvm::K_INT) {
    HIPCHK(c, hipMemcpyAsync(host, c->field[field].p, sizeof(int) * (size_t)np, hipMemcpyDeviceToHost, c->stream));
  } else {
    const size_t bytes = sizeof(double) * (size_t)fi.width * (size_t)np;
    int rc = ensure(c, c->stage, bytes);
    if (rc) return rc;
    if (fi.kind == fvm::K_STORE && c->prm.storage == FVM_STORAGE_SHIPPED)
      LAUNCH(c, fvm::k_soa_to_aos<float>, np, (const float *)c->field[field].p, (double *)c->stage.p, np, fi.width);
    else
      LAUNCH(c, fvm::k_soa_to_aos<double>, np, (const double *)c->field[field].p, (double *)c->stage.p, np, fi.width);
    HIPCHK(c, hipMemcpyAsync(host, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

// face-end lists of every cell over the current face_active_list
static int build_face_ends(fvm_ctx *c) {
  const int np = c->np, nfa = c->nfa;
  int rc;
  if ((rc = ensure(c, c->m_soff, sizeof(int) * (size_t)(np + 1))) || (rc = ensure(c, c->m_sent, sizeof(int) * 2 * (size_t)nfa + 4)) ||
      (rc = ensure(c, c->tmp0, sizeof(int) * (size_t)(np + 1))))
    return rc;
  int *cnt = (int *)c->tmp0.p;
  HIPCHK(c, hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)(np + 1), c->stream));
  fvm::MeshView M = meshview(c);
  LAUNCH(c, fvm::k_end_count, nfa, M, cnt);
  rc = fvm::exclusive_scan(c->scan, cnt, (int *)c->m_soff.p, np + 1, c->stream);
  if (rc) return fail(c, FVM_EHIP, "scan failed");
  HIPCHK(c, hipMemcpyAsync(cnt, c->m_soff.p, sizeof(int) * (size_t)(np + 1), hipMemcpyDeviceToDevice, c->stream));
  M = meshview(c);
  LAUNCH(c, fvm::k_end_fill, nfa, M, cnt, (int *)c->m_sent.p);
  LAUNCH(c, fvm::k_end_sort, np, (const int *)c->m_soff.p, (int *)c->m_sent.p, np);
  return FVM_OK;
}

int fvm_get_active_faces(fvm_ctx *c, int *nface_active) {
  if (!c || !nface_active) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "fvm_get_active_faces before fvm_upload_mesh");
  (void)hipSetDevice(c->device);
  const int nref = c->nref;
  int rc;
  if ((rc = ensure(c, c->tmp1, sizeof(int) * ((size_t)std::max(c->nface, nref) + 1))) ||
      (rc = ensure(c, c->tmp2, sizeof(int) * ((size_t)nref + 1))) || (rc = ensure(c, c->tmp0, sizeof(int) * ((size_t)nref + 1))) ||
      (rc = ensure(c, c->m_fact, sizeof(int) * ((size_t)c->nface + 1))))
    return rc;
  int *firstpos = (int *)c->tmp1.p, *flag = (int *)c->tmp2.p, *pos = (int *)c->tmp0.p;
  fvm::MeshView M = meshview(c);
  if (c->lists_complete) {
    LAUNCH(c, fvm::k_face_firstflag_direct, c->nactive, M, (const int *)c->field[FVM_FIELD_BND].p, c->nactive, flag);
  } else {
    LAUNCH(c, fvm::k_fill_int, c->nface, firstpos, INT_MAX, c->nface);
    LAUNCH(c, fvm::k_face_firstpos, c->nactive, M, (const int *)c->field[FVM_FIELD_BND].p, c->nactive, firstpos);
    LAUNCH(c, fvm::k_face_firstflag, c->nactive, M, (const int *)c->field[FVM_FIELD_BND].p, c->nactive, firstpos, flag);
  }
  int total = 0;
  if (nref > 0) {
    rc = fvm::exclusive_scan_flags(c->scan, flag, pos, nref, &total, c->stream);
    if (rc) return fail(c, FVM_EHIP, "flag scan failed");
    LAUNCH(c, fvm::k_face_compact, nref, M, flag, pos, nref, (int *)c->m_fact.p);
  }
  c->nfa = total;
  rc = build_face_ends(c);
  if (rc) return rc;
  if ((rc = finish(c))) return rc;
  *nface_active = total;
  return FVM_OK;
}

int fvm_set_active_faces(fvm_ctx *c, int nfa, const int *face_active) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "fvm_set_active_faces before fvm_upload_mesh");
  if (nfa < 0 || nfa > c->nface || (nfa > 0 && !face_active)) return fail(c, FVM_EARG, "fvm_set_active_faces: bad size");
  for (int a = 0; a < nfa; a++)
    if (face_active[a] < 0 || face_active[a] >= c->nface) return fail(c, FVM_EARG, "fvm_set_active_faces: index out of range");
  (void)hipSetDevice(c->device);
  int rc = ensure(c, c->m_fact, sizeof(int) * ((size_t)nfa + 1));
  if (rc) return rc;
  if (nfa > 0) HIPCHK(c, hipMemcpyAsync(c->m_fact.p, face_active, sizeof(int) * (size_t)nfa, hipMemcpyHostToDevice, c->stream));
  c->nfa = nfa;
  rc = build_face_ends(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_download_active_faces(fvm_ctx *c, int *face_active, int n) {
  if (!c) return FVM_EARG;
  if (c->nfa < 0) return fail(c, FVM_ESTATE, "no active face list");
  if (n != c->nfa || (n > 0 && !face_active)) return fail(c, FVM_EARG, "fvm_download_active_faces: size mismatch");
  (void)hipSetDevice(c->device);
  if (n > 0) HIPCHK(c, hipMemcpyAsync(face_active, c->m_fact.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

#define DISPATCH(c, fn, ...)                                   \
  ((c)->prm.storage == FVM_STORAGE_SHIPPED ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

template <typename S> static int do_reconstruction(fvm_ctx *c) {
  if (c->nactive >= fvm::ONE_THREAD_MIN_CELLS)
    LAUNCH_T(c, FVM_KT_RECONSTRUCTION, fvm::k_reconstruction<S>, c->nactive, cells<S>(c), meshview(c), phys(c));
  else
    LAUNCH_T(c, FVM_KT_RECONSTRUCTION, fvm::k_reconstruction_v<S>, fvm::cell_var_threads(c->nactive), cells<S>(c),
             meshview(c), phys(c));
  return finish(c);
}
int fvm_compute_reconstruction(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "no mesh");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_reconstruction, c);
}

template <typename S> static int do_limiter(fvm_ctx *c) {
  if (c->np >= fvm::ONE_THREAD_MIN_CELLS)
    LAUNCH_T(c, FVM_KT_SLOPE_LIMITER, fvm::k_slope_limiter<S>, c->np, cells<S>(c), meshview(c), phys(c),
             (unsigned long long *)c->diag.p);
  else
    LAUNCH_T(c, FVM_KT_SLOPE_LIMITER, fvm::k_slope_limiter_v<S>, fvm::cell_var_threads(c->np), cells<S>(c), meshview(c),
             phys(c), (unsigned long long *)c->diag.p);
  return finish(c);
}
int fvm_slope_limiter(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh || c->nfa < 0) return fail(c, FVM_ESTATE, "no mesh or no active face list");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_limiter, c);
}

static int read_diag(fvm_ctx *c, unsigned long long *h) {
  HIPCHK(c, hipMemcpyAsync(h, c->diag.p, sizeof(unsigned long long) * FVM_NDIAG, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

template <typename S> static int do_fluid_update(fvm_ctx *c) {
  const int nfa = c->nfa;
  int rc;
  if ((rc = ensure(c, c->rec, sizeof(double) * fvm::NREC_P * (size_t)nfa)) || (rc = ensure(c, c->swp, sizeof(int) * (size_t)nfa)))
    return rc;
  unsigned long long before[FVM_NDIAG], after[FVM_NDIAG];
  if (!c->defer_sync && (rc = read_diag(c, before))) return rc;
  if (c->face32)
    LAUNCH_T(c, FVM_KT_FACE_FLUX, (fvm::k_face_flux<S, float>), nfa, cells<S>(c), meshview(c), phys(c), (double *)c->rec.p,
             (int *)c->swp.p, (unsigned long long *)c->diag.p);
  else
    LAUNCH_T(c, FVM_KT_FACE_FLUX, (fvm::k_face_flux<S, double>), nfa, cells<S>(c), meshview(c), phys(c), (double *)c->rec.p,
             (int *)c->swp.p, (unsigned long long *)c->diag.p);
  if (c->np >= fvm::ONE_THREAD_MIN_CELLS)
    LAUNCH_T(c, FVM_KT_SCATTER, fvm::k_scatter<S>, c->np, cells<S>(c), meshview(c), (const double *)c->rec.p,
             (const int *)c->swp.p);
  else
    LAUNCH_T(c, FVM_KT_SCATTER, fvm::k_scatter_v<S>, fvm::cell_var_threads(c->np), cells<S>(c), meshview(c),
             (const double *)c->rec.p, (const int *)c->swp.p);
  if (c->defer_sync) return FVM_OK;  // the substep compares the positivity count itself
  if ((rc = read_diag(c, after))) return rc;
  if (after[FVM_DIAG_POSITIVITY] != before[FVM_DIAG_POSITIVITY])
    return fail(c, FVM_EPOS, "computeFluidUpdate: non-positive density or thermal energy at a face");
  return FVM_OK;
}
int fvm_set_face_precision(fvm_ctx *c, int bits) {
  if (!c) return FVM_EARG;
  if (bits != 32 && bits != 64) return fail(c, FVM_EARG, "fvm_set_face_precision: bits must be 32 or 64");
  c->face32 = bits == 32;
  return FVM_OK;
}
int fvm_compute_fluid_update(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh || c->nfa < 0) return fail(c, FVM_ESTATE, "no mesh or no active face list");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_fluid_update, c);
}

template <typename S>
static int do_face_states(fvm_ctx *c, int first, int count, double *wij, double *Wi, double *Wj, int *jbnd,
                          double *normal) {
  const size_t N = (size_t)count;
  const size_t nd = (3 + 10 + 10 + 3) * N;
  int rc = ensure(c, c->stage, sizeof(double) * nd + sizeof(int) * N);
  if (rc) return rc;
  double *d_wij = (double *)c->stage.p, *d_Wi = d_wij + 3 * N, *d_Wj = d_Wi + 10 * N, *d_n = d_Wj + 10 * N;
  int *d_jb = (int *)(d_n + 3 * N);
  LAUNCH(c, fvm::k_face_states<S>, count, cells<S>(c), meshview(c), phys(c), first, count, d_wij, d_Wi, d_Wj, d_jb,
         d_n);
  HIPCHK(c, hipMemcpyAsync(wij, d_wij, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Wi, d_Wi, sizeof(double) * 10 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Wj, d_Wj, sizeof(double) * 10 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(normal, d_n, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(jbnd, d_jb, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}
int fvm_get_face_states(fvm_ctx *c, int first, int count, double *wij, double *Wi, double *Wj, int *jbnd,
                        double *normal) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh || c->nfa < 0) return fail(c, FVM_ESTATE, "no mesh or no active face list");
  if (first < 0 || count < 0 || first + count > c->nfa) return fail(c, FVM_EARG, "face range outside the active list");
  if (count == 0) return FVM_OK;
  if (!wij || !Wi || !Wj || !jbnd || !normal) return fail(c, FVM_EARG, "NULL array");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_face_states, c, first, count, wij, Wi, Wj, jbnd, normal);
}

template <typename S> static int do_primitives(fvm_ctx *c, int flag) {
  LAUNCH_T(c, FVM_KT_PRIMITIVES, fvm::k_primitives<S>, c->nactive, cells<S>(c), meshview(c), phys(c), flag, (unsigned long long *)c->diag.p);
  return finish(c);
}
int fvm_compute_primitives(fvm_ctx *c, int predictor_flag) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "no mesh");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_primitives, c, predictor_flag);
}

template <typename S> static int do_timestep(fvm_ctx *c, double *dt_min) {
  unsigned long long *bits = (unsigned long long *)c->scalars.p;
  LAUNCH(c, fvm::k_set_u64, 1, bits, 0x7FF0000000000000ULL);  // +inf
  LAUNCH_T(c, FVM_KT_TIMESTEP, fvm::k_timestep<S>, c->nactive, cells<S>(c), meshview(c), phys(c), bits);
  if (c->defer_sync && !dt_min) return FVM_OK;
  unsigned long long h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, bits, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (dt_min) std::memcpy(dt_min, &h, sizeof(double));
  return FVM_OK;
}
int fvm_compute_timestep(fvm_ctx *c, double *dt_min) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "no mesh");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_timestep, c, dt_min);
}

int fvm_compute_timestep_limiter(fvm_ctx *c) {
  // Sequential neighbour sweep of computeTimestep.cpp:64-110 (its result
  // depends on the visiting order): runs on the host next to the Scheduler.
  if (!c) return FVM_EARG;
  if (!c->have_mesh) return fail(c, FVM_ESTATE, "no mesh");
  try {
    const int np = c->np, nactive = c->nactive;
    std::vector<double> dt_new(np), tend(np), tbeg(np);
    std::vector<int> is_ngb(np), is_active(np), coff(nactive + 1), cface(c->nref), s1(c->nface), s2(c->nface);
    int rc;
    if ((rc = fvm_download_field(c, FVM_FIELD_DT_NEW, dt_new.data(), np)) || (rc = fvm_download_field(c, FVM_FIELD_TEND, tend.data(), np)) ||
        (rc = fvm_download_field(c, FVM_FIELD_TBEG, tbeg.data(), np)) || (rc = fvm_download_field(c, FVM_FIELD_IS_NGB, is_ngb.data(), np)) ||
        (rc = fvm_download_field(c, FVM_FIELD_IS_ACTIVE, is_active.data(), np)))
      return rc;
    HIPCHK(c, hipMemcpy(coff.data(), c->m_coff.p, sizeof(int) * (size_t)(nactive + 1), hipMemcpyDeviceToHost));
    if (c->nref) HIPCHK(c, hipMemcpy(cface.data(), c->m_cface.p, sizeof(int) * (size_t)c->nref, hipMemcpyDeviceToHost));
    if (c->nface) {
      HIPCHK(c, hipMemcpy(s1.data(), c->m_s1.p, sizeof(int) * (size_t)c->nface, hipMemcpyDeviceToHost));
      HIPCHK(c, hipMemcpy(s2.data(), c->m_s2.p, sizeof(int) * (size_t)c->nface, hipMemcpyDeviceToHost));
    }
    std::vector<double> dti(np, -fvm::REF_HUGE);
    for (int i = 0; i < nactive; i++) dti[i] = dt_new[i];
    for (int i = nactive; i < np; i++)
      if (is_ngb[i]) dti[i] = is_active[i] ? dt_new[i] : tend[i] - tbeg[i];
    const double scale_fac = 2.0;
    for (int i = 0; i < nactive; i++)
      for (int e = coff[i]; e < coff[i + 1]; e++) {
        const int f = cface[e];
        const int j = (i == s1[f]) ? s2[f] : s1[f];
        dti[j] = fvm::smin_h(dti[j], scale_fac * dti[i]);
        dti[i] = fvm::smin_h(dti[i], scale_fac * dti[j]);
      }
    for (int i = 0; i < np; i++)
      if (i < nactive || is_ngb[i]) dt_new[i] = dti[i];
    return fvm_upload_field(c, FVM_FIELD_DT_NEW, dt_new.data(), np);
  } catch (...) {
    return fail(c, FVM_EARG, "fvm_compute_timestep_limiter: host allocation failed");
  }
}

template <typename S> static int do_predictor(fvm_ctx *c, double dtmax, int min_rung) {
  const int nfa = c->nfa;
  int rc;
  if ((rc = ensure(c, c->rec, sizeof(double) * fvm::NREC_P * (size_t)nfa)) || (rc = ensure(c, c->swp, sizeof(int) * (size_t)nfa)))
    return rc;
  if (c->face32)
    LAUNCH_T(c, FVM_KT_PREDICTOR_FACE, (fvm::k_predictor_face<S, float>), nfa, cells<S>(c), meshview(c), phys(c),
             (double *)c->rec.p, (int *)c->swp.p, (unsigned long long *)c->diag.p);
  else
    LAUNCH_T(c, FVM_KT_PREDICTOR_FACE, (fvm::k_predictor_face<S, double>), nfa, cells<S>(c), meshview(c), phys(c),
             (double *)c->rec.p, (int *)c->swp.p, (unsigned long long *)c->diag.p);
  if (c->nactive >= fvm::ONE_THREAD_MIN_CELLS)
    LAUNCH_T(c, FVM_KT_PREDICTOR_CELL, fvm::k_predictor_cell<S>, c->nactive, cells<S>(c), meshview(c), phys(c),
             (const double *)c->rec.p, (const int *)c->swp.p, dtmax, min_rung, (unsigned long long *)c->diag.p);
  else
    LAUNCH_T(c, FVM_KT_PREDICTOR_CELL, fvm::k_predictor_cell_v<S>, fvm::cell_var_threads(c->nactive), cells<S>(c),
             meshview(c), phys(c), (const double *)c->rec.p, (const int *)c->swp.p, dtmax, min_rung,
             (unsigned long long *)c->diag.p);
  return finish(c);
}
int fvm_compute_predictor(fvm_ctx *c, double dtmax, int min_rung) {
  if (!c) return FVM_EARG;
  if (!c->have_mesh || c->nfa < 0) return fail(c, FVM_ESTATE, "no mesh or no active face list");
  if (!(dtmax > 0.0) || min_rung < 0 || min_rung > 31) return fail(c, FVM_EARG, "bad dtmax/min_rung");
  (void)hipSetDevice(c->device);
  return DISPATCH(c, do_predictor, c, dtmax, min_rung);
}

// the batched computeFlux through device staging at base (room for n faces:
// FVM_FLUX_BYTES_PER_FACE each): inputs in, one launch, outputs out, synchronous
static int flux_batch_staged(fvm_ctx *c, char *base, int n, const double *wij, const double *Wi, const double *Wj,
                             const int *jbnd, const double *normal, double *psi, double *Bn, double *flux) {
  const size_t N = (size_t)n;
  double *d_wij = (double *)base, *d_Wi = d_wij + 3 * N, *d_Wj = d_Wi + 10 * N, *d_n = d_Wj + 10 * N, *d_psi = d_n + 3 * N,
         *d_Bn = d_psi + N, *d_flux = d_Bn + N;
  int *d_jb = (int *)(d_flux + 10 * N);
  HIPCHK(c, hipMemcpyAsync(d_wij, wij, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_Wi, Wi, sizeof(double) * 10 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_Wj, Wj, sizeof(double) * 10 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_n, normal, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_jb, jbnd, sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
  LAUNCH(c, fvm::k_flux_batch, n, n, c->prm.gamma_gas, d_wij, d_Wi, d_Wj, d_jb, d_n, d_psi, d_Bn, d_flux);
  HIPCHK(c, hipMemcpyAsync(psi, d_psi, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(Bn, d_Bn, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(flux, d_flux, sizeof(double) * 10 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_compute_flux(fvm_ctx *c, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                     const double *normal, double *psi, double *Bn, double *flux) {
  if (!c) return FVM_EARG;
  if (n < 0) return fail(c, FVM_EARG, "n < 0");
  if (n == 0) return FVM_OK;
  if (!wij || !Wi || !Wj || !jbnd || !normal || !psi || !Bn || !flux) return fail(c, FVM_EARG, "NULL array");
  (void)hipSetDevice(c->device);
  int rc = ensure(c, c->stage, (size_t)FVM_FLUX_BYTES_PER_FACE * (size_t)n);
  if (rc) return rc;
  return flux_batch_staged(c, (char *)c->stage.p, n, wij, Wi, Wj, jbnd, normal, psi, Bn, flux);
}

int fvm_compute_flux_noalloc(fvm_ctx *c, int n, const double *wij, const double *Wi, const double *Wj,
                             const int *jbnd, const double *normal, double *psi, double *Bn, double *flux) {
  if (!c) return FVM_EARG;
  if (n < 0) return fail(c, FVM_EARG, "n < 0");
  if (n > FVM_FLUX_NOALLOC_MAX)
    return fail(c, FVM_EARG, "fvm_compute_flux_noalloc: n = " + std::to_string(n) + " exceeds FVM_FLUX_NOALLOC_MAX = " +
                                 std::to_string(FVM_FLUX_NOALLOC_MAX) + " faces");
  if (n == 0) return FVM_OK;
  if (!wij || !Wi || !Wj || !jbnd || !normal || !psi || !Bn || !flux) return fail(c, FVM_EARG, "NULL array");
  (void)hipSetDevice(c->device);
  char *base = (char *)noalloc_scratch(c->device);
  if (!base) return fail(c, FVM_ESTATE, "fvm_compute_flux_noalloc: no staging address for this device");
  return flux_batch_staged(c, base, n, wij, Wi, Wj, jbnd, normal, psi, Bn, flux);
}

int fvm_get_diagnostics(fvm_ctx *c, long long *out, int n) {
  if (!c || !out || n < 0) return FVM_EARG;
  (void)hipSetDevice(c->device);
  unsigned long long h[FVM_NDIAG];
  int rc = read_diag(c, h);
  if (rc) return rc;
  for (int k = 0; k < n && k < FVM_NDIAG; k++) out[k] = (long long)h[k];
  return FVM_OK;
}

int fvm_reset_diagnostics(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemsetAsync(c->diag.p, 0, c->diag.bytes, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_synchronize(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

