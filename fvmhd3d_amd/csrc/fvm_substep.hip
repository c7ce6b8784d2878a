// Individual timesteps with the state resident on the device.
//
// The host driver of the reference builds, for every substep, a local index
// layout (active cells first, then the inactive cells that share a face with
// one; localMesh.cpp:216-322, faces oriented so that s1 is active,
// localMesh.cpp:497-591), predicts the inactive neighbours to the current
// time (Iterate.cpp:99-106), runs the path on that subset and adds the
// neighbours' dU back to their owners (computeFluidUpdate.cpp:320-345).  Here
// the global per-cell state stays in HBM in mesh order and the layout is built
// by marks and ballot scans on the device, so a substep moves only the
// handful of counts the host's rung clock needs.  The path itself runs through
// the same reference-order entry points as a host-built layout, and every
// step here is a copy, an integer compaction or the same FP64 expression as
// the host driver, so both give bit-identical states.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"

#define SFLD(p, k, i, n) (p)[(size_t)(k) * (size_t)(n) + (size_t)(i)]

namespace {

constexpr int kRungMax = 32;  // Scheduler.h: RUNGMAX
constexpr int kBlock = 256;

inline unsigned nblocks(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
__device__ inline int gidx() { return blockIdx.x * blockDim.x + threadIdx.x; }

// global fields that persist between substeps; the others are rebuilt for
// every substep (velocities and accelerations zero, limiter one, flags)
constexpr int kKept[] = {FVM_FIELD_W,         FVM_FIELD_GRAD_X, FVM_FIELD_GRAD_Y, FVM_FIELD_GRAD_Z,
                         FVM_FIELD_DWDT,      FVM_FIELD_POS,    FVM_FIELD_J,      FVM_FIELD_TEND,
                         FVM_FIELD_BND,       FVM_FIELD_MINMAX_LO, FVM_FIELD_MINMAX_HI, FVM_FIELD_DU,
                         FVM_FIELD_DGRADPSI,  FVM_FIELD_DDIVB,  FVM_FIELD_U,      FVM_FIELD_MESH_VOLUME,
                         FVM_FIELD_TBEG,      FVM_FIELD_DT_NEW};
// copied back to the owner for active cells
constexpr int kBack[] = {FVM_FIELD_W,         FVM_FIELD_U,      FVM_FIELD_GRAD_X, FVM_FIELD_GRAD_Y,
                         FVM_FIELD_GRAD_Z,    FVM_FIELD_DWDT,   FVM_FIELD_MINMAX_LO, FVM_FIELD_MINMAX_HI,
                         FVM_FIELD_J,         FVM_FIELD_MESH_VOLUME, FVM_FIELD_TBEG, FVM_FIELD_TEND,
                         FVM_FIELD_DT_NEW,    FVM_FIELD_DU,     FVM_FIELD_DGRADPSI, FVM_FIELD_DDIVB};
// accumulated into the owner for inactive neighbours
constexpr int kAccum[] = {FVM_FIELD_DU, FVM_FIELD_DGRADPSI, FVM_FIELD_DDIVB};

// field widths (per-cell values) of the field table
constexpr int kWidth[FVM_NFIELDS] = {10, 10, 10, 10, 10, 3, 3, 3, 1, 1, 1, 1, 10,
                                     10, 10, 10, 3, 1, 10, 1, 1, 3, 3, 3, 1, 1};
constexpr bool kInt[FVM_NFIELDS] = {false, false, false, false, false, false, false, false, false,
                                    true,  true,  true,  false, false, false, false, false, false,
                                    false, false, false, false, false, false, false, false};

bool kept(int f) {
  for (int k : kKept)
    if (k == f) return true;
  return false;
}

// ---------------------------------------------------------------------------
// kernels

// active flag from the rung bitmask; an active cell must end its step now
__global__ void k_sub_active(const int *__restrict__ rung, const double *__restrict__ tend, int n, unsigned mask,
                         double t_global, int check, int *__restrict__ act, int *__restrict__ err) {
  const int i = gidx();
  if (i >= n) return;
  const int a = (int)((mask >> rung[i]) & 1u);
  act[i] = a;
  if (check && a && tend[i] != t_global) atomicOr(err, 1);
}

// out[base + pos[i]] = i for the flagged entries (ascending order)
__global__ void k_sub_compact(const int *__restrict__ flag, const int *__restrict__ pos, int n, int base,
                          int *__restrict__ out) {
  const int i = gidx();
  if (i >= n || !flag[i]) return;
  out[base + pos[i]] = i;
}

__global__ void k_sub_fill_i(int *__restrict__ p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// mark the faces of the active cells
__global__ void k_sub_mark_faces(const int *__restrict__ A, int nA, const int *__restrict__ goff,
                             const int *__restrict__ gface, int *__restrict__ fmark) {
  const int j = gidx();
  if (j >= nA) return;
  const int g = A[j];
  for (int e = goff[g]; e < goff[g + 1]; e++) fmark[gface[e]] = 1;
}

// mark both cells of every marked face
__global__ void k_sub_mark_cells(const int *__restrict__ fl, int nfl, const int *__restrict__ gs1,
                             const int *__restrict__ gs2, int *__restrict__ cmark) {
  const int j = gidx();
  if (j >= nfl) return;
  const int f = fl[j];
  cmark[gs1[f]] = 1;
  cmark[gs2[f]] = 1;
}

__global__ void k_sub_inactive_flags(const int *__restrict__ cmark, const int *__restrict__ act, int n,
                                 int *__restrict__ iflag) {
  const int i = gidx();
  if (i >= n) return;
  iflag[i] = cmark[i] && !act[i];
}

// global -> local index of every cell in the layout
__global__ void k_sub_g2l(const int *__restrict__ act, const int *__restrict__ apos, const int *__restrict__ iflag,
                      const int *__restrict__ ipos, int n, int nA, int *__restrict__ g2l) {
  const int i = gidx();
  if (i >= n) return;
  g2l[i] = act[i] ? apos[i] : (iflag[i] ? nA + ipos[i] : -1);
}

// local faces: the faces of the active cells in ascending global order,
// turned so that s1 is active with the centroid moved to the periodic image
// next to the new s1 (the same FP64 expressions as the host layout)
__global__ void k_sub_local_faces(const int *__restrict__ fl, int nfl, const int *__restrict__ gs1,
                              const int *__restrict__ gs2, const double *__restrict__ gc,
                              const double *__restrict__ gn, int nface, const double *__restrict__ pos, int n,
                              const int *__restrict__ act, const int *__restrict__ g2l, double bx, double by,
                              double bz, int *__restrict__ s1, int *__restrict__ s2, double *__restrict__ c,
                              double *__restrict__ nn, double *__restrict__ cn) {
  const int j = gidx();
  if (j >= nfl) return;
  const int f = fl[j];
  const int a = gs1[f], b = gs2[f];
  const bool swap = !act[a] && act[b];
  s1[j] = g2l[swap ? b : a];
  s2[j] = g2l[swap ? a : b];
  const double box[3] = {bx, by, bz};
  for (int d = 0; d < 3; d++) {
    double cd = SFLD(gc, d, f, nface), nd = SFLD(gn, d, f, nface);
    if (swap) {
      const double p = SFLD(pos, d, b, n);
      double dd = cd - p;
      dd = dd - rint(dd / box[d]) * box[d];
      cd = p + dd;
      nd = -nd;
    }
    SFLD(c, d, j, nfl) = cd;
    SFLD(nn, d, j, nfl) = nd;
    cn[(size_t)j * 6 + d] = cd;
    cn[(size_t)j * 6 + 3 + d] = nd;
  }
}

__global__ void k_sub_face_counts(const int *__restrict__ A, int nA, const int *__restrict__ goff, int *__restrict__ len) {
  const int j = gidx();
  if (j >= nA) return;
  len[j] = goff[A[j] + 1] - goff[A[j]];
}

// local cell lists (local face ids, in the order of the global list) and volumes
__global__ void k_sub_local_cells(const int *__restrict__ A, int nA, const int *__restrict__ goff,
                              const int *__restrict__ gface, const int *__restrict__ fpos,
                              const double *__restrict__ gvol, const int *__restrict__ coff, int *__restrict__ cface,
                              double *__restrict__ cvol) {
  const int j = gidx();
  if (j >= nA) return;
  const int g = A[j];
  const int e0 = goff[g], ne = goff[g + 1] - e0, o = coff[j];
  for (int e = 0; e < ne; e++) cface[o + e] = fpos[gface[e0 + e]];
  cvol[j] = gvol[g];
}




// new rungs of the active cells from dt_new (Scheduler.h get_rung with the
// driver's clamp, Iterate.cpp:146-168): exponent difference to dtmax, at
// least min_rung, at most nrung-1; tbeg/tend of the new step
__global__ void k_sub_rungs(const double *__restrict__ dt_new, double *__restrict__ tbeg, double *__restrict__ tend,
                        const int *__restrict__ perm, int nA, int emax, int min_rung, int nrung, double dt_tick,
                        double t_global, int *__restrict__ rung, int *__restrict__ err) {
  const int i = gidx();
  if (i >= nA) return;
  const unsigned long long bits = (unsigned long long)__double_as_longlong(dt_new[i]);
  const int e = (int)((bits >> 52) & 0xFFFull);
  int r = emax - e;
  if (r < min_rung) r = min_rung;
  if (r >= kRungMax) atomicOr(err, 2);
  if (r > nrung - 1) r = nrung - 1;
  if (r < min_rung) r = min_rung;
  rung[perm[i]] = r;
  tbeg[i] = t_global;
  tend[i] = t_global + ldexp(dt_tick, kRungMax - 1 - r);
}



// The context's per-cell fields of one substep, all in one launch (one row
// of blocks per field).  Per field:
enum : signed char {
  F_COPY = 0,     // kept field, copied for every local cell
  F_ACTIVE = 1,   // kept field, copied for active cells, zero for neighbours (fresh dU)
  F_PREDICT = 2,  // W: copied, neighbours predicted to the current time w + Wt*(t - tbeg)
  F_ZERO = 3,     // rebuilt: zero
  F_ONE = 4,      // rebuilt: one (the limiter)
  F_NGB = 5,      // int: 1 for neighbours
  F_ACT = 6,      // int: 1 for active cells
  F_SKIP = 7
};
struct FieldTable {
  const void *G[FVM_NFIELDS];
  void *L[FVM_NFIELDS];
  signed char mode[FVM_NFIELDS], width[FVM_NFIELDS], is_int[FVM_NFIELDS];
};
__global__ void k_sub_gather_all(FieldTable T, int n, int npl, int nA, const int *__restrict__ perm, double t_global,
                                 const double *__restrict__ gwt, const double *__restrict__ gtbeg) {
  const int i = gidx();
  const int f = (int)blockIdx.y;  // one row of blocks per field
  if (i >= npl) return;
  const int g = perm[i];
  {
    const int w = T.width[f];
    switch (T.mode[f]) {
      case F_COPY:
      case F_ACTIVE:
        if (T.is_int[f]) {
          for (int k = 0; k < w; k++)
            SFLD((int *)T.L[f], k, i, npl) =
                (T.mode[f] == F_ACTIVE && i >= nA) ? 0 : SFLD((const int *)T.G[f], k, g, n);
        } else {
          for (int k = 0; k < w; k++)
            SFLD((double *)T.L[f], k, i, npl) =
                (T.mode[f] == F_ACTIVE && i >= nA) ? 0.0 : SFLD((const double *)T.G[f], k, g, n);
        }
        break;
      case F_PREDICT: {
        const double *G = (const double *)T.G[f];
        double *L = (double *)T.L[f];
        if (i < nA) {
          for (int k = 0; k < w; k++) SFLD(L, k, i, npl) = SFLD(G, k, g, n);
        } else {
          const double dtp = t_global - gtbeg[g];
          for (int k = 0; k < w; k++) SFLD(L, k, i, npl) = SFLD(G, k, g, n) + SFLD(gwt, k, g, n) * dtp;
        }
        break;
      }
      case F_ZERO:
      case F_ONE:
        for (int k = 0; k < w; k++) SFLD((double *)T.L[f], k, i, npl) = T.mode[f] == F_ONE ? 1.0 : 0.0;
        break;
      case F_NGB:
        ((int *)T.L[f])[i] = i >= nA;
        break;
      case F_ACT:
        ((int *)T.L[f])[i] = i < nA;
        break;
      default:
        break;
    }
  }
}

// Write-back of one substep in one launch (one row of blocks per field): the
// active cells' fields go to their owners, the neighbours' fresh dU, gradPsi
// and divB are added.
struct BackTable {
  double *G[FVM_NFIELDS];
  const double *L[FVM_NFIELDS];
  signed char back[FVM_NFIELDS], accum[FVM_NFIELDS], width[FVM_NFIELDS];
};
__global__ void k_sub_writeback(BackTable T, int n, int npl, int nA, const int *__restrict__ perm) {
  const int i = gidx();
  const int f = (int)blockIdx.y;  // one row of blocks per field
  if (i >= npl) return;
  const int g = perm[i];
  {
    const int w = T.width[f];
    if (i < nA && T.back[f]) {
      for (int k = 0; k < w; k++) SFLD(T.G[f], k, g, n) = SFLD(T.L[f], k, i, npl);
    } else if (i >= nA && T.accum[f]) {
      for (int k = 0; k < w; k++) SFLD(T.G[f], k, g, n) += SFLD(T.L[f], k, i, npl);
    }
  }
}

// rung histogram of every cell (block-local counts first)
__global__ void k_sub_rung_hist(const int *__restrict__ rung, int n, int *__restrict__ hist) {
  __shared__ int h[kRungMax];
  if (threadIdx.x < kRungMax) h[threadIdx.x] = 0;
  __syncthreads();
  // a fixed grid strides over the cells, so few blocks add into the global
  // bins; one LDS add per distinct rung of a wavefront (nearby cells share few)
  const int lane = (int)(threadIdx.x & 63);
  const int stride = (int)(gridDim.x * blockDim.x);
  const int first = (int)(blockIdx.x * blockDim.x);
  for (int base = first; base < n; base += stride) {
    const int i = base + (int)threadIdx.x;
    int r = -1;
    if (i < n) {
      r = rung[i];
      r = (r >= 0 && r < kRungMax) ? r : kRungMax - 1;
    }
    unsigned long long left = __ballot(r >= 0);
    while (left) {
      const int lead = __ffsll((long long)left) - 1;
      const int rv = __shfl(r, lead);
      const unsigned long long same = __ballot(r == rv);
      if (lane == lead) atomicAdd(&h[rv], __popcll(same));
      left &= ~same;
    }
  }
  __syncthreads();
  if (threadIdx.x < kRungMax && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// initial values of the timestep-limiter sweep in the substep layout
// (computeTimestep.cpp:64-110): dt_new for active cells, tend - tbeg for the
// inactive neighbours.  The sweep's final write gives every neighbour its
// value, so dt_new takes it here; the cells the sweep changes are patched after.
__global__ void k_sub_sweep_init(double *__restrict__ dt_new, const double *__restrict__ tend,
                             const double *__restrict__ tbeg, int npl, int nA, double *__restrict__ sdti) {
  const int l = gidx();
  if (l >= npl) return;
  const double v = l < nA ? dt_new[l] : tend[l] - tbeg[l];
  sdti[l] = v;
  if (l >= nA) dt_new[l] = v;
}

// active cells with a face whose visit changes a value at the initial values:
// the visit of (i, j) sets dti[j] when 2 dti[i] < dti[j] and dti[i] when
// 2 dti[j] < dti[i]; with neither it changes nothing
__global__ void k_sub_hot(const int *__restrict__ coff, const int *__restrict__ cface, const int *__restrict__ s1,
                      const int *__restrict__ s2, const double *__restrict__ sdti, int nA, int *__restrict__ hot) {
  const int l = gidx();
  if (l >= nA) return;
  const double di = sdti[l];
  int h = 0;
  for (int e = coff[l]; e < coff[l + 1]; e++) {
    const int f = cface[e];
    const int j = (l == s1[f]) ? s2[f] : s1[f];
    const double dj = sdti[j];
    h |= (2.0 * di < dj) || (2.0 * dj < di);
  }
  hot[l] = h;
}

// changed values, by mesh cell, into the local dt_new
__global__ void k_sub_apply_changes(const int *__restrict__ gidx_, const double *__restrict__ val, int n,
                                    const int *__restrict__ g2l, double *__restrict__ dt_new) {
  const int k = gidx();
  if (k >= n) return;
  dt_new[g2l[gidx_[k]]] = val[k];
}

// the sweep's initial values and active flags in mesh order (the flags of
// cells outside the layout were cleared before)
// the sweep's values and each active cell's place in the visiting order, in mesh order
__global__ void k_sub_sweep_global(const double *__restrict__ sdti, const int *__restrict__ perm, int npl, int nA,
                                   double *__restrict__ dtig, int *__restrict__ rankg) {
  const int l = gidx();
  if (l >= npl) return;
  const int g = perm[l];
  dtig[g] = sdti[l];
  rankg[g] = l < nA ? l : -1;
}

// Sort key of an active cell: finest rung first, then push order
// (Scheduler.h pull_active_list).  Every stamp lies within 2^32 below the next
// one to be given, so seq - base, base = next - 2^32, takes 32 bits; the rung
// (below 32) takes the 5 above them, and the sort looks at 37 bits.
constexpr int kKeyBits = 37;
__device__ __forceinline__ unsigned long long order_key(int rung, unsigned long long seq, unsigned long long base) {
  return ((unsigned long long)(kRungMax - 1 - rung) << 32) | ((seq - base) & 0xFFFFFFFFull);
}
__global__ void k_sub_order_keys(const int *__restrict__ perm, int nA, const int *__restrict__ rung,
                                 const unsigned long long *__restrict__ seq, unsigned long long base,
                                 unsigned long long *__restrict__ key) {
  const int k = gidx();
  if (k >= nA) return;
  const int g = perm[k];
  key[k] = order_key(rung[g], seq[g], base);
}

// the same order for a short list, in one workgroup: bitonic sort in LDS of
// the keys padded to a power of two (keys are distinct: the stamps are)
constexpr int kSmallSort = 4096;
__global__ void __launch_bounds__(1024) k_sub_small_sort(int *__restrict__ perm, int nA, const int *__restrict__ rung,
                                                        const unsigned long long *__restrict__ seq,
                                                        unsigned long long base) {
  __shared__ unsigned long long key[kSmallSort];
  __shared__ int val[kSmallSort];
  int P = 1;
  while (P < nA) P <<= 1;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    if (i < nA) {
      const int g = perm[i];
      key[i] = order_key(rung[g], seq[g], base);
      val[i] = g;
    } else {
      key[i] = ~0ull;
      val[i] = -1;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          if ((key[i] > key[ixj]) == up) {
            const unsigned long long tk = key[i];
            key[i] = key[ixj];
            key[ixj] = tk;
            const int tv = val[i];
            val[i] = val[ixj];
            val[ixj] = tv;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < nA; i += blockDim.x) perm[i] = val[i];
}

// position of each active cell in the active list
__global__ void k_sub_apos(const int *__restrict__ perm, int nA, int *__restrict__ apos) {
  const int k = gidx();
  if (k >= nA) return;
  apos[perm[k]] = k;
}

// the active cells are pushed again, in the order of the active list
// (Iterate.cpp:146-168): their push stamps follow the last one given
__global__ void k_sub_stamp(const int *__restrict__ perm, int nA, unsigned long long base,
                            unsigned long long *__restrict__ seq) {
  const int k = gidx();
  if (k >= nA) return;
  seq[perm[k]] = base + (unsigned long long)k;
}

__global__ void k_sub_iota64(unsigned long long *__restrict__ p, int n) {
  const int i = gidx();
  if (i < n) p[i] = (unsigned long long)i;
}

// hot cells as mesh cells
// The sequential sweep of computeTimestep.cpp:64-110 over a whole mesh:
// active cells in the order of the active list, each cell's neighbours in
// list order,
//   dti[j] = min(dti[j], 2 dti[i]);  dti[i] = min(dti[i], 2 dti[j]).
// A visit changes something only if one of the visited cell's faces exceeds
// ratio 2 at the values of that moment.  So the sweep visits, in list order,
// the cells with such a face at the initial values ("hot"), every active cell
// after the current one whose own value changed, and every one whose face
// with a changed cell then exceeds ratio 2; every other visit is a no-op.
// The result equals the full sweep's bit for bit.
// off/nbr: neighbour of every face-end (mesh cells); rank: the place of each
// active cell in the list, -1 for the others; order: the list (rank -> cell);
// hot: the ranks of the hot cells, ascending; pending (one bit per rank) and
// changed: zero on entry and on return; changed_list receives the cells whose
// value changed.  Every rank queued lies after the one being visited, so the
// queue is a bitset read forward from a word cursor.
void sparse_sweep(const int *off, const int *nbr, const int *rank, const int *order, const int *hot, int nhot,
                  double *dti, unsigned long long *pending, size_t nwords, unsigned char *changed,
                  std::vector<int> &changed_list) {
  auto push = [&](int r) { pending[(size_t)r >> 6] |= 1ull << (r & 63); };
  const double scale_fac = 2.0;
  // Invariant: every active cell after the one being visited that is not
  // pending has all its faces within ratio 2 at the current values, so its
  // visit would change nothing.  When the value of c changes while visiting
  // the cell of rank rg, c itself (if active and after it) is queued, and of
  // its active neighbours after it those whose face with c now exceeds ratio 2.
  auto touched = [&](int c, int rg) {
    if (!changed[c]) {
      changed[c] = 1;
      changed_list.push_back(c);
    }
    if (rank[c] > rg) push(rank[c]);
    const double dc = dti[c];
    for (int e = off[c]; e < off[c + 1]; e++) {
      const int a = nbr[e];
      if (rank[a] > rg && (scale_fac * dti[a] < dc || scale_fac * dc < dti[a])) push(rank[a]);
    }
  };
  for (int k = 0; k < nhot; k++) push(hot[k]);
  size_t w = nhot > 0 ? (size_t)hot[0] >> 6 : nwords;
  for (;;) {
    while (w < nwords && pending[w] == 0) w++;
    if (w >= nwords) break;
    const int b = __builtin_ctzll(pending[w]);
    pending[w] &= pending[w] - 1;
    const int rg = (int)(w * 64 + (size_t)b);
    const int g = order[rg];
    for (int e = off[g]; e < off[g + 1]; e++) {
      const int j = nbr[e];
      if (scale_fac * dti[g] < dti[j]) {
        dti[j] = scale_fac * dti[g];
        touched(j, rg);
      }
      if (scale_fac * dti[j] < dti[g]) {
        dti[g] = scale_fac * dti[j];
        touched(g, rg);
      }
    }
  }
  for (int c : changed_list) changed[c] = 0;
}

}  // namespace

#define SCHK(c, call) FVM_HIPCHK(c, call)
#define SLAUNCH(c, n, kern, ...)                                                           \
  do {                                                                                     \
    if ((n) > 0) {                                                                         \
      hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), 0, (c)->stream, __VA_ARGS__); \
      SCHK(c, hipGetLastError());                                                          \
    }                                                                                      \
  } while (0)

void fvm_substep_release(fvm_ctx *c) {
  SubstepState &s = c->sub;
  for (DBuf *b : {&s.gs1, &s.gs2, &s.gc, &s.gn, &s.goff, &s.gface, &s.gvol, &s.rung, &s.act, &s.apos, &s.fmark,
                  &s.fpos, &s.fl, &s.cmark, &s.iflag, &s.ipos, &s.g2l, &s.perm, &s.len, &s.hist, &s.err, &s.sdti,
                  &s.hot, &s.hpos, &s.hlist, &s.chg, &s.dtig, &s.rankg, &s.seq, &s.okey,
                  &s.okey2, &s.oval2, &s.otmp})
    fvm_release(*b);
  for (auto &b : s.g) fvm_release(b);
  for (void *h : {(void *)s.dtih, (void *)s.rankh, (void *)s.sdh, (void *)s.permh, (void *)s.hoth, (void *)s.chgi,
                  (void *)s.chgv, (void *)s.errh, (void *)s.posh})
    if (h) (void)hipHostFree(h);
  s = SubstepState{};
}

namespace {

template <typename T>
T *P(DBuf &b) {
  return (T *)b.p;
}

int ensure_all(fvm_ctx *c, std::initializer_list<std::pair<DBuf *, size_t>> l) {
  for (auto &e : l) {
    int rc = fvm_ensure(c, *e.first, e.second ? e.second : 4);
    if (rc) return rc;
  }
  return FVM_OK;
}

// Build the local layout of one substep in the context's mesh and fields:
// active cells (rung in `mask`) in the Scheduler's order, then the inactive
// cells that share a face with them, in mesh order; faces of the active
// cells, s1 active.
int build_layout(fvm_ctx *c, unsigned mask, double t_global, int check, int *counts) {
  SubstepState &s = c->sub;
  const int n = s.n, nface = s.nface;
  hipStream_t st = c->stream;
  SLAUNCH(c, 1, k_sub_fill_i, P<int>(s.err), 1, 0);
  SLAUNCH(c, n, k_sub_active, P<int>(s.rung), P<double>(s.g[FVM_FIELD_TEND]), n, mask, t_global, check, P<int>(s.act),
          P<int>(s.err));
  int nA = 0, nfl = 0, nI = 0, nref = 0;
  // the flag's copy completes with the scan's own synchronisation
  SCHK(c, hipMemcpyAsync(s.errh, s.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  if (fvm::exclusive_scan_flags(c->scan, P<int>(s.act), P<int>(s.apos), n, &nA, st))
    return fvm_fail(c, FVM_EHIP, "substep: scan of active cells failed");
  if (*s.errh & 1) return fvm_fail(c, FVM_ESTATE, "substep: an active cell does not end its step at the current time");
  SLAUNCH(c, n, k_sub_compact, P<int>(s.act), P<int>(s.apos), n, 0, P<int>(s.perm));
  // the active list in the Scheduler's order: finest rung first, each rung in
  // push order (Scheduler.h pull_active_list)
  if (nA > 1) {
    const unsigned long long base = s.next_seq - (1ull << 32);
    if (nA <= kSmallSort) {
      hipLaunchKernelGGL(k_sub_small_sort, dim3(1), dim3(1024), 0, st, P<int>(s.perm), nA, P<int>(s.rung),
                         P<unsigned long long>(s.seq), base);
      SCHK(c, hipGetLastError());
    } else {
      SLAUNCH(c, nA, k_sub_order_keys, P<int>(s.perm), nA, P<int>(s.rung), P<unsigned long long>(s.seq), base,
              P<unsigned long long>(s.okey));
      size_t tb = s.otmp_bytes;
      SCHK(c, rocprim::radix_sort_pairs(s.otmp.p, tb, P<unsigned long long>(s.okey), P<unsigned long long>(s.okey2),
                                        P<int>(s.perm), P<int>(s.oval2), (unsigned)nA, 0, kKeyBits, st));
      SCHK(c, hipMemcpyAsync(s.perm.p, s.oval2.p, 4 * (size_t)nA, hipMemcpyDeviceToDevice, st));
    }
    SLAUNCH(c, nA, k_sub_apos, P<int>(s.perm), nA, P<int>(s.apos));
  }
  SLAUNCH(c, nface, k_sub_fill_i, P<int>(s.fmark), nface, 0);
  SLAUNCH(c, nA, k_sub_mark_faces, P<int>(s.perm), nA, P<int>(s.goff), P<int>(s.gface), P<int>(s.fmark));
  if (fvm::exclusive_scan_flags(c->scan, P<int>(s.fmark), P<int>(s.fpos), nface, &nfl, st))
    return fvm_fail(c, FVM_EHIP, "substep: scan of faces failed");
  SLAUNCH(c, nface, k_sub_compact, P<int>(s.fmark), P<int>(s.fpos), nface, 0, P<int>(s.fl));
  SLAUNCH(c, n, k_sub_fill_i, P<int>(s.cmark), n, 0);
  SLAUNCH(c, nfl, k_sub_mark_cells, P<int>(s.fl), nfl, P<int>(s.gs1), P<int>(s.gs2), P<int>(s.cmark));
  SLAUNCH(c, n, k_sub_inactive_flags, P<int>(s.cmark), P<int>(s.act), n, P<int>(s.iflag));
  if (fvm::exclusive_scan_flags(c->scan, P<int>(s.iflag), P<int>(s.ipos), n, &nI, st))
    return fvm_fail(c, FVM_EHIP, "substep: scan of neighbour cells failed");
  SLAUNCH(c, n, k_sub_compact, P<int>(s.iflag), P<int>(s.ipos), n, nA, P<int>(s.perm));
  SLAUNCH(c, n, k_sub_g2l, P<int>(s.act), P<int>(s.apos), P<int>(s.iflag), P<int>(s.ipos), n, nA, P<int>(s.g2l));
  const int npl = nA + nI;
  // the context's mesh (capacity for the whole mesh was reserved at setup);
  // not valid until the layout is complete
  c->have_mesh = false;
  fvm_fast_release(c);
  SLAUNCH(c, nfl, k_sub_local_faces, P<int>(s.fl), nfl, P<int>(s.gs1), P<int>(s.gs2), P<double>(s.gc), P<double>(s.gn),
          nface, P<double>(s.g[FVM_FIELD_POS]), n, P<int>(s.act), P<int>(s.g2l), c->prm.domain_size[0],
          c->prm.domain_size[1], c->prm.domain_size[2], P<int>(c->m_s1), P<int>(c->m_s2), P<double>(c->m_c),
          P<double>(c->m_n), P<double>(c->m_cn));
  SLAUNCH(c, nA, k_sub_face_counts, P<int>(s.perm), nA, P<int>(s.goff), P<int>(s.len));
  if (nA > 0) {
    if (fvm::exclusive_scan(c->scan, P<int>(s.len), P<int>(c->m_coff), nA, st))
      return fvm_fail(c, FVM_EHIP, "substep: scan of face counts failed");
    SCHK(c, hipMemcpyAsync(P<int>(c->m_coff) + nA, c->scan.total, sizeof(int), hipMemcpyDeviceToDevice, st));
    SCHK(c, hipMemcpyAsync(&nref, c->scan.total, sizeof(int), hipMemcpyDeviceToHost, st));
  } else {
    SLAUNCH(c, 1, k_sub_fill_i, P<int>(c->m_coff), 1, 0);
  }
  SLAUNCH(c, nA, k_sub_local_cells, P<int>(s.perm), nA, P<int>(s.goff), P<int>(s.gface), P<int>(s.fpos), P<double>(s.gvol),
          P<int>(c->m_coff), P<int>(c->m_cface), P<double>(c->m_cvol));
  SCHK(c, hipStreamSynchronize(st));
  if (npl <= 0 || nref < 0 || nref > s.nref || nfl > nface)
    return fvm_fail(c, FVM_ESTATE, "substep: empty or inconsistent layout");
  c->np = npl;
  c->nactive = nA;
  c->nimport_loc = nI;
  c->nface = nfl;
  c->nref = nref;
  c->nfa = -1;
  c->lists_complete = s.complete;  // active cells list all their faces when the global lists are complete
  c->have_mesh = true;
  c->t_global = t_global;
  // the context's fields, one launch
  FieldTable T{};
  for (int f = 0; f < FVM_NFIELDS; f++) {
    T.L[f] = c->field[f].p;
    T.G[f] = kept(f) ? s.g[f].p : nullptr;
    T.width[f] = (signed char)kWidth[f];
    T.is_int[f] = kInt[f];
    if (f == FVM_FIELD_W)
      T.mode[f] = F_PREDICT;
    else if (kept(f))
      T.mode[f] = (f == FVM_FIELD_DU || f == FVM_FIELD_DGRADPSI || f == FVM_FIELD_DDIVB) ? F_ACTIVE : F_COPY;
    else if (f == FVM_FIELD_LIMITER)
      T.mode[f] = F_ONE;
    else if (f == FVM_FIELD_IS_NGB)
      T.mode[f] = F_NGB;
    else if (f == FVM_FIELD_IS_ACTIVE)
      T.mode[f] = F_ACT;
    else
      T.mode[f] = kInt[f] ? F_SKIP : F_ZERO;
  }
  hipLaunchKernelGGL(k_sub_gather_all, dim3(nblocks(npl), FVM_NFIELDS), dim3(kBlock), 0, c->stream, T, n, npl, nA,
                     P<int>(s.perm), t_global, P<double>(s.g[FVM_FIELD_DWDT]), P<double>(s.g[FVM_FIELD_TBEG]));
  SCHK(c, hipGetLastError());
  if (counts) {
    counts[0] = nA;
    counts[1] = nI;
    counts[2] = nfl;
  }
  return FVM_OK;
}

int exponent_of(double x) {
  unsigned long long b;
  std::memcpy(&b, &x, sizeof b);
  return (int)((b >> 52) & 0xFFFull);
}

// rungs of the active cells, their tbeg/tend, and the histogram of all rungs
int update_rungs(fvm_ctx *c, double dtmax, int min_rung, int nrung, double t_global, int *hist) {
  SubstepState &s = c->sub;
  const double dt_tick = dtmax / (double)(1u << (kRungMax - 1));
  SLAUNCH(c, c->nactive, k_sub_rungs, P<double>(c->field[FVM_FIELD_DT_NEW]), P<double>(c->field[FVM_FIELD_TBEG]),
          P<double>(c->field[FVM_FIELD_TEND]), P<int>(s.perm), c->nactive, exponent_of(dtmax), min_rung, nrung, dt_tick,
          t_global, P<int>(s.rung), P<int>(s.err));
  // read back with the substep's final synchronisation
  SCHK(c, hipMemcpyAsync(s.errh + 1, s.err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  return FVM_OK;
}

int rung_histogram(fvm_ctx *c, int *hist) {
  SubstepState &s = c->sub;
  SLAUNCH(c, kRungMax, k_sub_fill_i, P<int>(s.hist), kRungMax, 0);
  hipLaunchKernelGGL(k_sub_rung_hist, dim3(std::min(nblocks(s.n), 1024u)), dim3(kBlock), 0, c->stream, P<int>(s.rung), s.n,
                     P<int>(s.hist));
  SCHK(c, hipGetLastError());
  if (hist) SCHK(c, hipMemcpyAsync(hist, s.hist.p, sizeof(int) * kRungMax, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int write_back(fvm_ctx *c, bool full) {
  SubstepState &s = c->sub;
  const int n = s.n, npl = c->np, nA = c->nactive;
  BackTable T{};
  for (int f : kBack)
    if (full || f == FVM_FIELD_TBEG || f == FVM_FIELD_TEND) T.back[f] = 1;
  if (full)
    for (int f : kAccum) T.accum[f] = 1;
  for (int f = 0; f < FVM_NFIELDS; f++) {
    T.width[f] = (signed char)kWidth[f];
    if (T.back[f] || T.accum[f]) {
      T.G[f] = P<double>(s.g[f]);
      T.L[f] = P<double>(c->field[f]);
    }
  }
  hipLaunchKernelGGL(k_sub_writeback, dim3(nblocks(npl), FVM_NFIELDS), dim3(kBlock), 0, c->stream, T, n, npl, nA,
                     P<int>(s.perm));
  SCHK(c, hipGetLastError());
  return FVM_OK;  // completes with the rung histogram's synchronisation, which always follows
}

// the timestep limiter of a substep: initial values and hot cells on the
// device, the sparse sweep on the host over the global adjacency, the changed
// values back into dt_new.  Large layouts hand the values over in mesh order
// (one contiguous copy each into pinned memory); small ones per layout cell.
int substep_sweep(fvm_ctx *c) {
  SubstepState &s = c->sub;
  hipStream_t st = c->stream;
  const int npl = c->np, nA = c->nactive, n = s.n;
  SLAUNCH(c, npl, k_sub_sweep_init, P<double>(c->field[FVM_FIELD_DT_NEW]), P<double>(c->field[FVM_FIELD_TEND]),
          P<double>(c->field[FVM_FIELD_TBEG]), npl, nA, P<double>(s.sdti));
  SLAUNCH(c, nA, k_sub_hot, P<int>(c->m_coff), P<int>(c->m_cface), P<int>(c->m_s1), P<int>(c->m_s2), P<double>(s.sdti),
          nA, P<int>(s.hot));
  int nhot = 0;
  if (fvm::exclusive_scan_flags(c->scan, P<int>(s.hot), P<int>(s.hpos), nA, &nhot, st))
    return fvm_fail(c, FVM_EHIP, "substep: scan of the limiter's cells failed");
  if (nhot == 0) return FVM_OK;
  // the hot cells' places in the active list, ascending
  SLAUNCH(c, nA, k_sub_compact, P<int>(s.hot), P<int>(s.hpos), nA, 0, P<int>(s.hlist));
  const bool global = (long long)npl * 8 > n;
  SCHK(c, hipMemcpyAsync(s.hoth, s.hlist.p, 4 * (size_t)nhot, hipMemcpyDeviceToHost, st));
  SCHK(c, hipMemcpyAsync(s.permh, s.perm.p, 4 * (size_t)(global ? nA : npl), hipMemcpyDeviceToHost, st));
  if (global) {
    SLAUNCH(c, n, k_sub_fill_i, P<int>(s.rankg), n, -1);
    SLAUNCH(c, npl, k_sub_sweep_global, P<double>(s.sdti), P<int>(s.perm), npl, nA, P<double>(s.dtig),
            P<int>(s.rankg));
    SCHK(c, hipMemcpyAsync(s.dtih, s.dtig.p, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    SCHK(c, hipMemcpyAsync(s.rankh, s.rankg.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
  } else {
    SCHK(c, hipMemcpyAsync(s.sdh, s.sdti.p, 8 * (size_t)npl, hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
    for (int l = 0; l < npl; l++) {
      const int g = s.permh[l];
      s.dtih[g] = s.sdh[l];
      s.rankh[g] = l < nA ? l : -1;
    }
  }
  std::vector<int> chg;
  sparse_sweep(s.hoff.data(), s.hnbr.data(), s.rankh, s.permh, s.hoth, nhot, s.dtih, s.pending.data(),
               s.pending.size(), s.changed.data(), chg);
  // rankh is -1 everywhere between substeps
  for (int l = 0; l < nA; l++) s.rankh[s.permh[l]] = -1;
  const int nchg = (int)chg.size();
  if (nchg > 0) {
    for (int k = 0; k < nchg; k++) {
      s.chgi[k] = chg[k];
      s.chgv[k] = s.dtih[chg[k]];
    }
    int *di = P<int>(s.chg);
    double *dv = (double *)((char *)s.chg.p + 4 * (((size_t)nchg + 1) & ~(size_t)1));  // reserved at setup: 12n + 8 bytes
    SCHK(c, hipMemcpyAsync(di, s.chgi, 4 * (size_t)nchg, hipMemcpyHostToDevice, st));
    SCHK(c, hipMemcpyAsync(dv, s.chgv, 8 * (size_t)nchg, hipMemcpyHostToDevice, st));
    SLAUNCH(c, nchg, k_sub_apply_changes, di, dv, nchg, P<int>(s.g2l), P<double>(c->field[FVM_FIELD_DT_NEW]));
  }
  // the copies from chgi / chgv complete before the next substep's host sweep
  // writes them again: every substep synchronises at its end
  return FVM_OK;
}

int check_ready(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->sub.ready) return fvm_fail(c, FVM_ESTATE, "substep call before fvm_substep_setup");
  if (c->prm.storage != FVM_STORAGE_F64) return fvm_fail(c, FVM_EARG, "device substeps need FVM_STORAGE_F64");
  (void)hipSetDevice(c->device);
  return FVM_OK;
}

}  // namespace

extern "C" {

int fvm_substep_setup(fvm_ctx *c, int ncell, int nface, const fvm_face *faces, const int *cell_face_off,
                      const int *cell_face, const double *cell_volume) {
  if (!c) return FVM_EARG;
  if (c->prm.storage != FVM_STORAGE_F64) return fvm_fail(c, FVM_EARG, "device substeps need FVM_STORAGE_F64");
  if (ncell <= 0 || nface < 0 || (nface > 0 && !faces) || !cell_face_off || !cell_volume)
    return fvm_fail(c, FVM_EARG, "fvm_substep_setup: bad sizes or NULL array");
  const int nref = cell_face_off[ncell];
  if (cell_face_off[0] != 0 || nref < 0 || (nref > 0 && !cell_face))
    return fvm_fail(c, FVM_EARG, "fvm_substep_setup: bad cell_face_off");
  for (int i = 0; i < ncell; i++)
    if (cell_face_off[i + 1] < cell_face_off[i]) return fvm_fail(c, FVM_EARG, "fvm_substep_setup: cell_face_off not monotone");
  for (int e = 0; e < nref; e++)
    if (cell_face[e] < 0 || cell_face[e] >= nface) return fvm_fail(c, FVM_EARG, "fvm_substep_setup: face index out of range");
  for (int f = 0; f < nface; f++)
    if (faces[f].s1 < 0 || faces[f].s1 >= ncell || faces[f].s2 < 0 || faces[f].s2 >= ncell || faces[f].s1 == faces[f].s2)
      return fvm_fail(c, FVM_EARG, "fvm_substep_setup: face cell index out of range");
  for (int i = 0; i < ncell; i++)
    for (int e = cell_face_off[i]; e < cell_face_off[i + 1]; e++)
      if (faces[cell_face[e]].s1 != i && faces[cell_face[e]].s2 != i)
        return fvm_fail(c, FVM_EARG, "fvm_substep_setup: cell lists a face it does not own");
  const bool complete = fvm::lists_complete(ncell, nface, faces, cell_face_off, cell_face);
  (void)hipSetDevice(c->device);
  fvm_substep_release(c);
  fvm_fast_release(c);  // the context's mesh is replaced by the substep layouts
  SubstepState &s = c->sub;
  s.complete = complete;
  const size_t N = (size_t)ncell, F = (size_t)nface;
  int rc;
  if ((rc = ensure_all(c, {{&s.gs1, 4 * F}, {&s.gs2, 4 * F}, {&s.gc, 24 * F}, {&s.gn, 24 * F}, {&s.goff, 4 * (N + 1)},
                           {&s.gface, 4 * (size_t)nref}, {&s.gvol, 8 * N}, {&s.rung, 4 * N}, {&s.act, 4 * N},
                           {&s.apos, 4 * N}, {&s.fmark, 4 * F}, {&s.fpos, 4 * F}, {&s.fl, 4 * F}, {&s.cmark, 4 * N},
                           {&s.iflag, 4 * N}, {&s.ipos, 4 * N}, {&s.g2l, 4 * N}, {&s.perm, 4 * N}, {&s.len, 4 * N},
                           {&s.hist, 4 * kRungMax}, {&s.err, 4}, {&s.sdti, 8 * N}, {&s.hot, 4 * N},
                           {&s.hpos, 4 * N}, {&s.hlist, 4 * N}, {&s.dtig, 8 * N}, {&s.rankg, 4 * N},
                           {&s.chg, 12 * N + 8}, {&s.seq, 8 * N}, {&s.okey, 8 * N}, {&s.okey2, 8 * N},
                           {&s.oval2, 4 * N}})))
    return rc;
  {
    // scratch of the active-list sort at its largest (every cell active)
    size_t tb = 0;
    SCHK(c, rocprim::radix_sort_pairs(nullptr, tb, P<unsigned long long>(s.okey), P<unsigned long long>(s.okey2),
                                      P<int>(s.perm), P<int>(s.oval2), (unsigned)ncell, 0, kKeyBits, c->stream));
    if ((rc = fvm_ensure(c, s.otmp, std::max<size_t>(tb, 256)))) return rc;
    s.otmp_bytes = tb;
  }
  SCHK(c, hipHostMalloc((void **)&s.dtih, 8 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.rankh, 4 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.sdh, 8 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.permh, 4 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.hoth, 4 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.chgi, 4 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.chgv, 8 * N, hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.errh, 2 * sizeof(int), hipHostMallocDefault));
  SCHK(c, hipHostMalloc((void **)&s.posh, 2 * sizeof(unsigned long long), hipHostMallocDefault));
  s.errh[0] = s.errh[1] = 0;
  s.posh[0] = s.posh[1] = 0;
  for (size_t i = 0; i < N; i++) s.rankh[i] = -1;
  try {
    s.hoff.assign(cell_face_off, cell_face_off + N + 1);
    s.hnbr.resize((size_t)nref);
    for (int i = 0; i < ncell; i++)
      for (int e = cell_face_off[i]; e < cell_face_off[i + 1]; e++) {
        const fvm_face &Fc = faces[cell_face[e]];
        s.hnbr[e] = (i == Fc.s1) ? Fc.s2 : Fc.s1;
      }
    s.pending.assign((N + 63) / 64, 0ull);
    s.changed.assign(N, 0);
  } catch (...) {
    return fvm_fail(c, FVM_EARG, "fvm_substep_setup: host allocation failed");
  }
  // the context's mesh, fields and the path's scratch at their largest (every
  // cell local), so that no substep grows a buffer
  const size_t R = (size_t)nref;
  if ((rc = ensure_all(c, {{&c->m_s1, 4 * F}, {&c->m_s2, 4 * F}, {&c->m_c, 24 * F}, {&c->m_n, 24 * F}, {&c->m_cn, 48 * F},
                           {&c->m_coff, 4 * (N + 1)}, {&c->m_cface, 4 * R}, {&c->m_cvol, 8 * N},
                           {&c->m_soff, 4 * (N + 1)}, {&c->m_sent, 8 * F + 4}, {&c->m_fact, 4 * (F + 1)},
                           {&c->tmp0, 4 * (std::max(N, R) + 1)}, {&c->tmp1, 4 * (std::max(F, R) + 1)},
                           {&c->tmp2, 4 * (R + 1)}, {&c->rec, 8 * 15 * F} /* 15-double face records */,
                           {&c->swp, 4 * F}})))
    return rc;
  for (int f = 0; f < FVM_NFIELDS; f++) {
    const size_t bytes = (kInt[f] ? 4 : 8) * (size_t)kWidth[f] * N;
    if ((rc = fvm_ensure(c, c->field[f], bytes))) return rc;
    if (kept(f)) {
      if ((rc = fvm_ensure(c, s.g[f], bytes))) return rc;
      SCHK(c, hipMemsetAsync(s.g[f].p, 0, bytes, c->stream));
    }
  }
  SCHK(c, hipMemsetAsync(s.rung.p, 0, 4 * N, c->stream));
  SCHK(c, hipStreamSynchronize(c->stream));
  if (nface > 0) {
    // same conversion as the mesh upload: ints and SoA doubles
    std::vector<int> a(F), b(F);
    std::vector<double> cc(3 * F), nn(3 * F);
    for (size_t f = 0; f < F; f++) {
      a[f] = faces[f].s1;
      b[f] = faces[f].s2;
      for (int d = 0; d < 3; d++) {
        cc[d * F + f] = faces[f].centroid[d];
        nn[d * F + f] = faces[f].n[d];
      }
    }
    SCHK(c, hipMemcpy(s.gs1.p, a.data(), 4 * F, hipMemcpyHostToDevice));
    SCHK(c, hipMemcpy(s.gs2.p, b.data(), 4 * F, hipMemcpyHostToDevice));
    SCHK(c, hipMemcpy(s.gc.p, cc.data(), 24 * F, hipMemcpyHostToDevice));
    SCHK(c, hipMemcpy(s.gn.p, nn.data(), 24 * F, hipMemcpyHostToDevice));
  }
  SCHK(c, hipMemcpy(s.goff.p, cell_face_off, 4 * (N + 1), hipMemcpyHostToDevice));
  if (nref > 0) SCHK(c, hipMemcpy(s.gface.p, cell_face, 4 * (size_t)nref, hipMemcpyHostToDevice));
  SCHK(c, hipMemcpy(s.gvol.p, cell_volume, 8 * N, hipMemcpyHostToDevice));
  SCHK(c, hipStreamSynchronize(c->stream));
  // first use of the host handover buffers and of the copy paths happens
  // here, once (about 6 ms on the first medium-sized copy), not in a substep
  std::memset(s.sdh, 0, 8 * N);
  std::memset(s.permh, 0, 4 * N);
  std::memset(s.hoth, 0, 4 * N);
  std::memset(s.chgi, 0, 4 * N);
  std::memset(s.chgv, 0, 8 * N);
  std::memset(s.dtih, 0, 8 * N);
  // the same asynchronous copies on the context's stream as the substeps make
  const void *zero = s.g[FVM_FIELD_DT_NEW].p;  // 8n zero bytes
  std::vector<int> minus1(N, -1);
  SCHK(c, hipMemcpyAsync(s.dtih, zero, 8 * N, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipMemcpyAsync(s.sdh, zero, 8 * N, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipMemcpyAsync(s.permh, zero, 4 * N, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipMemcpyAsync(s.hoth, zero, 4 * N, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipMemcpyAsync(s.rankh, zero, 4 * N, hipMemcpyDeviceToHost, c->stream));
  SCHK(c, hipMemcpyAsync(s.chg.p, s.chgi, 4 * N, hipMemcpyHostToDevice, c->stream));
  SCHK(c, hipMemcpyAsync(s.chg.p, s.chgv, 8 * N, hipMemcpyHostToDevice, c->stream));
  SCHK(c, hipStreamSynchronize(c->stream));
  std::memcpy(s.rankh, minus1.data(), 4 * N);  // -1 everywhere between substeps
  s.n = ncell;
  s.nface = nface;
  s.nref = nref;
  s.ready = true;
  return FVM_OK;
}

int fvm_substep_upload(fvm_ctx *c, int field, const void *host, int ncell) {
  if (int rc = check_ready(c)) return rc;
  SubstepState &s = c->sub;
  if (ncell != s.n || !host) return fvm_fail(c, FVM_EARG, "fvm_substep_upload: ncell mismatch or NULL");
  if (field == FVM_SUBSTEP_RUNG) {
    const int *r = (const int *)host;
    for (int i = 0; i < ncell; i++)
      if (r[i] < 0 || r[i] >= kRungMax) return fvm_fail(c, FVM_EARG, "fvm_substep_upload: rung out of range");
    SCHK(c, hipMemcpy(s.rung.p, host, 4 * (size_t)ncell, hipMemcpyHostToDevice));
    return FVM_OK;
  }
  if (field < 0 || field >= FVM_NFIELDS || !kept(field))
    return fvm_fail(c, FVM_EARG, "fvm_substep_upload: field is not kept between substeps");
  if (kInt[field]) {
    SCHK(c, hipMemcpy(s.g[field].p, host, 4 * (size_t)ncell, hipMemcpyHostToDevice));
    return FVM_OK;
  }
  const int w = kWidth[field];
  std::vector<double> soa((size_t)w * ncell);
  const double *h = (const double *)host;
  for (int i = 0; i < ncell; i++)
    for (int k = 0; k < w; k++) soa[(size_t)k * ncell + i] = h[(size_t)i * w + k];
  SCHK(c, hipMemcpy(s.g[field].p, soa.data(), 8 * soa.size(), hipMemcpyHostToDevice));
  return FVM_OK;
}

int fvm_substep_download(fvm_ctx *c, int field, void *host, int ncell) {
  if (int rc = check_ready(c)) return rc;
  SubstepState &s = c->sub;
  if (ncell != s.n || !host) return fvm_fail(c, FVM_EARG, "fvm_substep_download: ncell mismatch or NULL");
  SCHK(c, hipStreamSynchronize(c->stream));
  if (field == FVM_SUBSTEP_RUNG) {
    SCHK(c, hipMemcpy(host, s.rung.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost));
    return FVM_OK;
  }
  if (field < 0 || field >= FVM_NFIELDS || !kept(field))
    return fvm_fail(c, FVM_EARG, "fvm_substep_download: field is not kept between substeps");
  if (kInt[field]) {
    SCHK(c, hipMemcpy(host, s.g[field].p, 4 * (size_t)ncell, hipMemcpyDeviceToHost));
    return FVM_OK;
  }
  const int w = kWidth[field];
  std::vector<double> soa((size_t)w * ncell);
  SCHK(c, hipMemcpy(soa.data(), s.g[field].p, 8 * soa.size(), hipMemcpyDeviceToHost));
  double *h = (double *)host;
  for (int i = 0; i < ncell; i++)
    for (int k = 0; k < w; k++) h[(size_t)i * w + k] = soa[(size_t)k * ncell + i];
  return FVM_OK;
}

int fvm_substep_init_rungs(fvm_ctx *c, double dtmax, int nrung, int *rung_hist) {
  if (int rc = check_ready(c)) return rc;
  if (!(dtmax > 0.0) || nrung < 1 || nrung > kRungMax - 1) return fvm_fail(c, FVM_EARG, "fvm_substep_init_rungs: bad dtmax or nrung");
  int rc;
  // every cell pushed in index order (System.cpp:300-315)
  SLAUNCH(c, c->sub.n, k_sub_iota64, P<unsigned long long>(c->sub.seq), c->sub.n);
  c->sub.next_seq = (unsigned long long)c->sub.n;
  if ((rc = build_layout(c, 0xFFFFFFFFu, 0.0, 0, nullptr))) return rc;
  double dtmin = 0.0;
  if ((rc = fvm_compute_timestep(c, &dtmin))) return rc;
  if ((rc = update_rungs(c, dtmax, 0, nrung, 0.0, nullptr))) return rc;
  if ((rc = write_back(c, false))) return rc;
  return rung_histogram(c, rung_hist);
}

int fvm_substep_run(fvm_ctx *c, unsigned rung_mask, double t_global, double dtmax, int min_rung, int nrung, int *counts,
                    int *rung_hist) {
  if (int rc = check_ready(c)) return rc;
  if (!(dtmax > 0.0) || nrung < 1 || nrung > kRungMax - 1 || min_rung < 0 || min_rung >= kRungMax)
    return fvm_fail(c, FVM_EARG, "fvm_substep_run: bad dtmax, nrung or min_rung");
  int cnt[3] = {0, 0, 0}, nfa = 0, rc;
  SubstepState &s = c->sub;
  // the entry points below leave their synchronisations to this substep: it
  // waits for the counts that size its launches and once at its end
  struct Defer {
    fvm_ctx *c;
    ~Defer() { c->defer_sync = false; }
  } defer{c};
  c->defer_sync = true;
  if ((rc = build_layout(c, rung_mask, t_global, 1, cnt))) return rc;
  if ((rc = fvm_get_active_faces(c, &nfa))) return rc;
  if ((rc = fvm_compute_primitives(c, 0))) return rc;
  if ((rc = fvm_compute_timestep(c, nullptr))) return rc;
  if ((rc = substep_sweep(c))) return rc;
  if ((rc = fvm_compute_predictor(c, dtmax, min_rung))) return rc;
  if ((rc = update_rungs(c, dtmax, min_rung, nrung, t_global, nullptr))) return rc;
  if ((rc = fvm_compute_reconstruction(c))) return rc;
  if ((rc = fvm_slope_limiter(c))) return rc;
  unsigned long long *diag = P<unsigned long long>(c->diag) + FVM_DIAG_POSITIVITY;
  SCHK(c, hipMemcpyAsync(s.posh, diag, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  if ((rc = fvm_compute_fluid_update(c))) return rc;
  SCHK(c, hipMemcpyAsync(s.posh + 1, diag, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  if ((rc = write_back(c, true))) return rc;
  // the active cells pushed again in the order of the active list
  SLAUNCH(c, cnt[0], k_sub_stamp, P<int>(s.perm), cnt[0], s.next_seq, P<unsigned long long>(s.seq));
  s.next_seq += (unsigned long long)cnt[0];
  if ((rc = rung_histogram(c, rung_hist))) return rc;  // synchronises
  if (s.errh[1] & 2) return fvm_fail(c, FVM_ESTATE, "substep: a timestep is below the finest rung of the clock");
  if (counts) {
    counts[0] = cnt[0];
    counts[1] = nfa;
    counts[2] = cnt[1];
    counts[3] = cnt[2];
    counts[4] = s.posh[1] != s.posh[0];
  }
  return FVM_OK;
}

int fvm_timestep_limiter_sweep(int ncell, const int *cell_off, const int *nbr, const unsigned char *active,
                               double *dti) {
  if (ncell <= 0 || !cell_off || !active || !dti || cell_off[0] != 0) return FVM_EARG;
  for (int i = 0; i < ncell; i++)
    if (cell_off[i + 1] < cell_off[i]) return FVM_EARG;
  const int nref = cell_off[ncell];
  if (nref > 0 && !nbr) return FVM_EARG;
  for (int e = 0; e < nref; e++)
    if (nbr[e] < 0 || nbr[e] >= ncell) return FVM_EARG;
  try {
    std::vector<int> hot;
    for (int i = 0; i < ncell; i++) {
      if (!active[i]) continue;
      bool h = false;
      for (int e = cell_off[i]; e < cell_off[i + 1]; e++) {
        const double di = dti[i], dj = dti[nbr[e]];
        h = h || (2.0 * di < dj) || (2.0 * dj < di);
      }
      if (h) hot.push_back(i);
    }
    // active cells visited in ascending index: rank = index
    std::vector<int> rank(ncell), order(ncell);
    for (int i = 0; i < ncell; i++) {
      rank[i] = active[i] ? i : -1;
      order[i] = i;
    }
    std::vector<unsigned char> changed(ncell, 0);
    std::vector<unsigned long long> pending(((size_t)ncell + 63) / 64, 0ull);
    std::vector<int> chg;
    sparse_sweep(cell_off, nbr, rank.data(), order.data(), hot.data(), (int)hot.size(), dti, pending.data(),
                 pending.size(), changed.data(), chg);
  } catch (...) {
    return FVM_EARG;
  }
  return FVM_OK;
}

}  // extern "C"
