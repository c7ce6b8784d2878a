// Learned HLLD surrogate (no counterpart in the reference; it stands in for
// riemannSolver_HLLD, computeFlux.cpp:178-425, inside computeFlux,
// computeFlux.cpp:6-148).
//
//   featurisation   the HLLD inputs of a face in the face frame (rotation and
//                   Dedner B_n/psi exactly as computeFlux, FP64), divided by two
//                   exact symmetries of the HLLD flux: the Galilean boost
//                   V = w_n n + u_t and the MHD scaling (rho_s, c_s); the
//                   network answers the residual over the HLL flux of the two
//                   sides per unit normalised jump (fvm_mlp.h)
//   training        Adam over bf16-operand MFMA GEMMs (v_mfma_f32_16x16x32_bf16,
//                   FP32 accumulation), one kernel per GEMM / elementwise step
//   inference       one fused kernel: each wavefront featurises 64 faces, runs
//                   the four layers as chained MFMAs with every activation kept
//                   in registers (the K order of each weight fragment is the
//                   order in which the previous layer's accumulator tiles land
//                   in the lanes, so no transpose is needed between layers),
//                   then gates and assembles the FP64 flux; weights stay in LDS
//   fallback        every face the gate rejects is appended to a list (wave
//                   ballot + one atomic per wave) and a second launch runs the
//                   FP64 computeFlux on it
//
// Compiled with -ffp-contract=off so the FP64 Dedner terms and the fallback
// are operation-for-operation the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "../../include/fvm_surrogate.h"
#include "fvm_internal.h"
#include "fvm_physics.h"
#include "fvm_mlp.h"

namespace fvm {
namespace sg {

using Frame = FrameT<double>;

__device__ __forceinline__ int gid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

// ---------------------------------------------------------------------------
// face frame (computeFlux.cpp:17-107), interior faces
// ---------------------------------------------------------------------------
__device__ __forceinline__ void frame_rows(const double *nrm, double A[3][3]) {
  const double ex = nrm[0], ey = nrm[1], ez = nrm[2];
  const double ds = dsqrt(ex * ex + ey * ey);
  const double ids = (ds != 0.0) ? 1.0 / ds : 0.0;
  const double cosph = (ids == 0.0) ? 1.0 : ex * ids;
  const double sinph = (ids == 0.0) ? 0.0 : ey * ids;
  const double costh = ez, sinth = ds;
  A[0][0] = cosph * sinth;
  A[0][1] = sinth * sinph;
  A[0][2] = costh;
  A[1][0] = -sinph;
  A[1][1] = cosph;
  A[1][2] = 0.0;
  A[2][0] = -costh * cosph;
  A[2][1] = -costh * sinph;
  A[2][2] = sinth;
}

__device__ __forceinline__ void to_frame(double gamma, const double *wij, const double *Wi, const double *Wj,
                                         const double *nrm, Frame &s) {
  frame_rows(nrm, s.A);
  for (int d = 0; d < 3; d++) {
    s.vL[d] = s.A[d][0] * Wi[VELX] + s.A[d][1] * Wi[VELY] + s.A[d][2] * Wi[VELZ];
    s.BL[d] = s.A[d][0] * Wi[BX] + s.A[d][1] * Wi[BY] + s.A[d][2] * Wi[BZ];
    s.vR[d] = s.A[d][0] * Wj[VELX] + s.A[d][1] * Wj[VELY] + s.A[d][2] * Wj[VELZ];
    s.BR[d] = s.A[d][0] * Wj[BX] + s.A[d][1] * Wj[BY] + s.A[d][2] * Wj[BZ];
  }
  s.rL = Wi[DENS];
  s.eL = Wi[ETHM];
  s.rR = Wj[DENS];
  s.eR = Wj[ETHM];
  s.psiL = Wi[PSI];
  s.psiR = Wj[PSI];
  s.entrL = Wi[ENTR];
  s.entrR = Wj[ENTR];
  s.pL = pressure(gamma, Wi[ETHM]);
  s.pR = pressure(gamma, Wj[ETHM]);
  s.wn = s.A[0][0] * wij[0] + s.A[0][1] * wij[1] + s.A[0][2] * wij[2];
  const double cfl2 = (gamma * s.pL + sqr(s.BL[0]) + sqr(s.BL[1]) + sqr(s.BL[2])) / s.rL;
  const double cfr2 = (gamma * s.pR + sqr(s.BR[0]) + sqr(s.BR[1]) + sqr(s.BR[2])) / s.rR;
  const double chL = dsqrt(cfl2 + 0.0 * sqr(s.vL[0] - s.wn));
  const double chR = dsqrt(cfr2 + 0.0 * sqr(s.vR[0] - s.wn));
  s.Bn = (s.BL[0] * chL + s.BR[0] * chR + (s.psiL - s.psiR)) / (chL + chR);
  s.psi = (s.psiL * chR + s.psiR * chL + chR * chL * (s.BL[0] - s.BR[0])) / (chL + chR);
}

__device__ __forceinline__ void load_face(int f, const double *wij, const double *Wi, const double *Wj,
                                          const double *nrm, double *w3, double *L, double *R, double *n3) {
  for (int d = 0; d < 3; d++) {
    w3[d] = wij[3 * (size_t)f + d];
    n3[d] = nrm[3 * (size_t)f + d];
  }
  for (int k = 0; k < NF; k++) {
    L[k] = Wi[(size_t)NF * f + k];
    R[k] = Wj[(size_t)NF * f + k];
  }
}

// ---------------------------------------------------------------------------
// training samples: features + normalised moving-frame HLLD fluxes (FP64 HLLD)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_samples(int n, double gamma, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                          const double *nrm, Gate G, float *Xs, float *Ts, int *flag) {
  const int f = gid();
  bool keep = false;
  float x[NIN], t[NOUT];
  if (f < n) {
    double w3[3], L[NF], R[NF], n3[3];
    load_face(f, wij, Wi, Wj, nrm, w3, L, R, n3);
    Frame s;
    to_frame(gamma, w3, L, R, n3, s);
    double Fc[7], J, sc[4];
    keep = features<double>(gamma, s, jbnd[f], G, x, Fc, J, sc);
    if (keep) {
      const double rs = sc[0], cs = sc[1], Ps = rs * cs * cs, Bs = dsqrt(Ps);
      double F[NF];
      riemann_hlld<double>(gamma, F, s.Bn, s.wn, s.rL, s.pL, s.eL, s.vL[0], s.vL[1], s.vL[2], s.BL[1], s.BL[2], s.rR,
                           s.pR, s.eR, s.vR[0], s.vR[1], s.vR[2], s.BR[1], s.BR[2]);
      // moving-frame flux: G' = G - boost(V), V = (w_n, u_t1, u_t2)
      const double V[3] = {s.wn, sc[2], sc[3]};
      const double Gm = F[MASS];
      const double Gp[3] = {F[MOMX] - V[0] * Gm, F[MOMY] - V[1] * Gm, F[MOMZ] - V[2] * Gm};
      const double GE = F[ENER] - (V[0] * Gp[0] + V[1] * Gp[1] + V[2] * Gp[2]) -
                        0.5 * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) * Gm;
      const double Gb1 = F[WBY] + s.Bn * V[1], Gb2 = F[WBZ] + s.Bn * V[2];
      // residual over the HLL base flux, per unit jump
      const double g[7] = {Gm / (rs * cs), Gp[0] / Ps, Gp[1] / Ps, Gp[2] / Ps, GE / (Ps * cs), Gb1 / (Bs * cs),
                           Gb2 / (Bs * cs)};
      for (int k = 0; k < 7; k++) t[k] = (float)((g[k] - Fc[k]) / J);
      t[7] = 0.0f;
      for (int k = 0; k < 7; k++) keep = keep && __builtin_isfinite(t[k]);
    }
  }
  // staged at the face's own index; k_pack_samples keeps the face order, so
  // the training set (and the minibatches drawn from it) is the same every run
  if (f < n) {
    flag[f] = keep ? 1 : 0;
    if (keep) {
      for (int k = 0; k < NIN; k++) Xs[(size_t)f * NIN + k] = x[k];
      for (int k = 0; k < NOUT; k++) Ts[(size_t)f * NOUT + k] = t[k];
    }
  }
}

// kept samples, in face order, appended at base (pos: exclusive scan of flag)
__global__ void k_pack_samples(int n, const int *flag, const int *pos, const float *Xs, const float *Ts, float *X,
                               float *T, long long base) {
  const int f = gid();
  if (f >= n || !flag[f]) return;
  const size_t i = (size_t)(base + pos[f]);
  for (int k = 0; k < NIN; k++) X[i * NIN + k] = Xs[(size_t)f * NIN + k];
  for (int k = 0; k < NOUT; k++) T[i * NOUT + k] = Ts[(size_t)f * NOUT + k];
}

// ---------------------------------------------------------------------------
// training GEMM: C[M][N] (+)= op(A)[M][K] op(B)[K][N], op(A)[m][k] = A[m*sam + k*sak],
// op(B)[k][n] = B[k*sbk + n*sbn]; operands rounded to bf16, FP32 accumulation.
// 64x64 block tile, 4 wavefronts of 32x32, K step 32.
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 32, GKP = GK + 8;  // LDS rows padded to 80 bytes

__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const float *__restrict__ A, long long sam,
                                              long long sak, const float *__restrict__ B, long long sbk,
                                              long long sbn, float *__restrict__ C, int ldc, int accumulate) {
  __shared__ __attribute__((aligned(16))) __bf16 As[GT][GKP];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[GT][GKP];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int g = lane >> 4, r16 = lane & 15;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int q = 0; q < (GT * GK) / 256; q++) {
      const int idx = t + 256 * q;
      int m, k;
      if (sak == 1) { m = idx / GK; k = idx % GK; }
      else { m = idx % GT; k = idx / GT; }
      const int gm = m0 + m, gk = k0 + k;
      const float v = (gm < M && gk < K) ? A[(long long)gm * sam + (long long)gk * sak] : 0.0f;
      As[m][k] = (__bf16)v;
      int n;
      if (sbn == 1) { n = idx % GT; k = idx / GT; }
      else { n = idx / GK; k = idx % GK; }
      const int gn = n0 + n, gk2 = k0 + k;
      const float u = (gn < N && gk2 < K) ? B[(long long)gk2 * sbk + (long long)gn * sbn] : 0.0f;
      Bs[n][k] = (__bf16)u;
    }
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * i + r16][8 * g]);
    for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * j + r16][8 * g]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D layout: lane holds column r16 and rows 4g..4g+3 of each 16x16 tile
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      const int col = n0 + wn + 16 * j + r16;
      for (int r = 0; r < 4; r++) {
        const int row = m0 + wm + 16 * i + 4 * g + r;
        if (row < M && col < N) {
          float *p = C + (long long)row * ldc + col;
          *p = accumulate ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
    }
}

// H[r][c] = act(H[r][c] + b[c])
__global__ void k_bias_act(float *H, const float *b, int rows, int cols, int relu) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  const float v = H[i] + b[i % cols];
  H[i] = relu ? fmaxf(v, 0.0f) : v;
}
// D *= (H > 0)
__global__ void k_relu_mask(float *D, const float *H, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(H[i] > 0.0f)) D[i] = 0.0f;
}
// out[c] = sum_r D[r][c]; 64 columns per block, fixed summation order
__global__ void __launch_bounds__(256) k_colsum(const float *D, int rows, int cols, float *out) {
  __shared__ float part[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + tx;
  float s = 0.0f;
  if (c < cols)
    for (int r = ty; r < rows; r += 4) s += D[(long long)r * cols + c];
  part[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && c < cols) out[c] = (part[0][tx] + part[1][tx]) + (part[2][tx] + part[3][tx]);
}
// loss terms and dL/dy for a batch (outputs 0..6 fluxes weighted by w, 7 confidence)
struct LossW {
  float w[7];
};
__global__ void k_loss(const float *Y, const float *T, int B, float lam, LossW lw, float *dY, float *loss_acc) {
  const int i = gid();
  float l = 0.0f;
  if (i < B) {
    float e2 = 0.0f, d[7];
    for (int k = 0; k < 7; k++) {
      d[k] = Y[(size_t)i * NOUT + k] - T[(size_t)i * NOUT + k];
      e2 += lw.w[k] * d[k] * d[k];
    }
    const float c = log10f(e2 + 1e-10f);
    const float dc = Y[(size_t)i * NOUT + 7] - c;
    const float iB = 1.0f / (float)B;
    for (int k = 0; k < 7; k++) dY[(size_t)i * NOUT + k] = 2.0f * lw.w[k] * d[k] * iB;
    dY[(size_t)i * NOUT + 7] = 2.0f * lam * dc * iB;
    l = (e2 + lam * dc * dc) * iB;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (__lane_id() == 0) atomicAdd(loss_acc, l);
}
__global__ void k_adam(float *P, const float *G, float *Mm, float *Vv, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = gid();
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * Mm[i] + (1.0f - b1) * g;
  const float v = b2 * Vv[i] + (1.0f - b2) * g * g;
  Mm[i] = m;
  Vv[i] = v;
  P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}
__device__ __forceinline__ unsigned long long splitmix(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ void k_gather(const float *X, const float *T, long long nsamp, int B, unsigned long long seed,
                         unsigned long long step, float *Xb, float *Tb) {
  const int i = gid();
  if (i >= B) return;
  const long long s = (long long)(splitmix(seed ^ splitmix((step << 32) ^ (unsigned long long)i)) % (unsigned long long)nsamp);
  for (int k = 0; k < NIN; k++) Xb[(size_t)i * NIN + k] = X[(size_t)s * NIN + k];
  for (int k = 0; k < NOUT; k++) Tb[(size_t)i * NOUT + k] = T[(size_t)s * NOUT + k];
}

// ---------------------------------------------------------------------------
// inference: weight fragments in the lane order of the MFMA A operand
//   L0 (16x16x16): frag[mt][lane] 4 x bf16 = W0[4g+e][16mt+r16]
//   L1, L2 (16x16x32): frag[mt][c][lane] 8 x bf16 = W[f(c,g,e)][16mt+r16]
//   L3: frag[c][lane] 8 x bf16 = W3[f(c,g,e)][r16] (r16 < 8, else 0)
// with f(c,g,e) = 32c + 4g + e (e < 4), 32c + 16 + 4g + e - 4 (e >= 4): the
// input feature that lane group g holds in K slot e after the previous layer's
// accumulator tiles 2c and 2c+1 are concatenated.
// ---------------------------------------------------------------------------
constexpr int IB = 512;                        // threads per inference block
constexpr int FEAT_B = 64 * NIN * 2;           // per wave: features (bf16), reused for outputs (f32)
constexpr int OUT_B = 64 * NOUT * 4;
struct FaceCtx {
  float nrm[3];
  float wn, ut1, ut2, rs, cs, psiL, psiR, entrL, entrR;
  float Fc[7], J;
  float pad[2];
  double Bn, psi;
};
constexpr int CTX_B = 64 * (int)sizeof(FaceCtx);
constexpr int WAVE_B = (FEAT_B > OUT_B ? FEAT_B : OUT_B) + CTX_B;
constexpr int LDS_B = P_BYTES + (IB / 64) * WAVE_B;
static_assert(sizeof(FaceCtx) == 104, "FaceCtx layout");
static_assert(LDS_B <= 160 * 1024, "LDS budget");

__global__ void k_pack(const float *P, __bf16 *frag, float *bias) {
  const int i = gid();
  if (i < P_BF) {
    float v = 0.0f;
    if (i < P_L0) {
      const int e = i & 3, lane = (i >> 2) & 63, mt = i >> 8;
      v = P[OW0 + (4 * (lane >> 4) + e) * NH + 16 * mt + (lane & 15)];
    } else if (i < P_L0 + 2 * P_L1) {
      const int j = (i - P_L0) % P_L1, layer = (i - P_L0) / P_L1;
      const int e = j & 7, lane = (j >> 3) & 63, c = (j >> 9) & 3, mt = j >> 11;
      v = P[(layer ? OW2 : OW1) + kfeat(c, lane >> 4, e) * NH + 16 * mt + (lane & 15)];
    } else {
      const int j = i - P_L0 - 2 * P_L1;
      const int e = j & 7, lane = (j >> 3) & 63, c = j >> 9;
      v = (lane & 15) < NOUT ? P[OW3 + kfeat(c, lane >> 4, e) * NOUT + (lane & 15)] : 0.0f;
    }
    frag[i] = (__bf16)v;
  }
  if (i < P_BIAS) {
    float v = 0.0f;
    if (i < NH) v = P[OB0 + i];
    else if (i < 2 * NH) v = P[OB1 + i - NH];
    else if (i < 3 * NH) v = P[OB2 + i - 2 * NH];
    else if (i < 3 * NH + NOUT) v = P[OB3 + i - 3 * NH];
    bias[i] = v;
  }
}

// MODE 0: computeFlux batch (features from the faces, gate, FP64 assembly,
// fallback list); MODE 1: raw features x[n][16] -> y[n][8]
template <int MODE>
__global__ void __launch_bounds__(IB) k_infer(int n, double gamma, const double *__restrict__ wij,
                                              const double *__restrict__ Wi, const double *__restrict__ Wj,
                                              const int *__restrict__ jbnd, const double *__restrict__ nrm,
                                              const __bf16 *__restrict__ pfrag, const float *__restrict__ pbias,
                                              Gate G, double *__restrict__ psi_out, double *__restrict__ Bn_out,
                                              double *__restrict__ flux, int *__restrict__ fb_list,
                                              unsigned long long *__restrict__ fb_count, const float *__restrict__ xin,
                                              float *__restrict__ yout) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __bf16 *wl = reinterpret_cast<__bf16 *>(smem);
  float *bias = reinterpret_cast<float *>(smem + P_BF * 2);
  {
    const int4 *src = reinterpret_cast<const int4 *>(pfrag);
    int4 *dst = reinterpret_cast<int4 *>(smem);
    for (int i = threadIdx.x; i < P_BF * 2 / 16; i += IB) dst[i] = src[i];
    for (int i = threadIdx.x; i < P_BIAS; i += IB) bias[i] = pbias[i];
  }
  __syncthreads();
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  char *wb = smem + P_BYTES + wv * WAVE_B;
  __bf16 *feat = reinterpret_cast<__bf16 *>(wb);
  float *out = reinterpret_cast<float *>(wb);
  FaceCtx *ctx = reinterpret_cast<FaceCtx *>(wb + (FEAT_B > OUT_B ? FEAT_B : OUT_B));
  const int nbatch = (n + 63) / 64;
  const int wstride = gridDim.x * (IB / 64);
  for (int bw = blockIdx.x * (IB / 64) + wv; bw < nbatch; bw += wstride) {
    const int f = bw * 64 + lane;
    const bool live = f < n;
    bool ok = false;
    float x[NIN];
    if constexpr (MODE == 1) {
      for (int k = 0; k < NIN; k++) x[k] = live ? xin[(size_t)f * NIN + k] : 0.0f;
      ok = live;
    } else {
      for (int k = 0; k < NIN; k++) x[k] = 0.0f;
      if (live) {
        double w3[3], L[NF], R[NF], n3[3];
        load_face(f, wij, Wi, Wj, nrm, w3, L, R, n3);
        Frame s;
        to_frame(gamma, w3, L, R, n3, s);
        float Fc[7], J, sc[4];
        ok = features<float>(gamma, s, jbnd[f], G, x, Fc, J, sc);
        FaceCtx &cx = ctx[lane];
        for (int k = 0; k < 7; k++) cx.Fc[k] = Fc[k];
        cx.J = J;
        for (int d = 0; d < 3; d++) cx.nrm[d] = (float)n3[d];
        cx.wn = (float)s.wn;
        cx.ut1 = sc[2];
        cx.ut2 = sc[3];
        cx.rs = sc[0];
        cx.cs = sc[1];
        cx.psiL = (float)s.psiL;
        cx.psiR = (float)s.psiR;
        cx.entrL = (float)s.entrL;
        cx.entrR = (float)s.entrR;
        cx.Bn = s.Bn;
        cx.psi = s.psi;
        if (!ok)
          for (int k = 0; k < NIN; k++) x[k] = 0.0f;
      }
    }
    {
      bf16x8 lo, hi;
      for (int k = 0; k < 8; k++) {
        lo[k] = (__bf16)x[k];
        hi[k] = (__bf16)x[8 + k];
      }
      *reinterpret_cast<bf16x8 *>(feat + lane * NIN) = lo;
      *reinterpret_cast<bf16x8 *>(feat + lane * NIN + 8) = hi;
    }
    wave_sync();
    mlp64(wl, bias, feat, out, lane);
    wave_sync();
    float y[NOUT];
    for (int k = 0; k < NOUT; k++) y[k] = out[lane * NOUT + k];
    wave_sync();  // outputs read before the next batch's features land
    if constexpr (MODE == 1) {
      if (live)
        for (int k = 0; k < NOUT; k++) yout[(size_t)f * NOUT + k] = y[k];
    } else {
      ok = ok && y[7] <= G.max_log10_err;
      if (live && ok) {
        const FaceCtx &cx = ctx[lane];
        const float rs = cx.rs, cs = cx.cs, Ps = rs * cs * cs, Bs = __builtin_sqrtf(Ps);
        const float V[3] = {cx.wn, cx.ut1, cx.ut2};
        float g[7];
        for (int k = 0; k < 7; k++) g[k] = cx.Fc[k] + cx.J * y[k];
        const float Gm = g[0] * rs * cs;
        const float Gp[3] = {g[1] * Ps, g[2] * Ps, g[3] * Ps};
        const float GE = g[4] * Ps * cs + (V[0] * Gp[0] + V[1] * Gp[1] + V[2] * Gp[2]) +
                         0.5f * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) * Gm;
        const float Bn = (float)cx.Bn;
        const float Gmom[3] = {Gp[0] + V[0] * Gm, Gp[1] + V[1] * Gm, Gp[2] + V[2] * Gm};
        const float GB[3] = {-cx.wn * Bn, g[5] * Bs * cs - Bn * V[1], g[6] * Bs * cs - Bn * V[2]};
        double A[3][3];
        const double n3[3] = {(double)cx.nrm[0], (double)cx.nrm[1], (double)cx.nrm[2]};
        frame_rows(n3, A);
        double *o = flux + (size_t)NF * f;
        o[MASS] = (double)Gm;
        o[ENER] = (double)GE;
        for (int d = 0; d < 3; d++) {
          o[MOMX + d] = A[0][d] * Gmom[0] + A[1][d] * Gmom[1] + A[2][d] * Gmom[2];
          o[WBX + d] = A[0][d] * GB[0] + A[1][d] * GB[1] + A[2][d] * GB[2];
        }
        o[MPSI] = (double)Gm * (double)(Gm > 0.0f ? cx.psiL : cx.psiR);
        o[MENTR] = (double)Gm * (double)(Gm > 0.0f ? cx.entrL : cx.entrR);
        psi_out[f] = cx.psi;
        Bn_out[f] = cx.Bn;
      }
      const bool rej = live && !ok;
      const unsigned long long m = __ballot(rej);
      unsigned long long b0 = 0;
      if (lane == 0 && m) b0 = atomicAdd(fb_count, (unsigned long long)__popcll(m));
      b0 = __shfl(b0, 0);
      if (rej) fb_list[b0 + (unsigned long long)__popcll(m & ((1ULL << lane) - 1ULL))] = f;
    }
  }
}

// FP64 computeFlux on the rejected faces
__global__ void __launch_bounds__(256) k_fallback(const int *list, const unsigned long long *count, double gamma, const double *wij,
                           const double *Wi, const double *Wj, const int *jbnd, const double *nrm, double *psi,
                           double *Bn, double *flux) {
  const int i = gid();
  if ((unsigned long long)i >= *count) return;
  const int f = list[i];
  double w3[3], L[NF], R[NF], n3[3], F[NF], p, b;
  load_face(f, wij, Wi, Wj, nrm, w3, L, R, n3);
  compute_flux<double>(gamma, w3, L, R, jbnd[f], n3, p, b, F);
  psi[f] = p;
  Bn[f] = b;
  for (int k = 0; k < NF; k++) flux[(size_t)NF * f + k] = F[k];
}

// computeFlux on every face in precision T (the comparison kernels)
template <typename T>
__global__ void __launch_bounds__(256) k_hlld(int n, double gamma, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                       const double *nrm, double *psi, double *Bn, double *flux) {
  const int f = gid();
  if (f >= n) return;
  double w3[3], L[NF], R[NF], n3[3];
  load_face(f, wij, Wi, Wj, nrm, w3, L, R, n3);
  T wt[3], Lt[NF], Rt[NF], nt[3], F[NF], p, b;
  for (int d = 0; d < 3; d++) {
    wt[d] = (T)w3[d];
    nt[d] = (T)n3[d];
  }
  for (int k = 0; k < NF; k++) {
    Lt[k] = (T)L[k];
    Rt[k] = (T)R[k];
  }
  compute_flux<T>((T)gamma, wt, Lt, Rt, jbnd[f], nt, p, b, F);
  psi[f] = (double)p;
  Bn[f] = (double)b;
  for (int k = 0; k < NF; k++) flux[(size_t)NF * f + k] = (double)F[k];
}

}  // namespace sg
}  // namespace fvm

// ===========================================================================
// host side
// ===========================================================================
namespace {
struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
};
}  // namespace

struct fvm_sg {
  int device = 0;
  hipStream_t stream = nullptr;
  double gamma = 5.0 / 3.0;
  std::string err;
  fvm::sg::Gate gate{0.5f, 1.0f, 3.0f, -3.0f};
  fvm::sg::LossW lossw{{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f}};
  Buf params, grads, am, av, pfrag, pbias;
  Buf X, T;
  long long nsamp = 0, cap = 0;
  Buf xb, tb, h[3], y, dy, d[2], loss;
  int bcap = 0;
  Buf stage, fbl, fbc;
  Buf sx, st, sflag, spos;  // sample staging of one add_faces call
  fvm::ScanScratch scan;
  int adam_t = 0;
  bool packed = false;
};

using namespace fvm::sg;

static int sg_fail(fvm_sg *s, int code, const std::string &m) {
  if (s) s->err = m;
  return code;
}
#define SGCHK(s, call)                                                             \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return sg_fail(s, FVM_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

static void sg_release(Buf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}
static int sg_ensure(fvm_sg *s, Buf &b, size_t bytes, bool keep = false) {
  if (b.bytes >= bytes && b.p) return FVM_OK;
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return sg_fail(s, FVM_EHIP, "hipMalloc failed");
  if (keep && b.p && b.bytes) (void)hipMemcpy(p, b.p, b.bytes, hipMemcpyDeviceToDevice);
  sg_release(b);
  b.p = p;
  b.bytes = std::max<size_t>(bytes, 256);
  return FVM_OK;
}
static unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static int gemm(fvm_sg *s, int M, int N, int K, const float *A, long long sam, long long sak, const float *B,
                long long sbk, long long sbn, float *C, int ldc, int accumulate) {
  dim3 grid((unsigned)((N + GT - 1) / GT), (unsigned)((M + GT - 1) / GT));
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s->stream, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate);
  SGCHK(s, hipGetLastError());
  return FVM_OK;
}

static int ensure_batch(fvm_sg *s, int B) {
  if (B <= s->bcap) return FVM_OK;
  int rc;
  const size_t b = (size_t)B;
  if ((rc = sg_ensure(s, s->xb, 4 * b * NIN)) || (rc = sg_ensure(s, s->tb, 4 * b * NOUT)) ||
      (rc = sg_ensure(s, s->h[0], 4 * b * NH)) || (rc = sg_ensure(s, s->h[1], 4 * b * NH)) ||
      (rc = sg_ensure(s, s->h[2], 4 * b * NH)) || (rc = sg_ensure(s, s->y, 4 * b * NOUT)) ||
      (rc = sg_ensure(s, s->dy, 4 * b * NOUT)) || (rc = sg_ensure(s, s->d[0], 4 * b * NH)) ||
      (rc = sg_ensure(s, s->d[1], 4 * b * NH)) || (rc = sg_ensure(s, s->loss, 4 * 1024)))
    return rc;
  s->bcap = B;
  return FVM_OK;
}

// forward + loss + backward on the batch in xb/tb; loss accumulated into loss[slot]
static int fwd_bwd(fvm_sg *s, int B, float lam, int slot) {
  const float *P = (const float *)s->params.p;
  float *G = (float *)s->grads.p;
  const float *X = (const float *)s->xb.p;
  float *H0 = (float *)s->h[0].p, *H1 = (float *)s->h[1].p, *H2 = (float *)s->h[2].p;
  float *Y = (float *)s->y.p, *dY = (float *)s->dy.p, *D0 = (float *)s->d[0].p, *D1 = (float *)s->d[1].p;
  int rc;
  const long long bh = (long long)B * NH;
  // forward
  if ((rc = gemm(s, B, NH, NIN, X, NIN, 1, P + OW0, NH, 1, H0, NH, 0))) return rc;
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(bh)), dim3(256), 0, s->stream, H0, P + OB0, B, NH, 1);
  if ((rc = gemm(s, B, NH, NH, H0, NH, 1, P + OW1, NH, 1, H1, NH, 0))) return rc;
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(bh)), dim3(256), 0, s->stream, H1, P + OB1, B, NH, 1);
  if ((rc = gemm(s, B, NH, NH, H1, NH, 1, P + OW2, NH, 1, H2, NH, 0))) return rc;
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(bh)), dim3(256), 0, s->stream, H2, P + OB2, B, NH, 1);
  if ((rc = gemm(s, B, NOUT, NH, H2, NH, 1, P + OW3, NOUT, 1, Y, NOUT, 0))) return rc;
  hipLaunchKernelGGL(k_bias_act, dim3(nblk((long long)B * NOUT)), dim3(256), 0, s->stream, Y, P + OB3, B, NOUT, 0);
  hipLaunchKernelGGL(k_loss, dim3(nblk(B)), dim3(256), 0, s->stream, Y, (const float *)s->tb.p, B, lam, s->lossw, dY,
                     (float *)s->loss.p + slot);
  // backward
  if ((rc = gemm(s, NH, NOUT, B, H2, 1, NH, dY, NOUT, 1, G + OW3, NOUT, 0))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(NOUT, 64)), dim3(256), 0, s->stream, dY, B, NOUT, G + OB3);
  if ((rc = gemm(s, B, NH, NOUT, dY, NOUT, 1, P + OW3, 1, NOUT, D0, NH, 0))) return rc;
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(bh)), dim3(256), 0, s->stream, D0, H2, bh);
  if ((rc = gemm(s, NH, NH, B, H1, 1, NH, D0, NH, 1, G + OW2, NH, 0))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(NH, 64)), dim3(256), 0, s->stream, D0, B, NH, G + OB2);
  if ((rc = gemm(s, B, NH, NH, D0, NH, 1, P + OW2, 1, NH, D1, NH, 0))) return rc;
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(bh)), dim3(256), 0, s->stream, D1, H1, bh);
  if ((rc = gemm(s, NH, NH, B, H0, 1, NH, D1, NH, 1, G + OW1, NH, 0))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(NH, 64)), dim3(256), 0, s->stream, D1, B, NH, G + OB1);
  if ((rc = gemm(s, B, NH, NH, D1, NH, 1, P + OW1, 1, NH, D0, NH, 0))) return rc;
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(bh)), dim3(256), 0, s->stream, D0, H0, bh);
  if ((rc = gemm(s, NIN, NH, B, X, 1, NIN, D0, NH, 1, G + OW0, NH, 0))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(NH, 64)), dim3(256), 0, s->stream, D0, B, NH, G + OB0);
  SGCHK(s, hipGetLastError());
  return FVM_OK;
}

static int pack(fvm_sg *s) {
  if (s->packed) return FVM_OK;
  int rc;
  if ((rc = sg_ensure(s, s->pfrag, (size_t)P_BF * 2)) || (rc = sg_ensure(s, s->pbias, (size_t)P_BIAS * 4))) return rc;
  hipLaunchKernelGGL(k_pack, dim3(nblk(P_BF)), dim3(256), 0, s->stream, (const float *)s->params.p,
                     (__bf16 *)s->pfrag.p, (float *)s->pbias.p);
  SGCHK(s, hipGetLastError());
  s->packed = true;
  return FVM_OK;
}

static unsigned infer_grid(int n) {
  const unsigned nb = (unsigned)(((long long)n + IB - 1) / IB);
  return std::max(1u, std::min(nb, 256u * 4u));
}

template <int MODE>
static int launch_infer(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                        const double *nrm, double *psi, double *Bn, double *flux, int *fbl,
                        unsigned long long *fbc, const float *xin, float *yout) {
  static bool attr = false;
  if (!attr) {
    SGCHK(s, hipFuncSetAttribute((const void *)k_infer<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_B));
    attr = true;
  }
  hipLaunchKernelGGL(k_infer<MODE>, dim3(infer_grid(n)), dim3(IB), LDS_B, s->stream, n, s->gamma, wij, Wi, Wj, jbnd,
                     nrm, (const __bf16 *)s->pfrag.p, (const float *)s->pbias.p, s->gate, psi, Bn, flux, fbl, fbc,
                     xin, yout);
  SGCHK(s, hipGetLastError());
  return FVM_OK;
}

int fvm_sg_device_weights(fvm_sg *s, const void **frag, const float **bias, float gate[4]) {
  if (!s || !frag || !bias || !gate) return FVM_EARG;
  (void)hipSetDevice(s->device);
  int rc = pack(s);
  if (rc) return rc;
  SGCHK(s, hipStreamSynchronize(s->stream));
  *frag = s->pfrag.p;
  *bias = (const float *)s->pbias.p;
  gate[0] = s->gate.max_log_rho;
  gate[1] = s->gate.max_log_p;
  gate[2] = s->gate.max_vn;
  gate[3] = s->gate.max_log10_err;
  return FVM_OK;
}

int fvm_sg_device_of(const fvm_sg *s) { return s ? s->device : -1; }

extern "C" {

int fvm_sg_create(fvm_sg **out, int device, double gamma_gas) {
  if (!out) return FVM_EARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FVM_EHIP;
  fvm_sg *s = new (std::nothrow) fvm_sg();
  if (!s) return FVM_EARG;
  s->device = device;
  s->gamma = gamma_gas;
  (void)hipSetDevice(device);
  if (hipStreamCreate(&s->stream) != hipSuccess) {
    delete s;
    return FVM_EHIP;
  }
  if (sg_ensure(s, s->params, 4 * NPARAM) || sg_ensure(s, s->grads, 4 * NPARAM) || sg_ensure(s, s->am, 4 * NPARAM) ||
      sg_ensure(s, s->av, 4 * NPARAM)) {
    fvm_sg_destroy(s);
    return FVM_EHIP;
  }
  (void)hipMemset(s->params.p, 0, 4 * NPARAM);
  (void)hipMemset(s->am.p, 0, 4 * NPARAM);
  (void)hipMemset(s->av.p, 0, 4 * NPARAM);
  *out = s;
  return FVM_OK;
}

void fvm_sg_destroy(fvm_sg *s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  for (Buf *b : {&s->params, &s->grads, &s->am, &s->av, &s->pfrag, &s->pbias, &s->X, &s->T, &s->xb, &s->tb, &s->h[0],
                 &s->h[1], &s->h[2], &s->y, &s->dy, &s->d[0], &s->d[1], &s->loss, &s->stage, &s->fbl, &s->fbc,
                 &s->sx, &s->st, &s->sflag, &s->spos})
    sg_release(*b);
  fvm::scan_release(s->scan);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char *fvm_sg_last_error(const fvm_sg *s) { return s ? s->err.c_str() : "null handle"; }

int fvm_sg_set_params(fvm_sg *s, const float *params, int n) {
  if (!s || !params || n != NPARAM) return sg_fail(s, FVM_EARG, "set_params: need FVM_SG_NPARAM floats");
  (void)hipSetDevice(s->device);
  SGCHK(s, hipMemcpyAsync(s->params.p, params, 4 * (size_t)NPARAM, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemsetAsync(s->am.p, 0, 4 * NPARAM, s->stream));
  SGCHK(s, hipMemsetAsync(s->av.p, 0, 4 * NPARAM, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  s->adam_t = 0;
  s->packed = false;
  return FVM_OK;
}

int fvm_sg_get_params(fvm_sg *s, float *params, int n) {
  if (!s || !params || n != NPARAM) return sg_fail(s, FVM_EARG, "get_params: need FVM_SG_NPARAM floats");
  (void)hipSetDevice(s->device);
  SGCHK(s, hipMemcpyAsync(params, s->params.p, 4 * (size_t)NPARAM, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return FVM_OK;
}

int fvm_sg_set_loss_weights(fvm_sg *s, const double *w, int n) {
  if (!s || !w || n != 7) return sg_fail(s, FVM_EARG, "set_loss_weights: need 7 weights");
  for (int k = 0; k < 7; k++) {
    if (!(w[k] > 0.0)) return sg_fail(s, FVM_EARG, "set_loss_weights: weights must be positive");
    s->lossw.w[k] = (float)w[k];
  }
  return FVM_OK;
}

int fvm_sg_set_gate(fvm_sg *s, const fvm_sg_gate *g) {
  if (!s || !g) return sg_fail(s, FVM_EARG, "set_gate: NULL");
  s->gate = Gate{(float)g->max_log_rho_jump, (float)g->max_log_p_jump, (float)g->max_vn, (float)g->max_log10_err};
  return FVM_OK;
}

static int stage_faces(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                       const double *normal, bool outputs, double **d) {
  const size_t N = (size_t)n;
  const size_t nd = (3 + 10 + 10 + 3 + (outputs ? 12 : 0)) * N;
  int rc = sg_ensure(s, s->stage, sizeof(double) * nd + sizeof(int) * N + 64);
  if (rc) return rc;
  double *b = (double *)s->stage.p;
  d[0] = b;
  d[1] = d[0] + 3 * N;
  d[2] = d[1] + 10 * N;
  d[3] = d[2] + 10 * N;
  d[4] = d[3] + 3 * N;  // outputs: psi, Bn, flux
  d[5] = d[4] + N;
  d[6] = d[5] + N;
  d[7] = (double *)((char *)s->stage.p + sizeof(double) * nd);  // jbnd
  SGCHK(s, hipMemcpyAsync(d[0], wij, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemcpyAsync(d[1], Wi, sizeof(double) * 10 * N, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemcpyAsync(d[2], Wj, sizeof(double) * 10 * N, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemcpyAsync(d[3], normal, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemcpyAsync(d[7], jbnd, sizeof(int) * N, hipMemcpyHostToDevice, s->stream));
  return FVM_OK;
}

int fvm_sg_add_faces(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                     const double *normal, int *n_kept) {
  if (!s || n < 0) return sg_fail(s, FVM_EARG, "add_faces: bad arguments");
  if (n == 0) {
    if (n_kept) *n_kept = 0;
    return FVM_OK;
  }
  if (!wij || !Wi || !Wj || !jbnd || !normal) return sg_fail(s, FVM_EARG, "add_faces: NULL array");
  (void)hipSetDevice(s->device);
  int rc;
  const long long need = s->nsamp + n;
  if (need > s->cap) {
    const long long cap = std::max(need, 2 * s->cap);
    if ((rc = sg_ensure(s, s->X, 4 * (size_t)cap * NIN, true)) || (rc = sg_ensure(s, s->T, 4 * (size_t)cap * NOUT, true)))
      return rc;
    s->cap = cap;
  }
  double *d[8];
  if ((rc = stage_faces(s, n, wij, Wi, Wj, jbnd, normal, false, d))) return rc;
  if ((rc = sg_ensure(s, s->sx, sizeof(float) * NIN * (size_t)n)) ||
      (rc = sg_ensure(s, s->st, sizeof(float) * NOUT * (size_t)n)) || (rc = sg_ensure(s, s->sflag, sizeof(int) * (size_t)n)) ||
      (rc = sg_ensure(s, s->spos, sizeof(int) * (size_t)n)))
    return rc;
  hipLaunchKernelGGL(k_samples, dim3(nblk(n)), dim3(256), 0, s->stream, n, s->gamma, d[0], d[1], d[2],
                     (const int *)d[7], d[3], s->gate, (float *)s->sx.p, (float *)s->st.p, (int *)s->sflag.p);
  SGCHK(s, hipGetLastError());
  int kept = 0;
  if (fvm::exclusive_scan_flags(s->scan, (const int *)s->sflag.p, (int *)s->spos.p, n, &kept, s->stream))
    return sg_fail(s, FVM_EHIP, "add_faces: sample scan failed");
  hipLaunchKernelGGL(k_pack_samples, dim3(nblk(n)), dim3(256), 0, s->stream, n, (const int *)s->sflag.p,
                     (const int *)s->spos.p, (const float *)s->sx.p, (const float *)s->st.p, (float *)s->X.p,
                     (float *)s->T.p, s->nsamp);
  SGCHK(s, hipGetLastError());
  SGCHK(s, hipStreamSynchronize(s->stream));
  s->nsamp += (long long)kept;
  if (n_kept) *n_kept = (int)kept;
  return FVM_OK;
}

int fvm_sg_clear_samples(fvm_sg *s) {
  if (!s) return FVM_EARG;
  s->nsamp = 0;
  return FVM_OK;
}

int fvm_sg_num_samples(fvm_sg *s, long long *n) {
  if (!s || !n) return FVM_EARG;
  *n = s->nsamp;
  return FVM_OK;
}

int fvm_sg_get_samples(fvm_sg *s, long long first, int n, float *x, float *t) {
  if (!s || n < 0 || first < 0 || first + n > s->nsamp || !x || !t) return sg_fail(s, FVM_EARG, "get_samples: range");
  (void)hipSetDevice(s->device);
  SGCHK(s, hipMemcpyAsync(x, (float *)s->X.p + first * NIN, 4 * (size_t)n * NIN, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipMemcpyAsync(t, (float *)s->T.p + first * NOUT, 4 * (size_t)n * NOUT, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return FVM_OK;
}

int fvm_sg_train(fvm_sg *s, int steps, int batch, double lr, double lam, unsigned long long seed, float *loss_hist) {
  if (!s || steps < 0 || batch < 64 || !(lr > 0.0)) return sg_fail(s, FVM_EARG, "train: bad arguments");
  if (s->nsamp < 1) return sg_fail(s, FVM_ESTATE, "train: no samples");
  (void)hipSetDevice(s->device);
  int rc;
  if ((rc = ensure_batch(s, batch)) || (rc = sg_ensure(s, s->loss, 4 * (size_t)std::max(steps, 1)))) return rc;
  SGCHK(s, hipMemsetAsync(s->loss.p, 0, 4 * (size_t)std::max(steps, 1), s->stream));
  const float b1 = 0.9f, b2 = 0.999f;
  for (int k = 0; k < steps; k++) {
    hipLaunchKernelGGL(k_gather, dim3(nblk(batch)), dim3(256), 0, s->stream, (const float *)s->X.p,
                       (const float *)s->T.p, s->nsamp, batch, seed, (unsigned long long)s->adam_t,
                       (float *)s->xb.p, (float *)s->tb.p);
    if ((rc = fwd_bwd(s, batch, (float)lam, k))) return rc;
    s->adam_t++;
    const float c1 = 1.0f - powf(b1, (float)s->adam_t), c2 = 1.0f - powf(b2, (float)s->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM)), dim3(256), 0, s->stream, (float *)s->params.p,
                       (const float *)s->grads.p, (float *)s->am.p, (float *)s->av.p, NPARAM, (float)lr, b1, b2, 1e-8f,
                       c1, c2);
    SGCHK(s, hipGetLastError());
  }
  s->packed = false;
  if (loss_hist && steps > 0)
    SGCHK(s, hipMemcpyAsync(loss_hist, s->loss.p, 4 * (size_t)steps, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return FVM_OK;
}

int fvm_sg_gradients(fvm_sg *s, int n, const float *x, const float *t, double lam, float *y, float *grads,
                     double *loss) {
  if (!s || n < 1 || !x || !t) return sg_fail(s, FVM_EARG, "gradients: bad arguments");
  (void)hipSetDevice(s->device);
  int rc;
  if ((rc = ensure_batch(s, n))) return rc;
  SGCHK(s, hipMemcpyAsync(s->xb.p, x, 4 * (size_t)n * NIN, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemcpyAsync(s->tb.p, t, 4 * (size_t)n * NOUT, hipMemcpyHostToDevice, s->stream));
  SGCHK(s, hipMemsetAsync(s->loss.p, 0, 4, s->stream));
  if ((rc = fwd_bwd(s, n, (float)lam, 0))) return rc;
  float l = 0.0f;
  if (y) SGCHK(s, hipMemcpyAsync(y, s->y.p, 4 * (size_t)n * NOUT, hipMemcpyDeviceToHost, s->stream));
  if (grads) SGCHK(s, hipMemcpyAsync(grads, s->grads.p, 4 * (size_t)NPARAM, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipMemcpyAsync(&l, s->loss.p, 4, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  if (loss) *loss = (double)l;
  return FVM_OK;
}

int fvm_sg_forward(fvm_sg *s, int n, const float *x, float *y) {
  if (!s || n < 0 || (n > 0 && (!x || !y))) return sg_fail(s, FVM_EARG, "forward: bad arguments");
  if (n == 0) return FVM_OK;
  (void)hipSetDevice(s->device);
  int rc;
  if ((rc = pack(s)) || (rc = sg_ensure(s, s->stage, 4 * (size_t)n * (NIN + NOUT)))) return rc;
  float *dx = (float *)s->stage.p, *dy = dx + (size_t)n * NIN;
  SGCHK(s, hipMemcpyAsync(dx, x, 4 * (size_t)n * NIN, hipMemcpyHostToDevice, s->stream));
  if ((rc = launch_infer<1>(s, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, dx, dy)))
    return rc;
  SGCHK(s, hipMemcpyAsync(y, dy, 4 * (size_t)n * NOUT, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return FVM_OK;
}

int fvm_sg_compute_flux_device(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj,
                               const int *jbnd, const double *normal, double *psi, double *Bn, double *flux,
                               int *n_fallback) {
  if (!s || n < 0) return sg_fail(s, FVM_EARG, "compute_flux: bad arguments");
  if (n == 0) {
    if (n_fallback) *n_fallback = 0;
    return FVM_OK;
  }
  if (!wij || !Wi || !Wj || !jbnd || !normal || !psi || !Bn || !flux)
    return sg_fail(s, FVM_EARG, "compute_flux: NULL array");
  (void)hipSetDevice(s->device);
  int rc;
  if ((rc = pack(s)) || (rc = sg_ensure(s, s->fbl, sizeof(int) * (size_t)n)) ||
      (rc = sg_ensure(s, s->fbc, sizeof(unsigned long long))))
    return rc;
  unsigned long long *cnt = (unsigned long long *)s->fbc.p;
  SGCHK(s, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s->stream));
  if ((rc = launch_infer<0>(s, n, wij, Wi, Wj, jbnd, normal, psi, Bn, flux, (int *)s->fbl.p, cnt, nullptr, nullptr)))
    return rc;
  hipLaunchKernelGGL(k_fallback, dim3(nblk(n)), dim3(256), 0, s->stream, (const int *)s->fbl.p, cnt, s->gamma, wij, Wi,
                     Wj, jbnd, normal, psi, Bn, flux);
  SGCHK(s, hipGetLastError());
  unsigned long long nf = 0;
  SGCHK(s, hipMemcpyAsync(&nf, cnt, sizeof nf, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  if (n_fallback) *n_fallback = (int)nf;
  return FVM_OK;
}

int fvm_sg_compute_flux(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                        const double *normal, double *psi, double *Bn, double *flux, int *n_fallback) {
  if (!s || n < 0) return sg_fail(s, FVM_EARG, "compute_flux: bad arguments");
  if (n == 0) {
    if (n_fallback) *n_fallback = 0;
    return FVM_OK;
  }
  if (!wij || !Wi || !Wj || !jbnd || !normal || !psi || !Bn || !flux)
    return sg_fail(s, FVM_EARG, "compute_flux: NULL array");
  (void)hipSetDevice(s->device);
  double *d[8];
  int rc = stage_faces(s, n, wij, Wi, Wj, jbnd, normal, true, d);
  if (rc) return rc;
  if ((rc = fvm_sg_compute_flux_device(s, n, d[0], d[1], d[2], (const int *)d[7], d[3], d[4], d[5], d[6], n_fallback)))
    return rc;
  const size_t N = (size_t)n;
  SGCHK(s, hipMemcpyAsync(psi, d[4], sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipMemcpyAsync(Bn, d[5], sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipMemcpyAsync(flux, d[6], sizeof(double) * 10 * N, hipMemcpyDeviceToHost, s->stream));
  SGCHK(s, hipStreamSynchronize(s->stream));
  return FVM_OK;
}

int fvm_sg_time_flux(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                     const double *normal, int iters, double *ms, int *n_fallback) {
  if (!s || n < 1 || iters < 1 || !ms) return sg_fail(s, FVM_EARG, "time_flux: bad arguments");
  (void)hipSetDevice(s->device);
  double *d[8];
  int rc = stage_faces(s, n, wij, Wi, Wj, jbnd, normal, true, d);
  if (rc) return rc;
  const int *jb = (const int *)d[7];
  if ((rc = fvm_sg_compute_flux_device(s, n, d[0], d[1], d[2], jb, d[3], d[4], d[5], d[6], n_fallback))) return rc;
  hipEvent_t e[4];
  for (auto &x : e) SGCHK(s, hipEventCreate(&x));
  unsigned long long *cnt = (unsigned long long *)s->fbc.p;
  SGCHK(s, hipEventRecord(e[0], s->stream));
  for (int k = 0; k < iters; k++) {
    SGCHK(s, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s->stream));
    if ((rc = launch_infer<0>(s, n, d[0], d[1], d[2], jb, d[3], d[4], d[5], d[6], (int *)s->fbl.p, cnt, nullptr,
                              nullptr)))
      return rc;
    hipLaunchKernelGGL(k_fallback, dim3(nblk(n)), dim3(256), 0, s->stream, (const int *)s->fbl.p, cnt, s->gamma, d[0],
                       d[1], d[2], jb, d[3], d[4], d[5], d[6]);
  }
  SGCHK(s, hipEventRecord(e[1], s->stream));
  for (int k = 0; k < iters; k++)
    hipLaunchKernelGGL(k_hlld<double>, dim3(nblk(n)), dim3(256), 0, s->stream, n, s->gamma, d[0], d[1], d[2], jb, d[3],
                       d[4], d[5], d[6]);
  SGCHK(s, hipEventRecord(e[2], s->stream));
  for (int k = 0; k < iters; k++)
    hipLaunchKernelGGL(k_hlld<float>, dim3(nblk(n)), dim3(256), 0, s->stream, n, s->gamma, d[0], d[1], d[2], jb, d[3],
                       d[4], d[5], d[6]);
  SGCHK(s, hipEventRecord(e[3], s->stream));
  SGCHK(s, hipGetLastError());
  SGCHK(s, hipEventSynchronize(e[3]));
  for (int k = 0; k < 3; k++) {
    float t = 0.0f;
    SGCHK(s, hipEventElapsedTime(&t, e[k], e[k + 1]));
    ms[k] = (double)t / iters;
  }
  for (auto &x : e) (void)hipEventDestroy(x);
  return FVM_OK;
}

}  // extern "C"
