// Building blocks of the learned HLLD surrogate shared by the batch kernels
// (fvm_surrogate.hip) and the surrogate flux step of the FP32 face path
// (fvm_sg_step.h): the network's shape and packed weight layout, the face
// features, and the fused four-layer MLP on the 64 faces of a wavefront
// (bf16 operands on the matrix cores, FP32 accumulation).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/fvm_surrogate.h"
#include "fvm_physics.h"

namespace fvm {
namespace sg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 16, NH = 128, NOUT = 8;
constexpr int OW0 = 0, OB0 = OW0 + NIN * NH, OW1 = OB0 + NH, OB1 = OW1 + NH * NH, OW2 = OB1 + NH,
              OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;
static_assert(NPARAM == FVM_SG_NPARAM, "parameter count");

struct Gate {
  float max_log_rho, max_log_p, max_vn, max_log10_err;
};

// the face frame of computeFlux.cpp:17-107 (rotation rows, the two states in
// the frame, face speed, Dedner B_n and psi), in precision R
template <typename R> struct FrameT {
  R A[3][3];
  R rL, pL, eL, vL[3], BL[3];
  R rR, pR, eR, vR[3], BR[3];
  R psiL, psiR, entrL, entrR, wn, Bn, psi;
};

// ideal-MHD flux through a face at rest, normalised units: mass, mom n, mom t1,
// mom t2, energy, B t1, B t2 (e = thermal energy)
template <typename TC>
__device__ __forceinline__ void physical_flux(TC r, TC u, TC v1, TC v2, TC p, TC e, TC bn, TC b1, TC b2, TC *F) {
  const TC h = TC(0.5);
  const TC B2 = bn * bn + b1 * b1 + b2 * b2;
  const TC PT = p + h * B2;
  const TC E = e + h * r * (u * u + v1 * v1 + v2 * v2) + h * B2;
  const TC vB = u * bn + v1 * b1 + v2 * b2;
  F[0] = r * u;
  F[1] = r * u * u + PT - bn * bn;
  F[2] = r * u * v1 - bn * b1;
  F[3] = r * u * v2 - bn * b2;
  F[4] = (E + PT) * u - bn * vB;
  F[5] = u * b1 - bn * v1;
  F[6] = u * b2 - bn * v2;
}
constexpr double J_EPS = 1e-6;
__device__ __forceinline__ float tlog(float x) { return __logf(x); }
__device__ __forceinline__ double tlog(double x) { return log(x); }

// Scales, boost and the 16 features of a face, and the base flux Fb of its
// normalised moving-frame states, in precision TC (float for inference: the
// features are rounded to bf16 anyway; double for training targets, which
// divide by J); returns whether the face is inside the domain part of the gate.
// sc = {rs, cs, ut1, ut2} for the inverse map.
//
// Features: the mean state of the two sides (v_n, p, B_n, B_t1, B_t2; the
// scaling fixes the mean density to 1 and the boost the mean tangential
// velocity to 0), the unit direction d/J of the jump d = R - L over (rho, v_n,
// v_t1, v_t2, p, B_t1, B_t2) and its size J, the log density and pressure
// ratios, and the larger fast speed.  A smooth-flow jump is a small difference
// of O(1) values that bf16 cannot hold; its direction and size can.
//
// Base flux: the HLL flux of the two normalised states at the face (Harten,
// Lax & van Leer 1983) with the wave speeds riemannSolver_HLLD starts from
// (computeFlux.cpp:218-229): S_L = min(v_nL, v_nR) - c_f,max, S_R = max(...) +
// c_f,max.  The network answers (G' - Fb)/J, what the HLLD's intermediate
// states add per unit jump.
template <typename TC, typename RS>
__device__ __forceinline__ bool features(double gamma64, const FrameT<RS> &s, int jbnd, const Gate &G, float *x, TC *Fc,
                                         TC &J, TC *sc) {
  const TC gamma = (TC)gamma64, h = TC(0.5);
  const TC rL = (TC)s.rL, rR = (TC)s.rR, pL = (TC)s.pL, pR = (TC)s.pR, wn = (TC)s.wn, Bn = (TC)s.Bn;
  const TC vL[3] = {(TC)s.vL[0], (TC)s.vL[1], (TC)s.vL[2]}, vR[3] = {(TC)s.vR[0], (TC)s.vR[1], (TC)s.vR[2]};
  const TC BL[3] = {(TC)s.BL[0], (TC)s.BL[1], (TC)s.BL[2]}, BR[3] = {(TC)s.BR[0], (TC)s.BR[1], (TC)s.BR[2]};
  const TC rs = h * (rL + rR);
  const TC B2 = h * ((BL[0] * BL[0] + BL[1] * BL[1] + BL[2] * BL[2]) + (BR[0] * BR[0] + BR[1] * BR[1] + BR[2] * BR[2]));
  const TC cs = dsqrt(fdiv(gamma * h * (pL + pR) + B2, rs));
  const TC Ps = rs * cs * cs, Bs = dsqrt(Ps);
  const TC ut1 = h * (vL[1] + vR[1]), ut2 = h * (vL[2] + vR[2]);
  sc[0] = rs;
  sc[1] = cs;
  sc[2] = ut1;
  sc[3] = ut2;
  const TC ic = frcp(cs), iP = frcp(Ps), iB = frcp(Bs), irs = frcp(rs);
  const TC vnL = (vL[0] - wn) * ic, vnR = (vR[0] - wn) * ic;
  const TC eL = gamma64 > 1.0 ? (TC)s.eL * iP : pL * iP, eR = gamma64 > 1.0 ? (TC)s.eR * iP : pR * iP;
  // normalised sides: rho, v_n, v_t1, v_t2, p, e (thermal energy), B_n, B_t1, B_t2
  const TC L[9] = {rL * irs, vnL, (vL[1] - ut1) * ic, (vL[2] - ut2) * ic, pL * iP, eL, Bn * iB, BL[1] * iB, BL[2] * iB};
  const TC R[9] = {rR * irs, vnR, (vR[1] - ut1) * ic, (vR[2] - ut2) * ic, pR * iP, eR, Bn * iB, BR[1] * iB, BR[2] * iB};
  TC FL[7], FR[7];
  physical_flux<TC>(L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], L[8], FL);
  physical_flux<TC>(R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], FR);
  // HLL at the face (at rest in the moving frame)
  const TC b2L = L[6] * L[6] + L[7] * L[7] + L[8] * L[8], b2R = R[6] * R[6] + R[7] * R[7] + R[8] * R[8];
  const TC cfm = dsqrt(smax(fdiv(gamma * L[4] + b2L, L[0]), fdiv(gamma * R[4] + b2R, R[0])));
  const TC SL = smin(L[1], R[1]) - cfm, SR = smax(L[1], R[1]) + cfm;
  if (SL >= TC(0)) {
    for (int k = 0; k < 7; k++) Fc[k] = FL[k];
  } else if (SR <= TC(0)) {
    for (int k = 0; k < 7; k++) Fc[k] = FR[k];
  } else {
    const TC EL = L[5] + h * L[0] * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) + h * b2L;
    const TC ER = R[5] + h * R[0] * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]) + h * b2R;
    const TC UL[7] = {L[0], L[0] * L[1], L[0] * L[2], L[0] * L[3], EL, L[7], L[8]};
    const TC UR[7] = {R[0], R[0] * R[1], R[0] * R[2], R[0] * R[3], ER, R[7], R[8]};
    const TC idS = frcp(SR - SL), SLR = SL * SR;
    for (int k = 0; k < 7; k++) Fc[k] = (SR * FL[k] - SL * FR[k] + SLR * (UR[k] - UL[k])) * idS;
  }
  const int jk[7] = {0, 1, 2, 3, 4, 7, 8};
  TC d[7], j2 = TC(0);
  for (int k = 0; k < 7; k++) {
    d[k] = R[jk[k]] - L[jk[k]];
    j2 += sqr(d[k]);
  }
  J = dsqrt(j2) + (TC)J_EPS;
  const TC iJ = frcp(J);
  x[0] = (float)(h * (L[1] + R[1]));
  x[1] = (float)(h * (L[4] + R[4]));
  x[2] = (float)L[6];
  x[3] = (float)(h * (L[7] + R[7]));
  x[4] = (float)(h * (L[8] + R[8]));
  for (int k = 0; k < 7; k++) x[5 + k] = (float)(d[k] * iJ);
  x[12] = (float)J;
  const TC lr = tlog(fdiv(rR, rL)), lp = tlog(fdiv(pR, pL));
  x[13] = (float)lr;
  x[14] = (float)lp;
  x[15] = (float)cfm;
  const bool pos = s.rL > 0.0 && s.rR > 0.0 && s.pL > 0.0 && s.pR > 0.0;
  bool ok = jbnd == NO_BOUNDARY && pos;
  ok = ok && dabs(lr) <= (TC)G.max_log_rho && dabs(lp) <= (TC)G.max_log_p;
  ok = ok && dabs(vnL) <= (TC)G.max_vn && dabs(vnR) <= (TC)G.max_vn;
  for (int k = 0; k < 16; k++) ok = ok && __builtin_isfinite(x[k]);
  return ok;
}

// ---------------------------------------------------------------------------
// packed weight fragments (layout described in fvm_surrogate.hip)
// ---------------------------------------------------------------------------
constexpr int P_L0 = 8 * 64 * 4, P_L1 = 8 * 4 * 64 * 8, P_L3 = 4 * 64 * 8;
constexpr int P_BF = P_L0 + 2 * P_L1 + P_L3;  // bf16 elements
constexpr int P_BIAS = 3 * NH + 16;           // floats
constexpr int P_BYTES = P_BF * 2 + P_BIAS * 4;
__device__ __forceinline__ int kfeat(int c, int g, int e) { return 32 * c + (e < 4 ? 4 * g + e : 16 + 4 * g + e - 4); }

__device__ __forceinline__ bf16x4 relu_pack(f32x4 v) {
  return bf16x4{(__bf16)fmaxf(v[0], 0.0f), (__bf16)fmaxf(v[1], 0.0f), (__bf16)fmaxf(v[2], 0.0f),
                (__bf16)fmaxf(v[3], 0.0f)};
}
__device__ __forceinline__ bf16x8 cat(bf16x4 a, bf16x4 b) {
  return bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the network on the 64 faces of one wavefront: features feat[64][16] (bf16)
// in, outputs out[64][8] (f32, may alias feat) out
__device__ __forceinline__ void mlp64(const __bf16 *wl, const float *bias, const __bf16 *feat, float *out, int lane) {
  const int g = lane >> 4, r16 = lane & 15;
  bf16x4 h[8][4];
  {
    bf16x4 xin[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) xin[nt] = *reinterpret_cast<const bf16x4 *>(feat + (16 * nt + r16) * NIN + 4 * g);
    const bf16x4 *fr = reinterpret_cast<const bf16x4 *>(wl);
#pragma unroll
    for (int mt = 0; mt < 8; mt++) {
      const bf16x4 a = fr[mt * 64 + lane];
      const f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 16 * mt + 4 * g);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) h[mt][nt] = relu_pack(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xin[nt], b, 0, 0, 0));
    }
  }
#pragma unroll
  for (int layer = 0; layer < 2; layer++) {
    bf16x8 hin[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) hin[c][nt] = cat(h[2 * c][nt], h[2 * c + 1][nt]);
    const bf16x8 *fr = reinterpret_cast<const bf16x8 *>(wl + P_L0 + layer * P_L1);
    const float *bl = bias + NH * (layer + 1);
#pragma unroll
    for (int mt = 0; mt < 8; mt++) {
      f32x4 acc[4];
      const f32x4 b = *reinterpret_cast<const f32x4 *>(bl + 16 * mt + 4 * g);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[nt] = b;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const bf16x8 a = fr[(mt * 4 + c) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hin[c][nt], acc[nt], 0, 0, 0);
      }
#pragma unroll
      for (int nt = 0; nt < 4; nt++) h[mt][nt] = relu_pack(acc[nt]);
    }
  }
  {
    bf16x8 hin[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int nt = 0; nt < 4; nt++) hin[c][nt] = cat(h[2 * c][nt], h[2 * c + 1][nt]);
    const bf16x8 *fr = reinterpret_cast<const bf16x8 *>(wl + P_L0 + 2 * P_L1);
    const f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 3 * NH + 4 * g);
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = b;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const bf16x8 a = fr[c * 64 + lane];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hin[c][nt], acc[nt], 0, 0, 0);
    }
    wave_sync();  // every lane has read its features before outputs overwrite them
    if (g < 2) {
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int r = 0; r < 4; r++) out[(16 * nt + r16) * NOUT + 4 * g + r] = acc[nt][r];
    }
  }
}

}  // namespace sg
}  // namespace fvm
