// The learned HLLD surrogate inside the FP32 face path's step (no counterpart
// in the reference; it stands in for riemannSolver_HLLD, computeFlux.cpp:178-425,
// at its call in computeFluidUpdate, computeFluidUpdate.cpp:211).
//
// With a surrogate set, the flux + apply launch of the FP32 face path becomes
// two launches that evaluate every face once:
//
//   k_sg_faces   one lane per face in canonical orientation (L = s1, R = s2):
//                gathers both cells' records, forms the two space-time
//                extrapolated face states exactly as the FP32 flux kernel does,
//                the face frame and Dedner B_n/psi in FP32, the 16 features,
//                then the four-layer network on the 64 faces of the wavefront
//                (bf16 MFMA, FP32 accumulation, weights in LDS).  Faces inside
//                the gate take the network's flux; every other face takes the
//                FP32 HLLD (compute_flux<float>) in the same launch.  Writes a
//                64-byte record per face: flux[10], psi, B_n, area, normal.
//   k_sg_cells   one lane per cell (SELL order): the FP64 sums of flux*dt*A over
//                the cell's face-ends in its face order, then computePrimitives
//                (FP64), as the FP32 flux kernel ends.
//
// A gate that accepts no face gives the FP32 face path's results bit for bit:
// the face states, the HLLD and the sums are the same operations.
//
// Included by fvm_fused_f32.hip after fvm_fused.h (internal linkage, FP32 only).
#pragma once
#include "fvm_mlp.h"

namespace fvm {
namespace {

constexpr int SGB = 512;                    // threads per face block (8 wavefronts)
constexpr int SG_FEAT_B = 64 * sg::NIN * 2;  // per wavefront: features (bf16), then outputs (f32)
static_assert(SG_FEAT_B >= 64 * sg::NOUT * 4, "output tile fits the feature tile");
constexpr int SG_LDS_B = sg::P_BYTES + (SGB / 64) * SG_FEAT_B;
static_assert(SG_LDS_B <= 160 * 1024, "LDS budget");

struct SgStepArgs {
  int nface;
  const int *faceL;                // [nface] L cell (internal index)
  const FaceEndRec<float> *face;   // [nface] the L face-end record (nbr = R, normal, drL, area)
  const int *fid;                  // [slots] face of each face-end slot; ~face on the R side
  float4 *flux;                    // [nface][4]: flux[10], psi, B_n | area, normal
  const __bf16 *pfrag;
  const float *pbias;
  sg::Gate gate;
};

// rotation rows of the face frame (computeFlux.cpp:17-40), as compute_flux<float> forms them
__device__ __forceinline__ void rows32(const float *nrm, float A[3][3]) {
  const float ex = nrm[0], ey = nrm[1], ez = nrm[2];
  const float s2 = ex * ex + ey * ey;
  const bool axis = !(s2 > 1.0e-30f);
  const float r = __builtin_amdgcn_rsqf(axis ? 1.0f : s2);
  const float ds = axis ? 0.0f : s2 * r, ids = axis ? 0.0f : r;
  const float cosph = (ids == 0.0f) ? 1.0f : ex * ids;
  const float sinph = (ids == 0.0f) ? 0.0f : ey * ids;
  const float costh = ez, sinth = ds;
  A[0][0] = cosph * sinth;
  A[0][1] = sinth * sinph;
  A[0][2] = costh;
  A[1][0] = -sinph;
  A[1][1] = cosph;
  A[1][2] = 0.0f;
  A[2][0] = -costh * cosph;
  A[2][1] = -costh * sinph;
  A[2][2] = sinth;
}

// FP32 face frame and Dedner B_n/psi of a static face (w = 0), the operations
// of compute_flux<float> (computeFlux.cpp:17-107)
__device__ __forceinline__ void frame32(float gamma, const float *WL, const float *WR, const float *nrm,
                                        sg::FrameT<float> &s) {
  float A[3][3];
  rows32(nrm, A);
  for (int a = 0; a < 3; a++) {
    for (int b = 0; b < 3; b++) s.A[a][b] = A[a][b];
    s.vL[a] = A[a][0] * WL[VELX] + A[a][1] * WL[VELY] + A[a][2] * WL[VELZ];
    s.BL[a] = A[a][0] * WL[BX] + A[a][1] * WL[BY] + A[a][2] * WL[BZ];
    s.vR[a] = A[a][0] * WR[VELX] + A[a][1] * WR[VELY] + A[a][2] * WR[VELZ];
    s.BR[a] = A[a][0] * WR[BX] + A[a][1] * WR[BY] + A[a][2] * WR[BZ];
  }
  s.rL = WL[DENS];
  s.eL = WL[ETHM];
  s.rR = WR[DENS];
  s.eR = WR[ETHM];
  s.psiL = WL[PSI];
  s.psiR = WR[PSI];
  s.entrL = WL[ENTR];
  s.entrR = WR[ENTR];
  s.pL = pressure(gamma, WL[ETHM]);
  s.pR = pressure(gamma, WR[ETHM]);
  s.wn = 0.0f;
  const float cfl2 = fdiv(gamma * s.pL + sqr(s.BL[0]) + sqr(s.BL[1]) + sqr(s.BL[2]), s.rL);
  const float cfr2 = fdiv(gamma * s.pR + sqr(s.BR[0]) + sqr(s.BR[1]) + sqr(s.BR[2]), s.rR);
  const float chL = fsqrt(cfl2), chR = fsqrt(cfr2);
  s.Bn = fdiv(s.BL[0] * chL + s.BR[0] * chR + (s.psiL - s.psiR), chL + chR);
  s.psi = fdiv(s.psiL * chR + s.psiR * chL + chR * chL * (s.BL[0] - s.BR[0]), chL + chR);
}

// one extrapolated face state from a gathered record (the flux kernel's operations)
__device__ __forceinline__ void face_state_rec(const float *__restrict__ rec, const float *dr, float *Wf) {
  float x[4 * GQ];
#pragma unroll
  for (int k = 0; k < GQ; k++) {
    const float4 v = reinterpret_cast<const float4 *>(rec)[k];
    x[4 * k] = v.x; x[4 * k + 1] = v.y; x[4 * k + 2] = v.z; x[4 * k + 3] = v.w;
  }
  for (int k = 0; k < NF; k++)
    Wf[k] = face_state32(x[k], x[R32_G + k], x[R32_G + NF + k], x[R32_G + 2 * NF + k], dr);
}

// Faces in batches of 64 (one per wavefront).  Block b runs on XCD b % 8 and
// the XCDs each take a contiguous eighth of the batches, so the neighbour
// records a block gathers stay in one XCD's L2.
__global__ void __launch_bounds__(SGB) k_sg_faces(FusedArgs<float> A, SgStepArgs S) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __bf16 *wl = reinterpret_cast<__bf16 *>(smem);
  float *bias = reinterpret_cast<float *>(smem + sg::P_BF * 2);
  {
    const int4 *src = reinterpret_cast<const int4 *>(S.pfrag);
    int4 *dst = reinterpret_cast<int4 *>(smem);
    for (int i = threadIdx.x; i < sg::P_BF * 2 / 16; i += SGB) dst[i] = src[i];
    for (int i = threadIdx.x; i < sg::P_BIAS; i += SGB) bias[i] = S.pbias[i];
  }
  __syncthreads();
  const int lane = (int)__lane_id(), wv = (int)threadIdx.x >> 6;
  __bf16 *feat = reinterpret_cast<__bf16 *>(smem + sg::P_BYTES + wv * SG_FEAT_B);
  float *out = reinterpret_cast<float *>(feat);

  const float g = (float)A.gamma;
  const int nbatch = (S.nface + 63) >> 6;
  const int bpx = (int)gridDim.x >> 3;  // blocks per XCD
  const int per = (nbatch + 7) >> 3;    // batches per XCD
  const int xcd = (int)blockIdx.x & 7;
  const int b0 = xcd * per, b1 = min(nbatch, b0 + per);
  int nbad = 0, nfb = 0;
  for (int bw = b0 + ((int)blockIdx.x >> 3) * (SGB / 64) + wv; bw < b1; bw += bpx * (SGB / 64)) {
    const int f = (bw << 6) + lane;
    const bool live = f < S.nface;
    // held through the network: the face geometry, the features' scales and
    // central flux, and the Dedner / passive-scalar face values (the two face
    // states are formed again for a face the gate sends to the HLLD)
    float nrm[3] = {0.0f, 0.0f, 1.0f}, area = 0.0f;
    float x[sg::NIN], Fc[7], J = 0.0f, sc[4] = {1.0f, 1.0f, 0.0f, 0.0f};
    float Bn = 0.0f, psi = 0.0f, psiL = 0.0f, psiR = 0.0f, entrL = 0.0f, entrR = 0.0f;
    bool ok = false;
    if (live) {
      const int L = S.faceL[f];
      FE<float> fe;
      decode_fe(S.face[f], fe);
      float WL[NF], WR[NF];
      face_state_rec(A.rec + (size_t)L * REC32, fe.drL, WL);
      face_state_rec(A.rec + (size_t)fe.j * REC32, fe.drR, WR);
      nbad += (!(WL[DENS] > 0.0f) || !(WL[ETHM] > 0.0f) || !(WR[DENS] > 0.0f) || !(WR[ETHM] > 0.0f)) ? 1 : 0;
      for (int d = 0; d < 3; d++) nrm[d] = fe.nrm[d];
      area = fe.area;
      sg::FrameT<float> s;
      frame32(g, WL, WR, nrm, s);
      ok = sg::features<float>(A.gamma, s, NO_BOUNDARY, S.gate, x, Fc, J, sc);
      Bn = s.Bn;
      psi = s.psi;
      psiL = s.psiL;
      psiR = s.psiR;
      entrL = s.entrL;
      entrR = s.entrR;
    }
    if (!ok)
      for (int k = 0; k < sg::NIN; k++) x[k] = 0.0f;
    {
      sg::bf16x8 lo, hi;
      for (int k = 0; k < 8; k++) {
        lo[k] = (__bf16)x[k];
        hi[k] = (__bf16)x[8 + k];
      }
      *reinterpret_cast<sg::bf16x8 *>(feat + lane * sg::NIN) = lo;
      *reinterpret_cast<sg::bf16x8 *>(feat + lane * sg::NIN + 8) = hi;
    }
    sg::wave_sync();
    sg::mlp64(wl, bias, feat, out, lane);
    sg::wave_sync();
    float y[sg::NOUT];
    for (int k = 0; k < sg::NOUT; k++) y[k] = out[lane * sg::NOUT + k];
    sg::wave_sync();  // outputs read before the next batch's features land
    if (!live) continue;
    ok = ok && y[7] <= S.gate.max_log10_err;
    float flux[NF];
    if (ok) {
      // normalised moving-frame fluxes back to the face frame (face speed 0),
      // then to the mesh frame
      float R3[3][3];
      rows32(nrm, R3);
      const float rs = sc[0], cs = sc[1], Ps = rs * cs * cs, Bs = fsqrt(Ps);
      const float V[3] = {0.0f, sc[2], sc[3]};
      float gg[7];
      for (int k = 0; k < 7; k++) gg[k] = Fc[k] + J * y[k];
      const float Gm = gg[0] * rs * cs;
      const float Gp[3] = {gg[1] * Ps, gg[2] * Ps, gg[3] * Ps};
      const float GE = gg[4] * Ps * cs + (V[0] * Gp[0] + V[1] * Gp[1] + V[2] * Gp[2]) +
                       0.5f * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) * Gm;
      const float Gmom[3] = {Gp[0] + V[0] * Gm, Gp[1] + V[1] * Gm, Gp[2] + V[2] * Gm};
      const float GB[3] = {0.0f, gg[5] * Bs * cs - Bn * V[1], gg[6] * Bs * cs - Bn * V[2]};
      flux[MASS] = Gm;
      flux[ENER] = GE;
      for (int d = 0; d < 3; d++) {
        flux[MOMX + d] = R3[0][d] * Gmom[0] + R3[1][d] * Gmom[1] + R3[2][d] * Gmom[2];
        flux[WBX + d] = R3[0][d] * GB[0] + R3[1][d] * GB[1] + R3[2][d] * GB[2];
      }
      flux[MPSI] = Gm * (Gm > 0.0f ? psiL : psiR);
      flux[MENTR] = Gm * (Gm > 0.0f ? entrL : entrR);
    } else {
      // the FP32 HLLD on the same two face states
      const int L = S.faceL[f];
      FE<float> fe;
      decode_fe(S.face[f], fe);
      float WL[NF], WR[NF];
      face_state_rec(A.rec + (size_t)L * REC32, fe.drL, WL);
      face_state_rec(A.rec + (size_t)fe.j * REC32, fe.drR, WR);
      const float wij[3] = {0.0f, 0.0f, 0.0f};
      compute_flux<float>(g, wij, WL, WR, NO_BOUNDARY, nrm, psi, Bn, flux);
      nfb++;
    }
    f4v_t *o = reinterpret_cast<f4v_t *>(S.flux + 4 * (size_t)f);
    nts(f4v_t{flux[0], flux[1], flux[2], flux[3]}, o);
    nts(f4v_t{flux[4], flux[5], flux[6], flux[7]}, o + 1);
    nts(f4v_t{flux[8], flux[9], psi, Bn}, o + 2);
    nts(f4v_t{area, nrm[0], nrm[1], nrm[2]}, o + 3);
  }
  wave_add(&A.diag[FVM_DIAG_POSITIVITY], nbad);
  wave_add(&A.diag[FVM_DIAG_SG_FALLBACK], nfb);
}

// per cell: the FP64 sums of the face records in the cell's face order
// (computeFluidUpdate.cpp:209-307), then computePrimitives
__global__ void __launch_bounds__(FB) k_sg_cells(FusedArgs<float> A, SgStepArgs S) {
  const int c = xcd_cell(A.n);
  if (c >= A.n) return;
  const StepTimes st = step_times(A);
  double dU[NF], gp[3], dv = 0.0;
  for (int k = 0; k < NF; k++) dU[k] = 0.0;
  gp[0] = gp[1] = gp[2] = 0.0;
  const int deg = A.deg[c];
  const int e0 = A.gbase[c >> 6] + (c & 63);
  int fn = deg > 0 ? S.fid[e0] : 0;
  for (int q = 0, e = e0; q < deg; q++, e += 64) {
    const int fc = fn;
    if (q + 1 < deg) fn = S.fid[e + 64];
    const bool isL = fc >= 0;
    const float4 *r = S.flux + 4 * (size_t)(isL ? fc : ~fc);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    const float flux[NF] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y};
    const float psi = r2.z, Bn = r2.w, nrm[3] = {r3.y, r3.z, r3.w};
    const double tarea = st.dtf * (double)r3.x;
    const double sa = isL ? -tarea : tarea;
    for (int k = 0; k < NF; k++) dU[k] = dU[k] + (double)flux[k] * sa;
    for (int d = 0; d < 3; d++) gp[d] = gp[d] - (double)nrm[d] * ((double)psi * sa);
    dv = dv - (double)Bn * sa;
  }
  float tc[NF];
  ldv<NF>(A.tv + (size_t)c * NF, tc);
  cell_update<float>(A, c, st.t0, st.dt_glob, dU, gp, dv, tc);
}

// ---- face list of the surrogate step, built once per fused setup ----------
// per face-end slot: the mesh face and whether this cell is its L side
__global__ void k_sg_slot_faces(int n, const int *perm, const int *coff, const int *cface, const int *s1,
                                const int *deg, const int *gbase, int *sface, int *isl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  for (int q = 0; q < deg[i]; q++) {
    const int slot = gbase[i >> 6] + 64 * q + (i & 63);
    const int f = cface[coff[o] + q];
    sface[slot] = f;
    isl[slot] = s1[f] == o ? 1 : 0;
  }
}
// L face-ends: their face id (scan position), L cell, record; mesh face -> id
__global__ void k_sg_build_faces(int n, const int *deg, const int *gbase, const int *sface, const int *isl,
                                 const int *pos, const FaceEndRec<float> *fe, int *faceL, FaceEndRec<float> *face,
                                 int *m2f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < deg[i]; q++) {
    const int slot = gbase[i >> 6] + 64 * q + (i & 63);
    if (!isl[slot]) continue;
    const int id = pos[slot];
    faceL[id] = i;
    face[id] = fe[slot];
    m2f[sface[slot]] = id;
  }
}
__global__ void k_sg_build_fid(int n, const int *deg, const int *gbase, const int *sface, const int *isl,
                               const int *m2f, int *fid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < deg[i]; q++) {
    const int slot = gbase[i >> 6] + 64 * q + (i & 63);
    const int id = m2f[sface[slot]];
    fid[slot] = isl[slot] ? id : ~id;
  }
}

int sg_step_setup(fvm_ctx *c) {
  FusedState &z = c->fz;
  const int n = c->np;
  const size_t slots = z.slots;
  if (slots > (size_t)0x7fffffff) return fvm_fail(c, FVM_EARG, "surrogate step: face-end slots exceed 2^31");
  DBuf sface, isl, pos, m2f;
  auto cleanup = [&]() {
    fvm_release(sface);
    fvm_release(isl);
    fvm_release(pos);
    fvm_release(m2f);
  };
  int rc;
  if ((rc = fvm_ensure(c, sface, sizeof(int) * slots)) || (rc = fvm_ensure(c, isl, sizeof(int) * slots)) ||
      (rc = fvm_ensure(c, pos, sizeof(int) * slots)) || (rc = fvm_ensure(c, m2f, sizeof(int) * (size_t)c->nface))) {
    cleanup();
    return rc;
  }
  const unsigned nb = (unsigned)((n + 255) / 256);
  FVM_HIPCHK(c, hipMemsetAsync(isl.p, 0, sizeof(int) * slots, c->stream));
  hipLaunchKernelGGL(k_sg_slot_faces, dim3(nb), dim3(256), 0, c->stream, n, (const int *)z.perm.p,
                     (const int *)c->m_coff.p, (const int *)c->m_cface.p, (const int *)c->m_s1.p, (const int *)z.deg.p,
                     (const int *)z.gbase.p, (int *)sface.p, (int *)isl.p);
  FVM_HIPCHK(c, hipGetLastError());
  int total = 0;
  if ((rc = exclusive_scan_flags(c->scan, (const int *)isl.p, (int *)pos.p, (int)slots, &total, c->stream))) {
    cleanup();
    return fvm_fail(c, rc, "surrogate step: scan failed");
  }
  if (total != c->nface) {
    cleanup();
    return fvm_fail(c, FVM_EARG, "surrogate step: every face must have its L cell among the cells");
  }
  if ((rc = fvm_ensure(c, z.sg_faceL, sizeof(int) * (size_t)total)) ||
      (rc = fvm_ensure(c, z.sg_face, sizeof(FaceEndRec<float>) * (size_t)total)) ||
      (rc = fvm_ensure(c, z.sg_fid, sizeof(int) * slots)) ||
      (rc = fvm_ensure(c, z.sg_flux, 4 * sizeof(float4) * (size_t)total))) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_sg_build_faces, dim3(nb), dim3(256), 0, c->stream, n, (const int *)z.deg.p,
                     (const int *)z.gbase.p, (const int *)sface.p, (const int *)isl.p, (const int *)pos.p,
                     (const FaceEndRec<float> *)z.fe.p, (int *)z.sg_faceL.p, (FaceEndRec<float> *)z.sg_face.p,
                     (int *)m2f.p);
  FVM_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sg_build_fid, dim3(nb), dim3(256), 0, c->stream, n, (const int *)z.deg.p,
                     (const int *)z.gbase.p, (const int *)sface.p, (const int *)isl.p, (const int *)m2f.p,
                     (int *)z.sg_fid.p);
  FVM_HIPCHK(c, hipGetLastError());
  FVM_HIPCHK(c, hipStreamSynchronize(c->stream));
  cleanup();
  z.sg_nface = total;
  return FVM_OK;
}

int sg_step_launch(fvm_ctx *c, const FusedArgs<float> &A) {
  FusedState &z = c->fz;
  SgStepArgs S;
  S.nface = z.sg_nface;
  S.faceL = (const int *)z.sg_faceL.p;
  S.face = (const FaceEndRec<float> *)z.sg_face.p;
  S.fid = (const int *)z.sg_fid.p;
  S.flux = (float4 *)z.sg_flux.p;
  const void *frag = nullptr;
  const float *bias = nullptr;
  float gate[4];
  int rc = fvm_sg_device_weights(z.sg, &frag, &bias, gate);
  if (rc) return fvm_fail(c, rc, "surrogate step: the surrogate's weights are not available");
  S.pfrag = (const __bf16 *)frag;
  S.pbias = bias;
  S.gate = sg::Gate{gate[0], gate[1], gate[2], gate[3]};
  static bool attr = false;
  if (!attr) {
    FVM_HIPCHK(c, hipFuncSetAttribute((const void *)k_sg_faces, hipFuncAttributeMaxDynamicSharedMemorySize, SG_LDS_B));
    attr = true;
  }
  // one block per CU (the weights fill half of its LDS), 32 per XCD
  hipLaunchKernelGGL(k_sg_faces, dim3(256), dim3(SGB), SG_LDS_B, c->stream, A, S);
  FVM_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_sg_cells, dim3(fused_grid(c->np)), dim3(FB), 0, c->stream, A, S);
  FVM_HIPCHK(c, hipGetLastError());
  return FVM_OK;
}

}  // namespace
}  // namespace fvm
