// Diagnostics and snapshots of the per-cell state on the device:
//   fvm_energy       the Energy totals (Energy.h:27-79), deterministic FP64
//                    block sums in a fixed order
//   fvm_dump_binary  System::dump_binary's per-chare float32 snapshot
//                    (IO.cpp:7-99), records built on the device, one write
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"
#include "fvm_physics.h"

namespace fvm {
namespace diag {

// accumulators in Energy's order (Energy.h:6-17); the four totals are formed at the end
enum { E_MASS = 0, E_ETHM, E_EPOT, E_ETOT, E_EKINX, E_EKINY, E_EKINZ, E_EKIN, E_EMAGX, E_EMAGY, E_EMAGZ, E_EMAG,
       E_MOMX, E_MOMY, E_MOMZ, E_MOM, E_ANMX, E_ANMY, E_ANMZ, E_ANM, E_VOLUME, E_NVAR };
static_assert(E_NVAR == FVM_ENERGY_NVAR, "Energy layout");
constexpr int EB = 256;

__device__ __forceinline__ void cell_terms(const double *U, const double *vol, const double *pos, int i, int np,
                                           double *e) {
  double Ui[NF], W[NF];
  for (int k = 0; k < NF; k++) Ui[k] = U[(size_t)k * np + i];
  const double V = vol[i];
  to_primitive(Ui, V, W);
  const double M = W[DENS] * V;
  const double x = pos[i], y = pos[(size_t)np + i], z = pos[2 * (size_t)np + i];
  for (int k = 0; k < E_NVAR; k++) e[k] = 0.0;
  e[E_MASS] = Ui[MASS];
  e[E_VOLUME] = V;
  e[E_EKINX] = M * sqr(W[VELX]) * 0.5;
  e[E_EKINY] = M * sqr(W[VELY]) * 0.5;
  e[E_EKINZ] = M * sqr(W[VELZ]) * 0.5;
  e[E_EMAGX] = V * sqr(W[BX]) * 0.5;
  e[E_EMAGY] = V * sqr(W[BY]) * 0.5;
  e[E_EMAGZ] = V * sqr(W[BZ]) * 0.5;
  e[E_ETHM] = W[ETHM] * V;
  e[E_MOMX] = M * W[VELX];
  e[E_MOMY] = M * W[VELY];
  e[E_MOMZ] = M * W[VELZ];
  e[E_ANMX] = M * (W[VELY] * z - W[VELZ] * y);
  e[E_ANMY] = M * (W[VELZ] * x - W[VELX] * z);
  // Energy.h:72 pairs VELY with pos.z here; kept as the reference computes it
  e[E_ANMZ] = M * (W[VELX] * y - W[VELY] * z);
}

// per block: sums of E_NVAR accumulators over a contiguous cell range, tree order
__global__ void __launch_bounds__(EB) k_energy_part(const double *U, const double *vol, const double *pos, int n,
                                                    int np, double *part) {
  __shared__ double s[E_NVAR][EB];
  const int i = blockIdx.x * EB + threadIdx.x;
  double e[E_NVAR];
  if (i < n) cell_terms(U, vol, pos, i, np, e);
  else
    for (int k = 0; k < E_NVAR; k++) e[k] = 0.0;
  for (int k = 0; k < E_NVAR; k++) s[k][threadIdx.x] = e[k];
  __syncthreads();
  for (int h = EB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < E_NVAR; k++) s[k][threadIdx.x] += s[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < E_NVAR) part[(size_t)blockIdx.x * E_NVAR + threadIdx.x] = s[threadIdx.x][0];
}
// one block: sums of the block partials (each thread a strided subset, then a tree)
__global__ void __launch_bounds__(EB) k_energy_final(const double *part, int nb, double *out) {
  __shared__ double s[E_NVAR][EB];
  double e[E_NVAR];
  for (int k = 0; k < E_NVAR; k++) e[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += EB)
    for (int k = 0; k < E_NVAR; k++) e[k] += part[(size_t)b * E_NVAR + k];
  for (int k = 0; k < E_NVAR; k++) s[k][threadIdx.x] = e[k];
  __syncthreads();
  for (int h = EB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < E_NVAR; k++) s[k][threadIdx.x] += s[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double d[E_NVAR];
    for (int k = 0; k < E_NVAR; k++) d[k] = s[k][0];
    d[E_EKIN] = d[E_EKINX] + d[E_EKINY] + d[E_EKINZ];
    d[E_EMAG] = d[E_EMAGX] + d[E_EMAGY] + d[E_EMAGZ];
    d[E_ETOT] = d[E_ETHM] + d[E_EKIN] + d[E_EMAG] + d[E_EPOT];
    d[E_MOM] = d[E_MOMX] + d[E_MOMY] + d[E_MOMZ];
    d[E_ANM] = d[E_ANMX] + d[E_ANMY] + d[E_ANMZ];
    for (int k = 0; k < E_NVAR; k++) out[k] = d[k];
  }
}

// snapshot record of cell i (IO.cpp:67-95): 26 words between two length markers
constexpr int SNAP_WORDS = 28;
template <typename S>
__global__ void __launch_bounds__(256) k_snapshot(const double *U, const double *vol, const double *pos, const S *vel,
                                                  const int *bnd, const double *divB, const int *idx,
                                                  const double *dtc, double dt_global, double gamma, int n, int np,
                                                  unsigned *words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double Ui[NF], W[NF];
  for (int k = 0; k < NF; k++) Ui[k] = U[(size_t)k * np + i];
  const double V = vol[i];
  to_primitive(Ui, V, W);
  const double L = pow(V * 3.0 / 4.0 / REF_PI, 1.0 / 3.0);
  unsigned *w = words + (size_t)SNAP_WORDS * i;
  auto fw = [&](int k, double v) { w[k] = __float_as_uint((float)v); };
  auto iw = [&](int k, int v) { w[k] = (unsigned)v; };
  iw(0, 26 * 4);
  iw(1, idx ? idx[i] : i);
  for (int d = 0; d < 3; d++) fw(2 + d, pos[(size_t)d * np + i]);
  for (int d = 0; d < 3; d++) fw(5 + d, (double)vel[(size_t)d * np + i]);
  fw(8, W[DENS]);
  fw(9, W[ETHM]);
  fw(10, pressure(gamma, W[ETHM]));
  fw(11, -1.0);
  iw(12, bnd[i]);
  fw(13, W[VELX]);
  fw(14, W[VELY]);
  fw(15, W[VELZ]);
  fw(16, W[BX]);
  fw(17, W[BY]);
  fw(18, W[BZ]);
  fw(19, dtc ? dtc[i] : dt_global);
  fw(20, V);
  fw(21, W[PSI]);
  fw(22, L * divB[i]);
  fw(23, W[ENTR]);
  fw(24, -1.0);
  fw(25, -1.0);
  fw(26, -1.0);
  iw(27, 26 * 4);
}

}  // namespace diag
}  // namespace fvm

using namespace fvm::diag;

int fvm_energy(fvm_ctx *c, double *out) {
  if (!c || !out) return FVM_EARG;
  if (!c->have_mesh) return fvm_fail(c, FVM_ESTATE, "fvm_energy before fvm_upload_mesh");
  (void)hipSetDevice(c->device);
  const int n = c->nactive, np = c->np;
  const int nb = std::max(1, (n + EB - 1) / EB);
  DBuf part;
  int rc = fvm_ensure(c, part, sizeof(double) * ((size_t)nb * E_NVAR + E_NVAR));
  if (rc) return rc;
  double *p = (double *)part.p;
  hipLaunchKernelGGL(k_energy_part, dim3(nb), dim3(EB), 0, c->stream, (const double *)c->field[FVM_FIELD_U].p,
                     (const double *)c->field[FVM_FIELD_MESH_VOLUME].p, (const double *)c->field[FVM_FIELD_POS].p, n,
                     np, p);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(EB), 0, c->stream, (const double *)p, nb, p + (size_t)nb * E_NVAR);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, p + (size_t)nb * E_NVAR, sizeof(double) * E_NVAR, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  fvm_release(part);
  if (e != hipSuccess) return fvm_fail(c, FVM_EHIP, hipGetErrorString(e));
  return FVM_OK;
}

template <typename S>
static int snapshot_words(fvm_ctx *c, const int *idx, const double *dtc, double dt_global, std::vector<unsigned> &h) {
  const int n = c->nactive, np = c->np;
  DBuf wb, ib, db;
  int rc = fvm_ensure(c, wb, sizeof(unsigned) * SNAP_WORDS * (size_t)std::max(n, 1));
  if (!rc && idx) rc = fvm_ensure(c, ib, sizeof(int) * (size_t)n);
  if (!rc && dtc) rc = fvm_ensure(c, db, sizeof(double) * (size_t)n);
  hipError_t e = hipSuccess;
  if (!rc) {
    if (idx) e = hipMemcpyAsync(ib.p, idx, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && dtc) e = hipMemcpyAsync(db.p, dtc, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(k_snapshot<S>, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                         (const double *)c->field[FVM_FIELD_U].p, (const double *)c->field[FVM_FIELD_MESH_VOLUME].p,
                         (const double *)c->field[FVM_FIELD_POS].p, (const S *)c->field[FVM_FIELD_VEL].p,
                         (const int *)c->field[FVM_FIELD_BND].p, (const double *)c->field[FVM_FIELD_DIVB].p,
                         idx ? (const int *)ib.p : nullptr, dtc ? (const double *)db.p : nullptr, dt_global,
                         c->prm.gamma_gas, n, np, (unsigned *)wb.p);
      e = hipGetLastError();
    }
    h.resize((size_t)SNAP_WORDS * n);
    if (e == hipSuccess && n > 0)
      e = hipMemcpyAsync(h.data(), wb.p, sizeof(unsigned) * h.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  fvm_release(wb);
  fvm_release(ib);
  fvm_release(db);
  if (rc) return rc;
  if (e != hipSuccess) return fvm_fail(c, FVM_EHIP, hipGetErrorString(e));
  return FVM_OK;
}

int fvm_dump_binary(fvm_ctx *c, const char *path, const fvm_snapshot_info *info, const int *idx, const double *dt_cell) {
  if (!c || !path || !info) return FVM_EARG;
  if (!c->have_mesh) return fvm_fail(c, FVM_ESTATE, "fvm_dump_binary before fvm_upload_mesh");
  (void)hipSetDevice(c->device);
  std::vector<unsigned> rec;
  const int rc = c->prm.storage == FVM_STORAGE_SHIPPED ? snapshot_words<float>(c, idx, dt_cell, info->dt_global, rec)
                                                       : snapshot_words<double>(c, idx, dt_cell, info->dt_global, rec);
  if (rc) return rc;
  // header (IO.cpp:29-63): 20 words between two length markers
  unsigned hd[22];
  auto fw = [](double v) { float f = (float)v; unsigned u; std::memcpy(&u, &f, 4); return u; };
  int k = 0;
  hd[k++] = 20 * 4;
  hd[k++] = (unsigned)info->chare_index;
  hd[k++] = (unsigned)info->num_elements;
  hd[k++] = (unsigned)info->num_elements;
  hd[k++] = (unsigned)(info->tsysU & 0xffffffffULL);
  hd[k++] = (unsigned)(info->tsysU >> 32);
  hd[k++] = (unsigned)info->global_n;
  hd[k++] = (unsigned)c->nactive;
  hd[k++] = 3u;
  hd[k++] = fw(c->t_global);
  hd[k++] = fw(info->dt_global);
  hd[k++] = (unsigned)info->iteration;
  hd[k++] = fw(c->prm.courant_no);
  hd[k++] = fw(c->prm.gamma_gas);
  hd[k++] = (unsigned)-1;
  for (int d = 0; d < 3; d++) hd[k++] = fw(info->rmin[d]);
  for (int d = 0; d < 3; d++) hd[k++] = fw(info->rmax[d]);
  hd[k++] = 20 * 4;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fvm_fail(c, FVM_EARG, std::string("cannot open ") + path);
  bool ok = std::fwrite(hd, sizeof(unsigned), 22, f) == 22;
  if (ok && !rec.empty()) ok = std::fwrite(rec.data(), sizeof(unsigned), rec.size(), f) == rec.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fvm_fail(c, FVM_EARG, std::string("write failed: ") + path);
  return FVM_OK;
}
