// C ABI of the fused stepping (the kernels and their host drivers are in
// fvm_fused.h, compiled once per face precision) and the stream-copy probe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"

void fvm_fast_release(fvm_ctx *c) {
  FusedState &z = c->fz;
  for (DBuf *b : {&z.perm, &z.inv, &z.deg, &z.gbase, &z.fe, &z.fdr, &z.fnb, &z.lmin, &z.U, &z.W64, &z.cvol, &z.mvel, &z.rec[0], &z.rec[1], &z.wc[0], &z.wc[1], &z.tv, &z.mm,
                  &z.dt_new, &z.dtmin, &z.tdev, &z.hist, &z.sg_faceL, &z.sg_face, &z.sg_fid, &z.sg_flux})
    fvm_release(*b);
  z.sg = nullptr;
  z.sg_nface = 0;
  for (auto &e : z.ev)
    if (e) (void)hipEventDestroy(e);
  z.ev.clear();
  z.ready = false;
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
  int rc;
  try {
    rc = precision == FVM_FUSED_FP64 ? fvm_fused_setup_f64(c) : fvm_fused_setup_f32(c);
  } catch (...) {
    return fvm_fail(c, FVM_EARG, "host allocation failed");
  }
  if (rc) return rc;
  c->fz.np = c->np;
  c->fz.ready = true;
  return FVM_OK;
}

int fvm_fused_run(fvm_ctx *c, int nsteps, double dtmax, double *t_out) {
  if (!c) return FVM_EARG;
  if (!c->fz.ready || !c->have_mesh || c->np != c->fz.np)
    return fvm_fail(c, FVM_ESTATE, "fvm_fused_run without a fvm_fused_setup on the current mesh");
  if (nsteps < 0 || !(dtmax > 0.0)) return fvm_fail(c, FVM_EARG, "bad nsteps/dtmax");
  (void)hipSetDevice(c->device);
  FusedState &z = c->fz;
  int rc = fvm_ensure(c, z.hist, sizeof(double) * (size_t)std::max(nsteps, 1));
  if (rc) return rc;
  rc = z.precision == FVM_FUSED_FP64 ? fvm_fused_run_f64(c, nsteps, dtmax) : fvm_fused_run_f32(c, nsteps, dtmax);
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

int fvm_fused_sync(fvm_ctx *c) {
  if (!c) return FVM_EARG;
  if (!c->fz.ready || !c->have_mesh || c->np != c->fz.np)
    return fvm_fail(c, FVM_ESTATE, "fvm_fused_sync without a fvm_fused_setup on the current mesh");
  (void)hipSetDevice(c->device);
  return c->fz.precision == FVM_FUSED_FP64 ? fvm_fused_sync_f64(c) : fvm_fused_sync_f32(c);
}

int fvm_fused_set_surrogate(fvm_ctx *c, fvm_sg *sg) {
  if (!c) return FVM_EARG;
  if (!sg) {
    c->fz.sg = nullptr;
    return FVM_OK;
  }
  if (!c->fz.ready || !c->have_mesh || c->np != c->fz.np)
    return fvm_fail(c, FVM_ESTATE, "fvm_fused_set_surrogate without a fvm_fused_setup on the current mesh");
  if (c->fz.precision != FVM_FUSED_FP32) return fvm_fail(c, FVM_EARG, "the surrogate step runs on the FP32 face path");
  if (c->fz.moving) return fvm_fail(c, FVM_EARG, "the surrogate step takes static meshes (face speed 0)");
  if (fvm_sg_device_of(sg) != c->device) return fvm_fail(c, FVM_EARG, "the surrogate lives on another device");
  (void)hipSetDevice(c->device);
  if (c->fz.sg_nface == 0) {
    int rc;
    try {
      rc = fvm_fused_sg_setup_f32(c);
    } catch (...) {
      return fvm_fail(c, FVM_EARG, "host allocation failed");
    }
    if (rc) return rc;
  }
  c->fz.sg = sg;
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

// ===========================================================================
// stream-copy bandwidth probe (the measured HBM ceiling next to the 8 TB/s
// data-sheet peak; also the calibration kernel for counter-based traffic)
// ===========================================================================
namespace fvm {
typedef float f4v_t __attribute__((ext_vector_type(4)));
// each thread moves UNR 16-byte words per iteration, all loads issued before
// the stores; non-temporal, so the stream does not linger in L2
constexpr int COPY_UNR = 4;
__global__ void __launch_bounds__(256) k_stream_copy(const f4v_t *__restrict__ src, f4v_t *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (COPY_UNR - 1) * stride < n; i += COPY_UNR * stride) {
    f4v_t v[COPY_UNR];
#pragma unroll
    for (int u = 0; u < COPY_UNR; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < COPY_UNR; u++) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
}  // namespace fvm

// the library's stream copy (runtime == false) or the HIP runtime's own
// device-to-device copy (runtime == true) of `bytes`, iters times after one
// untimed pass; read plus written bytes per second
static int copy_bandwidth(fvm_ctx *c, long long bytes, int iters, double *gbps, bool runtime) {
  if (!c || bytes < (1 << 20) || iters < 1 || !gbps) return FVM_EARG;
  (void)hipSetDevice(c->device);
  const size_t n = (size_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  FVM_HIPCHK(c, hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return fvm_fail(c, FVM_EHIP, "hipMalloc failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemsetAsync(a, 0, n * 16, c->stream);
  const unsigned grid = 256 * 8 * 4;
  for (int k = 0; k <= iters; k++) {
    if (k == 1) (void)hipEventRecord(e0, c->stream);
    if (runtime)
      (void)hipMemcpyAsync(b, a, n * 16, hipMemcpyDeviceToDevice, c->stream);
    else
      hipLaunchKernelGGL(fvm::k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (const fvm::f4v_t *)a, (fvm::f4v_t *)b, n);
  }
  (void)hipEventRecord(e1, c->stream);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) return fvm_fail(c, FVM_EHIP, hipGetErrorString(err));
  *gbps = 2.0 * (double)(n * 16) * iters / (ms * 1e-3) / 1e9;
  return FVM_OK;
}

int fvm_measure_copy_bandwidth(fvm_ctx *c, long long bytes, int iters, double *gbps) {
  return copy_bandwidth(c, bytes, iters, gbps, false);
}

int fvm_measure_runtime_copy_bandwidth(fvm_ctx *c, long long bytes, int iters, double *gbps) {
  return copy_bandwidth(c, bytes, iters, gbps, true);
}
