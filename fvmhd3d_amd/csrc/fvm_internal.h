// Internal declarations shared by the HIP translation units of libfvm_gpu.
#pragma once
#include <hip/hip_runtime.h>

struct fvm_ctx;

namespace fvm {

__host__ __device__ inline double smin_h(double a, double b) { return (b < a) ? b : a; }

// Scratch of the device prefix scans (grown on demand, owned by a context).
struct ScanScratch {
  int *block_sums = nullptr;
  int *block_offs = nullptr;
  int *total = nullptr;
  size_t nblocks = 0;
};

// out[i] = sum_{k<i} in[i] over n ints; out may alias nothing in in.
int exclusive_scan(ScanScratch &s, const int *in, int *out, int n, hipStream_t st);
// same for 0/1 flags with wavefront ballots; *total receives the count (host).
int exclusive_scan_flags(ScanScratch &s, const int *flags, int *out, int n, int *total, hipStream_t st);
void scan_release(ScanScratch &s);

}  // namespace fvm

// release of the fused-path buffers of a context (fvm_fast.hip)
void fvm_fast_release(fvm_ctx *c);
