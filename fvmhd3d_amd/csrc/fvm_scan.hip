// Device exclusive prefix scans over int arrays (three-pass: per-block totals,
// scan of the totals, per-block scan plus offset).  Blocks are 256 threads =
// four 64-wide wavefronts, each thread owning 4 consecutive elements; the
// wavefront-level scan uses DPP-backed shuffles, the 0/1 flag variant uses
// wavefront ballots and population counts.
#include <hip/hip_runtime.h>

#include "fvm_internal.h"

namespace fvm {

constexpr int SB = 256;           // threads per block
constexpr int SI = 4;             // elements per thread
constexpr int SPB = SB * SI;      // elements per block

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// exclusive scan of one value per thread across the block; returns the block total in *tot
__device__ __forceinline__ int block_excl_scan(int v, int *tot) {
  __shared__ int wsum[SB / 64];
  const int lane = __lane_id(), wid = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int woff = 0, total = 0;
  for (int w = 0; w < SB / 64; w++) {
    if (w < wid) woff += wsum[w];
    total += wsum[w];
  }
  __syncthreads();
  *tot = total;
  return woff + inc - v;
}

__global__ void __launch_bounds__(SB) k_scan_reduce(const int *__restrict__ in, int n, int *__restrict__ block_sums) {
  const long base = (long)blockIdx.x * SPB + (long)threadIdx.x * SI;
  int s = 0;
  for (int k = 0; k < SI; k++)
    if (base + k < n) s += in[base + k];
  int tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// one block scans all block totals, in chunks of SB * BSI: each thread sums
// BSI consecutive totals, the block scans the thread sums, and each thread
// writes its BSI offsets
constexpr int BSI = 16;
__global__ void __launch_bounds__(SB) k_scan_blocksums(const int *__restrict__ sums, int nb, int *__restrict__ offs,
                                           int *__restrict__ total) {
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += SB * BSI) {
    const int b = b0 + (int)threadIdx.x * BSI;
    int v[BSI];
    int s = 0;
#pragma unroll
    for (int k = 0; k < BSI; k++) {
      v[k] = b + k < nb ? sums[b + k] : 0;
      s += v[k];
    }
    int tot;
    int run = carry + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < BSI; k++) {
      if (b + k < nb) offs[b + k] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(SB) k_scan_final(const int *__restrict__ in, int n, const int *__restrict__ offs,
                                       int *__restrict__ out) {
  const long base = (long)blockIdx.x * SPB + (long)threadIdx.x * SI;
  int v[SI];
  int s = 0;
  for (int k = 0; k < SI; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int run = block_excl_scan(s, &tot) + offs[blockIdx.x];
  for (int k = 0; k < SI; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// flags: element e of a block is handled in round r = e / SB by thread e % SB,
// so each round is one ballot per wavefront
__global__ void __launch_bounds__(SB) k_flag_reduce(const int *__restrict__ flags, int n, int *__restrict__ block_sums) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int r = 0; r < SI; r++) {
    const long e = (long)blockIdx.x * SPB + (long)r * SB + threadIdx.x;
    const bool f = e < n && flags[e] != 0;
    const unsigned long long m = __ballot(f);
    if (__lane_id() == 0) c += __popcll(m);
  }
  if (__lane_id() == 0) atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) block_sums[blockIdx.x] = cnt;
}

__global__ void __launch_bounds__(SB) k_flag_final(const int *__restrict__ flags, int n, const int *__restrict__ offs,
                                       int *__restrict__ out) {
  __shared__ int wcnt[SB / 64];
  const int lane = __lane_id(), wid = threadIdx.x >> 6;
  int run = offs[blockIdx.x];
  const unsigned long long lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
  for (int r = 0; r < SI; r++) {
    const long e = (long)blockIdx.x * SPB + (long)r * SB + threadIdx.x;
    const bool f = e < n && flags[e] != 0;
    const unsigned long long m = __ballot(f);
    if (lane == 0) wcnt[wid] = __popcll(m);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < SB / 64; w++) {
      if (w < wid) woff += wcnt[w];
      tot += wcnt[w];
    }
    if (e < n) out[e] = run + woff + __popcll(m & lt);
    run += tot;
    __syncthreads();
  }
}

static int grow(ScanScratch &s, size_t nb) {
  if (s.nblocks >= nb && s.block_sums) return 0;
  scan_release(s);
  if (hipMalloc(&s.block_sums, sizeof(int) * nb) != hipSuccess) return -1;
  if (hipMalloc(&s.block_offs, sizeof(int) * nb) != hipSuccess) return -1;
  if (hipMalloc(&s.total, sizeof(int) * 2) != hipSuccess) return -1;
  s.nblocks = nb;
  return 0;
}

void scan_release(ScanScratch &s) {
  if (s.block_sums) (void)hipFree(s.block_sums);
  if (s.block_offs) (void)hipFree(s.block_offs);
  if (s.total) (void)hipFree(s.total);
  s.block_sums = s.block_offs = s.total = nullptr;
  s.nblocks = 0;
}

int exclusive_scan(ScanScratch &s, const int *in, int *out, int n, hipStream_t st) {
  if (n <= 0) return 0;
  const size_t nb = ((size_t)n + SPB - 1) / SPB;
  if (grow(s, nb)) return -1;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SB), 0, st, in, n, s.block_sums);
  hipLaunchKernelGGL(k_scan_blocksums, dim3(1), dim3(SB), 0, st, s.block_sums, (int)nb, s.block_offs, s.total);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(SB), 0, st, in, n, s.block_offs, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int exclusive_scan_flags(ScanScratch &s, const int *flags, int *out, int n, int *total, hipStream_t st) {
  if (n <= 0) {
    if (total) *total = 0;
    return 0;
  }
  const size_t nb = ((size_t)n + SPB - 1) / SPB;
  if (grow(s, nb)) return -1;
  hipLaunchKernelGGL(k_flag_reduce, dim3(nb), dim3(SB), 0, st, flags, n, s.block_sums);
  hipLaunchKernelGGL(k_scan_blocksums, dim3(1), dim3(SB), 0, st, s.block_sums, (int)nb, s.block_offs, s.total);
  hipLaunchKernelGGL(k_flag_final, dim3(nb), dim3(SB), 0, st, flags, n, s.block_offs, out);
  if (hipGetLastError() != hipSuccess) return -1;
  if (total) {
    if (hipMemcpyAsync(total, s.total, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
  }
  return 0;
}

}  // namespace fvm
