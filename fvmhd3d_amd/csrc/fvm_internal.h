// Internal declarations shared by the HIP translation units of libfvm_gpu.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fvm_gpu.h"
#include "../../include/fvm_surrogate.h"

namespace fvm {

__host__ __device__ inline double smin_h(double a, double b) { return (b < a) ? b : a; }

// Scratch of the device prefix scans (grown on demand, owned by a context).
struct ScanScratch {
  int *block_sums = nullptr;
  int *block_offs = nullptr;
  int *total = nullptr;
  size_t nblocks = 0;
};

// out[i] = sum_{k<i} in[k] over n ints.
int exclusive_scan(ScanScratch &s, const int *in, int *out, int n, hipStream_t st);
// same for 0/1 flags with wavefront ballots; *total receives the count (host).
int exclusive_scan_flags(ScanScratch &s, const int *flags, int *out, int n, int *total, hipStream_t st);
void scan_release(ScanScratch &s);

// every face listed once under each of its active cells (fvm_gpu.hip)
bool lists_complete(int nactive, int nface, const fvm_face *faces, const int *cell_face_off, const int *cell_face);

}  // namespace fvm

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// device-resident stepping state (fvm_fast.hip)
struct FusedState {
  bool ready = false;
  int np = 0;  // cells of the mesh the state was set up on
  int precision = 0;  // FVM_FUSED_FP64 / FVM_FUSED_FP32
  int moving = 0;
  int E = 0;
  int cur = 0;
  // internal (window-sorted) cell order: perm[i] = mesh cell of internal cell i
  DBuf perm, inv, deg, gbase, fe;
  DBuf fdr;  // FP32 path: own-side centroid offset per face-end slot (3 floats)
  DBuf fnb;  // FP32 path: the face-end records' neighbour field, compact (one int per slot)
  DBuf U, W64, cvol, mvel;  // FP64 state and geometry in internal order
  DBuf lmin;                // (3V/4pi)^(1/3) per cell (the volume is fixed on this path)
  DBuf rec[2], wc[2], tv, mm, dt_new, dtmin, tdev, hist;
  size_t slots = 0;
  // the learned HLLD surrogate in the flux step (FP32 path, fvm_sg_step.h):
  // every face once, L face-end records, face of each slot, flux records
  fvm_sg *sg = nullptr;
  int sg_nface = 0;
  DBuf sg_faceL, sg_face, sg_fid, sg_flux;
  bool timing = false;
  std::vector<hipEvent_t> ev;  // 5 per timed step
  double ms[4] = {0, 0, 0, 0};
  long long launches[4] = {0, 0, 0, 0};
};

// individual-timestep state in mesh order (fvm_substep.hip): the global
// mesh, the per-cell fields that persist between substeps, and the layout
// scratch of one substep
struct SubstepState {
  bool ready = false;
  bool complete = false;  // the global cell lists list every face under each of its cells
  int n = 0, nface = 0, nref = 0;
  DBuf gs1, gs2, gc, gn, goff, gface, gvol;  // global mesh (SoA)
  DBuf g[FVM_NFIELDS];                        // global fields, SoA [k*n + i]
  DBuf rung;                                  // int per cell
  DBuf act, apos, fmark, fpos, fl, cmark, iflag, ipos, g2l, perm, len, hist, err;
  DBuf sdti, hot, hpos, hlist, chg;  // timestep-limiter sweep: local dti, hot cells, changes
  DBuf dtig, rankg;                  // the same in mesh order (large layouts)
  // the Scheduler's order of the active list (Scheduler.h): a push stamp per
  // cell, the next stamp, and the sort of one substep's active cells
  DBuf seq, okey, okey2, oval2, otmp;
  size_t otmp_bytes = 0;
  unsigned long long next_seq = 0;
  // host side of the sweep: the global adjacency (neighbour per face-end),
  // pinned mirrors of dti and of each active cell's place in the active list
  // in mesh order (-1 elsewhere between substeps), and scratch
  std::vector<int> hoff, hnbr;
  std::vector<unsigned char> changed;
  std::vector<unsigned long long> pending;  // one bit per cell
  double *dtih = nullptr;       // pinned, n
  int *rankh = nullptr;         // pinned, n
  double *sdh = nullptr;        // pinned, n: the sweep's values per layout cell
  int *permh = nullptr;         // pinned, n: the layout's cells
  int *hoth = nullptr;          // pinned, n: the hot cells
  int *chgi = nullptr;          // pinned, n: changed cells
  double *chgv = nullptr;       // pinned, n: their values
  int *errh = nullptr;          // pinned, 2: the layout's error flag, the rung update's
  unsigned long long *posh = nullptr;  // pinned, 2: positivity count before and after the fluid update
};

struct fvm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  fvm_params prm{};
  std::string err;
  double t_global = 0.0;
  int np = 0, nactive = 0, nimport_loc = 0, nface = 0, nref = 0, nfa = -1;
  bool have_mesh = false;
  bool lists_complete = false;  // every face listed under each of its active cells
  // set while fvm_substep_run drives the entry points: they leave their final
  // synchronisation (and the fluid update its positivity read-back) to the
  // substep, which synchronises once at its end
  bool defer_sync = false;
  // FP32 face arithmetic in the face kernels of the reference-order calls
  // (fvm_set_face_precision)
  bool face32 = false;
  DBuf field[FVM_NFIELDS];
  DBuf m_s1, m_s2, m_c, m_n, m_coff, m_cface, m_cvol, m_fact, m_soff, m_sent;
  DBuf m_cn;  // per-face geometry record (centroid, n) of the same faces as m_c / m_n
  DBuf rec, swp, tmp0, tmp1, tmp2, stage;
  DBuf diag;     // unsigned long long[FVM_NDIAG]
  DBuf scalars;  // unsigned long long[8]
  fvm::ScanScratch scan;
  FusedState fz;
  SubstepState sub;
  // HIP-event timing of the reference-order kernels (fvm_kernel_timing): per
  // kernel FVM_KT_*, total milliseconds and launches since timing was enabled
  struct {
    bool on = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms[FVM_NKT] = {};
    long long launches[FVM_NKT] = {};
  } kt;
};

int fvm_fail(fvm_ctx *c, int code, const std::string &msg);
int fvm_ensure(fvm_ctx *c, DBuf &b, size_t bytes);
void fvm_release(DBuf &b);

// release of the fused-path buffers of a context (fvm_fast.hip)
void fvm_fast_release(fvm_ctx *c);
// release of the individual-timestep buffers of a context (fvm_substep.hip)
void fvm_substep_release(fvm_ctx *c);
// fused stepping per face precision (fvm_fused_f32.hip, fvm_fused_f64.hip)
int fvm_fused_setup_f32(fvm_ctx *c);
int fvm_fused_setup_f64(fvm_ctx *c);
int fvm_fused_run_f32(fvm_ctx *c, int nsteps, double dtmax);
int fvm_fused_run_f64(fvm_ctx *c, int nsteps, double dtmax);
int fvm_fused_sync_f32(fvm_ctx *c);
int fvm_fused_sync_f64(fvm_ctx *c);
// face list of the surrogate flux step of the FP32 path (fvm_fused_f32.hip)
int fvm_fused_sg_setup_f32(fvm_ctx *c);
// the surrogate's packed weights on its device (packed on demand, stream
// synchronised) and its gate {max |ln rho jump|, max |ln p jump|, max |v_n|/c_s,
// max log10 error}; the device it lives on (fvm_surrogate.hip)
int fvm_sg_device_weights(fvm_sg *s, const void **frag, const float **bias, float gate[4]);
int fvm_sg_device_of(const fvm_sg *s);

#define FVM_HIPCHK(c, call)                                                                \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fvm_fail((c), FVM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)
