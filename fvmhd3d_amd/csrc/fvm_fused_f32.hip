// FP32 face path of the fused stepping.  Built with -ffp-contract=fast: the
// face arithmetic (extrapolation, HLLD, limiters) uses fused multiply-add; the
// FP64 sums, pressure recovery and CFL stay FP64.
#include "fvm_fused.h"
#include "fvm_sg_step.h"

int fvm_fused_setup_f32(fvm_ctx *c) { return fvm::fused_setup_t<float>(c); }

int fvm_fused_run_f32(fvm_ctx *c, int nsteps, double dtmax) {
  return c->fz.moving ? fvm::fused_run_t<float, true>(c, nsteps, dtmax) : fvm::fused_run_t<float, false>(c, nsteps, dtmax);
}

int fvm_fused_sync_f32(fvm_ctx *c) {
  return c->prm.storage == FVM_STORAGE_SHIPPED ? fvm::fused_sync_t<float, float>(c) : fvm::fused_sync_t<float, double>(c);
}

int fvm_fused_sg_setup_f32(fvm_ctx *c) { return fvm::sg_step_setup(c); }
