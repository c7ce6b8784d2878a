// Plain FP64 path of the fused stepping.  Built with -ffp-contract=off, like
// every reference-order kernel: each operation rounds as the reference's does.
#include "fvm_fused.h"

int fvm_fused_setup_f64(fvm_ctx *c) { return fvm::fused_setup_t<double>(c); }

int fvm_fused_run_f64(fvm_ctx *c, int nsteps, double dtmax) {
  return c->fz.moving ? fvm::fused_run_t<double, true>(c, nsteps, dtmax) : fvm::fused_run_t<double, false>(c, nsteps, dtmax);
}

int fvm_fused_sync_f64(fvm_ctx *c) {
  return c->prm.storage == FVM_STORAGE_SHIPPED ? fvm::fused_sync_t<double, float>(c) : fvm::fused_sync_t<double, double>(c);
}
