// Fused FP32-face / FP64-state pipeline (device-resident stepping).
#include <hip/hip_runtime.h>

#include "../../include/fvm_gpu.h"
#include "fvm_internal.h"

void fvm_fast_release(fvm_ctx *c) { (void)c; }
