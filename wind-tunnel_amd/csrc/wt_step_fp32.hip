// The FP32 single-tunnel step kernels (k_step, k_step_adv for FP32 storage) in
// a translation unit of their own, so that the build can compile them with the
// compiler's max-ILP scheduling (-mllvm -amdgpu-sched-strategy=max-ilp) while
// every other kernel keeps the default. Everything else in wt_step.hip is left
// out of this unit; the main unit declares these instantiations extern.
#define WT_STEP_FP32_UNIT
#include "wt_step.hip"
