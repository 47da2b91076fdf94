// The persistent batched step (k_step_batch_persist) in a translation unit of
// its own, so that the build can compile it with machine loop-invariant code
// motion off (-mllvm -disable-machine-licm) and its thread index read through an
// empty asm: otherwise the compiler hoists the per-lane constants and addresses
// of the per-item code out of the item loop and keeps them live across it (141
// instead of 63 registers for FP32). Everything else in wt_step.hip is left out
// of this unit; the main unit declares these instantiations extern.
#define WT_STEP_PERSIST_UNIT
#include "wt_step.hip"
