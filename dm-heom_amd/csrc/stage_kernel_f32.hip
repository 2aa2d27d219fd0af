// stage_kernel_f32.hip -- the one-workgroup-per-tile stage kernel for complex64 state
// and FP32 arithmetic (the same source as stage_kernel.hip, compiled for float).
#define HEOM_STAGE_PRECISION 1
#include "stage_kernel.hip"
