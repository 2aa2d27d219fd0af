// stage_pipelined_f32.hip -- the pipelined stage kernel for complex64 state and FP32
// arithmetic (the same source as stage_pipelined.hip, compiled for float).
#define HEOM_STAGE_PRECISION 1
#include "stage_pipelined.hip"
