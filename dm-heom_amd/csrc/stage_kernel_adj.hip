// stage_kernel_adj.hip -- the one-workgroup-per-tile stage kernel applying the adjoint
// operator L^dag (the same source as stage_kernel.hip, compiled with ADJ = true), FP64.
#define HEOM_STAGE_ADJOINT 1
#include "stage_kernel.hip"
