// subsample kernels reading f32 images.
#define IN_T float
#define IN_IDX 1
#include "inst_sub.inc"
