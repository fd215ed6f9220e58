// subsample kernels reading f16 images.
#define IN_T f16_t
#define IN_IDX 2
#include "inst_sub.inc"
