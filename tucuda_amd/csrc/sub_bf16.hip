// subsample kernels reading bf16 images.
#define IN_T bf16_t
#define IN_IDX 3
#include "inst_sub.inc"
