// subsample kernels reading u8 images.
#define IN_T uint8_t
#define IN_IDX 0
#include "inst_sub.inc"
