// Fast arithmetic build of the tracer kernels (FMA contraction allowed).
#define BH_FAST 1
#define BH_NS fast
#include "bh_kernels.inc"
