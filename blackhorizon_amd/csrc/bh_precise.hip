// Precise arithmetic build of the tracer kernels: compiled with
// -ffp-contract=off so each expression rounds once per operation, in the
// order NumPy evaluates the reference's array expressions.
#define BH_FAST 0
#define BH_NS precise
#include "bh_kernels.inc"
