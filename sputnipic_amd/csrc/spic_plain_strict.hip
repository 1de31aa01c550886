// Caller-order kernels in the reference's rounding order; this translation
// unit is compiled with -ffp-contract=off (see Makefile).
#define SPIC_NS spic_strict
#define SPIC_STRICT 1
#include "spic_plain.inl"
