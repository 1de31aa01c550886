// Binned phase kernel in the reference's rounding order; compiled with
// -ffp-contract=off (see Makefile).
#define SPIC_NS spic_strict
#define SPIC_STRICT 1
#include "spic_binned.inl"
