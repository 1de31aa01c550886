// Binned phase kernel, production arithmetic (FMA contraction allowed).
#define SPIC_NS spic_fast
#define SPIC_STRICT 0
#include "spic_binned.inl"
