/* FP64 instantiation of the oracle (test infrastructure only). */
#define SPIC_ORACLE_FP64 1
#include "spic_oracle.c"
