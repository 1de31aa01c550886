/* FP32 instantiation of the oracle, the reference's default precision
 * (test infrastructure only). */
#include "spic_oracle.c"
