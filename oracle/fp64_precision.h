/* Force-included (-include) ahead of the reference sources for the FP64
 * reference build.  It selects the double option that the reference's
 * include/PrecisionTypes.h:4-10 documents ("choose float or double") by
 * defining that header's include guard with the three typedefs set to
 * double, so the reference files are compiled unmodified.  Test
 * infrastructure only. */
#ifndef PRECISIONTYPES_H
#define PRECISIONTYPES_H
typedef double FPpart;
typedef double FPfield;
typedef double FPinterp;
#endif
