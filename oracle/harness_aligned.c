/*
 * Oracle harness built with single precision logf / expf evaluated as a
 * rounded double precision logarithm / exponential, the evaluation the
 * device engine uses for the FP32 DCS splines (src/pumas.c:11298-11346).
 * Everything else is the unmodified reference (src/pumas.c). Comparing the
 * device with this variant of the oracle isolates the C library's float
 * rounding from the transport algorithm.
 */
#include <math.h>

static float aligned_logf(float x) { return (float)log((double)x); }
static float aligned_expf(float x) { return (float)exp((double)x); }

#define logf aligned_logf
#define expf aligned_expf

#include "harness.c"
