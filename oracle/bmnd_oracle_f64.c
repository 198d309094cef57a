/*
 * FP64 build of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * The same source as bmnd_oracle.c with every float32 quantity held and
 * computed in float64 (the transforms, variances, shrinkage, weights and
 * aggregation sums; the deviance and risks are float64 already).  It is the
 * float64 run of the same algorithm against which the error of the float32
 * path (the oracle and the GPU, which agree bit for bit) is stated.  The
 * float32 tie grid of the matching keys is kept: it is part of the
 * reference's selection rule, not a precision choice.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define float double
#define fmaf fma
#define sqrtf sqrt
#define fabsf fabs

#include "bmnd_oracle.c"
