/*
 * TEST INFRASTRUCTURE ONLY — instantiates the oracle FFT for float and double.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "offt.h"

#define OFFT_R float
#define OFFT_S f
#define OFFT_CMPLX(a, b) CMPLXF((a), (b))
#define OFFT_CREAL crealf
#define OFFT_CIMAG cimagf
#include "offt_impl.h"
#undef OFFT_R
#undef OFFT_S
#undef OFFT_CMPLX
#undef OFFT_CREAL
#undef OFFT_CIMAG

#define OFFT_R double
#define OFFT_S d
#define OFFT_CMPLX(a, b) CMPLX((a), (b))
#define OFFT_CREAL creal
#define OFFT_CIMAG cimag
#include "offt_impl.h"
