/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle's FFT (stands in for FFTW, which the
 * reference links in src/FFTapps.c:2 and which is absent here).  Never linked
 * into the product library.
 */
#ifndef ORACLE_OFFT_H
#define ORACLE_OFFT_H
#include <complex.h>
#include <stddef.h>

typedef float complex offt_cpx_f;
typedef double complex offt_cpx_d;

#define OFFT_PLAN_FIELDS(CPX, PT)                                                   \
    int n, inverse, blue, m;                                                         \
    int factors[64];                                                                 \
    CPX *tw;                                                                         \
    struct PT *sub_fwd, *sub_inv;                                                    \
    CPX *chirp, *bfilt, *work;

typedef struct offt_plan_f { OFFT_PLAN_FIELDS(offt_cpx_f, offt_plan_f) } offt_plan_f;
typedef struct offt_plan_d { OFFT_PLAN_FIELDS(offt_cpx_d, offt_plan_d) } offt_plan_d;

offt_plan_f *offt_make_f(int n, int inverse);
offt_plan_d *offt_make_d(int n, int inverse);
void offt_destroy_f(offt_plan_f *p);
void offt_destroy_d(offt_plan_d *p);
/* in-place allowed; allocates a temporary when needed (not for hot loops) */
void offt_execute_f(const offt_plan_f *p, const offt_cpx_f *in, offt_cpx_f *out);
void offt_execute_d(const offt_plan_d *p, const offt_cpx_d *in, offt_cpx_d *out);
/* thread-safe, allocation-free: ws must hold offt_scratch_len_*(p) elements */
size_t offt_scratch_len_f(const offt_plan_f *p);
size_t offt_scratch_len_d(const offt_plan_d *p);
void offt_execute_ws_f(const offt_plan_f *p, const offt_cpx_f *in, offt_cpx_f *out, offt_cpx_f *ws);
void offt_execute_ws_d(const offt_plan_d *p, const offt_cpx_d *in, offt_cpx_d *out, offt_cpx_d *ws);

#endif
