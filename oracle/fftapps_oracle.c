/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle for the interstation-correlation path.
 *
 * Restates, step by step, the CPU branches of src/FFTapps.c (sergiventosa/FastPCC):
 * the analytic signal, the two phasor regularisers, PCC2, PCC1, PCC-v, GNCC,
 * 1-bit GNCC and WPCC, each with the reference's own precision and rounding
 * points (FP32 where it calls fftwf_*, FP64 where it calls fftw_*).  FFTW is
 * replaced by the oracle's FFT (offt.c).  OpenMP runs over the windows (Tr) of
 * one pair, as in the reference.  Only tests/, the smoke check and bench.py's
 * CPU-baseline leg may load this library; the product never does.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>
#include "offt.h"
#include "oracle.h"

#ifndef PI
#define PI 3.14159265358979323846
#endif

int orc_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------- */
/* analytic signal                                                           */
/* ------------------------------------------------------------------------- */

/* AnalyticSignal_plan_float, src/FFTapps.c:104-124.  y = IDFT_N(mask . DFT_N(x)),
 * mask = 1/N at DC, 2/N on 0<k<ceil(N/2), 1/N at Nyquist (even N), 0 elsewhere;
 * da1 is a float holding 1/(double)N and 2*da1 is formed in float. */
static void analytic_f_plan(float complex *y, const float *x, unsigned N,
                            const offt_plan_f *fwd, const offt_plan_f *bwd, float complex *ws)
{
    for (unsigned n = 0; n < N; ++n) y[n] = CMPLXF(x[n], 0.0f);
    offt_execute_ws_f(fwd, y, y, ws);
    float da1 = 1 / (double)N;
    unsigned n;
    y[0] *= da1;
    for (n = 1; n < (N + 1) >> 1; n++) y[n] *= 2 * da1;
    if (!(N & 1)) y[n++] *= da1;
    for (; n < N; n++) y[n] = 0;
    offt_execute_ws_f(bwd, y, y, ws);
}

/* AnalyticSignal, src/FFTapps.c:45-80 (FP64 twin, same mask). */
static void analytic_d_plan(double complex *y, const double *x, unsigned N,
                            const offt_plan_d *fwd, const offt_plan_d *bwd, double complex *ws)
{
    for (unsigned n = 0; n < N; ++n) y[n] = CMPLX(x[n], 0.0);
    offt_execute_ws_d(fwd, y, y, ws);
    double da1 = 1 / (double)N;
    unsigned n;
    y[0] *= da1;
    for (n = 1; n < (N + 1) >> 1; n++) y[n] *= 2 * da1;
    if (!(N & 1)) y[n++] *= da1;
    for (; n < N; n++) y[n] = 0;
    offt_execute_ws_d(bwd, y, y, ws);
}

int orc_analytic_f(float complex *y, const float *x, unsigned N)
{
    if (!x || !y || !N) return -1;
    offt_plan_f *f = offt_make_f(N, 0), *b = offt_make_f(N, 1);
    size_t wl = offt_scratch_len_f(f) > offt_scratch_len_f(b) ? offt_scratch_len_f(f) : offt_scratch_len_f(b);
    float complex *ws = malloc(sizeof(float complex) * wl);
    analytic_f_plan(y, x, N, f, b, ws);
    free(ws);
    offt_destroy_f(f);
    offt_destroy_f(b);
    return 0;
}

int orc_analytic_d(double complex *y, const double *x, unsigned N)
{
    if (!x || !y || !N) return -1;
    offt_plan_d *f = offt_make_d(N, 0), *b = offt_make_d(N, 1);
    size_t wl = offt_scratch_len_d(f) > offt_scratch_len_d(b) ? offt_scratch_len_d(f) : offt_scratch_len_d(b);
    double complex *ws = malloc(sizeof(double complex) * wl);
    analytic_d_plan(y, x, N, f, b, ws);
    free(ws);
    offt_destroy_d(f);
    offt_destroy_d(b);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* phasor regularisers                                                       */
/* ------------------------------------------------------------------------- */

/* AmpNorm (FP64), src/FFTapps.c:147-175: eps = 1e-6*max|z|^2, z /= sqrt(|z|^2+eps),
 * zero where the reciprocal is not a normal number. */
void orc_ampnorm(double complex *y, unsigned N)
{
    double *pd1 = (double *)y, eps = 0, da1, da2, da3;
    for (unsigned n = 0; n < 2 * N; n += 2) {
        da1 = pd1[n];
        da2 = pd1[n + 1];
        da3 = da1 * da1 + da2 * da2;
        if (eps < da3) eps = da3;
    }
    eps *= 1e-6;
    for (unsigned n = 0; n < 2 * N; n += 2) {
        da1 = pd1[n];
        da2 = pd1[n + 1];
        da3 = 1 / sqrt(da1 * da1 + da2 * da2 + eps);
        if (isnormal(da3)) {
            pd1[n] *= da3;
            pd1[n + 1] *= da3;
        } else {
            pd1[n] = 0;
            pd1[n + 1] = 0;
        }
    }
}

/* AmpNormf (FP32), src/FFTapps.c:178-205: eps = 1e-6*sqrt(max|z|^2) (the sqrt and
 * the product are evaluated in double, then stored as float), z *= 1/(|z|+eps). */
void orc_ampnormf(float complex *y, unsigned N)
{
    float *pf1 = (float *)y, eps = 0, fa1, fa2, fa3;
    for (unsigned n = 0; n < 2 * N; n += 2) {
        fa1 = pf1[n];
        fa2 = pf1[n + 1];
        fa3 = fa1 * fa1 + fa2 * fa2;
        if (eps < fa3) eps = fa3;
    }
    eps = 1e-6 * sqrt(eps);
    for (unsigned n = 0; n < 2 * N; n += 2) {
        fa1 = pf1[n];
        fa2 = pf1[n + 1];
        fa3 = 1 / (sqrtf(fa1 * fa1 + fa2 * fa2) + eps);
        if (isnormal(fa3)) {
            pf1[n] *= fa3;
            pf1[n + 1] *= fa3;
        } else {
            pf1[n] = 0;
            pf1[n + 1] = 0;
        }
    }
}

/* AmpNormf's formula with every operation in double (all-FP64 oracle). */
void orc_ampnormf_d(double complex *y, unsigned N)
{
    double *pd1 = (double *)y, eps = 0, da1, da2, da3;
    for (unsigned n = 0; n < 2 * N; n += 2) {
        da1 = pd1[n];
        da2 = pd1[n + 1];
        da3 = da1 * da1 + da2 * da2;
        if (eps < da3) eps = da3;
    }
    eps = 1e-6 * sqrt(eps);
    for (unsigned n = 0; n < 2 * N; n += 2) {
        da1 = pd1[n];
        da2 = pd1[n + 1];
        da3 = 1 / (sqrt(da1 * da1 + da2 * da2) + eps);
        if (isnormal(da3)) {
            pd1[n] *= da3;
            pd1[n + 1] *= da3;
        } else {
            pd1[n] = 0;
            pd1[n + 1] = 0;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* helpers                                                                   */
/* ------------------------------------------------------------------------- */

static unsigned nz_rule(unsigned N, unsigned M)
{
    /* src/FFTapps.c:805: Nz = 1 << ceil(log2(N+M)) */
    return 1u << (unsigned)ceil(log2((double)(N + M)));
}

static size_t maxz(size_t a, size_t b) { return a > b ? a : b; }

/* real -> half spectrum (Nz/2+1 bins) via a complex FFT; full is scratch of Nz */
static void r2c_f(float complex *half, const float *in, unsigned Nz, const offt_plan_f *fwd,
                  float complex *full, float complex *ws)
{
    for (unsigned n = 0; n < Nz; ++n) full[n] = CMPLXF(in[n], 0.0f);
    offt_execute_ws_f(fwd, full, full, ws);
    memcpy(half, full, sizeof(float complex) * (Nz / 2 + 1));
}

static void r2c_d(double complex *half, const double *in, unsigned Nz, const offt_plan_d *fwd,
                  double complex *full, double complex *ws)
{
    for (unsigned n = 0; n < Nz; ++n) full[n] = CMPLX(in[n], 0.0);
    offt_execute_ws_d(fwd, full, full, ws);
    memcpy(half, full, sizeof(double complex) * (Nz / 2 + 1));
}

/* half spectrum -> real, unnormalised (FFTW c2r: imaginary parts of DC and
 * Nyquist ignored, Hermitian extension implied) */
static void c2r_f(float *out, const float complex *half, unsigned Nz, const offt_plan_f *bwd,
                  float complex *full, float complex *ws)
{
    full[0] = CMPLXF(crealf(half[0]), 0.0f);
    full[Nz / 2] = CMPLXF(crealf(half[Nz / 2]), 0.0f);
    for (unsigned k = 1; k < Nz / 2; ++k) {
        full[k] = half[k];
        full[Nz - k] = conjf(half[k]);
    }
    offt_execute_ws_f(bwd, full, full, ws);
    for (unsigned n = 0; n < Nz; ++n) out[n] = crealf(full[n]);
}

static void c2r_d(double *out, const double complex *half, unsigned Nz, const offt_plan_d *bwd,
                  double complex *full, double complex *ws)
{
    full[0] = CMPLX(creal(half[0]), 0.0);
    full[Nz / 2] = CMPLX(creal(half[Nz / 2]), 0.0);
    for (unsigned k = 1; k < Nz / 2; ++k) {
        full[k] = half[k];
        full[Nz - k] = conj(half[k]);
    }
    offt_execute_ws_d(bwd, full, full, ws);
    for (unsigned n = 0; n < Nz; ++n) out[n] = creal(full[n]);
}

/* ------------------------------------------------------------------------- */
/* time-domain PCC lowlevels                                                 */
/* ------------------------------------------------------------------------- */

/* pcc1f_lowlevel, src/FFTapps.c:251-278.  Called as (y, fxa2, fxa1): a = phi2[n+lag],
 * b = phi1[n] over the overlap; FP32 terms, double accumulator, (float)sum/(2*count). */
static void pcc1f_lowlevel(float *y, const float complex *xan1, const float complex *xan2,
                           int N, int Lag1, int Lag2)
{
    int L = Lag2 - Lag1 + 1, lag;
    unsigned n, n1, n2, l, l1;
    if (Lag2 > N) L -= (Lag2 - N);
    l1 = (Lag1 >= -N) ? 0 : -(Lag1 + N);
    for (l = l1; (int)l < L; l++) {
        lag = Lag1 + (int)l;
        if (lag < 0) { n1 = -lag; n2 = N; }
        else         { n1 = 0;    n2 = N - lag; }
        const float *pf1 = (const float *)(xan1 + n1 + lag);
        const float *pf2 = (const float *)(xan2 + n1);
        double da1 = 0;
        float fa2[2], fa3[2];
        for (n = 0; n < 2 * (n2 - n1); n += 2) {
            fa2[0] = pf1[n] + pf2[n];
            fa2[1] = pf1[n + 1] + pf2[n + 1];
            fa3[0] = pf1[n] - pf2[n];
            fa3[1] = pf1[n + 1] - pf2[n + 1];
            da1 += sqrtf(fa2[0] * fa2[0] + fa2[1] * fa2[1]) - sqrtf(fa3[0] * fa3[0] + fa3[1] * fa3[1]);
        }
        y[l] = (float)da1 / (2 * (n2 - n1));
    }
}

/* pccf_lowlevel, src/FFTapps.c:309-336: terms powf(|a+-b|^2, v/2) in FP32, double
 * accumulator, divided by pow(2, v/2)*count in double. */
static void pccf_lowlevel(float *y, const float complex *xan1, const float complex *xan2,
                          int N, double v, int Lag1, int Lag2)
{
    float V = v / 2;
    int L = Lag2 - Lag1 + 1, lag;
    unsigned n, n1, n2, l, l1;
    if (Lag2 > N) L -= (Lag2 - N);
    l1 = (Lag1 >= -N) ? 0 : -(Lag1 + N);
    for (l = l1; (int)l < L; l++) {
        lag = Lag1 + (int)l;
        if (lag < 0) { n1 = -lag; n2 = N; }
        else         { n1 = 0;    n2 = N - lag; }
        const float *pf1 = (const float *)(xan1 + n1 + lag);
        const float *pf2 = (const float *)(xan2 + n1);
        double da1 = 0;
        float fa2[2], fa3[2];
        for (n = 0; n < 2 * (n2 - n1); n += 2) {
            fa2[0] = pf1[n] + pf2[n];
            fa2[1] = pf1[n + 1] + pf2[n + 1];
            fa3[0] = pf1[n] - pf2[n];
            fa3[1] = pf1[n + 1] - pf2[n + 1];
            da1 += powf(fa2[0] * fa2[0] + fa2[1] * fa2[1], V) - powf(fa3[0] * fa3[0] + fa3[1] * fa3[1], V);
        }
        y[l] = da1 / (pow(2, V) * (n2 - n1));
    }
}

/* FP64 twin of both lowlevels (all-FP64 oracle): v == 1 gives PCC1's divisor
 * 2*count (src/FFTapps.c:276), otherwise pow(2, v/2)*count (:334). */
static void pcc_lowlevel_d(double *y, const double complex *xan1, const double complex *xan2,
                           int N, double v, int pcc1, int Lag1, int Lag2)
{
    double V = v / 2;
    int L = Lag2 - Lag1 + 1, lag;
    unsigned n, n1, n2, l, l1;
    if (Lag2 > N) L -= (Lag2 - N);
    l1 = (Lag1 >= -N) ? 0 : -(Lag1 + N);
    for (l = l1; (int)l < L; l++) {
        lag = Lag1 + (int)l;
        if (lag < 0) { n1 = -lag; n2 = N; }
        else         { n1 = 0;    n2 = N - lag; }
        const double complex *a = xan1 + n1 + lag, *b = xan2 + n1;
        double acc = 0;
        for (n = 0; n < n2 - n1; ++n) {
            double p = cabs(a[n] + b[n]), m = cabs(a[n] - b[n]);
            acc += pcc1 ? (p - m) : (pow(p * p, V) - pow(m * m, V));
        }
        y[l] = pcc1 ? acc / (2.0 * (n2 - n1)) : acc / (pow(2, V) * (n2 - n1));
    }
}

/* ------------------------------------------------------------------------- */
/* GN normalisation (src/FFTapps.c:372-429)                                  */
/* ------------------------------------------------------------------------- */

/* gn_lowlevel: divides y[n] (lags lag..lag+L-1) by sqrt(E1*E2), the overlap
 * energies maintained incrementally in double. */
static void gn_lowlevel_d(double *y, const double *x1, const double *x2, double norm1, double norm2,
                          unsigned N, unsigned L, int lag)
{
    unsigned n, n11 = 0, n21 = 0, n12 = N, n22 = N;
    if (lag > 0) { n21 = (unsigned)lag; n12 -= (unsigned)lag; }
    else         { n11 = (unsigned)(-lag); n22 -= (unsigned)(-lag); }
    for (n = 0; n < L - 1 && n22 < N; n++) {
        y[n] /= sqrt(norm1 * norm2);
        n11--;
        norm1 += x1[n11] * x1[n11];
        norm2 += x2[n22] * x2[n22];
        n22++;
    }
    for (; n < L - 1; n++) {
        y[n] /= sqrt(norm1 * norm2);
        n12--;
        norm1 -= x1[n12] * x1[n12];
        norm2 -= x2[n21] * x2[n21];
        n21++;
    }
    y[n] /= sqrt(norm1 * norm2);
}

/* gnf_lowlevel: the same on float y and float x (energies still in double). */
static void gn_lowlevel_f(float *y, const float *x1, const float *x2, double norm1, double norm2,
                          unsigned N, unsigned L, int lag)
{
    unsigned n, n11 = 0, n21 = 0, n12 = N, n22 = N;
    if (lag > 0) { n21 = (unsigned)lag; n12 -= (unsigned)lag; }
    else         { n11 = (unsigned)(-lag); n22 -= (unsigned)(-lag); }
    for (n = 0; n < L - 1 && n22 < N; n++) {
        y[n] /= sqrt(norm1 * norm2);
        n11--;
        norm1 += x1[n11] * x1[n11];
        norm2 += x2[n22] * x2[n22];
        n22++;
    }
    for (; n < L - 1; n++) {
        y[n] /= sqrt(norm1 * norm2);
        n12--;
        norm1 -= x1[n12] * x1[n12];
        norm2 -= x2[n21] * x2[n21];
        n21++;
    }
    y[n] /= sqrt(norm1 * norm2);
}

/* Norm / Normf, src/FFTapps.c:207-221 */
static double norm_d(const double *x, unsigned n1, unsigned n2)
{
    double s = 0;
    for (unsigned n = n1; n < n2; n++) s += x[n] * x[n];
    return s;
}
static double norm_f(const float *x, unsigned n1, unsigned n2)
{
    double s = 0;
    for (unsigned n = n1; n < n2; n++) s += x[n] * x[n];
    return s;
}

/* ------------------------------------------------------------------------- */
/* PCC2 (src/FFTapps.c:785-881)                                              */
/* ------------------------------------------------------------------------- */

/* Lag window of L values starting at lag.  The reference copies the window's
 * negative lags with `for (n=0; n<-lag; n++)` (src/FFTapps.c:350, :367 and the
 * same loop in pcc2_set and tspcc2_set), which writes past y's L values when the
 * whole window lies below -L; every copy here also stops at n < L.  The L values
 * written are the reference's. */
static int lag_setup(int Lag1, int Lag2, int *L, int *lag)
{
    if (Lag2 >= Lag1) { *L = Lag2 - Lag1 + 1; *lag = Lag1; }
    else              { *L = Lag1 - Lag2 + 1; *lag = Lag2; }
    return 0;
}

int orc_pcc2_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2)
{
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = nz_rule(N, M);
    offt_plan_f *fN = offt_make_f(N, 0), *bN = offt_make_f(N, 1);
    offt_plan_f *fZ = offt_make_f(Nz, 0), *bZ = offt_make_f(Nz, 1);
    size_t wl = maxz(maxz(offt_scratch_len_f(fN), offt_scratch_len_f(bN)),
                     maxz(offt_scratch_len_f(fZ), offt_scratch_len_f(bZ)));
#pragma omp parallel
    {
        float complex *xa1 = malloc(sizeof(float complex) * Nz);
        float complex *xa2 = malloc(sizeof(float complex) * Nz);
        float complex *out = malloc(sizeof(float complex) * Nz);
        float complex *ws = malloc(sizeof(float complex) * wl);
        float fa1 = 1 / ((float)Nz * (float)N);
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            analytic_f_plan(xa1, x1[tr], N, fN, bN, ws);
            orc_ampnormf(xa1, N);
            for (unsigned n = N; n < Nz; n++) xa1[n] = 0;
            offt_execute_ws_f(fZ, xa1, xa1, ws);

            analytic_f_plan(xa2, x2[tr], N, fN, bN, ws);
            orc_ampnormf(xa2, N);
            for (unsigned n = N; n < Nz; n++) xa2[n] = 0;
            offt_execute_ws_f(fZ, xa2, xa2, ws);

            /* conj() is the double-complex one in the reference (:851): the
             * product is formed in double and rounded to float */
            for (unsigned n = 0; n < Nz; n++)
                out[n] = (float complex)(conj((double complex)xa1[n]) * (double complex)xa2[n]);
            offt_execute_ws_f(bZ, out, out, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) y[tr][n] = crealf(fa1 * out[n + Nz + lag]);
            for (; n < L; n++) y[tr][n] = crealf(fa1 * out[n + lag]);
        }
        free(xa1); free(xa2); free(out); free(ws);
    }
    offt_destroy_f(fN); offt_destroy_f(bN); offt_destroy_f(fZ); offt_destroy_f(bZ);
    return 0;
}

/* all-FP64 PCC2: AnalyticSignal (FP64) + AmpNormf's formula in FP64 + FP64 FFTs */
static int orc64_pcc2(double **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2)
{
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = nz_rule(N, M);
    offt_plan_d *fN = offt_make_d(N, 0), *bN = offt_make_d(N, 1);
    offt_plan_d *fZ = offt_make_d(Nz, 0), *bZ = offt_make_d(Nz, 1);
    size_t wl = maxz(maxz(offt_scratch_len_d(fN), offt_scratch_len_d(bN)),
                     maxz(offt_scratch_len_d(fZ), offt_scratch_len_d(bZ)));
#pragma omp parallel
    {
        double *xd = malloc(sizeof(double) * N);
        double complex *xa1 = malloc(sizeof(double complex) * Nz);
        double complex *xa2 = malloc(sizeof(double complex) * Nz);
        double complex *out = malloc(sizeof(double complex) * Nz);
        double complex *ws = malloc(sizeof(double complex) * wl);
        double fa1 = 1.0 / ((double)Nz * (double)N);
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            for (unsigned n = 0; n < N; ++n) xd[n] = x1[tr][n];
            analytic_d_plan(xa1, xd, N, fN, bN, ws);
            orc_ampnormf_d(xa1, N);
            for (unsigned n = N; n < Nz; n++) xa1[n] = 0;
            offt_execute_ws_d(fZ, xa1, xa1, ws);
            for (unsigned n = 0; n < N; ++n) xd[n] = x2[tr][n];
            analytic_d_plan(xa2, xd, N, fN, bN, ws);
            orc_ampnormf_d(xa2, N);
            for (unsigned n = N; n < Nz; n++) xa2[n] = 0;
            offt_execute_ws_d(fZ, xa2, xa2, ws);
            for (unsigned n = 0; n < Nz; n++) out[n] = conj(xa1[n]) * xa2[n];
            offt_execute_ws_d(bZ, out, out, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) y[tr][n] = fa1 * creal(out[n + Nz + lag]);
            for (; n < L; n++) y[tr][n] = fa1 * creal(out[n + lag]);
        }
        free(xd); free(xa1); free(xa2); free(out); free(ws);
    }
    offt_destroy_d(fN); offt_destroy_d(bN); offt_destroy_d(fZ); offt_destroy_d(bZ);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* PCC1 / PCC-v (src/FFTapps.c:475-608 and 611-783, CPU branches)            */
/* ------------------------------------------------------------------------- */

static int pcc_time_f(float **y, float **x1, float **x2, int N, unsigned Tr, double v, int pcc1,
                      int Lag1, int Lag2)
{
    int L = Lag2 - Lag1 + 1;
    if (Lag1 > N || Lag2 < -N) return 0;
    if (x1 == NULL || x2 == NULL || L < 0) return -1;
    float complex *fxa1 = malloc(sizeof(float complex) * (size_t)Tr * N);
    float complex *fxa2 = malloc(sizeof(float complex) * (size_t)Tr * N);
    offt_plan_f *fN = offt_make_f(N, 0), *bN = offt_make_f(N, 1);
    size_t wl = maxz(offt_scratch_len_f(fN), offt_scratch_len_f(bN));
#pragma omp parallel
    {
        float complex *ws = malloc(sizeof(float complex) * wl);
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            analytic_f_plan(fxa1 + (size_t)tr * N, x1[tr], N, fN, bN, ws);
            analytic_f_plan(fxa2 + (size_t)tr * N, x2[tr], N, fN, bN, ws);
        }
        free(ws);
    }
#pragma omp parallel for schedule(static)
    for (unsigned tr = 0; tr < Tr; tr++) {
        orc_ampnormf(fxa1 + (size_t)tr * N, N);
        orc_ampnormf(fxa2 + (size_t)tr * N, N);
        memset(y[tr], 0, L * sizeof(float));
        if (pcc1) pcc1f_lowlevel(y[tr], fxa2 + (size_t)tr * N, fxa1 + (size_t)tr * N, N, Lag1, Lag2);
        else pccf_lowlevel(y[tr], fxa2 + (size_t)tr * N, fxa1 + (size_t)tr * N, N, v, Lag1, Lag2);
    }
    offt_destroy_f(fN); offt_destroy_f(bN);
    free(fxa1); free(fxa2);
    return 0;
}

int orc_pcc1_set(float **y, float **x1, float **x2, int N, unsigned Tr, int Lag1, int Lag2)
{
    return pcc_time_f(y, x1, x2, N, Tr, 1.0, 1, Lag1, Lag2);
}

int orc_pcc_set(float **y, float **x1, float **x2, int N, unsigned Tr, double v, int Lag1, int Lag2)
{
    return pcc_time_f(y, x1, x2, N, Tr, v, 0, Lag1, Lag2);
}

static int orc64_pcc_time(double **y, float **x1, float **x2, int N, unsigned Tr, double v, int pcc1,
                          int Lag1, int Lag2)
{
    int L = Lag2 - Lag1 + 1;
    if (Lag1 > N || Lag2 < -N) return 0;
    if (x1 == NULL || x2 == NULL || L < 0) return -1;
    offt_plan_d *fN = offt_make_d(N, 0), *bN = offt_make_d(N, 1);
    size_t wl = maxz(offt_scratch_len_d(fN), offt_scratch_len_d(bN));
#pragma omp parallel
    {
        double *xd = malloc(sizeof(double) * N);
        double complex *a1 = malloc(sizeof(double complex) * N), *a2 = malloc(sizeof(double complex) * N);
        double complex *ws = malloc(sizeof(double complex) * wl);
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            for (int n = 0; n < N; ++n) xd[n] = x1[tr][n];
            analytic_d_plan(a1, xd, N, fN, bN, ws);
            orc_ampnormf_d(a1, N);
            for (int n = 0; n < N; ++n) xd[n] = x2[tr][n];
            analytic_d_plan(a2, xd, N, fN, bN, ws);
            orc_ampnormf_d(a2, N);
            memset(y[tr], 0, L * sizeof(double));
            pcc_lowlevel_d(y[tr], a2, a1, N, v, pcc1, Lag1, Lag2);
        }
        free(xd); free(a1); free(a2); free(ws);
    }
    offt_destroy_d(fN); offt_destroy_d(bN);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* GNCC (src/FFTapps.c:883-978), FP64 throughout                             */
/* ------------------------------------------------------------------------- */

static int ccgn_core(float **yf, double **yd, float **x1, float **x2, unsigned N, unsigned Tr,
                     int Lag1, int Lag2)
{
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned n11 = 0, n21 = 0, n12 = N, n22 = N;
    if (lag > 0) { n21 = (unsigned)lag; n12 -= (unsigned)lag; }
    else         { n11 = (unsigned)(-lag); n22 -= (unsigned)(-lag); }
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = nz_rule(N, M), Nh = Nz / 2 + 1;
    offt_plan_d *fZ = offt_make_d(Nz, 0), *bZ = offt_make_d(Nz, 1);
    size_t wl = maxz(offt_scratch_len_d(fZ), offt_scratch_len_d(bZ));
#pragma omp parallel
    {
        double *in1 = malloc(sizeof(double) * Nz), *in2 = malloc(sizeof(double) * Nz);
        double *out = malloc(sizeof(double) * Nz), *ybuf = malloc(sizeof(double) * L);
        double complex *fin1 = malloc(sizeof(double complex) * Nh), *fin2 = malloc(sizeof(double complex) * Nh);
        double complex *fout = malloc(sizeof(double complex) * Nh);
        double complex *full = malloc(sizeof(double complex) * Nz), *ws = malloc(sizeof(double complex) * wl);
        const double da1 = 1. / (double)Nz;
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            for (unsigned n = 0; n < N; n++) in1[n] = (double)x1[tr][n];
            for (unsigned n = N; n < Nz; n++) in1[n] = 0;
            r2c_d(fin1, in1, Nz, fZ, full, ws);
            for (unsigned n = 0; n < N; n++) in2[n] = (double)x2[tr][n];
            for (unsigned n = N; n < Nz; n++) in2[n] = 0;
            r2c_d(fin2, in2, Nz, fZ, full, ws);
            /* cc_lowlevel, :338-353 */
            for (unsigned n = 0; n < Nh; n++) fout[n] = conj(fin1[n]) * fin2[n];
            c2r_d(out, fout, Nz, bZ, full, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) ybuf[n] = da1 * out[n + Nz + lag];
            for (; n < L; n++) ybuf[n] = da1 * out[n + lag];
            double norm1 = norm_d(in1, n11, n12), norm2 = norm_d(in2, n21, n22);
            gn_lowlevel_d(ybuf, in1, in2, norm1, norm2, N, L, lag);
            if (yf) for (n = 0; n < L; n++) yf[tr][n] = (float)ybuf[n];
            else memcpy(yd[tr], ybuf, sizeof(double) * L);
        }
        free(in1); free(in2); free(out); free(ybuf); free(fin1); free(fin2); free(fout); free(full); free(ws);
    }
    offt_destroy_d(fZ); offt_destroy_d(bZ);
    return 0;
}

int orc_ccgn_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2)
{
    return ccgn_core(y, NULL, x1, x2, N, Tr, Lag1, Lag2);
}

/* ------------------------------------------------------------------------- */
/* 1-bit GNCC (src/FFTapps.c:980-1073)                                       */
/* ------------------------------------------------------------------------- */

int orc_cc1b_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2)
{
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned n11 = 0, n21 = 0, n12 = N, n22 = N;
    if (lag > 0) { n21 = (unsigned)lag; n12 -= (unsigned)lag; }
    else         { n11 = (unsigned)(-lag); n22 -= (unsigned)(-lag); }
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = nz_rule(N, M), Nh = Nz / 2 + 1;
    offt_plan_f *fZ = offt_make_f(Nz, 0), *bZ = offt_make_f(Nz, 1);
    size_t wl = maxz(offt_scratch_len_f(fZ), offt_scratch_len_f(bZ));
#pragma omp parallel
    {
        float *in1 = malloc(sizeof(float) * Nz), *in2 = malloc(sizeof(float) * Nz);
        float *out = malloc(sizeof(float) * Nz);
        float complex *fin1 = malloc(sizeof(float complex) * Nh), *fin2 = malloc(sizeof(float complex) * Nh);
        float complex *fout = malloc(sizeof(float complex) * Nh);
        float complex *full = malloc(sizeof(float complex) * Nz), *ws = malloc(sizeof(float complex) * wl);
        const double da1 = 1. / (double)Nz;
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            const float *pf1 = x1[tr];
            for (unsigned n = 0; n < N; n++) in1[n] = (pf1[n] >= 0) ? 1 : -1; /* :1037 */
            for (unsigned n = N; n < Nz; n++) in1[n] = 0;
            r2c_f(fin1, in1, Nz, fZ, full, ws);
            pf1 = x2[tr];
            for (unsigned n = 0; n < N; n++) in2[n] = (pf1[n] > 0) ? 1 : -1;  /* :1042 */
            for (unsigned n = N; n < Nz; n++) in2[n] = 0;
            r2c_f(fin2, in2, Nz, fZ, full, ws);
            /* ccf_lowlevel, :355-370 */
            for (unsigned n = 0; n < Nh; n++) fout[n] = conjf(fin1[n]) * fin2[n];
            c2r_f(out, fout, Nz, bZ, full, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) y[tr][n] = da1 * out[n + Nz + lag];
            for (; n < L; n++) y[tr][n] = da1 * out[n + lag];
            double norm1 = norm_f(in1, n11, n12), norm2 = norm_f(in2, n21, n22);
            gn_lowlevel_f(y[tr], in1, in2, norm1, norm2, N, L, lag);
        }
        free(in1); free(in2); free(out); free(fin1); free(fin2); free(fout); free(full); free(ws);
    }
    offt_destroy_f(fZ); offt_destroy_f(bZ);
    return 0;
}

/* all-FP64 1-bit: same signs, FP64 FFTs, exact overlap counts */
static int orc64_cc1b(double **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2)
{
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned n11 = 0, n21 = 0, n12 = N, n22 = N;
    if (lag > 0) { n21 = (unsigned)lag; n12 -= (unsigned)lag; }
    else         { n11 = (unsigned)(-lag); n22 -= (unsigned)(-lag); }
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = nz_rule(N, M), Nh = Nz / 2 + 1;
    offt_plan_d *fZ = offt_make_d(Nz, 0), *bZ = offt_make_d(Nz, 1);
    size_t wl = maxz(offt_scratch_len_d(fZ), offt_scratch_len_d(bZ));
#pragma omp parallel
    {
        double *in1 = malloc(sizeof(double) * Nz), *in2 = malloc(sizeof(double) * Nz);
        double *out = malloc(sizeof(double) * Nz);
        double complex *fin1 = malloc(sizeof(double complex) * Nh), *fin2 = malloc(sizeof(double complex) * Nh);
        double complex *fout = malloc(sizeof(double complex) * Nh);
        double complex *full = malloc(sizeof(double complex) * Nz), *ws = malloc(sizeof(double complex) * wl);
        const double da1 = 1. / (double)Nz;
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            for (unsigned n = 0; n < N; n++) in1[n] = (x1[tr][n] >= 0) ? 1 : -1;
            for (unsigned n = N; n < Nz; n++) in1[n] = 0;
            r2c_d(fin1, in1, Nz, fZ, full, ws);
            for (unsigned n = 0; n < N; n++) in2[n] = (x2[tr][n] > 0) ? 1 : -1;
            for (unsigned n = N; n < Nz; n++) in2[n] = 0;
            r2c_d(fin2, in2, Nz, fZ, full, ws);
            for (unsigned n = 0; n < Nh; n++) fout[n] = conj(fin1[n]) * fin2[n];
            c2r_d(out, fout, Nz, bZ, full, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) y[tr][n] = da1 * out[n + Nz + lag];
            for (; n < L; n++) y[tr][n] = da1 * out[n + lag];
            double norm1 = norm_d(in1, n11, n12), norm2 = norm_d(in2, n21, n22);
            gn_lowlevel_d(y[tr], in1, in2, norm1, norm2, N, L, lag);
        }
        free(in1); free(in2); free(out); free(fin1); free(fin2); free(fout); free(full); free(ws);
    }
    offt_destroy_d(fZ); offt_destroy_d(bZ);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* WPCC (src/FFTapps.c:1076-1255, CPU branch), FP64 throughout               */
/* ------------------------------------------------------------------------- */

/* The correlation of tspcc2_set (:1116-1252) with a given filter bank: S
 * wavelets wc[s] of Ls[s] samples with origin center[s], scales scale[s], the
 * record placed at Ls0 = Ls[S-1] in Nz zeros, output C * Re IFFT(Y) over L lags
 * from lag.  cuda_tspcc2_set_freq (src/ccs_cuda.cu:898-1036) computes the same
 * steps with a caller-built bank and C; the phasor here is the CPU AmpNorm. */
static void wpcc_with_bank(float **yf, double **yd, float **x1, float **x2, unsigned N, unsigned Tr,
                           unsigned Nz, unsigned S, const double *scale, double complex *const *wc,
                           const unsigned *Ls_, const int *center, double C, int lag, int L)
{
    const unsigned Ls0 = Ls_[S - 1];
    /* FFTs of the wavelet family (:1137-1149): w[c..Ls) at 0, w[0..c) at Nz-c */
    offt_plan_d *fZ = offt_make_d(Nz, 0), *bZ = offt_make_d(Nz, 1);
    size_t wl = maxz(offt_scratch_len_d(fZ), offt_scratch_len_d(bZ));
    double complex *fw = calloc((size_t)S * Nz, sizeof(double complex));
    {
        double complex *ws = malloc(sizeof(double complex) * wl);
        for (unsigned s = 0; s < S; s++) {
            double complex *f = fw + (size_t)s * Nz;
            const double complex *pc = wc[s];
            int c = center[s], Ls = (int)Ls_[s];
            memcpy(f, pc + c, (Ls - c) * sizeof(double complex));
            memcpy(f + Nz - c, pc, c * sizeof(double complex));
            offt_execute_ws_d(fZ, f, f, ws);
        }
        free(ws);
    }
    const double da2 = 1. / (double)Nz;
#pragma omp parallel
    {
        double complex *in1 = malloc(sizeof(double complex) * Nz), *in2 = malloc(sizeof(double complex) * Nz);
        double complex *x1w = malloc(sizeof(double complex) * Nz), *x2w = malloc(sizeof(double complex) * Nz);
        double complex *ywt = malloc(sizeof(double complex) * Nz), *ws = malloc(sizeof(double complex) * wl);
#pragma omp for schedule(static)
        for (unsigned tr = 0; tr < Tr; tr++) {
            memset(in1, 0, sizeof(double complex) * Nz);
            for (unsigned n = 0; n < N; n++) in1[Ls0 + n] = da2 * x1[tr][n];
            offt_execute_ws_d(fZ, in1, in1, ws);
            memset(in2, 0, sizeof(double complex) * Nz);
            for (unsigned n = 0; n < N; n++) in2[Ls0 + n] = da2 * x2[tr][n];
            offt_execute_ws_d(fZ, in2, in2, ws);
            memset(ywt, 0, sizeof(double complex) * Nz);
            for (unsigned s = 0; s < S; s++) {
                const double complex *pc1 = fw + (size_t)s * Nz;
                for (unsigned n = 0; n < Nz; n++) x1w[n] = in1[n] * pc1[n];
                offt_execute_ws_d(bZ, x1w, x1w, ws);
                orc_ampnorm(x1w, Nz);
                offt_execute_ws_d(fZ, x1w, x1w, ws);
                for (unsigned n = 0; n < Nz; n++) x2w[n] = in2[n] * pc1[n];
                offt_execute_ws_d(bZ, x2w, x2w, ws);
                orc_ampnorm(x2w, Nz);
                offt_execute_ws_d(fZ, x2w, x2w, ws);
                double da3 = 1. / (scale[s] * (double)Nz);
                for (unsigned n = 0; n < Nz; n++) ywt[n] += da3 * conj(x1w[n]) * x2w[n];
            }
            offt_execute_ws_d(bZ, ywt, ywt, ws);
            int n;
            for (n = 0; n < -lag && n < L; n++) {
                double v = C * creal(ywt[n + Nz + lag]);
                if (yf) yf[tr][n] = (float)v; else yd[tr][n] = v;
            }
            for (; n < L; n++) {
                double v = C * creal(ywt[n + lag]);
                if (yf) yf[tr][n] = (float)v; else yd[tr][n] = v;
            }
        }
        free(in1); free(in2); free(x1w); free(x2w); free(ywt); free(ws);
    }
    free(fw);
    offt_destroy_d(fZ); offt_destroy_d(bZ);
}

static int tspcc2_core(float **yf, double **yd, float **x1, float **x2, unsigned N, unsigned Tr,
                       int Lag1, int Lag2, double pmin, double pmax, unsigned V, int type, double op1)
{
    int L = abs(Lag2 - Lag1 + 1); /* as written at :1083 */
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    for (unsigned tr = 0; tr < Tr; tr++) {
        if (yf) memset(yf[tr], 0, L * sizeof(float));
        else memset(yd[tr], 0, L * sizeof(double));
    }
    orc_wfamily *wf = orc_wpcc_family(N, Lag1, Lag2, pmin, pmax, V, type, op1);
    if (!wf) return -3;
    const int lag = (Lag2 >= Lag1) ? Lag1 : Lag2;
    wpcc_with_bank(yf, yd, x1, x2, N, Tr, wf->Nz, wf->S, wf->scale, wf->wc, wf->Ls, wf->center,
                   1. / (wf->K0 * (double)N), lag, L);
    orc_wfamily_free(wf);
    return 0;
}

int orc_wpcc_bank_f64(double *y, const float *x1, const float *x2, unsigned N, unsigned Tr, unsigned Nz,
                      int Lag1, int Lag2, unsigned S, const double *scale, const double *wflat,
                      const unsigned *Ls, const int *center, double C)
{
    if (!S || !Tr) return -1;
    const int L = Lag2 >= Lag1 ? Lag2 - Lag1 + 1 : Lag1 - Lag2 + 1;
    const int lag = Lag2 >= Lag1 ? Lag1 : Lag2;
    double complex **wc = malloc(sizeof(double complex *) * S);
    size_t off = 0;
    for (unsigned s = 0; s < S; s++) {
        wc[s] = (double complex *)(wflat + 2 * off);
        off += Ls[s];
    }
    double **py = malloc(sizeof(double *) * Tr);
    float **p1 = malloc(sizeof(float *) * Tr), **p2 = malloc(sizeof(float *) * Tr);
    for (unsigned t = 0; t < Tr; ++t) {
        py[t] = y + (size_t)t * L;
        p1[t] = (float *)x1 + (size_t)t * N;
        p2[t] = (float *)x2 + (size_t)t * N;
    }
    wpcc_with_bank(NULL, py, p1, p2, N, Tr, Nz, S, scale, wc, Ls, center, C, lag, L);
    free(wc); free(py); free(p1); free(p2);
    return 0;
}

int orc_tspcc2_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2,
                   double pmin, double pmax, unsigned V, int type, double op1)
{
    return tspcc2_core(y, NULL, x1, x2, N, Tr, Lag1, Lag2, pmin, pmax, V, type, op1);
}

/* ------------------------------------------------------------------------- */
/* Phase-weighted stack of one pair's per-window correlations, all FP64       */
/* ------------------------------------------------------------------------- */

/* The reference stops at the per-window correlations and leaves their stacking
 * to the external ts-PWS program (src/PCC_fullpair_1b.c:568-569, README.md:10-11).
 * The phase-weighted stack (Schimmel & Paulssen 1997, GJI 130:497-505) of W
 * windows' correlations c_w(l) is
 *     PWS(l) = (1/W sum_w c_w(l)) * |1/W sum_w u_w(l)|^power,
 * u_w(l) = a_w(l)/|a_w(l)| (0 where a_w(l) = 0), a_w the analytic signal of the
 * window's correlation.  Here a_w is AnalyticSignal (src/FFTapps.c:45-80) of the
 * whole circular correlation r_w[0..Nz) (the real sequence the lag window is cut
 * from), read at the window's lags (lag0 + l) mod Nz: the Hilbert transform sees
 * the correlation beyond the lag window, so the window's ends carry no truncation
 * artefacts.  c_w is the kind's per-window output (all-FP64 build); r_w is, per
 * kind, Re IFFT(conj F1 F2) of the PCC2 phasor spectra, IFFT(conj S1 S2) of the
 * 1-bit sign spectra, IFFT(conj X1 X2) of GNCC's records (unnormalised: every
 * kind's normalisation is a positive factor, which leaves u unchanged). */
static void circ_corr_d(int kind, double *r, const float *x1, const float *x2, unsigned N, unsigned Nz)
{
    offt_plan_d *fN = offt_make_d(N, 0), *bN = offt_make_d(N, 1);
    offt_plan_d *fZ = offt_make_d(Nz, 0), *bZ = offt_make_d(Nz, 1);
    size_t wl = maxz(maxz(offt_scratch_len_d(fN), offt_scratch_len_d(bN)),
                     maxz(offt_scratch_len_d(fZ), offt_scratch_len_d(bZ)));
    double complex *a = calloc(Nz, sizeof(double complex)), *b = calloc(Nz, sizeof(double complex));
    double complex *ws = malloc(sizeof(double complex) * wl);
    double *xd = malloc(sizeof(double) * N);
    const float *xs[2] = {x1, x2};
    double complex *ab[2] = {a, b};
    for (int j = 0; j < 2; j++) {
        double complex *v = ab[j];
        if (kind == ORC_PCC2) {
            for (unsigned n = 0; n < N; ++n) xd[n] = xs[j][n];
            analytic_d_plan(v, xd, N, fN, bN, ws);
            orc_ampnormf_d(v, N);
        } else if (kind == ORC_CC1B) {
            /* :1037, :1042 */
            for (unsigned n = 0; n < N; n++) v[n] = j == 0 ? (xs[j][n] >= 0 ? 1 : -1) : (xs[j][n] > 0 ? 1 : -1);
        } else {
            for (unsigned n = 0; n < N; n++) v[n] = xs[j][n];
        }
        for (unsigned n = N; n < Nz; n++) v[n] = 0;
        offt_execute_ws_d(fZ, v, v, ws);
    }
    for (unsigned n = 0; n < Nz; n++) a[n] = conj(a[n]) * b[n];
    offt_execute_ws_d(bZ, a, a, ws);
    for (unsigned n = 0; n < Nz; n++) r[n] = creal(a[n]);
    free(a); free(b); free(ws); free(xd);
    offt_destroy_d(fN); offt_destroy_d(bN); offt_destroy_d(fZ); offt_destroy_d(bZ);
}

int orc_pws_f64(int kind, double *y, const float *x, unsigned S, unsigned W, unsigned N, int Lag1, int Lag2,
                unsigned pair_begin, unsigned pair_end, double power)
{
    if (kind != ORC_PCC2 && kind != ORC_CC1B && kind != ORC_GNCC) return -1;
    int L, lag;
    lag_setup(Lag1, Lag2, &L, &lag);
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    const unsigned Nz = nz_rule(N, ua1 > ua2 ? ua1 : ua2), Ntot = W * N, P = S * (S - 1) / 2;
    if (pair_end > P) pair_end = P;
    double *c = malloc(sizeof(double) * (size_t)W * L);  /* c_w(l) of every window */
    double *r = malloc(sizeof(double) * Nz);
    double complex *z = malloc(sizeof(double complex) * Nz);
    double *sc = malloc(sizeof(double) * L);
    double complex *su = malloc(sizeof(double complex) * L);
    unsigned p = 0;
    int rc = 0;
    for (unsigned i = 0; i < S && p < pair_end; ++i)
        for (unsigned j = i + 1; j < S && p < pair_end; ++j, ++p) {
            if (p < pair_begin) continue;
            const float *x1 = x + (size_t)i * Ntot, *x2 = x + (size_t)j * Ntot;
            rc = orc_run_f64(kind, c, x1, x2, N, W, Lag1, Lag2, 2.0, 0, 0, 8, -1, 0);
            if (rc) goto done;
            for (int l = 0; l < L; ++l) { sc[l] = 0; su[l] = 0; }
            for (unsigned w = 0; w < W; ++w) {
                circ_corr_d(kind, r, x1 + (size_t)w * N, x2 + (size_t)w * N, N, Nz);
                orc_analytic_d(z, r, Nz);
                for (int l = 0; l < L; ++l) {
                    const double complex v = z[((long)lag + l + Nz) % Nz];
                    const double m = cabs(v);
                    sc[l] += c[(size_t)w * L + l];
                    if (m > 0) su[l] += v / m;
                }
            }
            double *dst = y + (size_t)(p - pair_begin) * L;
            for (int l = 0; l < L; ++l) {
                const double coh = power == 0 ? 1.0 : pow(cabs(su[l] / W), power);
                dst[l] = sc[l] / W * coh;
            }
        }
done:
    free(c); free(r); free(z); free(sc); free(su);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* flat wrappers                                                             */
/* ------------------------------------------------------------------------- */

static int out_len(int kind, int Lag1, int Lag2)
{
    if (kind == ORC_WPCC) return abs(Lag2 - Lag1 + 1);
    return Lag2 >= Lag1 ? Lag2 - Lag1 + 1 : Lag1 - Lag2 + 1;
}

int orc_run_f32(int kind, float *y, const float *x1, const float *x2, unsigned N, unsigned Tr,
                int Lag1, int Lag2, double v, double pmin, double pmax, unsigned V, int type, double op1)
{
    int L = out_len(kind, Lag1, Lag2);
    float **py = malloc(sizeof(float *) * Tr), **p1 = malloc(sizeof(float *) * Tr), **p2 = malloc(sizeof(float *) * Tr);
    for (unsigned t = 0; t < Tr; ++t) {
        py[t] = y + (size_t)t * L;
        p1[t] = (float *)x1 + (size_t)t * N;
        p2[t] = (float *)x2 + (size_t)t * N;
    }
    int r;
    switch (kind) {
    case ORC_PCC2: r = orc_pcc2_set(py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_PCC1: r = orc_pcc1_set(py, p1, p2, (int)N, Tr, Lag1, Lag2); break;
    case ORC_PCCV: r = orc_pcc_set(py, p1, p2, (int)N, Tr, v, Lag1, Lag2); break;
    case ORC_GNCC: r = orc_ccgn_set(py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_CC1B: r = orc_cc1b_set(py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_WPCC: r = orc_tspcc2_set(py, p1, p2, N, Tr, Lag1, Lag2, pmin, pmax, V, type, op1); break;
    default: r = -1;
    }
    free(py); free(p1); free(p2);
    return r;
}

int orc_run_f64(int kind, double *y, const float *x1, const float *x2, unsigned N, unsigned Tr,
                int Lag1, int Lag2, double v, double pmin, double pmax, unsigned V, int type, double op1)
{
    int L = out_len(kind, Lag1, Lag2);
    double **py = malloc(sizeof(double *) * Tr);
    float **p1 = malloc(sizeof(float *) * Tr), **p2 = malloc(sizeof(float *) * Tr);
    for (unsigned t = 0; t < Tr; ++t) {
        py[t] = y + (size_t)t * L;
        p1[t] = (float *)x1 + (size_t)t * N;
        p2[t] = (float *)x2 + (size_t)t * N;
    }
    int r;
    switch (kind) {
    case ORC_PCC2: r = orc64_pcc2(py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_PCC1: r = orc64_pcc_time(py, p1, p2, (int)N, Tr, 1.0, 1, Lag1, Lag2); break;
    case ORC_PCCV: r = orc64_pcc_time(py, p1, p2, (int)N, Tr, v, 0, Lag1, Lag2); break;
    case ORC_GNCC: r = ccgn_core(NULL, py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_CC1B: r = orc64_cc1b(py, p1, p2, N, Tr, Lag1, Lag2); break;
    case ORC_WPCC: r = tspcc2_core(NULL, py, p1, p2, N, Tr, Lag1, Lag2, pmin, pmax, V, type, op1); break;
    default: r = -1;
    }
    free(py); free(p1); free(p2);
    return r;
}

/* All pairs, one *_set call per pair over its W windows; stacked = mean over windows. */
#define ALLPAIRS_BODY(T, RUN)                                                                  \
    int L = out_len(kind, Lag1, Lag2);                                                         \
    unsigned Ntot = W * N, P = S * (S - 1) / 2;                                                \
    if (pair_end > P) pair_end = P;                                                            \
    T *ybuf = malloc(sizeof(T) * (size_t)W * L);                                               \
    float *x1 = malloc(sizeof(float) * (size_t)W * N), *x2 = malloc(sizeof(float) * (size_t)W * N); \
    unsigned p = 0;                                                                            \
    int rc = 0;                                                                                \
    for (unsigned i = 0; i < S && p < pair_end; ++i)                                           \
        for (unsigned j = i + 1; j < S && p < pair_end; ++j, ++p) {                            \
            if (p < pair_begin) continue;                                                      \
            memcpy(x1, x + (size_t)i * Ntot, sizeof(float) * (size_t)W * N);                   \
            memcpy(x2, x + (size_t)j * Ntot, sizeof(float) * (size_t)W * N);                   \
            int r = RUN(kind, ybuf, x1, x2, N, W, Lag1, Lag2, v, pmin, pmax, V, type, op1);    \
            if (r) { rc = r; goto done; }                                                      \
            T *dst = y + (size_t)(p - pair_begin) * (stacked ? 1 : W) * L;                     \
            if (stacked) {                                                                     \
                for (int l = 0; l < L; ++l) {                                                  \
                    double acc = 0;                                                            \
                    for (unsigned w = 0; w < W; ++w) acc += ybuf[(size_t)w * L + l];           \
                    dst[l] = (T)(acc / W);                                                     \
                }                                                                              \
            } else {                                                                           \
                memcpy(dst, ybuf, sizeof(T) * (size_t)W * L);                                  \
            }                                                                                  \
        }                                                                                      \
done:                                                                                          \
    free(ybuf); free(x1); free(x2);                                                            \
    return rc;

int orc_allpairs_f32(int kind, float *y, const float *x, unsigned S, unsigned W, unsigned N,
                     int Lag1, int Lag2, int stacked, unsigned pair_begin, unsigned pair_end,
                     double v, double pmin, double pmax, unsigned V, int type, double op1)
{
    ALLPAIRS_BODY(float, orc_run_f32)
}

int orc_allpairs_f64(int kind, double *y, const float *x, unsigned S, unsigned W, unsigned N,
                     int Lag1, int Lag2, int stacked, unsigned pair_begin, unsigned pair_end,
                     double v, double pmin, double pmax, unsigned V, int type, double op1)
{
    ALLPAIRS_BODY(double, orc_run_f64)
}
