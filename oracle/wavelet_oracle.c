/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle of the WPCC filter bank.
 *
 * Restates tspcc2_set's parameter setup (src/FFTapps.c:1076-1115) and
 * CreateWaveletFamily with its helpers (src/FWTa/wavelet_def_v7.c:21-131,
 * 204-339) in FP64: scales s0*2^(m/V) by repeated multiplication, lengths
 * min(2*ceil(5*s)+1, Nz), centre Ls/2, complex Morlet / complete Morlet /
 * complex Mexican hat samples.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include "oracle.h"

#define WPI 3.14159265358979328 /* src/FWTa/wavelet_v7.h:16 (same double as M_PI) */
#define NSIGMAS 5               /* src/FWTa/wavelet_v7.h:19 */

/* MorletFun, wavelet_def_v7.c:38-52 */
static void morlet(double complex *w, int *center, double w0, unsigned Ls, double scale)
{
    double k = 1 / sqrt(sqrt(WPI) * scale), daux, w1;
    int u, u0 = Ls / 2;
    *center = Ls / 2;
    scale = 1 / scale;
    for (u = -u0; u < (int)Ls - u0; u++) {
        daux = scale * u;
        w1 = w0 * daux;
        daux *= daux;
        *w++ = k * cexp(I * w1) * exp(-0.5 * daux);
    }
}

/* Complete_MorletFun, wavelet_def_v7.c:72-88 */
static void complete_morlet(double complex *w, int *center, double w0, unsigned Ls, double scale)
{
    const double ze = exp((-w0 * w0) / 2);
    double k = 1 / sqrt(sqrt(WPI) * scale), daux, w1;
    int u, u0 = Ls / 2;
    *center = Ls / 2;
    scale = 1 / scale;
    for (u = -u0; u < (int)Ls - u0; u++) {
        daux = scale * u;
        w1 = w0 * daux;
        daux *= daux;
        daux = k * exp(-0.5 * daux);
        *w++ = daux * (cexp(I * w1) - ze);
    }
}

/* erfi, wavelet_def_v7.c:104-117: 500-term Maclaurin series */
double orc_erfi(double z)
{
    double da1 = z, out = z, zz = z * z;
    for (unsigned n = 1; n < 500; n++) {
        da1 *= zz / n;
        out += da1 / (2 * n + 1);
    }
    out *= 2 / sqrt(WPI);
    return out;
}

/* MexicanHatFun, wavelet_def_v7.c:119-131 */
static void mexhat(double complex *w, int *center, double dump, unsigned Ls, double scale)
{
    (void)dump;
    double k = 2 / sqrt(3 * sqrt(WPI) * scale), daux;
    int u, u0 = Ls / 2;
    *center = Ls / 2;
    scale = 1 / scale;
    for (u = -u0; u < (int)Ls - u0; u++) {
        daux = scale * u;
        *w++ = k * ((daux * daux - 1) * exp(-daux * daux / 2) * (1 + I * orc_erfi(daux / sqrt(2)))
                    - I * sqrt(2 / WPI) * daux);
    }
}

void orc_wfamily_free(orc_wfamily *w)
{
    if (!w) return;
    if (w->wc) free(w->wc[0]);
    free(w->wc);
    free(w->scale);
    free(w->Ls);
    free(w->center);
    free(w);
}

orc_wfamily *orc_wpcc_family(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                             int type, double op1)
{
    unsigned ua1 = abs(Lag1), ua2 = abs(Lag2);
    if (ua1 >= N || ua2 >= N || V == 0) return NULL;
    unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = 1u << (unsigned)ceil(log2((double)(N + M)));
    double s0;
    /* src/FFTapps.c:1097-1106 */
    if (type == -3) {
        s0 = pmin / (sqrt(2) * M_PI);
    } else if (type == -1 || type == -2) {
        if (op1 == 0) op1 = M_PI * sqrt(2 / log(2));
        s0 = pmin * op1 / (2 * M_PI);
    } else {
        return NULL;
    }
    unsigned J = (unsigned)round(1. / (double)V + log(pmax / pmin) / log(2));
    if (J == 0) return NULL;
    orc_wfamily *w = calloc(1, sizeof(orc_wfamily));
    w->J = J;
    w->V = V;
    w->S = J * V;
    w->s0 = s0;
    w->op1 = op1;
    const unsigned S = w->S;
    if (S == 0) { /* V = 0: no scales */
        free(w);
        return NULL;
    }
    w->scale = malloc(sizeof(double) * S);
    w->Ls = malloc(sizeof(unsigned) * S);
    w->center = malloc(sizeof(int) * S);
    /* setscales0, wavelet_def_v7.c:299-309 (a0 = 2) */
    double da1 = s0, da2 = pow(2, 1 / (double)V);
    for (unsigned s = 0; s < S; s++) {
        w->scale[s] = da1;
        da1 *= da2;
    }
    /* setwaveletlength0, :312-322 with MaxL = Nz (before doubling) */
    size_t nsmpl = 0;
    for (unsigned s = 0; s < S; s++) {
        unsigned ua = 2 * (unsigned)ceil(NSIGMAS * w->scale[s]) + 1;
        w->Ls[s] = ua > Nz ? Nz : ua;
        nsmpl += w->Ls[s];
    }
    w->wc = malloc(sizeof(double complex *) * S);
    w->wc[0] = malloc(sizeof(double complex) * nsmpl);
    for (unsigned s = 1; s < S; s++) w->wc[s] = w->wc[s - 1] + w->Ls[s - 1];
    for (unsigned s = 0; s < S; s++) {
        if (type == -1) morlet(w->wc[s], w->center + s, op1, w->Ls[s], w->scale[s]);
        else if (type == -2) complete_morlet(w->wc[s], w->center + s, op1, w->Ls[s], w->scale[s]);
        else mexhat(w->wc[s], w->center + s, op1, w->Ls[s], w->scale[s]);
    }
    /* src/FFTapps.c:1110-1115 */
    w->Ls0 = w->Ls[S - 1];
    if (Nz - (N + M) < w->Ls0) Nz *= 2;
    w->Nz = Nz;
    double K0 = 0;
    for (unsigned m = 0; m < S; m++) K0 += 1 / w->scale[m];
    w->K0 = K0;
    return w;
}

int orc_wpcc_family_export(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                           int type, double op1, unsigned *out_meta, double *out_s0_K0_op1,
                           double *scale, unsigned *Ls, int *center, double *wbuf, unsigned wbuf_len)
{
    orc_wfamily *w = orc_wpcc_family(N, Lag1, Lag2, pmin, pmax, V, type, op1);
    if (!w) return -3;
    out_meta[0] = w->S;
    out_meta[1] = w->J;
    out_meta[2] = w->V;
    out_meta[3] = w->Nz;
    out_meta[4] = w->Ls0;
    out_s0_K0_op1[0] = w->s0;
    out_s0_K0_op1[1] = w->K0;
    out_s0_K0_op1[2] = w->op1;
    size_t total = 0;
    for (unsigned s = 0; s < w->S; ++s) total += w->Ls[s];
    if (scale) memcpy(scale, w->scale, sizeof(double) * w->S);
    if (Ls) memcpy(Ls, w->Ls, sizeof(unsigned) * w->S);
    if (center) memcpy(center, w->center, sizeof(int) * w->S);
    if (wbuf && wbuf_len >= 2 * total) memcpy(wbuf, w->wc[0], sizeof(double complex) * total);
    int S = (int)w->S;
    orc_wfamily_free(w);
    return S;
}
