/*
 * TEST INFRASTRUCTURE ONLY — CPU restatement of the reference's per-station
 * preprocessing (the steps PCC_fullpair_1b runs before the correlations), used
 * by the tests as the checker of the GPU preprocessing.  Never linked into the
 * product library.
 *
 *   orc_qabsmedian / orc_qmedian   src/PCC_fullpair_1b.c:1179-1214, 1216-1251
 *   orc_clipping                   src/PCC_fullpair_1b.c:881-891
 *   zero-trace rule                src/ReadManySacs.c:529-559 (RemoveZeroTraces)
 *   trace rms                      src/PCC_fullpair_1b.c:355-363
 *   outlier rule                   src/PCC_fullpair_1b.c:893-925 (RemoveOutlierTraces)
 *   orc_avewhite                   src/PCC_fullpair_1b.c:1065-1177 (AveWhite)
 *   polarity                       src/rotlib.c:14-34 (CorrectRevesedPolarity), as a
 *                                  per-trace sign factor
 *
 * FFTW is replaced by the oracle's FP64 FFT (offt.c), so the whitening is
 * tolerance-level, not bitwise.  AveWhite never writes HS[5] (its smoothing
 * loops cover n = 5..9 and n >= 11, src/PCC_fullpair_1b.c:1134-1145); the
 * buffer is fresh fftw_malloc memory, which this restatement takes as 0.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "offt.h"
#include "oracle.h"

#define ORC_PI 3.14159265358979323846

/* k-th smallest by Hoare partitioning, then the reference's even-length rule:
 * the mean of a[k] and the largest value left of it. */
static float select_median(float *a, int n)
{
    int k = n / 2, l = 0, m = n - 1;
    while (l < m) {
        float x = a[k];
        int i = l, j = m;
        do {
            while (a[i] < x) i++;
            while (x < a[j]) j--;
            if (i <= j) {
                float t = a[i];
                a[i] = a[j];
                a[j] = t;
                i++;
                j--;
            }
        } while (i <= j);
        if (j < k) l = i;
        if (k < i) m = j;
    }
    if (n % 2) return a[k];
    float t = -FLT_MAX;
    for (int i = 0; i < k; i++)
        if (a[i] <= a[k] && t < a[i]) t = a[i];
    return (a[k] + t) / 2;
}

float orc_qabsmedian(const float *x, int n)
{
    if (n < 1) return 0.f;
    float *a = malloc(sizeof(float) * n);
    if (!a) return 0.f;
    for (int i = 0; i < n; i++) a[i] = fabsf(x[i]);
    float r = select_median(a, n);
    free(a);
    return r;
}

float orc_qmedian(const float *x, int n)
{
    if (n < 1) return 0.f;
    float *a = malloc(sizeof(float) * n);
    if (!a) return 0.f;
    memcpy(a, x, sizeof(float) * n);
    float r = select_median(a, n);
    free(a);
    return r;
}

float orc_clip_threshold(const float *x, unsigned N)
{
    return (float)(4 * orc_qabsmedian(x, (int)N) / 0.6745);
}

void orc_clipping(float *x, unsigned N)
{
    const float fa1 = orc_clip_threshold(x, N);
    for (unsigned n = 0; n < N; n++) {
        const float fa2 = x[n];
        if (fabsf(fa2) > fa1) x[n] = (fa2 > fa1) ? fa1 : -fa1;
    }
}

float orc_trace_rms(const float *x, unsigned N)
{
    double da2 = 0;
    for (unsigned n = 0; n < N; n++) da2 += x[n] * x[n];
    return (float)sqrt(da2 / N);
}

/* The whitening filter HS[0..Nh) of the averaged amplitude spectrum H (modified
 * in place), as AveWhite builds it from n1, n2 (src/PCC_fullpair_1b.c:1100-1145). */
void orc_white_filter(double *H, double *HS, unsigned Nh, unsigned Nz, unsigned N, unsigned n1, unsigned n2)
{
    double da1, mn, mx, win[11];
    unsigned n, m, M;
    da1 = 1. / H[n1];
    for (n = 0; n < n1; n++) H[n] = da1;
    for (; n < n2; n++) H[n] = 1. / H[n];
    da1 = 1. / H[n2];
    for (; n < Nh; n++) H[n] = da1;
    mn = mx = H[n1];
    for (n = n1 + 1; n < n2; n++) {
        da1 = H[n];
        if (da1 > mx) mx = da1;
        else if (da1 < mn) mn = da1;
    }
    if (mx > 1000 * mn) mx = 1000 * mn;
    for (n = 0; n < Nh; n++)
        if (H[n] > mx) H[n] = mx;
    da1 = 1 / (mn * (double)Nz);
    for (n = 0; n < Nh; n++) H[n] *= da1;

    for (m = 0; m < 11; m++) {
        da1 = 2 * ORC_PI * m / (N - 1);
        win[m] = 0.42 - 0.5 * cos(da1) + 0.08 * cos(2 * da1);
    }
    da1 = 0;
    for (m = 0; m < 11; m++) da1 += win[m];
    da1 = 1 / da1;
    for (m = 0; m < 11; m++) win[m] *= da1;

    for (n = 0; n < Nh; n++) HS[n] = 0.0;
    for (n = 5; n < 10; n++) {
        da1 = 0;
        for (m = 0; m <= n; m++) da1 += win[m] * H[n - m];
        for (; m < 11; m++) da1 += win[m] * H[m - n];
        HS[n - 5] = da1;
    }
    for (n = 11; n < Nh; n++) {
        da1 = 0;
        for (m = 0; m < 11; m++) da1 += win[m] * H[n - m];
        HS[n - 5] = da1;
    }
    for (n = Nh; n < Nh + 5; n++) {
        da1 = 0;
        M = n - (Nh - 1);
        for (m = 0; m < M; m++) da1 += win[m] * H[Nh - 1 + m - M];
        for (; m < 11; m++) da1 += win[m] * H[n - m];
        HS[n - 5] = da1;
    }
}

unsigned orc_white_nz(unsigned N) { return 1u << (unsigned)ceil(log2((double)N)); }

/* AveWhite over Tr traces of N samples (in place). */
int orc_avewhite(float **x0, unsigned N, unsigned Tr, const double freq[2], double dt)
{
    const unsigned Nz = orc_white_nz(N), Nh = Nz / 2 + 1;
    const unsigned n1 = freq[0] * dt * (double)Nz, n2 = freq[1] * dt * (double)Nz;
    if (n2 >= Nh || n1 > n2 || Tr == 0) return -1;
    offt_plan_d *pf = offt_make_d((int)Nz, 0), *pi = offt_make_d((int)Nz, 1);
    double complex *X = malloc(sizeof(double complex) * Nz);
    double *H = calloc(Nh, sizeof(double)), *HS = calloc(Nh, sizeof(double));
    if (!pf || !pi || !X || !H || !HS) {
        free(X);
        free(H);
        free(HS);
        if (pf) offt_destroy_d(pf);
        if (pi) offt_destroy_d(pi);
        return -1;
    }
    for (unsigned tr = 0; tr < Tr; tr++) {
        for (unsigned n = 0; n < Nz; n++) X[n] = n < N ? (double)x0[tr][n] : 0.0;
        offt_execute_d(pf, X, X);
        for (unsigned n = 0; n < Nh; n++) H[n] += cabs(X[n]);
    }
    double da1 = 1. / (double)Tr;
    for (unsigned n = 0; n < Nh; n++) H[n] *= da1;
    orc_white_filter(H, HS, Nh, Nz, N, n1, n2);
    da1 = 1. / (double)Nz;
    for (unsigned tr = 0; tr < Tr; tr++) {
        for (unsigned n = 0; n < Nz; n++) X[n] = n < N ? (double)x0[tr][n] : 0.0;
        offt_execute_d(pf, X, X);
        /* c2r of the Hermitian half spectrum: DC and Nyquist imaginary parts are ignored */
        for (unsigned n = 0; n < Nh; n++) X[n] *= da1 * HS[n];
        X[0] = creal(X[0]);
        X[Nz / 2] = creal(X[Nz / 2]);
        for (unsigned n = 1; n < Nz / 2; n++) X[Nz - n] = conj(X[n]);
        offt_execute_d(pi, X, X);
        for (unsigned n = 0; n < N; n++) x0[tr][n] = (float)creal(X[n]);
    }
    free(X);
    free(H);
    free(HS);
    offt_destroy_d(pf);
    offt_destroy_d(pi);
    return 0;
}

/* The preprocessing of each station's windows, in the order PCC_fullpair_1b
 * runs it (src/PCC_fullpair_1b.c:330-424, 469-470): polarity signs, zero-window
 * removal, rms, clipping, outlier removal (rms > nstd * median rms of the
 * station's remaining windows), then whitening of the kept windows.  common != 0
 * keeps a window only where every station kept it (the paired lists of a
 * two-station run, src/PCC_fullpair_1b.c:399-401, with windows aligned by index).
 * Dropped windows are zeroed.  x is [S][stride], window w of station s at
 * s*stride + w*N; valid is [S][W]; rms is [S][W] (before clipping). */
int orc_preprocess(float *x, unsigned S, unsigned W, unsigned N, unsigned long long stride, int clip, float nstd,
                   double f1, double f2, double dt, int common, const float *sign, unsigned char *valid, float *rms)
{
    float *vals = malloc(sizeof(float) * (W ? W : 1));
    float **rows = malloc(sizeof(float *) * (W ? W : 1));
    if (!vals || !rows) {
        free(vals);
        free(rows);
        return -1;
    }
    for (unsigned s = 0; s < S; s++) {
        for (unsigned w = 0; w < W; w++) {
            float *px = x + s * stride + (unsigned long long)w * N;
            if (sign && sign[s * W + w] < 0)
                for (unsigned n = 0; n < N; n++) px[n] = -px[n];
            int nz = 0;
            for (unsigned n = 0; n < N; n++)
                if (0 != px[n]) {
                    nz = 1;
                    break;
                }
            valid[s * W + w] = (unsigned char)nz;
            rms[s * W + w] = orc_trace_rms(px, N);
            if (clip && nz) orc_clipping(px, N);
        }
        if (nstd > 0) {
            unsigned Tr = 0;
            for (unsigned w = 0; w < W; w++)
                if (valid[s * W + w]) vals[Tr++] = rms[s * W + w];
            const float fa1 = orc_qmedian(vals, (int)Tr);
            for (unsigned w = 0; w < W; w++)
                if (valid[s * W + w] && rms[s * W + w] > nstd * fa1) valid[s * W + w] = 0;
        }
    }
    if (common)
        for (unsigned w = 0; w < W; w++) {
            unsigned char all = 1;
            for (unsigned s = 0; s < S; s++) all &= valid[s * W + w];
            for (unsigned s = 0; s < S; s++) valid[s * W + w] = all;
        }
    int rc = 0;
    for (unsigned s = 0; s < S; s++) {
        unsigned Tr = 0;
        for (unsigned w = 0; w < W; w++) {
            float *px = x + s * stride + (unsigned long long)w * N;
            if (valid[s * W + w]) rows[Tr++] = px;
            else memset(px, 0, sizeof(float) * N);
        }
        if (f1 > 0 && f2 > f1 && Tr > 0) {
            const double band[2] = {f1, f2};
            if (orc_avewhite(rows, N, Tr, band, dt)) rc = -1;
        }
    }
    free(vals);
    free(rows);
    return rc;
}
