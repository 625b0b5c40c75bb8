/*
 * TEST INFRASTRUCTURE ONLY — part of the CPU oracle, never part of the product.
 *
 * Mixed-radix complex FFT (Cooley–Tukey decimation in time, radices 4/2/3/5 and a
 * generic odd-prime butterfly) with Bluestein's chirp-z algorithm for lengths that
 * have a prime factor above 61.  It stands in for FFTW, which the reference links
 * (src/FFTapps.c:2, src/makefile:7) and which is absent here, so FFTW's exact
 * rounding is not reproduced: parity with the reference is at tolerance level.
 *
 * Conventions follow FFTW: forward = exp(-2*pi*i*n*k/N), backward = exp(+...),
 * both unnormalised.
 *
 * This file is a type template: it is included by offt.c once with
 * OFFT_R=float and once with OFFT_R=double.  Twiddles and chirps are always
 * generated in double and rounded to OFFT_R.
 */
#ifndef OFFT_R
#error "define OFFT_R and OFFT_S before including offt_impl.h"
#endif

#define OFFT_CAT_(a, b) a##b
#define OFFT_CAT(a, b) OFFT_CAT_(a, b)
#define CPX OFFT_CAT(offt_cpx_, OFFT_S)
#define PLAN OFFT_CAT(offt_plan_, OFFT_S)
#define FN(name) OFFT_CAT(OFFT_CAT(offt_, name), OFFT_CAT(_, OFFT_S))

static void FN(kf_bfly_generic)(CPX *out, size_t fstride, const PLAN *p, int m, int p_radix, CPX *scratch)
{
    const CPX *tw = p->tw;
    int n = p->n;
    for (int u = 0; u < m; ++u) {
        for (int q1 = 0, k = u; q1 < p_radix; ++q1, k += m) scratch[q1] = out[k];
        for (int q1 = 0, k = u; q1 < p_radix; ++q1, k += m) {
            size_t twidx = 0;
            CPX acc = scratch[0];
            for (int q = 1; q < p_radix; ++q) {
                twidx += fstride * (size_t)k;
                if (twidx >= (size_t)n) twidx -= (size_t)n * (twidx / (size_t)n);
                acc += scratch[q] * tw[twidx];
            }
            out[k] = acc;
        }
    }
}

static void FN(kf_work)(CPX *out, const CPX *in, size_t fstride, size_t in_stride,
                        const int *factors, const PLAN *p, CPX *scratch)
{
    CPX *out_beg = out;
    const int radix = *factors++;
    const int m = *factors++;
    const CPX *out_end = out + radix * m;

    if (m == 1) {
        do {
            *out = *in;
            in += fstride * in_stride;
        } while (++out != out_end);
    } else {
        do {
            FN(kf_work)(out, in, fstride * radix, in_stride, factors, p, scratch);
            in += fstride * in_stride;
            out += m;
        } while (out != out_end);
    }
    out = out_beg;

    const CPX *tw = p->tw;
    switch (radix) {
    case 2: {
        for (int k = 0; k < m; ++k) {
            CPX t = out[k + m] * tw[k * fstride];
            out[k + m] = out[k] - t;
            out[k] += t;
        }
    } break;
    case 4: {
        /* j = +/- i depending on direction */
        for (int k = 0; k < m; ++k) {
            CPX s0 = out[k + m] * tw[k * fstride];
            CPX s1 = out[k + 2 * m] * tw[k * fstride * 2];
            CPX s2 = out[k + 3 * m] * tw[k * fstride * 3];
            CPX s5 = out[k] - s1;
            CPX a0 = out[k] + s1;
            CPX s3 = s0 + s2;
            CPX s4 = s0 - s2;
            CPX rot; /* s4 * (-i) forward, s4 * (+i) inverse */
            if (p->inverse) rot = OFFT_CMPLX(-OFFT_CIMAG(s4), OFFT_CREAL(s4));
            else            rot = OFFT_CMPLX(OFFT_CIMAG(s4), -OFFT_CREAL(s4));
            out[k + 2 * m] = a0 - s3;
            out[k] = a0 + s3;
            out[k + m] = s5 + rot;
            out[k + 3 * m] = s5 - rot;
        }
    } break;
    case 3: {
        const CPX w1 = tw[fstride * m]; /* exp(sign*2*pi*i/3) */
        const OFFT_R c = OFFT_CREAL(w1), s = OFFT_CIMAG(w1);
        for (int k = 0; k < m; ++k) {
            CPX a1 = out[k + m] * tw[k * fstride];
            CPX a2 = out[k + 2 * m] * tw[k * fstride * 2];
            CPX sum = a1 + a2, dif = a1 - a2;
            CPX x0 = out[k];
            out[k] = x0 + sum;
            CPX base = x0 + c * sum;
            CPX rot = OFFT_CMPLX(-s * OFFT_CIMAG(dif), s * OFFT_CREAL(dif)); /* i*s*dif */
            out[k + m] = base + rot;
            out[k + 2 * m] = base - rot;
        }
    } break;
    case 5: {
        const CPX ya = tw[fstride * m], yb = tw[fstride * 2 * m];
        for (int u = 0; u < m; ++u) {
            CPX s0 = out[u];
            CPX s1 = out[u + m] * tw[u * fstride];
            CPX s2 = out[u + 2 * m] * tw[2 * u * fstride];
            CPX s3 = out[u + 3 * m] * tw[3 * u * fstride];
            CPX s4 = out[u + 4 * m] * tw[4 * u * fstride];
            CPX s7 = s1 + s4, s10 = s1 - s4, s8 = s2 + s3, s9 = s2 - s3;
            out[u] = s0 + s7 + s8;
            CPX s5 = s0 + OFFT_CREAL(ya) * s7 + OFFT_CREAL(yb) * s8;
            CPX s6 = OFFT_CMPLX(OFFT_CIMAG(ya) * OFFT_CREAL(s10) + OFFT_CIMAG(yb) * OFFT_CREAL(s9),
                                OFFT_CIMAG(ya) * OFFT_CIMAG(s10) + OFFT_CIMAG(yb) * OFFT_CIMAG(s9));
            /* multiply s6 by i */
            CPX is6 = OFFT_CMPLX(-OFFT_CIMAG(s6), OFFT_CREAL(s6));
            out[u + m] = s5 + is6;
            out[u + 4 * m] = s5 - is6;
            CPX s11 = s0 + OFFT_CREAL(yb) * s7 + OFFT_CREAL(ya) * s8;
            CPX s12 = OFFT_CMPLX(OFFT_CIMAG(yb) * OFFT_CREAL(s10) - OFFT_CIMAG(ya) * OFFT_CREAL(s9),
                                 OFFT_CIMAG(yb) * OFFT_CIMAG(s10) - OFFT_CIMAG(ya) * OFFT_CIMAG(s9));
            /* X2 = s11 + i*s12 ; X3 = s11 - i*s12 */
            CPX is12 = OFFT_CMPLX(-OFFT_CIMAG(s12), OFFT_CREAL(s12));
            out[u + 2 * m] = s11 + is12;
            out[u + 3 * m] = s11 - is12;
        }
    } break;
    default:
        FN(kf_bfly_generic)(out, fstride, p, m, radix, scratch);
    }
}

static int FN(factorize)(int n, int *facbuf, int *maxp)
{
    int p = 4, nf = 0;
    *maxp = 1;
    do {
        while (n % p) {
            switch (p) {
            case 4: p = 2; break;
            case 2: p = 3; break;
            default: p += 2; break;
            }
            if (p * p > n) p = n;
        }
        n /= p;
        facbuf[2 * nf] = p;
        facbuf[2 * nf + 1] = n;
        if (p > *maxp) *maxp = p;
        ++nf;
    } while (n > 1);
    return nf;
}

void FN(execute)(const PLAN *p, const CPX *in, CPX *out);

PLAN *FN(make)(int n, int inverse)
{
    if (n <= 0) return NULL;
    PLAN *p = (PLAN *)calloc(1, sizeof(PLAN));
    p->n = n;
    p->inverse = inverse;
    int maxp = 1;
    FN(factorize)(n, p->factors, &maxp);
    const double sgn = inverse ? 1.0 : -1.0;
    if (maxp > 61) {
        /* Bluestein: X[k] = c[k] * sum_n (x[n] c[n]) conj(c[k-n]),  c[n] = exp(sgn*i*pi*n^2/n) */
        p->blue = 1;
        int m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->sub_fwd = FN(make)(m, 0);
        p->sub_inv = FN(make)(m, 1);
        p->chirp = (CPX *)malloc(sizeof(CPX) * n);
        p->bfilt = (CPX *)calloc(m, sizeof(CPX));
        p->work = (CPX *)malloc(sizeof(CPX) * m);
        for (int k = 0; k < n; ++k) {
            long long q = ((long long)k * k) % (2LL * n); /* exact n^2 mod 2N */
            double ang = sgn * M_PI * (double)q / (double)n;
            p->chirp[k] = OFFT_CMPLX((OFFT_R)cos(ang), (OFFT_R)sin(ang));
        }
        CPX *b = (CPX *)calloc(m, sizeof(CPX));
        for (int k = 0; k < n; ++k) {
            CPX c = OFFT_CMPLX(OFFT_CREAL(p->chirp[k]), -OFFT_CIMAG(p->chirp[k]));
            b[k] = c;
            if (k) b[m - k] = c;
        }
        FN(execute)(p->sub_fwd, b, p->bfilt);
        free(b);
        return p;
    }
    p->tw = (CPX *)malloc(sizeof(CPX) * n);
    for (int k = 0; k < n; ++k) {
        double ang = sgn * 2.0 * M_PI * (double)k / (double)n;
        p->tw[k] = OFFT_CMPLX((OFFT_R)cos(ang), (OFFT_R)sin(ang));
    }
    return p;
}

void FN(destroy)(PLAN *p)
{
    if (!p) return;
    free(p->tw);
    free(p->chirp);
    free(p->bfilt);
    free(p->work);
    if (p->sub_fwd) FN(destroy)(p->sub_fwd);
    if (p->sub_inv) FN(destroy)(p->sub_inv);
    free(p);
}

/* out-of-place, in != out required unless a scratch copy is made */
static void FN(execute_oop)(const PLAN *p, const CPX *in, CPX *out, CPX *blue_work)
{
    if (!p->blue) {
        CPX scratch[64]; /* generic radices are primes <= 61 (larger ones go to Bluestein) */
        FN(kf_work)(out, in, 1, 1, p->factors, p, scratch);
        return;
    }
    const int n = p->n, m = p->m;
    CPX *a = blue_work;            /* m */
    CPX *tmp = blue_work + m;      /* m */
    for (int k = 0; k < n; ++k) a[k] = in[k] * p->chirp[k];
    for (int k = n; k < m; ++k) a[k] = 0;
    FN(execute_oop)(p->sub_fwd, a, tmp, NULL);
    for (int k = 0; k < m; ++k) tmp[k] *= p->bfilt[k];
    FN(execute_oop)(p->sub_inv, tmp, a, NULL);
    const OFFT_R inv_m = (OFFT_R)(1.0 / (double)m);
    for (int k = 0; k < n; ++k) out[k] = a[k] * p->chirp[k] * inv_m;
}

/* in-place allowed (in == out) */
void FN(execute)(const PLAN *p, const CPX *in, CPX *out)
{
    if (p->blue) {
        CPX *work = (CPX *)malloc(sizeof(CPX) * 2 * p->m);
        FN(execute_oop)(p, in, out, work);
        free(work);
        return;
    }
    if (in == out) {
        CPX *tmp = (CPX *)malloc(sizeof(CPX) * p->n);
        memcpy(tmp, in, sizeof(CPX) * p->n);
        FN(execute_oop)(p, tmp, out, NULL);
        free(tmp);
    } else {
        FN(execute_oop)(p, in, out, NULL);
    }
}

/* Thread-safe execution with a caller-provided scratch of offt_scratch_len(p) elements. */
size_t FN(scratch_len)(const PLAN *p) { return p->blue ? (size_t)2 * p->m : (size_t)p->n; }

void FN(execute_ws)(const PLAN *p, const CPX *in, CPX *out, CPX *ws)
{
    if (p->blue) {
        FN(execute_oop)(p, in, out, ws);
        return;
    }
    if (in == out) {
        memcpy(ws, in, sizeof(CPX) * p->n);
        FN(execute_oop)(p, ws, out, NULL);
    } else {
        FN(execute_oop)(p, in, out, NULL);
    }
}

#undef CPX
#undef PLAN
#undef FN
