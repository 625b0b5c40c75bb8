// Exact correlation over short overlaps (GNCC and 1-bit on FP32 transforms).
//
// At a lag whose overlap m = N - |lag| is a small part of the window, the
// normalisation (1/sqrt(E1 E2) or 1/m) divides the transform's absolute rounding
// error by a small number.  Those lags are recomputed as exact FP64 sums over the
// overlap from the first and last K samples of every window, kept beside the
// spectra: edge[(station * Wall + window) * 2K + t], t < K: x[t]; t >= K:
// x[N - 2K + t].
#pragma once
#include <hip/hip_runtime.h>
#include "pairs.hpp"

namespace fpcc {

// sum over windows [w0, w1) of sum_n x1[n] x2[n + lag] (m = N - |lag| terms),
// x1 of station i, x2 of station j; signs: x1 >= 0 and x2 > 0 as +-1 (1-bit)
__device__ inline double edge_corr(const float *__restrict__ edge, int K, int Wall, int i, int j, int w0, int w1,
                                   int lag, int m, bool signs)
{
    const int K2 = 2 * K;
    double c = 0.0;
    for (int w = w0; w < w1; ++w) {
        const float *e1 = edge + (long long)(i * Wall + w) * K2;
        const float *e2 = edge + (long long)(j * Wall + w) * K2;
        // lag >= 0: x1[0, m) x2[N-m, N); lag < 0: x1[N-m, N) x2[0, m)
        const float *u = lag >= 0 ? e1 : e1 + K2 - m;
        const float *v = lag >= 0 ? e2 + K2 - m : e2;
        for (int t = 0; t < m; ++t) {
            double a = u[t], b = v[t];
            if (signs) {
                a = a >= 0.0 ? 1.0 : -1.0;
                b = b > 0.0 ? 1.0 : -1.0;
            }
            c += a * b;
        }
    }
    return c;
}

}  // namespace fpcc
