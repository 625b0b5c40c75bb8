// Pair index of the all-pairs layout: pair p of S stations is (i, j), i < j, in
// row-major order of the upper triangle, p = i S - i (i + 1) / 2 + (j - i - 1).
#pragma once
#include <hip/hip_runtime.h>

namespace fpcc {

__device__ __forceinline__ void pair_index(int p, int S, int &i, int &j)
{
    const double b = 2.0 * S - 1.0;
    int ii = (int)floor((b - sqrt(b * b - 8.0 * p)) * 0.5);  // the root of the row quadratic, then exact steps
    if (ii < 0) ii = 0;
    while (ii > 0 && ii * S - ii * (ii + 1) / 2 > p) --ii;
    while ((ii + 1) * S - (ii + 1) * (ii + 2) / 2 <= p) ++ii;
    i = ii;
    j = p - (ii * S - ii * (ii + 1) / 2) + ii + 1;
}

}  // namespace fpcc
