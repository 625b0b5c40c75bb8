// FP64 analytic signal, replaces AnalyticSignal (src/FFTapps.c:45-80):
//   y = IDFT_N( mask . DFT_N(x) ), mask = 1/N at DC, 2/N for 0 < k < (N+1)/2,
//   1/N at k = N/2 for even N, 0 above (the unnormalised inverse of the reference's
//   fftw_plan_dft_1d(FFTW_BACKWARD) with its da1 = 1/N folded into the mask).
// 7-smooth N <= 4096: one workgroup, FP64 Stockham FFTs in LDS.  Any other N: a
// direct FP64 DFT over the half spectrum the mask keeps, phases indexed exactly
// ((n k) mod N, stepped by integer additions) into a table of N FP64 roots.
#include <cmath>
#include "../../include/fastpcc.h"
#include "engine.hpp"
#include "profile.hpp"

namespace fpcc {
namespace {

__device__ __forceinline__ double analytic_mask(long long k, int N)
{
    const double da1 = 1.0 / (double)N;
    if (k == 0) return da1;
    if (k < (N + 1) / 2) return 2 * da1;
    if (!(N & 1) && k == N / 2) return da1;
    return 0.0;
}

__global__ void __launch_bounds__(512)
k_analytic64_lds(const double *__restrict__ x, int N, FftPlan plan, const cf64 *__restrict__ tw, cf64 *__restrict__ y)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf64 *b0 = reinterpret_cast<cf64 *>(smem_raw), *b1 = b0 + N;
    for (int n = threadIdx.x; n < N; n += blockDim.x) b0[n] = cf64{x[n], 0.0};
    __syncthreads();
    cf64 *X = fft_lds<double, -1>(b0, b1, plan, tw);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const double g = analytic_mask(k, N);
        X[k] = cf64{X[k].x * g, X[k].y * g};
    }
    __syncthreads();
    cf64 *z = fft_lds<double, +1>(X, X == b0 ? b1 : b0, plan, tw);
    for (int n = threadIdx.x; n < N; n += blockDim.x) y[n] = z[n];
}

// w[m] = exp(-2 pi i m / N)
__global__ void k_roots64(int N, cf64 *__restrict__ w)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)N, &s, &c);
    w[m] = cf64{c, s};
}

// Y[k] = mask_k sum_n x[n] w[(n k) mod N], k < K = N/2 + 1
__global__ void __launch_bounds__(256)
k_dft64_half(const double *__restrict__ x, int N, int K, const cf64 *__restrict__ w, cf64 *__restrict__ Y)
{
    __shared__ double xs[1024];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double re = 0.0, im = 0.0;
    int m = 0;  // (n k) mod N
    for (int n0 = 0; n0 < N; n0 += 1024) {
        const int cnt = min(1024, N - n0);
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) xs[i] = x[n0 + i];
        __syncthreads();
        if (k < K)
            for (int i = 0; i < cnt; ++i) {
                const cf64 t = w[m];
                re = fma(xs[i], t.x, re);
                im = fma(xs[i], t.y, im);
                m += k;
                if (m >= N) m -= N;
            }
    }
    if (k < K) {
        const double g = analytic_mask(k, N);
        Y[k] = cf64{re * g, im * g};
    }
}

// y[n] = sum_{k<K} Y[k] conj(w[(n k) mod N])
__global__ void __launch_bounds__(256)
k_idft64_half(const cf64 *__restrict__ Y, int N, int K, const cf64 *__restrict__ w, cf64 *__restrict__ y)
{
    __shared__ cf64 ys[512];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    double re = 0.0, im = 0.0;
    int m = 0;  // (n k) mod N
    for (int k0 = 0; k0 < K; k0 += 512) {
        const int cnt = min(512, K - k0);
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) ys[i] = Y[k0 + i];
        __syncthreads();
        if (n < N)
            for (int i = 0; i < cnt; ++i) {
                const cf64 t = w[m], v = ys[i];
                // v * conj(t)
                re = fma(v.x, t.x, fma(v.y, t.y, re));
                im = fma(v.y, t.x, fma(-v.x, t.y, im));
                m += n;
                if (m >= N) m -= N;
            }
    }
    if (n < N) y[n] = cf64{re, im};
}

struct Buf {
    void *p = nullptr;
    ~Buf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace fpcc

using namespace fpcc;

extern "C" int AnalyticSignal(double _Complex *y, double *x, unsigned int N)
{
    if (!x || !y || N < 1 || N > (1u << 30)) return -1;
    const int n = (int)N;
    Buf dx, dy;
    FPCC_CHECK(hipMalloc(&dx.p, sizeof(double) * n));
    FPCC_CHECK(hipMalloc(&dy.p, sizeof(cf64) * n));
    FPCC_CHECK(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
    const DevPlan *pl = n <= 4096 ? &get_plan(n, true) : nullptr;
    if (pl && pl->ok) {
        const size_t lds = 2 * sizeof(cf64) * (size_t)n;
        if (lds > 65536)
            (void)hipFuncSetAttribute((const void *)k_analytic64_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
        ProfScope prof("analytic64_lds", nullptr);
        hipLaunchKernelGGL(k_analytic64_lds, dim3(1), dim3(512), lds, nullptr, (const double *)dx.p, n, pl->plan,
                           (const cf64 *)pl->tw, (cf64 *)dy.p);
        FPCC_CHECK(hipGetLastError());
    } else {
        const int K = n / 2 + 1;
        Buf dw, dY;
        FPCC_CHECK(hipMalloc(&dw.p, sizeof(cf64) * n));
        FPCC_CHECK(hipMalloc(&dY.p, sizeof(cf64) * K));
        hipLaunchKernelGGL(k_roots64, dim3((n + 255) / 256), dim3(256), 0, nullptr, n, (cf64 *)dw.p);
        FPCC_CHECK(hipGetLastError());
        {
            ProfScope prof("analytic64_dft", nullptr);
            hipLaunchKernelGGL(k_dft64_half, dim3((K + 255) / 256), dim3(256), 0, nullptr, (const double *)dx.p, n,
                               K, (const cf64 *)dw.p, (cf64 *)dY.p);
            FPCC_CHECK(hipGetLastError());
        }
        {
            ProfScope prof("analytic64_idft", nullptr);
            hipLaunchKernelGGL(k_idft64_half, dim3((n + 255) / 256), dim3(256), 0, nullptr, (const cf64 *)dY.p, n,
                               K, (const cf64 *)dw.p, (cf64 *)dy.p);
            FPCC_CHECK(hipGetLastError());
        }
        FPCC_CHECK(hipMemcpy(y, dy.p, sizeof(cf64) * n, hipMemcpyDeviceToHost));
        return 0;
    }
    FPCC_CHECK(hipMemcpy(y, dy.p, sizeof(cf64) * n, hipMemcpyDeviceToHost));
    return 0;
}
