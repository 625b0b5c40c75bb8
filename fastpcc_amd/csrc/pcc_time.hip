// Time-domain PCC1 and PCC-v (src/FFTapps.c:251-278, 309-336, 475-783) on gfx950.
//
// For each lag, with a = phi2[n+lag] and b = phi1[n] over the overlap
// n in [max(0,-lag), min(N,N-lag)) (the lowlevels are called with the phasors of
// x2 first, src/FFTapps.c:598, :735):
//   PCC1:  y = sum(|a+b| - |a-b|) / (2 count)               (sqrtf terms)
//   PCC-v: y = sum(|a+b|^v - |a-b|^v) / (2^(v/2) count)     (powf(|.|^2, v/2) terms)
// FP32 terms; per 256-sample tile an FP32 partial sum, accumulated in double
// across tiles (the reference accumulates every term in double).
//
// Workgroup: 256 consecutive lags of one pair (one thread per lag) over one
// window (per-window output) or all windows (mean).  Each 256-sample tile of
// phi1 and the matching 511 samples of phi2 are staged in LDS; phi1[n] is read
// by the whole wave at once (broadcast).
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "pairs.hpp"

namespace fpcc {

constexpr int kPtLags = 256;
constexpr int kPtTile = 256;

// y rows start at out_off within rows of out_ld floats.
// Hardware square root and power (v_sqrt_f32; v_log_f32 and v_exp_f32, i.e.
// 2^(V log2 x)), about 1 ulp and a few ulp per term, as the reference's CUDA
// engine compiles its terms (fast math); the correctly rounded library versions
// cost several times more instructions.  Sums over N terms of either sign keep
// the output within 1e-8 of the correctly rounded terms.  x^V for V <= 0 keeps powf.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_pow(float x, float V)
{
    return __builtin_amdgcn_exp2f(V * __builtin_amdgcn_logf(x));
}

__global__ void __launch_bounds__(kPtLags)
k_pcc_time(const cf32 *__restrict__ phi, int S, int W, int N, int lag_first, int L, int pcc1, float V,
           double div_pow, int mode, float *__restrict__ y, int out_ld, int out_off)
{
    __shared__ cf32 s1[kPtTile];
    __shared__ cf32 s2[kPtTile + kPtLags - 1];
    const int lb = blockIdx.x;
    int p, w0, w1;
    if (mode == 0) {
        p = blockIdx.y / W;
        w0 = blockIdx.y % W;
        w1 = w0 + 1;
    } else {
        p = blockIdx.y;
        w0 = 0;
        w1 = W;
    }
    int i, j;
    pair_index(p, S, i, j);
    const int tid = threadIdx.x;
    const int lagA = lag_first + lb * kPtLags;
    const int lagB = min(lagA + kPtLags - 1, lag_first + L - 1);
    const int lag = lagA + tid;
    const bool valid = lag <= lagB;
    const int n_lo = max(0, -lag), n_hi = min(N, N - lag);
    const int nstart = max(0, -lagB), nend = min(N, N - lagA);
    float mean = 0.0f;
    for (int w = w0; w < w1; ++w) {
        const cf32 *p1 = phi + (long long)(i * W + w) * N;
        const cf32 *p2 = phi + (long long)(j * W + w) * N;
        double acc = 0.0;
        for (int n0 = nstart; n0 < nend; n0 += kPtTile) {
            __syncthreads();
            for (int t = tid; t < kPtTile; t += kPtLags) {
                const int n = n0 + t;
                s1[t] = (n < N) ? p1[n] : cf32{0.f, 0.f};
            }
            for (int t = tid; t < kPtTile + kPtLags - 1; t += kPtLags) {
                const int m = n0 + lagA + t;
                s2[t] = (m >= 0 && m < N) ? p2[m] : cf32{0.f, 0.f};
            }
            __syncthreads();
            if (!valid) continue;
            const int lo = max(n_lo, n0) - n0, hi = min(n_hi, n0 + kPtTile) - n0;
            float part = 0.0f;
            // full tiles (the common case) as a fixed-count loop the compiler
            // unrolls, so the LDS reads of several terms are in flight together
            auto run = [&](auto term) {
                if (lo == 0 && hi == kPtTile) {
#pragma unroll 16
                    for (int t = 0; t < kPtTile; ++t) part += term(s2[t + tid], s1[t]);
                } else {
                    for (int t = lo; t < hi; ++t) part += term(s2[t + tid], s1[t]);
                }
            };
            // PCC1, and PCC-v at v = 1 (|.|^2 to the power 1/2), take square roots; PCC-v
            // at v = 2 takes the squared magnitudes themselves (powf(x, 1) = x)
            if (pcc1 || V == 0.5f) {
                run([](cf32 a, cf32 b) {
                    const float sx = a.x + b.x, sy = a.y + b.y, dx = a.x - b.x, dy = a.y - b.y;
                    return hw_sqrt(sx * sx + sy * sy) - hw_sqrt(dx * dx + dy * dy);
                });
            } else if (V == 1.0f) {
                run([](cf32 a, cf32 b) {
                    const float sx = a.x + b.x, sy = a.y + b.y, dx = a.x - b.x, dy = a.y - b.y;
                    return (sx * sx + sy * sy) - (dx * dx + dy * dy);
                });
            } else if (V > 0.f) {
                run([V](cf32 a, cf32 b) {
                    const float sx = a.x + b.x, sy = a.y + b.y, dx = a.x - b.x, dy = a.y - b.y;
                    return hw_pow(sx * sx + sy * sy, V) - hw_pow(dx * dx + dy * dy, V);
                });
            } else {
                run([V](cf32 a, cf32 b) {
                    const float sx = a.x + b.x, sy = a.y + b.y, dx = a.x - b.x, dy = a.y - b.y;
                    return powf(sx * sx + sy * sy, V) - powf(dx * dx + dy * dy, V);
                });
            }
            acc += (double)part;
        }
        const int count = n_hi - n_lo;
        float v;
        if (pcc1) v = (float)acc / (float)(2 * count);  // src/FFTapps.c:276
        else v = (float)(acc / (div_pow * count));      // :334
        if (mode == 0) {
            if (valid) y[((long long)p * W + w) * out_ld + out_off + lb * kPtLags + tid] = v;
        } else {
            mean += v;
        }
    }
    if (mode != 0 && valid) y[(long long)p * out_ld + out_off + lb * kPtLags + tid] = mean / (float)W;
}

int launch_pcc_time(const cf32 *d_phi, int S, int W, int N, int lag_first, int L, double v, int pcc1, int mode,
                    float *d_y, int out_ld, int out_off, hipStream_t st)
{
    if (L <= 0) return 0;
    const int P = S * (S - 1) / 2;
    dim3 grid((L + kPtLags - 1) / kPtLags, mode == 0 ? P * W : P);
    const float V = (float)(v / 2);
    const double div_pow = pow(2.0, (double)V);
    ProfScope prof(pcc1 ? "pcc1_time" : "pccv_time", st);
    hipLaunchKernelGGL(k_pcc_time, grid, dim3(kPtLags), 0, st, d_phi, S, W, N, lag_first, L, pcc1, V, div_pow, mode,
                       d_y, out_ld, out_off);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc
