// Per-station preprocessing on gfx950: the steps the reference's driver runs on
// each station's traces before correlating them (here a trace is one window of a
// job's station row).
//
//   zero windows     src/ReadManySacs.c:529-559     a window of zeros is dropped
//   rms              src/PCC_fullpair_1b.c:355-363  sqrt(sum (float)(x*x) in FP64 / N)
//   clipping         src/PCC_fullpair_1b.c:881-891  |x| > 4 median|x| / 0.6745 -> +-threshold
//   outliers         src/PCC_fullpair_1b.c:893-925  rms > nstd * median rms -> dropped
//   whitening        src/PCC_fullpair_1b.c:1065-1177 (AveWhite), FP64
//   polarity         src/rotlib.c:14-34 as a per-window sign factor, applied first
//                    (every step is odd-symmetric, so the order does not change a bit)
//
// Kernels:
//   k_pre_window   one workgroup per window: the window (cached in LDS up to 8192
//                  samples), zero test, FP64 rms, exact median of |x| by a 4-pass
//                  radix select on the float bit patterns (+ one pass for the lower
//                  middle when N is even), clip threshold, signed/clipped write.
//   k_pre_valid    one workgroup per station: median of its windows' rms by rank
//                  counting, outlier rule.
//   k_pre_common   window kept only where every station kept it.
//   k_pre_zero     dropped windows are zeroed.
//   k_white_amp    per kept window: FP64 real FFT of the zero-padded window (half-
//                  length complex FFT in LDS), amplitude |X[k]| (hypot).
//   k_white_filter per station: average amplitude over its kept windows (in
//                  window order), the reference's filter design and 11-tap
//                  mirrored Blackman smoothing (bin 5 left at 0, as AveWhite leaves it).
//   k_white_apply  per kept window: FP64 real FFT, X *= HS / Nz, FP64 real
//                  inverse FFT, first N samples back as float.
// Windows with Nz > 8192 take the same steps with the half-length complex FFT
// (M = Nz/2 points) as an FP64 four-step transform through HBM: M = n1 * n2
// (n2 = 1024, n1 <= 256), k_w64_cols (n1-point column FFTs, 16 columns a
// workgroup, twiddles W_M^{c k1}) then k_w64_rows (n2-point row FFTs, 4 rows a
// workgroup) leave bin k = k1 + n1 k2 at k1 n2 + k2; the inverse runs rows,
// then columns, back to natural order.
#include <cmath>
#include "../../include/fastpcc.h"
#include "engine.hpp"
#include "profile.hpp"

namespace fpcc {
namespace {

constexpr int kPreThreads = 256;
constexpr int kPreCacheN = 8192;     // windows up to this length are cached in LDS
constexpr int kMaxOutlierW = 8192;   // windows per station (rank median in LDS)
constexpr int kWhiteMaxNz = 8192;    // FP64 whitening in one workgroup: Nz/2 complex points in LDS
constexpr int kWhiteLongMaxNz = 1 << 19;  // FP64 four-step whitening: M = Nz/2 <= 256 x 1024
constexpr int kW64Rows = 1024;       // row length of the four-step split
constexpr int kW64ColsPerWg = 16;
constexpr int kW64RowsPerWg = 4;

// Wave64 reductions (cross-lane shuffles), then one exchange through LDS between
// the waves: two barriers a reduction instead of one per tree level.  red holds
// one slot a wave; the leading barrier keeps the previous reduction's readers
// ahead of this one's writers.
template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, T *red, Op op)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = red[0];
    for (int i = 1; i < nw; ++i) r = op(r, red[i]);
    return r;
}
struct OpAdd {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
    __device__ unsigned operator()(unsigned a, unsigned b) const { return a > b ? a : b; }
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
struct OpMin {
    __device__ double operator()(double a, double b) const { return fmin(a, b); }
};

// Bit pattern of |v|: for non-negative floats the unsigned order is the value order.
__device__ __forceinline__ unsigned abs_bits(float v) { return __float_as_uint(v) & 0x7fffffffu; }

__global__ void __launch_bounds__(kPreThreads)
k_pre_window(float *__restrict__ x, long long stride, int W, int N, const float *__restrict__ sign, int clip,
             unsigned char *__restrict__ nzflag, float *__restrict__ rms, float *__restrict__ thr_out)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double redd[kPreThreads / 64];
    __shared__ int redi[kPreThreads / 64];
    __shared__ unsigned redu[kPreThreads / 64];
    __shared__ unsigned hist[256];
    __shared__ unsigned sel[2];
    float *cache = reinterpret_cast<float *>(smem_raw);
    const bool cached = N <= kPreCacheN;
    const int sw = blockIdx.x;
    const int s = sw / W, w = sw - s * W;
    float *px = x + (long long)s * stride + (long long)w * N;
    const float sg = sign ? sign[sw] : 1.0f;

    double part = 0.0;
    int nz = 0;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        const float v = px[n] * sg;
        if (cached) cache[n] = v;
        nz |= (v != 0.0f);
        part += (double)(v * v);
    }
    const double sumsq = block_reduce(part, redd, OpAdd());
    nz = block_reduce(nz, redi, OpAdd()) > 0;
    if (threadIdx.x == 0) {
        nzflag[sw] = (unsigned char)nz;
        rms[sw] = (float)sqrt(sumsq / (double)N);
    }
    float thr = 0.0f;
    const bool do_clip = clip && nz;
    if (do_clip) {
        // k-th smallest |x| (k = N/2): 8 bits at a time from the top
        int k = N / 2;
        unsigned prefix = 0, mask = 0;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0;
            __syncthreads();
            for (int n = threadIdx.x; n < N; n += blockDim.x) {
                const unsigned u = abs_bits(cached ? cache[n] : px[n]);
                if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (threadIdx.x < 64) {
                // the digit: the first bin whose cumulative count passes k, found by a
                // wave-wide scan of four bins a lane
                const int l = threadIdx.x;
                const unsigned h0 = hist[4 * l], h1 = hist[4 * l + 1], h2 = hist[4 * l + 2], h3 = hist[4 * l + 3];
                const unsigned tot = h0 + h1 + h2 + h3;
                unsigned inc = tot;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const unsigned o = __shfl_up(inc, off, 64);
                    if (l >= off) inc += o;
                }
                const unsigned exc = inc - tot, ku = (unsigned)k;
                if (exc <= ku && ku < inc) {
                    unsigned below = exc;
                    int d = 4 * l;
                    if (below + h0 <= ku) {
                        below += h0;
                        ++d;
                        if (below + h1 <= ku) {
                            below += h1;
                            ++d;
                            if (below + h2 <= ku) {
                                below += h2;
                                ++d;
                            }
                        }
                    }
                    sel[0] = (unsigned)d;
                    sel[1] = ku - below;
                }
            }
            __syncthreads();
            prefix |= sel[0] << shift;
            mask |= 255u << shift;
            k = (int)sel[1];
            __syncthreads();
        }
        const float vk = __uint_as_float(prefix);
        float med = vk;
        if (!(N & 1)) {
            // lower middle: vk again if it occurs before position N/2, else the
            // largest value below vk
            int less = 0;
            unsigned mless = 0;
            for (int n = threadIdx.x; n < N; n += blockDim.x) {
                const unsigned u = abs_bits(cached ? cache[n] : px[n]);
                if (u < prefix) {
                    ++less;
                    mless = max(mless, u);
                }
            }
            less = block_reduce(less, redi, OpAdd());
            mless = block_reduce(mless, redu, OpMax());
            const float vk1 = less <= N / 2 - 1 ? vk : __uint_as_float(mless);
            med = (vk + vk1) / 2;
        }
        thr = (float)((double)(4.0f * med) / 0.6745);
        if (threadIdx.x == 0 && thr_out) thr_out[sw] = thr;
    } else if (threadIdx.x == 0 && thr_out) {
        thr_out[sw] = 0.0f;
    }
    if (do_clip || sign) {
        for (int n = threadIdx.x; n < N; n += blockDim.x) {
            float v = cached ? cache[n] : px[n] * sg;
            if (do_clip && fabsf(v) > thr) v = v > thr ? thr : -thr;
            px[n] = v;
        }
    }
}

// One workgroup per station: the median of the rms of its non-zero windows (by
// rank: rank_i = #{v_j < v_i} + #{j < i, v_j == v_i}), the outlier rule.
__global__ void __launch_bounds__(kPreThreads)
k_pre_valid(const unsigned char *__restrict__ nzflag, const float *__restrict__ rms, int W, float nstd,
            unsigned char *__restrict__ valid)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int redi[kPreThreads / 64];
    __shared__ float midv[2];
    float *v = reinterpret_cast<float *>(smem_raw);
    unsigned char *f = reinterpret_cast<unsigned char *>(v + W);
    const int s = blockIdx.x;
    int cnt = 0;
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
        v[w] = rms[(long long)s * W + w];
        f[w] = nzflag[(long long)s * W + w];
        cnt += f[w];
    }
    const int n = block_reduce(cnt, redi, OpAdd());
    float thr = 0.0f;
    const bool rule = nstd > 0.0f && n > 0;
    if (rule) {
        const int k = n / 2;
        for (int i = threadIdx.x; i < W; i += blockDim.x) {
            if (!f[i]) continue;
            const float vi = v[i];
            int r = 0;
            for (int j = 0; j < W; ++j)
                if (f[j]) r += (v[j] < vi) || (v[j] == vi && j < i);
            if (r == k) midv[0] = vi;
            if (r == k - 1) midv[1] = vi;
        }
        __syncthreads();
        const float med = (n & 1) ? midv[0] : (midv[0] + midv[1]) / 2;
        thr = nstd * med;
    }
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
        unsigned char ok = f[w];
        if (rule && ok && v[w] > thr) ok = 0;
        valid[(long long)s * W + w] = ok;
    }
}

__global__ void k_pre_common(unsigned char *__restrict__ valid, int S, int W)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    unsigned char all = 1;
    for (int s = 0; s < S; ++s) all &= valid[(long long)s * W + w];
    for (int s = 0; s < S; ++s) valid[(long long)s * W + w] = all;
}

__global__ void k_pre_zero(float *__restrict__ x, long long stride, int W, int N,
                           const unsigned char *__restrict__ valid)
{
    const int sw = blockIdx.x;
    if (valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    float *px = x + (long long)s * stride + (long long)w * N;
    for (int n = threadIdx.x; n < N; n += blockDim.x) px[n] = 0.0f;
}

// FP64 half spectrum X[0..M] of the window zero-padded to Nz = 2M, from the
// half-length complex FFT z[m] = x[2m] + i x[2m+1]; returns Z (in LDS).
__device__ cf64 *white_fft(const float *__restrict__ px, int N, int M, const FftPlan &planM,
                           const cf64 *__restrict__ twM, cf64 *buf0, cf64 *buf1)
{
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const int n0 = 2 * m, n1 = 2 * m + 1;
        buf0[m] = cf64{n0 < N ? (double)px[n0] : 0.0, n1 < N ? (double)px[n1] : 0.0};
    }
    __syncthreads();
    return fft_lds<double, -1>(buf0, buf1, planM, twM);
}

__device__ __forceinline__ cf64 half_bin(const cf64 *Z, int k, int M, const cf64 *__restrict__ twc)
{
    const cf64 zk = Z[k == M ? 0 : k];
    const cf64 zm = Z[k == 0 ? 0 : M - k];
    const cf64 zc = cf64{zm.x, -zm.y};
    const cf64 ev = cf64{(zk.x + zc.x) * 0.5, (zk.y + zc.y) * 0.5};
    const cf64 d = cf64{zk.x - zc.x, zk.y - zc.y};
    const cf64 od = cf64{d.y * 0.5, -d.x * 0.5};
    const cf64 wk = (k == M) ? cf64{-1.0, 0.0} : cf64{twc[k].x, -twc[k].y};
    return ev + cmul(wk, od);
}

__global__ void __launch_bounds__(512)
k_white_amp(const float *__restrict__ x, long long stride, int W, int N, int M, FftPlan planM,
            const cf64 *__restrict__ twM, const cf64 *__restrict__ twc, const unsigned char *__restrict__ valid,
            double *__restrict__ A)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    cf64 *buf0 = reinterpret_cast<cf64 *>(smem_raw), *buf1 = buf0 + M;
    const cf64 *Z = white_fft(x + (long long)s * stride + (long long)w * N, N, M, planM, twM, buf0, buf1);
    double *a = A + (long long)sw * (M + 1);
    for (int k = threadIdx.x; k <= M; k += blockDim.x) {
        const cf64 X = half_bin(Z, k, M, twc);
        a[k] = hypot(X.x, X.y);
    }
}

// Windows of 2049..4096 samples (Nz = 4096, M = 2048; the hourly windows at
// 1 Hz): the same two kernels with compile-time FP64 transforms in place in one
// padded LDS buffer (35 KB instead of two 32 KB buffers: four workgroups a CU
// instead of two, and no runtime radix dispatch), radix order 4 x 8 x 8 x 8 with
// its own FP64 twiddle table.  Bin k sits at lds[ct_pad(k)].
constexpr int kW2M = 2048, kW2Threads = 256;

__device__ __forceinline__ void white_fft_2048(const float *__restrict__ px, int N, cf64 *lds,
                                               const cf64 *__restrict__ twM)
{
    for (int m = threadIdx.x; m < kW2M; m += kW2Threads) {
        const int n0 = 2 * m, n1 = 2 * m + 1;
        lds[ct_pad(m)] = cf64{n0 < N ? (double)px[n0] : 0.0, n1 < N ? (double)px[n1] : 0.0};
    }
    __syncthreads();
    ct_fft<double, kW2M, kW2Threads, -1, 4, 8, 8, 8>(lds, twM);
}

__device__ __forceinline__ cf64 half_bin_ct(const cf64 *Z, int k, const cf64 *__restrict__ twc)
{
    constexpr int M = kW2M;
    const cf64 zk = Z[ct_pad(k == M ? 0 : k)];
    const cf64 zm = Z[ct_pad(k == 0 ? 0 : M - k)];
    const cf64 zc = cf64{zm.x, -zm.y};
    const cf64 ev = cf64{(zk.x + zc.x) * 0.5, (zk.y + zc.y) * 0.5};
    const cf64 d = cf64{zk.x - zc.x, zk.y - zc.y};
    const cf64 od = cf64{d.y * 0.5, -d.x * 0.5};
    const cf64 wk = (k == M) ? cf64{-1.0, 0.0} : cf64{twc[k].x, -twc[k].y};
    return ev + cmul(wk, od);
}

__global__ void __launch_bounds__(kW2Threads)
k_white_amp_2048(const float *__restrict__ x, long long stride, int W, int N, const cf64 *__restrict__ twM,
                 const cf64 *__restrict__ twc, const unsigned char *__restrict__ valid, double *__restrict__ A)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    cf64 *lds = reinterpret_cast<cf64 *>(smem_raw);
    white_fft_2048(x + (long long)s * stride + (long long)w * N, N, lds, twM);
    double *a = A + (long long)sw * (kW2M + 1);
    for (int k = threadIdx.x; k <= kW2M; k += kW2Threads) {
        const cf64 X = half_bin_ct(lds, k, twc);
        a[k] = hypot(X.x, X.y);
    }
}

__global__ void __launch_bounds__(kW2Threads)
k_white_apply_2048(float *__restrict__ x, long long stride, int W, int N, const cf64 *__restrict__ twM,
                   const cf64 *__restrict__ twc, const unsigned char *__restrict__ valid,
                   const double *__restrict__ HS)
{
    constexpr int M = kW2M;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    float *px = x + (long long)s * stride + (long long)w * N;
    cf64 *lds = reinterpret_cast<cf64 *>(smem_raw);
    white_fft_2048(px, N, lds, twM);
    const double *hs = HS + (long long)s * (M + 1);
    const double da1 = 1. / (double)(2 * M);
    // filtered bins, then the c2r packing F[k] = P + i e^{2 pi i k/Nz} Q, in place:
    // the thread of pair {k, M-k} reads only Z[k] and Z[M-k] and writes F[k], F[M-k]
    auto pack = [&](cf64 a, cf64 c, int k) {
        const cf64 P = cf64{a.x + c.x, a.y - c.y};
        const cf64 Q = cf64{a.x - c.x, a.y + c.y};
        const cf64 wq = cmul(twc[k], Q);
        return cf64{P.x - wq.y, P.y + wq.x};
    };
    for (int k = threadIdx.x; k <= M / 2; k += kW2Threads) {
        cf64 a = half_bin_ct(lds, k, twc), c = half_bin_ct(lds, M - k, twc);
        const double ga = da1 * hs[k], gc = da1 * hs[M - k];
        a = cf64{a.x * ga, a.y * ga};
        c = cf64{c.x * gc, c.y * gc};
        if (k == 0) {  // c2r ignores the imaginary parts of DC and Nyquist
            a.y = 0.0;
            c.y = 0.0;
        }
        const cf64 fk = pack(a, c, k);
        if (k > 0 && k < M / 2) lds[ct_pad(M - k)] = pack(c, a, M - k);
        lds[ct_pad(k)] = fk;
    }
    __syncthreads();
    ct_fft<double, M, kW2Threads, +1, 4, 8, 8, 8>(lds, twM);
    for (int n = threadIdx.x; n < N; n += kW2Threads) {
        const cf64 v = lds[ct_pad(n >> 1)];
        px[n] = (float)((n & 1) ? v.y : v.x);
    }
}

// AveWhite's filter for station s from its kept windows' amplitudes (one
// workgroup a station: 1024 threads keep enough of the window average's loads in
// flight).
constexpr int kWhiteFilterThreads = 1024;
__global__ void __launch_bounds__(kWhiteFilterThreads)
k_white_filter(const double *__restrict__ A, const unsigned char *__restrict__ valid, int W, int M, int Nz,
               int N, int n1, int n2, double *__restrict__ Hs, double *__restrict__ HS)
{
    __shared__ double redd[kWhiteFilterThreads / 64];
    __shared__ double win[11];
    __shared__ int redi[kWhiteFilterThreads / 64];
    const int s = blockIdx.x;
    const int Nh = M + 1;
    double *H = Hs + (long long)s * Nh;  // this station's scratch row
    const unsigned char *vs = valid + (long long)s * W;
    int cnt = 0;
    for (int w = threadIdx.x; w < W; w += blockDim.x) cnt += vs[w];
    const int Tr = block_reduce(cnt, redi, OpAdd());
    if (Tr == 0) return;
    const double rTr = 1. / (double)Tr;
    for (int k = threadIdx.x; k < Nh; k += blockDim.x) {
        double h = 0.0;
        for (int w = 0; w < W; ++w)
            if (vs[w]) h += A[((long long)s * W + w) * Nh + k];
        H[k] = h * rTr;
    }
    if (threadIdx.x == 0) {
        double t[11], sum = 0;
        for (int m = 0; m < 11; ++m) {
            const double a = 2 * 3.14159265358979323846 * m / (N - 1);
            t[m] = 0.42 - 0.5 * cos(a) + 0.08 * cos(2 * a);
        }
        for (int m = 0; m < 11; ++m) sum += t[m];
        const double r = 1 / sum;
        for (int m = 0; m < 11; ++m) win[m] = t[m] * r;
    }
    __syncthreads();
    const double in1 = 1. / H[n1], in2 = 1. / H[n2];
    __syncthreads();
    for (int k = threadIdx.x; k < Nh; k += blockDim.x) H[k] = k < n1 ? in1 : (k < n2 ? 1. / H[k] : in2);
    __syncthreads();
    // minimum and maximum gain over [n1, n2) (H[n1] when the band is empty)
    double mn = H[n1], mx = H[n1];
    for (int k = n1 + 1 + threadIdx.x; k < n2; k += blockDim.x) {
        mn = fmin(mn, H[k]);
        mx = fmax(mx, H[k]);
    }
    mn = block_reduce(mn, redd, OpMin());
    mx = block_reduce(mx, redd, OpMax());
    if (mx > 1000 * mn) mx = 1000 * mn;
    const double scale = 1 / (mn * (double)Nz);
    for (int k = threadIdx.x; k < Nh; k += blockDim.x) {
        double h = H[k];
        if (h > mx) h = mx;
        H[k] = h * scale;
    }
    __syncthreads();
    // HS[n-5] = sum_m win[m] H[n-m], mirrored at 0 and at Nh-1; bin 5 is not written
    double *hs = HS + (long long)s * Nh;
    for (int c = threadIdx.x; c < Nh; c += blockDim.x) {
        const int n = c + 5;
        double acc = 0.0;
        if (n == 10) {
            hs[c] = 0.0;
            continue;
        }
        if (n < 10) {
            int m = 0;
            for (; m <= n; ++m) acc += win[m] * H[n - m];
            for (; m < 11; ++m) acc += win[m] * H[m - n];
        } else if (n < Nh) {
            for (int m = 0; m < 11; ++m) acc += win[m] * H[n - m];
        } else {
            const int Mm = n - (Nh - 1);
            int m = 0;
            for (; m < Mm; ++m) acc += win[m] * H[Nh - 1 + m - Mm];
            for (; m < 11; ++m) acc += win[m] * H[n - m];
        }
        hs[c] = acc;
    }
}

__global__ void __launch_bounds__(512)
k_white_apply(float *__restrict__ x, long long stride, int W, int N, int M, FftPlan planM,
              const cf64 *__restrict__ twM, const cf64 *__restrict__ twc, const unsigned char *__restrict__ valid,
              const double *__restrict__ HS)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    float *px = x + (long long)s * stride + (long long)w * N;
    cf64 *buf0 = reinterpret_cast<cf64 *>(smem_raw), *buf1 = buf0 + M;
    cf64 *Z = white_fft(px, N, M, planM, twM, buf0, buf1);
    cf64 *F = Z == buf0 ? buf1 : buf0;
    const double *hs = HS + (long long)s * (M + 1);
    const double da1 = 1. / (double)(2 * M);
    // filtered bins k and M-k, then the c2r packing F[k] = P + i e^{2 pi i k/Nz} Q
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        cf64 a = half_bin(Z, k, M, twc), c = half_bin(Z, M - k, M, twc);
        const double ga = da1 * hs[k], gc = da1 * hs[M - k];
        a = cf64{a.x * ga, a.y * ga};
        c = cf64{c.x * gc, c.y * gc};
        if (k == 0) {  // c2r ignores the imaginary parts of DC and Nyquist
            a.y = 0.0;
            c.y = 0.0;
        }
        const cf64 P = cf64{a.x + c.x, a.y - c.y};
        const cf64 Q = cf64{a.x - c.x, a.y + c.y};
        const cf64 wq = cmul(twc[k], Q);
        F[k] = cf64{P.x - wq.y, P.y + wq.x};
    }
    __syncthreads();
    const cf64 *z = fft_lds<double, +1>(F, Z, planM, twM);
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        const cf64 v = z[n >> 1];
        px[n] = (float)((n & 1) ? v.y : v.x);
    }
}

// ---- FP64 four-step transform of M = n1 * n2 points, batch row sw at sw * M ----

__device__ __forceinline__ long long w64_pos(long long k, int n1, int n2) { return (k % n1) * n2 + k / n1; }

// z[m] = x[2m] + i x[2m+1] of the zero-padded window (kept windows only)
__global__ void k_w64_pack(const float *__restrict__ x, long long stride, int W, int N, int M,
                           const unsigned char *__restrict__ valid, cf64 *__restrict__ Z)
{
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    const float *px = x + (long long)s * stride + (long long)w * N;
    cf64 *z = Z + (long long)sw * M;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const int a = 2 * m, b = 2 * m + 1;
        z[m] = cf64{a < N ? (double)px[a] : 0.0, b < N ? (double)px[b] : 0.0};
    }
}

// Column pass: for the 16 columns c0.. of transform sw, Y[k1][c] = FFT_n1 over r
// of data[r n2 + c].  Forward: times W_M^{c k1} on the way out; inverse: times
// W_M^{-c k1} on the way in (the twiddle of the preceding row pass).
template <int DIR>
__global__ void __launch_bounds__(256)
k_w64_cols(cf64 *__restrict__ data, int n1, int n2, FftPlan p1, const cf64 *__restrict__ tw1,
           const cf64 *__restrict__ twM, const unsigned char *__restrict__ valid)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const int ld = n1 + 1;
    const long long M = (long long)n1 * n2;
    cf64 *buf0 = reinterpret_cast<cf64 *>(smem_raw), *buf1 = buf0 + kW64ColsPerWg * ld;
    cf64 *d = data + (long long)sw * M;
    const int c0 = blockIdx.x * kW64ColsPerWg;
    for (int idx = threadIdx.x; idx < kW64ColsPerWg * n1; idx += blockDim.x) {
        const int t = idx % kW64ColsPerWg, r = idx / kW64ColsPerWg, c = c0 + t;
        cf64 v = d[(long long)r * n2 + c];
        if (DIR > 0) {
            const cf64 w = twM[((long long)c * r) % M];
            v = cmul(v, cf64{w.x, -w.y});
        }
        buf0[t * ld + r] = v;
    }
    __syncthreads();
    cf64 *res = fft_lds_batch<double, DIR>(buf0, buf1, p1, tw1, kW64ColsPerWg, ld);
    for (int idx = threadIdx.x; idx < kW64ColsPerWg * n1; idx += blockDim.x) {
        const int t = idx % kW64ColsPerWg, k1 = idx / kW64ColsPerWg, c = c0 + t;
        cf64 v = res[t * ld + k1];
        if (DIR < 0) v = cmul(v, twM[((long long)c * k1) % M]);
        d[(long long)k1 * n2 + c] = v;
    }
}

// Row pass: n2-point FFTs of rows r0..r0+3 of transform sw, in place.
template <int DIR>
__global__ void __launch_bounds__(256)
k_w64_rows(cf64 *__restrict__ data, int n1, int n2, FftPlan p2, const cf64 *__restrict__ tw2,
           const unsigned char *__restrict__ valid)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const long long M = (long long)n1 * n2;
    cf64 *buf0 = reinterpret_cast<cf64 *>(smem_raw), *buf1 = buf0 + kW64RowsPerWg * n2;
    cf64 *d = data + (long long)sw * M + (long long)blockIdx.x * kW64RowsPerWg * n2;
    for (int idx = threadIdx.x; idx < kW64RowsPerWg * n2; idx += blockDim.x) buf0[idx] = d[idx];
    __syncthreads();
    cf64 *res = fft_lds_batch<double, DIR>(buf0, buf1, p2, tw2, kW64RowsPerWg, n2);
    for (int idx = threadIdx.x; idx < kW64RowsPerWg * n2; idx += blockDim.x) d[idx] = res[idx];
}

// Bin k of the real window's spectrum from the permuted half-length transform.
__device__ __forceinline__ cf64 half_bin_perm(const cf64 *Z, long long k, int M, int n1, int n2,
                                              const cf64 *__restrict__ twc)
{
    const cf64 zk = Z[w64_pos(k == M ? 0 : k, n1, n2)];
    const cf64 zm = Z[w64_pos(k == 0 ? 0 : M - k, n1, n2)];
    const cf64 zc = cf64{zm.x, -zm.y};
    const cf64 ev = cf64{(zk.x + zc.x) * 0.5, (zk.y + zc.y) * 0.5};
    const cf64 dd = cf64{zk.x - zc.x, zk.y - zc.y};
    const cf64 od = cf64{dd.y * 0.5, -dd.x * 0.5};
    const cf64 wk = (k == M) ? cf64{-1.0, 0.0} : cf64{twc[k].x, -twc[k].y};
    return ev + cmul(wk, od);
}

__global__ void k_w64_amp(const cf64 *__restrict__ Z, int M, int n1, int n2, const cf64 *__restrict__ twc,
                          const unsigned char *__restrict__ valid, double *__restrict__ A)
{
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const cf64 *z = Z + (long long)sw * M;
    double *a = A + (long long)sw * (M + 1);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= M; k += gridDim.x * blockDim.x) {
        const cf64 X = half_bin_perm(z, k, M, n1, n2, twc);
        a[k] = hypot(X.x, X.y);
    }
}

// Filtered bins, packed for the half-length inverse, stored in permuted order.
__global__ void k_w64_filter(const cf64 *__restrict__ Z, int M, int n1, int n2, const cf64 *__restrict__ twc,
                             const unsigned char *__restrict__ valid, int W, const double *__restrict__ HS,
                             cf64 *__restrict__ F)
{
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const cf64 *z = Z + (long long)sw * M;
    cf64 *f = F + (long long)sw * M;
    const double *hs = HS + (long long)(sw / W) * (M + 1);
    const double da1 = 1. / (double)(2 * M);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < M; k += gridDim.x * blockDim.x) {
        cf64 a = half_bin_perm(z, k, M, n1, n2, twc), c = half_bin_perm(z, M - k, M, n1, n2, twc);
        const double ga = da1 * hs[k], gc = da1 * hs[M - k];
        a = cf64{a.x * ga, a.y * ga};
        c = cf64{c.x * gc, c.y * gc};
        if (k == 0) {
            a.y = 0.0;
            c.y = 0.0;
        }
        const cf64 P = cf64{a.x + c.x, a.y - c.y};
        const cf64 Q = cf64{a.x - c.x, a.y + c.y};
        const cf64 wq = cmul(twc[k], Q);
        f[w64_pos(k, n1, n2)] = cf64{P.x - wq.y, P.y + wq.x};
    }
}

__global__ void k_w64_unpack(const cf64 *__restrict__ F, int M, float *__restrict__ x, long long stride, int W,
                             int N, const unsigned char *__restrict__ valid)
{
    const int sw = blockIdx.y;
    if (!valid[sw]) return;
    const int s = sw / W, w = sw - s * W;
    float *px = x + (long long)s * stride + (long long)w * N;
    const cf64 *z = F + (long long)sw * M;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
        const cf64 v = z[n >> 1];
        px[n] = (float)((n & 1) ? v.y : v.x);
    }
}

void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

struct PreDims {
    int S, W, N, M, Nz, n1, n2;
    long long stride;
    bool white, lng;  // lng: four-step transforms (Nz > 8192)
    int c1, c2;       // four-step split M = c1 * c2
};

int pre_dims(const fpcc_job *job, const fpcc_preproc *pp, PreDims &d)
{
    if (!job || !pp || job->n_stations < 1 || job->n_windows < 1 || job->win_len < 2) return -1;
    if (job->station_stride < (unsigned long long)job->n_windows * job->win_len) return -1;
    if (job->n_stations > (1u << 20) || job->n_windows > (1u << 20) ||
        (unsigned long long)job->n_stations * job->n_windows > 65535ull * 1024)
        return -1;
    d.S = (int)job->n_stations;
    d.W = (int)job->n_windows;
    d.N = (int)job->win_len;
    d.stride = (long long)job->station_stride;
    if (d.W > kMaxOutlierW) {
        fprintf(stderr, "fastpcc: preprocessing takes at most %d windows per station\n", kMaxOutlierW);
        return -2;
    }
    d.white = pp->band[0] > 0 && pp->band[1] > pp->band[0];
    d.Nz = (int)(1u << (unsigned)ceil(log2((double)d.N)));
    d.M = d.Nz / 2;
    d.n1 = d.n2 = 0;
    d.lng = false;
    d.c1 = d.c2 = 0;
    if (d.white) {
        if (!(pp->dt > 0)) return -1;
        if (d.Nz > kWhiteLongMaxNz || d.Nz < 32) {
            fprintf(stderr, "fastpcc: whitening takes windows of 17 to %d samples\n", kWhiteLongMaxNz);
            return -2;
        }
        d.lng = d.Nz > kWhiteMaxNz;
        if (d.lng) {
            d.c2 = kW64Rows;
            d.c1 = d.M / kW64Rows;
            if ((size_t)d.S * d.W > 65535) {
                fprintf(stderr, "fastpcc: whitening of windows over %d samples takes at most 65535 windows\n",
                        kWhiteMaxNz);
                return -2;
            }
        }
        const double a = pp->band[0] * pp->dt * (double)d.Nz, b = pp->band[1] * pp->dt * (double)d.Nz;
        if (!(b < (double)(d.M + 1))) {
            fprintf(stderr, "fastpcc: whitening band reaches the Nyquist frequency\n");
            return -1;
        }
        d.n1 = (int)(unsigned)a;
        d.n2 = (int)(unsigned)b;
    }
    return 0;
}

// Workspace regions, each 256-byte aligned: zero flags [SW], clip thresholds
// [SW], amplitudes [SW][M+1], filters [S][M+1], filter scratch [S][M+1], and
// for four-step whitening the spectra and filtered spectra [SW][M].
struct PreLayout {
    size_t nz, thr, A, HS, Hs, Z, F, total;
};

PreLayout pre_layout(const PreDims &d)
{
    const size_t SW = (size_t)d.S * d.W;
    PreLayout l{};
    auto take = [&](size_t bytes) {
        const size_t at = l.total;
        l.total += (bytes + 255) / 256 * 256;
        return at;
    };
    l.nz = take(SW);
    l.thr = take(SW * sizeof(float));
    if (d.white) {
        l.A = take(sizeof(double) * SW * (size_t)(d.M + 1));
        l.HS = take(sizeof(double) * (size_t)d.S * (d.M + 1));
        l.Hs = take(sizeof(double) * (size_t)d.S * (d.M + 1));
    }
    if (d.lng) {
        l.Z = take(sizeof(cf64) * SW * (size_t)d.M);
        l.F = take(sizeof(cf64) * SW * (size_t)d.M);
    }
    return l;
}

size_t pre_workspace(const PreDims &d) { return pre_layout(d).total; }

int white_filter(const PreDims &d, const double *A, const unsigned char *valid, double *Hs, double *HS,
                 hipStream_t st)
{
    ProfScope prof("white_filter", st);
    hipLaunchKernelGGL(k_white_filter, dim3(d.S), dim3(kWhiteFilterThreads), 0, st, A, valid, d.W, d.M, d.Nz, d.N, d.n1, d.n2,
                       Hs, HS);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

int preprocess_device(const fpcc_job *job, const fpcc_preproc *pp, float *d_x, const float *d_sign,
                      unsigned char *d_valid, float *d_rms, void *d_ws, size_t ws_bytes, hipStream_t st)
{
    PreDims d;
    int r = pre_dims(job, pp, d);
    if (r) return r;
    if (!d_x || !d_valid || !d_rms || !d_ws || ws_bytes < pre_workspace(d)) return -1;
    const size_t SW = (size_t)d.S * d.W;
    const PreLayout l = pre_layout(d);
    char *ws = (char *)d_ws;
    unsigned char *nz = (unsigned char *)(ws + l.nz);
    float *thr = (float *)(ws + l.thr);
    double *A = (double *)(ws + l.A), *HS = (double *)(ws + l.HS), *Hs = (double *)(ws + l.Hs);
    cf64 *Zl = (cf64 *)(ws + l.Z), *Fl = (cf64 *)(ws + l.F);
    {
        ProfScope prof("pre_window", st);
        const size_t lds = d.N <= kPreCacheN ? sizeof(float) * (size_t)d.N : 0;
        allow_lds((const void *)k_pre_window, lds);
        hipLaunchKernelGGL(k_pre_window, dim3((unsigned)SW), dim3(kPreThreads), lds, st, d_x, d.stride, d.W, d.N,
                           d_sign, pp->clip ? 1 : 0, nz, d_rms, thr);
        FPCC_CHECK(hipGetLastError());
    }
    {
        ProfScope prof("pre_valid", st);
        const size_t lds = (sizeof(float) + 1) * (size_t)d.W + 16;
        allow_lds((const void *)k_pre_valid, lds);
        hipLaunchKernelGGL(k_pre_valid, dim3(d.S), dim3(kPreThreads), lds, st, nz, d_rms, d.W, pp->std_max, d_valid);
        FPCC_CHECK(hipGetLastError());
        if (pp->common_windows) {
            hipLaunchKernelGGL(k_pre_common, dim3((d.W + 255) / 256), dim3(256), 0, st, d_valid, d.S, d.W);
            FPCC_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_pre_zero, dim3((unsigned)SW), dim3(256), 0, st, d_x, d.stride, d.W, d.N, d_valid);
        FPCC_CHECK(hipGetLastError());
    }
    if (d.white && !d.lng && d.M == kW2M) {
        const void *twc = get_c2r_twiddles(d.Nz, true), *twW = get_order_twiddles(kW2M, {4, 8, 8, 8}, true);
        if (!twc || !twW) return -1;
        const size_t lds = sizeof(cf64) * ct_lds_elems<kW2M>();
        {
            ProfScope prof("white_amp", st);
            allow_lds((const void *)k_white_amp_2048, lds);
            hipLaunchKernelGGL(k_white_amp_2048, dim3((unsigned)SW), dim3(kW2Threads), lds, st, (const float *)d_x,
                               d.stride, d.W, d.N, (const cf64 *)twW, (const cf64 *)twc,
                               (const unsigned char *)d_valid, A);
            FPCC_CHECK(hipGetLastError());
        }
        if ((r = white_filter(d, A, d_valid, Hs, HS, st))) return r;
        ProfScope prof("white_apply", st);
        allow_lds((const void *)k_white_apply_2048, lds);
        hipLaunchKernelGGL(k_white_apply_2048, dim3((unsigned)SW), dim3(kW2Threads), lds, st, d_x, d.stride, d.W,
                           d.N, (const cf64 *)twW, (const cf64 *)twc, (const unsigned char *)d_valid,
                           (const double *)HS);
        FPCC_CHECK(hipGetLastError());
    } else if (d.white && !d.lng) {
        const DevPlan &pM = get_plan(d.M, true);
        const void *twc = get_c2r_twiddles(d.Nz, true);
        if (!pM.ok || !twc) return -1;
        const size_t lds = sizeof(cf64) * 2 * (size_t)d.M;
        {
            ProfScope prof("white_amp", st);
            allow_lds((const void *)k_white_amp, lds);
            hipLaunchKernelGGL(k_white_amp, dim3((unsigned)SW), dim3(512), lds, st, (const float *)d_x, d.stride,
                               d.W, d.N, d.M, pM.plan, (const cf64 *)pM.tw, (const cf64 *)twc,
                               (const unsigned char *)d_valid, A);
            FPCC_CHECK(hipGetLastError());
        }
        if ((r = white_filter(d, A, d_valid, Hs, HS, st))) return r;
        {
            ProfScope prof("white_apply", st);
            allow_lds((const void *)k_white_apply, lds);
            hipLaunchKernelGGL(k_white_apply, dim3((unsigned)SW), dim3(512), lds, st, d_x, d.stride, d.W, d.N, d.M,
                               pM.plan, (const cf64 *)pM.tw, (const cf64 *)twc, (const unsigned char *)d_valid,
                               (const double *)HS);
            FPCC_CHECK(hipGetLastError());
        }
    } else if (d.white) {
        const DevPlan &q1 = get_plan(d.c1, true), &q2 = get_plan(d.c2, true);
        const cf64 *twM = (const cf64 *)get_full_twiddles(d.M, true);
        const cf64 *twc = (const cf64 *)get_c2r_twiddles(d.Nz, true);
        if (!q1.ok || !q2.ok || !twM || !twc) return -1;
        const unsigned char *v = d_valid;
        const unsigned ew = (unsigned)std::min<long long>(((long long)d.M + 255) / 256, 1024);
        const size_t lc = sizeof(cf64) * 2 * kW64ColsPerWg * (size_t)(d.c1 + 1);
        const size_t lr = sizeof(cf64) * 2 * kW64RowsPerWg * (size_t)d.c2;
        const dim3 gc(d.c2 / kW64ColsPerWg, (unsigned)SW), gr(d.c1 / kW64RowsPerWg, (unsigned)SW);
        allow_lds((const void *)k_w64_cols<-1>, lc);
        allow_lds((const void *)k_w64_cols<1>, lc);
        allow_lds((const void *)k_w64_rows<-1>, lr);
        allow_lds((const void *)k_w64_rows<1>, lr);
        {
            ProfScope prof("white_amp", st);
            hipLaunchKernelGGL(k_w64_pack, dim3(ew, (unsigned)SW), dim3(256), 0, st, (const float *)d_x, d.stride,
                               d.W, d.N, d.M, v, Zl);
            hipLaunchKernelGGL(k_w64_cols<-1>, gc, dim3(256), lc, st, Zl, d.c1, d.c2, q1.plan, (const cf64 *)q1.tw,
                               twM, v);
            hipLaunchKernelGGL(k_w64_rows<-1>, gr, dim3(256), lr, st, Zl, d.c1, d.c2, q2.plan, (const cf64 *)q2.tw, v);
            hipLaunchKernelGGL(k_w64_amp, dim3(ew, (unsigned)SW), dim3(256), 0, st, (const cf64 *)Zl, d.M, d.c1, d.c2,
                               twc, v, A);
            FPCC_CHECK(hipGetLastError());
        }
        if ((r = white_filter(d, A, d_valid, Hs, HS, st))) return r;
        {
            ProfScope prof("white_apply", st);
            hipLaunchKernelGGL(k_w64_filter, dim3(ew, (unsigned)SW), dim3(256), 0, st, (const cf64 *)Zl, d.M, d.c1,
                               d.c2, twc, v, d.W, (const double *)HS, Fl);
            hipLaunchKernelGGL(k_w64_rows<1>, gr, dim3(256), lr, st, Fl, d.c1, d.c2, q2.plan, (const cf64 *)q2.tw, v);
            hipLaunchKernelGGL(k_w64_cols<1>, gc, dim3(256), lc, st, Fl, d.c1, d.c2, q1.plan, (const cf64 *)q1.tw,
                               twM, v);
            hipLaunchKernelGGL(k_w64_unpack, dim3(ew, (unsigned)SW), dim3(256), 0, st, (const cf64 *)Fl, d.M, d_x,
                               d.stride, d.W, d.N, v);
            FPCC_CHECK(hipGetLastError());
        }
    }
    return 0;
}

size_t preprocess_workspace(const fpcc_job *job, const fpcc_preproc *pp)
{
    PreDims d;
    if (pre_dims(job, pp, d)) return 0;
    return pre_workspace(d);
}

}  // namespace fpcc

using namespace fpcc;

extern "C" size_t fpcc_preprocess_workspace_size(const fpcc_job *job, const fpcc_preproc *pp)
{
    return preprocess_workspace(job, pp);
}

extern "C" int fpcc_preprocess_device(const fpcc_job *job, const fpcc_preproc *pp, float *d_x, const float *d_sign,
                                      unsigned char *d_valid, float *d_rms, void *d_ws, size_t ws_bytes,
                                      void *stream)
{
    return preprocess_device(job, pp, d_x, d_sign, d_valid, d_rms, d_ws, ws_bytes, (hipStream_t)stream);
}

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" int fpcc_preprocess(const fpcc_job *job, const fpcc_preproc *pp, float *h_x, const float *h_sign,
                               unsigned char *h_valid, float *h_rms)
{
    PreDims d;
    const int rd = pre_dims(job, pp, d);
    if (rd) return rd;
    if (!h_x || !h_valid) return -1;
    const size_t ws = pre_workspace(d);
    const size_t SW = (size_t)job->n_stations * job->n_windows;
    const size_t nx = (size_t)(job->n_stations - 1) * job->station_stride + (size_t)job->n_windows * job->win_len;
    DevBuf dx, dsg, dv, dr, dws;
    FPCC_CHECK(hipMalloc(&dx.p, sizeof(float) * nx));
    FPCC_CHECK(hipMalloc(&dv.p, SW));
    FPCC_CHECK(hipMalloc(&dr.p, sizeof(float) * SW));
    FPCC_CHECK(hipMalloc(&dws.p, ws));
    if (h_sign) {
        FPCC_CHECK(hipMalloc(&dsg.p, sizeof(float) * SW));
        FPCC_CHECK(hipMemcpy(dsg.p, h_sign, sizeof(float) * SW, hipMemcpyHostToDevice));
    }
    FPCC_CHECK(hipMemcpy(dx.p, h_x, sizeof(float) * nx, hipMemcpyHostToDevice));
    int r = preprocess_device(job, pp, (float *)dx.p, (const float *)dsg.p, (unsigned char *)dv.p, (float *)dr.p,
                              dws.p, ws, nullptr);
    if (r) return r;
    FPCC_CHECK(hipMemcpy(h_x, dx.p, sizeof(float) * nx, hipMemcpyDeviceToHost));
    FPCC_CHECK(hipMemcpy(h_valid, dv.p, SW, hipMemcpyDeviceToHost));
    if (h_rms) FPCC_CHECK(hipMemcpy(h_rms, dr.p, sizeof(float) * SW, hipMemcpyDeviceToHost));
    return 0;
}
