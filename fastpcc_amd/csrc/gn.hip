// GNCC and 1-bit GNCC (src/FFTapps.c:883-1073) on gfx950.
//
// Station stage, one workgroup per (station, window): samples (GNCC) or their
// signs (1-bit; x >= 0 -> +1 for the virtual source, x > 0 -> +1 for the
// receiver, src/FFTapps.c:1037 and :1042) are zero-padded to Nz and transformed
// with a half-length complex FFT: z[m] = x[2m] + i x[2m+1], Z = FFT_{Nz/2}(z),
//   X[k] = (Z[k] + conj Z[M-k])/2 + e^{-2 pi i k/Nz} (Z[k] - conj Z[M-k])/(2i),
// k in [0, M], M = Nz/2 (Z[M] = Z[0]).  GNCC runs in FP64 like the reference
// (fftw_plan_dft_r2c_1d) and also writes prefix sums of x^2 in FP64 for the
// overlap energies; 1-bit runs in FP32 (fftwf_plan_dft_r2c_1d).
//
// Pair stage: H = conj(X1) X2 (Hermitian half spectrum of a real correlation),
// c2r through the same half-length trick, lag window y[(lag0+l) mod Nz] / Nz,
// then division by sqrt(E1(lag) E2(lag)) with E1 = sum of x1^2 over
// [max(0,-lag), min(N,N-lag)), E2 = sum of x2^2 over [max(0,lag), min(N,N+lag))
// (gn_lowlevel, :372-429); for 1-bit both energies equal the overlap length.
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "edge.hpp"
#include "pairs.hpp"

namespace fpcc {

template <typename T>
__device__ __forceinline__ T sample_val(float v, int sign_rule)
{
    if (sign_rule == 1) return v >= 0.0f ? T(1) : T(-1);
    if (sign_rule == 2) return v > 0.0f ? T(1) : T(-1);
    return (T)v;
}

// One station-window: half spectrum of the zero-padded samples (or signs) into
// out[0..M], and (P != nullptr) the exclusive FP64 prefix sums of x^2 into P[0..N].
template <typename T>
__device__ void station_real_row(const float *__restrict__ in, int N, int Nz, int sign_rule, const FftPlan &planM,
                                 const C2<T> *__restrict__ twM, const C2<T> *__restrict__ twc, C2<T> *__restrict__ out,
                                 double *__restrict__ P, unsigned char *smem_raw)
{
    const int M = Nz >> 1;
    C2<T> *buf0 = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *buf1 = buf0 + M;
    double *scan = reinterpret_cast<double *>(buf1 + M);  // blockDim.x partial sums
    const int T_ = blockDim.x;
    for (int m = threadIdx.x; m < M; m += T_) {
        const int n0 = 2 * m, n1 = 2 * m + 1;
        T e = n0 < N ? sample_val<T>(in[n0], sign_rule) : T(0);
        T o = n1 < N ? sample_val<T>(in[n1], sign_rule) : T(0);
        buf0[m] = C2<T>{e, o};
    }
    __syncthreads();
    C2<T> *Z = fft_lds<T, -1>(buf0, buf1, planM, twM);
    for (int k = threadIdx.x; k <= M; k += T_) {
        C2<T> zk = Z[k == M ? 0 : k];
        C2<T> zm = Z[k == 0 ? 0 : M - k];
        C2<T> zc = C2<T>{zm.x, -zm.y};
        C2<T> ev = C2<T>{(zk.x + zc.x) * T(0.5), (zk.y + zc.y) * T(0.5)};
        C2<T> d = C2<T>{zk.x - zc.x, zk.y - zc.y};
        C2<T> od = C2<T>{d.y * T(0.5), -d.x * T(0.5)};  // d / (2i)
        C2<T> wk = (k == M) ? C2<T>{T(-1), T(0)} : C2<T>{twc[k].x, -twc[k].y};
        out[k] = ev + cmul(wk, od);
    }
    if (P) {
        // exclusive prefix sums of x^2 (double), P[0..N]
        const int per = (N + T_ - 1) / T_;
        const int a = threadIdx.x * per, b = min(N, a + per);
        double part = 0.0;
        for (int n = a; n < b; ++n) {
            double v = (double)in[n];
            part += v * v;
        }
        scan[threadIdx.x] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            double run = 0.0;
            for (int t = 0; t < T_; ++t) {
                double v = scan[t];
                scan[t] = run;
                run += v;
            }
        }
        __syncthreads();
        double run = scan[threadIdx.x];
        for (int n = a; n < b; ++n) {
            P[n] = run;
            double v = (double)in[n];
            run += v * v;
        }
        if (b == N && a < N) P[N] = run;
        if (N == 0 && threadIdx.x == 0) P[0] = 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(512)
k_station_real(const float *__restrict__ x, long long station_stride, int W, int N, int Nz, int sign_rule,
               FftPlan planM, const C2<T> *__restrict__ twM, const C2<T> *__restrict__ twc,
               C2<T> *__restrict__ X, long long ldx, double *__restrict__ E, long long lde)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    const int s = sw / W, w = sw - s * W;
    const float *in = x + (long long)s * station_stride + (long long)w * N;
    station_real_row<T>(in, N, Nz, sign_rule, planM, twM, twc, X + (long long)sw * ldx,
                        E ? E + (long long)sw * lde : nullptr, smem_raw);
}

// The same for the station-windows of several jobs with one Nz (uneven station
// counts, window counts and lengths) in one launch: block b belongs to job j with
// b0[j] <= b < b0[j+1].
template <typename T>
__global__ void __launch_bounds__(512)
k_station_real_batch(RealBatch jb, int Nz, int sign_rule, FftPlan planM, const C2<T> *__restrict__ twM,
                     const C2<T> *__restrict__ twc, long long ldx)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int j = 0;
    while (j + 1 < jb.n && (int)blockIdx.x >= jb.b0[j + 1]) ++j;
    const int sw = blockIdx.x - jb.b0[j];
    const int s = sw / jb.W[j], w = sw - s * jb.W[j];
    const float *in = jb.x[j] + (long long)s * jb.stride[j] + (long long)w * jb.N[j];
    station_real_row<T>(in, jb.N[j], Nz, sign_rule, planM, twM, twc, (C2<T> *)jb.X[j] + (long long)sw * ldx,
                        jb.E[j] ? jb.E[j] + (long long)sw * jb.lde[j] : nullptr, smem_raw);
}

// mode 0: y[p][w][L]; 1: mean over windows.  norm 0: energies (GNCC); 1: overlap count.
// CT4096: Nz = 8192 (M = 4096) with a compile-time radix-8^4 inverse in place in
// one padded LDS buffer (bin k at ct_pad(k); twM the twiddles of that radix
// order): 69 KB instead of two 64 KB buffers in FP64, two workgroups a CU instead
// of one, and the fold's loads unrolled; else the runtime plan with two buffers.
// BATCH: workgroups b0[jj] <= blockIdx.x < b0[jj+1] run job jj of the table (its
// spectra, energies, counts, lag window, output and window edges).
template <typename T, bool CT4096, bool BATCH>
__global__ void __launch_bounds__(512)
k_pair_gn(const C2<T> *__restrict__ Xa, const C2<T> *__restrict__ Xb, long long ldx,
          const double *__restrict__ E, long long lde, int S, int W, int N, int Nz, FftPlan planM,
          const C2<T> *__restrict__ twM, const C2<T> *__restrict__ twc, int lag0, int L, int norm, int mode,
          float *__restrict__ y, const float *__restrict__ edge, int K, GnBatch jb)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int bid = blockIdx.x;
    if (BATCH) {
        int jj = 0;
        while (jj + 1 < jb.n && bid >= jb.b0[jj + 1]) ++jj;
        bid -= jb.b0[jj];
        Xa = (const C2<T> *)jb.Xa[jj];
        Xb = (const C2<T> *)jb.Xb[jj];
        E = jb.E[jj];
        lde = jb.lde[jj];
        S = jb.S[jj];
        W = jb.W[jj];
        N = jb.N[jj];
        lag0 = jb.lag0[jj];
        L = jb.L[jj];
        y = jb.y[jj];
        edge = jb.edge[jj];
        K = jb.K[jj];
    }
    const int M = CT4096 ? 4096 : Nz >> 1;
    C2<T> *buf0 = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *buf1 = buf0 + M;
    float *acc = reinterpret_cast<float *>(CT4096 ? buf0 + ct_lds_elems<4096>() : buf1 + M);
    auto zi = [&](int k) { return CT4096 ? ct_pad(k) : k; };
    int p, w0, w1;
    if (mode == 0) {
        p = bid / W;
        w0 = bid % W;
        w1 = w0 + 1;
    } else {
        p = bid;
        w0 = 0;
        w1 = W;
    }
    int i, j;
    pair_index(p, S, i, j);
    const int T_ = blockDim.x;
    const T inv_nz = T(1) / (T)Nz;
    // every lag within +-1024 (the 4096-point inverse): thread t's butterflies take
    // the bins it folds, t + 512 q, from registers, and its last one gives the
    // transform values t and t + 3584 (lags 2t, 2t + 1 and 2t - 1024, 2t - 1023),
    // normalised and summed over the windows in registers
    // (FP64 only: the FP32 kernel keeps four waves a SIMD only without it)
    constexpr bool kRegs = CT4096 && sizeof(T) == 8;
    const bool ends = kRegs && lag0 >= -1024 && lag0 + L - 1 <= 1023;
    float ar[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!ends)
        for (int l = threadIdx.x; l < L; l += T_) acc[l] = 0.0f;
    // the normalised correlation at lag (lag0 + l) of window w from its transform value
    auto lagval = [&](int l, T vt, int w, const double *P1, const double *P2) -> float {
        const int lag = lag0 + l;
        double c = (double)(vt * inv_nz);
        const int m = N - abs(lag);
        if (edge && m > 0 && m < K) c = edge_corr(edge, K, W, i, j, w, w + 1, lag, m, norm != 0);
        double den;
        if (norm == 0) {
            const int a1 = max(0, -lag), b1 = min(N, N - lag);
            const int a2 = max(0, lag), b2 = min(N, N + lag);
            double e1 = P1[b1] - P1[a1], e2 = P2[b2] - P2[a2];
            den = sqrt(e1 * e2);
        } else {
            den = (double)(N - abs(lag));
        }
        return (float)(c / den);
    };
    for (int w = w0; w < w1; ++w) {
        const C2<T> *X1 = Xa + (long long)(i * W + w) * ldx;
        const C2<T> *X2 = Xb + (long long)(j * W + w) * ldx;
        auto foldv = [&](int k) {
            C2<T> a = cmulc(X1[k], X2[k]);
            C2<T> c = cmulc(X1[M - k], X2[M - k]);
            if (k == 0) {  // c2r ignores the imaginary parts of DC and Nyquist
                a.y = T(0);
                c.y = T(0);
            }
            C2<T> P = C2<T>{a.x + c.x, a.y - c.y};
            C2<T> Q = C2<T>{a.x - c.x, a.y + c.y};
            C2<T> wq = cmul(twc[k], Q);
            return C2<T>{P.x - wq.y, P.y + wq.x};
        };
        auto fold = [&](int k) { buf0[zi(k)] = foldv(k); };
        const double *P1 = E ? E + (long long)(i * W + w) * lde : nullptr;
        const double *P2 = E ? E + (long long)(j * W + w) * lde : nullptr;
        if constexpr (kRegs) {
            if (ends) {
                C2<T> v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    v[q] = foldv(threadIdx.x + 512 * q);
                    // two bins' loads at a time keep FP64 within two workgroups a CU
                    if (q & 1) __builtin_amdgcn_sched_barrier(0);
                }
                ct_first_pass_regs<T, 4096, 512, +1, 8, 1>(v, buf0);
                ct_pass<T, 4096, 512, +1, 8, 8, 1, 1>(buf0, twM + 7);
                ct_pass<T, 4096, 512, +1, 8, 64, 1, 1>(buf0, twM + 7 + 56);
                ct_last_pass_regs<T, 4096, 512, +1, 8, 1>(buf0, v, twM + 7 + 56 + 448);
                const int t2 = 2 * threadIdx.x;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int l = (e < 2 ? t2 + e : t2 - 1024 + (e - 2)) - lag0;
                    const T vt = e == 0 ? v[0].x : e == 1 ? v[0].y : e == 2 ? v[7].x : v[7].y;
                    if (l >= 0 && l < L) ar[e] += lagval(l, vt, w, P1, P2);
                }
                __syncthreads();  // every lane has read buf0 before the next window's first pass
                continue;
            }
        }
        C2<T> *z;
        if constexpr (CT4096) {
            // two bins' loads at a time keep FP64 within two workgroups a CU
#pragma unroll 2
            for (int q = 0; q < 4096 / 512; ++q) fold(threadIdx.x + 512 * q);
            __syncthreads();
            ct_fft<T, 4096, 512, +1, 8, 8, 8, 8>(buf0, twM);
            z = buf0;
        } else {
            for (int k = threadIdx.x; k < M; k += T_) fold(k);
            __syncthreads();
            z = fft_lds<T, +1>(buf0, buf1, planM, twM);
        }
        for (int l = threadIdx.x; l < L; l += T_) {
            const int n = (lag0 + l) & (Nz - 1);
            C2<T> v = z[zi(n >> 1)];
            acc[l] += lagval(l, (n & 1) ? v.y : v.x, w, P1, P2);
        }
        __syncthreads();
    }
    const float sc = mode == 0 ? 1.0f : 1.0f / (float)W;
    float *dst = mode == 0 ? y + ((long long)p * W + w0) * L : y + (long long)p * L;
    if (ends) {
        const int t2 = 2 * threadIdx.x;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int l = (e < 2 ? t2 + e : t2 - 1024 + (e - 2)) - lag0;
            if (l >= 0 && l < L) dst[l] = mode == 0 ? ar[e] : ar[e] * sc;
        }
    } else {
        for (int l = threadIdx.x; l < L; l += T_) dst[l] = mode == 0 ? acc[l] : acc[l] * sc;
    }
}

static void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

long long real_ld(int Nz) { return (long long)(Nz / 2 + 1 + 1) / 2 * 2; }
long long energy_ld(int N) { return (long long)(N + 1 + 1) / 2 * 2; }

int launch_station_real(const float *d_x, long long station_stride, int S, int W, int N, int Nz, int sign_rule,
                        bool f64, void *d_X, double *d_E, hipStream_t st)
{
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, f64);
    const void *twc = get_c2r_twiddles(Nz, f64);
    if (!pM.ok || !twc) return -1;
    const int threads = 512;
    const size_t elem = f64 ? sizeof(cf64) : sizeof(cf32);
    size_t lds = elem * 2 * (size_t)M + sizeof(double) * threads;
    if (lds > 160 * 1024) {
        fprintf(stderr, "fastpcc: Nz=%d exceeds the single-workgroup limit for real spectra\n", Nz);
        return -1;
    }
    ProfScope prof(f64 ? "station_real_f64" : "station_real_f32", st);
    if (f64) {
        allow_lds((const void *)k_station_real<double>, lds);
        hipLaunchKernelGGL(k_station_real<double>, dim3(S * W), dim3(threads), lds, st, d_x, station_stride, W, N,
                           Nz, sign_rule, pM.plan, (const cf64 *)pM.tw, (const cf64 *)twc, (cf64 *)d_X, real_ld(Nz),
                           d_E, energy_ld(N));
    } else {
        allow_lds((const void *)k_station_real<float>, lds);
        hipLaunchKernelGGL(k_station_real<float>, dim3(S * W), dim3(threads), lds, st, d_x, station_stride, W, N,
                           Nz, sign_rule, pM.plan, (const cf32 *)pM.tw, (const cf32 *)twc, (cf32 *)d_X, real_ld(Nz),
                           d_E, energy_ld(N));
    }
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_station_real_batch(const RealBatch &jb, int Nz, int sign_rule, bool f64, hipStream_t st)
{
    if (jb.n < 1 || jb.n > kMaxBatchJobs) return -1;
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, f64);
    const void *twc = get_c2r_twiddles(Nz, f64);
    if (!pM.ok || !twc) return -1;
    const int threads = 512;
    const size_t elem = f64 ? sizeof(cf64) : sizeof(cf32);
    size_t lds = elem * 2 * (size_t)M + sizeof(double) * threads;
    if (lds > 160 * 1024) return -1;
    const int total = jb.b0[jb.n];
    if (total <= 0) return 0;
    ProfScope prof(f64 ? "station_real_f64_batch" : "station_real_f32_batch", st);
    if (f64) {
        allow_lds((const void *)k_station_real_batch<double>, lds);
        hipLaunchKernelGGL(k_station_real_batch<double>, dim3(total), dim3(threads), lds, st, jb, Nz, sign_rule,
                           pM.plan, (const cf64 *)pM.tw, (const cf64 *)twc, real_ld(Nz));
    } else {
        allow_lds((const void *)k_station_real_batch<float>, lds);
        hipLaunchKernelGGL(k_station_real_batch<float>, dim3(total), dim3(threads), lds, st, jb, Nz, sign_rule,
                           pM.plan, (const cf32 *)pM.tw, (const cf32 *)twc, real_ld(Nz));
    }
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_pair_gn(const void *d_Xa, const void *d_Xb, const double *d_E, bool f64, int S, int W, int N, int Nz,
                   int lag0, int L, int norm, int mode, float *d_y, hipStream_t st, const float *d_edge)
{
    const int K = d_edge ? edge_len(N) : 0;
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, f64);
    const void *twc = get_c2r_twiddles(Nz, f64);
    if (!pM.ok || !twc) return -1;
    const int P = S * (S - 1) / 2;
    const int blocks = mode == 0 ? P * W : P;
    if (!blocks) return 0;
    const size_t elem = f64 ? sizeof(cf64) : sizeof(cf32);
    // Nz 8192: the compile-time in-place inverse and its own twiddle order
    const bool ct = M == 4096;
    const void *tw = ct ? get_order_twiddles(M, {8, 8, 8, 8}, f64) : pM.tw;
    if (!tw) return -1;
    size_t lds = elem * (ct ? (size_t)ct_lds_elems<4096>() : 2 * (size_t)M) + sizeof(float) * (size_t)L;
    if (lds > 160 * 1024) {
        fprintf(stderr, "fastpcc: lag window too long for the single-workgroup pair stage\n");
        return -1;
    }
    ProfScope prof(f64 ? "pair_gn_f64" : "pair_gn_f32", st);
    if (f64) {
        const auto fn = ct ? k_pair_gn<double, true, false> : k_pair_gn<double, false, false>;
        allow_lds((const void *)fn, lds);
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(512), lds, st, (const cf64 *)d_Xa, (const cf64 *)d_Xb,
                           real_ld(Nz), d_E, energy_ld(N), S, W, N, Nz, pM.plan, (const cf64 *)tw,
                           (const cf64 *)twc, lag0, L, norm, mode, d_y, d_edge, K, GnBatch{});
    } else {
        const auto fn = ct ? k_pair_gn<float, true, false> : k_pair_gn<float, false, false>;
        allow_lds((const void *)fn, lds);
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(512), lds, st, (const cf32 *)d_Xa, (const cf32 *)d_Xb,
                           real_ld(Nz), d_E, energy_ld(N), S, W, N, Nz, pM.plan, (const cf32 *)tw,
                           (const cf32 *)twc, lag0, L, norm, mode, d_y, d_edge, K, GnBatch{});
    }
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_pair_gn_batch(GnBatch jb, bool f64, int Nz, int norm, int mode, hipStream_t st)
{
    if (jb.n < 1 || jb.n > kMaxBatchJobs || mode < 0 || mode > 1) return -1;
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, f64);
    const void *twc = get_c2r_twiddles(Nz, f64);
    if (!pM.ok || !twc) return -1;
    int Lmax = 1;
    long long blocks = 0;
    for (int j = 0; j < jb.n; ++j) {
        const long long P = (long long)jb.S[j] * (jb.S[j] - 1) / 2;
        jb.b0[j] = (int)blocks;
        blocks += mode == 0 ? P * jb.W[j] : P;
        Lmax = std::max(Lmax, jb.L[j]);
        jb.lde[j] = energy_ld(jb.N[j]);
        jb.K[j] = jb.edge[j] ? edge_len(jb.N[j]) : 0;
    }
    if (blocks > 0x7fffffff) return -1;
    jb.b0[jb.n] = (int)blocks;
    if (!blocks) return 0;
    const size_t elem = f64 ? sizeof(cf64) : sizeof(cf32);
    const bool ct = M == 4096;
    const void *tw = ct ? get_order_twiddles(M, {8, 8, 8, 8}, f64) : pM.tw;
    if (!tw) return -1;
    size_t lds = elem * (ct ? (size_t)ct_lds_elems<4096>() : 2 * (size_t)M) + sizeof(float) * (size_t)Lmax;
    if (lds > 160 * 1024) return -1;
    ProfScope prof(f64 ? "pair_gn_f64_batch" : "pair_gn_f32_batch", st);
    if (f64) {
        const auto fn = ct ? k_pair_gn<double, true, true> : k_pair_gn<double, false, true>;
        allow_lds((const void *)fn, lds);
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(512), lds, st, (const cf64 *)nullptr,
                           (const cf64 *)nullptr, real_ld(Nz), (const double *)nullptr, 0LL, 0, 0, 0, Nz, pM.plan,
                           (const cf64 *)tw, (const cf64 *)twc, 0, 0, norm, mode, (float *)nullptr,
                           (const float *)nullptr, 0, jb);
    } else {
        const auto fn = ct ? k_pair_gn<float, true, true> : k_pair_gn<float, false, true>;
        allow_lds((const void *)fn, lds);
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(512), lds, st, (const cf32 *)nullptr,
                           (const cf32 *)nullptr, real_ld(Nz), (const double *)nullptr, 0LL, 0, 0, 0, Nz, pM.plan,
                           (const cf32 *)tw, (const cf32 *)twc, 0, 0, norm, mode, (float *)nullptr,
                           (const float *)nullptr, 0, jb);
    }
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc
