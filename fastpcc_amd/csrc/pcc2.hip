// PCC2 on gfx950: station stage (phasor spectra) and pair stage (cross-correlation).
//
// Station stage, one workgroup per (station, window), all in LDS:
//   x[N] -> DFT_N -> analytic mask -> IDFT_N -> AmpNormf -> zero-pad -> DFT_Nz
// which is AnalyticSignal_plan_float + AmpNormf + the zero-padded forward FFT of
// pcc2_set (src/FFTapps.c:104-124, 178-205, 838-848), computed once per station
// window instead of once per pair window.
//
// Pair stage: only Re(IFFT_Nz(conj(F1) F2)) is needed (src/FFTapps.c:851-856), so
// the cross spectrum Y is folded to its Hermitian part and inverted as a real
// signal with a half-length complex IFFT:
//   Z[k] = P_k + i e^{+2 pi i k/Nz} Q_k,  k < Nz/2,
//   P_k = (Y[k] + Y[k+M] + conj(Y[-k] + Y[M-k])) / 2,
//   Q_k = (Y[k] - Y[k+M] + conj(Y[-k] - Y[M-k])) / 2,   M = Nz/2,
//   z = IFFT_M(Z),  y[2m] = Re z[m],  y[2m+1] = Im z[m].
// The lag window y[(lag0 + l) mod Nz], l < L, is scaled by 1/((float)Nz*(float)N).
#include <initializer_list>
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "pairs.hpp"

namespace fpcc {

// One station-window: analytic signal, AmpNormf, zero-padding to Nz and FFT_Nz
// into Frow (or, phirow != nullptr, the length-N phasors).
__device__ void station_pcc_row(const float *__restrict__ in, int N, int Nz, const FftPlan &planN,
                                const cf32 *__restrict__ twN, const FftPlan &planZ, const cf32 *__restrict__ twZ,
                                float da1, cf32 *__restrict__ Frow, cf32 *__restrict__ phirow,
                                unsigned char *smem_raw)
{
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + Nz;
    float *red = reinterpret_cast<float *>(buf1 + Nz);
    const int T = blockDim.x;

    for (int n = threadIdx.x; n < N; n += T) buf0[n] = cf32{in[n], 0.0f};
    __syncthreads();
    cf32 *X = fft_lds<float, -1>(buf0, buf1, planN, twN);
    // analytic mask (AnalyticSignal_plan_float, :113-117)
    const int half = (N + 1) >> 1;
    const float two_da1 = 2 * da1;
    for (int k = threadIdx.x; k < N; k += T) {
        cf32 v = X[k];
        float g;
        if (k == 0) g = da1;
        else if (k < half) g = two_da1;
        else if (!(N & 1) && k == (N >> 1)) g = da1;
        else g = 0.0f;
        X[k] = cf32{v.x * g, v.y * g};
    }
    __syncthreads();
    cf32 *other = (X == buf0) ? buf1 : buf0;
    cf32 *z = fft_lds<float, +1>(X, other, planN, twN);
    // AmpNormf (:178-205): eps = 1e-6*sqrt(max|z|^2) in double, stored as float
    float m = 0.0f;
    for (int n = threadIdx.x; n < N; n += T) {
        cf32 v = z[n];
        float a = v.x * v.x + v.y * v.y;
        m = fmaxf(m, a);
    }
    m = block_max(m, red);
    const float eps = (float)(1e-6 * sqrt((double)m));
    if (phirow) {  // phasor-only output (time-domain PCC1 / PCC-v)
        for (int n = threadIdx.x; n < N; n += T) {
            cf32 v = z[n];
            float r = 1.0f / (sqrtf(v.x * v.x + v.y * v.y) + eps);
            phirow[n] = __builtin_isnormal(r) ? cf32{v.x * r, v.y * r} : cf32{0.0f, 0.0f};
        }
        return;
    }
    cf32 *dst = (z == buf0) ? buf1 : buf0;
    for (int n = threadIdx.x; n < Nz; n += T) {
        cf32 o = cf32{0.0f, 0.0f};
        if (n < N) {
            cf32 v = z[n];
            float r = 1.0f / (sqrtf(v.x * v.x + v.y * v.y) + eps);
            if (__builtin_isnormal(r)) o = cf32{v.x * r, v.y * r};
        }
        dst[n] = o;
    }
    __syncthreads();
    cf32 *Fz = fft_lds<float, -1>(dst, z, planZ, twZ);
    for (int k = threadIdx.x; k < Nz; k += T) Frow[k] = Fz[k];
}

__global__ void __launch_bounds__(512)
k_station_pcc(const float *__restrict__ x, long long station_stride, int W, int N, int Nz, FftPlan planN,
              const cf32 *__restrict__ twN, FftPlan planZ, const cf32 *__restrict__ twZ, float da1,
              cf32 *__restrict__ F, cf32 *__restrict__ phi)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int sw = blockIdx.x;
    const int s = sw / W, w = sw - s * W;
    const float *in = x + (long long)s * station_stride + (long long)w * N;
    station_pcc_row(in, N, Nz, planN, twN, planZ, twZ, da1, F ? F + (long long)sw * Nz : nullptr,
                    phi ? phi + (long long)sw * N : nullptr, smem_raw);
}

// The station-windows of several PCC2 jobs with one Nz (uneven station counts,
// window counts, lengths and strides) in one launch: block b belongs to job j
// with b0[j] <= b < b0[j+1] and runs exactly k_station_pcc's code for it.
__global__ void __launch_bounds__(512)
k_station_pcc_batch(PccBatch jb, int Nz, FftPlan planZ, const cf32 *__restrict__ twZ)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int j = 0;
    while (j + 1 < jb.n && (int)blockIdx.x >= jb.b0[j + 1]) ++j;
    const int sw = blockIdx.x - jb.b0[j];
    const int s = sw / jb.W[j], w = sw - s * jb.W[j];
    const float *in = jb.x[j] + (long long)s * jb.stride[j] + (long long)w * jb.N[j];
    station_pcc_row(in, jb.N[j], Nz, jb.planN[j], jb.twN[j], planZ, twZ, jb.da1[j], jb.F[j] + (long long)sw * Nz,
                    nullptr, smem_raw);
}

constexpr int kSpN = 3600, kSpNz = 8192, kSpM = 4096, kSpT = 512;

// Station stage specialised for N = 3600, Nz = 8192 (one day at 1 Hz cut in
// hourly windows): compile-time FFTs out of place between two LDS buffers of 4096
// values (64 KB), 512 threads (the 4096-point transforms in radix-8 passes, a
// butterfly a thread; 116 registers and 64 KB allow two workgroups, 16 waves, a CU:
// 6% faster than 256 threads with radix-16 passes at 147 registers and 12 waves;
// out of place, 2.5% faster than in place in one padded buffer).  The window occupies
// samples < 3600 < 4096 of the zero-padded 8192, so the 8192-point spectrum
// splits into two 4096-point transforms:
//   F[2m] = FFT_4096(z)[m],  F[2m+1] = FFT_4096(z[n] e^{-2 pi i n/8192})[m];
// the normalised window z waits in registers for the second one, the even bins
// for the odd ones, and each lane writes (F[2m], F[2m+1]) as one 16-byte store.
__global__ void __launch_bounds__(kSpT)
k_station_pcc_3600(const float *__restrict__ x, long long station_stride, int W, const cf32 *__restrict__ twN,
                   const cf32 *__restrict__ twI, const cf32 *__restrict__ twM, const cf32 *__restrict__ twZ, float da1,
                   cf32 *__restrict__ F)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *lds = reinterpret_cast<cf32 *>(smem_raw);
    float *red = reinterpret_cast<float *>(lds + 2 * kSpM);
    const int sw = blockIdx.x, t = threadIdx.x;
    const int s = sw / W, w = sw - s * W;
    const float *in = x + (long long)s * station_stride + (long long)w * kSpN;
    cf32 *alt = lds + kSpM;
    constexpr int kB = kSpN / 16, kB25 = kSpN / 25;  // butterflies of the radix-16 and radix-25 passes
    cf32 v[25];
    // forward transform, radices 16, 9, 25: the window straight from global memory
    // into the first pass's registers (thread b < 225: samples b + 225 r), the last
    // pass's outputs kept in registers (thread b < 144: bins b + 144 r)
    if (t < kB) {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = cf32{in[t + kB * r], 0.0f};
    }
    ct_first_pass_regs<float, kSpN, kSpT, -1, 16, 3>(v, alt);
    ct_pass_oop<float, kSpN, kSpT, -1, 9, 16, 3, 0>(alt, lds, twN + 15);
    ct_last_pass_regs<float, kSpN, kSpT, -1, 25, 0>(lds, v, twN + 15 + 8 * 16);
    // the analytic-signal gain in registers: bins 1 .. 1799 doubled, 0 and 1800 kept,
    // the rest zero
    constexpr int half = (kSpN + 1) >> 1;
    const float two_da1 = 2 * da1;
    if (t < kB25) {
#pragma unroll
        for (int r = 0; r < 25; ++r) {
            const int k = t + kB25 * r;
            const float g = (k == 0 || k == (kSpN >> 1)) ? da1 : k < half ? two_da1 : 0.0f;
            v[r] = cf32{v[r].x * g, v[r].y * g};
        }
    }
    // inverse transform, radices 25, 9, 16 (twiddles twI): thread b < 144's first
    // butterfly takes bins b + 144 r, its own registers, and writes alt (last read
    // by the forward's second pass, before its barrier); the last pass leaves
    // samples b + 225 r (b < 225) of the analytic signal in registers
    ct_first_pass_regs<float, kSpN, kSpT, +1, 25, 0>(v, alt);
    ct_pass_oop<float, kSpN, kSpT, +1, 9, 25, 0, 0>(alt, lds, twI + 24);
    ct_last_pass_regs<float, kSpN, kSpT, +1, 16, 0>(lds, v, twI + 24 + 8 * 25);
    // AmpNormf: max |z|^2 from registers; the samples go through alt (last read by
    // the inverse's second pass) to the 4096-point transforms' distribution, and
    // the one barrier of the block maximum orders those writes before the reads
    // (red is written once in this kernel, so no barrier guards its earlier use)
    float m = 0.0f;
    if (t < kB) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            m = fmaxf(m, v[r].x * v[r].x + v[r].y * v[r].y);
            alt[t + kB * r] = v[r];
        }
    }
    m = wave_max(m);
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSpT / 64; ++i) m = fmaxf(m, red[i]);
    const float eps = (float)(1e-6 * sqrt((double)m));
    constexpr int kZ = kSpM / kSpT;  // samples a thread (n = t + kSpT i), zero past 3600
    cf32 *row = F + (long long)sw * kSpNz;
    cf32 z[kZ];
#pragma unroll
    for (int i = 0; i < kZ; ++i) {
        const int n = t + kSpT * i;
        z[i] = cf32{0.0f, 0.0f};
        if (n < kSpN) {
            const cf32 a = alt[n];
            const float q = 1.0f / (sqrtf(a.x * a.x + a.y * a.y) + eps);
            if (__builtin_isnormal(q)) z[i] = cf32{a.x * q, a.y * q};
        }
    }
    // the 4096-point transforms out of place between lds and alt, one barrier a
    // pass, the stages between passes 1-2 and 2-3 XOR-swizzled (every wave's LDS
    // accesses free of bank conflicts).  Thread t's radix-8 butterfly of the first
    // pass takes elements t + 512 i, its own z[i], from registers, and that of the
    // last pass gives elements t + 512 i, the bins it stores, in registers.  Even
    // bins: registers -> lds -> alt -> lds -> registers (lds was last read by the
    // inverse's last pass, before the barrier above; alt, holding the analytic
    // signal, is written only after the first pass's barrier, when every thread
    // has its z)
#pragma unroll
    for (int i = 0; i < kZ; ++i) v[i] = z[i];
    ct_first_pass_regs<float, kSpM, kSpT, -1, 8, 2>(v, lds);
    ct_pass_oop<float, kSpM, kSpT, -1, 8, 8, 2, 2>(lds, alt, twM + 7);
    ct_pass_oop<float, kSpM, kSpT, -1, 8, 64, 2, 0>(alt, lds, twM + 7 + 56);
    cf32 ev[kSpM / kSpT];
    ct_last_pass_regs<float, kSpM, kSpT, -1, 8, 0>(lds, ev, twM + 7 + 56 + 448);
    // odd bins: registers -> alt -> lds -> alt -> registers (alt was last read by
    // the even transform's third pass, before its barrier; lds is written only
    // after the first pass's barrier, when every thread has read the last pass's
    // inputs from it)
#pragma unroll
    for (int i = 0; i < kZ; ++i) {
        const int n = t + kSpT * i;
        v[i] = n < kSpN ? cmul(z[i], twZ[n]) : cf32{0.0f, 0.0f};
    }
    ct_first_pass_regs<float, kSpM, kSpT, -1, 8, 2>(v, alt);
    ct_pass_oop<float, kSpM, kSpT, -1, 8, 8, 2, 2>(alt, lds, twM + 7);
    ct_pass_oop<float, kSpM, kSpT, -1, 8, 64, 2, 0>(lds, alt, twM + 7 + 56);
    ct_last_pass_regs<float, kSpM, kSpT, -1, 8, 0>(alt, v, twM + 7 + 56 + 448);
    float4 *out = reinterpret_cast<float4 *>(row);
#pragma unroll
    for (int i = 0; i < kSpM / kSpT; ++i) out[t + kSpT * i] = float4{ev[i].x, ev[i].y, v[i].x, v[i].y};
}

// AmpNormf (src/FFTapps.c:178-205) in place on rows of N complex values.
__global__ void __launch_bounds__(256) k_ampnormf_rows(cf32 *__restrict__ z, int N)
{
    __shared__ float red[4];
    cf32 *row = z + (long long)blockIdx.x * N;
    float m = 0.0f;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        cf32 v = row[n];
        m = fmaxf(m, v.x * v.x + v.y * v.y);
    }
    m = block_max(m, red);
    const float eps = (float)(1e-6 * sqrt((double)m));
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
        cf32 v = row[n];
        float r = 1.0f / (sqrtf(v.x * v.x + v.y * v.y) + eps);
        row[n] = __builtin_isnormal(r) ? cf32{v.x * r, v.y * r} : cf32{0.0f, 0.0f};
    }
}

// Fold conj(F1)F2 into the half-length spectrum Z (see file header) in LDS.
__device__ __forceinline__ void fold_cross(const cf32 *__restrict__ F1, const cf32 *__restrict__ F2, int Nz,
                                           const cf32 *__restrict__ twc, cf32 *Z, bool accumulate)
{
    const int M = Nz >> 1;
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        const int kn = (Nz - k) & (Nz - 1);
        const int km = M - k;
        cf32 y0 = cmulc(F1[k], F2[k]);
        cf32 y1 = cmulc(F1[k + M], F2[k + M]);
        cf32 y2 = cmulc(F1[kn], F2[kn]);
        cf32 y3 = cmulc(F1[km], F2[km]);
        cf32 P = cf32{0.5f * (y0.x + y1.x + y2.x + y3.x), 0.5f * (y0.y + y1.y - y2.y - y3.y)};
        cf32 Q = cf32{0.5f * (y0.x - y1.x + y2.x - y3.x), 0.5f * (y0.y - y1.y - y2.y + y3.y)};
        cf32 wq = cmul(twc[k], Q);
        cf32 zk = cf32{P.x - wq.y, P.y + wq.x};
        if (accumulate) {
            cf32 a = Z[k];
            Z[k] = cf32{a.x + zk.x, a.y + zk.y};
        } else {
            Z[k] = zk;
        }
    }
}

__device__ __forceinline__ float real_out(const cf32 *z, int n)
{
    cf32 v = z[n >> 1];
    return (n & 1) ? v.y : v.x;
}

// mode 0: per-window outputs y[p][w][L]; 1: mean of per-window IFFTs y[p][L];
// 2: IFFT of the window-summed folded spectrum, y[p][L] (identical by linearity).
__global__ void __launch_bounds__(512)
k_pair_pcc2(const cf32 *__restrict__ F, int S, int W, int Nz, FftPlan planM, const cf32 *__restrict__ twM,
            const cf32 *__restrict__ twc, int lag0, int L, float scale, int mode, int pair_base,
            float *__restrict__ y)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = Nz >> 1;
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + M;
    float *acc = reinterpret_cast<float *>(buf1 + M);
    int p, w0, w1;
    if (mode == 0) {
        p = blockIdx.x / W + pair_base;
        w0 = blockIdx.x % W;
        w1 = w0 + 1;
    } else {
        p = blockIdx.x + pair_base;
        w0 = 0;
        w1 = W;
    }
    int i, j;
    pair_index(p, S, i, j);
    const int T = blockDim.x;
    if (mode == 2) {
        for (int w = w0; w < w1; ++w) {
            fold_cross(F + (long long)(i * W + w) * Nz, F + (long long)(j * W + w) * Nz, Nz, twc, buf0, w > w0);
            __syncthreads();
        }
        cf32 *z = fft_lds<float, +1>(buf0, buf1, planM, twM);
        const float sc = scale / (float)W;
        float *dst = y + (long long)(p - pair_base) * L;
        for (int l = threadIdx.x; l < L; l += T) {
            int n = (lag0 + l) & (Nz - 1);
            dst[l] = sc * real_out(z, n);
        }
        return;
    }
    for (int l = threadIdx.x; l < L; l += T) acc[l] = 0.0f;
    for (int w = w0; w < w1; ++w) {
        fold_cross(F + (long long)(i * W + w) * Nz, F + (long long)(j * W + w) * Nz, Nz, twc, buf0, false);
        __syncthreads();
        cf32 *z = fft_lds<float, +1>(buf0, buf1, planM, twM);
        for (int l = threadIdx.x; l < L; l += T) {
            int n = (lag0 + l) & (Nz - 1);
            acc[l] += scale * real_out(z, n);
        }
        __syncthreads();
    }
    if (mode == 0) {
        float *dst = y + ((long long)(p - pair_base) * W + w0) * L;
        for (int l = threadIdx.x; l < L; l += T) dst[l] = acc[l];
    } else {
        const float inv = 1.0f / (float)W;
        float *dst = y + (long long)(p - pair_base) * L;
        for (int l = threadIdx.x; l < L; l += T) dst[l] = acc[l] * inv;
    }
}

// gfx950 workgroups may address up to 160 KB of LDS; opt in above 64 KB.
static void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int launch_station_pcc(const float *d_x, long long station_stride, int S, int W, int N, int Nz, cf32 *d_F,
                       hipStream_t st)
{
    const DevPlan &pN = get_plan(N, false);
    const DevPlan &pZ = get_plan(Nz, false);
    if (!pN.ok || !pZ.ok) {
        if (!tables_missing()) fprintf(stderr, "fastpcc: unsupported FFT length (N=%d, Nz=%d)\n", N, Nz);
        return -1;
    }
    if (Nz > 8192) {
        fprintf(stderr, "fastpcc: Nz=%d exceeds the single-workgroup limit (8192)\n", Nz);
        return -1;
    }
    const float da1 = (float)(1 / (double)N);
    if (station_pcc_specialised(N, Nz)) {
        const void *twN = get_order_twiddles(kSpN, {16, 9, 25}), *twI = get_order_twiddles(kSpN, {25, 9, 16});
        const void *twM = get_order_twiddles(kSpM, {8, 8, 8, 8});
        const void *twZ = get_full_twiddles(kSpNz, false);
        if (!twN || !twI || !twM || !twZ) return -1;
        size_t lds = sizeof(cf32) * 2 * kSpM + sizeof(float) * 8;
        allow_lds((const void *)k_station_pcc_3600, lds);
        ProfScope prof("station_pcc", st);
        hipLaunchKernelGGL(k_station_pcc_3600, dim3(S * W), dim3(kSpT), lds, st, d_x, station_stride, W,
                           (const cf32 *)twN, (const cf32 *)twI, (const cf32 *)twM, (const cf32 *)twZ, da1, d_F);
        FPCC_CHECK(hipGetLastError());
        return 0;
    }
    const int threads = 512;
    size_t lds = sizeof(cf32) * 2 * (size_t)Nz + sizeof(float) * 16;
    allow_lds((const void *)k_station_pcc, lds);
    ProfScope prof("station_pcc", st);
    hipLaunchKernelGGL(k_station_pcc, dim3(S * W), dim3(threads), lds, st, d_x, station_stride, W, N, Nz,
                       pN.plan, (const cf32 *)pN.tw, pZ.plan, (const cf32 *)pZ.tw, da1, d_F, (cf32 *)nullptr);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

// The compile-time kernel takes its own twiddle tables, so only the lengths matter.
bool station_pcc_specialised(int N, int Nz) { return N == kSpN && Nz == kSpNz; }

int station_pcc_batch_job(int N, int Nz, cf32 *d_F, PccBatch &jb, int j)
{
    const DevPlan &pN = get_plan(N, false);
    if (!pN.ok || Nz > 8192 || N > Nz) return -1;
    jb.planN[j] = pN.plan;
    jb.twN[j] = (const cf32 *)pN.tw;
    jb.N[j] = N;
    jb.da1[j] = (float)(1 / (double)N);
    jb.F[j] = d_F;
    return 0;
}

int launch_station_pcc_batch(const PccBatch &jb, int Nz, hipStream_t st)
{
    if (jb.n < 1 || jb.n > kMaxBatchJobs) return -1;
    const DevPlan &pZ = get_plan(Nz, false);
    if (!pZ.ok || Nz > 8192) return -1;
    const int blocks = jb.b0[jb.n];
    if (blocks <= 0) return 0;
    const size_t lds = sizeof(cf32) * 2 * (size_t)Nz + sizeof(float) * 16;
    allow_lds((const void *)k_station_pcc_batch, lds);
    ProfScope prof("station_pcc_batch", st);
    hipLaunchKernelGGL(k_station_pcc_batch, dim3(blocks), dim3(512), lds, st, jb, Nz, pZ.plan, (const cf32 *)pZ.tw);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_station_phasor(const float *d_x, long long station_stride, int S, int W, int N, cf32 *d_phi,
                          hipStream_t st)
{
    const DevPlan &pN = get_plan(N, false);
    if (!pN.ok || N > 8192) {
        fprintf(stderr, "fastpcc: unsupported analytic-signal length N=%d\n", N);
        return -1;
    }
    const float da1 = (float)(1 / (double)N);
    const int threads = 512;
    size_t lds = sizeof(cf32) * 2 * (size_t)N + sizeof(float) * 16;
    allow_lds((const void *)k_station_pcc, lds);
    ProfScope prof("station_phasor", st);
    hipLaunchKernelGGL(k_station_pcc, dim3(S * W), dim3(threads), lds, st, d_x, station_stride, W, N, N, pN.plan,
                       (const cf32 *)pN.tw, pN.plan, (const cf32 *)pN.tw, da1, (cf32 *)nullptr, d_phi);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_pair_pcc2(const cf32 *d_F, int S, int W, int N, int Nz, int lag0, int L, int mode, int pair_base,
                     int npairs, float *d_y, hipStream_t st)
{
    if (Nz == 8192 && pair_base == 0 && npairs == S * (S - 1) / 2)
        return launch_pair_pcc2_4096(d_F, S, W, N, lag0, L, mode, d_y, st);
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, false);
    const void *twc = get_c2r_twiddles(Nz, false);
    if (!pM.ok || !twc) return -1;
    const float scale = 1 / ((float)Nz * (float)N);
    const int threads = 512;
    size_t lds = sizeof(cf32) * 2 * (size_t)M + sizeof(float) * (size_t)L;
    int blocks = mode == 0 ? npairs * W : npairs;
    if (blocks == 0) return 0;
    allow_lds((const void *)k_pair_pcc2, lds);
    ProfScope prof(mode == 2 ? "pair_pcc2_zsum" : "pair_pcc2_plain", st);
    hipLaunchKernelGGL(k_pair_pcc2, dim3(blocks), dim3(threads), lds, st, d_F, S, W, Nz, pM.plan,
                       (const cf32 *)pM.tw, (const cf32 *)twc, lag0, L, scale, mode, pair_base, d_y);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc

namespace fpcc {
int launch_ampnormf_rows(cf32 *d_z, int rows, int N, hipStream_t st)
{
    if (rows <= 0) return 0;
    ProfScope prof("ampnormf_rows", st);
    hipLaunchKernelGGL(k_ampnormf_rows, dim3(rows), dim3(256), 0, st, d_z, N);
    FPCC_CHECK(hipGetLastError());
    return 0;
}
}  // namespace fpcc
