// Phase-weighted stack over windows (Schimmel & Paulssen 1997) of all pairs'
// per-window correlations, the step after the reference's path (its output goes
// to the external ts-PWS program, src/PCC_fullpair_1b.c:568-569):
//     PWS(l) = (1/W sum_w c_w(l)) * |1/W sum_w u_w(l)|^power,
// u_w(l) = a_w(l) / |a_w(l)| (0 where a_w(l) = 0), a_w the analytic signal of
// window w's whole circular correlation r_w (Nz lags), read at the lag window.
//
// One workgroup per pair loops over the windows.  From the station spectra it
// forms the one-sided spectrum A of a_w: with Y = conj(X1) X2,
//   PCC2 (complex phasor spectra F, r = Re IFFT(Y)):  A(k) = Y(k) + conj Y(Nz-k),
//   1-bit / GNCC (half spectra of real records):     A(k) = 2 Y(k),
// for 0 < k < M = Nz/2, A(0) = Re Y(0), A(M) = Re Y(M), zero above M.  The
// Nz-point inverse of a spectrum that vanishes above M is two M-point inverses:
//   a[2m] = IFFT_M(A)[m] + A(M),   a[2m+1] = IFFT_M(A(k) e^{2 pi i k/Nz})[m] - A(M),
// both run in LDS (the spectrum of one window stays in registers between them).
// Re a_w = r_w, so PCC2's linear stack comes from the same transforms (scaled by
// 1/(Nz N)); 1-bit and GNCC have their linear stacks (overlap and energy
// normalisations, exact short-overlap sums) written first by their plain pair
// stages, which this kernel then multiplies by the coherence in place.
#include <type_traits>
#include "engine.hpp"
#include "kernels.hpp"
#include "pairs.hpp"
#include "profile.hpp"

namespace fpcc {

namespace {

constexpr int kPwsThreads = 512;

enum { SRC_FULL = 0, SRC_HALF32 = 1, SRC_HALF64 = 2 };

// One complex bin through a buffer resource at a byte offset; the 64-bit result is
// reinterpreted whole (taking its two 32-bit lanes apart compiles to one dword load).
__device__ __forceinline__ cf32 ld_bin32(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(cf32, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// conj(a) b
template <typename T>
__device__ __forceinline__ cf32 conj_mul(const C2<T> &a, const C2<T> &b)
{
    return cf32{(float)(a.x * b.x + a.y * b.y), (float)(a.x * b.y - a.y * b.x)};
}

// CT4096: Nz = 8192 (M = 4096) with the compile-time radix-8 x 8 x 8 x 8 inverse of the
// pair IFFT (out of place between the two buffers, the stage layouts plain or
// XOR-swizzled; tw the twiddles of that radix order), else the runtime plan.
template <int SRC, int LPT, bool CT4096, bool ENDS = false>
__global__ void __launch_bounds__(kPwsThreads)
k_pair_pws(const void *__restrict__ Xa, const void *__restrict__ Xb, long long ld, int S, int W, int Nz,
           FftPlan plan, const cf32 *__restrict__ tw, const cf32 *__restrict__ twc, int lag0, int L, float cscale,
           float power, int multiply, float *__restrict__ y)
{
    constexpr int NT = kPwsThreads, KPT = 8;  // M <= 4096 = KPT NT bins
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = Nz >> 1;
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw), *buf1 = buf0 + M;
    const int t = threadIdx.x;
    const int p = blockIdx.x;
    int i = 0, j = 0;
    pair_index(p, S, i, j);
    float sc[LPT];
    cf32 su[LPT];
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
        sc[q] = 0.f;
        su[q] = cf32{0.f, 0.f};
    }
    for (int w = 0; w < W; ++w) {
        cf32 A[KPT];
        float AM;
        if (SRC == SRC_FULL) {
            // the two rows as buffer resources (scalar registers) read at 32-bit lane
            // offsets: one 64-bit address a load would hold most of the registers
            const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc(
                (void *)((const cf32 *)Xa + ((long long)i * W + w) * ld), 0, Nz * 8, 0x00020000);
            const __amdgpu_buffer_rsrc_t rj = __builtin_amdgcn_make_buffer_rsrc(
                (void *)((const cf32 *)Xa + ((long long)j * W + w) * ld), 0, Nz * 8, 0x00020000);
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                const int k = t + NT * q;
                if (k < M) {
                    const cf32 Y = conj_mul(ld_bin32(ri, 8u * k), ld_bin32(rj, 8u * k));
                    if (k == 0) {
                        A[q] = cf32{Y.x, 0.f};
                    } else {
                        const cf32 Yc = conj_mul(ld_bin32(ri, 8u * (Nz - k)), ld_bin32(rj, 8u * (Nz - k)));
                        A[q] = cf32{Y.x + Yc.x, Y.y - Yc.y};
                    }
                }
            }
            AM = conj_mul(ld_bin32(ri, 8u * M), ld_bin32(rj, 8u * M)).x;
        } else {
            typedef typename std::conditional<SRC == SRC_HALF64, cf64, cf32>::type CT;
            const CT *Hi = (const CT *)Xa + ((long long)i * W + w) * ld;
            const CT *Hj = (const CT *)Xb + ((long long)j * W + w) * ld;
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                const int k = t + NT * q;
                if (k < M) {
                    const cf32 Y = conj_mul(Hi[k], Hj[k]);
                    A[q] = k == 0 ? cf32{Y.x, 0.f} : cf32{2.f * Y.x, 2.f * Y.y};
                }
            }
            AM = conj_mul(Hi[M], Hj[M]).x;
        }
        if constexpr (CT4096 && ENDS && LPT == 4) {
            // every lag within +-1024: thread t's first butterflies take its bins
            // t + 512 q from registers (A, then A e^{2 pi i k/Nz}) and its last ones
            // give transform values t and t + 3584 of each half, i.e. lags 2t, 2t + 1,
            // 2t - 1024 and 2t - 1023, summed into registers.  Half 0 runs
            // registers -> buf1 -> buf0 -> buf1 -> registers and half 1 registers ->
            // buf0 -> buf1 -> buf0 -> registers: each first pass writes the buffer
            // the previous transform's third pass read (before its barrier), and the
            // second pass writes the one the previous last pass read only after the
            // first pass's barrier.  Six barriers a window.
            constexpr int MC = 4096;
            static_assert(KPT * NT == MC, "a first-pass butterfly's bins a thread");
            cf32 e[2][2];
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                cf32 v[KPT];
#pragma unroll
                for (int q = 0; q < KPT; ++q) v[q] = half ? cmul(A[q], twc[t + NT * q]) : A[q];
                cf32 *b0 = half ? buf0 : buf1, *b1 = half ? buf1 : buf0;
                ct_first_pass_regs<float, MC, NT, +1, 8, 2>(v, b0);
                ct_pass_oop<float, MC, NT, +1, 8, 8, 2, 2>(b0, b1, tw + 7);
                ct_pass_oop<float, MC, NT, +1, 8, 64, 2, 0>(b1, b0, tw + 7 + 56);
                ct_last_pass_ends_4096_regs(b0, tw + 7 + 56 + 448, e[half][0], e[half][1]);
            }
            // lag 2t + half from value t, lag 2t - 1024 + half from value t + 3584
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int half = g & 1, hi = g >> 1;
                const int l = 2 * t - (hi ? 1024 : 0) + half - lag0;
                if (l >= 0 && l < L) {
                    const cf32 v = e[half][hi];
                    const cf32 zz = cf32{v.x + (half ? -AM : AM), v.y};
                    sc[g] += zz.x * cscale;
                    const float m2 = zz.x * zz.x + zz.y * zz.y;
                    if (m2 > 0.f) {
                        const float r = rsqrtf(m2);
                        su[g].x += zz.x * r;
                        su[g].y += zz.y * r;
                    }
                }
            }
            continue;
        }
        cf32 z[LPT];
#pragma unroll
        for (int q = 0; q < LPT; ++q) z[q] = cf32{0.f, 0.f};
        for (int half = 0; half < 2; ++half) {
            __syncthreads();  // the previous transform's outputs have been read
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                const int k = t + NT * q;
                if (k < M) buf0[k] = half ? cmul(A[q], twc[k]) : A[q];
            }
            __syncthreads();
            const cf32 *res;
            if constexpr (CT4096) {
                constexpr int MC = 4096;
                ct_pass_oop<float, MC, NT, +1, 8, 1, 0, 2>(buf0, buf1, tw);
                ct_pass_oop<float, MC, NT, +1, 8, 8, 2, 2>(buf1, buf0, tw + 7);
                ct_pass_oop<float, MC, NT, +1, 8, 64, 2, 0>(buf0, buf1, tw + 7 + 56);
                // lags within +-1024: only the outputs they read
                if (ENDS) ct_last_pass_ends_4096(buf1, buf0, tw + 7 + 56 + 448);
                else ct_pass_oop<float, MC, NT, +1, 8, 512, 0, 0>(buf1, buf0, tw + 7 + 56 + 448);
                res = buf0;
            } else {
                res = fft_lds<float, +1>(buf0, buf1, plan, tw);
            }
            const float edge = half ? -AM : AM;
#pragma unroll
            for (int q = 0; q < LPT; ++q) {
                const int l = t + NT * q;
                if (l < L) {
                    const int n = (lag0 + l) & (Nz - 1);
                    if ((n & 1) == half) {
                        const cf32 v = res[n >> 1];
                        z[q] = cf32{v.x + edge, v.y};
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < LPT; ++q) {
            if (t + NT * q < L) {
                sc[q] += z[q].x * cscale;
                const float m2 = z[q].x * z[q].x + z[q].y * z[q].y;
                if (m2 > 0.f) {
                    const float r = rsqrtf(m2);
                    su[q].x += z[q].x * r;
                    su[q].y += z[q].y * r;
                }
            }
        }
    }
    const float invW = 1.f / (float)W;
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
        // the lag of accumulator q: l = t + NT q, or with the register path above
        // (lags within +-1024) 2t + (q & 1) - 1024 (q >> 1) - lag0
        const int l = (CT4096 && ENDS && LPT == 4) ? 2 * t + (q & 1) - ((q >> 1) ? 1024 : 0) - lag0 : t + NT * q;
        if (l < 0 || l >= L) continue;
        const float ux = su[q].x * invW, uy = su[q].y * invW;
        const float m2 = ux * ux + uy * uy;
        const float coh = power == 0.f ? 1.f : (m2 > 0.f ? powf(m2, 0.5f * power) : 0.f);
        float *dst = y + (long long)p * L + l;
        *dst = multiply ? *dst * coh : sc[q] * invW * coh;
    }
}

template <int SRC>
int launch_lpt(int lpt, bool ct, dim3 grid, size_t lds, hipStream_t st, const void *Xa, const void *Xb, long long ld,
               int S, int W, int Nz, const FftPlan &plan, const cf32 *tw, const cf32 *twc, int lag0, int L,
               float cscale, float power, int multiply, float *y)
{
    // lags within +-1024: the register-ended kernel, four lags a thread
    const bool ends = ct && lag0 >= -1024 && lag0 + L - 1 <= 1023;
#define FPCC_PWS_CASE(N)                                                                                          \
    case N: {                                                                                                     \
        const auto fn = ends ? k_pair_pws<SRC, 4, true, true>                                                     \
                             : ct ? k_pair_pws<SRC, N, true> : k_pair_pws<SRC, N, false>;                        \
        if (lds > 65536) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                   (int)lds);                                                     \
        hipLaunchKernelGGL(fn, grid, dim3(kPwsThreads), lds, st, Xa, Xb, ld, S, W, Nz, plan, tw, twc, lag0, L,  \
                           cscale, power, multiply, y);                                                          \
        break;                                                                                                    \
    }
    switch (lpt) {
        FPCC_PWS_CASE(1)
        FPCC_PWS_CASE(2)
        FPCC_PWS_CASE(4)
        FPCC_PWS_CASE(8)
        FPCC_PWS_CASE(16)
    default: return -1;
    }
#undef FPCC_PWS_CASE
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

bool pws_supported(int Nz) { return Nz >= 16 && Nz <= 8192 && !(Nz & (Nz - 1)); }

int launch_pair_pws(int src, const void *d_Xa, const void *d_Xb, long long ld, int S, int W, int Nz, int lag0, int L,
                    float cscale, float power, int multiply, float *d_y, hipStream_t st)
{
    if (!pws_supported(Nz) || L < 1 || L > Nz || S < 2 || W < 1 || !(power >= 0.f)) return -1;
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, false);
    const void *twc = get_c2r_twiddles(Nz, false);
    if (!pM.ok || !twc) return -1;
    int lpt = 1;
    while (lpt * kPwsThreads < L) lpt *= 2;
    const long long P = (long long)S * (S - 1) / 2;
    if (P > 0x7fffffff) return -1;
    const size_t lds = sizeof(cf32) * 2 * (size_t)M;
    ProfScope prof("pair_pws", st);
    const dim3 grid((unsigned)P);
    // Nz 8192: the compile-time radix-8^4 inverse and its own twiddle order
    const bool ct = M == 4096;
    const cf32 *tw = ct ? (const cf32 *)get_order_twiddles(M, {8, 8, 8, 8}) : (const cf32 *)pM.tw;
    if (!tw) return -1;
    const cf32 *tc = (const cf32 *)twc;
    switch (src) {
    case SRC_FULL:
        return launch_lpt<SRC_FULL>(lpt, ct, grid, lds, st, d_Xa, d_Xb, ld, S, W, Nz, pM.plan, tw, tc, lag0, L, cscale,
                                    power, multiply, d_y);
    case SRC_HALF32:
        return launch_lpt<SRC_HALF32>(lpt, ct, grid, lds, st, d_Xa, d_Xb, ld, S, W, Nz, pM.plan, tw, tc, lag0, L,
                                      cscale, power, multiply, d_y);
    case SRC_HALF64:
        return launch_lpt<SRC_HALF64>(lpt, ct, grid, lds, st, d_Xa, d_Xb, ld, S, W, Nz, pM.plan, tw, tc, lag0, L,
                                      cscale, power, multiply, d_y);
    }
    return -1;
}

}  // namespace fpcc
