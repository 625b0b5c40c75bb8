// Device FFT core for gfx950: register butterflies + LDS-staged Stockham passes.
//
// A length-n transform runs inside one workgroup.  n is factored into radices
// from {16, 8, 4, 2, 9, 3, 25, 5, 7}; pass p reads in[b + r*n/R], applies the
// twiddles W_{Ns*R}^{r*k} (k = b mod Ns, Ns = product of earlier radices), does
// a radix-R DFT in registers and writes out[(b/Ns)*Ns*R + k + r*Ns] (Stockham
// autosort: natural order in and out, no bit reversal).  Passes ping-pong between
// two LDS buffers.  Twiddles come from a per-plan table generated in FP64 on the
// host and rounded once to the working precision; entry (r-1)*Ns + k of pass p
// holds W_{Ns*R}^{r*k} of the forward transform (the inverse uses its conjugate).
#pragma once
#include <utility>
#include "cplx.hpp"

namespace fpcc {

constexpr int kMaxPasses = 10;

// Plan of one transform length, passed to kernels by value.
struct FftPlan {
    int n;
    int npass;
    int radix[kMaxPasses];
    int tw_off[kMaxPasses];  // offset of each pass's twiddles in the table
};

// ---- compile-time loop ----
template <int... Is, class F>
__host__ __device__ __forceinline__ void sfor_impl(std::integer_sequence<int, Is...>, F &&f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ void sfor(F &&f)
{
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// multiply by the constant W_R^M = exp(DIR*2*pi*i*M/R), folding the trivial cases
template <int DIR, int R, int M, typename T>
__host__ __device__ __forceinline__ C2<T> twc(C2<T> a)
{
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (2 * m == R) {
        return C2<T>{-a.x, -a.y};
    } else if constexpr (4 * m == R) {
        return rot<DIR>(a);
    } else if constexpr (4 * m == 3 * R) {
        return rot<-DIR>(a);
    } else {
        constexpr T c = (T)ct_cos2pi(m, R);
        constexpr T s = (T)(DIR * ct_sin2pi(m, R));
        return C2<T>{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

constexpr int first_factor(int r)
{
    return (r % 4 == 0 && r != 4) ? 4 : (r % 2 == 0 && r != 2) ? 2 : (r % 3 == 0 && r != 3) ? 3 : (r % 5 == 0 && r != 5) ? 5 : r;
}

// ---- in-register DFTs ----
template <int R, int DIR>
struct Dft {
    // composite: R = A*B, Cooley-Tukey in registers with constant twiddles
    static constexpr int A = first_factor(R);
    static constexpr int B = R / A;
    static_assert(A > 1 && B > 1, "unsupported radix");
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        C2<T> t[R];
        sfor<B>([&](auto b) {
            C2<T> col[A];
            sfor<A>([&](auto a) { col[a()] = v[B * a() + b()]; });
            Dft<A, DIR>::run(col);
            sfor<A>([&](auto k1) { t[b() * A + k1()] = twc<DIR, R, b() * k1()>(col[k1()]); });
        });
        sfor<A>([&](auto k1) {
            C2<T> row[B];
            sfor<B>([&](auto b) { row[b()] = t[b() * A + k1()]; });
            Dft<B, DIR>::run(row);
            sfor<B>([&](auto k2) { v[k1() + A * k2()] = row[k2()]; });
        });
    }
};

template <int DIR>
struct Dft<2, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        C2<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <int DIR>
struct Dft<4, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        C2<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
        C2<T> s13 = v[1] + v[3], d13 = rot<DIR>(v[1] - v[3]);
        v[0] = s02 + s13;
        v[2] = s02 - s13;
        v[1] = d02 + d13;
        v[3] = d02 - d13;
    }
};

template <int DIR>
struct Dft<8, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        // radix-2 over two radix-4 halves (even / odd samples)
        C2<T> e[4] = {v[0], v[2], v[4], v[6]};
        C2<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        constexpr T h = (T)0.70710678118654752440;
        C2<T> o1 = C2<T>{(o[1].x - DIR * o[1].y) * h, (o[1].y + DIR * o[1].x) * h};  // * W8
        C2<T> o2 = rot<DIR>(o[2]);                                                  // * W8^2
        C2<T> o3 = C2<T>{(-o[3].x - DIR * o[3].y) * h, (-o[3].y + DIR * o[3].x) * h}; // * W8^3
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
};

template <int DIR>
struct Dft<3, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        constexpr T s3 = (T)0.86602540378443864676;
        C2<T> s = v[1] + v[2];
        C2<T> d = s3 * (v[1] - v[2]);
        C2<T> t = C2<T>{v[0].x - (T)0.5 * s.x, v[0].y - (T)0.5 * s.y};
        v[0] = v[0] + s;
        C2<T> rd = rot<DIR>(d);
        v[1] = t + rd;
        v[2] = t - rd;
    }
};

template <int DIR>
struct Dft<5, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        constexpr T c1 = (T)0.30901699437494742410;   // cos(2pi/5)
        constexpr T c2 = (T)-0.80901699437494742410;  // cos(4pi/5)
        constexpr T q1 = (T)(DIR * 0.95105651629515357212);  // DIR*sin(2pi/5)
        constexpr T q2 = (T)(DIR * 0.58778525229247312917);  // DIR*sin(4pi/5)
        C2<T> a0 = v[0];
        C2<T> s7 = v[1] + v[4], s10 = v[1] - v[4], s8 = v[2] + v[3], s9 = v[2] - v[3];
        v[0] = a0 + s7 + s8;
        C2<T> s5 = a0 + c1 * s7 + c2 * s8;
        C2<T> s6 = q1 * s10 + q2 * s9;
        C2<T> s11 = a0 + c2 * s7 + c1 * s8;
        C2<T> s12 = q2 * s10 - q1 * s9;
        C2<T> is6 = C2<T>{-s6.y, s6.x}, is12 = C2<T>{-s12.y, s12.x};
        v[1] = s5 + is6;
        v[4] = s5 - is6;
        v[2] = s11 + is12;
        v[3] = s11 - is12;
    }
};

template <int DIR>
struct Dft<7, DIR> {
    template <typename T>
    __host__ __device__ __forceinline__ static void run(C2<T> *v)
    {
        C2<T> a[7];
        sfor<7>([&](auto j) { a[j()] = v[j()]; });
        C2<T> sp[3], sm[3];
        sfor<3>([&](auto j) {
            sp[j()] = a[j() + 1] + a[6 - j()];
            sm[j()] = a[j() + 1] - a[6 - j()];
        });
        v[0] = a[0] + sp[0] + sp[1] + sp[2];
        sfor<6>([&](auto km1) {
            constexpr int k = km1() + 1;
            C2<T> re = a[0], im = C2<T>{0, 0};
            sfor<3>([&](auto j) {
                constexpr T c = (T)ct_cos2pi((j() + 1) * k, 7);
                constexpr T s = (T)(DIR * ct_sin2pi((j() + 1) * k, 7));
                re = re + c * sp[j()];
                im = im + s * sm[j()];
            });
            v[k] = re + C2<T>{-im.y, im.x};
        });
    }
};

// ---- one Stockham pass over LDS (ping-pong: in != out) ----
template <typename T, int R, int DIR>
__device__ __forceinline__ void stockham_pass(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int n,
                                              int Ns, const C2<T> *__restrict__ tw)
{
    const int nb = n / R;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[b + r * nb];
        const int k = b % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2<T> w = tw[(r - 1) * Ns + k];
                if (DIR > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Dft<R, DIR>::run(v);
        const int base = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
    }
}

// Runs all passes of `plan` on data in buf0 (LDS), ping-ponging with buf1.
// Must be called by the whole workgroup; ends with a barrier.  Returns the
// buffer that holds the result.
template <typename T, int DIR>
__device__ C2<T> *fft_lds(C2<T> *buf0, C2<T> *buf1, const FftPlan &plan, const C2<T> *__restrict__ tw)
{
    C2<T> *in = buf0, *out = buf1;
    int Ns = 1;
    for (int p = 0; p < plan.npass; ++p) {
        const int R = plan.radix[p];
        const C2<T> *twp = tw + plan.tw_off[p];
        switch (R) {
        case 2: stockham_pass<T, 2, DIR>(in, out, plan.n, Ns, twp); break;
        case 3: stockham_pass<T, 3, DIR>(in, out, plan.n, Ns, twp); break;
        case 4: stockham_pass<T, 4, DIR>(in, out, plan.n, Ns, twp); break;
        case 5: stockham_pass<T, 5, DIR>(in, out, plan.n, Ns, twp); break;
        case 7: stockham_pass<T, 7, DIR>(in, out, plan.n, Ns, twp); break;
        case 8: stockham_pass<T, 8, DIR>(in, out, plan.n, Ns, twp); break;
        case 9: stockham_pass<T, 9, DIR>(in, out, plan.n, Ns, twp); break;
        case 16: stockham_pass<T, 16, DIR>(in, out, plan.n, Ns, twp); break;
        case 25: stockham_pass<T, 25, DIR>(in, out, plan.n, Ns, twp); break;
        default: break;
        }
        __syncthreads();
        C2<T> *t = in;
        in = out;
        out = t;
        Ns *= R;
    }
    return in;
}

// ---- batched variant: nb transforms of length plan.n, transform t at t*ld ----
template <typename T, int R, int DIR>
__device__ __forceinline__ void stockham_pass_batch(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, int n,
                                                    int Ns, const C2<T> *__restrict__ tw, int nbatch, int ld)
{
    const int nb = n / R;
    const int total = nb * nbatch;
    for (int bb = threadIdx.x; bb < total; bb += blockDim.x) {
        const int t = bb / nb, b = bb - t * nb;
        const C2<T> *src = in + t * ld;
        C2<T> *dst = out + t * ld;
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[b + r * nb];
        const int k = b % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C2<T> w = tw[(r - 1) * Ns + k];
                if (DIR > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Dft<R, DIR>::run(v);
        const int base = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
}

template <typename T, int DIR>
__device__ C2<T> *fft_lds_batch(C2<T> *buf0, C2<T> *buf1, const FftPlan &plan, const C2<T> *__restrict__ tw,
                                int nbatch, int ld)
{
    C2<T> *in = buf0, *out = buf1;
    int Ns = 1;
    for (int p = 0; p < plan.npass; ++p) {
        const int R = plan.radix[p];
        const C2<T> *twp = tw + plan.tw_off[p];
        switch (R) {
        case 2: stockham_pass_batch<T, 2, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 3: stockham_pass_batch<T, 3, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 4: stockham_pass_batch<T, 4, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 5: stockham_pass_batch<T, 5, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 7: stockham_pass_batch<T, 7, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 8: stockham_pass_batch<T, 8, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 9: stockham_pass_batch<T, 9, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 16: stockham_pass_batch<T, 16, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        case 25: stockham_pass_batch<T, 25, DIR>(in, out, plan.n, Ns, twp, nbatch, ld); break;
        default: break;
        }
        __syncthreads();
        C2<T> *t = in;
        in = out;
        out = t;
        Ns *= R;
    }
    return in;
}

// ---- compile-time FFT: radices, strides and thread count fixed at compile time ----
// In place in one LDS buffer: each pass reads all of a thread's butterfly inputs
// into registers, waits, then writes; LDS index i lives at i + i/16 (one pad slot
// per 16 complex values breaks the power-of-two strides of the early passes).
// The twiddle table is the runtime plan's (same radix order, checked by the host).
__host__ __device__ constexpr int ct_pad(int i) { return i + (i >> 4); }
template <int N>
constexpr int ct_lds_elems()
{
    return N + N / 16 + 1;
}

// LDS layouts of a stage: 0 identity, 1 ct_pad, 2 the low five bits of the index
// XOR-ed with bits 3..7 (a permutation of every aligned block of 256), 3 the low
// four bits XOR-ed with bits 5..8 (a permutation of every aligned block of 16, so
// any multiple of 16 is a valid length)
template <int LAY>
__host__ __device__ constexpr int ct_lay(int i)
{
    if constexpr (LAY == 1) return ct_pad(i);
    else if constexpr (LAY == 2) return i ^ ((i >> 3) & 31);
    else if constexpr (LAY == 3) return i ^ ((i >> 5) & 15);
    else return i;
}

// One pass reading its inputs in layout LIN and writing its outputs in layout LOUT.
template <typename T, int N, int NT, int DIR, int R, int NS, int LIN = 1, int LOUT = 1>
__device__ __forceinline__ void ct_pass(C2<T> *lds, const C2<T> *__restrict__ tw)
{
    constexpr int NB = N / R;
    constexpr int BPT = (NB + NT - 1) / NT;
    C2<T> v[BPT][R];
    // the twiddles are loaded with the LDS reads, ahead of the barrier (which the
    // compiler does not move global loads across), so their latency overlaps it
    C2<T> wt[BPT][R > 1 && NS > 1 ? R - 1 : 1];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = lds[ct_lay<LIN>(b + r * NB)];
            if (NS > 1) {
                const int k = b % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) wt[i][r - 1] = tw[(r - 1) * NS + k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            const int k = b % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C2<T> w = wt[i][r - 1];
                    if (DIR > 0) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            Dft<R, DIR>::run(v[i]);
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[ct_lay<LOUT>(base + r * NS)] = v[i][r];
        }
    }
    __syncthreads();
}

// The same pass out of place: inputs from `in` (layout LIN), outputs to `out`
// (layout LOUT), one barrier at the end (no lane writes what another still reads).
template <typename T, int N, int NT, int DIR, int R, int NS, int LIN, int LOUT>
__device__ __forceinline__ void ct_pass_oop(const C2<T> *in, C2<T> *out, const C2<T> *__restrict__ tw)
{
    constexpr int NB = N / R;
    constexpr int BPT = (NB + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int b = threadIdx.x + i * NT;
        if (NB % NT == 0 || b < NB) {
            C2<T> v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = in[ct_lay<LIN>(b + r * NB)];
            const int k = b % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C2<T> w = tw[(r - 1) * NS + k];
                    if (DIR > 0) w.y = -w.y;
                    v[r] = cmul(v[r], w);
                }
            }
            Dft<R, DIR>::run(v);
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) out[ct_lay<LOUT>(base + r * NS)] = v[r];
        }
    }
    __syncthreads();
}

// The first pass of ct_pass_oop (stride 1, no twiddles), one butterfly a thread,
// with thread b's inputs (elements b + r N/R, r < R) already in its registers v:
// the transform starts without staging its input in LDS.  Outputs to `out` in
// layout LOUT, then a barrier.
template <typename T, int N, int NT, int DIR, int R, int LOUT>
__device__ __forceinline__ void ct_first_pass_regs(C2<T> *v, C2<T> *out)
{
    static_assert(N / R <= NT, "one butterfly a thread");
    const int b = threadIdx.x;
    if (N / R == NT || b < N / R) {
        Dft<R, DIR>::run(v);
#pragma unroll
        for (int r = 0; r < R; ++r) out[ct_lay<LOUT>(b * R + r)] = v[r];
    }
    __syncthreads();
}

// The last pass of ct_pass_oop (stride N/R), at most one butterfly a thread, with
// its outputs kept in registers: thread b < N/R holds elements b + r N/R (r < R),
// the positions of its inputs (threads past N/R leave v alone).  No barrier: the
// caller orders any later write of `in` after every thread's reads.
template <typename T, int N, int NT, int DIR, int R, int LIN>
__device__ __forceinline__ void ct_last_pass_regs(const C2<T> *in, C2<T> *v, const C2<T> *__restrict__ tw)
{
    constexpr int NS = N / R;
    static_assert(NS <= NT, "one butterfly a thread");
    const int b = threadIdx.x;
    if (NS == NT || b < NS) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[ct_lay<LIN>(b + r * NS)];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            C2<T> w = tw[(r - 1) * NS + b];
            if (DIR > 0) w.y = -w.y;
            v[r] = cmul(v[r], w);
        }
        Dft<R, DIR>::run(v);
    }
}

// ct_last_pass_ends_4096 (below) with its two outputs kept in registers: lo =
// value b, hi = value b + 3584 of the transform, for thread b.  No barrier.
__device__ __forceinline__ void ct_last_pass_ends_4096_regs(const C2<float> *in, const C2<float> *__restrict__ tw,
                                                            C2<float> &lo, C2<float> &hi)
{
    constexpr int NS = 512;
    const int b = threadIdx.x;
    C2<float> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = in[b + r * NS];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        C2<float> w = tw[(r - 1) * NS + b];
        w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    const C2<float> es02 = v[0] + v[4], ed02 = v[0] - v[4];
    const C2<float> es13 = v[2] + v[6], ed13 = rot<+1>(v[2] - v[6]);
    const C2<float> os02 = v[1] + v[5], od02 = v[1] - v[5];
    const C2<float> os13 = v[3] + v[7], od13 = rot<+1>(v[3] - v[7]);
    const C2<float> e0 = es02 + es13, e3 = ed02 - ed13;
    const C2<float> o0 = os02 + os13, o3 = od02 - od13;
    constexpr float h = 0.70710678118654752440f;
    const C2<float> o3w = C2<float>{(-o3.x - o3.y) * h, (-o3.y + o3.x) * h};  // * W8^3
    lo = e0 + o0;
    hi = e3 - o3w;
}

template <typename T, int N, int NT, int DIR, int NS, int OFF, int R, int... Rest>
__device__ __forceinline__ void ct_passes(C2<T> *lds, const C2<T> *__restrict__ tw)
{
    ct_pass<T, N, NT, DIR, R, NS>(lds, tw + OFF);
    if constexpr (sizeof...(Rest) > 0) ct_passes<T, N, NT, DIR, NS * R, OFF + (R - 1) * NS, Rest...>(lds, tw);
}

// Length N = product of Rs, NT threads, data at lds[ct_pad(i)]; ends with a barrier.
template <typename T, int N, int NT, int DIR, int... Rs>
__device__ __forceinline__ void ct_fft(C2<T> *lds, const C2<T> *__restrict__ tw)
{
    static_assert((Rs * ... * 1) == N, "radices must multiply to N");
    ct_passes<T, N, NT, DIR, 1, 0, Rs...>(lds, tw);
}

// The last radix-8 pass of the 4096-point inverse (stride 512, plain layouts) for
// lags within +-1024 of zero: only the outputs b and b + 3584 (real lags 2b, 2b + 1
// and 7168 + 2b, 7169 + 2b) are read, so thread b of a 512-thread workgroup forms
// only outputs 0 and 7 of its butterfly, with the operations of the full radix-8
// DFT that lead to them.
__device__ __forceinline__ void ct_last_pass_ends_4096(const C2<float> *in, C2<float> *out,
                                                       const C2<float> *__restrict__ tw)
{
    constexpr int NS = 512;
    const int b = threadIdx.x;
    C2<float> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = in[b + r * NS];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        C2<float> w = tw[(r - 1) * NS + b];
        w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    // even and odd radix-4 halves: outputs 0 and 3 of each
    const C2<float> es02 = v[0] + v[4], ed02 = v[0] - v[4];
    const C2<float> es13 = v[2] + v[6], ed13 = rot<+1>(v[2] - v[6]);
    const C2<float> os02 = v[1] + v[5], od02 = v[1] - v[5];
    const C2<float> os13 = v[3] + v[7], od13 = rot<+1>(v[3] - v[7]);
    const C2<float> e0 = es02 + es13, e3 = ed02 - ed13;
    const C2<float> o0 = os02 + os13, o3 = od02 - od13;
    constexpr float h = 0.70710678118654752440f;
    const C2<float> o3w = C2<float>{(-o3.x - o3.y) * h, (-o3.y + o3.x) * h};  // * W8^3
    out[b] = e0 + o0;
    out[b + 7 * NS] = e3 - o3w;
    __syncthreads();
}

// ---- workgroup reductions (wave64) ----
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// red must hold blockDim.x/64 elements; result broadcast to all threads
template <typename T>
__device__ __forceinline__ T block_max(T v, T *red)
{
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = red[0];
    for (int i = 1; i < nw; ++i) r = fmax(r, red[i]);
    return r;
}
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = red[0];
    for (int i = 1; i < nw; ++i) r += red[i];
    return r;
}

}  // namespace fpcc
