// Frequency-domain stack of PCC2 over windows (no counterpart in the reference,
// which never stacks; this is the linear stack the reference leaves to ts-PWS).
//
// The stacked correlation of pair (i, j) is, by linearity of the inverse FFT,
//   mean_w y_w = Re IFFT_Nz( sum_w conj(F_iw) F_jw ) / (W * Nz * N).
// Only the Hermitian part of the summed cross spectrum matters (real output), and
// for k in [0, Nz/2] it is half of
//   G_ij[k] = sum_{m < 2W} conj(U_k[m][i]) U_k[m][j],
//   U_k[m][s] = F_sm[k] (m < W),  conj(F_s(m-W)[Nz-k]) (m >= W),
// i.e. per frequency a Hermitian rank-2W Gram of the station matrix U_k.  It runs
// on MFMA as two real GEMMs with K = 4W:
//   Re G = sum_m ur_i ur_j + ui_i ui_j,   Im G = sum_m ur_i ui_j - ui_i ur_j.
// Stages:
//   k_spectra_to_u   F[s*W+w][k] -> U[k][s][m], k <= Nz/2  (LDS-tiled transpose)
//   k_gram, k_gram16_dma, k_gram_small  U -> G[p][k] (pair-major rows of ld >= Nz/2+1
//                    bins, or runs of 128 bins)
//   k_pair_c2r       G[p] -> half-length complex IFFT -> lag window -> y[p][L]
#include <algorithm>
#include <atomic>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "edge.hpp"

namespace fpcc {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

// hi = fp16(v), lo = fp16(v - hi) of one scaled component.
__device__ __forceinline__ unsigned short split_hi_lo(float v, unsigned short &lo)
{
    const _Float16 h = (_Float16)v;
    const _Float16 l = (_Float16)(v - (float)h);
    lo = __builtin_bit_cast(unsigned short, l);
    return __builtin_bit_cast(unsigned short, h);
}

// One U element: FP32 complex, or (sigma > 0) the split pair
// {hi.re | hi.im << 16, lo.re | lo.im << 16} of sigma * value.
__device__ __forceinline__ uint2 u_element(cf32 v, float sigma)
{
    if (sigma == 0.f) return uint2{__float_as_uint(v.x), __float_as_uint(v.y)};
    unsigned short lr, li;
    const unsigned short hr = split_hi_lo(v.x * sigma, lr);
    const unsigned short hi = split_hi_lo(v.y * sigma, li);
    return uint2{(unsigned)hr | ((unsigned)hi << 16), (unsigned)lr | ((unsigned)li << 16)};
}

// ---- F[sw][Nz] -> U[k][s][2W], k in [0, M] ----
// Tile: 32 frequencies x 64 station-windows, staged in LDS for both the direct
// and the mirrored frequency.
// mirror = 1 (analytic-signal spectra, Nz bins per row): m in [0, 2W) with the
// conjugated mirror frequency in the second half; mirror = 0 (half spectra of real
// signals, ldF >= Nz/2+1 bins per row): m in [0, W).
__global__ void __launch_bounds__(256)
k_spectra_to_u(const cf32 *__restrict__ F, long long ldF, int SW, int W, int Nz, int nk, int mirror, float sigma,
               int grouped, uint2 *__restrict__ U)
{
    __shared__ cf32 td[64][33];
    __shared__ cf32 tm[64][33];
    // R rows a tile: whole stations when W <= 64, so every (bin, station) row of U
    // is written by one tile (whole 128-byte lines)
    const int R = W <= 64 ? 64 / W * W : 64;
    const int k0 = blockIdx.x * 32;
    const int sw0 = blockIdx.y * R;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // all of a thread's loads first (R <= 64: at most 8 rows a thread), then the LDS
    // writes: a loop that loads and stores one row at a time waits for every load
    {
        const int k = k0 + tx;
        cf32 d[8], m[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = ty + 8 * i, sw = sw0 + r;
            d[i] = m[i] = cf32{0.f, 0.f};
            if (r < R && sw < SW && k < nk) {
                d[i] = F[(long long)sw * ldF + k];
                if (mirror) m[i] = F[(long long)sw * ldF + ((Nz - k) & (Nz - 1))];
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = ty + 8 * i;
            if (r < R) {
                td[r][tx] = d[i];
                tm[r][tx] = cf32{m[i].x, -m[i].y};
            }
        }
    }
    __syncthreads();
    // write: for each k of the tile, 64 station-windows (contiguous w runs per station)
    const int W2 = mirror ? 2 * W : W;
    const int S = SW / W;
    if (grouped && (W & 3) == 0) {
        // a thread converts 4 consecutive windows of one station (a group: never split
        // across tiles or stations when 4 | W) and writes its 4 hi words, then 4 lo words
        for (int idx = threadIdx.x; idx < 32 * 16; idx += 256) {
            const int kk = idx >> 4, g = idx & 15;
            const int k = k0 + kk, sw = sw0 + 4 * g;
            if (k >= nk || sw >= SW || 4 * g >= R) continue;
            const int s = sw / W, w = sw - s * W;
            uint4 *row = reinterpret_cast<uint4 *>(U + ((long long)k * S + s) * W2);
            uint2 e0 = u_element(td[4 * g][kk], sigma), e1 = u_element(td[4 * g + 1][kk], sigma);
            uint2 e2 = u_element(td[4 * g + 2][kk], sigma), e3 = u_element(td[4 * g + 3][kk], sigma);
            row[2 * (w >> 2)] = uint4{e0.x, e1.x, e2.x, e3.x};
            row[2 * (w >> 2) + 1] = uint4{e0.y, e1.y, e2.y, e3.y};
            if (mirror) {
                const int m = W + w;
                e0 = u_element(tm[4 * g][kk], sigma);
                e1 = u_element(tm[4 * g + 1][kk], sigma);
                e2 = u_element(tm[4 * g + 2][kk], sigma);
                e3 = u_element(tm[4 * g + 3][kk], sigma);
                row[2 * (m >> 2)] = uint4{e0.x, e1.x, e2.x, e3.x};
                row[2 * (m >> 2) + 1] = uint4{e0.y, e1.y, e2.y, e3.y};
            }
        }
        return;
    }
    for (int idx = threadIdx.x; idx < 32 * 64; idx += 256) {
        const int kk = idx >> 6, r = idx & 63;
        const int k = k0 + kk, sw = sw0 + r;
        if (k >= nk || sw >= SW || r >= R) continue;
        const int s = sw / W, w = sw - s * W;
        uint2 *row = U + ((long long)k * S + s) * W2;
        if (grouped) {
            // split elements in groups of four: 4 hi words, then their 4 lo words
            unsigned *rw = reinterpret_cast<unsigned *>(row);
            uint2 e = u_element(td[r][kk], sigma);
            rw[8 * (w >> 2) + (w & 3)] = e.x;
            rw[8 * (w >> 2) + 4 + (w & 3)] = e.y;
            if (mirror) {
                const int m = W + w;
                e = u_element(tm[r][kk], sigma);
                rw[8 * (m >> 2) + (m & 3)] = e.x;
                rw[8 * (m >> 2) + 4 + (m & 3)] = e.y;
            }
        } else {
            row[w] = u_element(td[r][kk], sigma);
            if (mirror) row[W + w] = u_element(tm[r][kk], sigma);
        }
    }
}

// ---- Gram on MFMA ----
// Workgroup (512 threads, 8 waves): a 64 x 64 block of stations (rows of Ua x
// rows of Ub) for a chunk of kGF consecutive frequencies.  Per step (one
// frequency, up to kGK complex K values) both station blocks are staged through
// LDS (double-buffered, register prefetch of the next step); wave (ti, cj)
// owns the 16 x 32 sub-block rows 16ti.., columns 32cj.. and keeps its
// accumulators for all kGF frequencies in registers, so each pair's kGF bins
// leave as one 64-byte segment (staged through LDS in quarters of the block).
// SPLIT = 0: FP32 MFMA 16x16x4, A[l%16][l/16], B[l/16][l%16], D[4(l/16)+r][l%16];
//   K index (m, component), lane quarter q holds m = 8mm + 2q + e (e = 0, 1).
// SPLIT = 1: U holds sigma-scaled FP16 hi and lo parts; FP16 MFMA 16x16x32 with
//   A[l%16][8(l/16)+e], B[8(l/16)+e][l%16]; lane quarter q holds the 4 complex
//   values m = 16st + 4q + u as 8 halves (re, im interleaved).
//   Re G = Ah.Bh + Ah.Bl + Al.Bh, Im G from the rotated A' = (-ui, ur); the
//   dropped lo.lo term is below 2^-22 relative.  Results are scaled by oscale.
// LDS row (dwords): FP32: 2*kGK components; split: kGK hi words then kGK lo words;
// + 4 dwords of padding, which puts the 16 rows of a 16-byte operand read on
// distinct banks.
constexpr int kGB = 64;
constexpr int kGF = 8;
constexpr int kGK = 64;
constexpr int kGRow = 2 * kGK + 4;
constexpr int kGBuf = 2 * kGB * kGRow;  // dwords per staging buffer (A rows, then B rows)
constexpr int kGOut4 = 2 * kGF + 4;     // dwords per staged pair in the epilogue (16-byte aligned)
constexpr int kGThreads = 512;
constexpr size_t kGramLds = (size_t)4 *
    (2 * kGBuf > kGB * (kGB / 4) * kGOut4 ? 2 * kGBuf : kGB * (kGB / 4) * kGOut4);

__device__ __forceinline__ u4v rot_cplx16(u4v v)
{
    // (re, im) -> (-im, re) on each packed FP16 complex value
    u4v r;
#pragma unroll
    for (int e = 0; e < 4; ++e) r[e] = __builtin_amdgcn_alignbit(v[e], v[e], 16) ^ 0x8000u;
    return r;
}

__device__ __forceinline__ long long gaddr(const GLayout &gl, long long p, int k)
{
    return (long long)(k >> gl.sh) * gl.gs + p * gl.pl + (k & ((1 << gl.sh) - 1));
}

template <int SPLIT>
__global__ void __launch_bounds__(kGThreads)
k_gram(const uint2 *__restrict__ Ua, const uint2 *__restrict__ Ub, int S, int W2, int m0, int K, int nk,
       const int2 *__restrict__ blocks, int nblk, int nchunks, float oscale, cf32 *__restrict__ G, GLayout gl)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned *lds = reinterpret_cast<unsigned *>(smem_raw);
    // XCD-aware order: blocks b, b+8, ... run on one XCD; each XCD takes pairs of
    // adjacent frequency chunks (one 128-byte line of every G row) and all
    // station blocks of the pair back to back, sharing its U rows in L2; the two
    // chunks of a line go to consecutive workgroups, so both halves of each G line
    // are written while it is in L2 and leave as one line.
    const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
    const int half = slot & 1, rest = slot >> 1;
    const int line = rest / nblk, blk = rest - line * nblk;
    const int chunk = (line * 8 + xcd) * 2 + half;
    if (chunk >= nchunks) return;
    const int bi = blocks[blk].x, bj = blocks[blk].y;
    const bool same = bi == bj && Ua == Ub;
    const int k0 = chunk * kGF;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int ti = wave & 3, cj = wave >> 2;
    // skip sub-blocks below a diagonal block's diagonal, and those whose rows or
    // columns are all past the last station
    const bool active = !(bi == bj && 2 * cj + 1 < ti) && bi * kGB + 16 * ti < S && bj * kGB + 32 * cj < S;
    const int nkc = (K + kGK - 1) / kGK;
    const int nsteps = kGF * nkc;
    const int nper = same ? kGB * kGK / kGThreads : 2 * kGB * kGK / kGThreads;  // elements per thread
    constexpr int gran = SPLIT ? 16 : 8;

    uint2 pre[2 * kGB * kGK / kGThreads];
    auto fetch = [&](int s) {
        const int f = s / nkc, kc = s - f * nkc;
        const int k = k0 + f;
        const int kcc = min(kGK, K - kc * kGK);
#pragma unroll
        for (int r = 0; r < 2 * kGB * kGK / kGThreads; ++r) {
            if (r < nper) {
                const int e = t + kGThreads * r, row = e >> 6, mm = e & 63;
                const int side = row >> 6;
                const int st = (side ? bj : bi) * kGB + (row & 63);
                const uint2 *U = side ? Ub : Ua;
                uint2 v = uint2{0u, 0u};
                if (k < nk && st < S && mm < kcc) v = U[((long long)k * S + st) * W2 + m0 + kc * kGK + mm];
                pre[r] = v;
            }
        }
    };
    auto stash = [&](int s, int buf) {
        const int f = s / nkc, kc = s - f * nkc;
        (void)f;
        const int kcc = min(kGK, K - kc * kGK);
        const int kcg = (kcc + gran - 1) / gran * gran;
        unsigned *base = lds + buf * kGBuf;
#pragma unroll
        for (int r = 0; r < 2 * kGB * kGK / kGThreads; ++r) {
            if (r < nper) {
                const int e = t + kGThreads * r, row = e >> 6, mm = e & 63;
                if (mm < kcg) {
                    unsigned *rp = base + row * kGRow;
                    if (SPLIT) {
                        rp[mm] = pre[r].x;
                        rp[kGK + mm] = pre[r].y;
                    } else {
                        *reinterpret_cast<uint2 *>(rp + 2 * mm) = pre[r];
                    }
                }
            }
        }
    };

    f4v acc[kGF][4];
#pragma unroll
    for (int f = 0; f < kGF; ++f)
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[f][a] = f4v{0.f, 0.f, 0.f, 0.f};

    fetch(0);
    stash(0, 0);
    __syncthreads();
    const int arow = 16 * ti + li;
    const int brow = (same ? 0 : kGB) + 32 * cj + li;
#pragma unroll
    for (int f = 0; f < kGF; ++f) {
        for (int kc = 0; kc < nkc; ++kc) {
            const int s = f * nkc + kc;
            const int buf = s & 1;
            if (s + 1 < nsteps) fetch(s + 1);
            const int kcc = min(kGK, K - kc * kGK);
            const int kcg = (kcc + gran - 1) / gran * gran;
            const unsigned *base = lds + buf * kGBuf;
            if (active) {
                f4v r0 = acc[f][0], i0 = acc[f][1], r1 = acc[f][2], i1 = acc[f][3];
                if (SPLIT) {
                    const unsigned *pa = base + arow * kGRow + 4 * q;
                    const unsigned *pb0 = base + brow * kGRow + 4 * q;
                    const unsigned *pb1 = pb0 + 16 * kGRow;
                    for (int stp = 0; stp < kcg / 16; ++stp) {
                        const u4v ah = *reinterpret_cast<const u4v *>(pa + 16 * stp);
                        const u4v al = *reinterpret_cast<const u4v *>(pa + kGK + 16 * stp);
                        const u4v bh0 = *reinterpret_cast<const u4v *>(pb0 + 16 * stp);
                        const u4v bl0 = *reinterpret_cast<const u4v *>(pb0 + kGK + 16 * stp);
                        const u4v bh1 = *reinterpret_cast<const u4v *>(pb1 + 16 * stp);
                        const u4v bl1 = *reinterpret_cast<const u4v *>(pb1 + kGK + 16 * stp);
                        const h8v Ah = __builtin_bit_cast(h8v, ah), Al = __builtin_bit_cast(h8v, al);
                        const h8v Rh = __builtin_bit_cast(h8v, rot_cplx16(ah));
                        const h8v Rl = __builtin_bit_cast(h8v, rot_cplx16(al));
                        const h8v Bh0 = __builtin_bit_cast(h8v, bh0), Bl0 = __builtin_bit_cast(h8v, bl0);
                        const h8v Bh1 = __builtin_bit_cast(h8v, bh1), Bl1 = __builtin_bit_cast(h8v, bl1);
                        r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh0, r0, 0, 0, 0);
                        i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bh0, i0, 0, 0, 0);
                        r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh1, r1, 0, 0, 0);
                        i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bh1, i1, 0, 0, 0);
                        r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl0, r0, 0, 0, 0);
                        i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bl0, i0, 0, 0, 0);
                        r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl1, r1, 0, 0, 0);
                        i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bl1, i1, 0, 0, 0);
                        r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh0, r0, 0, 0, 0);
                        i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rl, Bh0, i0, 0, 0, 0);
                        r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh1, r1, 0, 0, 0);
                        i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rl, Bh1, i1, 0, 0, 0);
                    }
                } else {
                    const float *pa = reinterpret_cast<const float *>(base) + arow * kGRow + 4 * q;
                    const float *pb0 = reinterpret_cast<const float *>(base) + brow * kGRow + 4 * q;
                    const float *pb1 = pb0 + 16 * kGRow;
                    for (int mm = 0; mm < kcg / 8; ++mm) {
                        const f4v a = *reinterpret_cast<const f4v *>(pa + 16 * mm);  // (re, im) of 2 values
                        const f4v b0 = *reinterpret_cast<const f4v *>(pb0 + 16 * mm);
                        const f4v b1 = *reinterpret_cast<const f4v *>(pb1 + 16 * mm);
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            const float ar = a[2 * e], ai = a[2 * e + 1];
                            r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b0[2 * e], r0, 0, 0, 0);
                            i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b0[2 * e + 1], i0, 0, 0, 0);
                            r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b1[2 * e], r1, 0, 0, 0);
                            i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b1[2 * e + 1], i1, 0, 0, 0);
                            r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b0[2 * e + 1], r0, 0, 0, 0);
                            i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b0[2 * e], i0, 0, 0, 0);
                            r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b1[2 * e + 1], r1, 0, 0, 0);
                            i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b1[2 * e], i1, 0, 0, 0);
                        }
                    }
                }
                acc[f][0] = r0;
                acc[f][1] = i0;
                acc[f][2] = r1;
                acc[f][3] = i1;
            }
            if (s + 1 < nsteps) stash(s + 1, buf ^ 1);
            __syncthreads();
        }
    }

    // epilogue in quarters (columns 16h..16h+15, held by the waves with cj = h/2,
    // tt = h%2): pair (il, jl) at dwords (il*16 + jl)*kGOut4 in LDS, then 16 bytes
    // (two bins) a lane, four lanes a pair: 64-byte segments of G rows
    float *out = reinterpret_cast<float *>(smem_raw);
    for (int h = 0; h < 4; ++h) {
        const int tt = h & 1;
        if (cj == (h >> 1) && active) {
#pragma unroll
            for (int f = 0; f < kGF; ++f)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int il = 16 * ti + 4 * q + r, jl = li;
                    const f4v re = tt ? acc[f][2] : acc[f][0], im = tt ? acc[f][3] : acc[f][1];
                    *reinterpret_cast<float2 *>(out + (il * 16 + jl) * kGOut4 + 2 * f) =
                        float2{re[r] * oscale, im[r] * oscale};
                }
        }
        __syncthreads();
        for (int e = t; e < kGB * 16 * 4; e += kGThreads) {
            const int pl = e >> 2, qq = e & 3;
            const int i = bi * kGB + (pl >> 4), j = bj * kGB + 16 * h + (pl & 15);
            if (i >= j || j >= S) continue;
            const long long p = (long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1);
            *reinterpret_cast<f4v *>(G + gaddr(gl, p, k0 + 2 * qq)) =
                *reinterpret_cast<const f4v *>(out + pl * kGOut4 + 4 * qq);
        }
        __syncthreads();
    }
}

// ---- split Gram with operands staged by direct global -> LDS loads ----
// For rows of K = W2 split elements, K a multiple of 16, stored in groups of four
// (4 hi words, then the 4 lo words): the same 64 x 64 blocks, chunks of kGF bins
// and epilogue as k_gram<1>, but each bin's rows are copied by the memory unit
// straight into LDS (buffer loads with the LDS flag, 16 bytes a lane, bounds-
// checked by the buffer resource: out-of-range units read as zero) while the
// MFMAs of an earlier bin run; no registers are spent on staging.  With NB
// staging buffers, bins f+1 .. f+NB-1 are in flight while bin f is multiplied:
// each wave waits (vmcnt) only for its own copies of bin f+1, then a bare
// s_barrier publishes them (__syncthreads would add a wait for every copy in
// flight).  An LDS row holds the K/2 units of a station plus one or two pad
// units (an odd number of 16-byte units), so the 16 rows of an operand read fall
// on distinct banks.  The operand reads are inline ds_read_b128 that wait for
// themselves, which keeps the compiler from waiting for the later bins' copies
// before every LDS read.
__host__ __device__ constexpr int dma_row_units(int K) { return K / 2 + ((K / 2) % 2 ? 2 : 1); }

__device__ __forceinline__ unsigned lds_offset(const void *p)
{
    return (unsigned)(size_t)((__attribute__((address_space(3))) const void *)p);
}

// wait until at most n of this wave's vector-memory operations are in flight (n uniform)
__device__ __forceinline__ void wait_vm(int n)
{
    switch (n) {
    case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    case 1: asm volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
    case 2: asm volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
    case 3: asm volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
    case 4: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    case 5: asm volatile("s_waitcnt vmcnt(5)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
    case 7: asm volatile("s_waitcnt vmcnt(7)" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(9)" ::: "memory"); break;
    }
}

// workgroup barrier over LDS only: this wave's LDS operations complete, then the
// barrier; its global stores stay in flight (__syncthreads would wait for them)
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int NB, bool F32>
__global__ void __launch_bounds__(kGThreads)
k_gram16_dma(const uint2 *__restrict__ Ua, const uint2 *__restrict__ Ub, int S, int K, int nk,
             const int2 *__restrict__ blocks, int nblk, int nchunks, float oscale, cf32 *__restrict__ G,
             GLayout gl)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the XCD-aware order of k_gram: the two chunks of a G line on consecutive
    // workgroups of one XCD
    const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
    const int half = slot & 1, rest = slot >> 1;
    const int line = rest / nblk, blk = rest - line * nblk;
    const int chunk = (line * 8 + xcd) * 2 + half;
    if (chunk >= nchunks) return;
    const int bi = blocks[blk].x, bj = blocks[blk].y;
    const bool same = bi == bj && Ua == Ub;
    const int k0 = chunk * kGF;
    const int t = threadIdx.x, lane = t & 63;
    // the wave index is uniform: taking it through a scalar register keeps the
    // buffer resources and LDS destinations of the copies scalar
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int li = lane & 15, q = lane >> 4;
    const int ti = wave & 3, cj = wave >> 2;
    const bool active = !(bi == bj && 2 * cj + 1 < ti) && bi * kGB + 16 * ti < S && bj * kGB + 32 * cj < S;
    const int su = dma_row_units(K), upr = K / 2;
    const int nchunk64 = (same ? kGB : 2 * kGB) * su / 64;  // 64-unit pieces per bin (rows are 64 per side)
    const int bufb = 2 * kGB * su * 16;                     // bytes per staging buffer
    // resources over this chunk's bins (bins past nk read as zero); the host keeps
    // a chunk under 2^31 bytes, so a pad offset of 2^31 plus any bin's offset is
    // still past the end
    const long long binb = (long long)S * K * 8;
    const unsigned nrec = (unsigned)((long long)min(kGF, nk - k0) * binb);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc((void *)(Ua + (long long)k0 * S * K), 0, nrec, 0x00020000);
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc((void *)(Ub + (long long)k0 * S * K), 0, nrec, 0x00020000);
    unsigned char *lds = smem_raw;
    // each lane's source offset within bin 0 for its (at most 9) copies of a bin
    unsigned voff0[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int c64 = j * 8 + wave;
        const int side = c64 >= su;  // 64 rows x su units = su pieces per side
        const int p = c64 * 64 + lane;
        const int row = p / su, pu = p - row * su;
        const int st = (side ? bj : bi) * kGB + (row & 63);
        voff0[j] = pu < upr ? (unsigned)(((long long)st * K + 2 * pu) * 8) : 0x80000000u;
    }

    // copies holding rows of real stations: a side's rows 0 .. n-1 fill its first
    // ceil(n su / 64) copies, and only the last block ends past the last station, as
    // side b (bi < bj: side a is whole) or as the one side of a diagonal block, so
    // those copies are the first `lim` of the workgroup's; a block that ends past the
    // last station copies only those (its other rows feed only pairs that are never
    // stored).  A prefix keeps the copy loops' early exit (a test of each copy
    // against both sides' counts ran the headline Gram 6% slower).
    const int lim = same ? (min(kGB, S - bi * kGB) * su + 63) / 64 : su + (min(kGB, S - bj * kGB) * su + 63) / 64;

    auto dma = [&](int f, int buf) {
        const unsigned boff = (unsigned)(f * binb);
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int c64 = j * 8 + wave;  // scalar
            if (c64 >= lim) break;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(c64 >= su ? rb : ra,
                                                     (__attribute__((address_space(3))) void *)(lds + buf * bufb +
                                                                                                c64 * 1024),
                                                     16, voff0[j] + boff, 0, 0, 0);
        }
    };

    f4v acc[kGF][4];
#pragma unroll
    for (int f = 0; f < kGF; ++f)
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[f][a] = f4v{0.f, 0.f, 0.f, 0.f};

    // copies this wave issues per bin
    int nw = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) nw += j * 8 + wave < lim;
    // NB buffers: bins f + 1 .. f + NB - 1 in flight while bin f is multiplied
    dma(0, 0);
#pragma unroll
    for (int f = 1; f < NB - 1; ++f)
        if (f < kGF) dma(f, f);
    wait_vm(NB > 2 && kGF > 1 ? nw * (NB - 2) : 0);
    __builtin_amdgcn_s_barrier();  // this wave's copies waited for above
    const int arow = 16 * ti + li;
    const int brow = (same ? 0 : kGB) + 32 * cj + li;
    const unsigned base = lds_offset(smem_raw);
#pragma unroll
    for (int f = 0; f < kGF; ++f) {
        const int buf = f % NB;
        if (f + NB - 1 < kGF) dma(f + NB - 1, (f + NB - 1) % NB);
        if (active) {
            f4v r0 = acc[f][0], i0 = acc[f][1], r1 = acc[f][2], i1 = acc[f][3];
            unsigned pa = base + buf * bufb + arow * su * 16 + (F32 ? 16 : 32) * q;
            unsigned pb0 = base + buf * bufb + brow * su * 16 + (F32 ? 16 : 32) * q;
            unsigned pb1 = pb0 + 16 * su * 16;
            // FP32 rows (k_gram<0>'s operands): 8 complex K values (64 bytes) a step,
            // lane quarter q holding values 2q and 2q + 1, on 16x16x4 FP32 MFMAs
            for (int mm = 0; F32 && mm < K / 8; ++mm) {
                f4v a, b0, b1;
                asm volatile(
                    "ds_read_b128 %0, %3\n\t"
                    "ds_read_b128 %1, %4\n\t"
                    "ds_read_b128 %2, %5\n\t"
                    "s_waitcnt lgkmcnt(0)"
                    : "=&v"(a), "=&v"(b0), "=&v"(b1)
                    : "v"(pa), "v"(pb0), "v"(pb1)
                    : "memory");
                pa += 64;
                pb0 += 64;
                pb1 += 64;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const float ar = a[2 * e], ai = a[2 * e + 1];
                    r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b0[2 * e], r0, 0, 0, 0);
                    i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b0[2 * e + 1], i0, 0, 0, 0);
                    r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b1[2 * e], r1, 0, 0, 0);
                    i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b1[2 * e + 1], i1, 0, 0, 0);
                    r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b0[2 * e + 1], r0, 0, 0, 0);
                    i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b0[2 * e], i0, 0, 0, 0);
                    r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b1[2 * e + 1], r1, 0, 0, 0);
                    i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b1[2 * e], i1, 0, 0, 0);
                }
            }
            for (int stp = 0; !F32 && stp < K / 16; ++stp) {
                u4v ah, al, bh0, bl0, bh1, bl1;
                asm volatile(
                    "ds_read_b128 %0, %6\n\t"
                    "ds_read_b128 %1, %6 offset:16\n\t"
                    "ds_read_b128 %2, %7\n\t"
                    "ds_read_b128 %3, %7 offset:16\n\t"
                    "ds_read_b128 %4, %8\n\t"
                    "ds_read_b128 %5, %8 offset:16\n\t"
                    "s_waitcnt lgkmcnt(0)"
                    : "=&v"(ah), "=&v"(al), "=&v"(bh0), "=&v"(bl0), "=&v"(bh1), "=&v"(bl1)
                    : "v"(pa), "v"(pb0), "v"(pb1)
                    : "memory");
                pa += 128;
                pb0 += 128;
                pb1 += 128;
                const h8v Ah = __builtin_bit_cast(h8v, ah), Al = __builtin_bit_cast(h8v, al);
                const h8v Rh = __builtin_bit_cast(h8v, rot_cplx16(ah));
                const h8v Rl = __builtin_bit_cast(h8v, rot_cplx16(al));
                const h8v Bh0 = __builtin_bit_cast(h8v, bh0), Bl0 = __builtin_bit_cast(h8v, bl0);
                const h8v Bh1 = __builtin_bit_cast(h8v, bh1), Bl1 = __builtin_bit_cast(h8v, bl1);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh0, r0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bh0, i0, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh1, r1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bh1, i1, 0, 0, 0);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl0, r0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bl0, i0, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl1, r1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bl1, i1, 0, 0, 0);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh0, r0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rl, Bh0, i0, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh1, r1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rl, Bh1, i1, 0, 0, 0);
            }
            acc[f][0] = r0;
            acc[f][1] = i0;
            acc[f][2] = r1;
            acc[f][3] = i1;
        }
        // bin f + 1 has landed (the later bins' copies may still be in flight)
        wait_vm(nw * max(0, min(NB - 2, kGF - 2 - f)));
        __builtin_amdgcn_s_barrier();  // this wave's copies waited for above
    }

    // epilogue in quarters (columns 16h..16h+15, held by the waves with cj = h/2,
    // tt = h%2): pair pl = il*16 + jl as 64 bytes at 64 pl in LDS, its 16-byte units
    // (two bins each) permuted by u ^ ((pl >> 1 ^ pl >> 2) & 3), which puts the
    // 16 lanes of a write (16 pairs, one unit) and of a read (4 pairs, 4 units) on
    // distinct banks; then 16 bytes a lane, four lanes a pair: 64-byte segments of G rows
    float *out = reinterpret_cast<float *>(smem_raw);
    const int swl = ((li >> 1) ^ (li >> 2)) & 3;          // writer lane: pl & 15 = li
    const int swr = ((t >> 3) ^ (t >> 4)) & 3;            // reader lane: pl & 15 = (t >> 2) & 15
    for (int h = 0; h < 4; ++h) {
        const int tt = h & 1;
        if (cj == (h >> 1) && active) {
            const unsigned ob = lds_offset(smem_raw) + (16 * (16 * ti + 4 * q) + li) * 64;
#pragma unroll
            for (int f = 0; f < kGF; f += 2)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const f4v re0 = tt ? acc[f][2] : acc[f][0], im0 = tt ? acc[f][3] : acc[f][1];
                    const f4v re1 = tt ? acc[f + 1][2] : acc[f + 1][0], im1 = tt ? acc[f + 1][3] : acc[f + 1][1];
                    const f4v v{re0[r] * oscale, im0[r] * oscale, re1[r] * oscale, im1[r] * oscale};
                    // staged with inline stores: no copy to LDS is in flight here, and
                    // the global stores of the previous quarter need not complete first
                    asm volatile("ds_write_b128 %0, %1" ::"v"(ob + r * 16 * 64 + (((f >> 1) ^ swl) << 4)), "v"(v)
                                 : "memory");
                }
        }
        lds_barrier();
        // each thread's LDS reads issued together, then its stores
        constexpr int kE = kGB * 16 * 4 / kGThreads;
        f4v ev[kE];
#pragma unroll
        for (int u = 0; u < kE; ++u) {
            const int e = t + u * kGThreads, pl = e >> 2, qq = e & 3;
            ev[u] = *reinterpret_cast<const f4v *>(out + pl * 16 + 4 * (qq ^ swr));
        }
#pragma unroll
        for (int u = 0; u < kE; ++u) {
            const int e = t + u * kGThreads, pl = e >> 2, qq = e & 3;
            const int i = bi * kGB + (pl >> 4), j = bj * kGB + 16 * h + (pl & 15);
            if (i < j && j < S) {
                const long long p = (long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1);
                *reinterpret_cast<f4v *>(G + gaddr(gl, p, k0 + 2 * qq)) = ev[u];
            }
        }
        if (h < 3) lds_barrier();
    }
}

// Gram of a network of at most 32 stations (one diagonal block, where the 64 x 64
// kernel keeps two of its eight waves busy): a workgroup a chunk of 8 bins, every
// bin's rows copied into LDS at once (each bin's region a whole number of 1 KB
// pieces), wave f multiplying bin f's 32 x 32 block (its three upper 16 x 16
// tiles) with the operand reads and MFMA sequences of k_gram16_dma, then the
// results staged through LDS (64 bytes a pair, 16-byte units permuted by the pair
// index) into 64-byte runs of G.
template <bool F32>
__global__ void __launch_bounds__(kGThreads)
k_gram_small(const uint2 *__restrict__ U, int S, int K, int nk, int nchunks, float oscale, cf32 *__restrict__ G,
             GLayout gl)
{
    static_assert(kGThreads / 64 == kGF, "a wave a bin");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
    const int chunk = ((slot >> 1) * 8 + xcd) * 2 + (slot & 1);  // a G line's two chunks on one XCD
    if (chunk >= nchunks) return;
    const int k0 = chunk * kGF;
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // the bin this wave multiplies
    const int li = lane & 15, q = lane >> 4;
    const int su = dma_row_units(K), upr = K / 2;
    const int binlds = (su + 1) / 2 * 1024;  // 32 rows of su 16-byte units, rounded up to 1 KB pieces
    const int npb = (S * su + 63) / 64;       // pieces holding the rows of real stations
    const long long binb = (long long)S * K * 8;
    const unsigned nrec = (unsigned)((long long)min(kGF, nk - k0) * binb);
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void *)(U + (long long)k0 * S * K), 0, nrec, 0x00020000);
    // wave f copies bin f's rows and multiplies them as soon as its own copies have
    // landed: no barrier before the products (no wave reads another's bin)
    for (int c = 0; c < npb; ++c) {  // uniform
        const int p = c * 64 + lane, row = p / su, pu = p - row * su;
        const unsigned off = pu < upr ? (unsigned)(wave * binb + ((long long)row * K + 2 * pu) * 8) : 0x80000000u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            r, (__attribute__((address_space(3))) void *)(smem_raw + wave * binlds + c * 1024), 16, off, 0, 0, 0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

    // tiles (0,0), (0,1), (1,1): rows 16 ti .., columns 16 tj ..
    f4v acc[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a][0] = acc[a][1] = f4v{0.f, 0.f, 0.f, 0.f};
    const unsigned base = lds_offset(smem_raw) + wave * binlds;
    unsigned p0 = base + li * su * 16 + (F32 ? 16 : 32) * q;
    unsigned p1 = p0 + 16 * su * 16;
    for (int mm = 0; F32 && mm < K / 8; ++mm) {
        f4v x0, x1;
        asm volatile(
            "ds_read_b128 %0, %2\n\t"
            "ds_read_b128 %1, %3\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(x0), "=&v"(x1)
            : "v"(p0), "v"(p1)
            : "memory");
        p0 += 64;
        p1 += 64;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const f4v &A = a == 2 ? x1 : x0, &B = a == 0 ? x0 : x1;
                const float ar = A[2 * e], ai = A[2 * e + 1];
                f4v re = acc[a][0], im = acc[a][1];
                re = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, B[2 * e], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, B[2 * e + 1], im, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, B[2 * e + 1], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, B[2 * e], im, 0, 0, 0);
                acc[a][0] = re;
                acc[a][1] = im;
            }
        }
    }
    for (int stp = 0; !F32 && stp < K / 16; ++stp) {
        u4v h0, l0, h1, l1;
        asm volatile(
            "ds_read_b128 %0, %4\n\t"
            "ds_read_b128 %1, %4 offset:16\n\t"
            "ds_read_b128 %2, %5\n\t"
            "ds_read_b128 %3, %5 offset:16\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(h0), "=&v"(l0), "=&v"(h1), "=&v"(l1)
            : "v"(p0), "v"(p1)
            : "memory");
        p0 += 128;
        p1 += 128;
        const h8v H0 = __builtin_bit_cast(h8v, h0), L0 = __builtin_bit_cast(h8v, l0);
        const h8v H1 = __builtin_bit_cast(h8v, h1), L1 = __builtin_bit_cast(h8v, l1);
        const h8v RH0 = __builtin_bit_cast(h8v, rot_cplx16(h0)), RL0 = __builtin_bit_cast(h8v, rot_cplx16(l0));
        const h8v RH1 = __builtin_bit_cast(h8v, rot_cplx16(h1)), RL1 = __builtin_bit_cast(h8v, rot_cplx16(l1));
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const h8v &Ah = a == 2 ? H1 : H0, &Al = a == 2 ? L1 : L0;
            const h8v &Rh = a == 2 ? RH1 : RH0, &Rl = a == 2 ? RL1 : RL0;
            const h8v &Bh = a == 0 ? H0 : H1, &Bl = a == 0 ? L0 : L1;
            f4v re = acc[a][0], im = acc[a][1];
            re = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bh, re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bh, im, 0, 0, 0);
            re = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah, Bl, re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rh, Bl, im, 0, 0, 0);
            re = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al, Bh, re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x32_f16(Rl, Bh, im, 0, 0, 0);
            acc[a][0] = re;
            acc[a][1] = im;
        }
    }
    __builtin_amdgcn_s_barrier();  // every wave's operand reads waited for above

    // slot ps = 32 i + j: bins at 64 ps, unit u = f >> 1 at ((u ^ sw(ps)) << 4)
    const unsigned eb = lds_offset(smem_raw);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ti = a == 2 ? 1 : 0, tj = a == 0 ? 0 : 1;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int i = 16 * ti + 4 * q + rr, j = 16 * tj + li, ps = 32 * i + j;
            const int sw = ((ps >> 1) ^ (ps >> 2)) & 3;
            const float2 v{acc[a][0][rr] * oscale, acc[a][1][rr] * oscale};
            asm volatile("ds_write_b64 %0, %1" ::"v"(eb + ps * 64 + ((((wave >> 1) ^ sw) << 4) | ((wave & 1) << 3))),
                         "v"(v)
                         : "memory");
        }
    }
    lds_barrier();
    constexpr int kE = 32 * 32 * 4 / kGThreads;
    f4v ev[kE];
#pragma unroll
    for (int u = 0; u < kE; ++u) {
        const int e = t + u * kGThreads, ps = e >> 2, qq = e & 3;
        const int sw = ((ps >> 1) ^ (ps >> 2)) & 3;
        ev[u] = *reinterpret_cast<const f4v *>(smem_raw + ps * 64 + ((qq ^ sw) << 4));
    }
#pragma unroll
    for (int u = 0; u < kE; ++u) {
        const int e = t + u * kGThreads, ps = e >> 2, qq = e & 3;
        const int i = ps >> 5, j = ps & 31;
        if (i < j && j < S) {
            const long long p = (long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1);
            *reinterpret_cast<f4v *>(G + gaddr(gl, p, k0 + 2 * qq)) = ev[u];
        }
    }
}

// ---- per pair: folded spectrum -> real lags ----
// 1-bit: window edge samples (edge.hpp) for the exact short-overlap lags
struct C2rEdge {
    const float *edge = nullptr;
    int K = 0, S = 0, W = 0;
};

// one output lag: r (scaled transform value), then for counts (1-bit) the exact
// sum over a short overlap and the division by the overlap count
__device__ __forceinline__ float c2r_value(float r, float scale, int count_norm, int N, int Nz, int lag,
                                           const C2rEdge &ed, int pi, int pj)
{
    if (!count_norm) return r;
    const int m = N - abs(lag);
    double d = (double)r;
    if (ed.edge && m > 0 && m < ed.K)
        d = edge_corr(ed.edge, ed.K, ed.W, pi, pj, 0, ed.W, lag, m, true) * (double)scale * (double)Nz;
    return (float)(d / (double)m);
}
// Z[k] = P + i e^{+2 pi i k/Nz} Q, P = G[k] + conj(G[M-k]), Q = G[k] - conj(G[M-k])
__global__ void __launch_bounds__(512)
k_pair_c2r(const cf32 *__restrict__ G, GLayout gl, int Nz, FftPlan planM, const cf32 *__restrict__ twM,
           const cf32 *__restrict__ twc, int lag0, int L, float scale, int count_norm, int N,
           float *__restrict__ y, C2rEdge ed)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = Nz >> 1;
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + M;
    const long long p = blockIdx.x;
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        cf32 a = G[gaddr(gl, p, k)], c = G[gaddr(gl, p, M - k)];
        if (k == 0) {  // DC and Nyquist of a real signal's spectrum are real
            a.y = 0.f;
            c.y = 0.f;
        }
        cf32 P = cf32{a.x + c.x, a.y - c.y};
        cf32 Q = cf32{a.x - c.x, a.y + c.y};
        cf32 wq = cmul(twc[k], Q);
        buf0[k] = cf32{P.x - wq.y, P.y + wq.x};
    }
    __syncthreads();
    cf32 *z = fft_lds<float, +1>(buf0, buf1, planM, twM);
    int pi = 0, pj = 0;
    if (ed.edge) pair_index((int)p, ed.S, pi, pj);
    float *dst = y + p * L;
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        const int n = (lag0 + l) & (Nz - 1);
        cf32 v = z[n >> 1];
        dst[l] = c2r_value(scale * ((n & 1) ? v.y : v.x), scale, count_norm, N, Nz, lag0 + l, ed, pi, pj);
    }
}

// Specialisation for Nz = 8192 (M = 4096): compile-time in-place FFT (radix 8^4,
// NT threads), persistent workgroups each walking pairs p, p + grid, ...  Thread t
// loads bins k = t + NT i (k < M/2) and their mirrors M - k, so the fold into the
// half-length spectrum Z runs in registers as the spectrum is written to LDS
// (Z[k] and Z[M-k] both come from G[k] and G[M-k]); the next pair's bins load
// while the current pair is transformed.
__device__ __forceinline__ cf32 c2r_fold(cf32 a, cf32 c, cf32 w)
{
    const cf32 P = cf32{a.x + c.x, a.y - c.y};
    const cf32 Q = cf32{a.x - c.x, a.y + c.y};
    const cf32 wq = cmul(w, Q);
    return cf32{P.x - wq.y, P.y + wq.x};
}

template <int NT, bool COUNT, bool ENDS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
k_pair_c2r_4096(const cf32 *__restrict__ G, GLayout gl, int P, const cf32 *__restrict__ twM,
                const cf32 *__restrict__ twc, int lag0, int L, float scale, int N, float *__restrict__ y,
                C2rEdge ed)
{
    constexpr int M = 4096, Nz = 8192;
    constexpr int H = M / 2 / NT;    // bins k < M/2 a thread
    static_assert(NT == 512, "one radix-8 butterfly a thread per pass");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // two buffers: each pass reads one and writes the other; the next pair's fold
    // goes to the buffer its predecessor's last pass read
    cf32 *lds = reinterpret_cast<cf32 *>(smem_raw), *alt = lds + M;
    const int t = threadIdx.x;
    long long p = blockIdx.x;
    if (p >= P) return;
    // bin j NT + t of pair q sits at q pl + pos(j NT) + vo, bin j NT - t at
    // q pl + pos((j - 1) NT) + vm (NT is a multiple of the group size, checked by the host)
    const bool rows = gl.sh == 30;
    const int mask = rows ? 0 : (1 << gl.sh) - 1;
    const unsigned vo = rows ? (unsigned)t : (unsigned)((t >> gl.sh) * gl.gs + (t & mask));
    const unsigned vm = rows ? (unsigned)(NT - t)
                             : (unsigned)(((NT >> gl.sh) - ((t + mask) >> gl.sh)) * gl.gs + ((-t) & mask));
    auto pos = [&](int k) -> long long { return rows ? (long long)k : (long long)(k >> gl.sh) * gl.gs; };
    cf32 pa[H], pc[H], ph = cf32{0.f, 0.f};
    const cf32 w0 = twc[t];  // twc[k] = exp(2 pi i k / Nz); twc[t + NT i] = w0 twc[NT i]
    auto fetch = [&](long long q) {
        const cf32 *gq = G + q * gl.pl;
#pragma unroll
        for (int i = 0; i < H; ++i) {
            pa[i] = gq[pos(NT * i) + vo];
            pc[i] = gq[pos(M - NT * (i + 1)) + vm];  // bin M - t - NT i
        }
        if (t == 0) ph = gq[pos(M / 2)];
    };
    fetch(p);
    for (;;) {
#pragma unroll
        for (int i = 0; i < H; ++i) {
            cf32 a = pa[i], c = pc[i];
            if (i == 0 && t == 0) {  // DC and Nyquist of a real signal's spectrum are real
                a.y = 0.f;
                c.y = 0.f;
            }
            const cf32 w = i == 0 ? w0 : cmul(w0, twc[NT * i]);
            lds[t + NT * i] = c2r_fold(a, c, w);
            // twc[M - k] = -conj(twc[k]); Z[M] is Z[0] (not written again)
            if (i > 0 || t > 0) lds[M - t - NT * i] = c2r_fold(c, a, cf32{-w.x, w.y});
        }
        if (t == 0) lds[M / 2] = cf32{2.f * ph.x, -2.f * ph.y};  // the fold at k = M/2 (twiddle i)
        const long long pn = p + gridDim.x;
        if (pn < P) fetch(pn);  // in flight during this pair's work
        __syncthreads();
        // radix 8 x 8 x 8 x 8; the stages between passes 1-2 and 2-3 are XOR-swizzled,
        // the others plain: every LDS access of a wave is then free of bank conflicts
        ct_pass_oop<float, M, NT, +1, 8, 1, 0, 2>(lds, alt, twM);
        ct_pass_oop<float, M, NT, +1, 8, 8, 2, 2>(alt, lds, twM + 7);
        ct_pass_oop<float, M, NT, +1, 8, 64, 2, 0>(lds, alt, twM + 7 + 56);
        if (ENDS) ct_last_pass_ends_4096(alt, lds, twM + 7 + 56 + 448);
        else ct_pass_oop<float, M, NT, +1, 8, 512, 0, 0>(alt, lds, twM + 7 + 56 + 448);
        float *dst = y + p * L;
        // lags whose overlap is under ed.K samples come from the edge sums below
        for (int l = t; l < L; l += NT) {
            const int n = (lag0 + l) & (Nz - 1);
            const cf32 v = lds[n >> 1];
            const float r = scale * ((n & 1) ? v.y : v.x);
            if (!COUNT) {
                dst[l] = r;
            } else {
                const int m = N - abs(lag0 + l);
                if (m >= (ed.edge ? ed.K : 1)) dst[l] = (float)((double)r / (double)m);
            }
        }
        if (COUNT && ed.edge) {  // lags -(N-1) .. -(N-K+1) and N-K+1 .. N-1
            int pi = 0, pj = 0;
            pair_index((int)p, ed.S, pi, pj);
            for (int e = t; e < 2 * (ed.K - 1); e += NT) {
                const int lag = e < ed.K - 1 ? e - (N - 1) : N - ed.K + 1 + (e - (ed.K - 1));
                const int l = lag - lag0;
                if (l >= 0 && l < L) dst[l] = c2r_value(0.f, scale, 1, N, Nz, lag, ed, pi, pj);
            }
        }
        if (pn >= P) break;
        // the next fold writes the buffer the last pass read (every lane is past
        // that pass's barrier); this pair's outputs are read before the next
        // fold's barrier, ahead of the next first pass writing them over
        cf32 *const tmp = lds;
        lds = alt;
        alt = tmp;
        p = pn;
    }
}

// The per pair-window PCC2 pair stage for Nz = 8192 on the same transform.
// Persistent workgroups walk the pairs; for each window w of pair (i, j) thread t
// forms, for its bins k = t + NT q (k < M/2) and their mirrors M - k, the folded
// cross spectrum Gf(k) = Y(k) + conj Y(Nz-k), Y = conj(F_i,w) F_j,w, from the bins
// k, Nz-k, M-k and M+k of both stations' spectra, folds it into the half-length
// spectrum in registers as it writes LDS, and runs the radix-8^4 inverse; the
// next window's bins load while the current one transforms.  MODE 0 writes every
// window's lags, MODE 1 the mean of the windows' lags (accumulated in registers),
// MODE 2 sums the folded spectra over the windows and transforms once a pair.
// One complex bin through a buffer resource at a byte offset (0 beyond the row).
// The 64-bit result is reinterpreted whole: taking its two 32-bit lanes apart
// compiled to one dword load read twice.
__device__ __forceinline__ cf32 ld_bin(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(cf32, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <int NT, int LPT, int MODE, bool ENDS, bool BATCH>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
k_pair_plain_4096(const cf32 *__restrict__ F, int S, int W, int P, const int *__restrict__ order,
                  const cf32 *__restrict__ twM, const cf32 *__restrict__ twc, int lag0, int L, float scale,
                  float *__restrict__ y, PlainBatch jb)
{
    constexpr int M = 4096, Nz = 8192;
    constexpr int H = M / 2 / NT;  // bins k < M/2 a thread
    static_assert(NT == 512, "one radix-8 butterfly a thread per pass");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *lds = reinterpret_cast<cf32 *>(smem_raw), *alt = lds + M;
    const int t = threadIdx.x;
    const cf32 w0 = twc[t];  // twc[k] = exp(2 pi i k / Nz); twc[t + NT q] = w0 twc[NT q]
    cf32 sa[MODE == 2 ? H : 1], sc[MODE == 2 ? H : 1], sph = cf32{0.f, 0.f};
    float acc[LPT];
    for (int u = blockIdx.x; u < P; u += gridDim.x) {
        int uj = u;
        if (BATCH) {  // job j's pairs are u0[j] <= u < u0[j+1]: its spectra, order, lags and output
            int j = 0;
            while (j + 1 < jb.n && u >= jb.u0[j + 1]) ++j;
            uj = u - jb.u0[j];
            F = jb.F[j];
            S = jb.S[j];
            W = jb.W[j];
            order = jb.order[j];
            lag0 = jb.lag0[j];
            L = jb.L[j];
            scale = jb.scale[j];
            y = jb.y[j];
        }
        // pairs in tiles of stations (order[u]): the workgroups running together read
        // the spectra of a few tiles' stations, which stay in the caches
        const int p = __builtin_amdgcn_readfirstlane(order[uj]);
        int i = 0, j = 0;
        pair_index(p, S, i, j);
        // uniform across the workgroup: scalar registers hold the row addresses
        i = __builtin_amdgcn_readfirstlane(i);
        j = __builtin_amdgcn_readfirstlane(j);
#pragma unroll
        for (int q = 0; q < LPT; ++q) acc[q] = 0.f;
        for (int w = 0; w < W; ++w) {
            // the two rows as buffer resources (scalar registers) read at 32-bit lane
            // offsets: plain pointers with lane indices made the compiler keep one
            // 64-bit address per load across the loop and spill
            const __amdgpu_buffer_rsrc_t ri =
                __builtin_amdgcn_make_buffer_rsrc((void *)(F + ((long long)i * W + w) * Nz), 0, Nz * 8, 0x00020000);
            const __amdgpu_buffer_rsrc_t rj =
                __builtin_amdgcn_make_buffer_rsrc((void *)(F + ((long long)j * W + w) * Nz), 0, Nz * 8, 0x00020000);
            // Gf(k) = Y(k) + conj Y(Nz - k) at k (a) and M - k (c), Y = conj(F_i) F_j,
            // folded into LDS as soon as each bin group is in (MODE 2 sums the windows)
            auto fold_store = [&](int h, cf32 av, cf32 cv) {
                if (h == 0 && t == 0) {  // DC and Nyquist of a real signal's spectrum are real
                    av.y = 0.f;
                    cv.y = 0.f;
                }
                const cf32 wk = h == 0 ? w0 : cmul(w0, twc[NT * h]);
                lds[t + NT * h] = c2r_fold(av, cv, wk);
                if (h > 0 || t > 0) lds[M - t - NT * h] = c2r_fold(cv, av, cf32{-wk.x, wk.y});
            };
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const unsigned k = (unsigned)t + NT * h;
                const unsigned kn = (Nz - k) & (Nz - 1), km = M - k, kp = M + k;
                const cf32 y0 = cmulc(ld_bin(ri, 8 * k), ld_bin(rj, 8 * k));
                const cf32 y1 = cmulc(ld_bin(ri, 8 * kn), ld_bin(rj, 8 * kn));
                const cf32 y2 = cmulc(ld_bin(ri, 8 * km), ld_bin(rj, 8 * km));
                const cf32 y3 = cmulc(ld_bin(ri, 8 * kp), ld_bin(rj, 8 * kp));
                const cf32 av = cf32{y0.x + y1.x, y0.y - y1.y}, cv = cf32{y2.x + y3.x, y2.y - y3.y};
                if (MODE == 2) {
                    sa[h] = w == 0 ? av : cf32{sa[h].x + av.x, sa[h].y + av.y};
                    sc[h] = w == 0 ? cv : cf32{sc[h].x + cv.x, sc[h].y + cv.y};
                    if (w + 1 == W) fold_store(h, sa[h], sc[h]);
                } else {
                    fold_store(h, av, cv);
                }
                // two of the four bin groups' loads in flight at a time
                if (h & 1) __builtin_amdgcn_sched_barrier(0);
            }
            cf32 ph = cf32{0.f, 0.f};
            if (t == 0) {
                const cf32 u0 = cmulc(ld_bin(ri, 8 * (M / 2)), ld_bin(rj, 8 * (M / 2)));
                const cf32 u1 = cmulc(ld_bin(ri, 8 * (Nz - M / 2)), ld_bin(rj, 8 * (Nz - M / 2)));
                ph = cf32{u0.x + u1.x, u0.y - u1.y};
            }
            if (MODE == 2) {
                sph = w == 0 ? ph : cf32{sph.x + ph.x, sph.y + ph.y};
                ph = sph;
                if (w + 1 < W) continue;
            }
            if (t == 0) lds[M / 2] = cf32{2.f * ph.x, -2.f * ph.y};
            // an opaque copy of the twiddle table's address each window keeps the
            // compiler from holding every pass's twiddles in registers across the loop
            const cf32 *tw = twM;
            asm volatile("" : "+s"(tw));
            __syncthreads();
            ct_pass_oop<float, M, NT, +1, 8, 1, 0, 2>(lds, alt, tw);
            ct_pass_oop<float, M, NT, +1, 8, 8, 2, 2>(alt, lds, tw + 7);
            ct_pass_oop<float, M, NT, +1, 8, 64, 2, 0>(lds, alt, tw + 7 + 56);
            if (ENDS) ct_last_pass_ends_4096(alt, lds, tw + 7 + 56 + 448);
            else ct_pass_oop<float, M, NT, +1, 8, 512, 0, 0>(alt, lds, tw + 7 + 56 + 448);
            float *dst = y + (MODE == 0 ? ((long long)p * W + w) * L : (long long)p * L);
#pragma unroll
            for (int q = 0; q < LPT; ++q) {
                const int l = t + NT * q;
                if (l < L) {
                    const int n = (lag0 + l) & (Nz - 1);
                    const cf32 v = lds[n >> 1];
                    const float val = scale * ((n & 1) ? v.y : v.x);
                    if (MODE == 1) acc[q] += val;
                    else dst[l] = val;
                }
            }
            if (MODE == 1 && w + 1 == W) {
#pragma unroll
                for (int q = 0; q < LPT; ++q) {
                    const int l = t + NT * q;
                    if (l < L) dst[l] = acc[q] / (float)W;
                }
            }
            // the next fold writes the buffer the last pass read (every lane is past
            // that pass's barrier); these outputs are read before the next
            // transform's first barrier, ahead of its first pass writing them over
            cf32 *const tmp = lds;
            lds = alt;
            alt = tmp;
        }
    }
}

// ---- host side ----
static void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

long long gram_ld(int nk) { return ((long long)nk + 15) / 16 * 16; }

#ifndef FPCC_GSHIFT
#define FPCC_GSHIFT 7
#endif
constexpr int kGGroup = 1 << FPCC_GSHIFT;  // bins per group of the grouped G
// an odd number of pair slots per group keeps a pair's pieces off one memory channel
GLayout gram_grouped(long long P) { return GLayout{(P | 1) * kGGroup, kGGroup, FPCC_GSHIFT}; }

size_t gram_grouped_elems(long long P, int nk)
{
    const long long bins = ((long long)nk + kGF - 1) / kGF * kGF;  // the Gram writes whole chunks
    return (size_t)((bins + kGGroup - 1) / kGGroup) * (size_t)(P | 1) * kGGroup;
}

// G takes bins [0, nchunks * kGF) of pairs [0, S(S-1)/2) with 16-byte stores
static bool layout_ok(const GLayout &gl, int S, int nchunks)
{
    const long long P = (long long)S * (S - 1) / 2;
    if (gl.sh == 30) return gl.gs == 0 && gl.pl >= (long long)nchunks * kGF && !(gl.pl & 1);
    return gl.sh >= 3 && gl.sh < 30 && gl.pl == (1LL << gl.sh) && gl.gs >= P * gl.pl;
}

// Upper-triangle blocks (bi <= bj) of B stations, each holding at least one pair
// i < j, as a device table cached per device and station count.
static const int2 *block_list(int S, int B, int &ntiles)
{
    static std::map<std::tuple<int, int, int>, std::pair<const int2 *, int>> cache;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, S, B});
    if (it != cache.end()) {
        ntiles = it->second.second;
        return it->second.first;
    }
    if (!table_builds_allowed()) return nullptr;
    std::vector<int2> t;
    for (int bi = 0; bi * B < S; ++bi)
        for (int bj = bi; bj * B < S; ++bj) t.push_back(int2{bi, bj});
    void *d = nullptr;
    if (hipMalloc(&d, sizeof(int2) * t.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, t.data(), sizeof(int2) * t.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    cache[{dev, S, B}] = {(const int2 *)d, (int)t.size()};
    ntiles = (int)t.size();
    return (const int2 *)d;
}

const int2 *gram_tiles(int S, int &ntiles) { return block_list(S, kGB, ntiles); }

// The pairs in tiles of T x T stations (tiles row-major over the upper triangle of
// blocks, pairs row-major in a tile) as a device table cached per device and
// station count: order[u] is the u-th pair's row-major index p.
static const int *pair_tile_order(int S, int T)
{
    static std::map<std::tuple<int, int, int>, const int *> cache;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, S, T});
    if (it != cache.end()) return it->second;
    if (!table_builds_allowed()) return nullptr;
    std::vector<int> o;
    o.reserve((size_t)S * (S - 1) / 2);
    for (int bi = 0; bi * T < S; ++bi)
        for (int bj = bi; bj * T < S; ++bj)
            for (int i = bi * T; i < std::min(S, (bi + 1) * T); ++i)
                for (int j = std::max(i + 1, bj * T); j < std::min(S, (bj + 1) * T); ++j)
                    o.push_back((int)((long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1)));
    void *d = nullptr;
    if (o.empty() || hipMalloc(&d, sizeof(int) * o.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, o.data(), sizeof(int) * o.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    cache[{dev, S, T}] = (const int *)d;
    return (const int *)d;
}

float gram16_sigma(int bound)
{
    int e = 0;
    while ((1 << e) < bound && e < 30) ++e;
    return ldexpf(1.0f, 14 - e);
}

int launch_spectra_to_u(const cf32 *d_F, long long ldF, int S, int W, int Nz, int nk, int mirror, void *d_U,
                        hipStream_t st, float sigma)
{
    const int SW = S * W;
    const int R = W <= 64 ? 64 / W * W : 64;  // rows a tile (whole stations), as the kernel
    dim3 grid((nk + 31) / 32, (SW + R - 1) / R);
    ProfScope prof("spectra_to_u", st);
    const int W2 = mirror ? 2 * W : W;
    hipLaunchKernelGGL(k_spectra_to_u, grid, dim3(256), 0, st, d_F, ldF, SW, W, Nz, nk, mirror, sigma,
                       (sigma > 0.f && gram16_grouped(W2)) ? 1 : 0, (uint2 *)d_U);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

static int launch_gram(bool split, const void *d_Ua, const void *d_Ub, int S, int W2, int m0, int K, int nk,
                       const int2 *d_tiles, int ntiles, float oscale, cf32 *d_G, GLayout gl, hipStream_t st)
{
    if (S < 2 || K < 1 || nk < 1 || m0 < 0 || m0 + K > W2) return -1;
    const int nchunks = (nk + kGF - 1) / kGF;
    if (!layout_ok(gl, S, nchunks)) return -1;
    const long long slots = (long long)(nchunks + 15) / 16 * 2 * ntiles;
    if (slots * 8 > 0x7fffffffLL) return -1;
    const void *fn = split ? (const void *)k_gram<1> : (const void *)k_gram<0>;
    allow_lds(fn, kGramLds);
    ProfScope prof(split ? "gram_split16" : "gram_f32", st);
    if (split)
        hipLaunchKernelGGL(k_gram<1>, dim3((unsigned)(slots * 8)), dim3(kGThreads), kGramLds, st,
                           (const uint2 *)d_Ua, (const uint2 *)d_Ub, S, W2, m0, K, nk, d_tiles, ntiles, nchunks,
                           oscale, d_G, gl);
    else
        hipLaunchKernelGGL(k_gram<0>, dim3((unsigned)(slots * 8)), dim3(kGThreads), kGramLds, st,
                           (const uint2 *)d_Ua, (const uint2 *)d_Ub, S, W2, m0, K, nk, d_tiles, ntiles, nchunks,
                           oscale, d_G, gl);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

static int launch_gram_dma(bool f32, const void *d_Ua, const void *d_Ub, int S, int K, int nk, const int2 *d_tiles,
                           int ntiles, float oscale, cf32 *d_G, GLayout gl, hipStream_t st);

int launch_gram_f32(const cf32 *d_Ua, const cf32 *d_Ub, int S, int W2, int m0, int K, int nk, const int2 *d_tiles,
                    int ntiles, cf32 *d_G, GLayout gl, hipStream_t st)
{
    // whole rows of a multiple of 8 values (up to 70): staged by the memory unit straight into LDS
    if (m0 == 0 && K == W2 && K % 8 == 0 && dma_row_units(K) <= 36 && (long long)S * K * 8 * kGF < 0x80000000LL)
        return launch_gram_dma(true, d_Ua, d_Ub, S, K, nk, d_tiles, ntiles, 1.0f, d_G, gl, st);
    return launch_gram(false, d_Ua, d_Ub, S, W2, m0, K, nk, d_tiles, ntiles, 1.0f, d_G, gl, st);
}

bool gram16_grouped(int W2) { return W2 % 16 == 0 && W2 <= kGK; }

// The DMA-staged Gram on rows of K 8-byte elements (f32: FP32 complex; else the
// grouped split words), K a multiple of 8 (16 for the split words) and at most 70:
// three staging buffers when they fit the 160 KB of LDS, else two.  -1 when the
// shape does not qualify.
static int launch_gram_dma(bool f32, const void *d_Ua, const void *d_Ub, int S, int K, int nk, const int2 *d_tiles,
                           int ntiles, float oscale, cf32 *d_G, GLayout gl, hipStream_t st)
{
    if (S < 2 || nk < 1 || K < 8 || K % (f32 ? 8 : 16) || dma_row_units(K) > 36) return -1;
    const int nchunks = (nk + kGF - 1) / kGF;
    if (!layout_ok(gl, S, nchunks)) return -1;
    if ((long long)S * K * 8 * kGF >= 0x80000000LL) return -1;  // chunk offsets stay below the pad offset
    if (S <= 32 && d_Ua == d_Ub) {
        // one diagonal block: every wave on a bin of its own
        const long long slots = (long long)(nchunks + 15) / 16 * 2;
        const size_t lds = std::max((size_t)32 * 32 * 64, (size_t)kGF * ((dma_row_units(K) + 1) / 2) * 1024);
        const void *fn = f32 ? (const void *)k_gram_small<true> : (const void *)k_gram_small<false>;
        allow_lds(fn, lds);
        ProfScope prof(f32 ? "gram_f32" : "gram_split16", st);
        if (f32)
            hipLaunchKernelGGL(k_gram_small<true>, dim3((unsigned)(slots * 8)), dim3(kGThreads), lds, st,
                               (const uint2 *)d_Ua, S, K, nk, nchunks, oscale, d_G, gl);
        else
            hipLaunchKernelGGL(k_gram_small<false>, dim3((unsigned)(slots * 8)), dim3(kGThreads), lds, st,
                               (const uint2 *)d_Ua, S, K, nk, nchunks, oscale, d_G, gl);
        FPCC_CHECK(hipGetLastError());
        return 0;
    }
    const long long slots = (long long)(nchunks + 15) / 16 * 2 * ntiles;
    if (slots * 8 > 0x7fffffffLL) return -1;
    const size_t bufb = (size_t)2 * kGB * dma_row_units(K) * 16;
    const size_t epi = (size_t)kGB * 16 * 64;  // an epilogue quarter: 64 x 16 pairs of 64 bytes
    const bool three = 3 * bufb <= 160 * 1024;
    const size_t lds = std::max(epi, (three ? 3 : 2) * bufb);
    const void *fn = f32 ? (three ? (const void *)k_gram16_dma<3, true> : (const void *)k_gram16_dma<2, true>)
                         : (three ? (const void *)k_gram16_dma<3, false> : (const void *)k_gram16_dma<2, false>);
    allow_lds(fn, lds);
    ProfScope prof(f32 ? "gram_f32" : "gram_split16", st);
#define FPCC_GRAM_DMA(NB, F)                                                                                       \
    hipLaunchKernelGGL((k_gram16_dma<NB, F>), dim3((unsigned)(slots * 8)), dim3(kGThreads), lds, st,             \
                       (const uint2 *)d_Ua, (const uint2 *)d_Ub, S, K, nk, d_tiles, ntiles, nchunks, oscale, d_G, gl)
    if (f32) {
        if (three) FPCC_GRAM_DMA(3, true);
        else FPCC_GRAM_DMA(2, true);
    } else {
        if (three) FPCC_GRAM_DMA(3, false);
        else FPCC_GRAM_DMA(2, false);
    }
#undef FPCC_GRAM_DMA
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_gram_split16(const void *d_Ua, const void *d_Ub, int S, int W2, int m0, int K, int nk,
                        const int2 *d_tiles, int ntiles, float sigma, cf32 *d_G, GLayout gl, hipStream_t st)
{
    if (!(sigma > 0.f)) return -1;
    if (!gram16_grouped(W2))
        return launch_gram(true, d_Ua, d_Ub, S, W2, m0, K, nk, d_tiles, ntiles, 1.0f / (sigma * sigma), d_G, gl,
                           st);
    // grouped rows (written so by launch_spectra_to_u): the whole row, DMA-staged
    if (m0 != 0 || K != W2) return -1;
    static_assert(2 * 2 * kGB * dma_row_units(kGK) * 16 <= kGramLds, "DMA staging fits the Gram's LDS");
    return launch_gram_dma(false, d_Ua, d_Ub, S, K, nk, d_tiles, ntiles, 1.0f / (sigma * sigma), d_G, gl, st);
}

// 512 threads, radix 8 x 8 x 8 x 8: eight values a thread per pass and the next
// pair's bins in flight keep the kernel within 128 registers, two workgroups a CU
// (256 threads with radix 16 x 16 x 16 and no prefetch spill and run 19% slower)
constexpr int kC2rThreads = 512;

template <bool COUNT>
static int launch_c2r_4096(const cf32 *d_G, const GLayout &gl, int P, const cf32 *twr, const cf32 *twc, int lag0,
                           int L, float scale, int N, float *d_y, const C2rEdge &ed, hipStream_t st)
{
    constexpr int kT = kC2rThreads;
    // lags within +-1024: the last pass forms only the outputs they read
    const bool ends = lag0 >= -1024 && lag0 + L - 1 <= 1023;
    const auto fn = ends ? k_pair_c2r_4096<kT, COUNT, true> : k_pair_c2r_4096<kT, COUNT, false>;
    const size_t lds = sizeof(cf32) * 2 * 4096;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return -1;
    // one persistent workgroup per resident slot (the query is the same on every
    // call and device of this build; concurrent first calls store the same value)
    static std::atomic<int> per_cu_cached[2] = {{0}, {0}};
    int per_cu = per_cu_cached[ends].load(std::memory_order_relaxed);
    if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, kT, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        per_cu_cached[ends].store(per_cu, std::memory_order_relaxed);
    }
    const int nwg = P < per_cu * ncu ? P : per_cu * ncu;
    ProfScope prof("pair_c2r", st);
    hipLaunchKernelGGL(fn, dim3(nwg), dim3(kT), lds, st, d_G, gl, P, twr, twc, lag0, L, scale, N, d_y, ed);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

template <int LPT, int MODE, bool ENDS>
static int launch_plain_4096_t(const cf32 *d_F, int S, int W, int P, const cf32 *twr, const cf32 *twc, int lag0,
                               int L, float scale, float *d_y, hipStream_t st, const char *name)
{
    // tiles of 16 stations: two tiles' spectra (32 stations x W windows x 64 KB, 49 MB
    // at W 24) cover the pairs the resident workgroups run at once
    const int *order = pair_tile_order(S, 16);
    if (!order) return -1;
    constexpr int kT = kC2rThreads;
    const auto fn = k_pair_plain_4096<kT, LPT, MODE, ENDS, false>;
    const size_t lds = sizeof(cf32) * 2 * 4096;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return -1;
    static std::atomic<int> per_cu_cached{0};  // the same on every call and device of this build
    int per_cu = per_cu_cached.load(std::memory_order_relaxed);
    if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, kT, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        per_cu_cached.store(per_cu, std::memory_order_relaxed);
    }
    const int nwg = P < per_cu * ncu ? P : per_cu * ncu;
    ProfScope prof(name, st);
    hipLaunchKernelGGL(fn, dim3(nwg), dim3(kT), lds, st, d_F, S, W, P, order, twr, twc, lag0, L, scale, d_y,
                       PlainBatch{});
    FPCC_CHECK(hipGetLastError());
    return 0;
}

template <int MODE>
static int launch_plain_4096_m(int lpt, const cf32 *d_F, int S, int W, int P, const cf32 *twr, const cf32 *twc,
                               int lag0, int L, float scale, float *d_y, hipStream_t st, const char *name)
{
    // lags within +-1024: the last pass forms only the outputs they read
    const bool ends = lag0 >= -1024 && lag0 + L - 1 <= 1023;
#define FPCC_PLAIN(LPT, E) launch_plain_4096_t<LPT, MODE, E>(d_F, S, W, P, twr, twc, lag0, L, scale, d_y, st, name)
    switch (lpt) {
    case 1: return ends ? FPCC_PLAIN(1, true) : FPCC_PLAIN(1, false);
    case 2: return ends ? FPCC_PLAIN(2, true) : FPCC_PLAIN(2, false);
    case 4: return ends ? FPCC_PLAIN(4, true) : FPCC_PLAIN(4, false);
    case 8: return FPCC_PLAIN(8, false);
    case 16: return FPCC_PLAIN(16, false);
    }
#undef FPCC_PLAIN
    return -1;
}

int launch_pair_pcc2_4096(const cf32 *d_F, int S, int W, int N, int lag0, int L, int mode, float *d_y,
                          hipStream_t st)
{
    constexpr int Nz = 8192;
    const long long P = (long long)S * (S - 1) / 2;
    if (P < 1 || P > 0x7fffffff || W < 1 || L < 1 || L > Nz || mode < 0 || mode > 2) return -1;
    const void *twc = get_c2r_twiddles(Nz, false);
    const cf32 *twr = (const cf32 *)get_order_twiddles(4096, {8, 8, 8, 8});
    if (!twc || !twr) return -1;
    int lpt = 1;
    while (lpt * kC2rThreads < L) lpt *= 2;
    // the folded spectra are twice the half spectrum of the real correlation: 1/2
    // times the reference's 1/((float)Nz*(float)N) (mode 2 also takes the mean's 1/W)
    float scale = 0.5f * (1 / ((float)Nz * (float)N));
    if (mode == 2) scale /= (float)W;
    const cf32 *tc = (const cf32 *)twc;
    switch (mode) {
    case 0: return launch_plain_4096_m<0>(lpt, d_F, S, W, (int)P, twr, tc, lag0, L, scale, d_y, st, "pair_pcc2_none");
    case 1: return launch_plain_4096_m<1>(lpt, d_F, S, W, (int)P, twr, tc, lag0, L, scale, d_y, st, "pair_pcc2_plain");
    default: return launch_plain_4096_m<2>(lpt, d_F, S, W, (int)P, twr, tc, lag0, L, scale, d_y, st, "pair_pcc2_zsum");
    }
}

// The per-window PCC2 pair stages of several jobs at Nz 8192 (one stack mode) in
// one launch of the persistent pair kernel: a job table of spectra, station and
// window counts, lag windows and outputs, the pairs of all jobs walked as one
// sequence (each job's in its own tile order).
template <int LPT, int MODE, bool ENDS>
static int launch_plain_4096_batch_t(const PlainBatch &jb, int P, const cf32 *twr, const cf32 *twc, hipStream_t st)
{
    constexpr int kT = kC2rThreads;
    const auto fn = k_pair_plain_4096<kT, LPT, MODE, ENDS, true>;
    const size_t lds = sizeof(cf32) * 2 * 4096;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return -1;
    static std::atomic<int> per_cu_cached{0};
    int per_cu = per_cu_cached.load(std::memory_order_relaxed);
    if (!per_cu) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, kT, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        per_cu_cached.store(per_cu, std::memory_order_relaxed);
    }
    const int nwg = P < per_cu * ncu ? P : per_cu * ncu;
    ProfScope prof(MODE == 0 ? "pair_pcc2_none_batch" : MODE == 1 ? "pair_pcc2_plain_batch" : "pair_pcc2_zsum_batch",
                   st);
    hipLaunchKernelGGL(fn, dim3(nwg), dim3(kT), lds, st, (const cf32 *)nullptr, 0, 0, P, (const int *)nullptr, twr,
                       twc, 0, 0, 0.f, (float *)nullptr, jb);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

template <int MODE>
static int launch_plain_4096_batch_m(int lpt, bool ends, const PlainBatch &jb, int P, const cf32 *twr,
                                     const cf32 *twc, hipStream_t st)
{
#define FPCC_PLAINB(LPT, E) launch_plain_4096_batch_t<LPT, MODE, E>(jb, P, twr, twc, st)
    switch (lpt) {
    case 1: return ends ? FPCC_PLAINB(1, true) : FPCC_PLAINB(1, false);
    case 2: return ends ? FPCC_PLAINB(2, true) : FPCC_PLAINB(2, false);
    case 4: return ends ? FPCC_PLAINB(4, true) : FPCC_PLAINB(4, false);
    case 8: return FPCC_PLAINB(8, false);
    case 16: return FPCC_PLAINB(16, false);
    }
#undef FPCC_PLAINB
    return -1;
}

int launch_pair_pcc2_4096_batch(int n, const cf32 *const *d_F, const int *S, const int *W, const int *N,
                                const int *lag0, const int *L, int mode, float *const *d_y, hipStream_t st)
{
    constexpr int Nz = 8192;
    if (n < 1 || n > kMaxBatchJobs || mode < 0 || mode > 2) return -1;
    const void *twc = get_c2r_twiddles(Nz, false);
    const cf32 *twr = (const cf32 *)get_order_twiddles(4096, {8, 8, 8, 8});
    if (!twc || !twr) return -1;
    PlainBatch jb{};
    jb.n = n;
    long long P = 0;
    int Lmax = 1;
    bool ends = true;
    for (int j = 0; j < n; ++j) {
        const long long Pj = (long long)S[j] * (S[j] - 1) / 2;
        if (Pj < 1 || W[j] < 1 || L[j] < 1 || L[j] > Nz) return -1;
        jb.order[j] = pair_tile_order(S[j], 16);
        if (!jb.order[j]) return -1;
        jb.u0[j] = (int)P;
        P += Pj;
        jb.F[j] = d_F[j];
        jb.S[j] = S[j];
        jb.W[j] = W[j];
        jb.lag0[j] = lag0[j];
        jb.L[j] = L[j];
        // as launch_pair_pcc2_4096: 1/2 times 1/((float)Nz*(float)N), mode 2 also 1/W
        float sc = 0.5f * (1 / ((float)Nz * (float)N[j]));
        if (mode == 2) sc /= (float)W[j];
        jb.scale[j] = sc;
        jb.y[j] = d_y[j];
        Lmax = std::max(Lmax, L[j]);
        ends = ends && lag0[j] >= -1024 && lag0[j] + L[j] - 1 <= 1023;
    }
    if (P > 0x7fffffff) return -1;
    jb.u0[n] = (int)P;
    int lpt = 1;
    while (lpt * kC2rThreads < Lmax) lpt *= 2;
    const cf32 *tc = (const cf32 *)twc;
    switch (mode) {
    case 0: return launch_plain_4096_batch_m<0>(lpt, ends, jb, (int)P, twr, tc, st);
    case 1: return launch_plain_4096_batch_m<1>(lpt, ends, jb, (int)P, twr, tc, st);
    default: return launch_plain_4096_batch_m<2>(lpt, ends, jb, (int)P, twr, tc, st);
    }
}

int launch_pair_c2r(const cf32 *d_G, GLayout gl, int P, int Nz, int N, int W, int lag0, int L, float scale,
                    int count_norm, float *d_y, hipStream_t st, const float *d_edge, int S)
{
    C2rEdge ed;
    if (d_edge && count_norm) {
        ed.edge = d_edge;
        ed.K = edge_len(N);
        ed.S = S;
        ed.W = W;
    }
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, false);
    const void *twc = get_c2r_twiddles(Nz, false);
    if (!pM.ok || !twc) return -1;
    // k_pair_c2r_4096 splits a bin's address into a scalar part and a 32-bit offset
    // below (kT >> sh) groups
    constexpr int kT = kC2rThreads;
    const bool split_ok = gl.sh == 30 || ((1 << gl.sh) <= kT && (long long)(kT >> gl.sh) * gl.gs + kT < (1LL << 29));
    if (M == 4096 && split_ok) {
        const cf32 *twr = (const cf32 *)get_order_twiddles(4096, {8, 8, 8, 8});
        if (!twr) return -1;
        return count_norm ? launch_c2r_4096<true>(d_G, gl, P, twr, (const cf32 *)twc, lag0, L, scale, N, d_y, ed, st)
                          : launch_c2r_4096<false>(d_G, gl, P, twr, (const cf32 *)twc, lag0, L, scale, N, d_y, ed, st);
    }
    size_t lds = sizeof(cf32) * 2 * (size_t)M;
    allow_lds((const void *)k_pair_c2r, lds);
    ProfScope prof("pair_c2r", st);
    hipLaunchKernelGGL(k_pair_c2r, dim3(P), dim3(512), lds, st, d_G, gl, Nz, pM.plan,
                       (const cf32 *)pM.tw, (const cf32 *)twc, lag0, L, scale, count_norm, N, d_y, ed);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc
