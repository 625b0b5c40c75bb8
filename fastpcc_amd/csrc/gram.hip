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
//   k_gram_*         U -> G[p][k] (pair-major rows of ld >= Nz/2+1 bins)
//   k_pair_c2r       G[p] -> half-length complex IFFT -> lag window -> y[p][L]
#include <map>
#include <mutex>
#include <vector>
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"

namespace fpcc {

typedef float f4v __attribute__((ext_vector_type(4)));

// ---- F[sw][Nz] -> U[k][s][2W], k in [0, M] ----
// Tile: 32 frequencies x 64 station-windows, staged in LDS for both the direct
// and the mirrored frequency.
// mirror = 1 (analytic-signal spectra, Nz bins per row): m in [0, 2W) with the
// conjugated mirror frequency in the second half; mirror = 0 (half spectra of real
// signals, ldF >= Nz/2+1 bins per row): m in [0, W).
__global__ void __launch_bounds__(256)
k_spectra_to_u(const cf32 *__restrict__ F, long long ldF, int SW, int W, int Nz, int nk, int mirror,
               cf32 *__restrict__ U, float *__restrict__ umax)
{
    __shared__ cf32 td[64][33];
    __shared__ cf32 tm[64][33];
    __shared__ float red[4];
    float amax = 0.f;
    const int k0 = blockIdx.x * 32;
    const int sw0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 64; r += 8) {
        const int sw = sw0 + r;
        const int k = k0 + tx;
        cf32 d = cf32{0.f, 0.f}, m = cf32{0.f, 0.f};
        if (sw < SW && k < nk) {
            d = F[(long long)sw * ldF + k];
            if (mirror) {
                cf32 mm = F[(long long)sw * ldF + ((Nz - k) & (Nz - 1))];
                m = cf32{mm.x, -mm.y};
            }
        }
        td[r][tx] = d;
        tm[r][tx] = m;
        amax = fmaxf(amax, fmaxf(fmaxf(fabsf(d.x), fabsf(d.y)), fmaxf(fabsf(m.x), fabsf(m.y))));
    }
    __syncthreads();
    if (umax) {  // largest |component| of U, for the split-FP16 Gram's scaling
        amax = block_max(amax, red);
        if (threadIdx.x == 0) atomicMax(reinterpret_cast<int *>(umax), __float_as_int(amax));
    }
    // write: for each k of the tile, 64 station-windows (contiguous w runs per station)
    const int W2 = mirror ? 2 * W : W;
    const int S = SW / W;
    for (int idx = threadIdx.x; idx < 32 * 64; idx += 256) {
        const int kk = idx >> 6, r = idx & 63;
        const int k = k0 + kk, sw = sw0 + r;
        if (k >= nk || sw >= SW) continue;
        const int s = sw / W, w = sw - s * W;
        cf32 *row = U + ((long long)k * S + s) * W2;
        row[w] = td[r][kk];
        if (mirror) row[W + w] = tm[r][kk];
    }
}

// ---- Gram on MFMA f32 16x16x4 ----
// Workgroup: station tile I (16) x J (32) for a chunk of 16 frequencies; wave q
// computes frequencies k0+q, k0+q+4, ...  Operand layouts (verified on gfx950):
// A[l%16][l/16], B[l/16][l%16], D[4*(l/16)+r][l%16].
// K index: (m, component) with m = 4*mm + l/16 over 2W = 4*nmm complex values.
constexpr int kGramKc = 16;

__global__ void __launch_bounds__(256)
k_gram_f32(const cf32 *__restrict__ Ua, const cf32 *__restrict__ Ub, int S, int W2, int m0, int K, int nk,
           const int2 *__restrict__ tiles, int ntiles, int nchunks, cf32 *__restrict__ G, long long ldg)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *stage = reinterpret_cast<cf32 *>(smem_raw);  // [16*32 pairs][kGramKc]
    // XCD-aware order: the 8 XCDs each sweep their own frequency chunks so that a
    // chunk's U rows are fetched into one L2 and shared by all tiles of the chunk.
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int chunk = (slot / ntiles) * 8 + xcd;
    const int tile = slot % ntiles;
    if (chunk >= nchunks) return;
    const int bi = tiles[tile].x, bj = tiles[tile].y;
    const int k0 = chunk * kGramKc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int si = bi * 16 + li, sj0 = bj * 32 + li, sj1 = sj0 + 16;
    const int nmm = (K + 3) / 4;  // K padded with zeros to a multiple of 4 complex values

    for (int t = 0; t < kGramKc / 4; ++t) {
        const int kl = wave + 4 * t;
        const int k = k0 + kl;
        f4v r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
        if (k < nk) {
            const cf32 *pa = Ua + ((long long)k * S + si) * W2 + m0;
            const cf32 *pb0 = Ub + ((long long)k * S + sj0) * W2 + m0;
            const cf32 *pb1 = Ub + ((long long)k * S + sj1) * W2 + m0;
            const bool va = si < S, vb0 = sj0 < S, vb1 = sj1 < S;
            for (int mm = 0; mm < nmm; ++mm) {
                const int m = 4 * mm + q;
                const bool vm = m < K;
                cf32 a = (va && vm) ? pa[m] : cf32{0.f, 0.f};
                cf32 b0 = (vb0 && vm) ? pb0[m] : cf32{0.f, 0.f};
                cf32 b1 = (vb1 && vm) ? pb1[m] : cf32{0.f, 0.f};
                r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b0.x, r0, 0, 0, 0);
                r0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b0.y, r0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b0.y, i0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, -b0.x, i0, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b1.x, r1, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b1.y, r1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b1.y, i1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, -b1.x, i1, 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int il = 4 * q + r;
            stage[(il * 32 + li) * kGramKc + kl] = cf32{r0[r], i0[r]};
            stage[(il * 32 + li + 16) * kGramKc + kl] = cf32{r1[r], i1[r]};
        }
    }
    __syncthreads();
    // write pair rows: 16 consecutive threads store the 16 bins of one pair
    const int kk = threadIdx.x & 15;
    for (int pl = threadIdx.x >> 4; pl < 16 * 32; pl += 16) {
        const int i = bi * 16 + (pl >> 5), j = bj * 32 + (pl & 31);
        if (i >= j || j >= S) continue;
        const long long p = (long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1);
        G[p * ldg + k0 + kk] = stage[pl * kGramKc + kk];
    }
}

// ---- Gram on MFMA f16 16x16x32 with split hi+lo operands ----
// Every U component v (scaled by sigma = 2^e, exact) is split into
// hi = fp16(v), lo = fp16(v - hi); products use hi*hi + hi*lo + lo*hi with FP32
// accumulation (the dropped lo*lo is below 2^-22 relative).  sigma puts the
// largest |component| near 2^14, inside FP16's range with normal lo parts down
// to 2^-8 of it; results are multiplied by 1/sigma^2.  Operand layouts (verified
// on gfx950): A[l%16][8*(l/16)+e], B[8*(l/16)+e][l%16]; the K index (m, c)
// interleaves components so a lane's 8 values are 4 consecutive complex U[m].
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split8(const float *v, float sigma, h8v &hi, h8v &lo)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float s = v[e] * sigma;
        const _Float16 h = (_Float16)s;
        hi[e] = h;
        lo[e] = (_Float16)(s - (float)h);
    }
}

__global__ void __launch_bounds__(256)
k_gram_split16(const cf32 *__restrict__ Ua, const cf32 *__restrict__ Ub, int S, int W2, int m0, int K, int nk,
               const int2 *__restrict__ tiles, int ntiles, int nchunks, const float *__restrict__ umax,
               cf32 *__restrict__ G, long long ldg)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *stage = reinterpret_cast<cf32 *>(smem_raw);  // [16*32 pairs][kGramKc]
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int chunk = (slot / ntiles) * 8 + xcd;
    const int tile = slot % ntiles;
    if (chunk >= nchunks) return;
    const int bi = tiles[tile].x, bj = tiles[tile].y;
    const int k0 = chunk * kGramKc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int si = bi * 16 + li, sj0 = bj * 32 + li, sj1 = sj0 + 16;
    const bool va = si < S, vb0 = sj0 < S, vb1 = sj1 < S;
    // sigma = 2^(14 - ceil(log2 max)), exact; inv2 = 1/sigma^2
    const float mx = umax[0];
    int ex = 0;
    if (mx > 0.f) (void)frexpf(mx, &ex);
    const float sigma = ldexpf(1.0f, 14 - ex), inv2 = ldexpf(1.0f, 2 * (ex - 14));
    const int nsteps = (K + 15) / 16;  // 16 complex values (K = 32 reals) per MFMA

    for (int t = 0; t < kGramKc / 4; ++t) {
        const int kl = wave + 4 * t;
        const int k = k0 + kl;
        f4v r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
        if (k < nk) {
            const cf32 *pa = Ua + ((long long)k * S + si) * W2 + m0;
            const cf32 *pb0 = Ub + ((long long)k * S + sj0) * W2 + m0;
            const cf32 *pb1 = Ub + ((long long)k * S + sj1) * W2 + m0;
            for (int stp = 0; stp < nsteps; ++stp) {
                float av[8], b0v[8], b1v[8], b0r[8], b1r[8];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int m = 16 * stp + 4 * q + u;
                    const bool vm = m < K;
                    const cf32 a = (va && vm) ? pa[m] : cf32{0.f, 0.f};
                    const cf32 c0 = (vb0 && vm) ? pb0[m] : cf32{0.f, 0.f};
                    const cf32 c1 = (vb1 && vm) ? pb1[m] : cf32{0.f, 0.f};
                    av[2 * u] = a.x;
                    av[2 * u + 1] = a.y;
                    b0v[2 * u] = c0.x;
                    b0v[2 * u + 1] = c0.y;
                    b1v[2 * u] = c1.x;
                    b1v[2 * u + 1] = c1.y;
                    b0r[2 * u] = c0.y;  // Im G: (ur_i, ui_i) . (ui_j, -ur_j)
                    b0r[2 * u + 1] = -c0.x;
                    b1r[2 * u] = c1.y;
                    b1r[2 * u + 1] = -c1.x;
                }
                h8v ah, al, bh, bl;
                split8(av, sigma, ah, al);
                split8(b0v, sigma, bh, bl);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, r0, 0, 0, 0);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, r0, 0, 0, 0);
                r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, r0, 0, 0, 0);
                split8(b0r, sigma, bh, bl);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, i0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, i0, 0, 0, 0);
                i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, i0, 0, 0, 0);
                split8(b1v, sigma, bh, bl);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, r1, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, r1, 0, 0, 0);
                r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, r1, 0, 0, 0);
                split8(b1r, sigma, bh, bl);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, i1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, i1, 0, 0, 0);
                i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, i1, 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int il = 4 * q + r;
            stage[(il * 32 + li) * kGramKc + kl] = cf32{r0[r] * inv2, i0[r] * inv2};
            stage[(il * 32 + li + 16) * kGramKc + kl] = cf32{r1[r] * inv2, i1[r] * inv2};
        }
    }
    __syncthreads();
    const int kk = threadIdx.x & 15;
    for (int pl = threadIdx.x >> 4; pl < 16 * 32; pl += 16) {
        const int i = bi * 16 + (pl >> 5), j = bj * 32 + (pl & 31);
        if (i >= j || j >= S) continue;
        const long long p = (long long)i * S - (long long)i * (i + 1) / 2 + (j - i - 1);
        G[p * ldg + k0 + kk] = stage[pl * kGramKc + kk];
    }
}

// ---- per pair: folded spectrum -> real lags ----
// Z[k] = P + i e^{+2 pi i k/Nz} Q, P = G[k] + conj(G[M-k]), Q = G[k] - conj(G[M-k])
__global__ void __launch_bounds__(512)
k_pair_c2r(const cf32 *__restrict__ G, long long ldg, int Nz, FftPlan planM, const cf32 *__restrict__ twM,
           const cf32 *__restrict__ twc, int lag0, int L, float scale, int count_norm, int N,
           float *__restrict__ y)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = Nz >> 1;
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + M;
    const long long p = blockIdx.x;
    const cf32 *g = G + p * ldg;
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        cf32 a = g[k], c = g[M - k];
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
    float *dst = y + p * L;
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        const int n = (lag0 + l) & (Nz - 1);
        cf32 v = z[n >> 1];
        float r = scale * ((n & 1) ? v.y : v.x);
        if (count_norm) r = (float)((double)r / (double)(N - abs(lag0 + l)));
        dst[l] = r;
    }
}

// Specialisation for Nz = 8192 (M = 4096 = 16^3): compile-time in-place FFT, 35 KB
// of LDS, four workgroups per CU.
__global__ void __launch_bounds__(256)
k_pair_c2r_4096(const cf32 *__restrict__ G, long long ldg, const cf32 *__restrict__ twM,
                const cf32 *__restrict__ twc, int lag0, int L, float scale, int count_norm, int N,
                float *__restrict__ y)
{
    constexpr int M = 4096, Nz = 8192;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *lds = reinterpret_cast<cf32 *>(smem_raw);
    const long long p = blockIdx.x;
    const cf32 *g = G + p * ldg;
    for (int k = threadIdx.x; k < M; k += 256) {
        cf32 a = g[k], c = g[M - k];
        if (k == 0) {
            a.y = 0.f;
            c.y = 0.f;
        }
        cf32 P = cf32{a.x + c.x, a.y - c.y};
        cf32 Q = cf32{a.x - c.x, a.y + c.y};
        cf32 wq = cmul(twc[k], Q);
        lds[ct_pad(k)] = cf32{P.x - wq.y, P.y + wq.x};
    }
    __syncthreads();
    ct_fft<float, M, 256, +1, 16, 16, 16>(lds, twM);
    float *dst = y + p * L;
    for (int l = threadIdx.x; l < L; l += 256) {
        const int n = (lag0 + l) & (Nz - 1);
        cf32 v = lds[ct_pad(n >> 1)];
        float r = scale * ((n & 1) ? v.y : v.x);
        if (count_norm) r = (float)((double)r / (double)(N - abs(lag0 + l)));
        dst[l] = r;
    }
}

// ---- host side ----
static void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

long long gram_ld(int nk) { return ((long long)nk + kGramKc - 1) / kGramKc * kGramKc; }

// Station tiles (16 rows x 32 columns) holding at least one pair i < j, as a
// device table cached per device and station count.
const int2 *gram_tiles(int S, int &ntiles)
{
    static std::map<std::pair<int, int>, std::pair<const int2 *, int>> cache;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, S});
    if (it != cache.end()) {
        ntiles = it->second.second;
        return it->second.first;
    }
    std::vector<int2> t;
    for (int bi = 0; bi * 16 < S; ++bi)
        for (int bj = 0; bj * 32 < S; ++bj)
            if (bj * 32 + 31 > bi * 16) t.push_back(int2{bi, bj});
    void *d = nullptr;
    if (hipMalloc(&d, sizeof(int2) * t.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, t.data(), sizeof(int2) * t.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    cache[{dev, S}] = {(const int2 *)d, (int)t.size()};
    ntiles = (int)t.size();
    return (const int2 *)d;
}

int launch_spectra_to_u(const cf32 *d_F, long long ldF, int S, int W, int Nz, int nk, int mirror, cf32 *d_U,
                        hipStream_t st, float *d_umax)
{
    const int SW = S * W;
    dim3 grid((nk + 31) / 32, (SW + 63) / 64);
    ProfScope prof("spectra_to_u", st);
    hipLaunchKernelGGL(k_spectra_to_u, grid, dim3(256), 0, st, d_F, ldF, SW, W, Nz, nk, mirror, d_U, d_umax);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_gram_f32(const cf32 *d_Ua, const cf32 *d_Ub, int S, int W2, int m0, int K, int nk, const int2 *d_tiles,
                    int ntiles, cf32 *d_G, long long ldg, hipStream_t st)
{
    const int nchunks = (nk + kGramKc - 1) / kGramKc;
    if (ldg < (long long)nchunks * kGramKc) return -1;
    const int slots = (nchunks + 7) / 8 * ntiles;
    size_t lds = sizeof(cf32) * 16 * 32 * kGramKc;
    allow_lds((const void *)k_gram_f32, lds);
    ProfScope prof("gram_f32", st);
    hipLaunchKernelGGL(k_gram_f32, dim3(slots * 8), dim3(256), lds, st, d_Ua, d_Ub, S, W2, m0, K, nk, d_tiles,
                       ntiles, nchunks, d_G, ldg);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_gram_split16(const cf32 *d_Ua, const cf32 *d_Ub, int S, int W2, int m0, int K, int nk,
                        const int2 *d_tiles, int ntiles, const float *d_umax, cf32 *d_G, long long ldg,
                        hipStream_t st)
{
    const int nchunks = (nk + kGramKc - 1) / kGramKc;
    if (ldg < (long long)nchunks * kGramKc) return -1;
    const int slots = (nchunks + 7) / 8 * ntiles;
    size_t lds = sizeof(cf32) * 16 * 32 * kGramKc;
    allow_lds((const void *)k_gram_split16, lds);
    ProfScope prof("gram_split16", st);
    hipLaunchKernelGGL(k_gram_split16, dim3(slots * 8), dim3(256), lds, st, d_Ua, d_Ub, S, W2, m0, K, nk, d_tiles,
                       ntiles, nchunks, d_umax, d_G, ldg);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_pair_c2r(const cf32 *d_G, int P, int Nz, int N, int W, int lag0, int L, float scale, int count_norm,
                    float *d_y, hipStream_t st)
{
    const int M = Nz / 2;
    const DevPlan &pM = get_plan(M, false);
    const void *twc = get_c2r_twiddles(Nz, false);
    if (!pM.ok || !twc) return -1;
    if (M == 4096 && pM.plan.npass == 3 && pM.plan.radix[0] == 16 && pM.plan.radix[1] == 16 &&
        pM.plan.radix[2] == 16) {
        size_t lds = sizeof(cf32) * ct_lds_elems<4096>();
        ProfScope prof("pair_c2r", st);
        hipLaunchKernelGGL(k_pair_c2r_4096, dim3(P), dim3(256), lds, st, d_G, gram_ld(Nz / 2 + 1),
                           (const cf32 *)pM.tw, (const cf32 *)twc, lag0, L, scale, count_norm, N, d_y);
        FPCC_CHECK(hipGetLastError());
        return 0;
    }
    size_t lds = sizeof(cf32) * 2 * (size_t)M;
    allow_lds((const void *)k_pair_c2r, lds);
    ProfScope prof("pair_c2r", st);
    hipLaunchKernelGGL(k_pair_c2r, dim3(P), dim3(512), lds, st, d_G, gram_ld(Nz / 2 + 1), Nz, pM.plan,
                       (const cf32 *)pM.tw, (const cf32 *)twc, lag0, L, scale, count_norm, N, d_y);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc
