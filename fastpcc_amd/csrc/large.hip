// Station and pair stages for windows longer than one workgroup's LDS
// (N > 8192 or Nz > 8192) and for analytic signals of lengths that are not 7-smooth,
// on the two-pass FFTs of bigfft.hip.
//
// PCC2: analytic signal as DFT_N -> mask -> IDFT_N (7-smooth N: four-step with the
//   mask and both row passes fused; any other N: Bluestein chirp convolutions on a
//   power-of-two grid), max|z|^2 per window, AmpNormf and zero pad
//   fused into the first pass of FFT_Nz; spectra stay in permuted bin order.
// PCC1 / PCC-v: the same analytic signal, AmpNormf per row, then the time-domain kernel.
// GNCC / 1-bit: complex FFT_Nz of the zero-padded samples (or signs); energies as
//   FP64 prefix sums.  (GNCC runs its large-N FFTs in FP32.)
// Pair stage: inverse row pass with the cross spectrum conj(F_i) F_j (summed over
//   windows when the normalisation allows it) as prologue, then the inverse column
//   pass writing the normalised lag window; PCC2 and 1-bit can instead stack through
//   the MFMA Gram over all Nz bins.
#include <algorithm>
#include "../../include/fastpcc.h"
#include "bigfft.hpp"
#include "kernels.hpp"
#include "profile.hpp"

namespace fpcc {

// GNCC and 1-bit on two-pass transforms (FP32): lags whose overlap is shorter than
// edge_len(N) samples are summed exactly from the first and last edge_len(N)
// samples of each window, kept beside the spectra (k_big_cols, COL_INV_LAGWIN).
// Below N/64 samples of overlap the transform's absolute error, divided by the
// overlap's energy, could exceed 1e-6; above it the error is at most 64 times its
// full-overlap size (a few 1e-9 measured).
int edge_len(int N) { return std::min(N, std::max(64, (N + 63) / 64)); }

__global__ void __launch_bounds__(256)
k_edge_copy(const float *__restrict__ x, long long stride, int W, int N, int K, float *__restrict__ edge)
{
    const int sw = blockIdx.x, s = sw / W, w = sw - s * W;
    const float *in = x + (long long)s * stride + (long long)w * N;
    float *out = edge + (long long)sw * 2 * K;
    for (int i = threadIdx.x; i < 2 * K; i += blockDim.x) out[i] = in[i < K ? i : N - 2 * K + i];
}

int large_edge_copy(const float *x, long long stride, int S, int W, int N, float *edge, hipStream_t st)
{
    if (S * W <= 0) return 0;
    ProfScope prof("edge_copy", st);
    hipLaunchKernelGGL(k_edge_copy, dim3(S * W), dim3(256), 0, st, x, stride, W, N, edge_len(N), edge);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

namespace {
size_t al(size_t v) { return (v + 255) / 256 * 256; }

struct LargeLayout {
    size_t Z = 0, F = 0, F2 = 0, mx = 0, E = 0, X = 0, C = 0, U = 0, U2 = 0, G = 0, B = 0, total = 0;
};

// Analytic signals of lengths without a two-pass split run as Bluestein
// convolutions on a power-of-two grid; 0 when N has a split.
int bluestein_grid(int N) { return big_plan(N).ok ? 0 : bluestein_n(N); }

// Bluestein scratch: the SW rows of the grid, then the chirp (N) and kernel
// spectrum (n) tables.
size_t bluestein_bytes(int SW, int N)
{
    const int n = bluestein_grid(N);
    return n ? sizeof(cf32) * ((size_t)SW * n + N + n) : 0;
}

size_t take(size_t &t, size_t bytes)
{
    size_t o = t;
    t += al(bytes);
    return o;
}

// Batches per launch of the direct pair inverse: bounded by the launch limit and a
// 2 GiB scratch; per-window batches come in whole pairs (multiples of W).  The
// workspace layout and the launcher both size with this function.
long long pair_chunk(int Nz, int W, bool per_window, long long nb)
{
    long long c = (2LL << 30) / ((long long)Nz * (long long)sizeof(cf32));
    if (c < 1) c = 1;
    if (c > 65535) c = 65535;
    if (per_window) {
        c = c / W * W;
        if (c < W) c = W;
        if (c > 65535) c = 65535 / W * W;
    }
    return c < nb ? c : nb;
}

LargeLayout large_layout(int kind, int stack, int S, int W, int N, int Nz)
{
    LargeLayout l;
    size_t t = 0;
    const size_t SW = (size_t)S * W, P = (size_t)S * (S - 1) / 2;
    const bool gram = stack == FPCC_STACK_GRAM || stack == FPCC_STACK_GRAM16;
    switch (kind) {
    case FPCC_PCC2:
        l.Z = take(t, sizeof(cf32) * SW * N);
        l.F = take(t, sizeof(cf32) * SW * Nz);
        l.mx = take(t, sizeof(float) * SW);
        l.B = take(t, bluestein_bytes((int)SW, N));
        break;
    case FPCC_PCC1:
    case FPCC_PCCV:
        l.Z = take(t, sizeof(cf32) * SW * N);
        l.mx = take(t, sizeof(float) * SW);
        l.B = take(t, bluestein_bytes((int)SW, N));
        break;
    case FPCC_GNCC:
        l.F = take(t, sizeof(cf32) * SW * Nz);
        l.E = take(t, sizeof(double) * SW * (size_t)energy_ld(N));
        l.X = take(t, sizeof(float) * SW * 2 * (size_t)edge_len(N));
        break;
    case FPCC_CC1B:
        l.F = take(t, sizeof(cf32) * SW * Nz);
        l.F2 = take(t, sizeof(cf32) * SW * Nz);
        l.X = take(t, sizeof(float) * SW * 2 * (size_t)edge_len(N));
        break;
    }
    if (kind == FPCC_PCC2 || kind == FPCC_GNCC || kind == FPCC_CC1B) {
        if (gram) {
            l.U = take(t, sizeof(cf32) * (size_t)Nz * SW);
            if (kind == FPCC_CC1B) l.U2 = take(t, sizeof(cf32) * (size_t)Nz * SW);
            l.G = take(t, sizeof(cf32) * P * (size_t)gram_ld(Nz));
        } else {
            const bool per_window = stack == FPCC_STACK_NONE;
            const long long nb = per_window ? (long long)P * W : (long long)P;
            l.C = take(t, sizeof(cf32) * (size_t)pair_chunk(Nz, W, per_window, nb) * Nz);
        }
    }
    l.total = t;
    return l;
}
}  // namespace

bool large_supported(int kind, int N, int Nz)
{
    const bool analytic = kind == FPCC_PCC2 || kind == FPCC_PCC1 || kind == FPCC_PCCV;
    if (analytic && !big_plan(N).ok && !big_plan(bluestein_n(N)).ok) return false;
    if (kind == FPCC_PCC1 || kind == FPCC_PCCV) return true;  // time-domain lags: no Nz transforms
    return big_plan(Nz).ok;
}

size_t large_workspace(int kind, int stack, int S, int W, int N, int Nz)
{
    return large_layout(kind, stack, S, W, N, Nz).total;
}

int large_station_regions(int kind, int stack, int S, int W, int N, int Nz, size_t *off, size_t *per)
{
    const LargeLayout l = large_layout(kind, stack, S, W, N, Nz);
    const size_t spec = sizeof(cf32) * (size_t)W * Nz;
    switch (kind) {
    case FPCC_PCC2: off[0] = l.F; per[0] = spec; return 1;
    case FPCC_PCC1:
    case FPCC_PCCV: off[0] = l.Z; per[0] = sizeof(cf32) * (size_t)W * N; return 1;
    case FPCC_GNCC:
        off[0] = l.F;
        per[0] = spec;
        off[1] = l.E;
        per[1] = sizeof(double) * (size_t)W * energy_ld(N);
        off[2] = l.X;
        per[2] = sizeof(float) * (size_t)W * 2 * edge_len(N);
        return 3;
    case FPCC_CC1B:
        off[0] = l.F;
        per[0] = spec;
        off[1] = l.F2;
        per[1] = spec;
        off[2] = l.X;
        per[2] = sizeof(float) * (size_t)W * 2 * edge_len(N);
        return 3;
    }
    return 0;
}

// Bluestein: z = IDFT_N(mask DFT_N(x)) with both N-point DFTs as chirp
// convolutions on the n-point grid (five two-pass stages through the scratch B).
static int analytic_bluestein(const float *x, long long stride, int S, int W, int N, cf32 *Z, float *mx, cf32 *B,
                              hipStream_t st)
{
    const int n = bluestein_n(N);
    const BigPlan &bp = big_plan(n);
    if (!bp.ok) {
        fprintf(stderr, "fastpcc: no analytic-signal plan for N=%d\n", N);
        return -1;
    }
    const int SW = S * W;
    cf32 *chirp = B + (size_t)SW * n, *T = chirp + N;
    int r = bluestein_tables(N, n, chirp, T, st);
    if (r) return r;
    if ((r = fill_zero(mx, SW, st))) return r;
    ColArgs ca;
    ca.op = COL_FWD_CHIRP;
    ca.dst = B;
    ca.bstride = n;
    ca.x = x;
    ca.x_stride = stride;
    ca.W = W;
    ca.N = N;
    ca.chirp = chirp;
    if ((r = run_cols(bp, ca, SW, st, "bluestein_fwd_cols"))) return r;
    RowArgs ra;
    ra.op = ROW_MUL_INV;
    ra.src = B;
    ra.dst = B;
    ra.bstride = n;
    ra.T = T;
    if ((r = run_rows(bp, ra, SW, st, "bluestein_conv_rows"))) return r;
    ca = ColArgs{};
    ca.op = COL_CHIRP_MID;
    ca.src = B;
    ca.dst = B;
    ca.bstride = n;
    ca.N = N;
    ca.da1 = (float)(1 / (double)N);
    if ((r = run_cols(bp, ca, SW, st, "bluestein_mid_cols"))) return r;
    if ((r = run_rows(bp, ra, SW, st, "bluestein_conv_rows"))) return r;
    ca = ColArgs{};
    ca.op = COL_INV_CHIRP_MAX;
    ca.src = B;
    ca.bstride = n;
    ca.dst = Z;
    ca.dst_bstride = N;
    ca.N = N;
    ca.chirp = chirp;
    ca.maxbuf = mx;
    return run_cols(bp, ca, SW, st, "bluestein_inv_cols");
}

// analytic signal of every station-window into Z (natural order), max|z|^2 into mx
static int analytic_big(const float *x, long long stride, int S, int W, int N, cf32 *Z, float *mx, cf32 *B,
                        hipStream_t st)
{
    const BigPlan &bp = big_plan(N);
    if (!bp.ok) return analytic_bluestein(x, stride, S, W, N, Z, mx, B, st);
    const int SW = S * W;
    int r = fill_zero(mx, SW, st);
    if (r) return r;
    ColArgs ca;
    ca.op = COL_FWD_INPUT;
    ca.dst = Z;
    ca.bstride = N;
    ca.x = x;
    ca.x_stride = stride;
    ca.W = W;
    ca.N = N;
    if ((r = run_cols(bp, ca, SW, st, "analytic_fwd_cols"))) return r;
    RowArgs ra;
    ra.op = ROW_MASK_INV;
    ra.src = Z;
    ra.dst = Z;
    ra.bstride = N;
    ra.N = N;
    ra.da1 = (float)(1 / (double)N);
    if ((r = run_rows(bp, ra, SW, st, "analytic_mask_rows"))) return r;
    ca = ColArgs{};
    ca.op = COL_INV_MAX;
    ca.src = Z;
    ca.dst = Z;
    ca.bstride = N;
    ca.maxbuf = mx;
    return run_cols(bp, ca, SW, st, "analytic_inv_cols");
}

int large_station_stage(int kind, const float *x, long long stride, int S, int W, int N, int Nz, int stack,
                        char *ws, hipStream_t st)
{
    LargeLayout l = large_layout(kind, stack, S, W, N, Nz);
    const BigPlan &bz = big_plan(Nz);
    const int SW = S * W;
    int r;
    switch (kind) {
    case FPCC_PCC2: {
        cf32 *Z = (cf32 *)(ws + l.Z), *F = (cf32 *)(ws + l.F);
        float *mx = (float *)(ws + l.mx);
        if ((r = analytic_big(x, stride, S, W, N, Z, mx, (cf32 *)(ws + l.B), st))) return r;
        ColArgs ca;
        ca.op = COL_FWD_AMPNORMF;
        ca.src = Z;
        ca.src_bstride = N;
        ca.nvalid = N;
        ca.maxbuf = mx;
        ca.dst = F;
        ca.bstride = Nz;
        if ((r = run_cols(bz, ca, SW, st, "pcc2_pad_cols"))) return r;
        RowArgs ra;
        ra.op = ROW_FWD_PLAIN;
        ra.src = F;
        ra.dst = F;
        ra.bstride = Nz;
        return run_rows(bz, ra, SW, st, "pcc2_pad_rows");
    }
    case FPCC_PCC1:
    case FPCC_PCCV: {
        cf32 *Z = (cf32 *)(ws + l.Z);
        if ((r = analytic_big(x, stride, S, W, N, Z, (float *)(ws + l.mx), (cf32 *)(ws + l.B), st))) return r;
        return launch_ampnormf_rows(Z, SW, N, st);
    }
    case FPCC_GNCC:
    case FPCC_CC1B: {
        for (int pass = 0; pass < (kind == FPCC_CC1B ? 2 : 1); ++pass) {
            cf32 *F = (cf32 *)(ws + (pass ? l.F2 : l.F));
            ColArgs ca;
            ca.op = COL_FWD_INPUT;
            ca.dst = F;
            ca.bstride = Nz;
            ca.x = x;
            ca.x_stride = stride;
            ca.W = W;
            ca.N = N;
            ca.sign_rule = kind == FPCC_CC1B ? pass + 1 : 0;
            if ((r = run_cols(bz, ca, SW, st, "real_fwd_cols"))) return r;
            RowArgs ra;
            ra.op = ROW_FWD_PLAIN;
            ra.src = F;
            ra.dst = F;
            ra.bstride = Nz;
            if ((r = run_rows(bz, ra, SW, st, "real_fwd_rows"))) return r;
        }
        if ((r = large_edge_copy(x, stride, S, W, N, (float *)(ws + l.X), st))) return r;
        if (kind == FPCC_GNCC)
            return launch_energy_prefix(x, stride, S, W, N, (double *)(ws + l.E), energy_ld(N), st);
        return 0;
    }
    }
    return -1;
}

int large_pcc2_station_batch(int n, const float *const *x, const long long *stride, const int *S, const int *W,
                             const int *N, const int *stack, int Nz, char *const *ws, hipStream_t st)
{
    if (n < 1 || n > kMaxBatchJobs) return -1;
    RealBatch jb{};
    jb.n = n;
    jb.b0[0] = 0;
    int r;
    for (int g = 0; g < n; ++g) {
        // each job's analytic signal on the plan of its own window length
        const LargeLayout l = large_layout(FPCC_PCC2, stack[g], S[g], W[g], N[g], Nz);
        cf32 *Z = (cf32 *)(ws[g] + l.Z);
        float *mx = (float *)(ws[g] + l.mx);
        if ((r = analytic_big(x[g], stride[g], S[g], W[g], N[g], Z, mx, (cf32 *)(ws[g] + l.B), st))) return r;
        jb.Xin[g] = Z;
        jb.N[g] = N[g];
        jb.mx[g] = mx;
        jb.X[g] = ws[g] + l.F;
        jb.b0[g + 1] = jb.b0[g] + S[g] * W[g];
    }
    if (jb.b0[n] > 65535) return -1;
    // AmpNormf and the zero-padded forward transform of every job's station-windows
    const BigPlan &bz = big_plan(Nz);
    ColArgs ca;
    ca.op = COL_FWD_AMPNORMF;
    ca.bstride = Nz;
    if ((r = run_cols_batch(bz, ca, jb, st, "pcc2_pad_cols_batch"))) return r;
    RowArgs ra;
    ra.op = ROW_FWD_PLAIN;
    ra.bstride = Nz;
    return run_rows_batch(bz, ra, jb, st, "pcc2_pad_rows_batch");
}

int large_pair_stage(int kind, int stack, int S, int W, int N, int Nz, int lag0, int L, double v, char *ws, float *y,
                     hipStream_t st)
{
    LargeLayout l = large_layout(kind, stack, S, W, N, Nz);
    const int P = S * (S - 1) / 2;
    int r;
    if (kind == FPCC_PCC1 || kind == FPCC_PCCV) {
        const int mode = stack == FPCC_STACK_NONE ? 0 : 1;
        return launch_pcc_time((const cf32 *)(ws + l.Z), S, W, N, lag0, L, kind == FPCC_PCC1 ? 1.0 : v,
                               kind == FPCC_PCC1, mode, y, L, 0, st);
    }
    const BigPlan &bz = big_plan(Nz);
    // scale and normalisation of each type (pcc2_set :832, ccf/cc_lowlevel 1/Nz, gn(f)_lowlevel)
    const float base_scale = kind == FPCC_PCC2 ? 1 / ((float)Nz * (float)N) : (float)(1.0 / (double)Nz);
    const int norm = kind == FPCC_GNCC ? 1 : kind == FPCC_CC1B ? 2 : 0;
    const cf32 *F = (const cf32 *)(ws + l.F);
    const cf32 *F2 = kind == FPCC_CC1B ? (const cf32 *)(ws + l.F2) : F;
    if (stack == FPCC_STACK_GRAM || stack == FPCC_STACK_GRAM16) {  // PCC2, 1-bit: Gram over all Nz bins
        int ntiles = 0;
        const int2 *tiles = gram_tiles(S, ntiles);
        if (!tiles) return -1;
        cf32 *U = (cf32 *)(ws + l.U), *U2 = kind == FPCC_CC1B ? (cf32 *)(ws + l.U2) : U;
        cf32 *G = (cf32 *)(ws + l.G);
        const bool s16 = stack == FPCC_STACK_GRAM16;
        // |F| <= N: spectra of N unit-bounded phasors or signs
        const float sigma = s16 ? gram16_sigma(N) : 0.f;
        const long long ldg = gram_ld(Nz);
        if ((r = launch_spectra_to_u(F, Nz, S, W, Nz, Nz, 0, U, st, sigma))) return r;
        if (kind == FPCC_CC1B && (r = launch_spectra_to_u(F2, Nz, S, W, Nz, Nz, 0, U2, st, sigma))) return r;
        r = s16 ? launch_gram_split16(U, U2, S, W, 0, W, Nz, tiles, ntiles, sigma, G, gram_rows(ldg), st)
                : launch_gram_f32(U, U2, S, W, 0, W, Nz, tiles, ntiles, G, gram_rows(ldg), st);
        if (r) return r;
        RowArgs ra;
        ra.op = ROW_INV_PLAIN;
        ra.src = G;
        ra.dst = G;
        ra.bstride = ldg;
        for (int p0 = 0; p0 < P; p0 += 65535) {
            const int np = std::min(65535, P - p0);
            ra.src = G + (long long)p0 * ldg;
            ra.dst = G + (long long)p0 * ldg;
            if ((r = run_rows(bz, ra, np, st, "pair_inv_rows"))) return r;
            ColArgs ca;
            ca.op = COL_INV_LAGWIN;
            ca.src = G + (long long)p0 * ldg;
            ca.bstride = ldg;
            ca.y = y;  // pair p0 + b at y[(p0 + b) L]
            ca.lag0 = lag0;
            ca.L = L;
            ca.scale = base_scale / (float)W;
            ca.norm = norm;
            ca.Nsig = N;
            ca.S = S;
            ca.pair0 = p0;
            if (kind == FPCC_CC1B) {
                ca.edge = (const float *)(ws + l.X);
                ca.K = edge_len(N);
                ca.Wall = W;
                ca.esum = 1;
            }
            if ((r = run_cols(bz, ca, np, st, "pair_inv_cols"))) return r;
        }
        return 0;
    }
    // direct cross spectra: per (pair, window), or per pair summed over windows
    const bool per_window = stack == FPCC_STACK_NONE;
    const bool rounds = !per_window && kind == FPCC_GNCC;  // normalise each window, then average
    const bool wsum = !per_window && !rounds;
    cf32 *C = (cf32 *)(ws + l.C);
    const long long nb = per_window ? (long long)P * W : P;
    const long long chunk = pair_chunk(Nz, W, per_window, nb);
    if (W > 65535 && per_window) return -1;
    if (rounds && (r = fill_zero(y, (long long)P * L, st))) return r;
    const int nround = rounds ? W : 1;
    for (int wr = 0; wr < nround; ++wr) {
        for (long long b0 = 0; b0 < nb; b0 += chunk) {
            const int nbt = (int)std::min<long long>(chunk, nb - b0);
            // batches b0 .. b0+nbt-1: whole pairs from pair0 on (b0 is a multiple of W)
            RowArgs ra;
            ra.op = ROW_INV_CROSS;
            ra.dst = C;
            ra.bstride = Nz;
            ra.F = F;
            ra.F2 = F2;
            ra.Sst = S;
            ra.Wfull = W;
            ra.Wb = per_window ? W : 1;
            ra.wsel = rounds ? wr : 0;
            ra.wsum = wsum;
            ColArgs ca;
            ca.op = COL_INV_LAGWIN;
            ca.src = C;
            ca.bstride = Nz;
            ca.lag0 = lag0;
            ca.L = L;
            ca.S = S;
            ca.Wb = per_window ? W : 1;
            ca.wsel = rounds ? wr : 0;
            ca.Wfull = per_window ? W : 1;
            ca.accumulate = rounds;
            ca.scale = per_window ? base_scale : base_scale / (float)W;
            ca.norm = norm;
            ca.Nsig = N;
            ca.E = (const double *)(ws + l.E);
            ca.lde = energy_ld(N);
            if (rounds) ca.Wfull = W;  // energies are indexed by (station, window)
            if (kind != FPCC_PCC2) {
                ca.edge = (const float *)(ws + l.X);
                ca.K = edge_len(N);
                ca.Wall = W;
                ca.esum = wsum;
            }
            const int pair0 = (int)(per_window ? b0 / W : b0);
            ra.pair0 = pair0;
            ca.pair0 = pair0;
            ca.y = y;
            if ((r = run_rows(bz, ra, nbt, st, "pair_cross_rows"))) return r;
            if ((r = run_cols(bz, ca, nbt, st, "pair_inv_cols"))) return r;
        }
    }
    return 0;
}

}  // namespace fpcc
