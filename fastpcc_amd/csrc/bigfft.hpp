// FFTs longer than one workgroup's LDS: two passes through HBM (four-step).
//
// n = N1*N2 (7-smooth).  Element (r, c) of a length-n sequence sits at r*N2 + c.
// Forward: column pass (FFT_N1 over r for each column c, then W_n^{c k1}), row
// pass (FFT_N2 over c) -> bin k1 + N1*k2 stored at k1*N2 + k2 ("permuted" order).
// Inverse: row pass (inverse FFT_N2, then W_n^{-n2 k1}) on permuted data, then
// column pass (inverse FFT_N1) -> natural order.  Both passes take an op code
// selecting a fused prologue/epilogue.
#pragma once
#include "engine.hpp"

namespace fpcc {

constexpr int kBigCols = 16;  // columns per workgroup (128-byte row segments)
constexpr int kBigRows = 4;   // rows per workgroup

struct BigPlan {
    int n = 0, n1 = 0, n2 = 0;
    FftPlan p1{}, p2{};
    const cf32 *tw1 = nullptr, *tw2 = nullptr, *twN = nullptr;  // twN: get_split_twiddles(n)
    bool ok = false;
};

// Cached per device; ok = false if n has no admissible split.
const BigPlan &big_plan(int n);

enum ColOp {
    COL_FWD_INPUT = 0,       // x[g - off] * in_scale (or its sign), zero outside [0, N)
    COL_FWD_PLAIN = 1,       // src
    COL_FWD_AMPNORM = 2,     // src / sqrt(|src|^2 + 1e-6 max)            (AmpNorm, FP64 formula)
    COL_FWD_AMPNORMF = 3,    // src / (|src| + 1e-6 sqrt(max)), 0 past nvalid (AmpNormf + zero pad)
    COL_INV_MAX = 4,         // inverse, natural out, per-batch max|z|^2 into maxbuf
    COL_INV_LAGWIN = 5,      // inverse, Re of the lag window into y (normalised)
    // Bluestein analytic signal of N samples on an n-point grid (n >= 2N-1), chirp
    // w_m = exp(-i pi m^2 / N):
    COL_FWD_CHIRP = 6,       // x[g] * w_g (g < N), forward
    COL_CHIRP_MID = 7,       // inverse -> mask_g * conj(conv_g) (g < N) -> forward
    COL_INV_CHIRP_MAX = 8,   // inverse -> z_g = conj(w_g conv_g), g < N, into dst rows of dst_bstride; max|z|^2
};

enum RowOp {
    ROW_FWD_PLAIN = 0,
    ROW_FWD_SCALE = 1,  // * wscale[b % S]
    ROW_FWD_MASK = 2,   // * analytic mask of an N-point DFT (AnalyticSignal_plan_float)
    ROW_INV_MUL = 3,    // inverse of X[b / S] * fw[b % S]
    ROW_INV_PLAIN = 4,
    ROW_INV_CROSS = 5,  // inverse of conj(F[i,w]) F2[j,w] for the batch's pair and window (or summed over w)
    ROW_MASK_INV = 6,   // forward row FFT, analytic mask, inverse row FFT (both row passes of the analytic signal)
    ROW_MUL_INV = 7,    // forward row FFT, * T (same permuted position), inverse row FFT (Bluestein convolution)
};

struct ColArgs {
    int op = 0;
    const cf32 *src = nullptr;
    cf32 *dst = nullptr;
    long long bstride = 0;      // dst (and src) elements between batches
    long long src_bstride = 0;  // src stride if different (0: bstride)
    // COL_FWD_INPUT: batch = station-window sw (station sw / W, window sw % W)
    const float *x = nullptr;
    long long x_stride = 0;
    int W = 1, N = 0, off = 0, sign_rule = 0;  // sign_rule 1: x >= 0, 2: x > 0 -> +-1
    float in_scale = 1.f;
    // COL_INV_MAX writes, COL_FWD_AMPNORM(F) read max|z|^2 per batch
    float *maxbuf = nullptr;
    int nvalid = 0;  // COL_FWD_AMPNORMF: valid samples per src row
    // Bluestein ops: chirp table (N values), mask scale, output row stride
    const cf32 *chirp = nullptr;
    float da1 = 0.f;
    long long dst_bstride = 0;
    // COL_INV_LAGWIN: batch b -> pair p = b / Wb, window w = wsel + b % Wb
    float *y = nullptr;
    int lag0 = 0, L = 0;
    float scale = 1.f;
    int S = 2, Wb = 1, wsel = 0, Wfull = 1, pair0 = 0;  // pair p = pair0 + b / Wb
    int accumulate = 0;  // 0: y[(p*Wfull + w)*L + l] = v ; 1: y[p*L + l] += v
    int norm = 0;        // 0: none, 1: 1/sqrt(E1 E2) from prefix sums E, 2: 1/(Nsig - |lag|)
    const double *E = nullptr;
    long long lde = 0;
    int Nsig = 0;
    // lags whose overlap m = Nsig - |lag| is below K: the correlation as an exact
    // FP64 sum over the overlap from the first and last K samples of each window
    // (edge[(station * Wall + window) * 2K + ...], head then tail); esum: the
    // output sums all Wall windows (else window w); norm 2 takes the samples'
    // signs (x1 >= 0, x2 > 0)
    const float *edge = nullptr;
    int K = 0, Wall = 1, esum = 0;
};

struct RowArgs {
    int op = 0;
    const cf32 *src = nullptr;
    cf32 *dst = nullptr;
    long long bstride = 0;
    const cf32 *X = nullptr, *fw = nullptr;  // ROW_INV_MUL
    int S = 1;
    const float *wscale = nullptr;  // ROW_FWD_SCALE
    int N = 0;                      // ROW_FWD_MASK
    float da1 = 0.f;
    const cf32 *T = nullptr;        // ROW_MUL_INV: n values in the two-pass (permuted) order
    // ROW_INV_CROSS: spectra F[s*Wfull + w][n] (station i) and F2 (station j, F if null);
    // batch b -> pair p = b / Wb, window wsel + b % Wb; wsum != 0 sums over all Wfull windows
    const cf32 *F = nullptr, *F2 = nullptr;
    int Sst = 2, Wb = 1, wsel = 0, Wfull = 1, wsum = 0, pair0 = 0;  // pair p = pair0 + b / Wb
};

// Bluestein grid for an N-point DFT: the smallest power of two >= max(64, 2N-1).
int bluestein_n(int N);
// Tables of the N-point DFT on the n-point grid, computed on the stream into
// caller memory: chirp w_m = exp(-i pi m^2/N) (m < N, FP64 phases) and
// T = FFT_n(b)/n in the two-pass (permuted) order, b = conj(w) wrapped circularly.
int bluestein_tables(int N, int n, cf32 *chirp, cf32 *T, hipStream_t st);

struct RealBatch;
int run_cols(const BigPlan &bp, const ColArgs &a, int nbatch, hipStream_t st, const char *name);
// COL_FWD_INPUT / ROW_FWD_PLAIN over the station-windows of several jobs (x, stride,
// W, N and destination rows from the job table; the other fields from a).
int run_cols_batch(const BigPlan &bp, const ColArgs &a, const RealBatch &jb, hipStream_t st, const char *name);
int run_rows_batch(const BigPlan &bp, const RowArgs &a, const RealBatch &jb, hipStream_t st, const char *name);
int launch_energy_prefix_batch(const RealBatch &jb, hipStream_t st);
int run_rows(const BigPlan &bp, const RowArgs &a, int nbatch, hipStream_t st, const char *name);
int fill_zero(float *p, long long n, hipStream_t st);
// Exclusive prefix sums of x^2 in FP64, P[sw][0..N] (rows of lde), for the GNCC energies.
int launch_energy_prefix(const float *d_x, long long station_stride, int S, int W, int N, double *d_E, long long lde,
                         hipStream_t st);

}  // namespace fpcc
