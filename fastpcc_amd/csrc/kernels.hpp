// Launchers of the device stages (host-callable, stream-ordered, no allocation).
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/fastpcc.h"
#include "cplx.hpp"

namespace fpcc {

// PCC2 station stage: F[s*W+w][Nz] = FFT_Nz(pad(AmpNormf(analytic_N(x_sw))))
int launch_station_pcc(const float *d_x, long long station_stride, int S, int W, int N, int Nz, cf32 *d_F,
                       hipStream_t st);
// AmpNormf in place on `rows` rows of N complex values.
int launch_ampnormf_rows(cf32 *d_z, int rows, int N, hipStream_t st);
// Phasors only (analytic signal + AmpNormf), phi[s*W+w][N].
int launch_station_phasor(const float *d_x, long long station_stride, int S, int W, int N, cf32 *d_phi,
                          hipStream_t st);
// Time-domain PCC1 (pcc1 = 1) or PCC-v over lags lag_first .. lag_first+L-1;
// output element l of row r at y[r*out_ld + out_off + l] (rows: pair*W+w or pair).
int launch_pcc_time(const cf32 *d_phi, int S, int W, int N, int lag_first, int L, double v, int pcc1, int mode,
                    float *d_y, int out_ld, int out_off, hipStream_t st);
// PCC2 pair stage over pairs [pair_base, pair_base+npairs) of S stations.
// mode 0: y[p][w][L]; 1: mean over windows of per-window IFFTs; 2: IFFT of the
// window sum (same result by linearity).
int launch_pair_pcc2(const cf32 *d_F, int S, int W, int N, int Nz, int lag0, int L, int mode, int pair_base,
                     int npairs, float *d_y, hipStream_t st);

// Phase-weighted stack over windows (pws.hip) for Nz <= 8192 (pws_supported).
// src 0: PCC2 phasor spectra F[s*W+w][ld]; 1: FP32 half spectra (1-bit: Xa the
// x >= 0 signs, Xb the x > 0 signs); 2: FP64 half spectra (GNCC, Xa = Xb).
// multiply 0: y[p][l] = mean over windows of Re(a_w) * cscale, times the
// coherence; 1: y[p][l] (a linear stack written before) times the coherence.
bool pws_supported(int Nz);
int launch_pair_pws(int src, const void *d_Xa, const void *d_Xb, long long ld, int S, int W, int Nz, int lag0, int L,
                    float cscale, float power, int multiply, float *d_y, hipStream_t st);

// The same for Nz = 8192 and all pairs (gram.hip: persistent workgroups, the fold
// in registers, the radix-8^4 inverse of the Gram path's pair stage).
int launch_pair_pcc2_4096(const cf32 *d_F, int S, int W, int N, int lag0, int L, int mode, float *d_y,
                          hipStream_t st);

// Stack over windows as a per-frequency Gram (gram.hip).
long long gram_ld(int nk);                       // row length of G for nk bins
// Layout of G: bin k of pair p at (k >> sh) * gs + p * pl + (k & (2^sh - 1)).
// Rows (pair-major, ld bins): {0, ld, 30}.  Grouped: bins in groups of 128, each
// group holding 128 bins of every pair ({(P | 1) * 128, 128, 7}; an odd number
// of pair slots a group), so a Gram chunk of bins writes pairs 1 KB apart and a
// pair's spectrum is read in 1 KB pieces.
struct GLayout {
    long long gs, pl;
    int sh;
};
inline GLayout gram_rows(long long ld) { return GLayout{0, ld, 30}; }
GLayout gram_grouped(long long P);
size_t gram_grouped_elems(long long P, int nk);  // complex values of a grouped G
const int2 *gram_tiles(int S, int &ntiles);      // cached device tile table
// U[k][s][m] for k < nk from rows F[s*W+w] of ldF bins; mirror 1 appends the
// conjugated bin Nz-k as m = W+w.  sigma = 0: FP32 complex elements; sigma > 0:
// 8-byte split elements {fp16 hi re|im, fp16 lo re|im} of sigma * value for
// launch_gram_split16.
int launch_spectra_to_u(const cf32 *d_F, long long ldF, int S, int W, int Nz, int nk, int mirror, void *d_U,
                        hipStream_t st, float sigma = 0.f);
// G[p][k] = sum_{m0 <= m < m0+K} conj(Ua[k][i][m]) Ub[k][j][m] for pairs i < j, k < nk
// (U rows hold W2 values per station); bins nk .. of the last chunk of 8 are zeroed.
int launch_gram_f32(const cf32 *d_Ua, const cf32 *d_Ub, int S, int W2, int m0, int K, int nk, const int2 *d_tiles,
                    int ntiles, cf32 *d_G, GLayout gl, hipStream_t st);
// The same Gram on FP16 MFMA from split elements scaled by sigma (results are
// divided by sigma^2).
// Split U rows of W2 elements are stored in groups of four (4 hi words, then 4
// lo words) when this holds, as launch_spectra_to_u writes them and
// launch_gram_split16 reads them (which then takes m0 = 0, K = W2).
bool gram16_grouped(int W2);
int launch_gram_split16(const void *d_Ua, const void *d_Ub, int S, int W2, int m0, int K, int nk,
                        const int2 *d_tiles, int ntiles, float sigma, cf32 *d_G, GLayout gl, hipStream_t st);
// Power-of-two split scale for |component| <= bound: largest scaled value <= 2^14,
// inside FP16's range with lo parts normal down to 2^-8 of it.
float gram16_sigma(int bound);
int fill_zero(float *p, long long n, hipStream_t st);
// y[p][l] = scale * c2r(G[p])[(lag0+l) mod Nz], divided by N-|lag| if count_norm
// d_edge (1-bit): window edge samples of edge_len(N) (see edge.hpp) for the exact
// short-overlap lags; S stations
int launch_pair_c2r(const cf32 *d_G, GLayout gl, int P, int Nz, int N, int W, int lag0, int L, float scale,
                    int count_norm, float *d_y, hipStream_t st, const float *d_edge = nullptr, int S = 0);

// Station-windows of several jobs in one launch (uneven station counts, window
// counts and lengths; one transform length): job j owns launch blocks
// [b0[j], b0[j+1]), one per station-window, reading x[j] rows of stride[j] and
// writing the rows of X[j] (and E[j] energy prefix sums of lde[j], if not null).
constexpr int kMaxBatchJobs = 16;
struct RealBatch {
    int n;
    int b0[kMaxBatchJobs + 1];
    const float *x[kMaxBatchJobs];
    long long stride[kMaxBatchJobs];
    int W[kMaxBatchJobs], N[kMaxBatchJobs];
    void *X[kMaxBatchJobs];
    double *E[kMaxBatchJobs];
    long long lde[kMaxBatchJobs];
    // per-(station-window, scale) passes of WPCC: X holds job j's per-scale rows,
    // Xin its input spectra (ROW_INV_MUL), mx its per-row maxima (COL_INV_MAX,
    // COL_FWD_AMPNORM)
    const void *Xin[kMaxBatchJobs];
    float *mx[kMaxBatchJobs];
};

// Job table of a PCC2 station-stage launch over several jobs of one Nz <= 8192
// (pcc2.hip): per job its samples, stride, windows, window length, that
// length's FFT plan and twiddles, 1/N and the spectra rows.
struct PccBatch {
    int n;
    int b0[kMaxBatchJobs + 1];
    const float *x[kMaxBatchJobs];
    long long stride[kMaxBatchJobs];
    int W[kMaxBatchJobs], N[kMaxBatchJobs];
    FftPlan planN[kMaxBatchJobs];
    const cf32 *twN[kMaxBatchJobs];
    float da1[kMaxBatchJobs];
    cf32 *F[kMaxBatchJobs];
};
// Job table of a per-window PCC2 pair-stage launch over several jobs at Nz 8192
// (gram.hip): job j's pairs are u0[j] <= u < u0[j+1] of the launch's sequence.
struct PlainBatch {
    int n;
    int u0[kMaxBatchJobs + 1];
    const cf32 *F[kMaxBatchJobs];
    const int *order[kMaxBatchJobs];
    int S[kMaxBatchJobs], W[kMaxBatchJobs], lag0[kMaxBatchJobs], L[kMaxBatchJobs];
    float scale[kMaxBatchJobs];
    float *y[kMaxBatchJobs];
};
// The pair stages of n PCC2 jobs at Nz 8192 with one stack mode (0 none, 1 plain,
// 2 zsum): spectra d_F[j] of S[j] stations x W[j] windows, window length N[j],
// lags lag0[j] .. lag0[j] + L[j] - 1, output d_y[j]; the same results as
// launch_pair_pcc2 per job.
int launch_pair_pcc2_4096_batch(int n, const cf32 *const *d_F, const int *S, const int *W, const int *N,
                                const int *lag0, const int *L, int mode, float *const *d_y, hipStream_t st);
// True when (N, Nz) runs the compile-time specialised station kernel (those
// jobs are not batched: the batch kernel is the generic one).
bool station_pcc_specialised(int N, int Nz);
// Fills job j's plan, twiddles, length, 1/N and output rows (-1: unsupported).
int station_pcc_batch_job(int N, int Nz, cf32 *d_F, PccBatch &jb, int j);
int launch_station_pcc_batch(const PccBatch &jb, int Nz, hipStream_t st);

// GNCC / 1-bit (gn.hip).  Half spectra rows hold real_ld(Nz) bins, energy rows
// energy_ld(N) prefix sums.  sign_rule 0: samples, 1: x >= 0, 2: x > 0.
long long real_ld(int Nz);
long long energy_ld(int N);
int launch_station_real(const float *d_x, long long station_stride, int S, int W, int N, int Nz, int sign_rule,
                        bool f64, void *d_X, double *d_E, hipStream_t st);
int launch_station_real_batch(const RealBatch &jb, int Nz, int sign_rule, bool f64, hipStream_t st);
// norm 0: sqrt(E1 E2) from prefix sums (GNCC); 1: overlap count (1-bit).
// mode 0: y[p][w][L]; 1: mean over windows.
int launch_pair_gn(const void *d_Xa, const void *d_Xb, const double *d_E, bool f64, int S, int W, int N, int Nz,
                   int lag0, int L, int norm, int mode, float *d_y, hipStream_t st, const float *d_edge = nullptr);
// Job table of a GNCC / 1-bit pair-stage launch over several jobs of one Nz, norm
// and mode: per job the spectra (Xa, Xb), energies, counts, lag window, output and
// window edges; b0, lde and K are filled by launch_pair_gn_batch.
struct GnBatch {
    int n;
    int b0[kMaxBatchJobs + 1];
    const void *Xa[kMaxBatchJobs];
    const void *Xb[kMaxBatchJobs];
    const double *E[kMaxBatchJobs];
    long long lde[kMaxBatchJobs];
    int S[kMaxBatchJobs], W[kMaxBatchJobs], N[kMaxBatchJobs], lag0[kMaxBatchJobs], L[kMaxBatchJobs],
        K[kMaxBatchJobs];
    float *y[kMaxBatchJobs];
    const float *edge[kMaxBatchJobs];
};
// The pair stages of the jobs of jb in one launch, the same results as
// launch_pair_gn per job.
int launch_pair_gn_batch(GnBatch jb, bool f64, int Nz, int norm, int mode, hipStream_t st);

// Windows longer than one workgroup's LDS (large.hip).
bool large_supported(int kind, int N, int Nz);
size_t large_workspace(int kind, int stack, int S, int W, int N, int Nz);
int large_station_stage(int kind, const float *x, long long stride, int S, int W, int N, int Nz, int stack,
                        char *ws, hipStream_t st);
// The long-window PCC2 station stage of n jobs of one Nz (window lengths, station
// and window counts and strides may differ): each job's analytic signal on its own
// plan, then AmpNormf and the zero-padded forward transform of all their
// station-windows as shared launches.
int large_pcc2_station_batch(int n, const float *const *x, const long long *stride, const int *S, const int *W,
                             const int *N, const int *stack, int Nz, char *const *ws, hipStream_t st);
int large_pair_stage(int kind, int stack, int S, int W, int N, int Nz, int lag0, int L, double v, char *ws, float *y,
                     hipStream_t st);
// Workspace regions holding the station stage's per-station output (what the pair
// stage reads): station s at off[k] + s*per[k]; returns the region count.
int large_station_regions(int kind, int stack, int S, int W, int N, int Nz, size_t *off, size_t *per);
// GNCC / 1-bit on two-pass transforms: the first and last edge samples of every
// window into the region large_station_regions lists third.
int large_edge_copy(const float *x, long long stride, int S, int W, int N, float *edge, hipStream_t st);
// edge samples kept per window for the exact short-overlap lags (edge.hpp)
int edge_len(int N);

// WPCC (wpcc.hip): workspace bytes and the two stages for S stations x W windows.
struct WFamily;
size_t wpcc_workspace(int S, int W, int N, int Nz, int nscales);
int wpcc_station_stage(const float *d_x, long long station_stride, int S, int W, int N, const WFamily &f,
                       char *ws, hipStream_t st);
int wpcc_station_regions(int S, int W, int Nz, int nscales, size_t *off, size_t *per);
// Station stages of n WPCC jobs with one filter bank (wpcc_batchable) in shared
// launches: job j has S[j] stations of W[j] windows of N[j] samples at x[j]
// (rows stride[j] apart) and its own workspace ws[j]; each job's rows equal
// wpcc_station_stage's.
int wpcc_station_stage_batch(int n, const float *const *x, const long long *stride, const int *S, const int *W,
                             const int *N, const WFamily &f, char *const *ws, hipStream_t st);
// Whether two families share the filter bank (type, Nz, scales, voices, s0, w0).
bool wpcc_same_bank(const WFamily &a, const WFamily &b);
// per_window: y[p][w][L] (one Gram per window), else mean over windows y[p][L].
int wpcc_pair_stage(int S, int W, int N, const WFamily &f, int lag0, int L, bool per_window, char *ws, float *d_y,
                    hipStream_t st);

// Station-output regions of a job's workspace (any kind and path), as
// large_station_regions; 0 if the job is invalid.  Stations are independent: a
// station's rows are the same in any job with the same kind, windows, N and Nz
// (and, for WPCC, filter parameters).
int job_station_regions(const fpcc_job *job, size_t *off, size_t *per);
// Whether the job's station stage can share launches with other jobs of the same
// key (fpcc_station_stage_batch_device).
bool job_batch_key(const fpcc_job *job, long long *key);
// Whether the job's pair stage can share a launch with other jobs of the same mode
// (fpcc_pair_stage_batch_device: PCC2 windows in LDS at Nz 8192, stacks none,
// plain and zsum); mode 0, 1 or 2.
bool job_pair_batch_mode(const fpcc_job *job, int *mode);

}  // namespace fpcc
