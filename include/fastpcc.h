/*
 * fastpcc — MI355X (gfx950) interstation correlation library, C ABI.
 *
 * Drop-in for the correlation entry points of sergiventosa/FastPCC: the
 * functions below keep the reference's names, argument meanings, ownership and
 * return codes, so that PCC_fullpair_1b.c (or any C host) links against
 * libfastpcc.so instead of FFTapps.o/ccs_cuda.o.  All arithmetic runs in
 * hand-written HIP kernels on the current HIP device; there is no CPU fallback:
 * a GPU failure is returned as an error code.
 *
 * Return codes (as the reference):
 *    0  success
 *   -1  invalid argument (NULL pointer, negative length) or a GPU failure
 *   -2  unsupported option (cuda_tspcc2_set_freq with Down_smp != 1)
 *   -3  |Lag1| or |Lag2| >= N, or an unsupported wavelet type
 * Diagnostics are printed to stderr.
 *
 * Layout: y, x1, x2 are arrays of Tr row pointers owned by the caller; x rows
 * hold N samples, y rows hold L = |Lag2-Lag1|+1 lags (abs(Lag2-Lag1+1) for
 * tspcc2_set, as the reference).  y[tr][l] is the correlation at lag
 * min(Lag1,Lag2)+l, positive lags meaning x2 later than x1.
 */
#ifndef FASTPCC_H
#define FASTPCC_H

#include <stddef.h>
#include <semaphore.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Reference entry points (src/FFTapps.h:6-16).  Host buffers in and out.
 * ------------------------------------------------------------------------- */

/* FP64 analytic signal, replaces AnalyticSignal, src/FFTapps.c:45-80:
 * y[n] = sum_k mask_k X_k e^{2 pi i k n/N}, X = DFT_N(x), mask = 1/N at DC, 2/N for
 * 0 < k < (N+1)/2, 1/N at N/2 for even N, 0 above.  y and x hold N values (host
 * memory); -1 if x or y is NULL or on a GPU failure.  Only dead branches of the
 * reference call it (src/FFTapps.c:432-473, 743-779). */
int AnalyticSignal(double _Complex *y, double *x, unsigned int N);

/* PCC of power 2 (Ventosa et al. 2019), replaces pcc2_set, src/FFTapps.c:785-881.
 * FP32 like the reference's fftwf path. */
int pcc2_set(float **const y, float **const x1, float **const x2, const unsigned int N,
             const unsigned int Tr, const int Lag1, const int Lag2);
/* Geometrically normalised correlation, replaces ccgn_set, src/FFTapps.c:883-978.
 * FP64 FFTs and overlap energies like the reference's fftw path.  Above a zero-pad
 * length of 8192 the transforms are FP32, and the lags whose overlap is under N/64
 * samples are summed exactly in FP64 from the windows' edge samples instead (the
 * FP32 transform's error divided by a short overlap's small energy would dominate
 * there). */
int ccgn_set(float **const y, float **const x1, float **const x2, const unsigned int N,
             const unsigned int Tr, const int Lag1, const int Lag2);
/* 1-bit GNCC, replaces cc1b_set, src/FFTapps.c:980-1073.  Signs x1 >= 0 -> +1 and
 * x2 > 0 -> +1 (the reference's asymmetric rule, :1037 and :1042); FP32 FFTs.
 * The lags whose overlap m is under max(64, N/64) samples (at most N) are exact
 * integer sums of sign products from the windows' edge samples, divided by m: there
 * the result is the exact value, which the reference's FP32 transform misses by up
 * to about 1.5e-5 near |lag| = N-1 (the transform's absolute error over a small m). */
int cc1b_set(float **const y, float **const x1, float **const x2, const unsigned int N,
             const unsigned int Tr, const int Lag1, const int Lag2);

/* PCC of power 1, replaces pcc1_set, src/FFTapps.c:611-783 (CPU branch):
 * analytic signal + AmpNormf in FP32, then the time-domain sum of
 * |phi2[n+lag]+phi1[n]| - |phi2[n+lag]-phi1[n]| over the overlap / (2 (N-|lag|)).
 * Lag handling as the reference: returns 0 without output if Lag1 > N or
 * Lag2 < -N, -1 if Lag2 < Lag1, rows zeroed outside |lag| <= N. */
int pcc1_set(float **const y, float **const x1, float **const x2, const int N, const unsigned int Tr,
             const int Lag1, const int Lag2);
/* PCC of power v, replaces pcc_set, src/FFTapps.c:475-608 (CPU branch): terms
 * |.|^v, divided by 2^(v/2) (N-|lag|).  Lag handling as pcc1_set. */
int pcc_set(float **const y, float **const x1, float **const x2, const int N, const unsigned int Tr,
            const double v, const int Lag1, const int Lag2);

/* Wavelet phase cross-correlation, replaces tspcc2_set, src/FFTapps.c:1076-1255.
 * pmin/pmax are periods in samples; type -1 Morlet, -2 complete Morlet, -3
 * complex Mexican hat; op1 is the Morlet w0 (0 = pi sqrt(2/ln 2)).  The filter
 * bank is built in FP64 as CreateWaveletFamily (src/FWTa/wavelet_def_v7.c:204-295);
 * the FFTs of Nz points run in FP32 on the GPU.  Rows hold abs(Lag2-Lag1+1) lags
 * starting at min(Lag1, Lag2) and are zeroed first; -3 for an unknown type. */
int tspcc2_set(float **const y, float **const x1, float **const x2, const unsigned int N, const unsigned int Tr,
               const int Lag1, const int Lag2, double pmin, double pmax, unsigned int V, int type, double op1);

/* ---------------------------------------------------------------------------
 * Reference GPU-engine entry points (src/ccs_cuda.h:8-19).  x1/x2 of the
 * *_highlevel pair are the un-normalised analytic signals (Tr rows of N complex
 * values) that the reference's CPU producer computes; when anok is not NULL,
 * anok[tr] is waited on (sem_wait) before row tr is read, in trace order, as the
 * reference's consumer does.  The phasor regulariser and the PCC-v divisor follow
 * the reference's CPU path (AmpNormf with 1e-6 max|z|; 2^(v/2) (N-|lag|)), not
 * its CUDA kernels, so every entry gives the same numbers as pcc1_set/pcc_set.
 * Only lags l in [l1, L') are written (rows are not zeroed), as the reference.
 * ------------------------------------------------------------------------- */
int pcc1_highlevel(float **y, _Complex float **x1, _Complex float **x2, int N, unsigned int Tr, int Lag1, int Lag2,
                   sem_t *anok);
int pcc_highlevel(float **y, _Complex float **x1, _Complex float **x2, int N, unsigned int Tr, float v, int Lag1,
                  int Lag2, sem_t *anok);
/* Real inputs, PCC1 end to end on the GPU; never called by the reference
 * (src/FFTapps.c:745 sits in a dead branch).  Errors are printed. */
void pcc1_highlevel2(float **y, float **x1, float **x2, int N, unsigned int Tr, int Lag1, int Lag2);
/* WPCC with a caller-built filter bank (src/ccs_cuda.cu:898-1036): Nz (power of
 * two), S wavelets wc[s] of Ls[s] FP64 complex samples with origin center[s],
 * scales, and C = 1/(K0 N).  -2 if any Down_smp[s] != 1.  Writes |Lag2-Lag1|+1
 * lags per window from min(Lag1, Lag2).  The phasor regulariser is the CPU one
 * (z / sqrt(|z|^2 + 1e-6 max|z|^2)), not the reference CUDA threshold. */
int cuda_tspcc2_set_freq(float **y0, float **x01, float **x02, unsigned int N, unsigned int Tr, unsigned int Nz,
                         const int Lag1, const int Lag2, unsigned int S, double *scale, _Complex double **wc,
                         unsigned int *Ls, int *center, unsigned int *Down_smp, float C);

/* ---------------------------------------------------------------------------
 * Batched all-pairs entry (no counterpart in the reference, which correlates one
 * pair per process run; src/PCC_fullpair_1b.c:467-499).
 *
 * Input x: n_stations rows, row s starting at x + s*station_stride, each holding
 * n_windows consecutive windows of win_len samples (window w = samples
 * [w*win_len, (w+1)*win_len)).  Pairs p <-> (i<j) in row-major upper-triangle
 * order, p = i*S - i*(i+1)/2 + (j-i-1); station i is x1 (virtual source), j is x2.
 * Output y: stack == FPCC_STACK_NONE -> [P][n_windows][L], else [P][L] holding
 * the arithmetic mean over windows.  Stack modes per kind: PCC2 all; GNCC NONE,
 * PLAIN (its per-window energies forbid stacking spectra); 1-bit NONE, PLAIN, GRAM,
 * GRAM16;
 * PCC1 and PCC-v NONE, PLAIN (time-domain sums, N <= 8192); WPCC NONE, PLAIN
 * (= GRAM: the sum over scales and windows is one per-frequency Gram).
 *
 * FPCC_STACK_PWS (PCC2, GNCC and 1-bit, zero-pad length Nz <= 8192): the
 * phase-weighted stack (Schimmel & Paulssen 1997) of the per-window correlations
 * c_w, the stacking the reference leaves to the external ts-PWS program
 * (src/PCC_fullpair_1b.c:568-569):
 *     y[p][l] = (1/W sum_w c_w(l)) * |1/W sum_w u_w(l)|^pws_power,
 * u_w(l) = a_w(l)/|a_w(l)| (0 where a_w(l) = 0), a_w(l) the analytic signal of
 * window w's whole circular correlation (all Nz lags, before the lag window is
 * cut; as AnalyticSignal computes it, src/FFTapps.c:45-80) at lag l.  pws_power 0
 * gives the linear stack; 2 is the usual choice.
 * ------------------------------------------------------------------------- */
enum { FPCC_PCC2 = 0, FPCC_PCC1 = 1, FPCC_PCCV = 2, FPCC_GNCC = 3, FPCC_CC1B = 4, FPCC_WPCC = 5 };
enum {
    FPCC_STACK_NONE = 0,  /* per-window outputs                                       */
    FPCC_STACK_PLAIN = 1, /* mean of per-window correlations (one IFFT per pair-window) */
    FPCC_STACK_ZSUM = 2,  /* sum of folded cross spectra over windows, one IFFT per pair */
    FPCC_STACK_GRAM = 3,  /* per-frequency Hermitian rank-W Gram (FP32 MFMA), one IFFT per pair */
    FPCC_STACK_GRAM16 = 4, /* the same Gram on FP16 MFMA with split hi+lo operands           */
    FPCC_STACK_PWS = 5     /* phase-weighted stack of the per-window correlations (below)    */
};

typedef struct {
    int kind;                        /* FPCC_PCC2 ...                                  */
    int stack;                       /* FPCC_STACK_*                                   */
    unsigned int n_stations;         /* S                                              */
    unsigned int n_windows;          /* W                                              */
    unsigned int win_len;            /* N, samples per window                          */
    unsigned long long station_stride; /* elements between station rows (>= W*N)      */
    int lag1, lag2;                  /* lag window, lag1 <= lag2, |lag| < N            */
    double v;                        /* PCC-v power (kind FPCC_PCCV)                   */
    double pmin, pmax;               /* WPCC periods in samples                        */
    unsigned int V;                  /* WPCC voices per octave                         */
    int wtype;                       /* WPCC wavelet: -1 Morlet, -2 complete Morlet, -3 MexHat */
    double op1;                      /* WPCC Morlet w0 (0 = pi*sqrt(2/ln 2))           */
    double pws_power;                /* FPCC_STACK_PWS: power of the phase coherence (>= 0) */
} fpcc_job;

/* Number of output floats for a job (0 if the job is invalid). */
size_t fpcc_output_len(const fpcc_job *job);
/* Device workspace bytes fpcc_allpairs_device needs for a job. */
size_t fpcc_workspace_size(const fpcc_job *job);
/* Device buffers, stream-ordered on `stream` (a hipStream_t, NULL = default
 * stream); no allocation, no synchronisation.  Returns 0 or an error code. */
int fpcc_allpairs_device(const fpcc_job *job, const float *d_x, float *d_y, void *d_ws, size_t ws_bytes,
                         void *stream);
/* fpcc_allpairs_device without a workspace argument and without allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a fixed device array of
 * FPCC_NOALLOC_SCRATCH_BYTES inside the library, one per device (held from the
 * first call on a device until the process ends); the call makes no hipMalloc or
 * hipFree.  Every call on one device shares that array: calls that overlap on a
 * device (other streams, other threads) race and give wrong results.  The job's
 * constant tables (FFT plans and twiddles, Gram tile lists, filter banks) must
 * exist on the device: one fpcc_allpairs_device call of a job with the same
 * kind, stations, window length, lags, stack and filter parameters builds them.
 * Returns 0; -1 / -3 as fpcc_allpairs_device; -5 if the tables are not built
 * yet; -6 if the job's workspace (fpcc_workspace_size) exceeds the scratch. */
#define FPCC_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int fpcc_allpairs_device_noalloc(const fpcc_job *job, const float *d_x, float *d_y, void *stream);
/* Host buffers: copies x in, runs, copies y back, through pinned staging moved
 * on up to 8 host threads (no more than the CPUs the process may run on; pageable
 * copies if pinned memory is refused).  Each host-entry call (this one and the
 * reference-style ones) takes a buffer set of its own, so concurrent calls do not
 * wait for each other; the sets' pinned staging and device buffers are kept per
 * device between calls while the device's kept totals stay within the limits
 * below, and calls beyond them allocate and free their own. */
int fpcc_allpairs(const fpcc_job *job, const float *h_x, float *h_y);
/* Limits of the pinned host and device bytes the host entries keep per device
 * between calls (defaults 1 GiB and 8 GiB; 0 keeps nothing).  Buffers already
 * kept stay until fpcc_release_buffers. */
void fpcc_set_kept_limits(size_t host_bytes, size_t device_bytes);
/* Frees every kept buffer on every device (the reference frees its buffers at the
 * end of each call, src/ccs_cuda.cu:600-603; here that happens on request).
 * Buffers held by calls still running are freed when those calls return.
 * Returns the number of such calls. */
int fpcc_release_buffers(void);
/* The two halves of fpcc_allpairs_device.  The station stage writes every
 * (station, window) spectrum into the workspace; the pair stage reads them and
 * writes y.  Spectra left in the workspace may be reused by later pair stages of
 * jobs with the same kind, stations, windows, N and lags. */
int fpcc_station_stage_device(const fpcc_job *job, const float *d_x, void *d_ws, size_t ws_bytes, void *stream);
int fpcc_pair_stage_device(const fpcc_job *job, void *d_ws, size_t ws_bytes, float *d_y, void *stream);
/* Station stages of njobs jobs on one stream, each into its own workspace
 * (d_x[j], d_ws[j], ws_bytes[j]).  Jobs of uneven station count, window count,
 * window length and stride share kernel launches when their transforms agree
 * (GNCC and 1-bit GNCC of equal Nz; PCC2 of equal Nz <= 8192, each job with its
 * own window-length FFT; PCC2 of equal Nz > 8192, whose AmpNormf and zero-padded
 * transforms share launches after each job's analytic signal on its own plan;
 * WPCC whose families share the filter bank: wavelet type, Nz, scales, voices,
 * first scale and w0; up to 16 jobs per launch); the others run one by one.  The
 * results equal fpcc_station_stage_device per job. */
int fpcc_station_stage_batch_device(const fpcc_job *jobs, int njobs, const float *const *d_x, void *const *d_ws,
                                    const size_t *ws_bytes, void *stream);
/* Pair stages of njobs jobs on one stream, after their station stages, each from
 * its own workspace into its own output (d_ws[j], ws_bytes[j], d_y[j]).  Jobs of
 * uneven station count, window count, window length and lags share launches (up
 * to 16 a launch): PCC2 jobs whose windows fit one workgroup at Nz 8192
 * (2048 < N <= 4096) with stacks NONE, PLAIN or ZSUM, one launch of the
 * persistent pair kernel per stack; GNCC and 1-bit GNCC jobs whose windows fit one
 * workgroup (Nz <= 8192) with stacks NONE or PLAIN, one launch per kind, Nz and
 * stack.  The others run one by one.  The results equal fpcc_pair_stage_device
 * per job. */
int fpcc_pair_stage_batch_device(const fpcc_job *jobs, int njobs, void *const *d_ws, const size_t *ws_bytes,
                                 float *const *d_y, void *stream);

/* ---------------------------------------------------------------------------
 * Per-station preprocessing (the steps src/PCC_fullpair_1b.c runs on each
 * station's traces before correlating them; here a trace is one window of the
 * job's station rows, x[s*station_stride + w*win_len + n]).  In place on d_x, in
 * this order:
 *   sign       optional [S][W] polarity factors (+1/-1), the per-trace result of
 *              CorrectRevesedPolarity (src/rotlib.c:14-34); every later step is
 *              odd-symmetric, so applying it first changes no bit;
 *   zero       windows of zeros are dropped (RemoveZeroTraces, src/ReadManySacs.c:529-559);
 *   rms        [S][W] sqrt(sum of x^2 in FP64 / N), before clipping (:355-363);
 *   clip       |x| > 4*median|x|/0.6745 -> +-that threshold (clipping, :881-891);
 *   std_max    > 0: windows with rms > std_max * (median rms of the station's
 *              non-zero windows) are dropped (RemoveOutlierTraces, :893-925);
 *   common_windows != 0: a window is kept only where every station kept it (the
 *              paired trace lists of a two-station run, :399-401);
 *   dropped windows are zeroed; valid[S][W] = 1 for kept windows;
 *   band, dt   band[0] > 0 and band[1] > band[0] (Hz, dt in s): FP64 spectral
 *              whitening of each station's kept windows by their averaged
 *              amplitude spectrum (AveWhite, :1065-1177); windows of 17 to 8192
 *              samples (-2 otherwise).
 * n_stations, n_windows (at most 8192), win_len and station_stride of the job
 * are used; the kind, lags and stack are not.  Returns 0, -1 (invalid) or -2 (unsupported).
 * No counterpart in the reference's library API: these are static functions
 * of its driver.
 * ------------------------------------------------------------------------- */
typedef struct fpcc_preproc {
    int clip;            /* clipping on/off                                    */
    float std_max;       /* outlier factor (0: off)                            */
    double band[2];      /* whitening band in Hz (0, 0: off)                   */
    double dt;           /* sampling interval in s (whitening)                 */
    int common_windows;  /* keep only windows kept at every station            */
} fpcc_preproc;
size_t fpcc_preprocess_workspace_size(const fpcc_job *job, const fpcc_preproc *pp);
int fpcc_preprocess_device(const fpcc_job *job, const fpcc_preproc *pp, float *d_x, const float *d_sign,
                           unsigned char *d_valid, float *d_rms, void *d_ws, size_t ws_bytes, void *stream);
/* Host buffers: x is [n_stations][station_stride] floats (updated in place), sign
 * [S][W] or NULL, valid [S][W] out, rms [S][W] out or NULL. */
int fpcc_preprocess(const fpcc_job *job, const fpcc_preproc *pp, float *h_x, const float *h_sign,
                    unsigned char *h_valid, float *h_rms);

/* ---------------------------------------------------------------------------
 * Graph replay (no counterpart in the reference): one job's station and pair
 * stages captured once as a HIP graph on `stream` (not the default stream) and
 * replayed with one launch.  fpcc_graph_create runs the job eagerly first (so
 * every plan and table exists), then captures it; each fpcc_graph_launch
 * recomputes d_y from the current contents of d_x, with the same buffers.
 * ------------------------------------------------------------------------- */
typedef struct fpcc_graph fpcc_graph;
int fpcc_graph_create(const fpcc_job *job, const float *d_x, float *d_y, void *d_ws, size_t ws_bytes, void *stream,
                      fpcc_graph **graph);
int fpcc_graph_launch(fpcc_graph *graph, void *stream);
int fpcc_graph_destroy(fpcc_graph *graph);

/* ---------------------------------------------------------------------------
 * Request server (no counterpart in the reference, whose driver correlates one
 * pair per process run, src/PCC_fullpair_1b.c:467-499): continuous batching of
 * all-pairs jobs on the current HIP device.
 *
 * fpcc_server_create reserves one device arena (device_bytes; 0 = 40% of free
 * HBM) and one pinned host staging arena (pinned_bytes; 0 = 512 MiB).  A request
 * is admitted when its input, output and workspace (fpcc_workspace_size) fit in
 * the device arena and its pageable buffers fit in the staging arena; requests
 * are admitted in submission order, except that later ones may pass a request
 * that does not fit yet, at most max_bypass times (0 = 8).  Up to max_inflight
 * (0 = 16) requests run at once on n_streams (0 = 4) HIP streams, overlapping
 * their copies and kernels.  Pinned caller buffers are copied by DMA; pageable
 * ones through the staging arena.  x must stay valid and unchanged, and y
 * untouched, until the request reaches a final state.  fail_every > 0 makes
 * every fail_every-th request fail before its kernels launch (fault injection
 * for tests).
 *
 * Resident station spectra: with cache_bytes > 0, a request that names its
 * stations (station_keys, one 64-bit key per station row; equal keys promise
 * equal samples in the job's windows) reuses the station-stage output (spectra,
 * GNCC energies, 1-bit sign spectra, WPCC per-scale spectra) of earlier requests
 * with the same kind, window length, window count, zero-pad length and filter
 * parameters: only missing stations are computed (gathered into one job), their
 * rows are kept in a separate cache arena of cache_bytes, and every station's
 * rows are copied into the request's workspace before its pair stage.  Entries
 * in use by running requests stay; others are evicted least recently used.
 *
 * Uneven jobs in one launch: when requests whose station stages can share
 * kernels (fpcc_station_stage_batch_device: GNCC or 1-bit of one zero-pad length,
 * PCC2 of one zero-pad length, WPCC of one filter bank; without station keys) are
 * next in line together, they are admitted as a group and their
 * station stages run as one set of launches, their pair stages too where
 * fpcc_pair_stage_batch_device shares them.
 *
 * Return codes: 0 ok; -1 invalid argument or GPU failure; -3 |lag| >= N;
 * -4 the request's footprint exceeds the server's arenas (never admissible);
 * fpcc_server_wait returns 1 on timeout, fpcc_server_cancel 1 if the request had
 * already finished.  Cancelling a running request drops its result (y is not
 * written through the staging arena; a pinned y may already hold it).
 * Latencies run from submission to the result in y.  All functions are
 * thread-safe; destroy cancels queued requests and waits for running ones.
 * ------------------------------------------------------------------------- */
enum { FPCC_REQ_QUEUED = 0, FPCC_REQ_RUNNING = 1, FPCC_REQ_DONE = 2, FPCC_REQ_FAILED = 3, FPCC_REQ_CANCELLED = 4 };

typedef struct {
    size_t device_bytes;  /* device arena for request buffers (0: 40% of free HBM) */
    size_t pinned_bytes;  /* pinned staging arena for pageable buffers (0: 512 MiB) */
    int max_inflight;     /* requests on the device at once (0: 16)               */
    int n_streams;        /* HIP streams (0: 4)                                   */
    int max_bypass;       /* times a waiting request may be passed (0: 8)          */
    int fail_every;       /* fault injection: fail every k-th request (0: off)    */
    size_t cache_bytes;   /* resident station-spectra cache (0: off)              */
} fpcc_server_config;

typedef struct {
    fpcc_job job;   /* one all-pairs job (layout as fpcc_allpairs)        */
    const float *x; /* host input: n_stations rows of station_stride floats */
    float *y;       /* host output: fpcc_output_len(&job) floats            */
    const unsigned long long *station_keys; /* per-station identity for the cache, or NULL */
} fpcc_request;

typedef struct {
    int state;                     /* FPCC_REQ_*                              */
    int rc;                        /* error code of a failed request          */
    double t_submit, t_start, t_done; /* seconds since the server was created */
} fpcc_request_status;

typedef struct {
    unsigned long long submitted, completed, failed, cancelled, rejected;
    unsigned long long cache_hits, cache_misses; /* stations found / computed for keyed requests */
    size_t cache_used;                           /* bytes of resident station rows              */
    unsigned long long batched_requests, batch_launches; /* requests whose station stages shared a launch; groups */
    unsigned long long queued, inflight, inflight_peak;
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests (p50, p99: of the
                                               last 2^20 of them)                          */
    double pair_correlations;              /* sum of pairs x windows delivered        */
    double busy_s;                         /* first admission to last completion      */
    size_t device_bytes, device_used, device_peak;
} fpcc_serving_stats;

typedef struct fpcc_server fpcc_server;

/* Pinned (page-locked) host memory from the library's own HIP runtime, for
 * request buffers the server copies by DMA (NULL on failure). */
void *fpcc_host_alloc(size_t bytes);
void fpcc_host_free(void *p);

int fpcc_server_create(const fpcc_server_config *cfg, fpcc_server **server);
int fpcc_server_submit(fpcc_server *server, const fpcc_request *req, unsigned long long *id);
/* n requests entering the queue together (ids[k] for reqs[k]; on an error the
 * requests from the failing one on are not queued and their ids are 0). */
int fpcc_server_submit_many(fpcc_server *server, const fpcc_request *reqs, int n, unsigned long long *ids);
int fpcc_server_poll(fpcc_server *server, unsigned long long id, fpcc_request_status *status);
/* timeout_ms < 0 waits without limit */
int fpcc_server_wait(fpcc_server *server, unsigned long long id, int timeout_ms, fpcc_request_status *status);
int fpcc_server_cancel(fpcc_server *server, unsigned long long id);
/* Drops the record of a finished request (its id becomes unknown). */
int fpcc_server_forget(fpcc_server *server, unsigned long long id);
int fpcc_server_stats(fpcc_server *server, fpcc_serving_stats *stats);
int fpcc_server_destroy(fpcc_server *server);

/* ---------------------------------------------------------------------------
 * Autotuning (no counterpart in the reference).  Transforms longer than one
 * workgroup's LDS run as two passes, n = n1 * n2 (n1 <= 512 column points, n2 <=
 * 2048 row points, both 7-smooth).  The split comes from the table
 * tuning/fft_splits_gfx950.txt beside the library (or $FASTPCC_FFT_TUNING), else
 * from a heuristic.  fpcc_autotune_fft times every admissible split on `batch`
 * forward transforms of n points (reps repetitions, HIP events), returns the
 * fastest (and the heuristic's split and time, for comparison) and uses it for
 * plans of n first built later in the process; fpcc_fft_split returns the split
 * in use (0 if n has none).  Returns 0, or -1 if n has no admissible split.
 * ------------------------------------------------------------------------- */
int fpcc_autotune_fft(int n, int batch, int reps, int *best_n1, double *best_ms, int *default_n1,
                      double *default_ms);
int fpcc_fft_split(int n);
/* Radix order of the single-workgroup FP32 LDS FFTs (lengths up to 8192; 3600,
 * 4096 and 8192 keep the order their compile-time kernels use): from the table
 * tuning/fft_radix_gfx950.txt (or $FASTPCC_RADIX_TUNING), else the default
 * factorisation.  fpcc_autotune_radix times every order of every multiset of
 * supported radices (2, 3, 4, 5, 7, 8, 9, 16, 25) with product n, largest-first
 * and smallest-first, on `batch` transforms (reps repetitions), returns the
 * fastest in radices[npass] (and the default order's time) and uses it for plans
 * of n first built later in the process; fpcc_fft_radices returns the pass count
 * of the plan in use and its radices (0 if n has none). */
int fpcc_autotune_radix(int n, int batch, int reps, int *radices, int *npass, double *best_ms, double *default_ms);
int fpcc_fft_radices(int n, int *radices);

/* ---------------------------------------------------------------------------
 * Kernel timing for benchmarks (no counterpart in the reference).  While
 * enabled, each kernel launch records a HIP event pair on its stream;
 * fpcc_profile_report waits for them and returns, per kernel name (written
 * newline-separated into names), the summed milliseconds and launch counts.
 * Returns the number of kernels reported.
 * ------------------------------------------------------------------------- */
void fpcc_profile_enable(int on);
void fpcc_profile_reset(void);
int fpcc_profile_report(char *names, size_t names_len, double *total_ms, long long *counts, int max_kernels);

#ifdef __cplusplus
}
#endif
#endif
