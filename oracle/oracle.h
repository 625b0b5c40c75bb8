/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle.
 *
 * A from-scratch CPU restatement of the reference's correlation engine
 * (src/FFTapps.c and src/FWTa/wavelet_def_v7.c of sergiventosa/FastPCC), used by
 * the tests as the checker and by bench.py as the timed CPU baseline.  Nothing
 * in the product library links, loads or calls it.
 *
 * Two builds of every correlation are provided:
 *   orc_*    "reference precision": FP32 where the reference uses fftwf
 *            (PCC2, PCC1, PCC-v, 1-bit), FP64 where it uses fftw (GNCC, WPCC),
 *            with the same mixed-precision accumulators and roundings;
 *   orc64_*  "all-FP64": the same algorithm with every step in double, the
 *            yardstick against which each output type's error is stated.
 * FFTW is replaced by the oracle's own FFT (offt.c), so FFTW's rounding is not
 * reproduced: parity is at tolerance level, never bitwise, for FFT outputs.
 */
#ifndef FASTPCC_ORACLE_H
#define FASTPCC_ORACLE_H
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc_max_threads(void); /* OpenMP threads the drivers use */

enum { ORC_PCC2 = 0, ORC_PCC1 = 1, ORC_PCCV = 2, ORC_GNCC = 3, ORC_CC1B = 4, ORC_WPCC = 5 };

/* ---- building blocks (src/FFTapps.c) ---- */
int orc_analytic_f(float complex *y, const float *x, unsigned N);            /* :104-124 */
int orc_analytic_d(double complex *y, const double *x, unsigned N);          /* :45-80   */
void orc_ampnormf(float complex *y, unsigned N);                             /* :178-205 */
void orc_ampnorm(double complex *y, unsigned N);                             /* :147-175 */
void orc_ampnormf_d(double complex *y, unsigned N);   /* AmpNormf's formula evaluated in FP64 */

/* ---- drivers with the reference's signatures (src/FFTapps.h:11-16) ---- */
int orc_pcc_set(float **y, float **x1, float **x2, int N, unsigned Tr, double v, int Lag1, int Lag2);
int orc_pcc1_set(float **y, float **x1, float **x2, int N, unsigned Tr, int Lag1, int Lag2);
int orc_pcc2_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2);
int orc_ccgn_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2);
int orc_cc1b_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2);
int orc_tspcc2_set(float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2,
                   double pmin, double pmax, unsigned V, int type, double op1);

/* ---- flat-array entry points for ctypes: x1/x2 are [Tr][N], y is [Tr][L] ----
 * kind is one of ORC_*.  v is PCC-v's power; pmin/pmax/V/type/op1 are WPCC's. */
int orc_run_f32(int kind, float *y, const float *x1, const float *x2, unsigned N, unsigned Tr,
                int Lag1, int Lag2, double v, double pmin, double pmax, unsigned V, int type, double op1);
int orc_run_f64(int kind, double *y, const float *x1, const float *x2, unsigned N, unsigned Tr,
                int Lag1, int Lag2, double v, double pmin, double pmax, unsigned V, int type, double op1);

/* All station pairs (i<j, p = i*S - i*(i+1)/2 + (j-i-1)), x is [S][W*N], window w
 * = samples [w*N, (w+1)*N).  stacked != 0: y is [P][L], the mean over windows;
 * else y is [P][W][L].  One pair per *_set call, OpenMP over windows inside it,
 * the way the reference's one-pair-per-run CLI would be driven. */
int orc_allpairs_f32(int kind, float *y, const float *x, unsigned S, unsigned W, unsigned N,
                     int Lag1, int Lag2, int stacked, unsigned pair_begin, unsigned pair_end,
                     double v, double pmin, double pmax, unsigned V, int type, double op1);
int orc_allpairs_f64(int kind, double *y, const float *x, unsigned S, unsigned W, unsigned N,
                     int Lag1, int Lag2, int stacked, unsigned pair_begin, unsigned pair_end,
                     double v, double pmin, double pmax, unsigned V, int type, double op1);

/* Phase-weighted stack over W windows (Schimmel & Paulssen 1997) of the all-FP64
 * per-window correlations of pairs [pair_begin, pair_end): y [pairs][L].  The
 * instantaneous phase is that of AnalyticSignal (:45-80) of each window's whole
 * circular correlation (Nz lags), read at the lag window.  kind: ORC_PCC2,
 * ORC_CC1B or ORC_GNCC.  power 0 gives the linear stack. */
int orc_pws_f64(int kind, double *y, const float *x, unsigned S, unsigned W, unsigned N, int Lag1, int Lag2,
                unsigned pair_begin, unsigned pair_end, double power);

/* ---- wavelet filter bank (src/FWTa/wavelet_def_v7.c:204-339) ---- */
typedef struct {
    unsigned S;           /* number of scales = J*V             */
    unsigned J, V;
    unsigned Nz;          /* FFT length after the doubling rule */
    unsigned Ls0;         /* Ls[S-1]                            */
    double s0, K0, op1;
    double *scale;        /* [S]                                 */
    unsigned *Ls;         /* [S]                                 */
    int *center;          /* [S]                                 */
    double complex **wc;  /* [S] -> Ls[s] samples                */
} orc_wfamily;

/* Reproduces tspcc2_set's parameter setup (src/FFTapps.c:1076-1115): returns NULL
 * for an unsupported type or lags, otherwise the family with Nz already doubled. */
orc_wfamily *orc_wpcc_family(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                             int type, double op1);
void orc_wfamily_free(orc_wfamily *w);
/* The WPCC correlation with a caller-built filter bank (the semantics of
 * cuda_tspcc2_set_freq, src/ccs_cuda.cu:898-1036, with the CPU AmpNorm phasor of
 * src/FFTapps.c:1116-1252), all FP64: x1/x2 [Tr][N], y [Tr][|Lag2-Lag1|+1] from
 * min(Lag1, Lag2); wflat holds the S wavelets concatenated, interleaved re,im. */
int orc_wpcc_bank_f64(double *y, const float *x1, const float *x2, unsigned N, unsigned Tr, unsigned Nz,
                      int Lag1, int Lag2, unsigned S, const double *scale, const double *wflat,
                      const unsigned *Ls, const int *center, double C);
double orc_erfi(double z);
/* Flat export for ctypes: fills scale/Ls/center (length S) and the wavelet samples
 * concatenated (interleaved re,im) into wbuf (length 2*sum Ls); returns S or <0. */
int orc_wpcc_family_export(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                           int type, double op1, unsigned *out_meta /* S,J,V,Nz,Ls0 */,
                           double *out_s0_K0_op1, double *scale, unsigned *Ls, int *center,
                           double *wbuf, unsigned wbuf_len);

/* ---- per-station preprocessing (src/PCC_fullpair_1b.c, src/ReadManySacs.c, src/rotlib.c) ---- */
float orc_qabsmedian(const float *x, int n);                                 /* :1179-1214 */
float orc_qmedian(const float *x, int n);                                    /* :1216-1251 */
float orc_clip_threshold(const float *x, unsigned N);                        /* :885       */
void orc_clipping(float *x, unsigned N);                                     /* :881-891   */
float orc_trace_rms(const float *x, unsigned N);                             /* :355-363   */
unsigned orc_white_nz(unsigned N);                                           /* :1072      */
void orc_white_filter(double *H, double *HS, unsigned Nh, unsigned Nz, unsigned N, unsigned n1, unsigned n2);
int orc_avewhite(float **x0, unsigned N, unsigned Tr, const double freq[2], double dt); /* :1065-1177 */
int orc_preprocess(float *x, unsigned S, unsigned W, unsigned N, unsigned long long stride, int clip, float nstd,
                   double f1, double f2, double dt, int common, const float *sign, unsigned char *valid, float *rms);

#ifdef __cplusplus
}
#endif
#endif
