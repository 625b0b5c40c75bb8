// Two-pass (four-step) FFT kernels with fused pro- and epilogues; see bigfft.hpp.
#include <cmath>
#include <map>
#include <mutex>
#include <cstdlib>
#include <string>
#include "../../include/fastpcc.h"
#include "bigfft.hpp"
#include "edge.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "pairs.hpp"

namespace fpcc {

// Split n = n1*n2 with n1 <= 512 (column pass in LDS), n2 <= 2048 (row pass),
// both 7-smooth.  The split comes from the autotuning table when it lists n,
// else n1 is the admissible divisor closest to sqrt(n), with n2 >= 16 when
// possible.
static bool valid_split(int n, int d)
{
    if (d < 4 || d > 512 || d > n || n % d) return false;
    const int e = n / d;
    FftPlan t1, t2;
    return e <= 2048 && factor_plan(d, t1) && factor_plan(e, t2);
}

static int heuristic_split(int n)
{
    int best = 0;
    double bestscore = 1e300;
    for (int d = 4; d <= 512 && d <= n; ++d) {
        if (!valid_split(n, d)) continue;
        double score = std::fabs(std::log((double)d) - 0.5 * std::log((double)n));
        if (n / d < kBigCols) score += 100;  // partial column tiles waste lanes
        if (score < bestscore) {
            bestscore = score;
            best = d;
        }
    }
    return best;
}

// Tuning table: lines "n n1 [ms]" ('#' starts a comment), read once from
// $FASTPCC_FFT_TUNING or tuning/fft_splits_<arch>.txt beside the library.
static std::mutex tune_mu;
static std::map<int, int> &tune_table()
{
    static std::map<int, int> t;
    static bool loaded = false;
    if (loaded) return t;
    loaded = true;
    const char *e = getenv("FASTPCC_FFT_TUNING");
    const std::string path = e ? std::string(e) : tuning_file("fft_splits_gfx950.txt");
    if (FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r")) {
        char line[256];
        while (fgets(line, sizeof line, f)) {
            if (line[0] == '#') continue;
            int n = 0, d = 0;
            if (sscanf(line, "%d %d", &n, &d) == 2 && valid_split(n, d)) t[n] = d;
        }
        fclose(f);
    }
    return t;
}

static int chosen_split(int n)
{
    std::lock_guard<std::mutex> lk(tune_mu);
    auto &t = tune_table();
    auto it = t.find(n);
    return it != t.end() ? it->second : heuristic_split(n);
}

static BigPlan build_big_plan(int n, int n1)
{
    BigPlan b;
    if (n < 64 || !n1) return b;
    b.n = n;
    b.n1 = n1;
    b.n2 = n / n1;
    const DevPlan &a = get_plan(b.n1, false), &c = get_plan(b.n2, false);
    if (!a.ok || !c.ok) return b;
    b.p1 = a.plan;
    b.p2 = c.plan;
    b.tw1 = (const cf32 *)a.tw;
    b.tw2 = (const cf32 *)c.tw;
    b.twN = (const cf32 *)get_split_twiddles(n);
    b.ok = b.twN != nullptr;
    return b;
}

static BigPlan make_big_plan(int n) { return n < 64 ? BigPlan{} : build_big_plan(n, chosen_split(n)); }

const BigPlan &big_plan(int n)
{
    static std::map<std::pair<int, int>, BigPlan> cache;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, n});
    if (it != cache.end()) return it->second;
    static const BigPlan none{};
    if (!table_builds_allowed()) return none;
    return cache.emplace(std::make_pair(dev, n), make_big_plan(n)).first->second;
}

// index of W_n^{c r} in the full table: c < N2 and r < N1, so c*r < n
// W_n^{c r} from the two-factor table of get_split_twiddles (c r < n): two small
// cached lookups instead of a gather over an n-entry table
__device__ __forceinline__ cf32 tw_cr(const cf32 *__restrict__ twN, int c, int r)
{
    const unsigned m = (unsigned)c * (unsigned)r;
    return cmul(twN[kTwLo + m / kTwLo], twN[m % kTwLo]);
}

// Job-table launch (BATCH): batch index b runs over the rows of several jobs; job j
// (jb.b0[j] <= b < jb.b0[j+1]) supplies x, its stride, W, N and the destination
// rows (COL_FWD_INPUT), or its rows and maxima (COL_INV_MAX, COL_FWD_AMPNORM), or
// its rows (row passes; ROW_INV_MUL also its input spectra).
__device__ __forceinline__ int batch_job(const RealBatch &jb, int b)
{
    int j = 0;
    while (j + 1 < jb.n && b >= jb.b0[j + 1]) ++j;
    return j;
}

// Compile-time batched transforms for the common splits (2^17 .. 2^19 as 256 x
// 512 .. 256 x 2048): NTR transforms of NPT points transform in place in one LDS
// buffer, transform t at t * LDT, point i at ct_pad(i) (one pad slot per 16
// values), 256 threads, each pass reading all of a thread's butterflies before
// writing.  Columns: 16 x 256 points, radix 16 x 16, stride 274; rows: 4 x 512
// (8 x 8 x 8, stride 544) or 4 x 1024 (16 x 8 x 8, stride 1088).  The strides keep
// the loads, passes and stores within 1.5x of conflict-free LDS access.
constexpr int kCtN1 = 256, kCtLD = 274;

template <int NPT, int NTR, int LDT, int DIR, int R, int NS>
__device__ __forceinline__ void ctb_pass(cf32 *lds, const cf32 *__restrict__ tw)
{
    constexpr int NBF = NPT / R, TOT = NTR * NBF, BPT = (TOT + 255) / 256;
    cf32 v[BPT][R];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int bb = threadIdx.x + 256 * i;
        if (TOT % 256 == 0 || bb < TOT) {
            const int t = bb / NBF, b = bb % NBF;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = lds[t * LDT + ct_pad(b + r * NBF)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int bb = threadIdx.x + 256 * i;
        if (TOT % 256 == 0 || bb < TOT) {
            const int t = bb / NBF, b = bb % NBF;
            const int k = b % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cf32 w = tw[(r - 1) * NS + k];
                    if (DIR > 0) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            Dft<R, DIR>::run(v[i]);
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[t * LDT + ct_pad(base + r * NS)] = v[i][r];
        }
    }
    __syncthreads();
}

// Register-ended passes of a tile's in-place transforms (transform t's point p at
// lds[t * LDT + ct_pad(p)]), one butterfly b of transform t a call.  First pass
// (stride 1): the butterfly's inputs, points b + r NPT/R, are in v; its outputs go
// to LDS (the caller puts a barrier after every thread's call).
template <int NPT, int LDT, int DIR, int R>
__device__ __forceinline__ void ctb_first_regs(cf32 *v, int t, int b, cf32 *lds)
{
    Dft<R, DIR>::run(v);
#pragma unroll
    for (int r = 0; r < R; ++r) lds[t * LDT + ct_pad(b * R + r)] = v[r];
}
// Last pass (stride NPT/R): the butterfly's outputs, points b + r NPT/R, stay in v.
template <int NPT, int LDT, int DIR, int R>
__device__ __forceinline__ void ctb_last_regs(const cf32 *lds, int t, int b, cf32 *v, const cf32 *__restrict__ tw)
{
    constexpr int NS = NPT / R;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = lds[t * LDT + ct_pad(b + r * NS)];
#pragma unroll
    for (int r = 1; r < R; ++r) {
        cf32 w = tw[(r - 1) * NS + b];
        if (DIR > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
    }
    Dft<R, DIR>::run(v);
}

template <int NPT, int NTR, int LDT, int DIR, int NS, int OFF, int R, int... Rest>
__device__ __forceinline__ void ctb_passes(cf32 *lds, const cf32 *__restrict__ tw)
{
    ctb_pass<NPT, NTR, LDT, DIR, R, NS>(lds, tw + OFF);
    if constexpr (sizeof...(Rest) > 0) ctb_passes<NPT, NTR, LDT, DIR, NS * R, OFF + (R - 1) * NS, Rest...>(lds, tw);
}

// tw: the per-pass twiddles of the plan with radices Rs (in that order)
template <int NPT, int NTR, int LDT, int DIR, int... Rs>
__device__ __forceinline__ cf32 *ctb_fft(cf32 *lds, const cf32 *__restrict__ tw)
{
    static_assert((Rs * ... * 1) == NPT, "radices must multiply to NPT");
    ctb_passes<NPT, NTR, LDT, DIR, 1, 0, Rs...>(lds, tw);
    return lds;
}

template <int DIR>
__device__ __forceinline__ cf32 *ct_cols_fft(cf32 *lds, const cf32 *__restrict__ tw)
{
    return ctb_fft<kCtN1, kBigCols, kCtLD, DIR, 16, 16>(lds, tw);
}

// rows: stride of the in-place layout for N2 (0: not specialised)
__host__ __device__ constexpr int ct_row_ld(int n2) { return n2 == 512 ? 544 : n2 == 1024 ? 1088 : 0; }

template <int CTN2, int DIR>
__device__ __forceinline__ cf32 *ct_rows_fft(cf32 *lds, const cf32 *__restrict__ tw)
{
    if constexpr (CTN2 == 512) return ctb_fft<512, kBigRows, ct_row_ld(512), DIR, 8, 8, 8>(lds, tw);
    else return ctb_fft<1024, kBigRows, ct_row_ld(1024), DIR, 16, 8, 8>(lds, tw);
}

template <bool BATCH, bool CT>
__global__ void __launch_bounds__(256)
k_big_cols(ColArgs a, int N1, int N2, FftPlan p1, const cf32 *__restrict__ tw1, const cf32 *__restrict__ twN,
           RealBatch jb)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // column t, point r at cix(t, r): padded column stride N1 + 1 with ping-pong
    // buffers, or the in-place layout of ct_cols_fft (N1 = 256, tw1 its twiddles)
    const int ld = CT ? kCtLD : N1 + 1;
    auto cix = [&](int t, int r) { return CT ? t * kCtLD + ct_pad(r) : t * ld + r; };
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + kBigCols * ld;
    float *red = reinterpret_cast<float *>(CT ? buf1 : buf1 + kBigCols * ld);
    int b = blockIdx.y;
    if (BATCH) {
        const int j = batch_job(jb, b);
        b -= jb.b0[j];
        if (a.op == COL_FWD_INPUT) {
            a.x = jb.x[j];
            a.x_stride = jb.stride[j];
            a.W = jb.W[j];
            a.N = jb.N[j];
            a.dst = (cf32 *)jb.X[j];
        } else if (a.op == COL_FWD_AMPNORMF) {  // job j's analytic rows (Xin, N long) into its rows X
            a.src = (const cf32 *)jb.Xin[j];
            a.src_bstride = jb.N[j];
            a.nvalid = jb.N[j];
            a.maxbuf = jb.mx[j];
            a.dst = (cf32 *)jb.X[j];
        } else {  // COL_INV_MAX, COL_FWD_AMPNORM: job j's rows and maxima
            a.src = a.dst = (cf32 *)jb.X[j];
            a.maxbuf = jb.mx[j];
        }
    }
    const int c0 = blockIdx.x * kBigCols;
    const long long base = (long long)b * a.bstride;
    const long long sbase = (long long)b * (a.src_bstride ? a.src_bstride : a.bstride);
    const int n = N1 * N2;
    const int op = a.op;
    float eps = 0.f;
    if (op == COL_FWD_AMPNORM) eps = 1e-6f * a.maxbuf[b];
    if (op == COL_FWD_AMPNORMF) eps = (float)(1e-6 * sqrt((double)a.maxbuf[b]));
    const float *xr = nullptr;
    if (op == COL_FWD_INPUT || op == COL_FWD_CHIRP) {
        const int s = b / a.W, w = b - s * a.W;
        xr = a.x + (long long)s * a.x_stride + (long long)w * a.N;
    }
    // the value a point of the tile starts with, from its loaded source (u: the
    // complex row element, xv: the input sample, in: inside the input)
    auto start = [&](long long g, cf32 u, float xv, bool in) -> cf32 {
        if (op == COL_FWD_INPUT || op == COL_FWD_CHIRP) {
            if (!in) return cf32{0.f, 0.f};
            float sv;
            if (a.sign_rule == 1) sv = xv >= 0.f ? 1.f : -1.f;
            else if (a.sign_rule == 2) sv = xv > 0.f ? 1.f : -1.f;
            else sv = xv * a.in_scale;
            if (op == COL_FWD_CHIRP) {
                const cf32 ch = a.chirp[g - a.off];
                return cf32{sv * ch.x, sv * ch.y};
            }
            return cf32{sv, 0.f};
        }
        if (op == COL_FWD_AMPNORM) {
            const float rr = 1.0f / sqrtf(u.x * u.x + u.y * u.y + eps);
            return __builtin_isnormal(rr) ? cf32{u.x * rr, u.y * rr} : cf32{0.f, 0.f};
        }
        if (op == COL_FWD_AMPNORMF) {
            if (!in) return cf32{0.f, 0.f};
            const float rr = 1.0f / (sqrtf(u.x * u.x + u.y * u.y) + eps);
            return __builtin_isnormal(rr) ? cf32{u.x * rr, u.y * rr} : cf32{0.f, 0.f};
        }
        return u;
    };
    const bool inv = op == COL_INV_MAX || op == COL_INV_LAGWIN || op == COL_CHIRP_MID || op == COL_INV_CHIRP_MAX;
    const bool fwd_out = !inv || op == COL_CHIRP_MID;
    float m = 0.f;
    // LAGWIN batch decode
    int p = 0, w = 0, i = 0, j = 0;  // pair, window, stations (LAGWIN)
    if (op == COL_INV_LAGWIN) {
        const int q = b / a.Wb;
        w = a.wsel + (b - q * a.Wb);
        p = a.pair0 + q;
        if (a.norm == 1) pair_index(p, a.S, i, j);
    }
    // one output point of the transformed tile: column t, row r, value v
    auto emit = [&](int t, int r, cf32 v) {
        const int c = c0 + t;
        const long long g = (long long)r * N2 + c;
        if (fwd_out) {
            a.dst[base + g] = cmul(v, tw_cr(twN, c, r));  // W_n^{n2 k1}
        } else if (op == COL_INV_CHIRP_MAX) {
            if (g < a.N) {
                const cf32 d = cmul(a.chirp[g], v);
                const cf32 z = cf32{d.x, -d.y};
                a.dst[(long long)b * a.dst_bstride + g] = z;
                m = fmaxf(m, z.x * z.x + z.y * z.y);
            }
        } else if (op == COL_INV_MAX) {
            a.dst[base + g] = v;
            m = fmaxf(m, v.x * v.x + v.y * v.y);
        } else {
            long long d = g - a.lag0;  // in (-n, 2n): lag0 is in (-n, n)
            if (d < 0) d += n;
            if (d >= n) d -= n;
            const int l = (int)d;
            if (l < a.L) {
                const int lag = a.lag0 + l;
                double o = (double)(a.scale * v.x);
                const int m = a.Nsig - abs(lag);
                if (a.edge && m > 0 && m < a.K) {
                    // short overlap: the FP32 transform's absolute error, divided by
                    // the overlap's small energy, would dominate; sum it exactly
                    int ei = i, ej = j;
                    if (a.norm != 1) pair_index(p, a.S, ei, ej);
                    const int w0 = a.esum ? 0 : w, w1 = a.esum ? a.Wall : w + 1;
                    const double c = edge_corr(a.edge, a.K, a.Wall, ei, ej, w0, w1, lag, m, a.norm == 2);
                    o = c * (double)a.scale * (double)n;  // the transform's scale: v.x is n c
                }
                if (a.norm == 1) {
                    const double *P1 = a.E + (long long)(i * a.Wfull + w) * a.lde;
                    const double *P2 = a.E + (long long)(j * a.Wfull + w) * a.lde;
                    const int N = a.Nsig;
                    const int a1 = max(0, -lag), b1 = min(N, N - lag);
                    const int a2 = max(0, lag), b2 = min(N, N + lag);
                    o = o / sqrt((P1[b1] - P1[a1]) * (P2[b2] - P2[a2]));
                } else if (a.norm == 2) {
                    o = o / (double)(a.Nsig - abs(lag));
                }
                if (a.accumulate) a.y[(long long)p * a.L + l] += (float)o;
                else a.y[((long long)p * a.Wfull + w) * a.L + l] = (float)o;
            }
        }
    };
    bool regs = false;  // transformed in registers and emitted
    if constexpr (CT) {
        // all of a thread's loads first (16 points of 256 columns-rows), then the work:
        // a loop that loads, computes and stores one point at a time waits for every load
        constexpr int PER = kBigCols * kCtN1 / 256;
        const bool xin = op == COL_FWD_INPUT || op == COL_FWD_CHIRP;
        cf32 u[PER];
        float xv[PER];
        bool in[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = threadIdx.x + 256 * i, t = idx % kBigCols, r = idx / kBigCols, c = c0 + t;
            const long long g = (long long)r * N2 + c;
            u[i] = cf32{0.f, 0.f};
            xv[i] = 0.f;
            in[i] = false;
            if (c < N2) {
                if (xin) {
                    const long long m = g - a.off;
                    in[i] = m >= 0 && m < a.N;
                    if (in[i]) xv[i] = xr[m];
                } else {
                    in[i] = op != COL_FWD_AMPNORMF || g < a.nvalid;
                    if (in[i]) u[i] = a.src[sbase + g];
                }
            }
        }
        cf32 v[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = threadIdx.x + 256 * i, t = idx % kBigCols, r = idx / kBigCols, c = c0 + t;
            v[i] = c < N2 ? start((long long)r * N2 + c, u[i], xv[i], in[i]) : cf32{0.f, 0.f};
        }
        if (op != COL_CHIRP_MID) {
            // thread (column t, q) holds rows q + 16 i: the inputs of the first radix-16
            // butterfly q of column t; the second pass's butterfly q gives rows q + 16 i
            // back, the points this thread stores.  One barrier between the passes.
            const int t = threadIdx.x % kBigCols, q = threadIdx.x / kBigCols, c = c0 + t;
            if (inv) ctb_first_regs<kCtN1, kCtLD, +1, 16>(v, t, q, buf0);
            else ctb_first_regs<kCtN1, kCtLD, -1, 16>(v, t, q, buf0);
            __syncthreads();
            if (inv) ctb_last_regs<kCtN1, kCtLD, +1, 16>(buf0, t, q, v, tw1 + 15);
            else ctb_last_regs<kCtN1, kCtLD, -1, 16>(buf0, t, q, v, tw1 + 15);
            if (fwd_out) {
                // the step twiddles of all of a thread's points first, then the stores
                cf32 wt[PER];
#pragma unroll
                for (int i = 0; i < PER; ++i) wt[i] = c < N2 ? tw_cr(twN, c, q + 16 * i) : cf32{0.f, 0.f};
#pragma unroll
                for (int i = 0; i < PER; ++i)
                    if (c < N2) a.dst[base + (long long)(q + 16 * i) * N2 + c] = cmul(v[i], wt[i]);  // W_n^{n2 k1}
            } else {
#pragma unroll
                for (int i = 0; i < PER; ++i)
                    if (c < N2) emit(t, q + 16 * i, v[i]);
            }
            regs = true;
        } else {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = threadIdx.x + 256 * i, t = idx % kBigCols, r = idx / kBigCols;
                buf0[cix(t, r)] = v[i];
            }
        }
    }
    if (!regs) {
        for (int idx = threadIdx.x; !CT && idx < kBigCols * N1; idx += blockDim.x) {
            const int t = idx % kBigCols, r = idx / kBigCols;
            const int c = c0 + t;
            cf32 v = cf32{0.f, 0.f};
            if (c < N2) {
                const long long g = (long long)r * N2 + c;
                if (op == COL_FWD_INPUT || op == COL_FWD_CHIRP) {
                    const long long m = g - a.off;
                    if (m >= 0 && m < a.N) {
                        const float xv = xr[m];
                        float s;
                        if (a.sign_rule == 1) s = xv >= 0.f ? 1.f : -1.f;
                        else if (a.sign_rule == 2) s = xv > 0.f ? 1.f : -1.f;
                        else s = xv * a.in_scale;
                        v = cf32{s, 0.f};
                        if (op == COL_FWD_CHIRP) v = cf32{s * a.chirp[m].x, s * a.chirp[m].y};
                    }
                } else if (op == COL_FWD_AMPNORM) {
                    const cf32 u = a.src[sbase + g];
                    const float rr = 1.0f / sqrtf(u.x * u.x + u.y * u.y + eps);
                    if (__builtin_isnormal(rr)) v = cf32{u.x * rr, u.y * rr};
                } else if (op == COL_FWD_AMPNORMF) {
                    if (g < a.nvalid) {
                        const cf32 u = a.src[sbase + g];
                        const float rr = 1.0f / (sqrtf(u.x * u.x + u.y * u.y) + eps);
                        if (__builtin_isnormal(rr)) v = cf32{u.x * rr, u.y * rr};
                    }
                } else {
                    v = a.src[sbase + g];
                }
            }
            buf0[cix(t, r)] = v;
        }
        __syncthreads();
        cf32 *res;
        if (CT) res = inv ? ct_cols_fft<+1>(buf0, tw1) : ct_cols_fft<-1>(buf0, tw1);
        else
            res = inv ? fft_lds_batch<float, +1>(buf0, buf1, p1, tw1, kBigCols, ld)
                      : fft_lds_batch<float, -1>(buf0, buf1, p1, tw1, kBigCols, ld);
        if (op == COL_CHIRP_MID) {
            // second transform's chirped input: conj(mask X) w = mask conj(conv) (|w| = 1)
            const int half = (a.N + 1) >> 1;
            for (int idx = threadIdx.x; idx < kBigCols * N1; idx += blockDim.x) {
                const int t = idx % kBigCols, r = idx / kBigCols;
                const long long g = (long long)r * N2 + c0 + t;
                float gm = 0.f;
                if (c0 + t < N2 && g < a.N) {
                    if (g == 0) gm = a.da1;
                    else if (g < half) gm = 2 * a.da1;
                    else if (!(a.N & 1) && g == (a.N >> 1)) gm = a.da1;
                }
                const cf32 v = res[cix(t, r)];
                res[cix(t, r)] = cf32{gm * v.x, -gm * v.y};
            }
            __syncthreads();
            res = CT ? ct_cols_fft<-1>(res, tw1) : fft_lds_batch<float, -1>(res, res == buf0 ? buf1 : buf0, p1, tw1, kBigCols, ld);
        }
        bool done = false;
        if constexpr (CT) {
            if (fwd_out) {
                // the step twiddles of all of a thread's points first, then the stores
                constexpr int PER = kBigCols * kCtN1 / 256;
                cf32 w[PER];
    #pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int idx = threadIdx.x + 256 * i, t = idx % kBigCols, r = idx / kBigCols, c = c0 + t;
                    w[i] = c < N2 ? tw_cr(twN, c, r) : cf32{0.f, 0.f};
                }
    #pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int idx = threadIdx.x + 256 * i, t = idx % kBigCols, r = idx / kBigCols, c = c0 + t;
                    if (c < N2) a.dst[base + (long long)r * N2 + c] = cmul(res[cix(t, r)], w[i]);  // W_n^{n2 k1}
                }
                done = true;
            }
        }
        for (int idx = threadIdx.x; !done && idx < kBigCols * N1; idx += blockDim.x) {
            const int t = idx % kBigCols, r = idx / kBigCols;
            if (c0 + t >= N2) continue;
            emit(t, r, res[cix(t, r)]);
        }
    }
    if (op == COL_INV_MAX || op == COL_INV_CHIRP_MAX) {
        m = block_max(m, red);
        if (threadIdx.x == 0) atomicMax(reinterpret_cast<int *>(a.maxbuf + b), __float_as_int(m));
    }
}

// BATCH: the rows of job j's destination (in place, ROW_FWD_PLAIN only).
// CTN2 = 512 or 1024: N2 = CTN2 transforms in place (ct_rows_fft, tw2 its
// twiddles); 0: the runtime plan with ping-pong buffers.
template <bool BATCH, int CTN2>
__global__ void __launch_bounds__(256)
k_big_rows(RowArgs a, int N1, int N2, FftPlan p2, const cf32 *__restrict__ tw2, const cf32 *__restrict__ twN,
           RealBatch jb)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // row rr, point c at rix(rr, c)
    auto rix = [&](int rr, int c) { return CTN2 ? rr * ct_row_ld(CTN2) + ct_pad(c) : rr * N2 + c; };
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw);
    cf32 *buf1 = buf0 + kBigRows * N2;
    int b = blockIdx.y;
    if (BATCH) {
        const int j = batch_job(jb, b);
        b -= jb.b0[j];
        a.src = a.dst = (cf32 *)jb.X[j];
        if (a.op == ROW_INV_MUL) a.X = (const cf32 *)jb.Xin[j];
    }
    const int r0 = blockIdx.x * kBigRows;
    const long long base = (long long)b * a.bstride;
    const long long n = (long long)N1 * N2;
    const int op = a.op;
    const cf32 *Xs = nullptr, *Fs = nullptr;
    if (op == ROW_INV_MUL) {
        const int sw = b / a.S, s = b - sw * a.S;
        Xs = a.X + sw * n;
        Fs = a.fw + s * n;
    }
    int ci = 0, cj = 0, cw0 = 0, cw1 = 0;
    const cf32 *F2 = a.F2 ? a.F2 : a.F;
    if (op == ROW_INV_CROSS) {
        const int q = b / a.Wb;
        pair_index(a.pair0 + q, a.Sst, ci, cj);
        cw0 = a.wsum ? 0 : a.wsel + (b - q * a.Wb);
        cw1 = a.wsum ? a.Wfull : cw0 + 1;
    }
    bool loaded = false;
    if constexpr (CTN2 != 0) {
        if (op == ROW_FWD_PLAIN || op == ROW_FWD_SCALE || op == ROW_INV_MUL || op == ROW_INV_PLAIN) {
            // transformed in registers between a first and a last pass: thread (row t, q)
            // loads points q + 64 k of its row (coalesced across q), the inputs of the
            // first pass's butterfly q; the last pass's butterflies give points
            // b + (CTN2/8) k back, which the thread stores
            constexpr int LD = ct_row_ld(CTN2), R1 = CTN2 == 512 ? 8 : 16, NB1 = CTN2 / R1;
            constexpr int NSL = CTN2 / 8, BPL = kBigRows * NSL / 256;  // last pass: radix 8, stride NSL
            static_assert(kBigRows * NB1 == 256, "one first-pass butterfly a thread");
            const bool inv = op == ROW_INV_MUL || op == ROW_INV_PLAIN;
            const int t = threadIdx.x / NB1, q = threadIdx.x % NB1, r = r0 + t;
            cf32 v[R1], f[R1];
#pragma unroll
            for (int k = 0; k < R1; ++k) {
                const long long g = (long long)r * CTN2 + q + NB1 * k;
                v[k] = f[k] = cf32{0.f, 0.f};
                if (r < N1) {
                    if (op == ROW_INV_MUL) {
                        v[k] = Xs[g];
                        f[k] = Fs[g];
                    } else {
                        v[k] = a.src[base + g];
                    }
                }
            }
            if (op == ROW_INV_MUL) {
#pragma unroll
                for (int k = 0; k < R1; ++k) v[k] = cmul(v[k], f[k]);
            }
            if (inv) ctb_first_regs<CTN2, LD, +1, R1>(v, t, q, buf0);
            else ctb_first_regs<CTN2, LD, -1, R1>(v, t, q, buf0);
            __syncthreads();
            if (inv) ctb_pass<CTN2, kBigRows, LD, +1, 8, R1>(buf0, tw2 + (R1 - 1));
            else ctb_pass<CTN2, kBigRows, LD, -1, 8, R1>(buf0, tw2 + (R1 - 1));
            cf32 o[BPL][8];
#pragma unroll
            for (int i = 0; i < BPL; ++i) {
                const int bb = threadIdx.x + 256 * i;
                if (inv) ctb_last_regs<CTN2, LD, +1, 8>(buf0, bb / NSL, bb % NSL, o[i], tw2 + (R1 - 1) + 7 * R1);
                else ctb_last_regs<CTN2, LD, -1, 8>(buf0, bb / NSL, bb % NSL, o[i], tw2 + (R1 - 1) + 7 * R1);
            }
            const float sc = op == ROW_FWD_SCALE ? a.wscale[b % a.S] : 1.0f;
            // the step twiddles of all of a thread's points first, then the stores
            cf32 w[BPL][8];
#pragma unroll
            for (int i = 0; i < BPL; ++i) {
                const int bb = threadIdx.x + 256 * i, rl = r0 + bb / NSL, bl = bb % NSL;
#pragma unroll
                for (int k = 0; k < 8; ++k) w[i][k] = inv && rl < N1 ? tw_cr(twN, bl + NSL * k, rl) : cf32{0.f, 0.f};
            }
#pragma unroll
            for (int i = 0; i < BPL; ++i) {
                const int bb = threadIdx.x + 256 * i, rl = r0 + bb / NSL, bl = bb % NSL;
                if (rl >= N1) continue;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int c = bl + NSL * k;
                    const cf32 x = o[i][k];
                    a.dst[base + (long long)rl * CTN2 + c] =
                        inv ? cmulc(w[i][k], x) : op == ROW_FWD_SCALE ? cf32{x.x * sc, x.y * sc} : x;  // W_n^{-n2 k1}
                }
            }
            return;
        }
        if (op != ROW_INV_CROSS) {
            // all of a thread's loads first, then the LDS writes (a loop that loads and
            // stores one point at a time waits for every load)
            constexpr int PER = kBigRows * CTN2 / 256;
            cf32 u[PER], f[PER];
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = threadIdx.x + 256 * i, rr = idx / CTN2, c = idx - rr * CTN2, r = r0 + rr;
                const long long g = (long long)r * CTN2 + c;
                u[i] = f[i] = cf32{0.f, 0.f};
                if (r < N1) {
                    if (op == ROW_INV_MUL) {
                        u[i] = Xs[g];
                        f[i] = Fs[g];
                    } else {
                        u[i] = a.src[base + g];
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = threadIdx.x + 256 * i, rr = idx / CTN2, c = idx - rr * CTN2;
                buf0[rix(rr, c)] = op == ROW_INV_MUL ? cmul(u[i], f[i]) : u[i];
            }
            loaded = true;
        }
    }
    for (int idx = threadIdx.x; !loaded && idx < kBigRows * N2; idx += blockDim.x) {
        const int rr = idx / N2, c = idx - rr * N2;
        const int r = r0 + rr;
        cf32 v = cf32{0.f, 0.f};
        if (r < N1) {
            const long long g = (long long)r * N2 + c;
            if (op == ROW_INV_MUL) {
                v = cmul(Xs[g], Fs[g]);
            } else if (op == ROW_INV_CROSS) {
                for (int w = cw0; w < cw1; ++w) {
                    const cf32 t = cmulc(a.F[(long long)(ci * a.Wfull + w) * n + g], F2[(long long)(cj * a.Wfull + w) * n + g]);
                    v = cf32{v.x + t.x, v.y + t.y};
                }
            } else {
                v = a.src[base + g];
            }
        }
        buf0[rix(rr, c)] = v;
    }
    __syncthreads();
    const bool inv = op == ROW_INV_MUL || op == ROW_INV_PLAIN || op == ROW_INV_CROSS;
    cf32 *res;
    if constexpr (CTN2 != 0) res = inv ? ct_rows_fft<CTN2, +1>(buf0, tw2) : ct_rows_fft<CTN2, -1>(buf0, tw2);
    else
        res = (inv ? fft_lds_batch<float, +1>(buf0, buf1, p2, tw2, kBigRows, N2)
                   : fft_lds_batch<float, -1>(buf0, buf1, p2, tw2, kBigRows, N2));
    const float sc = (op == ROW_FWD_SCALE) ? a.wscale[b % a.S] : 1.0f;
    const int half = (a.N + 1) >> 1;
    if (op == ROW_MUL_INV) {
        for (int idx = threadIdx.x; idx < kBigRows * N2; idx += blockDim.x) {
            const int rr = idx / N2, c = idx - rr * N2;
            if (r0 + rr < N1) res[rix(rr, c)] = cmul(res[rix(rr, c)], a.T[(long long)(r0 + rr) * N2 + c]);
        }
        __syncthreads();
        if constexpr (CTN2 != 0) res = ct_rows_fft<CTN2, +1>(res, tw2);
        else res = fft_lds_batch<float, +1>(res, res == buf0 ? buf1 : buf0, p2, tw2, kBigRows, N2);
    }
    if (op == ROW_FWD_MASK || op == ROW_MASK_INV) {
        for (int idx = threadIdx.x; idx < kBigRows * N2; idx += blockDim.x) {
            const int rr = idx / N2, c = idx - rr * N2;
            const long long k = (r0 + rr) + (long long)N1 * c;  // natural bin
            float gm;
            if (k == 0) gm = a.da1;
            else if (k < half) gm = 2 * a.da1;
            else if (!(a.N & 1) && k == (a.N >> 1)) gm = a.da1;
            else gm = 0.f;
            res[rix(rr, c)] = cf32{res[rix(rr, c)].x * gm, res[rix(rr, c)].y * gm};
        }
        if (op == ROW_MASK_INV) {
            __syncthreads();
            if constexpr (CTN2 != 0) res = ct_rows_fft<CTN2, +1>(res, tw2);
            else res = fft_lds_batch<float, +1>(res, (res == buf0) ? buf1 : buf0, p2, tw2, kBigRows, N2);
        }
    }
    const bool inv_out = inv || op == ROW_MASK_INV || op == ROW_MUL_INV;
    bool stored = false;
    if constexpr (CTN2 != 0) {
        if (inv_out) {
            constexpr int PER = kBigRows * CTN2 / 256;
            cf32 w[PER];
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = threadIdx.x + 256 * i, rr = idx / CTN2, c = idx - rr * CTN2, r = r0 + rr;
                w[i] = r < N1 ? tw_cr(twN, c, r) : cf32{0.f, 0.f};
            }
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = threadIdx.x + 256 * i, rr = idx / CTN2, c = idx - rr * CTN2, r = r0 + rr;
                if (r < N1) a.dst[base + (long long)r * CTN2 + c] = cmulc(w[i], res[rix(rr, c)]);  // W_n^{-n2 k1}
            }
            stored = true;
        }
    }
    for (int idx = threadIdx.x; !stored && idx < kBigRows * N2; idx += blockDim.x) {
        const int rr = idx / N2, c = idx - rr * N2;
        const int r = r0 + rr;
        if (r >= N1) continue;
        const long long g = (long long)r * N2 + c;
        cf32 v = res[rix(rr, c)];
        if (inv_out) {
            v = cmulc(tw_cr(twN, c, r), v);  // W_n^{-n2 k1}
        } else if (op == ROW_FWD_SCALE) {
            v = cf32{v.x * sc, v.y * sc};
        }
        a.dst[base + g] = v;
    }
}

// One workgroup per station-window: P[n] = sum_{m<n} x[m]^2 in FP64, n in [0, N].
// BATCH: the station-windows of the jobs of jb (x, stride, W, N, E, lde per job).
template <bool BATCH>
__global__ void __launch_bounds__(512)
k_energy_prefix(const float *__restrict__ x, long long station_stride, int W, int N, double *__restrict__ E,
                long long lde, RealBatch jb)
{
    __shared__ double part[512];
    int sw = blockIdx.x;
    if (BATCH) {
        const int j = batch_job(jb, sw);
        sw -= jb.b0[j];
        x = jb.x[j];
        station_stride = jb.stride[j];
        W = jb.W[j];
        N = jb.N[j];
        E = jb.E[j];
        lde = jb.lde[j];
    }
    const int s = sw / W, w = sw - s * W;
    const float *in = x + (long long)s * station_stride + (long long)w * N;
    const int T = blockDim.x;
    const int per = (N + T - 1) / T;
    const int a0 = threadIdx.x * per, b0 = min(N, a0 + per);
    double acc = 0.0;
    for (int n = a0; n < b0; ++n) {
        const double v = (double)in[n];
        acc += v * v;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < T; ++t) {
            const double v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    double *P = E + (long long)sw * lde;
    for (int n = a0; n < b0; ++n) {
        P[n] = run;
        const double v = (double)in[n];
        run += v * v;
    }
    if (a0 < N && b0 == N) P[N] = run;
}

int launch_energy_prefix(const float *d_x, long long station_stride, int S, int W, int N, double *d_E, long long lde,
                         hipStream_t st)
{
    ProfScope prof("energy_prefix", st);
    RealBatch none{};
    hipLaunchKernelGGL(k_energy_prefix<false>, dim3(S * W), dim3(512), 0, st, d_x, station_stride, W, N, d_E, lde,
                       none);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int launch_energy_prefix_batch(const RealBatch &jb, hipStream_t st)
{
    const int total = jb.b0[jb.n];
    if (total <= 0) return 0;
    ProfScope prof("energy_prefix_batch", st);
    hipLaunchKernelGGL(k_energy_prefix<true>, dim3(total), dim3(512), 0, st, nullptr, 0LL, 1, 0, nullptr, 0LL, jb);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

__global__ void k_fill_zero(float *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

static void allow_lds(const void *fn, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int bluestein_n(int N)
{
    int n = 64;
    while (n < 2 * N - 1) n <<= 1;
    return n;
}

// b[m] = conj(w_m)/n for m < N and b[n-m] = conj(w_m)/n for 0 < m < N (n >= 2N-1
// keeps the two ranges apart), zero between; chirp[m] = w_m.
__global__ void k_bluestein_chirp(int N, int n, cf32 *__restrict__ chirp, cf32 *__restrict__ b)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int j = m < N ? m : (m > n - N ? n - m : -1);
    cf32 v = cf32{0.f, 0.f};
    if (j >= 0) {
        const long long q = ((long long)j * j) % (2LL * N);  // exact reduction of the phase
        double sn, cs;
        sincospi(-(double)q / (double)N, &sn, &cs);
        if (m < N) chirp[m] = cf32{(float)cs, (float)sn};
        const double inv = 1.0 / (double)n;
        v = cf32{(float)(cs * inv), (float)(-sn * inv)};
    }
    b[m] = v;
}

int bluestein_tables(int N, int n, cf32 *chirp, cf32 *T, hipStream_t st)
{
    const BigPlan &bp = big_plan(n);
    if (!bp.ok || n < 2 * N - 1) return -1;
    hipLaunchKernelGGL(k_bluestein_chirp, dim3((n + 255) / 256), dim3(256), 0, st, N, n, chirp, T);
    FPCC_CHECK(hipGetLastError());
    ColArgs ca;
    ca.op = COL_FWD_PLAIN;
    ca.src = T;
    ca.dst = T;
    ca.bstride = n;
    int r = run_cols(bp, ca, 1, st, "bluestein_table_cols");
    if (r) return r;
    RowArgs ra;
    ra.op = ROW_FWD_PLAIN;
    ra.src = T;
    ra.dst = T;
    ra.bstride = n;
    return run_rows(bp, ra, 1, st, "bluestein_table_rows");
}

template <bool BATCH>
static int launch_cols(const BigPlan &bp, const ColArgs &a, const RealBatch &jb, int nbatch, hipStream_t st,
                       const char *name)
{
    const dim3 grid((bp.n2 + kBigCols - 1) / kBigCols, nbatch);
    if (bp.n1 == kCtN1) {  // compile-time in-place column transform, one LDS buffer
        const cf32 *twc = (const cf32 *)get_order_twiddles(kCtN1, {16, 16});
        if (!twc) return -1;
        const size_t lds = sizeof(cf32) * kBigCols * (size_t)kCtLD + sizeof(float) * 8;
        ProfScope prof(name, st);
        hipLaunchKernelGGL((k_big_cols<BATCH, true>), grid, dim3(256), lds, st, a, bp.n1, bp.n2, bp.p1, twc, bp.twN,
                           jb);
    } else {
        const size_t lds = sizeof(cf32) * 2 * kBigCols * (size_t)(bp.n1 + 1) + sizeof(float) * 8;
        allow_lds((const void *)k_big_cols<BATCH, false>, lds);
        ProfScope prof(name, st);
        hipLaunchKernelGGL((k_big_cols<BATCH, false>), grid, dim3(256), lds, st, a, bp.n1, bp.n2, bp.p1, bp.tw1,
                           bp.twN, jb);
    }
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int run_cols(const BigPlan &bp, const ColArgs &a, int nbatch, hipStream_t st, const char *name)
{
    if (nbatch <= 0) return 0;
    if (!bp.ok) return -1;
    if (nbatch > 65535) {
        fprintf(stderr, "fastpcc: %d transforms exceed one launch\n", nbatch);
        return -1;
    }
    RealBatch none{};
    return launch_cols<false>(bp, a, none, nbatch, st, name);
}

int run_cols_batch(const BigPlan &bp, const ColArgs &a, const RealBatch &jb, hipStream_t st, const char *name)
{
    const int nbatch = jb.b0[jb.n];
    if (nbatch <= 0) return 0;
    if (!bp.ok || nbatch > 65535) return -1;
    if (a.op != COL_FWD_INPUT && a.op != COL_INV_MAX && a.op != COL_FWD_AMPNORM && a.op != COL_FWD_AMPNORMF)
        return -1;
    return launch_cols<true>(bp, a, jb, nbatch, st, name);
}

template <bool BATCH, int CTN2>
static int launch_rows_ct(const BigPlan &bp, const RowArgs &a, const RealBatch &jb, int nbatch, hipStream_t st,
                          const char *name, const cf32 *twc)
{
    const size_t lds = sizeof(cf32) * kBigRows * (size_t)ct_row_ld(CTN2);
    ProfScope prof(name, st);
    hipLaunchKernelGGL((k_big_rows<BATCH, CTN2>), dim3((bp.n1 + kBigRows - 1) / kBigRows, nbatch), dim3(256), lds,
                       st, a, bp.n1, bp.n2, bp.p2, twc, bp.twN, jb);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

template <bool BATCH>
static int launch_rows(const BigPlan &bp, const RowArgs &a, const RealBatch &jb, int nbatch, hipStream_t st,
                       const char *name)
{
    if (bp.n2 == 512 || bp.n2 == 1024) {  // compile-time in-place row transforms, one LDS buffer
        const cf32 *twc = (const cf32 *)(bp.n2 == 512 ? get_order_twiddles(512, {8, 8, 8})
                                                      : get_order_twiddles(1024, {16, 8, 8}));
        if (!twc) return -1;
        return bp.n2 == 512 ? launch_rows_ct<BATCH, 512>(bp, a, jb, nbatch, st, name, twc)
                            : launch_rows_ct<BATCH, 1024>(bp, a, jb, nbatch, st, name, twc);
    }
    const size_t lds = sizeof(cf32) * 2 * kBigRows * (size_t)bp.n2;
    allow_lds((const void *)k_big_rows<BATCH, 0>, lds);
    ProfScope prof(name, st);
    hipLaunchKernelGGL((k_big_rows<BATCH, 0>), dim3((bp.n1 + kBigRows - 1) / kBigRows, nbatch), dim3(256), lds, st,
                       a, bp.n1, bp.n2, bp.p2, bp.tw2, bp.twN, jb);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

int run_rows(const BigPlan &bp, const RowArgs &a, int nbatch, hipStream_t st, const char *name)
{
    if (nbatch <= 0) return 0;
    if (!bp.ok) return -1;
    if (nbatch > 65535) {
        fprintf(stderr, "fastpcc: %d transforms exceed one launch\n", nbatch);
        return -1;
    }
    RealBatch none{};
    return launch_rows<false>(bp, a, none, nbatch, st, name);
}

int run_rows_batch(const BigPlan &bp, const RowArgs &a, const RealBatch &jb, hipStream_t st, const char *name)
{
    const int nbatch = jb.b0[jb.n];
    if (nbatch <= 0) return 0;
    if (!bp.ok || nbatch > 65535) return -1;
    if (a.op != ROW_FWD_PLAIN && a.op != ROW_INV_MUL && a.op != ROW_FWD_SCALE) return -1;
    return launch_rows<true>(bp, a, jb, nbatch, st, name);
}

int tune_split(int n, int batch, int reps, int *best_n1, double *best_ms, int *heur_n1, double *heur_ms)
{
    if (n < 64 || batch < 1 || reps < 1) return -1;
    std::vector<int> cand;
    for (int d = 4; d <= 512 && d <= n; ++d)
        if (valid_split(n, d)) cand.push_back(d);
    if (cand.empty()) return -1;
    void *buf = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    double best = 1e300, heur = 0;
    int bestd = 0;
    const int hd = heuristic_split(n);
    if (hipMalloc(&buf, sizeof(cf32) * (size_t)n * batch) != hipSuccess) return -1;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess && hipMemsetAsync(buf, 0, sizeof(cf32) * (size_t)n * batch, st) == hipSuccess) {
        rc = 0;
        for (int d : cand) {
            BigPlan bp = build_big_plan(n, d);
            if (!bp.ok) continue;
            ColArgs ca;
            ca.op = COL_FWD_PLAIN;
            ca.src = ca.dst = (cf32 *)buf;
            ca.bstride = n;
            RowArgs ra;
            ra.op = ROW_FWD_PLAIN;
            ra.src = ra.dst = (cf32 *)buf;
            ra.bstride = n;
            if (run_cols(bp, ca, batch, st, "tune_cols") || run_rows(bp, ra, batch, st, "tune_rows")) continue;
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps; ++r)
                if (run_cols(bp, ca, batch, st, "tune_cols") || run_rows(bp, ra, batch, st, "tune_rows")) break;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = -1;
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (d == hd) heur = ms / reps;
            if (ms / reps < best) {
                best = ms / reps;
                bestd = d;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    if (rc || !bestd) return -1;
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        tune_table()[n] = bestd;  // used by plans built from now on
    }
    if (best_n1) *best_n1 = bestd;
    if (best_ms) *best_ms = best;
    if (heur_n1) *heur_n1 = hd;
    if (heur_ms) *heur_ms = heur;
    return 0;
}

int fill_zero(float *p, long long n, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    FPCC_CHECK(hipGetLastError());
    return 0;
}

}  // namespace fpcc

extern "C" int fpcc_autotune_fft(int n, int batch, int reps, int *best_n1, double *best_ms, int *default_n1,
                                 double *default_ms)
{
    return fpcc::tune_split(n, batch, reps, best_n1, best_ms, default_n1, default_ms);
}

extern "C" int fpcc_fft_split(int n)
{
    const fpcc::BigPlan &bp = fpcc::big_plan(n);
    return bp.ok ? bp.n1 : 0;
}
