// Wavelet phase cross-correlation (tspcc2_set, src/FFTapps.c:1076-1255) on gfx950.
//
// Per (station, window) with samples x[0..N): in = 0 except in[Ls0+n] = x[n]/Nz,
// X = FFT_Nz(in); per scale s: phi_s = AmpNorm(IFFT_Nz(X . W_s)) with
// AmpNorm(z) = z / sqrt(|z|^2 + 1e-6 max|z|^2) (src/FFTapps.c:147-175), and
// F_s = FFT_Nz(phi_s) * sqrt(1/(scale_s Nz)).  Per pair:
//   y[lag] = Re IFFT_Nz( sum_s conj(F_1s) F_2s ) [lag] / (K0 N),  K0 = sum_s 1/scale_s,
// with W_s = FFT_Nz of the wavelet circularly shifted by its centre (:1140-1147).
// The sum over scales (and over windows for the stack) is a per-frequency Gram
// of the station matrix, run by the MFMA Gram kernel.
//
// FFTs of Nz = N1*N2 points (powers of two, too long for one workgroup's LDS)
// run as two passes through HBM (four-step): a column pass (N2 columns of N1
// points, stride N2, then the twiddles W_Nz^{n2 k1}) and a row pass (N1 rows of
// N2 contiguous points).  The forward transform leaves bin k = k1 + N1 k2 at
// position k1*N2 + k2; every frequency-domain product and the Gram work on that
// order, and the inverse (rows first, conjugate twiddles, then columns) returns
// to natural order.  Pro- and epilogues are fused into the passes.
#include <map>
#include <mutex>
#include <tuple>
#include <vector>
#include "engine.hpp"
#include "kernels.hpp"
#include "profile.hpp"
#include "wavelet.hpp"
#include "bigfft.hpp"

namespace fpcc {

// Filter-bank spectra W_s (permuted order) and sqrt(1/(scale_s Nz)), cached per
// device and family (the wavelets are rounded to FP32 before their FFT).
struct FilterBank {
    cf32 *fw = nullptr;
    float *wscale = nullptr;
};

// Families built by the library (type -1/-2/-3) are cached; a caller-supplied
// family (type 0) gets fresh buffers that free_filter_bank releases.
static int filter_bank(const WFamily &f, hipStream_t st, FilterBank &out)
{
    // a library family is fixed by type, w0, first scale, voices, count and Nz
    typedef std::tuple<int, int, unsigned, unsigned, unsigned, double, double> Key;
    static std::map<Key, FilterBank> cache;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const bool cached = f.type != 0;
    const Key key(dev, f.type, f.Nz, f.S, f.V, f.s0, f.op1);
    std::unique_lock<std::mutex> lk(mu);
    if (cached) {
        auto it = cache.find(key);
        if (it != cache.end()) {
            out = it->second;
            return 0;
        }
    }
    if (!table_builds_allowed()) return -1;
    const BigPlan &bp = big_plan((int)f.Nz);
    if (!bp.ok) return -1;
    const size_t Nz = f.Nz;
    std::vector<cf32> h(Nz * f.S, cf32{0.f, 0.f});
    std::vector<float> ws(f.S);
    for (unsigned s = 0; s < f.S; ++s) {
        cf32 *row = h.data() + s * Nz;
        const int c = f.center[s], Ls = (int)f.Ls[s];
        for (int u = 0; u < Ls - c; ++u) row[u] = cf32{(float)f.wc[s][c + u].real(), (float)f.wc[s][c + u].imag()};
        for (int u = 0; u < c; ++u)
            row[Nz - c + u] = cf32{(float)f.wc[s][u].real(), (float)f.wc[s][u].imag()};
        ws[s] = (float)std::sqrt(1.0 / (f.scale[s] * (double)f.Nz));
    }
    FilterBank fb;
    FPCC_CHECK(hipMalloc(&fb.fw, sizeof(cf32) * h.size()));
    FPCC_CHECK(hipMalloc(&fb.wscale, sizeof(float) * ws.size()));
    FPCC_CHECK(hipMemcpy(fb.fw, h.data(), sizeof(cf32) * h.size(), hipMemcpyHostToDevice));
    FPCC_CHECK(hipMemcpy(fb.wscale, ws.data(), sizeof(float) * ws.size(), hipMemcpyHostToDevice));
    ColArgs ca{};
    ca.op = COL_FWD_PLAIN;
    ca.src = fb.fw;
    ca.dst = fb.fw;
    ca.bstride = (long long)Nz;
    int r = run_cols(bp, ca, (int)f.S, st, "wpcc_filter_cols");
    if (r) return r;
    RowArgs ra{};
    ra.op = ROW_FWD_PLAIN;
    ra.src = fb.fw;
    ra.dst = fb.fw;
    ra.bstride = (long long)Nz;
    ra.S = 1;
    if ((r = run_rows(bp, ra, (int)f.S, st, "wpcc_filter_rows"))) return r;
    FPCC_CHECK(hipStreamSynchronize(st));
    if (cached) cache[key] = fb;
    out = fb;
    return 0;
}

static void free_filter_bank(const WFamily &f, FilterBank &fb)
{
    if (f.type != 0) return;
    if (fb.fw) (void)hipFree(fb.fw);
    if (fb.wscale) (void)hipFree(fb.wscale);
    fb = FilterBank{};
}

// Workspace: X [SW][Nz] | B [SW*S][Nz] | max [SW*S] | U [Nz][S][W*S] | G [P][Nz]
struct WpccLayout {
    size_t X, B, mx, U, G, total;
};

static size_t al(size_t v) { return (v + 255) / 256 * 256; }

static WpccLayout wpcc_layout(int S, int W, int Nz, int nsc)
{
    WpccLayout l{};
    const size_t SW = (size_t)S * W, P = (size_t)S * (S - 1) / 2;
    l.X = 0;
    l.B = l.X + al(sizeof(cf32) * SW * Nz);
    l.mx = l.B + al(sizeof(cf32) * SW * nsc * Nz);
    l.U = l.mx + al(sizeof(float) * SW * nsc);
    l.G = l.U + al(sizeof(cf32) * (size_t)Nz * SW * nsc);
    l.total = l.G + al(sizeof(cf32) * P * (size_t)gram_ld(Nz));
    return l;
}

size_t wpcc_workspace(int S, int W, int N, int Nz, int nscales)
{
    (void)N;
    return wpcc_layout(S, W, Nz, nscales).total;
}

// The pair stage reads only the per-scale spectra B (rows sw*S + scale).
int wpcc_station_regions(int S, int W, int Nz, int nscales, size_t *off, size_t *per)
{
    off[0] = wpcc_layout(S, W, Nz, nscales).B;
    per[0] = sizeof(cf32) * (size_t)W * nscales * Nz;
    return 1;
}

int wpcc_station_stage(const float *d_x, long long station_stride, int S, int W, int N, const WFamily &f, char *ws,
                       hipStream_t st)
{
    const BigPlan &bp = big_plan((int)f.Nz);
    if (!bp.ok) {
        fprintf(stderr, "fastpcc: unsupported WPCC length Nz=%u\n", f.Nz);
        return -1;
    }
    FilterBank fb;
    int r = filter_bank(f, st, fb);
    if (r) return r;
    struct Release {
        const WFamily &f;
        FilterBank &fb;
        hipStream_t st;
        ~Release()
        {
            if (f.type == 0) {
                (void)hipStreamSynchronize(st);
                free_filter_bank(f, fb);
            }
        }
    } release{f, fb, st};
    const int SW = S * W, nsc = (int)f.S;
    WpccLayout l = wpcc_layout(S, W, (int)f.Nz, nsc);
    cf32 *X = (cf32 *)(ws + l.X), *B = (cf32 *)(ws + l.B);
    float *mx = (float *)(ws + l.mx);
    const long long Nz = f.Nz;
    // X = FFT(in), in[Ls0 + n] = x[n] / Nz
    ColArgs ca{};
    ca.op = COL_FWD_INPUT;
    ca.dst = X;
    ca.bstride = Nz;
    ca.x = d_x;
    ca.x_stride = station_stride;
    ca.W = W;
    ca.N = N;
    ca.off = (int)f.Ls0;
    ca.in_scale = (float)(1.0 / (double)Nz);
    if ((r = run_cols(bp, ca, SW, st, "wpcc_input_cols"))) return r;
    RowArgs ra{};
    ra.op = ROW_FWD_PLAIN;
    ra.src = X;
    ra.dst = X;
    ra.bstride = Nz;
    ra.S = 1;
    if ((r = run_rows(bp, ra, SW, st, "wpcc_input_rows"))) return r;
    // per (sw, s): IFFT(X W_s) with max|z|^2, AmpNorm, FFT, * sqrt(1/(scale Nz))
    const long long nb = (long long)SW * nsc;
    if ((r = fill_zero(mx, nb, st))) return r;
    ra = RowArgs{};
    ra.op = ROW_INV_MUL;
    ra.dst = B;
    ra.bstride = Nz;
    ra.X = X;
    ra.fw = fb.fw;
    ra.S = nsc;
    if ((r = run_rows(bp, ra, (int)nb, st, "wpcc_scale_inv_rows"))) return r;
    ca = ColArgs{};
    ca.op = COL_INV_MAX;
    ca.src = B;
    ca.dst = B;
    ca.bstride = Nz;
    ca.maxbuf = mx;
    if ((r = run_cols(bp, ca, (int)nb, st, "wpcc_scale_inv_cols"))) return r;
    ca.op = COL_FWD_AMPNORM;
    if ((r = run_cols(bp, ca, (int)nb, st, "wpcc_ampnorm_cols"))) return r;
    ra = RowArgs{};
    ra.op = ROW_FWD_SCALE;
    ra.src = B;
    ra.dst = B;
    ra.bstride = Nz;
    ra.S = nsc;
    ra.wscale = fb.wscale;
    return run_rows(bp, ra, (int)nb, st, "wpcc_fwd_rows");
}

bool wpcc_same_bank(const WFamily &a, const WFamily &b)
{
    // library families: the filter bank is fixed by these (as filter_bank's cache key)
    return a.type != 0 && a.type == b.type && a.Nz == b.Nz && a.S == b.S && a.V == b.V && a.s0 == b.s0 &&
           a.op1 == b.op1 && a.Ls0 == b.Ls0;
}

int wpcc_station_stage_batch(int n, const float *const *x, const long long *stride, const int *S, const int *W,
                             const int *N, const WFamily &f, char *const *ws, hipStream_t st)
{
    if (n < 1 || n > kMaxBatchJobs) return -1;
    const BigPlan &bp = big_plan((int)f.Nz);
    if (!bp.ok) return -1;
    FilterBank fb;
    int r = filter_bank(f, st, fb);
    if (r) return r;
    const int nsc = (int)f.S;
    const long long Nz = f.Nz;
    // job tables: station-windows (input passes), station-window scales (the rest)
    RealBatch jw{}, js{};
    jw.n = js.n = n;
    jw.b0[0] = js.b0[0] = 0;
    for (int g = 0; g < n; ++g) {
        const WpccLayout l = wpcc_layout(S[g], W[g], (int)Nz, nsc);
        const int SW = S[g] * W[g];
        jw.x[g] = x[g];
        jw.stride[g] = stride[g];
        jw.W[g] = W[g];
        jw.N[g] = N[g];
        jw.X[g] = ws[g] + l.X;
        jw.b0[g + 1] = jw.b0[g] + SW;
        js.X[g] = ws[g] + l.B;
        js.Xin[g] = ws[g] + l.X;
        js.mx[g] = (float *)(ws[g] + l.mx);
        js.b0[g + 1] = js.b0[g] + SW * nsc;
        if ((r = fill_zero(js.mx[g], (long long)SW * nsc, st))) return r;
    }
    if (js.b0[n] > 65535) return -1;
    ColArgs ca{};
    ca.op = COL_FWD_INPUT;  // X = FFT(in), in[Ls0 + n] = x[n] / Nz
    ca.bstride = Nz;
    ca.off = (int)f.Ls0;
    ca.in_scale = (float)(1.0 / (double)Nz);
    if ((r = run_cols_batch(bp, ca, jw, st, "wpcc_input_cols_batch"))) return r;
    RowArgs ra{};
    ra.op = ROW_FWD_PLAIN;
    ra.bstride = Nz;
    ra.S = 1;
    if ((r = run_rows_batch(bp, ra, jw, st, "wpcc_input_rows_batch"))) return r;
    ra = RowArgs{};
    ra.op = ROW_INV_MUL;  // IFFT(X W_s), then max|z|^2, AmpNorm, FFT, * sqrt(1/(scale Nz))
    ra.bstride = Nz;
    ra.fw = fb.fw;
    ra.S = nsc;
    if ((r = run_rows_batch(bp, ra, js, st, "wpcc_scale_inv_rows_batch"))) return r;
    ca = ColArgs{};
    ca.op = COL_INV_MAX;
    ca.bstride = Nz;
    if ((r = run_cols_batch(bp, ca, js, st, "wpcc_scale_inv_cols_batch"))) return r;
    ca.op = COL_FWD_AMPNORM;
    if ((r = run_cols_batch(bp, ca, js, st, "wpcc_ampnorm_cols_batch"))) return r;
    ra = RowArgs{};
    ra.op = ROW_FWD_SCALE;
    ra.bstride = Nz;
    ra.S = nsc;
    ra.wscale = fb.wscale;
    return run_rows_batch(bp, ra, js, st, "wpcc_fwd_rows_batch");
}

int wpcc_pair_stage(int S, int W, int N, const WFamily &f, int lag0, int L, bool per_window, char *ws, float *d_y,
                    hipStream_t st)
{
    const BigPlan &bp = big_plan((int)f.Nz);
    if (!bp.ok) return -1;
    const int nsc = (int)f.S, Nz = (int)f.Nz, P = S * (S - 1) / 2;
    WpccLayout l = wpcc_layout(S, W, Nz, nsc);
    cf32 *B = (cf32 *)(ws + l.B), *U = (cf32 *)(ws + l.U), *G = (cf32 *)(ws + l.G);
    int ntiles = 0;
    const int2 *tiles = gram_tiles(S, ntiles);
    if (!tiles) return -1;
    const long long ldg = gram_ld(Nz);
    // rows (station, window, scale) of Nz bins -> U[k][station][window*S + scale]
    int r = launch_spectra_to_u(B, Nz, S, W * nsc, Nz, Nz, 0, U, st);
    if (r) return r;
    const double C = 1. / (f.K0 * (double)N);
    const int nrounds = per_window ? W : 1;
    for (int w = 0; w < nrounds; ++w) {
        const int m0 = per_window ? w * nsc : 0, K = per_window ? nsc : W * nsc;
        if ((r = launch_gram_f32(U, U, S, W * nsc, m0, K, Nz, tiles, ntiles, G, gram_rows(ldg), st))) return r;
        RowArgs ra{};
        ra.op = ROW_INV_PLAIN;
        ra.src = G;
        ra.dst = G;
        ra.bstride = ldg;
        ra.S = 1;
        if ((r = run_rows(bp, ra, P, st, "wpcc_pair_inv_rows"))) return r;
        ColArgs ca{};
        ca.op = COL_INV_LAGWIN;
        ca.src = G;
        ca.bstride = ldg;
        ca.y = d_y;
        ca.lag0 = lag0;
        ca.L = L;
        ca.S = S;
        ca.Wb = 1;
        ca.wsel = per_window ? w : 0;
        ca.Wfull = per_window ? W : 1;
        ca.scale = (float)(per_window ? C : C / W);
        if ((r = run_cols(bp, ca, P, st, "wpcc_pair_inv_cols"))) return r;
    }
    return 0;
}

}  // namespace fpcc
