// Host side of the FFT core: radix factorisation, the radix-order tuning table
// and its autotuner, twiddle tables.
#include <dlfcn.h>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include "../../include/fastpcc.h"
#include "engine.hpp"

namespace fpcc {

unsigned nz_rule(unsigned N, unsigned M)
{
    return 1u << (unsigned)std::ceil(std::log2((double)(N + M)));
}

static const int kRadices[] = {2, 3, 4, 5, 7, 8, 9, 16, 25};

// Plan with the given pass radices (first pass first); false unless their product
// is n, every radix is one the LDS passes implement and they fit kMaxPasses.
bool plan_from_radices(int n, const std::vector<int> &r, FftPlan &p)
{
    p = FftPlan{};
    p.n = n;
    if (r.empty() || (int)r.size() > kMaxPasses) return false;
    long long prod = 1;
    for (int x : r) {
        if (std::find(std::begin(kRadices), std::end(kRadices), x) == std::end(kRadices)) return false;
        prod *= x;
    }
    if (prod != n) return false;
    p.npass = (int)r.size();
    int off = 0, Ns = 1;
    for (int i = 0; i < p.npass; ++i) {
        p.radix[i] = r[i];
        p.tw_off[i] = off;
        off += (r[i] - 1) * Ns;
        Ns *= r[i];
    }
    return true;
}

std::string tuning_file(const char *name)
{
    Dl_info info;
    if (dladdr((const void *)&plan_from_radices, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning/" + name;
    }
    return std::string();
}

// Lengths whose radices compile-time kernels rely on (kept at the default order).
static bool fixed_radix_length(int n) { return n == 3600 || n == 4096 || n == 8192; }

// Radix-order table of the FP32 LDS FFTs: lines "n r1 r2 ... [# comment]", read
// once from $FASTPCC_RADIX_TUNING or tuning/fft_radix_gfx950.txt beside the library.
static std::mutex radix_mu;
static std::map<int, std::vector<int>> &radix_table()
{
    static std::map<int, std::vector<int>> t;
    static bool loaded = false;
    if (loaded) return t;
    loaded = true;
    const char *e = getenv("FASTPCC_RADIX_TUNING");
    const std::string path = e ? std::string(e) : tuning_file("fft_radix_gfx950.txt");
    if (FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r")) {
        char line[512];
        while (fgets(line, sizeof line, f)) {
            if (char *h = strchr(line, '#')) *h = '\0';
            std::vector<int> v;
            char *c = line;
            for (;;) {
                char *end = nullptr;
                const long x = strtol(c, &end, 10);
                if (end == c) break;
                v.push_back((int)x);
                c = end;
            }
            if (v.size() < 2 || fixed_radix_length(v[0])) continue;
            FftPlan chk;
            std::vector<int> r(v.begin() + 1, v.end());
            if (plan_from_radices(v[0], r, chk)) t[v[0]] = r;
        }
        fclose(f);
    }
    return t;
}

bool factor_plan(int n, FftPlan &p)
{
    p = FftPlan{};
    p.n = n;
    if (n < 1) return false;
    int m = n, e2 = 0, e3 = 0, e5 = 0, e7 = 0;
    while (m % 2 == 0) { m /= 2; ++e2; }
    while (m % 3 == 0) { m /= 3; ++e3; }
    while (m % 5 == 0) { m /= 5; ++e5; }
    while (m % 7 == 0) { m /= 7; ++e7; }
    if (m != 1) return false;
    std::vector<int> r;
    while (e2 >= 4) { r.push_back(16); e2 -= 4; }
    if (e2 == 3) r.push_back(8);
    if (e2 == 2) r.push_back(4);
    if (e2 == 1) r.push_back(2);
    while (e3 >= 2) { r.push_back(9); e3 -= 2; }
    if (e3 == 1) r.push_back(3);
    while (e5 >= 2) { r.push_back(25); e5 -= 2; }
    if (e5 == 1) r.push_back(5);
    while (e7 >= 1) { r.push_back(7); --e7; }
    if ((int)r.size() > kMaxPasses) return false;
    p.npass = (int)r.size();
    int off = 0, Ns = 1;
    for (int i = 0; i < p.npass; ++i) {
        p.radix[i] = r[i];
        p.tw_off[i] = off;
        off += (r[i] - 1) * Ns;
        Ns *= r[i];
    }
    return true;
}

// W_Q^j = exp(-2*pi*i*j/Q), in double with exact integer reduction
static void tw_fwd(long j, long Q, double &c, double &s)
{
    long m = j % Q;
    double a = -2.0 * M_PI * (double)m / (double)Q;
    c = std::cos(a);
    s = std::sin(a);
}

template <typename T>
static const void *upload_table(const std::vector<double> &re_im)
{
    size_t cnt = re_im.size() / 2;
    std::vector<C2<T>> h(cnt ? cnt : 1);
    for (size_t i = 0; i < cnt; ++i) h[i] = C2<T>{(T)re_im[2 * i], (T)re_im[2 * i + 1]};
    void *d = nullptr;
    if (hipMalloc(&d, sizeof(C2<T>) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(C2<T>) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

static std::mutex g_mu;

namespace {
thread_local int t_nobuild = 0;
thread_local bool t_missing = false;
}  // namespace

NoBuildScope::NoBuildScope()
{
    if (t_nobuild++ == 0) t_missing = false;
}
NoBuildScope::~NoBuildScope() { --t_nobuild; }

bool table_builds_allowed()
{
    if (!t_nobuild) return true;
    t_missing = true;
    return false;
}

bool tables_missing() { return t_missing; }

// Per-pass twiddles W_{Ns R}^{r k}, r < R, k < Ns, in pass order.
static std::vector<double> plan_twiddles(const FftPlan &plan)
{
    std::vector<double> t;
    int Ns = 1;
    for (int p = 0; p < plan.npass; ++p) {
        const int R = plan.radix[p];
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < Ns; ++k) {
                double c, s;
                tw_fwd((long)r * k, (long)Ns * R, c, s);
                t.push_back(c);
                t.push_back(s);
            }
        Ns *= R;
    }
    return t;
}

const DevPlan &get_plan(int n, bool f64)
{
    static std::map<std::tuple<int, int, bool>, DevPlan> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, n, f64);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    static const DevPlan none{};
    if (!table_builds_allowed()) return none;
    DevPlan dp{};
    dp.ok = factor_plan(n, dp.plan);
    if (dp.ok && !f64) {
        std::lock_guard<std::mutex> lr(radix_mu);
        auto &t = radix_table();
        auto r = t.find(n);
        if (r != t.end()) dp.ok = plan_from_radices(n, r->second, dp.plan);
    }
    if (dp.ok) {
        const std::vector<double> t = plan_twiddles(dp.plan);
        dp.tw = f64 ? upload_table<double>(t) : upload_table<float>(t);
        dp.ok = dp.tw != nullptr;
    }
    return cache.emplace(key, dp).first->second;
}

const void *get_c2r_twiddles(int nz, bool f64)
{
    static std::map<std::tuple<int, int, bool>, const void *> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, nz, f64);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (!table_builds_allowed()) return nullptr;
    std::vector<double> t;
    for (int k = 0; k < nz / 2; ++k) {
        double c, s;
        tw_fwd(k, nz, c, s);
        t.push_back(c);
        t.push_back(-s); /* exp(+2*pi*i*k/nz) */
    }
    const void *d = f64 ? upload_table<double>(t) : upload_table<float>(t);
    cache.emplace(key, d);
    return d;
}

const void *get_order_twiddles(int n, const std::vector<int> &radices, bool f64)
{
    static std::map<std::pair<int, std::vector<int>>, const void *> cache;
    FftPlan p{};
    if (!plan_from_radices(n, radices, p)) return nullptr;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<int> key = radices;
    key.push_back(dev);
    key.push_back(f64 ? 1 : 0);
    auto it = cache.find({n, key});
    if (it != cache.end()) return it->second;
    if (!table_builds_allowed()) return nullptr;
    const void *d = f64 ? upload_table<double>(plan_twiddles(p)) : upload_table<float>(plan_twiddles(p));
    cache.emplace(std::make_pair(n, key), d);
    return d;
}

const void *get_split_twiddles(int n)
{
    static std::map<std::pair<int, int>, const void *> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = cache.find({dev, n});
    if (it != cache.end()) return it->second;
    if (!table_builds_allowed()) return nullptr;
    std::vector<double> t;
    const int hi = (n + kTwLo - 1) / kTwLo;
    t.reserve(2 * (size_t)(kTwLo + hi));
    for (int k = 0; k < kTwLo; ++k) {
        double c, s;
        tw_fwd(k, n, c, s);
        t.push_back(c);
        t.push_back(s);
    }
    for (int j = 0; j < hi; ++j) {
        double c, s;
        tw_fwd((long)kTwLo * j, n, c, s);
        t.push_back(c);
        t.push_back(s);
    }
    const void *d = upload_table<float>(t);
    cache.emplace(std::make_pair(dev, n), d);
    return d;
}

const void *get_full_twiddles(int n, bool f64)
{
    static std::map<std::tuple<int, int, bool>, const void *> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, n, f64);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (!table_builds_allowed()) return nullptr;
    std::vector<double> t;
    t.reserve(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double c, s;
        tw_fwd(k, n, c, s);
        t.push_back(c);
        t.push_back(s);
    }
    const void *d = f64 ? upload_table<double>(t) : upload_table<float>(t);
    cache.emplace(key, d);
    return d;
}


// ---- radix-order autotuner ----
namespace {

// One forward FFT of length n per workgroup, rows of data in place.
__global__ void __launch_bounds__(256) k_fft_rows(cf32 *__restrict__ data, FftPlan plan, const cf32 *__restrict__ tw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cf32 *buf0 = reinterpret_cast<cf32 *>(smem_raw), *buf1 = buf0 + plan.n;
    cf32 *row = data + (long long)blockIdx.x * plan.n;
    for (int i = threadIdx.x; i < plan.n; i += blockDim.x) buf0[i] = row[i];
    __syncthreads();
    const cf32 *res = fft_lds<float, -1>(buf0, buf1, plan, tw);
    for (int i = threadIdx.x; i < plan.n; i += blockDim.x) row[i] = res[i];
}

// Every multiset of supported radices with product n (at most kMaxPasses parts).
void radix_multisets(int n, int min_idx, std::vector<int> &cur, std::vector<std::vector<int>> &out)
{
    if (n == 1) {
        if (!cur.empty()) out.push_back(cur);
        return;
    }
    if ((int)cur.size() >= kMaxPasses) return;
    const int nr = (int)(sizeof(kRadices) / sizeof(kRadices[0]));
    for (int i = min_idx; i < nr; ++i) {
        if (n % kRadices[i]) continue;
        cur.push_back(kRadices[i]);
        radix_multisets(n / kRadices[i], i, cur, out);
        cur.pop_back();
    }
}

}  // namespace

int tune_radices(int n, int batch, int reps, std::vector<int> &best, double &best_ms, double &default_ms)
{
    FftPlan def;
    if (n < 2 || n > 8192 || batch < 1 || reps < 1 || fixed_radix_length(n) || !factor_plan(n, def)) return -1;
    std::vector<std::vector<int>> sets, cand;
    std::vector<int> cur;
    radix_multisets(n, 0, cur, sets);
    cand.push_back(std::vector<int>(def.radix, def.radix + def.npass));
    for (auto &m : sets) {
        std::vector<int> up = m, down = m;
        std::sort(up.begin(), up.end());
        std::sort(down.rbegin(), down.rend());
        for (auto *c : {&down, &up})
            if (std::find(cand.begin(), cand.end(), *c) == cand.end()) cand.push_back(*c);
    }
    void *buf = nullptr, *tw = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    best_ms = 1e300;
    default_ms = 0;
    const size_t bytes = sizeof(cf32) * (size_t)n * batch, lds = sizeof(cf32) * 2 * (size_t)n;
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)k_fft_rows, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
    if (hipMalloc(&buf, bytes) == hipSuccess && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        hipMemsetAsync(buf, 0, bytes, st) == hipSuccess) {
        rc = 0;
        for (size_t c = 0; c < cand.size(); ++c) {
            FftPlan p;
            if (!plan_from_radices(n, cand[c], p)) continue;
            const std::vector<double> t = plan_twiddles(p);
            std::vector<cf32> h(t.size() / 2);
            for (size_t i = 0; i < h.size(); ++i) h[i] = cf32{(float)t[2 * i], (float)t[2 * i + 1]};
            if (tw) (void)hipFree(tw);
            tw = nullptr;
            if (hipMalloc(&tw, sizeof(cf32) * std::max<size_t>(h.size(), 1)) != hipSuccess ||
                hipMemcpy(tw, h.data(), sizeof(cf32) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
                rc = -1;
                break;
            }
            hipLaunchKernelGGL(k_fft_rows, dim3(batch), dim3(256), lds, st, (cf32 *)buf, p, (const cf32 *)tw);
            (void)hipEventRecord(e0, st);
            for (int r = 0; r < reps; ++r)
                hipLaunchKernelGGL(k_fft_rows, dim3(batch), dim3(256), lds, st, (cf32 *)buf, p, (const cf32 *)tw);
            (void)hipEventRecord(e1, st);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = -1;
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = ms / reps;
            if (c == 0) default_ms = per;
            if (per < best_ms) {
                best_ms = per;
                best = cand[c];
            }
        }
    }
    if (tw) (void)hipFree(tw);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (buf) (void)hipFree(buf);
    if (rc || best.empty()) return -1;
    std::lock_guard<std::mutex> lr(radix_mu);
    radix_table()[n] = best;  // used by plans first built later in the process
    return 0;
}

}  // namespace fpcc

extern "C" int fpcc_autotune_radix(int n, int batch, int reps, int *radices, int *npass, double *best_ms,
                                   double *default_ms)
{
    std::vector<int> best;
    double b = 0, d = 0;
    const int r = fpcc::tune_radices(n, batch, reps, best, b, d);
    if (r) return r;
    if (npass) *npass = (int)best.size();
    if (radices)
        for (size_t i = 0; i < best.size(); ++i) radices[i] = best[i];
    if (best_ms) *best_ms = b;
    if (default_ms) *default_ms = d;
    return 0;
}

extern "C" int fpcc_fft_radices(int n, int *radices)
{
    const fpcc::DevPlan &p = fpcc::get_plan(n, false);
    if (!p.ok) return 0;
    if (radices)
        for (int i = 0; i < p.plan.npass; ++i) radices[i] = p.plan.radix[i];
    return p.plan.npass;
}
