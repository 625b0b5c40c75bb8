// Host side of the FFT core: radix factorisation and twiddle tables.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>
#include "engine.hpp"

namespace fpcc {

unsigned nz_rule(unsigned N, unsigned M)
{
    return 1u << (unsigned)std::ceil(std::log2((double)(N + M)));
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

const DevPlan &get_plan(int n, bool f64)
{
    static std::map<std::tuple<int, int, bool>, DevPlan> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, n, f64);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    DevPlan dp{};
    dp.ok = factor_plan(n, dp.plan);
    if (dp.ok) {
        std::vector<double> t;
        int Ns = 1;
        for (int p = 0; p < dp.plan.npass; ++p) {
            const int R = dp.plan.radix[p];
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    double c, s;
                    tw_fwd((long)r * k, (long)Ns * R, c, s);
                    t.push_back(c);
                    t.push_back(s);
                }
            Ns *= R;
        }
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

const void *get_full_twiddles(int n, bool f64)
{
    static std::map<std::tuple<int, int, bool>, const void *> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_tuple(dev, n, f64);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
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

}  // namespace fpcc
