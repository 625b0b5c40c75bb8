// WPCC filter bank (host, FP64): the parameter setup of tspcc2_set
// (src/FFTapps.c:1076-1115) and CreateWaveletFamily with its helpers
// (src/FWTa/wavelet_def_v7.c:21-131, 204-339), computed for the GPU path.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include "wavelet.hpp"

namespace fpcc {

namespace {
constexpr double kPiW = 3.14159265358979323846;
constexpr int kNsigmas = 5;  // src/FWTa/wavelet_v7.h:19

// erfi by its Maclaurin series, 500 terms (wavelet_def_v7.c:104-117)
double erfi_series(double z)
{
    double term = z, sum = z;
    const double zz = z * z;
    for (unsigned n = 1; n < 500; ++n) {
        term *= zz / n;
        sum += term / (2 * n + 1);
    }
    return sum * 2 / std::sqrt(kPiW);
}

// Samples u = -Ls/2 .. Ls-Ls/2-1 at t = u/scale of the complex mother wavelet:
//  -1 Morlet      (pi^-1/4 / sqrt(s)) e^{i w0 t} e^{-t^2/2}
//  -2 complete    (pi^-1/4 / sqrt(s)) (e^{i w0 t} - e^{-w0^2/2}) e^{-t^2/2}
//  -3 Mexican hat (2/sqrt(3 sqrt(pi) s)) ((t^2-1) e^{-t^2/2} (1 + i erfi(t/sqrt2)) - i sqrt(2/pi) t)
void sample_wavelet(int type, double w0, unsigned Ls, double scale, std::vector<std::complex<double>> &w)
{
    w.resize(Ls);
    const int u0 = (int)Ls / 2;
    const double inv = 1 / scale;
    if (type == -3) {
        const double k = 2 / std::sqrt(3 * std::sqrt(kPiW) * scale);
        for (int u = -u0, i = 0; u < (int)Ls - u0; ++u, ++i) {
            const double t = inv * u;
            const double g = (t * t - 1) * std::exp(-t * t / 2);
            w[i] = k * (std::complex<double>(g, g * erfi_series(t / std::sqrt(2.0))) -
                        std::complex<double>(0, std::sqrt(2 / kPiW) * t));
        }
        return;
    }
    const double k = 1 / std::sqrt(std::sqrt(kPiW) * scale);
    const double ze = type == -2 ? std::exp((-w0 * w0) / 2) : 0.0;
    for (int u = -u0, i = 0; u < (int)Ls - u0; ++u, ++i) {
        const double t = inv * u;
        const double env = k * std::exp(-0.5 * t * t);
        std::complex<double> osc = std::exp(std::complex<double>(0, w0 * t));
        if (type == -2) osc -= ze;
        w[i] = env * osc;
    }
}
}  // namespace

int build_wfamily(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V, int type, double op1,
                  WFamily &f)
{
    const unsigned ua1 = (unsigned)std::abs(Lag1), ua2 = (unsigned)std::abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    if (V == 0 || !(pmin > 0) || !(pmax > 0)) return -1;
    const unsigned M = ua1 > ua2 ? ua1 : ua2;
    unsigned Nz = 1u << (unsigned)std::ceil(std::log2((double)(N + M)));
    double s0;
    if (type == -3) {
        s0 = pmin / (std::sqrt(2.0) * kPiW);
    } else if (type == -1 || type == -2) {
        if (op1 == 0) op1 = kPiW * std::sqrt(2 / std::log(2.0));
        s0 = pmin * op1 / (2 * kPiW);
    } else {
        return -3;
    }
    const unsigned J = (unsigned)std::round(1. / (double)V + std::log(pmax / pmin) / std::log(2.0));
    if (J == 0) return -1;
    f.type = type;
    f.J = J;
    f.V = V;
    f.S = J * V;
    f.s0 = s0;
    f.op1 = op1;
    f.scale.resize(f.S);
    f.Ls.resize(f.S);
    f.center.resize(f.S);
    f.wc.resize(f.S);
    // scales s0 * 2^(m/V) by repeated multiplication (setscales0, :299-309)
    double sc = s0;
    const double step = std::pow(2.0, 1 / (double)V);
    for (unsigned s = 0; s < f.S; ++s) {
        f.scale[s] = sc;
        sc *= step;
    }
    for (unsigned s = 0; s < f.S; ++s) {
        const unsigned len = 2 * (unsigned)std::ceil(kNsigmas * f.scale[s]) + 1;  // setwaveletlength0
        f.Ls[s] = len > Nz ? Nz : len;
        f.center[s] = (int)(f.Ls[s] / 2);
        sample_wavelet(type, op1, f.Ls[s], f.scale[s], f.wc[s]);
    }
    f.Ls0 = f.Ls[f.S - 1];
    if (Nz - (N + M) < f.Ls0) Nz *= 2;  // src/FFTapps.c:1112
    f.Nz = Nz;
    double K0 = 0;
    for (unsigned s = 0; s < f.S; ++s) K0 += 1 / f.scale[s];
    f.K0 = K0;
    return 0;
}

std::shared_ptr<const WFamily> get_wfamily(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                                           int type, double op1, int *err)
{
    typedef std::tuple<unsigned, int, int, double, double, unsigned, int, double> Key;
    static std::map<Key, std::pair<unsigned long long, std::shared_ptr<const WFamily>>> cache;
    static unsigned long long tick = 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const Key key(N, Lag1, Lag2, pmin, pmax, V, type, op1);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *err = 0;
        it->second.first = ++tick;
        return it->second.second;
    }
    auto f = std::make_shared<WFamily>();
    *err = build_wfamily(N, Lag1, Lag2, pmin, pmax, V, type, op1, *f);
    if (*err) return nullptr;
    if (cache.size() >= kWFamilyCache) {  // drop the least recently used
        auto old = cache.begin();
        for (auto c = cache.begin(); c != cache.end(); ++c)
            if (c->second.first < old->second.first) old = c;
        cache.erase(old);
    }
    cache[key] = {++tick, f};
    return f;
}

}  // namespace fpcc
