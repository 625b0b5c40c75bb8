// WPCC filter bank (host side).
#pragma once
#include <complex>
#include <memory>
#include <vector>

namespace fpcc {

struct WFamily {
    int type = 0;
    unsigned S = 0, J = 0, V = 0, Nz = 0, Ls0 = 0;
    double s0 = 0, K0 = 0, op1 = 0;
    std::vector<double> scale;
    std::vector<unsigned> Ls;
    std::vector<int> center;
    std::vector<std::vector<std::complex<double>>> wc;
};

// Returns 0, -3 (|lag| >= N or unsupported type) or -1 (invalid periods/voices).
int build_wfamily(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V, int type, double op1,
                  WFamily &f);
// Cached build_wfamily (families are immutable once built; the cache keeps the
// most recent kWFamilyCache and a caller's reference outlives eviction); nullptr on
// error with *err set.
constexpr size_t kWFamilyCache = 64;
std::shared_ptr<const WFamily> get_wfamily(unsigned N, int Lag1, int Lag2, double pmin, double pmax, unsigned V,
                                           int type, double op1, int *err);

}  // namespace fpcc
