// Internal host-side interfaces of the fastpcc engine (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdio.h>
#include <string>
#include <vector>
#include "fft_core.hpp"

namespace fpcc {

// Returns a message on failure and keeps the first error of a call sequence.
#define FPCC_CHECK(expr)                                                                      \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            fprintf(stderr, "fastpcc: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                                      \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)

// Device-resident plan of a length-n complex FFT in precision T (float/double):
// the pass list and the twiddle table (FP64-generated, rounded once to T).
struct DevPlan {
    FftPlan plan;
    const void *tw;  // C2<T>*, on the current device
    bool ok;
};

// Plan for length n (cached per device and precision).  ok=false if n has a
// prime factor above 7 or needs more than kMaxPasses passes.
const DevPlan &get_plan(int n, bool f64);
// Table of exp(+2*pi*i*k/nz), k in [0, nz/2), for the half-length real-output
// inverse (cached per device and precision).
const void *get_c2r_twiddles(int nz, bool f64);
// Table of exp(-2*pi*i*k/n), k in [0, n) (cached per device and precision).
const void *get_full_twiddles(int n, bool f64);
// FP32 exp(-2*pi*i*m/n) for m in [0, n) in two factors: entries [0, kTwLo) hold
// W^i, entries kTwLo + j hold W^{kTwLo j}; W^m = T[kTwLo + m / kTwLo] T[m % kTwLo]
// (cached per device).
constexpr int kTwLo = 512;
const void *get_split_twiddles(int n);
// Per-pass twiddles of an FP32 plan with the given radix order (the compile-time
// FFT kernels fix their order; cached per device), nullptr if the order is invalid.
const void *get_order_twiddles(int n, const std::vector<int> &radices, bool f64 = false);
// Radix list for n (host); returns false if unsupported.
bool factor_plan(int n, FftPlan &p);
// Plan with the given pass radices (false unless valid for n).
bool plan_from_radices(int n, const std::vector<int> &radices, FftPlan &p);
// Path of a tuning table in tuning/ beside the library ("" if unknown).
std::string tuning_file(const char *name);

// Constant device tables (FFT plans and twiddles, Gram tile lists, filter banks)
// are built on first use, once per device, and kept.  Inside a NoBuildScope (on
// the constructing thread) a table that does not exist yet is not built: the
// lookup fails as an unsupported case would, and tables_missing() turns true.
struct NoBuildScope {
    NoBuildScope();
    ~NoBuildScope();
};
// True outside a NoBuildScope; inside one, false (and the miss is recorded).
bool table_builds_allowed();
bool tables_missing();

// Smallest power of two >= v
static inline unsigned pow2_ceil(unsigned v)
{
    unsigned r = 1;
    while (r < v) r <<= 1;
    return r;
}

// Nz rule of the reference (src/FFTapps.c:805)
unsigned nz_rule(unsigned N, unsigned M);

}  // namespace fpcc
