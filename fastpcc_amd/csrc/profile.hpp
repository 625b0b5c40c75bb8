// Per-kernel HIP-event timing used by benchmarks (see profile.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fpcc {
bool prof_on();
// Records an event on `st` at construction and at destruction when profiling is
// enabled; the elapsed time is attributed to `name`.
class ProfScope {
public:
    ProfScope(const char *name, hipStream_t st);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

private:
    const char *name_;
    hipStream_t st_;
    void *a_;
};
}  // namespace fpcc
