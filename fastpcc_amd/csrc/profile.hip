// Optional per-kernel timing with HIP events, for benchmarks.
//
// While enabled, each launch made through prof_begin()/prof_end() records an event
// pair on the launch's stream; fpcc_profile_report() waits for the events and
// returns, per kernel name, the summed elapsed time and the launch count.
#include <map>
#include <mutex>
#include <string>
#include <string.h>
#include <vector>
#include "engine.hpp"
#include "profile.hpp"

namespace fpcc {
namespace {
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
std::mutex mu;
bool enabled = false;
std::vector<Pending> pending;
std::vector<hipEvent_t> pool;
std::map<std::string, std::pair<double, long>> totals;

hipEvent_t get_event()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool prof_on() { return enabled; }

ProfScope::ProfScope(const char *name, hipStream_t st) : name_(name), st_(st), a_(nullptr)
{
    if (!enabled) return;
    std::lock_guard<std::mutex> lk(mu);
    a_ = get_event();
    if (a_) (void)hipEventRecord((hipEvent_t)a_, st_);
}

ProfScope::~ProfScope()
{
    if (!a_) return;
    std::lock_guard<std::mutex> lk(mu);
    hipEvent_t b = get_event();
    if (!b) return;
    (void)hipEventRecord(b, st_);
    pending.push_back(Pending{name_, (hipEvent_t)a_, b});
}

static void drain()
{
    for (auto &p : pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &t = totals[p.name];
            t.first += ms;
            t.second += 1;
        }
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
}

}  // namespace fpcc

using namespace fpcc;

extern "C" void fpcc_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(mu);
    enabled = on != 0;
    // events made up front: a timed run then only records them (creating events
    // between launches delays the launches)
    while (enabled && pool.size() < 2048) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) break;
        pool.push_back(e);
    }
}

extern "C" void fpcc_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(mu);
    drain();
    totals.clear();
}

extern "C" int fpcc_profile_report(char *names, size_t names_len, double *total_ms, long long *counts,
                                   int max_kernels)
{
    std::lock_guard<std::mutex> lk(mu);
    drain();
    int k = 0;
    size_t off = 0;
    if (names && names_len) names[0] = 0;
    for (auto &kv : totals) {
        if (k >= max_kernels) break;
        if (total_ms) total_ms[k] = kv.second.first;
        if (counts) counts[k] = kv.second.second;
        if (names) {
            size_t len = kv.first.size();
            if (off + len + 2 > names_len) break;
            memcpy(names + off, kv.first.c_str(), len);
            off += len;
            names[off++] = '\n';
            names[off] = 0;
        }
        ++k;
    }
    return k;
}
