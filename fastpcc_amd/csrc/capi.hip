// C ABI of libfastpcc (declared in include/fastpcc.h).
#include <errno.h>
#include <math.h>
#include <semaphore.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <memory>
#include <atomic>
#include <condition_variable>
#include <sched.h>
#include <mutex>
#include <thread>
#include <vector>
#include "../../include/fastpcc.h"
#include "bigfft.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "wavelet.hpp"

using namespace fpcc;

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct DeviceBuf {
    void *p = nullptr;
    ~DeviceBuf()
    {
        if (p) (void)hipFree(p);
    }
};
struct HostBuf {
    void *p = nullptr;
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

// Buffers of the host entry points (the reference-style *_set calls and
// fpcc_allpairs): pinned input and output staging, device input, output and
// workspace.  Each call takes one buffer set from its device's pool (a set no other
// call holds, or a new one), so concurrent callers never wait for each other.  A
// set's buffers are kept between calls, grown to the largest call that used the
// set, as long as the device's kept total stays within the limits (default 1 GiB
// pinned, 8 GiB device; fpcc_set_kept_limits); a call whose buffers would pass
// them allocates its own and frees them on return.  fpcc_release_buffers frees
// every kept buffer.
std::atomic<size_t> g_keep_host{(size_t)1 << 30}, g_keep_device{(size_t)8 << 30};

struct BufSet {
    void *p[5] = {};  // hx, hy (pinned host), dx, dy, dws (device)
    size_t n[5] = {};
    unsigned gen = 0;
};
constexpr bool kSetHost[5] = {true, true, false, false, false};

void free_set(BufSet &s)
{
    for (int i = 0; i < 5; ++i) {
        if (s.p[i]) (void)(kSetHost[i] ? hipHostFree(s.p[i]) : hipFree(s.p[i]));
        s.p[i] = nullptr;
        s.n[i] = 0;
    }
}

struct DevicePool {
    std::mutex mu;
    std::vector<std::unique_ptr<BufSet>> idle;
    size_t kept_h = 0, kept_d = 0;  // bytes held by every set of this device, idle or in use
    unsigned gen = 0;               // bumped by fpcc_release_buffers: sets of older generations are freed
    int busy = 0;
};

std::mutex g_pools_mu;
std::map<int, std::unique_ptr<DevicePool>> g_pools;

DevicePool &pool_for_device(int dev)
{
    std::lock_guard<std::mutex> lk(g_pools_mu);
    auto &p = g_pools[dev];
    if (!p) p.reset(new DevicePool);
    return *p;
}

class EntryBufs {
public:
    void *hx = nullptr, *hy = nullptr, *dx = nullptr, *dy = nullptr, *dws = nullptr;

    ~EntryBufs()
    {
        for (int i = 0; i < 5; ++i)
            if (own_[i]) (void)(kSetHost[i] ? hipHostFree(own_[i]) : hipFree(own_[i]));
        if (!set_) return;
        std::lock_guard<std::mutex> lk(pool_->mu);
        --pool_->busy;
        if (set_->gen != pool_->gen) {  // released while this call held it
            for (int i = 0; i < 5; ++i) (kSetHost[i] ? pool_->kept_h : pool_->kept_d) -= set_->n[i];
            free_set(*set_);
            return;
        }
        pool_->idle.push_back(std::move(set_));
    }

    // sizes in bytes (0: not needed); false if an allocation fails
    bool get(size_t bhx, size_t bhy, size_t bdx, size_t bdy, size_t bdws)
    {
        const size_t need[5] = {bhx, bhy, bdx, bdy, bdws};
        int dev = 0;
        (void)hipGetDevice(&dev);
        pool_ = &pool_for_device(dev);
        {
            std::lock_guard<std::mutex> lk(pool_->mu);
            // the idle set that needs the least growth
            size_t best = ~(size_t)0;
            int bi = -1;
            for (int i = 0; i < (int)pool_->idle.size(); ++i) {
                size_t grow = 0;
                for (int k = 0; k < 5; ++k) grow += need[k] > pool_->idle[i]->n[k] ? need[k] : 0;
                if (grow < best) best = grow, bi = i;
            }
            if (bi >= 0) {
                set_ = std::move(pool_->idle[bi]);
                pool_->idle.erase(pool_->idle.begin() + bi);
            } else {
                set_.reset(new BufSet);
                set_->gen = pool_->gen;
            }
            ++pool_->busy;
            // the growth this set may take, within the device's limits (reserved now,
            // corrected to what was allocated below)
            for (int k = 0; k < 5; ++k)
                if (need[k] > set_->n[k]) (kSetHost[k] ? res_h_ : res_d_) += need[k] - set_->n[k];
            keep_h_ = pool_->kept_h + res_h_ <= g_keep_host.load();
            keep_d_ = pool_->kept_d + res_d_ <= g_keep_device.load();
            if (!keep_h_) res_h_ = 0;
            if (!keep_d_) res_d_ = 0;
            pool_->kept_h += res_h_;
            pool_->kept_d += res_d_;
        }
        size_t before_h = 0, before_d = 0;
        for (int k = 0; k < 5; ++k) (kSetHost[k] ? before_h : before_d) += set_->n[k];
        const bool ok = fill(need);
        size_t after_h = 0, after_d = 0;
        for (int k = 0; k < 5; ++k) (kSetHost[k] ? after_h : after_d) += set_->n[k];
        std::lock_guard<std::mutex> lk(pool_->mu);
        pool_->kept_h = pool_->kept_h - res_h_ + after_h - before_h;
        pool_->kept_d = pool_->kept_d - res_d_ + after_d - before_d;
        return ok;
    }

private:
    static bool alloc(int k, size_t bytes, void *&p)
    {
        p = nullptr;
        const hipError_t e = kSetHost[k] ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return false;
        }
        return true;
    }
    // kept buffers grown in place (set_), the others allocated for this call (own_)
    bool fill(const size_t *need)
    {
        void **out[5] = {&hx, &hy, &dx, &dy, &dws};
        for (int k = 0; k < 5; ++k) {
            if (!need[k]) continue;
            if (kSetHost[k] ? keep_h_ : keep_d_) {
                if (need[k] > set_->n[k]) {
                    if (set_->p[k]) (void)(kSetHost[k] ? hipHostFree(set_->p[k]) : hipFree(set_->p[k]));
                    set_->p[k] = nullptr;
                    set_->n[k] = 0;
                    if (!alloc(k, need[k], set_->p[k])) return false;
                    set_->n[k] = need[k];
                }
                *out[k] = set_->p[k];
            } else {
                if (!alloc(k, need[k], own_[k])) return false;
                *out[k] = own_[k];
            }
        }
        return true;
    }
    DevicePool *pool_ = nullptr;
    std::unique_ptr<BufSet> set_;
    bool keep_h_ = false, keep_d_ = false;
    size_t res_h_ = 0, res_d_ = 0;
    void *own_[5] = {};
};

// Frees the idle kept sets of every device and marks the sets in use to be freed
// when their calls return.  Returns the number of sets still in use.
int release_kept()
{
    std::lock_guard<std::mutex> lk(g_pools_mu);
    int dev0 = 0, busy = 0;
    (void)hipGetDevice(&dev0);
    for (auto &kv : g_pools) {
        DevicePool &p = *kv.second;
        std::lock_guard<std::mutex> lp(p.mu);
        if (p.idle.empty() && !p.busy) continue;
        (void)hipSetDevice(kv.first);
        for (auto &s : p.idle) {
            for (int i = 0; i < 5; ++i) (kSetHost[i] ? p.kept_h : p.kept_d) -= s->n[i];
            free_set(*s);
        }
        p.idle.clear();
        ++p.gen;
        busy += p.busy;
    }
    (void)hipSetDevice(dev0);
    return busy;
}

// Host threads for the row gathers of the host entry points (pageable rows into
// pinned staging and back): one thread copies at a fraction of host memory
// bandwidth.  Inputs under two pieces stay on the calling thread.  The count is
// capped by the CPUs this process may run on (affinity mask, cgroup cpusets).
constexpr unsigned kGatherThreads = 8;
constexpr size_t kPieceBytes = (size_t)8 << 20;

static unsigned usable_cpus()
{
    static const unsigned n = [] {
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof(set), &set) == 0) return (unsigned)std::max(1, CPU_COUNT(&set));
        const unsigned hw = std::thread::hardware_concurrency();
        return hw ? hw : 1u;
    }();
    return n;
}

static unsigned gather_threads(size_t bytes)
{
    if (bytes < 2 * kPieceBytes) return 1;
    return std::max(1u, std::min(kGatherThreads, usable_cpus()));
}

// A counter that threads block on (no spinning) until it reaches a value.
class Gate {
public:
    void add(size_t d = 1)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            v_ += d;
        }
        cv_.notify_all();
    }
    void wait_for(size_t target)
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return v_ >= target; });
    }

private:
    std::mutex mu_;
    std::condition_variable cv_;
    size_t v_ = 0;
};

// fn(k) for k in [0, n) spread over up to kGatherThreads threads (the caller's
// included) when the rows hold at least 4 MB; returns once every call is done.
template <typename Fn>
static void rows_parallel(size_t n, size_t bytes, Fn fn)
{
    unsigned nt = bytes >= ((size_t)4 << 20) ? gather_threads(2 * kPieceBytes) : 1;
    if (nt > n) nt = (unsigned)std::max<size_t>(n, 1);
    std::vector<std::thread> pool;
    try {
        for (unsigned w = 1; w < nt; ++w)
            pool.emplace_back([&, w] {
                for (size_t k = w; k < n; k += nt) fn(k);
            });
    } catch (...) {  // the rows of threads that did not start are copied below
    }
    for (size_t k = 0; k < n; k += nt) fn(k);
    for (unsigned w = (unsigned)pool.size() + 1; w < nt; ++w)
        for (size_t k = w; k < n; k += nt) fn(k);
    for (auto &t : pool) t.join();
}

// `total` floats in rows of `row` (the last row may be shorter; src_of(k) is
// row k) gathered into pinned staging and copied to the device in pieces of
// about 8 MB on the null stream, so the gather of one piece overlaps the copy of
// the last; the rows of a piece are spread over the gather threads, which block
// (not spin) while they wait for a piece.  With `blocking_before`, before(k) may
// block (a semaphore wait): the rows are then copied on the calling thread alone,
// each one right after its before(k) returns, so no helper threads compete with
// the caller's producers for cores.  before(k) runs (on the calling thread) before
// row k < nbefore is read.
template <typename SrcOf, typename Before>
int upload_gather(size_t total, size_t row, SrcOf src_of, size_t nbefore, Before before, float *hx, float *dx,
                  bool blocking_before = false)
{
    if (!total) return 0;
    row = std::max<size_t>(row, 1);
    const size_t nrows = (total + row - 1) / row;
    const size_t prow = std::max<size_t>(1, kPieceBytes / (sizeof(float) * row));
    const size_t npieces = (nrows + prow - 1) / prow;
    auto copy_row = [&](size_t k) {
        memcpy(hx + k * row, src_of(k), sizeof(float) * std::min(row, total - k * row));
    };
    auto copy_piece_rows = [&](size_t pc, unsigned w, unsigned nw) {
        for (size_t k = pc * prow + w; k < std::min(nrows, (pc + 1) * prow); k += nw) copy_row(k);
    };
    auto send = [&](size_t pc) {
        const size_t f0 = pc * prow * row, f1 = std::min(total, (pc + 1) * prow * row);
        return hipMemcpyAsync(dx + f0, hx + f0, sizeof(float) * (f1 - f0), hipMemcpyHostToDevice, nullptr);
    };
    auto befores = [&](size_t pc) {
        for (size_t k = pc * prow; k < std::min(nbefore, (pc + 1) * prow); ++k) before(k);
    };
    const unsigned nt = blocking_before ? 1 : gather_threads(sizeof(float) * total);
    if (nt > 1) {
        // pieces are released in order once their before() calls have returned;
        // of nw copiers (the calling thread is copier 0), copier w takes rows
        // w, w + nw, ... of each released piece
        Gate released, finished;  // pieces released; (piece, helper) copies done
        std::atomic<unsigned> nw{0};
        std::vector<std::thread> pool;
        try {
            for (unsigned w = 1; w < nt; ++w)
                pool.emplace_back([&, w] {
                    for (size_t pc = 0; pc < npieces; ++pc) {
                        released.wait_for(pc + 1);
                        copy_piece_rows(pc, w, nw.load(std::memory_order_relaxed));
                        finished.add();
                    }
                });
        } catch (...) {  // fewer threads than asked for: those that started share the rows
        }
        const unsigned n = (unsigned)pool.size() + 1;
        nw.store(n, std::memory_order_relaxed);  // published to the helpers by the first release
        hipError_t err = hipSuccess;
        for (size_t pc = 0; pc < npieces; ++pc) {
            befores(pc);
            released.add();
            copy_piece_rows(pc, 0, n);
            // helpers take the pieces in order and piece pc + 1 is not released
            // yet, so (pc + 1)(n - 1) finished copies mean piece pc is complete
            finished.wait_for((pc + 1) * (n - 1));
            if (err == hipSuccess) err = send(pc);
        }
        for (auto &t : pool) t.join();
        if (err != hipSuccess) {
            fprintf(stderr, "fastpcc: %s\n", hipGetErrorString(err));
            return -1;
        }
        return 0;
    }
    for (size_t pc = 0; pc < npieces; ++pc) {
        for (size_t k = pc * prow; k < std::min(nrows, (pc + 1) * prow); ++k) {
            if (k < nbefore) before(k);
            copy_row(k);
        }
        FPCC_CHECK(send(pc));
    }
    return 0;
}

// Rows of two stations (x1 rows, then x2 rows, each of `row` floats), the
// reference's row-pointer inputs.  `before(tr)` runs before row tr of either
// station is read (it may block if blocking_before).  -1 on a null row.
template <typename Before>
int upload_rows(const float *const *x1, const float *const *x2, unsigned Tr, size_t row, float *hx, float *dx,
                Before before, bool blocking_before = false)
{
    for (unsigned tr = 0; tr < Tr; ++tr)
        if (!x1[tr] || !x2[tr]) return -1;
    // staging row k < Tr is x1[k], waited for by before(k); rows Tr + k (x2[k])
    // come after every before() has returned
    return upload_gather(
        2 * (size_t)Tr * row, row, [&](size_t k) { return k < Tr ? x1[k] : x2[k - Tr]; }, Tr,
        [&](size_t k) { before((unsigned)k); }, hx, dx, blocking_before);
}

// `total` floats from the device (dy) to the caller's pageable h_y through
// pinned staging (hy): the copies to staging go in pieces of about 8 MB on the
// null stream, and the gather threads move each piece on as soon as it lands.
static int download_scatter(size_t total, const float *dy, float *hy, float *h_y)
{
    if (!total) return 0;
    const size_t pf = kPieceBytes / sizeof(float), npieces = (total + pf - 1) / pf;
    const unsigned nt = gather_threads(sizeof(float) * total);
    std::vector<hipEvent_t> ev(npieces, nullptr);
    bool events = nt > 1;
    for (size_t pc = 0; events && pc < npieces; ++pc)
        events = hipEventCreateWithFlags(&ev[pc], hipEventDisableTiming) == hipSuccess;
    auto drop_events = [&] {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    if (!events) {
        drop_events();
        FPCC_CHECK(hipMemcpy(hy, dy, sizeof(float) * total, hipMemcpyDeviceToHost));
        memcpy(h_y, hy, sizeof(float) * total);
        return 0;
    }
    hipError_t err = hipSuccess;
    for (size_t pc = 0; pc < npieces && err == hipSuccess; ++pc) {
        const size_t f0 = pc * pf, f1 = std::min(total, f0 + pf);
        err = hipMemcpyAsync(hy + f0, dy + f0, sizeof(float) * (f1 - f0), hipMemcpyDeviceToHost, nullptr);
        if (err == hipSuccess) err = hipEventRecord(ev[pc], nullptr);
    }
    if (err != hipSuccess) {
        (void)hipDeviceSynchronize();
        drop_events();
        fprintf(stderr, "fastpcc: %s\n", hipGetErrorString(err));
        return -1;
    }
    // of nw copiers (the calling thread is copier 0), copier w moves the
    // quarter-megabyte blocks w, w + nw, ... of each piece that has landed
    constexpr size_t kBlock = (size_t)1 << 16;  // floats
    auto move_piece = [&](size_t pc, unsigned w, unsigned nw) {
        const size_t f0 = pc * pf, f1 = std::min(total, f0 + pf);
        for (size_t b0 = f0 + w * kBlock; b0 < f1; b0 += (size_t)nw * kBlock)
            memcpy(h_y + b0, hy + b0, sizeof(float) * std::min(kBlock, f1 - b0));
    };
    Gate landed;
    std::atomic<unsigned> nw{0};
    std::vector<std::thread> pool;
    try {
        for (unsigned w = 1; w < nt; ++w)
            pool.emplace_back([&, w] {
                for (size_t pc = 0; pc < npieces; ++pc) {
                    landed.wait_for(pc + 1);
                    move_piece(pc, w, nw.load(std::memory_order_relaxed));
                }
            });
    } catch (...) {  // fewer threads than asked for: those that started share the blocks
    }
    const unsigned n = (unsigned)pool.size() + 1;
    nw.store(n, std::memory_order_relaxed);  // published to the copiers by the first landing
    for (size_t pc = 0; pc < npieces; ++pc) {
        if (err == hipSuccess) err = hipEventSynchronize(ev[pc]);
        landed.add();  // on a failure the copiers still finish
        move_piece(pc, 0, n);
    }
    for (auto &t : pool) t.join();
    drop_events();
    if (err != hipSuccess) {
        fprintf(stderr, "fastpcc: %s\n", hipGetErrorString(err));
        return -1;
    }
    return 0;
}

// Resolved sizes of a job.
struct Dims {
    int S, W, N, L, lag0;
    unsigned M, Nz;
    size_t P;
    std::shared_ptr<const WFamily> fam;  // WPCC only
    bool large;          // windows beyond the single-workgroup kernels
};

// The single-workgroup kernels need Nz <= 8192 and, for the analytic signal, a
// 7-smooth N <= 8192; anything else goes through the two-pass FFTs (Bluestein for
// analytic signals of other lengths).
bool needs_large(int kind, int N, unsigned Nz)
{
    if (kind == FPCC_WPCC) return false;
    FftPlan p;
    const bool analytic = kind == FPCC_PCC2 || kind == FPCC_PCC1 || kind == FPCC_PCCV;
    if (analytic && (N > 8192 || !factor_plan(N, p))) return true;
    if (kind == FPCC_PCC1 || kind == FPCC_PCCV) return false;
    return Nz > 8192;
}

bool stack_supported(int kind, int stack)
{
    switch (kind) {
    case FPCC_PCC2: return stack >= FPCC_STACK_NONE && stack <= FPCC_STACK_PWS;
    case FPCC_GNCC: return stack == FPCC_STACK_NONE || stack == FPCC_STACK_PLAIN || stack == FPCC_STACK_PWS;
    case FPCC_CC1B:
        return stack == FPCC_STACK_NONE || stack == FPCC_STACK_PLAIN || stack == FPCC_STACK_GRAM ||
               stack == FPCC_STACK_GRAM16 || stack == FPCC_STACK_PWS;
    case FPCC_PCC1:
    case FPCC_PCCV: return stack == FPCC_STACK_NONE || stack == FPCC_STACK_PLAIN;
    case FPCC_WPCC: return stack == FPCC_STACK_NONE || stack == FPCC_STACK_PLAIN || stack == FPCC_STACK_GRAM;
    default: return false;
    }
}

// Lag window and Nz rule of pcc2_set/ccgn_set/cc1b_set (src/FFTapps.c:790-805).
int resolve(const fpcc_job *job, Dims &d)
{
    if (!job || job->n_stations < 2 || job->n_windows < 1 || job->win_len < 2) return -1;
    if (job->station_stride < (unsigned long long)job->n_windows * job->win_len) return -1;
    if (job->lag1 > job->lag2) return -1;
    const unsigned N = job->win_len;
    const unsigned ua1 = (unsigned)abs(job->lag1), ua2 = (unsigned)abs(job->lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    if (!stack_supported(job->kind, job->stack)) return -1;
    d.S = (int)job->n_stations;
    d.W = (int)job->n_windows;
    d.N = (int)N;
    d.L = job->lag2 - job->lag1 + 1;
    d.lag0 = job->lag1;
    d.M = ua1 > ua2 ? ua1 : ua2;
    d.Nz = nz_rule(N, d.M);
    d.P = (size_t)d.S * (d.S - 1) / 2;
    d.fam = nullptr;
    d.large = needs_large(job->kind, d.N, d.Nz);
    // The two-pass transforms start at 64 points.  A longer zero pad leaves the
    // correlations unchanged: the 1/(Nz N) and 1/Nz scales cancel the unnormalised
    // inverse FFT and Nz >= N + M keeps the lag window free of wrap-around.
    if (d.large && d.Nz < 64) d.Nz = 64;
    if (d.large && !large_supported(job->kind, d.N, (int)d.Nz)) {
        fprintf(stderr, "fastpcc: unsupported window length N=%d (Nz=%u)\n", d.N, d.Nz);
        return -1;
    }
    if (job->stack == FPCC_STACK_PWS && (d.large || !pws_supported((int)d.Nz) || !(job->pws_power >= 0))) {
        fprintf(stderr, "fastpcc: the phase-weighted stack needs a zero-pad length of at most 8192 (Nz=%u) and a "
                        "power >= 0\n", d.Nz);
        return -1;
    }
    if (job->kind == FPCC_WPCC) {
        int err = 0;
        d.fam = get_wfamily(N, job->lag1, job->lag2, job->pmin, job->pmax, job->V, job->wtype, job->op1, &err);
        if (!d.fam) return err ? err : -1;
        d.Nz = d.fam->Nz;
    }
    return 0;
}

// Workspace regions of a job (byte offsets, 256-B aligned).
struct Layout {
    size_t a = 0, b = 0, c = 0, d = 0, e = 0, f = 0, total = 0;
};

bool is_gram(int stack) { return stack == FPCC_STACK_GRAM || stack == FPCC_STACK_GRAM16; }

size_t take(size_t &total, size_t bytes)
{
    size_t off = total;
    total += align_up(bytes, 256);
    return off;
}

Layout layout(const fpcc_job *job, const Dims &d)
{
    Layout l;
    const size_t SW = (size_t)d.S * d.W, K1 = d.Nz / 2 + 1;
    if (d.large) {
        l.a = take(l.total, large_workspace(job->kind, job->stack, d.S, d.W, d.N, (int)d.Nz));
        return l;
    }
    switch (job->kind) {
    case FPCC_PCC2:
        l.a = take(l.total, sizeof(cf32) * SW * d.Nz);  // spectra
        if (is_gram(job->stack)) {
            l.b = take(l.total, sizeof(cf32) * K1 * SW * 2);                          // U
            l.c = take(l.total, sizeof(cf32) * gram_grouped_elems(d.P, (int)K1));  // G
        }
        break;
    case FPCC_GNCC:
        l.a = take(l.total, sizeof(cf64) * SW * (size_t)real_ld(d.Nz));    // half spectra
        l.b = take(l.total, sizeof(double) * SW * (size_t)energy_ld(d.N));  // energies
        break;
    case FPCC_CC1B:
        l.a = take(l.total, sizeof(cf32) * SW * (size_t)real_ld(d.Nz));  // x >= 0 signs
        l.b = take(l.total, sizeof(cf32) * SW * (size_t)real_ld(d.Nz));  // x > 0 signs
        l.f = take(l.total, sizeof(float) * SW * 2 * (size_t)edge_len(d.N));  // window edges (edge.hpp)
        if (is_gram(job->stack)) {
            l.c = take(l.total, sizeof(cf32) * K1 * SW);
            l.d = take(l.total, sizeof(cf32) * K1 * SW);
            l.e = take(l.total, sizeof(cf32) * gram_grouped_elems(d.P, (int)K1));
        }
        break;
    case FPCC_PCC1:
    case FPCC_PCCV:
        l.a = take(l.total, sizeof(cf32) * SW * (size_t)d.N);  // phasors
        break;
    case FPCC_WPCC:
        l.a = take(l.total, wpcc_workspace(d.S, d.W, d.N, (int)d.Nz, (int)d.fam->S));
        break;
    }
    return l;
}

}  // namespace

namespace fpcc {
int job_station_regions(const fpcc_job *job, size_t *off, size_t *per)
{
    Dims d;
    if (resolve(job, d)) return 0;
    const Layout l = layout(job, d);
    const size_t W = (size_t)d.W;
    if (d.large) {
        const int n = large_station_regions(job->kind, job->stack, d.S, d.W, d.N, (int)d.Nz, off, per);
        for (int k = 0; k < n; ++k) off[k] += l.a;
        return n;
    }
    switch (job->kind) {
    case FPCC_PCC2: off[0] = l.a; per[0] = sizeof(cf32) * W * d.Nz; return 1;
    case FPCC_GNCC:
        off[0] = l.a;
        per[0] = sizeof(cf64) * W * (size_t)real_ld(d.Nz);
        off[1] = l.b;
        per[1] = sizeof(double) * W * (size_t)energy_ld(d.N);
        return 2;
    case FPCC_CC1B:
        off[0] = l.a;
        off[1] = l.b;
        per[0] = per[1] = sizeof(cf32) * W * (size_t)real_ld(d.Nz);
        off[2] = l.f;
        per[2] = sizeof(float) * W * 2 * (size_t)edge_len(d.N);
        return 3;
    case FPCC_PCC1:
    case FPCC_PCCV: off[0] = l.a; per[0] = sizeof(cf32) * W * d.N; return 1;
    case FPCC_WPCC: {
        const int n = wpcc_station_regions(d.S, d.W, (int)d.Nz, (int)d.fam->S, off, per);
        for (int k = 0; k < n; ++k) off[k] += l.a;
        return n;
    }
    }
    return 0;
}
}  // namespace fpcc

namespace {

int station_stage(const fpcc_job *job, const Dims &d, const float *x, char *ws, hipStream_t st)
{
    Layout l = layout(job, d);
    const long long stride = (long long)job->station_stride;
    if (d.large) return large_station_stage(job->kind, x, stride, d.S, d.W, d.N, (int)d.Nz, job->stack, ws + l.a, st);
    switch (job->kind) {
    case FPCC_PCC2: return launch_station_pcc(x, stride, d.S, d.W, d.N, d.Nz, (cf32 *)(ws + l.a), st);
    case FPCC_GNCC:
        return launch_station_real(x, stride, d.S, d.W, d.N, d.Nz, 0, true, ws + l.a, (double *)(ws + l.b), st);
    case FPCC_CC1B: {
        int r = launch_station_real(x, stride, d.S, d.W, d.N, d.Nz, 1, false, ws + l.a, nullptr, st);
        if (r) return r;
        if ((r = launch_station_real(x, stride, d.S, d.W, d.N, d.Nz, 2, false, ws + l.b, nullptr, st))) return r;
        return large_edge_copy(x, stride, d.S, d.W, d.N, (float *)(ws + l.f), st);
    }
    case FPCC_PCC1:
    case FPCC_PCCV: return launch_station_phasor(x, stride, d.S, d.W, d.N, (cf32 *)(ws + l.a), st);
    case FPCC_WPCC: return wpcc_station_stage(x, stride, d.S, d.W, d.N, *d.fam, ws + l.a, st);
    }
    return -1;
}

int pair_stage(const fpcc_job *job, const Dims &d, char *ws, float *y, hipStream_t st)
{
    Layout l = layout(job, d);
    int r = 0;
    if (d.large) {
        const int stack = job->stack == FPCC_STACK_ZSUM ? FPCC_STACK_PLAIN : job->stack;
        return large_pair_stage(job->kind, stack, d.S, d.W, d.N, (int)d.Nz, d.lag0, d.L, job->v, ws + l.a, y, st);
    }
    const int mode = job->stack == FPCC_STACK_NONE ? 0 : 1;
    switch (job->kind) {
    case FPCC_PCC2:
        if (is_gram(job->stack)) {
            int ntiles = 0;
            const int2 *tiles = gram_tiles(d.S, ntiles);
            if (!tiles) return -1;
            cf32 *F = (cf32 *)(ws + l.a), *U = (cf32 *)(ws + l.b), *G = (cf32 *)(ws + l.c);
            const bool s16 = job->stack == FPCC_STACK_GRAM16;
            // |F| <= N: the spectra are sums of N unit-bounded phasors
            const float sigma = s16 ? gram16_sigma(d.N) : 0.f;
            const int nk = d.Nz / 2 + 1;
            if ((r = launch_spectra_to_u(F, d.Nz, d.S, d.W, d.Nz, nk, 1, U, st, sigma))) return r;
            const GLayout gl = gram_grouped(d.P);
            r = s16 ? launch_gram_split16(U, U, d.S, 2 * d.W, 0, 2 * d.W, nk, tiles, ntiles, sigma, G, gl, st)
                    : launch_gram_f32(U, U, d.S, 2 * d.W, 0, 2 * d.W, nk, tiles, ntiles, G, gl, st);
            if (r) return r;
            // 1/2 of the Gram's Hermitian fold, the reference's 1/((float)Nz*(float)N), the mean over W
            const float scale = 0.5f * (1 / ((float)d.Nz * (float)d.N)) / (float)d.W;
            return launch_pair_c2r(G, gl, (int)d.P, d.Nz, d.N, d.W, d.lag0, d.L, scale, 0, y, st);
        }
        if (job->stack == FPCC_STACK_PWS)  // the reference's 1/((float)Nz*(float)N) of each window
            return launch_pair_pws(0, ws + l.a, nullptr, d.Nz, d.S, d.W, (int)d.Nz, d.lag0, d.L,
                                   1 / ((float)d.Nz * (float)d.N), (float)job->pws_power, 0, y, st);
        return launch_pair_pcc2((const cf32 *)(ws + l.a), d.S, d.W, d.N, d.Nz, d.lag0, d.L, job->stack, 0,
                                (int)d.P, y, st);
    case FPCC_GNCC:
        r = launch_pair_gn(ws + l.a, ws + l.a, (const double *)(ws + l.b), true, d.S, d.W, d.N, d.Nz, d.lag0, d.L, 0,
                           mode, y, st);
        if (r || job->stack != FPCC_STACK_PWS) return r;
        return launch_pair_pws(2, ws + l.a, ws + l.a, real_ld(d.Nz), d.S, d.W, (int)d.Nz, d.lag0, d.L, 0.f,
                               (float)job->pws_power, 1, y, st);
    case FPCC_CC1B:
        if (is_gram(job->stack)) {
            int ntiles = 0;
            const int2 *tiles = gram_tiles(d.S, ntiles);
            if (!tiles) return -1;
            const long long ld = real_ld(d.Nz);
            cf32 *Ua = (cf32 *)(ws + l.c), *Ub = (cf32 *)(ws + l.d), *G = (cf32 *)(ws + l.e);
            const bool s16 = job->stack == FPCC_STACK_GRAM16;
            // |F| <= N: spectra of N signs
            const float sigma = s16 ? gram16_sigma(d.N) : 0.f;
            const int nk = d.Nz / 2 + 1;
            if ((r = launch_spectra_to_u((const cf32 *)(ws + l.a), ld, d.S, d.W, d.Nz, nk, 0, Ua, st, sigma)))
                return r;
            if ((r = launch_spectra_to_u((const cf32 *)(ws + l.b), ld, d.S, d.W, d.Nz, nk, 0, Ub, st, sigma)))
                return r;
            const GLayout gl = gram_grouped(d.P);
            r = s16 ? launch_gram_split16(Ua, Ub, d.S, d.W, 0, d.W, nk, tiles, ntiles, sigma, G, gl, st)
                    : launch_gram_f32(Ua, Ub, d.S, d.W, 0, d.W, nk, tiles, ntiles, G, gl, st);
            if (r) return r;
            // ccf_lowlevel's 1/Nz, the mean over W, then 1/(N-|lag|) (gnf_lowlevel with +-1 data)
            const float scale = (float)(1.0 / (double)d.Nz) / (float)d.W;
            return launch_pair_c2r(G, gl, (int)d.P, d.Nz, d.N, d.W, d.lag0, d.L, scale, 1, y, st,
                                   (const float *)(ws + l.f), d.S);
        }
        r = launch_pair_gn(ws + l.a, ws + l.b, nullptr, false, d.S, d.W, d.N, d.Nz, d.lag0, d.L, 1, mode, y, st,
                           (const float *)(ws + l.f));
        if (r || job->stack != FPCC_STACK_PWS) return r;
        return launch_pair_pws(1, ws + l.a, ws + l.b, real_ld(d.Nz), d.S, d.W, (int)d.Nz, d.lag0, d.L, 0.f,
                               (float)job->pws_power, 1, y, st);
    case FPCC_PCC1:
    case FPCC_PCCV:
        return launch_pcc_time((const cf32 *)(ws + l.a), d.S, d.W, d.N, d.lag0, d.L,
                               job->kind == FPCC_PCC1 ? 1.0 : job->v, job->kind == FPCC_PCC1, mode, y, d.L, 0, st);
    case FPCC_WPCC: return wpcc_pair_stage(d.S, d.W, d.N, *d.fam, d.lag0, d.L, mode == 0, ws + l.a, y, st);
    }
    return -1;
}

// WPCC of one pair over Tr windows with a given family (tspcc2_set and
// cuda_tspcc2_set_freq): x1 rows station 0, x2 rows station 1, L lags from lag0.
int wpcc_pair_host(const WFamily &f, float **y, float **x1, float **x2, unsigned N, unsigned Tr, int lag0, int L)
{
    if (Tr == 0 || L <= 0) return 0;
    if (!y || !x1 || !x2) return -1;
    const size_t xlen = (size_t)2 * Tr * N, ylen = (size_t)Tr * L;
    const size_t wsb = wpcc_workspace(2, (int)Tr, (int)N, (int)f.Nz, (int)f.S);
    for (unsigned tr = 0; tr < Tr; ++tr)
        if (!x1[tr] || !x2[tr] || !y[tr]) return -1;
    EntryBufs b;
    if (!b.get(sizeof(float) * xlen, sizeof(float) * ylen, sizeof(float) * xlen, sizeof(float) * ylen, wsb)) return -1;
    int r = upload_rows(x1, x2, Tr, N, (float *)b.hx, (float *)b.dx, [](unsigned) {});
    if (r) return r;
    r = wpcc_station_stage((const float *)b.dx, (long long)Tr * N, 2, (int)Tr, (int)N, f, (char *)b.dws, nullptr);
    if (r) return r;
    if ((r = wpcc_pair_stage(2, (int)Tr, (int)N, f, lag0, L, true, (char *)b.dws, (float *)b.dy, nullptr)))
        return r;
    FPCC_CHECK(hipMemcpy(b.hy, b.dy, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    rows_parallel(Tr, sizeof(float) * ylen,
                  [&](size_t tr) { memcpy(y[tr], (float *)b.hy + tr * L, sizeof(float) * L); });
    return 0;
}

// PCC1 / PCC-v for one pair over Tr windows (pcc1_set, src/FFTapps.c:611-783;
// pcc_set, :475-608).  Their lag handling differs from the FFT types: no -3, a
// lag range disjoint from [-N, N] returns 0 untouched, L < 0 returns -1, and only
// lags l in [l1, L') are computed (pcc1f_lowlevel, :251-278), the rest of each
// row being zero.  If complex_in != 0 the inputs are analytic signals (the
// *_highlevel entries): x1/x2 hold Tr complex rows and the phasor regulariser is
// applied on the GPU; `anok`, if given, is waited on per trace before reading it.
int pcc_time_host(int pcc1, float **y, const float *const *x1, const float *const *x2, int N, unsigned Tr,
                  double v, int Lag1, int Lag2, int complex_in, sem_t *anok, bool zero_rows)
{
    if (Lag1 > N || Lag2 < -N) return 0;
    const int L = Lag2 - Lag1 + 1;
    if (!x1 || !x2 || L < 0 || !y) return -1;
    if (Tr == 0 || L == 0 || N < 1) return 0;
    int Lp = L;
    if (Lag2 > N) Lp -= (Lag2 - N);
    const int l1 = (Lag1 >= -N) ? 0 : -(Lag1 + N);
    const size_t cw = complex_in ? 2 : 1;
    const size_t xlen = (size_t)2 * Tr * N * cw, ylen = (size_t)Tr * L;
    for (unsigned tr = 0; tr < Tr; ++tr)
        if (!x1[tr] || !x2[tr] || !y[tr]) return -1;
    FftPlan fpl;
    const bool small = N <= 8192 && factor_plan(N, fpl);
    if (!complex_in && !small && !large_supported(FPCC_PCC1, N, N)) {
        fprintf(stderr, "fastpcc: no analytic-signal plan for N=%d\n", N);
        return -1;
    }
    // phasors: in place (analytic input), rows of N complex values, or leading the
    // large-window workspace
    const size_t bphi = complex_in ? 0
                        : small    ? sizeof(cf32) * 2 * Tr * (size_t)N
                                   : large_workspace(FPCC_PCC1, FPCC_STACK_NONE, 2, (int)Tr, N, N);
    EntryBufs b;
    if (!b.get(sizeof(float) * xlen, sizeof(float) * ylen, sizeof(float) * xlen, sizeof(float) * ylen, bphi))
        return -1;
    int r = upload_rows(
        x1, x2, Tr, (size_t)N * cw, (float *)b.hx, (float *)b.dx,
        [&](unsigned tr) {
            if (anok)
                while (sem_wait(&anok[tr]) != 0 && errno == EINTR) {
                }
        },
        anok != nullptr);
    if (r) return r;
    FPCC_CHECK(hipMemsetAsync(b.dy, 0, sizeof(float) * ylen, nullptr));
    const cf32 *phi;
    if (complex_in) {
        if ((r = launch_ampnormf_rows((cf32 *)b.dx, 2 * (int)Tr, N, nullptr))) return r;
        phi = (const cf32 *)b.dx;
    } else if (small) {
        if ((r = launch_station_phasor((const float *)b.dx, (long long)Tr * N, 2, Tr, N, (cf32 *)b.dws, nullptr)))
            return r;
        phi = (const cf32 *)b.dws;
    } else {
        if ((r = large_station_stage(FPCC_PCC1, (const float *)b.dx, (long long)Tr * N, 2, (int)Tr, N, N,
                                     FPCC_STACK_NONE, (char *)b.dws, nullptr)))
            return r;
        phi = (const cf32 *)b.dws;
    }
    if ((r = launch_pcc_time(phi, 2, Tr, N, Lag1 + l1, Lp - l1, v, pcc1, 0, (float *)b.dy, L, l1, nullptr)))
        return r;
    FPCC_CHECK(hipMemcpy(b.hy, b.dy, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    rows_parallel(Tr, sizeof(float) * ylen, [&](size_t tr) {
        const float *src = (float *)b.hy + tr * L;
        if (zero_rows) memcpy(y[tr], src, sizeof(float) * L);
        else if (Lp > l1) memcpy(y[tr] + l1, src + l1, sizeof(float) * (Lp - l1));
    });
    return 0;
}

// One pair over Tr windows from host row pointers, as the reference's *_set:
// x1 rows become station 0, x2 rows station 1, outputs per window.
int pair_set_host(int kind, float **y, float **x1, float **x2, unsigned N, unsigned Tr, int Lag1, int Lag2,
                  double v)
{
    const unsigned ua1 = (unsigned)abs(Lag1), ua2 = (unsigned)abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    if (Tr == 0) return 0;
    if (!y || !x1 || !x2) return -1;
    fpcc_job job{};
    job.kind = kind;
    job.stack = FPCC_STACK_NONE;
    job.n_stations = 2;
    job.n_windows = Tr;
    job.win_len = N;
    job.station_stride = (unsigned long long)Tr * N;
    job.lag1 = std::min(Lag1, Lag2);
    job.lag2 = std::max(Lag1, Lag2);
    job.v = v;
    Dims d;
    int r = resolve(&job, d);
    if (r) return r;
    const size_t xlen = (size_t)2 * Tr * N, ylen = (size_t)Tr * d.L;
    for (unsigned tr = 0; tr < Tr; ++tr)
        if (!x1[tr] || !x2[tr] || !y[tr]) return -1;
    const size_t wsb = layout(&job, d).total;
    EntryBufs b;
    if (!b.get(sizeof(float) * xlen, sizeof(float) * ylen, sizeof(float) * xlen, sizeof(float) * ylen, wsb)) return -1;
    if ((r = upload_rows(x1, x2, Tr, N, (float *)b.hx, (float *)b.dx, [](unsigned) {}))) return r;
    if ((r = station_stage(&job, d, (const float *)b.dx, (char *)b.dws, nullptr))) return r;
    if ((r = pair_stage(&job, d, (char *)b.dws, (float *)b.dy, nullptr))) return r;
    FPCC_CHECK(hipMemcpy(b.hy, b.dy, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    rows_parallel(Tr, sizeof(float) * (size_t)Tr * d.L,
                  [&](size_t tr) { memcpy(y[tr], (float *)b.hy + tr * d.L, sizeof(float) * d.L); });
    return 0;
}

}  // namespace

extern "C" size_t fpcc_output_len(const fpcc_job *job)
{
    Dims d;
    if (resolve(job, d)) return 0;
    return d.P * (job->stack == FPCC_STACK_NONE ? d.W : 1) * (size_t)d.L;
}

extern "C" size_t fpcc_workspace_size(const fpcc_job *job)
{
    Dims d;
    if (resolve(job, d)) return 0;
    return layout(job, d).total;
}

extern "C" int fpcc_station_stage_device(const fpcc_job *job, const float *d_x, void *d_ws, size_t ws_bytes,
                                         void *stream)
{
    Dims d;
    int r = resolve(job, d);
    if (r) return r;
    if (!d_x || !d_ws || ws_bytes < layout(job, d).total) return -1;
    return station_stage(job, d, d_x, (char *)d_ws, (hipStream_t)stream);
}

extern "C" int fpcc_pair_stage_device(const fpcc_job *job, void *d_ws, size_t ws_bytes, float *d_y, void *stream)
{
    Dims d;
    int r = resolve(job, d);
    if (r) return r;
    if (!d_y || !d_ws || ws_bytes < layout(job, d).total) return -1;
    return pair_stage(job, d, (char *)d_ws, d_y, (hipStream_t)stream);
}

namespace {

// Pair stages that can share a launch: PCC2 windows in LDS at Nz 8192 with stacks
// none, plain and zsum (mode 0, 1, 2 of the persistent pair kernel); GNCC and 1-bit
// windows in LDS with stacks none and plain (mode 0, 1 of their pair kernel).  Jobs
// of one key (kind, Nz, mode) share.
bool pair_batch_mode(const fpcc_job *job, const Dims &d, int &mode)
{
    if (d.large) return false;
    if (job->kind == FPCC_PCC2) {
        if (d.Nz != 8192) return false;
        switch (job->stack) {
        case FPCC_STACK_NONE: mode = 0; return true;
        case FPCC_STACK_PLAIN: mode = 1; return true;
        case FPCC_STACK_ZSUM: mode = 2; return true;
        default: return false;
        }
    }
    if (job->kind == FPCC_GNCC || job->kind == FPCC_CC1B) {
        if (job->stack != FPCC_STACK_NONE && job->stack != FPCC_STACK_PLAIN) return false;
        mode = job->stack == FPCC_STACK_NONE ? 0 : 1;
        return true;
    }
    return false;
}

long long pair_batch_key(const fpcc_job *job, const Dims &d, int mode)
{
    return ((long long)d.Nz << 16) | ((long long)job->kind << 4) | mode;
}

}  // namespace

namespace fpcc {
bool job_pair_batch_mode(const fpcc_job *job, int *mode)
{
    Dims d;
    int m = 0;
    if (resolve(job, d) || !pair_batch_mode(job, d, m)) return false;
    if (mode) *mode = m;
    return true;
}
}  // namespace fpcc

extern "C" int fpcc_pair_stage_batch_device(const fpcc_job *jobs, int njobs, void *const *d_ws,
                                            const size_t *ws_bytes, float *const *d_y, void *stream)
{
    if (njobs < 0 || (njobs > 0 && (!jobs || !d_ws || !ws_bytes || !d_y))) return -1;
    hipStream_t st = (hipStream_t)stream;
    std::vector<Dims> D(njobs);
    std::vector<int> mode(njobs, -1);
    std::vector<long long> key(njobs, -1);
    for (int j = 0; j < njobs; ++j) {
        int r = resolve(&jobs[j], D[j]);
        if (r) return r;
        if (!d_ws[j] || !d_y[j] || ws_bytes[j] < layout(&jobs[j], D[j]).total) return -1;
        int m;
        if (pair_batch_mode(&jobs[j], D[j], m)) {
            mode[j] = m;
            key[j] = pair_batch_key(&jobs[j], D[j], m);
        }
    }
    std::vector<char> done(njobs, 0);
    for (int i = 0; i < njobs; ++i) {
        if (done[i]) continue;
        std::vector<int> grp{i};
        if (key[i] >= 0)
            for (int j = i + 1; j < njobs && (int)grp.size() < kMaxBatchJobs; ++j)
                if (!done[j] && key[j] == key[i]) grp.push_back(j);
        for (int g : grp) done[g] = 1;
        int r;
        if (grp.size() == 1) {
            r = pair_stage(&jobs[i], D[i], (char *)d_ws[i], d_y[i], st);
        } else if (jobs[i].kind != FPCC_PCC2) {
            const bool gn = jobs[i].kind == FPCC_GNCC;
            GnBatch jb{};
            jb.n = (int)grp.size();
            for (int k = 0; k < jb.n; ++k) {
                const int g = grp[k];
                const Layout l = layout(&jobs[g], D[g]);
                char *ws = (char *)d_ws[g];
                jb.Xa[k] = ws + l.a;
                jb.Xb[k] = gn ? ws + l.a : ws + l.b;
                jb.E[k] = gn ? (const double *)(ws + l.b) : nullptr;
                jb.edge[k] = gn ? nullptr : (const float *)(ws + l.f);
                jb.S[k] = D[g].S;
                jb.W[k] = D[g].W;
                jb.N[k] = D[g].N;
                jb.lag0[k] = D[g].lag0;
                jb.L[k] = D[g].L;
                jb.y[k] = d_y[g];
            }
            r = launch_pair_gn_batch(jb, gn, (int)D[i].Nz, gn ? 0 : 1, mode[i], st);
        } else {
            const int n = (int)grp.size();
            std::vector<const cf32 *> F(n);
            std::vector<int> S(n), W(n), N(n), lag0(n), L(n);
            std::vector<float *> y(n);
            for (int k = 0; k < n; ++k) {
                const int g = grp[k];
                F[k] = (const cf32 *)((char *)d_ws[g] + layout(&jobs[g], D[g]).a);
                S[k] = D[g].S;
                W[k] = D[g].W;
                N[k] = D[g].N;
                lag0[k] = D[g].lag0;
                L[k] = D[g].L;
                y[k] = d_y[g];
            }
            r = launch_pair_pcc2_4096_batch(n, F.data(), S.data(), W.data(), N.data(), lag0.data(), L.data(),
                                            mode[i], y.data(), st);
        }
        if (r) return r;
    }
    return 0;
}

namespace {

// Jobs whose station stages can share launches: GNCC and 1-bit (real-input
// transforms of Nz points; any station count, window count, window length and
// stride), grouped by kind, path and Nz.
bool batch_key(const fpcc_job *job, const Dims &d, long long &key)
{
    if (job->kind == FPCC_WPCC) {
        // jobs whose families share the filter bank (type, Nz, scales, voices, s0, w0):
        // a hash of those; groups are checked with wpcc_same_bank as well
        if (!d.fam || d.fam->type == 0) return false;
        const WFamily &f = *d.fam;
        unsigned long long h = 1469598103934665603ull;
        auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ull; };
        mix((unsigned long long)(long long)f.type);
        mix(f.Nz);
        mix(f.S);
        mix(f.V);
        mix(f.Ls0);
        mix(__builtin_bit_cast(unsigned long long, f.s0));
        mix(__builtin_bit_cast(unsigned long long, f.op1));
        key = (long long)((h << 8) | FPCC_WPCC);
        return true;
    }
    if (job->kind == FPCC_PCC2) {
        // the in-LDS station stage of any window length, except the specialised one;
        // long windows of one Nz share their zero-padded transforms
        if (!d.large && station_pcc_specialised(d.N, (int)d.Nz)) return false;
    } else if (job->kind != FPCC_GNCC && job->kind != FPCC_CC1B) {
        return false;
    }
    key = ((long long)d.Nz << 8) | ((long long)d.large << 4) | job->kind;
    return true;
}

// One group of batchable jobs (same key), at most kMaxBatchJobs.
int station_stage_group(const std::vector<const fpcc_job *> &jobs, const std::vector<Dims> &D,
                        const std::vector<const float *> &x, const std::vector<char *> &ws, hipStream_t st)
{
    const int n = (int)jobs.size();
    const int kind = jobs[0]->kind, Nz = (int)D[0].Nz;
    const bool large = D[0].large;
    RealBatch jb{};
    jb.n = n;
    jb.b0[0] = 0;
    for (int g = 0; g < n; ++g) {
        jb.x[g] = x[g];
        jb.stride[g] = (long long)jobs[g]->station_stride;
        jb.W[g] = D[g].W;
        jb.N[g] = D[g].N;
        jb.lde[g] = energy_ld(D[g].N);
        jb.b0[g + 1] = jb.b0[g] + D[g].S * D[g].W;
    }
    int r = 0;
    if (kind == FPCC_WPCC) {
        std::vector<long long> stride(n);
        std::vector<int> S(n), W(n), N(n);
        std::vector<char *> wsp(n);
        for (int g = 0; g < n; ++g) {
            stride[g] = (long long)jobs[g]->station_stride;
            S[g] = D[g].S;
            W[g] = D[g].W;
            N[g] = D[g].N;
            wsp[g] = ws[g] + layout(jobs[g], D[g]).a;
        }
        return wpcc_station_stage_batch(n, x.data(), stride.data(), S.data(), W.data(), N.data(), *D[0].fam,
                                        wsp.data(), st);
    }
    if (kind == FPCC_PCC2 && large) {
        if (jb.b0[n] > 65535) return -1;
        std::vector<long long> stride(n);
        std::vector<int> S(n), W(n), N(n), stack(n);
        std::vector<char *> wsp(n);
        for (int g = 0; g < n; ++g) {
            stride[g] = jb.stride[g];
            S[g] = D[g].S;
            W[g] = D[g].W;
            N[g] = D[g].N;
            stack[g] = jobs[g]->stack;  // the layout large_station_stage uses
            wsp[g] = ws[g] + layout(jobs[g], D[g]).a;
        }
        return large_pcc2_station_batch(n, x.data(), stride.data(), S.data(), W.data(), N.data(), stack.data(), Nz,
                                        wsp.data(), st);
    }
    if (kind == FPCC_PCC2) {
        PccBatch pb{};
        pb.n = n;
        for (int g = 0; g < n; ++g) {
            pb.b0[g] = jb.b0[g];
            pb.x[g] = jb.x[g];
            pb.stride[g] = jb.stride[g];
            pb.W[g] = jb.W[g];
            if ((r = station_pcc_batch_job(D[g].N, Nz, (cf32 *)(ws[g] + layout(jobs[g], D[g]).a), pb, g))) return r;
        }
        pb.b0[n] = jb.b0[n];
        return launch_station_pcc_batch(pb, Nz, st);
    }
    const int passes = kind == FPCC_CC1B ? 2 : 1;
    if (!large) {
        for (int p = 0; p < passes; ++p) {
            for (int g = 0; g < n; ++g) {
                const Layout l = layout(jobs[g], D[g]);
                jb.X[g] = ws[g] + (p ? l.b : l.a);
                jb.E[g] = kind == FPCC_GNCC ? (double *)(ws[g] + l.b) : nullptr;
            }
            if ((r = launch_station_real_batch(jb, Nz, kind == FPCC_CC1B ? p + 1 : 0, kind == FPCC_GNCC, st)))
                return r;
        }
        if (kind == FPCC_CC1B)  // window edges for the exact short-overlap lags
            for (int g = 0; g < n; ++g)
                if ((r = large_edge_copy(jb.x[g], jb.stride[g], D[g].S, D[g].W, D[g].N,
                                         (float *)(ws[g] + layout(jobs[g], D[g]).f), st)))
                    return r;
        return 0;
    }
    if (jb.b0[n] > 65535) return -1;
    const BigPlan &bz = big_plan(Nz);
    for (int p = 0; p < passes; ++p) {
        for (int g = 0; g < n; ++g) {
            size_t off[4], per[4];
            if (large_station_regions(kind, jobs[g]->stack, D[g].S, D[g].W, D[g].N, Nz, off, per) < 2) return -1;
            const Layout l = layout(jobs[g], D[g]);
            jb.X[g] = ws[g] + l.a + off[p];                                               // F, or F2 (1-bit)
            jb.E[g] = kind == FPCC_GNCC ? (double *)(ws[g] + l.a + off[1]) : nullptr;  // GNCC energies
        }
        ColArgs ca;
        ca.op = COL_FWD_INPUT;
        ca.bstride = Nz;
        ca.sign_rule = kind == FPCC_CC1B ? p + 1 : 0;
        if ((r = run_cols_batch(bz, ca, jb, st, "real_fwd_cols_batch"))) return r;
        RowArgs ra;
        ra.op = ROW_FWD_PLAIN;
        ra.bstride = Nz;
        if ((r = run_rows_batch(bz, ra, jb, st, "real_fwd_rows_batch"))) return r;
    }
    for (int g = 0; g < n; ++g) {  // window edges for the exact short-overlap lags
        size_t off[4], per[4];
        if (large_station_regions(kind, jobs[g]->stack, D[g].S, D[g].W, D[g].N, Nz, off, per) < 3) return -1;
        if ((r = large_edge_copy(jb.x[g], jb.stride[g], D[g].S, D[g].W, D[g].N,
                                 (float *)(ws[g] + layout(jobs[g], D[g]).a + off[2]), st)))
            return r;
    }
    if (kind == FPCC_GNCC) return launch_energy_prefix_batch(jb, st);
    return 0;
}

}  // namespace

namespace fpcc {
bool job_batch_key(const fpcc_job *job, long long *key)
{
    Dims d;
    if (resolve(job, d)) return false;
    return batch_key(job, d, *key);
}
}  // namespace fpcc

extern "C" int fpcc_station_stage_batch_device(const fpcc_job *jobs, int njobs, const float *const *d_x,
                                               void *const *d_ws, const size_t *ws_bytes, void *stream)
{
    if (njobs < 0 || (njobs > 0 && (!jobs || !d_x || !d_ws || !ws_bytes))) return -1;
    hipStream_t st = (hipStream_t)stream;
    std::vector<Dims> D(njobs);
    for (int j = 0; j < njobs; ++j) {
        int r = resolve(&jobs[j], D[j]);
        if (r) return r;
        if (!d_x[j] || !d_ws[j] || ws_bytes[j] < layout(&jobs[j], D[j]).total) return -1;
    }
    std::vector<char> done(njobs, 0);
    for (int i = 0; i < njobs; ++i) {
        if (done[i]) continue;
        long long ki;
        std::vector<int> grp{i};
        if (batch_key(&jobs[i], D[i], ki)) {
            // rows of one launch: station-windows (two-pass transforms) or, for WPCC,
            // station-window scales; a launch takes at most 65535
            const bool wpcc = jobs[i].kind == FPCC_WPCC;
            const long long per = wpcc ? (long long)D[i].fam->S : 1;
            long long sw = (long long)D[i].S * D[i].W * per;
            for (int j = i + 1; j < njobs && (int)grp.size() < kMaxBatchJobs; ++j) {
                long long kj;
                if (done[j] || !batch_key(&jobs[j], D[j], kj) || kj != ki) continue;
                if (wpcc && !wpcc_same_bank(*D[i].fam, *D[j].fam)) continue;
                const long long add = (long long)D[j].S * D[j].W * per;
                if ((D[i].large || wpcc) && sw + add > 65535) continue;
                sw += add;
                grp.push_back(j);
            }
        }
        for (int g : grp) done[g] = 1;
        int r;
        if (grp.size() == 1) {
            r = station_stage(&jobs[i], D[i], d_x[i], (char *)d_ws[i], st);
        } else {
            std::vector<const fpcc_job *> jp;
            std::vector<Dims> dg;
            std::vector<const float *> xg;
            std::vector<char *> wg;
            for (int g : grp) {
                jp.push_back(&jobs[g]);
                dg.push_back(D[g]);
                xg.push_back(d_x[g]);
                wg.push_back((char *)d_ws[g]);
            }
            r = station_stage_group(jp, dg, xg, wg, st);
        }
        if (r) return r;
    }
    return 0;
}

extern "C" int fpcc_allpairs_device(const fpcc_job *job, const float *d_x, float *d_y, void *d_ws, size_t ws_bytes,
                                    void *stream)
{
    int r = fpcc_station_stage_device(job, d_x, d_ws, ws_bytes, stream);
    if (r) return r;
    return fpcc_pair_stage_device(job, d_ws, ws_bytes, d_y, stream);
}

extern "C" int fpcc_allpairs(const fpcc_job *job, const float *h_x, float *h_y)
{
    Dims d;
    int r = resolve(job, d);
    if (r) return r;
    if (!h_x || !h_y) return -1;
    const size_t ylen = fpcc_output_len(job), ws = layout(job, d).total;
    const size_t xlen = (size_t)(job->n_stations - 1) * job->station_stride + (size_t)job->n_windows * job->win_len;
    {
        // through pinned staging when it can be had
        EntryBufs b;
        if (b.get(sizeof(float) * xlen, sizeof(float) * ylen, sizeof(float) * xlen, sizeof(float) * ylen, ws)) {
            constexpr size_t kRow = (size_t)1 << 18;  // floats a gathered piece of the input
            if ((r = upload_gather(
                     xlen, kRow, [&](size_t k) { return h_x + k * kRow; }, 0, [](size_t) {}, (float *)b.hx,
                     (float *)b.dx)))
                return r;
            r = fpcc_allpairs_device(job, (const float *)b.dx, (float *)b.dy, b.dws, ws, nullptr);
            if (r) return r;
            return download_scatter(ylen, (const float *)b.dy, (float *)b.hy, h_y);
        }
    }
    // else (pinned host memory refused) straight from and to the pageable buffers
    EntryBufs b;
    if (!b.get(0, 0, sizeof(float) * xlen, sizeof(float) * ylen, ws)) return -1;
    FPCC_CHECK(hipMemcpy(b.dx, h_x, sizeof(float) * xlen, hipMemcpyHostToDevice));
    r = fpcc_allpairs_device(job, (const float *)b.dx, (float *)b.dy, b.dws, ws, nullptr);
    if (r) return r;
    FPCC_CHECK(hipMemcpy(h_y, b.dy, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fpcc_release_buffers(void) { return release_kept(); }

extern "C" void fpcc_set_kept_limits(size_t host_bytes, size_t device_bytes)
{
    g_keep_host.store(host_bytes);
    g_keep_device.store(device_bytes);
}

// ---- reference-compatible entries (src/FFTapps.h:11-16) ----

extern "C" int pcc2_set(float **const y, float **const x1, float **const x2, const unsigned int N,
                        const unsigned int Tr, const int Lag1, const int Lag2)
{
    return pair_set_host(FPCC_PCC2, y, x1, x2, N, Tr, Lag1, Lag2, 2.0);
}

extern "C" int ccgn_set(float **const y, float **const x1, float **const x2, const unsigned int N,
                        const unsigned int Tr, const int Lag1, const int Lag2)
{
    return pair_set_host(FPCC_GNCC, y, x1, x2, N, Tr, Lag1, Lag2, 2.0);
}

extern "C" int cc1b_set(float **const y, float **const x1, float **const x2, const unsigned int N,
                        const unsigned int Tr, const int Lag1, const int Lag2)
{
    return pair_set_host(FPCC_CC1B, y, x1, x2, N, Tr, Lag1, Lag2, 2.0);
}

extern "C" int pcc1_set(float **const y, float **const x1, float **const x2, const int N, const unsigned int Tr,
                        const int Lag1, const int Lag2)
{
    return pcc_time_host(1, y, x1, x2, N, Tr, 1.0, Lag1, Lag2, 0, nullptr, true);
}

extern "C" int pcc_set(float **const y, float **const x1, float **const x2, const int N, const unsigned int Tr,
                       const double v, const int Lag1, const int Lag2)
{
    return pcc_time_host(0, y, x1, x2, N, Tr, v, Lag1, Lag2, 0, nullptr, true);
}

// ---- reference GPU-engine entries (src/ccs_cuda.h:8-19) ----

extern "C" int pcc1_highlevel(float **y, _Complex float **x1, _Complex float **x2, int N, unsigned int Tr, int Lag1,
                              int Lag2, sem_t *anok)
{
    return pcc_time_host(1, y, (const float *const *)x1, (const float *const *)x2, N, Tr, 1.0, Lag1, Lag2, 1, anok,
                         false);
}

extern "C" int pcc_highlevel(float **y, _Complex float **x1, _Complex float **x2, int N, unsigned int Tr, float v,
                             int Lag1, int Lag2, sem_t *anok)
{
    return pcc_time_host(0, y, (const float *const *)x1, (const float *const *)x2, N, Tr, (double)v, Lag1, Lag2, 1,
                         anok, false);
}

extern "C" void pcc1_highlevel2(float **y, float **x1, float **x2, int N, unsigned int Tr, int Lag1, int Lag2)
{
    int r = pcc_time_host(1, y, x1, x2, N, Tr, 1.0, Lag1, Lag2, 0, nullptr, false);
    if (r) fprintf(stderr, "fastpcc: pcc1_highlevel2 failed (%d)\n", r);
}

// WPCC, replaces tspcc2_set (src/FFTapps.c:1076-1255).  L = abs(Lag2-Lag1+1) and
// rows are zeroed before the family is built, as the reference.
extern "C" int tspcc2_set(float **const y, float **const x1, float **const x2, const unsigned int N,
                          const unsigned int Tr, const int Lag1, const int Lag2, double pmin, double pmax,
                          unsigned int V, int type, double op1)
{
    const int L = abs(Lag2 - Lag1 + 1);
    const unsigned ua1 = (unsigned)abs(Lag1), ua2 = (unsigned)abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    if (y)
        for (unsigned tr = 0; tr < Tr; ++tr)
            if (y[tr]) memset(y[tr], 0, sizeof(float) * L);
    int err = 0;
    const std::shared_ptr<const WFamily> f = get_wfamily(N, Lag1, Lag2, pmin, pmax, V, type, op1, &err);
    if (!f) return err ? err : -1;
    const int lag0 = Lag2 >= Lag1 ? Lag1 : Lag2;
    return wpcc_pair_host(*f, y, x1, x2, N, Tr, lag0, L);
}

// Reference GPU WPCC entry (src/ccs_cuda.h:12, src/ccs_cuda.cu:898-1036): the
// caller supplies Nz, the S wavelets (FP64, length Ls[s], origin center[s]) and
// C = 1/(K0 N).  Down_smp must be 1 (-2 otherwise).  Phasors follow the CPU
// regulariser z / sqrt(|z|^2 + 1e-6 max|z|^2), not the CUDA threshold.
extern "C" int cuda_tspcc2_set_freq(float **y0, float **x01, float **x02, unsigned int N, unsigned int Tr,
                                    unsigned int Nz, const int Lag1, const int Lag2, unsigned int S, double *scale,
                                    _Complex double **wc, unsigned int *Ls, int *center, unsigned int *Down_smp,
                                    float C)
{
    if (!scale || !wc || !Ls || !center || !Down_smp || S == 0) return -1;
    for (unsigned s = 0; s < S; ++s)
        if (Down_smp[s] != 1) return -2;
    if (Nz == 0 || (Nz & (Nz - 1)) || !(C > 0)) return -1;
    WFamily f;
    f.type = 0;
    f.S = S;
    f.Nz = Nz;
    f.Ls0 = Ls[S - 1];
    f.K0 = 1.0 / ((double)C * (double)N);
    f.scale.assign(scale, scale + S);
    f.Ls.assign(Ls, Ls + S);
    f.center.assign(center, center + S);
    f.wc.resize(S);
    for (unsigned s = 0; s < S; ++s) {
        if (!wc[s] || Ls[s] > Nz || center[s] < 0 || (unsigned)center[s] > Ls[s]) return -1;
        const double *p = (const double *)wc[s];
        f.wc[s].resize(Ls[s]);
        for (unsigned u = 0; u < Ls[s]; ++u) f.wc[s][u] = std::complex<double>(p[2 * u], p[2 * u + 1]);
    }
    const int Lags = Lag2 >= Lag1 ? Lag2 - Lag1 + 1 : Lag1 - Lag2 + 1;
    const int lag0 = Lag2 >= Lag1 ? Lag1 : Lag2;
    return wpcc_pair_host(f, y0, x01, x02, N, Tr, lag0, Lags);
}
