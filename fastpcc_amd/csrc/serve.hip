// Request server: continuous batching of all-pairs correlation jobs on one GPU
// (no counterpart in the reference, which correlates one station pair per process
// run, src/PCC_fullpair_1b.c:467-499).
//
// Requests are admitted by HBM footprint: each needs its input, output and
// workspace (fpcc_workspace_size) as blocks of one device arena reserved at
// creation, so serving never allocates device memory per request.  Admission is
// FIFO; a request that does not fit yet may be passed by later ones that do, at
// most max_bypass times, after which the queue waits for memory to drain.
// Admitted requests run on a ring of HIP streams, so their copies and kernels
// overlap on the device: H2D input, station stage (spectra computed once per
// station and shared by every pair), pair stage, D2H output, then a host function
// on the stream hands the request to the completion thread, which finishes
// pageable outputs, returns the memory and wakes the scheduler.  Pinned caller
// buffers are copied by DMA directly; pageable ones go through a pinned staging
// arena.  Latency = submit -> result in the caller's buffer.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>
#include "../../include/fastpcc.h"
#include "arena.hpp"
#include "engine.hpp"
#include "kernels.hpp"

namespace fpcc {

namespace {

using clock_t_ = std::chrono::steady_clock;

constexpr int kMaxRegions = 4;
// p50/p99 are taken over the latencies of the last kLatencyKeep completed requests
constexpr size_t kLatencyKeep = (size_t)1 << 20;
constexpr int kCopyStations = 128;  // stations per cache copy launch (offsets go by value)

// A HIP event shared by the request that fills cache entries and the requests
// that wait for them; destroyed with its last reference.
struct Event {
    hipEvent_t e = nullptr;
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// Station rows (every station-output region of one station) resident in the
// cache arena.  ev: recorded after the fill, until the filling request is done.
struct CacheEntry {
    size_t off = 0, bytes = 0;
    int pins = 0;
    unsigned long long tick = 0;
    std::shared_ptr<Event> ev;
};

typedef std::tuple<unsigned long long, int, unsigned, unsigned, size_t, size_t, double, double, unsigned, int, double>
    CacheKey;

struct Request {
    unsigned long long id = 0;
    fpcc_job job{};      // as submitted
    fpcc_job dev_job{};  // compact device layout: station_stride = W*N
    const float *x = nullptr;
    float *y = nullptr;
    size_t x_bytes = 0, y_bytes = 0, ws_bytes = 0;
    bool x_pinned = false, y_pinned = false;
    size_t off_x = 0, off_y = 0, off_ws = 0, off_sx = 0, off_sy = 0;  // arena offsets
    int state = FPCC_REQ_QUEUED, rc = 0;
    bool cancel = false;
    int bypassed = 0;
    int stream = -1;
    double t_submit = 0, t_start = 0, t_done = 0;
    unsigned long long pair_corr = 0;
    bool batchable = false;    // station stage may share launches (fpcc_station_stage_batch_device)
    long long batch_key = 0;
    // station-spectra cache (keyed requests)
    std::vector<unsigned long long> keys;
    int nreg = 0;
    size_t off[kMaxRegions] = {}, per[kMaxRegions] = {}, rbase[kMaxRegions] = {}, entry_bytes = 0;
    bool use_cache = false, fill_done = false;
    bool direct = false;                 // every station missing: compute in place, copy out once
    std::vector<CacheEntry *> entries;   // per station
    std::vector<int> misses;             // stations computed by this request
    std::vector<CacheKey> new_keys;      // entries this request created
    fpcc_job sub_job{};                  // the missing stations, gathered
    size_t sub_off[kMaxRegions] = {};
    size_t off_xs = 0, xs_bytes = 0, off_wss = 0, wss_bytes = 0, off_sxs = 0;
    std::shared_ptr<Event> fill_ev;
    std::vector<std::shared_ptr<Event>> waits;
};

struct Server;

struct DoneMsg {
    Server *srv;
    std::shared_ptr<Request> req;
};

struct Server {
    fpcc_server_config cfg{};
    int device = 0;
    void *dmem = nullptr, *hmem = nullptr, *cmem = nullptr;
    Arena darena, harena, carena;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;
    clock_t_::time_point t0;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_wait;
    std::deque<std::shared_ptr<Request>> queue;
    std::deque<std::shared_ptr<Request>> finished;  // handed over by stream host functions
    std::unordered_map<unsigned long long, std::shared_ptr<Request>> all;
    std::map<CacheKey, CacheEntry> cache;
    unsigned long long next_id = 1, tick = 0;
    int inflight = 0, inflight_peak = 0;
    size_t dev_peak = 0;
    bool stop = false;
    // statistics
    unsigned long long n_submit = 0, n_done = 0, n_fail = 0, n_cancel = 0, n_reject = 0;
    unsigned long long n_hit = 0, n_miss = 0, n_batched = 0, n_batch_launch = 0;
    // latencies of the last kLatencyKeep completed requests (a ring: memory stays
    // bounded in a long-running loop); sum and max cover every completed request
    std::vector<double> lat_ms;
    size_t lat_next = 0;
    double lat_sum = 0, lat_max = 0;
    double first_start = -1, last_done = 0;
    unsigned long long pair_corr_done = 0;

    std::thread sched, compl_;

    double now() const { return std::chrono::duration<double>(clock_t_::now() - t0).count(); }
};

bool is_pinned(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();  // pageable memory reports an error here; clear it
    return e == hipSuccess && a.type == hipMemoryTypeHost;
}

void stream_host_done(void *arg)
{
    // runs on the runtime's callback thread: no HIP calls, hand over and return
    DoneMsg *m = static_cast<DoneMsg *>(arg);
    Server *s = m->srv;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->finished.push_back(m->req);
    }
    s->cv_done.notify_one();
    delete m;
}

// ---- station rows between the cache and workspaces ----
// Row block (station t, region k): src_base + src_off[k] + t*src_step[k] (+ eoff[t]
// if src_entry), same for dst; per[k] bytes.  Grid: (chunks of 256 KiB, t*nreg + k).
struct CopyArgs {
    const char *src_base;
    char *dst_base;
    int nreg, nst, src_entry, dst_entry;
    size_t src_off[kMaxRegions], dst_off[kMaxRegions], src_step[kMaxRegions], dst_step[kMaxRegions];
    size_t per[kMaxRegions];
    size_t eoff[kCopyStations];
};

constexpr size_t kCopyChunk = 256 << 10;

__global__ void __launch_bounds__(256) k_copy_rows(CopyArgs a)
{
    const int t = blockIdx.y / a.nreg, k = blockIdx.y - t * a.nreg;
    const size_t bytes = a.per[k];
    const size_t c0 = (size_t)blockIdx.x * kCopyChunk;
    if (c0 >= bytes) return;
    const size_t c1 = bytes - c0 < kCopyChunk ? bytes : c0 + kCopyChunk;
    const char *src = a.src_base + a.src_off[k] + t * a.src_step[k] + (a.src_entry ? a.eoff[t] : 0);
    char *dst = a.dst_base + a.dst_off[k] + t * a.dst_step[k] + (a.dst_entry ? a.eoff[t] : 0);
    if ((((size_t)src | (size_t)dst | c0 | c1) & 15) == 0) {
        for (size_t i = c0 + 16 * threadIdx.x; i < c1; i += 16 * blockDim.x)
            *reinterpret_cast<uint4 *>(dst + i) = *reinterpret_cast<const uint4 *>(src + i);
    } else {  // every region holds 4-byte words
        for (size_t i = c0 + 4 * threadIdx.x; i < c1; i += 4 * blockDim.x)
            *reinterpret_cast<unsigned *>(dst + i) = *reinterpret_cast<const unsigned *>(src + i);
    }
}

int copy_rows(CopyArgs a, const std::vector<size_t> &eoff, hipStream_t st)
{
    size_t maxb = 0;
    for (int k = 0; k < a.nreg; ++k) maxb = std::max(maxb, a.per[k]);
    const unsigned chunks = (unsigned)((maxb + kCopyChunk - 1) / kCopyChunk);
    const int n = (int)eoff.size();
    for (int t0 = 0; t0 < n; t0 += kCopyStations) {
        a.nst = std::min(kCopyStations, n - t0);
        for (int t = 0; t < a.nst; ++t) a.eoff[t] = eoff[t0 + t];
        CopyArgs b = a;
        for (int k = 0; k < a.nreg; ++k) {  // shift the stepped side to station t0
            if (!a.src_entry) b.src_off[k] += (size_t)t0 * a.src_step[k];
            if (!a.dst_entry) b.dst_off[k] += (size_t)t0 * a.dst_step[k];
        }
        hipLaunchKernelGGL(k_copy_rows, dim3(chunks, a.nst * a.nreg), dim3(256), 0, st, b);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

// Async host->device copy of station rows `rows` of the caller's input into
// consecutive rows of dst (runs of consecutive stations as one 2D copy).
int copy_in_rows(Server *s, Request &r, const std::vector<int> &rows, float *dst, size_t off_stage, hipStream_t st)
{
    const fpcc_job &j = r.job;
    const size_t row = (size_t)j.n_windows * j.win_len * sizeof(float);
    if (!r.x_pinned) {
        char *hx = (char *)s->hmem + off_stage;
        for (size_t i = 0; i < rows.size(); ++i) memcpy(hx + i * row, r.x + (size_t)rows[i] * j.station_stride, row);
        return hipMemcpyAsync(dst, hx, rows.size() * row, hipMemcpyHostToDevice, st) == hipSuccess ? 0 : -1;
    }
    for (size_t i = 0; i < rows.size();) {
        size_t n = 1;
        while (i + n < rows.size() && rows[i + n] == rows[i] + (int)n) ++n;
        if (hipMemcpy2DAsync((char *)dst + i * row, row, r.x + (size_t)rows[i] * j.station_stride,
                             j.station_stride * sizeof(float), row, n, hipMemcpyHostToDevice, st) != hipSuccess)
            return -1;
        i += n;
    }
    return 0;
}

// H2D of an uncached request's whole input into its compact device rows.
int stage_input(Server *s, Request &r)
{
    hipStream_t st = s->streams[r.stream];
    const fpcc_job &j = r.job;
    const size_t row = (size_t)j.n_windows * j.win_len * sizeof(float);
    float *dx = (float *)((char *)s->dmem + r.off_x);
    if (r.x_pinned)
        return hipMemcpy2DAsync(dx, row, r.x, j.station_stride * sizeof(float), row, j.n_stations,
                                hipMemcpyHostToDevice, st) == hipSuccess ? 0 : -1;
    char *hx = (char *)s->hmem + r.off_sx;
    for (unsigned i = 0; i < j.n_stations; ++i) memcpy(hx + i * row, r.x + (size_t)i * j.station_stride, row);
    return hipMemcpyAsync(dx, hx, r.x_bytes, hipMemcpyHostToDevice, st) == hipSuccess ? 0 : -1;
}

// D2H of the output (into the caller's pinned y, or the staging block).
int copy_out(Server *s, Request &r)
{
    void *yh = r.y_pinned ? (void *)r.y : (void *)((char *)s->hmem + r.off_sy);
    return hipMemcpyAsync(yh, (char *)s->dmem + r.off_y, r.y_bytes, hipMemcpyDeviceToHost, s->streams[r.stream]) ==
                   hipSuccess ? 0 : -1;
}

bool injected_failure(Server *s, const Request &r)
{
    return s->cfg.fail_every > 0 && r.id % (unsigned long long)s->cfg.fail_every == 0;
}

// Launches one admitted request on its stream (scheduler thread, lock not held).
int issue(Server *s, Request &r)
{
    (void)hipSetDevice(s->device);
    hipStream_t st = s->streams[r.stream];
    char *dbase = (char *)s->dmem, *hbase = (char *)s->hmem;
    float *dy = (float *)(dbase + r.off_y);
    char *ws = dbase + r.off_ws;
    const fpcc_job &j = r.job;
    const size_t row = (size_t)j.n_windows * j.win_len * sizeof(float);
    if (injected_failure(s, r)) return -1;
    if (r.use_cache) {
        if (!r.misses.empty()) {
            std::vector<int> rows(r.misses);
            if (rows.size() < 2) rows.push_back(rows[0]);  // a job has at least two stations
            float *dxs = (float *)(dbase + r.off_xs);
            if (copy_in_rows(s, r, rows, dxs, r.off_sxs, st)) return -1;
            // direct: the request's own station stage; else the gathered missing stations
            char *wss = r.direct ? ws : dbase + r.off_wss;
            int rc = r.direct ? fpcc_station_stage_device(&r.dev_job, dxs, ws, r.ws_bytes, st)
                              : fpcc_station_stage_device(&r.sub_job, dxs, wss, r.wss_bytes, st);
            if (rc) return rc;
            CopyArgs a{};
            a.src_base = wss;
            a.dst_base = (char *)s->cmem;
            a.nreg = r.nreg;
            a.dst_entry = 1;
            for (int k = 0; k < r.nreg; ++k) {
                a.src_off[k] = r.direct ? r.off[k] : r.sub_off[k];
                a.src_step[k] = r.per[k];
                a.dst_off[k] = r.rbase[k];
                a.per[k] = r.per[k];
            }
            std::vector<size_t> eoff;
            for (int m : r.misses) eoff.push_back(r.entries[m]->off);
            if (copy_rows(a, eoff, st)) return -1;
            if (hipEventRecord(r.fill_ev->e, st) != hipSuccess) return -1;
            r.fill_done = true;
        }
        for (auto &w : r.waits)
            if (hipStreamWaitEvent(st, w->e, 0) != hipSuccess) return -1;
        if (r.direct) {  // the workspace already holds every station
            int rc = fpcc_pair_stage_device(&r.dev_job, ws, r.ws_bytes, dy, st);
            if (rc) return rc;
            void *yh = r.y_pinned ? (void *)r.y : (void *)(hbase + r.off_sy);
            return hipMemcpyAsync(yh, dy, r.y_bytes, hipMemcpyDeviceToHost, st) == hipSuccess ? 0 : -1;
        }
        CopyArgs a{};
        a.src_base = (const char *)s->cmem;
        a.dst_base = ws;
        a.nreg = r.nreg;
        a.src_entry = 1;
        for (int k = 0; k < r.nreg; ++k) {
            a.src_off[k] = r.rbase[k];
            a.dst_off[k] = r.off[k];
            a.dst_step[k] = r.per[k];
            a.per[k] = r.per[k];
        }
        std::vector<size_t> eoff;
        for (CacheEntry *e : r.entries) eoff.push_back(e->off);
        if (copy_rows(a, eoff, st)) return -1;
        int rc = fpcc_pair_stage_device(&r.dev_job, ws, r.ws_bytes, dy, st);
        if (rc) return rc;
    } else {
        if (stage_input(s, r)) return -1;
        int rc = fpcc_allpairs_device(&r.dev_job, (float *)(dbase + r.off_x), dy, ws, r.ws_bytes, st);
        if (rc) return rc;
    }
    (void)row;
    return copy_out(s, r);
}

// A group of uncached requests whose station stages share launches: inputs on
// their own streams, the grouped station stage on the first one's stream (joined
// by events both ways), the pair stages that share launches on that stream too,
// then every other pair stage and each output copy on its request's own stream.
int issue_group(Server *s, std::vector<Request *> &g)
{
    (void)hipSetDevice(s->device);
    char *dbase = (char *)s->dmem;
    hipStream_t bst = s->streams[g[0]->stream];
    std::vector<hipEvent_t> evs;
    auto cleanup = [&] {
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);  // released once complete
    };
    auto event = [&](hipEvent_t &e) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
        evs.push_back(e);
        return true;
    };
    std::vector<fpcc_job> jobs;
    std::vector<const float *> xs;
    std::vector<void *> wss;
    std::vector<size_t> wsb;
    for (Request *r : g) {
        if (stage_input(s, *r)) {
            cleanup();
            return -1;
        }
        if (r != g[0]) {
            hipEvent_t e;
            if (!event(e) || hipEventRecord(e, s->streams[r->stream]) != hipSuccess ||
                hipStreamWaitEvent(bst, e, 0) != hipSuccess) {
                cleanup();
                return -1;
            }
        }
        jobs.push_back(r->dev_job);
        xs.push_back((const float *)(dbase + r->off_x));
        wss.push_back(dbase + r->off_ws);
        wsb.push_back(r->ws_bytes);
    }
    int rc = fpcc_station_stage_batch_device(jobs.data(), (int)jobs.size(), xs.data(), wss.data(), wsb.data(), bst);
    // pair stages that can share launches (fpcc_pair_stage_batch_device) run together
    // on the group's stream, the others each on its own
    std::vector<char> shared(g.size(), 0);
    std::vector<fpcc_job> pj;
    std::vector<void *> pw;
    std::vector<size_t> pb;
    std::vector<float *> py;
    for (size_t k = 0; k < g.size(); ++k) {
        int m;
        if (fpcc::job_pair_batch_mode(&g[k]->dev_job, &m)) {
            shared[k] = 1;
            pj.push_back(g[k]->dev_job);
            pw.push_back(dbase + g[k]->off_ws);
            pb.push_back(g[k]->ws_bytes);
            py.push_back((float *)(dbase + g[k]->off_y));
        }
    }
    if (pj.size() < 2) std::fill(shared.begin(), shared.end(), 0);
    hipEvent_t done, pdone = nullptr;  // station stages done; shared pair stages done
    if (!rc && (!event(done) || hipEventRecord(done, bst) != hipSuccess)) rc = -1;
    if (!rc && pj.size() >= 2) {
        rc = fpcc_pair_stage_batch_device(pj.data(), (int)pj.size(), pw.data(), pb.data(), py.data(), bst);
        if (!rc && (!event(pdone) || hipEventRecord(pdone, bst) != hipSuccess)) rc = -1;
    }
    for (size_t k = 0; k < g.size(); ++k) {
        Request *r = g[k];
        if (rc) break;
        hipStream_t st = s->streams[r->stream];
        if (r != g[0] && hipStreamWaitEvent(st, shared[k] ? pdone : done, 0) != hipSuccess) rc = -1;
        if (!rc && !shared[k])
            rc = fpcc_pair_stage_device(&r->dev_job, dbase + r->off_ws, r->ws_bytes, (float *)(dbase + r->off_y), st);
        if (!rc) rc = copy_out(s, *r);
    }
    cleanup();
    return rc;
}

CacheKey cache_key(const Request &r, unsigned long long station)
{
    const fpcc_job &j = r.job;
    const bool w = j.kind == FPCC_WPCC;
    return CacheKey(station, j.kind, j.win_len, j.n_windows, r.per[0], r.nreg > 1 ? r.per[1] : 0, w ? j.pmin : 0,
                    w ? j.pmax : 0, w ? j.V : 0, w ? j.wtype : 0, w ? j.op1 : 0);
}

void release(Server *s, Request &r)
{
    if (r.use_cache) {
        for (CacheEntry *e : r.entries)
            if (e) e->pins--;
        if (r.fill_done)
            for (int m : r.misses)
                if (r.entries[m] && r.entries[m]->ev == r.fill_ev) r.entries[m]->ev.reset();  // filled
        if (!r.misses.empty()) {
            s->darena.release(r.off_xs, r.xs_bytes);
            if (r.wss_bytes) s->darena.release(r.off_wss, r.wss_bytes);
            if (!r.x_pinned) s->harena.release(r.off_sxs, r.xs_bytes);
        }
        r.entries.clear();
        r.waits.clear();
        r.fill_ev.reset();
    } else {
        s->darena.release(r.off_x, r.x_bytes);
        if (!r.x_pinned) s->harena.release(r.off_sx, r.x_bytes);
    }
    s->darena.release(r.off_y, r.y_bytes);
    s->darena.release(r.off_ws, r.ws_bytes);
    if (!r.y_pinned) s->harena.release(r.off_sy, r.y_bytes);
    s->stream_load[r.stream]--;
    s->inflight--;
}

// Evicts the least recently used unpinned entry (lock held); false if none.
bool evict_one(Server *s)
{
    auto victim = s->cache.end();
    for (auto it = s->cache.begin(); it != s->cache.end(); ++it)
        if (!it->second.pins && (victim == s->cache.end() || it->second.tick < victim->second.tick)) victim = it;
    if (victim == s->cache.end()) return false;
    s->carena.release(victim->second.off, victim->second.bytes);
    s->cache.erase(victim);
    return true;
}

// Cached admission (lock held): 1 admitted, 0 not now, -1 cache not usable for
// this request (too small; it then runs uncached).
int reserve_cached(Server *s, Request &r)
{
    const unsigned S = r.job.n_stations;
    r.entries.assign(S, nullptr);
    r.misses.clear();
    r.new_keys.clear();
    r.waits.clear();
    std::vector<int> dup(S, -1);  // a station repeating an earlier missing station's key
    for (unsigned i = 0; i < S; ++i) {
        auto it = s->cache.find(cache_key(r, r.keys[i]));
        if (it != s->cache.end()) {
            r.entries[i] = &it->second;
            continue;
        }
        for (int m : r.misses)
            if (r.keys[m] == r.keys[i]) {
                dup[i] = m;
                break;
            }
        if (dup[i] < 0) r.misses.push_back((int)i);
    }
    const int M = (int)r.misses.size();
    const size_t row = (size_t)r.job.n_windows * r.job.win_len * sizeof(float);
    r.xs_bytes = r.wss_bytes = 0;
    r.direct = M == (int)S;  // no hits and no repeated keys: the job itself is the gathered job
    if (r.direct) {
        r.xs_bytes = S * row;
    } else if (M) {
        r.sub_job = r.dev_job;
        r.sub_job.n_stations = std::max(2, M);
        r.xs_bytes = r.sub_job.n_stations * row;
        r.wss_bytes = fpcc_workspace_size(&r.sub_job);
        size_t so[kMaxRegions], sp[kMaxRegions];
        if (!r.wss_bytes || job_station_regions(&r.sub_job, so, sp) != r.nreg) return -1;
        for (int k = 0; k < r.nreg; ++k) r.sub_off[k] = so[k];
    }
    if ((size_t)M * r.entry_bytes > s->carena.size()) return -1;
    const size_t need_d[4] = {r.y_bytes, r.ws_bytes, r.xs_bytes, r.wss_bytes};
    const size_t need_h[2] = {r.x_pinned ? 0 : r.xs_bytes, r.y_pinned ? 0 : r.y_bytes};
    if (!s->darena.fits(need_d) || !s->harena.fits(need_h)) return 0;
    for (CacheEntry *e : r.entries)
        if (e) e->pins++;
    std::vector<size_t> noff;
    bool ok = true;
    for (int q = 0; q < M && ok; ++q) {
        size_t o = 0;
        while (!s->carena.alloc(r.entry_bytes, o))
            if (!evict_one(s)) {
                ok = false;
                break;
            }
        if (ok) noff.push_back(o);
    }
    std::shared_ptr<Event> ev;
    if (ok && M) {
        ev = std::make_shared<Event>();
        ok = hipEventCreateWithFlags(&ev->e, hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {  // pinned entries could not make room: run this one uncached
        for (size_t o : noff) s->carena.release(o, r.entry_bytes);
        for (CacheEntry *e : r.entries)
            if (e) e->pins--;
        r.entries.clear();
        r.misses.clear();
        return -1;
    }
    // commit
    for (unsigned i = 0; i < S; ++i)
        if (r.entries[i] && r.entries[i]->ev) r.waits.push_back(r.entries[i]->ev);  // still being filled
    r.fill_ev = ev;
    for (int q = 0; q < M; ++q) {
        const CacheKey key = cache_key(r, r.keys[r.misses[q]]);
        CacheEntry &e = s->cache[key];
        e.off = noff[q];
        e.bytes = r.entry_bytes;
        e.pins = 1;
        e.ev = ev;
        r.entries[r.misses[q]] = &e;
        r.new_keys.push_back(key);
    }
    for (unsigned i = 0; i < S; ++i)
        if (dup[i] >= 0) {
            r.entries[i] = r.entries[dup[i]];
            r.entries[i]->pins++;
        }
    for (CacheEntry *e : r.entries) e->tick = ++s->tick;
    s->darena.alloc(r.y_bytes, r.off_y);
    s->darena.alloc(r.ws_bytes, r.off_ws);
    if (M) {
        s->darena.alloc(r.xs_bytes, r.off_xs);
        if (r.wss_bytes) s->darena.alloc(r.wss_bytes, r.off_wss);
        if (!r.x_pinned) s->harena.alloc(r.xs_bytes, r.off_sxs);
    }
    if (!r.y_pinned) s->harena.alloc(r.y_bytes, r.off_sy);
    s->n_hit += S - M;
    s->n_miss += M;
    r.use_cache = true;
    r.fill_done = false;
    return 1;
}

// Reserves the request's blocks if they all fit (lock held).
bool try_reserve(Server *s, Request &r)
{
    if (s->cmem && !r.keys.empty() && r.nreg > 0) {
        const int c = reserve_cached(s, r);
        if (c >= 0) return c == 1;
    }
    r.use_cache = false;
    const size_t need_d[3] = {r.x_bytes, r.y_bytes, r.ws_bytes};
    if (!s->darena.fits(need_d)) return false;
    const size_t need_h[2] = {r.x_pinned ? 0 : r.x_bytes, r.y_pinned ? 0 : r.y_bytes};
    if (!s->harena.fits(need_h)) return false;
    s->darena.alloc(r.x_bytes, r.off_x);
    s->darena.alloc(r.y_bytes, r.off_y);
    s->darena.alloc(r.ws_bytes, r.off_ws);
    if (!r.x_pinned) s->harena.alloc(r.x_bytes, r.off_sx);
    if (!r.y_pinned) s->harena.alloc(r.y_bytes, r.off_sy);
    return true;
}

void scheduler(Server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_sched.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        if (s->stop && s->queue.empty()) return;
        // admit in FIFO order; later requests may pass one that does not fit yet
        std::shared_ptr<Request> pick;
        for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
            if (s->inflight >= s->cfg.max_inflight) break;
            Request &r = **it;
            if (try_reserve(s, r)) {
                pick = *it;
                s->queue.erase(it);
                break;
            }
            if (++r.bypassed > s->cfg.max_bypass) break;  // hold the line for this one
        }
        if (!pick) {
            // nothing fits: wait for a completion (or a new request / shutdown)
            s->cv_sched.wait(lk);
            continue;
        }
        auto start = [&](Request &q) {
            int best = 0;
            for (int i = 1; i < (int)s->streams.size(); ++i)
                if (s->stream_load[i] < s->stream_load[best]) best = i;
            q.stream = best;
            s->stream_load[best]++;
            s->inflight++;
            s->inflight_peak = std::max(s->inflight_peak, s->inflight);
            s->dev_peak = std::max(s->dev_peak, s->darena.used());
            q.state = FPCC_REQ_RUNNING;
            q.t_start = s->now();
            if (s->first_start < 0) s->first_start = q.t_start;
        };
        Request &r = *pick;
        start(r);
        // the requests next in line whose station stages can share launches with it
        std::vector<std::shared_ptr<Request>> group{pick};
        if (!r.use_cache && r.batchable && !injected_failure(s, r)) {
            while (!s->queue.empty() && (int)group.size() < kMaxBatchJobs && s->inflight < s->cfg.max_inflight) {
                Request &q = *s->queue.front();
                if (!q.batchable || q.batch_key != r.batch_key || injected_failure(s, q) || !try_reserve(s, q))
                    break;
                start(q);
                group.push_back(s->queue.front());
                s->queue.pop_front();
            }
        }
        if (group.size() > 1) {
            s->n_batched += group.size();
            s->n_batch_launch++;
            lk.unlock();
            std::vector<Request *> g;
            for (auto &q : group) g.push_back(q.get());
            const int rc = issue_group(s, g);
            lk.lock();
            for (auto &q : group)
                if (rc) q->rc = rc;
            lk.unlock();
            for (auto &q : group) {
                DoneMsg *m = new DoneMsg{s, q};
                hipStream_t st = s->streams[q->stream];
                if (hipLaunchHostFunc(st, stream_host_done, m) != hipSuccess) {
                    (void)hipStreamSynchronize(st);
                    stream_host_done(m);
                }
            }
            lk.lock();
            continue;
        }
        lk.unlock();
        const int rc = issue(s, r);
        lk.lock();
        if (rc) {
            r.rc = rc;
            // entries whose fill never reached the stream must not be found
            // (no admission ran since this request's, so nobody else holds them)
            if (r.use_cache && !r.fill_done) {
                for (const CacheKey &k : r.new_keys) {
                    auto it = s->cache.find(k);
                    if (it == s->cache.end()) continue;
                    CacheEntry *e = &it->second;
                    for (auto &p : r.entries)
                        if (p == e) p = nullptr;
                    s->carena.release(e->off, e->bytes);
                    s->cache.erase(it);
                }
                r.new_keys.clear();
            }
        }
        lk.unlock();
        // completion after everything already queued on the stream (also on failure:
        // earlier copies or kernels may still be reading the request's blocks)
        DoneMsg *m = new DoneMsg{s, pick};
        hipStream_t st = s->streams[r.stream];
        if (hipLaunchHostFunc(st, stream_host_done, m) != hipSuccess) {
            (void)hipStreamSynchronize(st);
            stream_host_done(m);
        }
        lk.lock();
    }
}

void completer(Server *s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_done.wait(lk, [&] { return !s->finished.empty() || (s->stop && s->inflight == 0); });
        if (s->finished.empty()) return;
        std::shared_ptr<Request> p = s->finished.front();
        s->finished.pop_front();
        Request &r = *p;
        if (!r.rc && !r.cancel && !r.y_pinned) {
            lk.unlock();
            memcpy(r.y, (char *)s->hmem + r.off_sy, r.y_bytes);
            lk.lock();
        }
        release(s, r);
        r.t_done = s->now();
        if (r.cancel) {
            r.state = FPCC_REQ_CANCELLED;
            s->n_cancel++;
        } else if (r.rc) {
            r.state = FPCC_REQ_FAILED;
            s->n_fail++;
        } else {
            r.state = FPCC_REQ_DONE;
            s->n_done++;
            const double ms = (r.t_done - r.t_submit) * 1e3;
            if (s->lat_ms.size() < kLatencyKeep) s->lat_ms.push_back(ms);
            else s->lat_ms[s->lat_next] = ms;
            s->lat_next = (s->lat_next + 1) % kLatencyKeep;
            s->lat_sum += ms;
            s->lat_max = std::max(s->lat_max, ms);
            s->pair_corr_done += r.pair_corr;
            s->last_done = r.t_done;
        }
        s->cv_sched.notify_one();
        s->cv_wait.notify_all();
    }
}

void fill_status(const Request &r, fpcc_request_status *st)
{
    if (!st) return;
    st->state = r.state;
    st->rc = r.rc;
    st->t_submit = r.t_submit;
    st->t_start = r.t_start;
    st->t_done = r.t_done;
}

bool final_state(int s) { return s == FPCC_REQ_DONE || s == FPCC_REQ_FAILED || s == FPCC_REQ_CANCELLED; }

// v sorted
double percentile(const std::vector<double> &v, double q)
{
    if (v.empty()) return 0;
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace
}  // namespace fpcc

using fpcc::Server;

struct fpcc_server {
    Server impl;
};

extern "C" int fpcc_server_create(const fpcc_server_config *cfg, fpcc_server **out)
{
    if (!out) return -1;
    *out = nullptr;
    auto *h = new fpcc_server;
    Server *s = &h->impl;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 16;
    if (s->cfg.n_streams <= 0) s->cfg.n_streams = 4;
    if (s->cfg.max_bypass <= 0) s->cfg.max_bypass = 8;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete h;
        return -1;
    }
    size_t dfree = 0, dtotal = 0;
    if (hipMemGetInfo(&dfree, &dtotal) != hipSuccess) {
        delete h;
        return -1;
    }
    if (!s->cfg.device_bytes) s->cfg.device_bytes = dfree / 10 * 4;
    if (!s->cfg.pinned_bytes) s->cfg.pinned_bytes = (size_t)512 << 20;
    if (s->cfg.device_bytes > dfree) {
        fprintf(stderr, "fastpcc: server arena of %zu bytes exceeds free device memory (%zu)\n", s->cfg.device_bytes,
                dfree);
        delete h;
        return -1;
    }
    if (hipMalloc(&s->dmem, s->cfg.device_bytes) != hipSuccess) {
        delete h;
        return -1;
    }
    if (hipHostMalloc(&s->hmem, s->cfg.pinned_bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(s->dmem);
        delete h;
        return -1;
    }
    if (s->cfg.cache_bytes && hipMalloc(&s->cmem, s->cfg.cache_bytes) != hipSuccess) {
        (void)hipHostFree(s->hmem);
        (void)hipFree(s->dmem);
        delete h;
        return -1;
    }
    s->darena.reset(s->cfg.device_bytes);
    s->harena.reset(s->cfg.pinned_bytes);
    s->carena.reset(s->cfg.cache_bytes);
    s->streams.resize(s->cfg.n_streams);
    s->stream_load.assign(s->cfg.n_streams, 0);
    for (auto &st : s->streams) {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (auto &t : s->streams)
                if (t) (void)hipStreamDestroy(t);
            if (s->cmem) (void)hipFree(s->cmem);
            (void)hipHostFree(s->hmem);
            (void)hipFree(s->dmem);
            delete h;
            return -1;
        }
    }
    s->t0 = fpcc::clock_t_::now();
    s->sched = std::thread(fpcc::scheduler, s);
    s->compl_ = std::thread(fpcc::completer, s);
    *out = h;
    return 0;
}

namespace fpcc {
namespace {
int make_request(Server *s, const fpcc_request *req, std::shared_ptr<Request> &out);
int enqueue(Server *s, std::shared_ptr<Request> &r, unsigned long long *id);
}  // namespace
}  // namespace fpcc

extern "C" int fpcc_server_submit(fpcc_server *h, const fpcc_request *req, unsigned long long *id)
{
    if (!h || !req || !req->x || !req->y) return -1;
    Server *s = &h->impl;
    std::shared_ptr<fpcc::Request> r;
    int rc = fpcc::make_request(s, req, r);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    return fpcc::enqueue(s, r, id);
}

extern "C" int fpcc_server_submit_many(fpcc_server *h, const fpcc_request *reqs, int n, unsigned long long *ids)
{
    if (!h || n < 0 || (n && (!reqs || !ids))) return -1;
    Server *s = &h->impl;
    std::vector<std::shared_ptr<fpcc::Request>> rs(n);
    for (int k = 0; k < n; ++k) {
        if (!reqs[k].x || !reqs[k].y) return -1;
        int rc = fpcc::make_request(s, &reqs[k], rs[k]);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(s->mu);  // all enter the queue together
    for (int k = 0; k < n; ++k) {
        int rc = fpcc::enqueue(s, rs[k], &ids[k]);
        if (rc) {
            for (int q = k; q < n; ++q) ids[q] = 0;
            return rc;
        }
    }
    return 0;
}

namespace fpcc {
namespace {

int make_request(Server *s, const fpcc_request *req, std::shared_ptr<Request> &out)
{
    (void)s;
    const fpcc_job &j = req->job;
    const size_t ylen = fpcc_output_len(&j);
    if (!ylen) {
        // resolve's own code: -3 for lags beyond the window, else -1
        const unsigned N = j.win_len;
        if (N >= 2 && ((unsigned)abs(j.lag1) >= N || (unsigned)abs(j.lag2) >= N)) return -3;
        return -1;
    }
    auto r = std::make_shared<Request>();
    r->job = j;
    r->dev_job = j;
    r->dev_job.station_stride = (unsigned long long)j.n_windows * j.win_len;
    r->x = req->x;
    r->y = req->y;
    r->x_bytes = (size_t)j.n_stations * r->dev_job.station_stride * sizeof(float);
    r->y_bytes = ylen * sizeof(float);
    r->ws_bytes = fpcc_workspace_size(&r->dev_job);
    const size_t last = (size_t)(j.n_stations - 1) * j.station_stride + (size_t)j.n_windows * j.win_len - 1;
    r->x_pinned = fpcc::is_pinned(req->x) && fpcc::is_pinned(req->x + last);
    r->y_pinned = fpcc::is_pinned(req->y) && fpcc::is_pinned(req->y + ylen - 1);
    r->pair_corr = (unsigned long long)j.n_stations * (j.n_stations - 1) / 2 * j.n_windows;
    r->batchable = !req->station_keys && fpcc::job_batch_key(&r->dev_job, &r->batch_key);
    if (req->station_keys) {
        r->keys.assign(req->station_keys, req->station_keys + j.n_stations);
        r->nreg = fpcc::job_station_regions(&r->dev_job, r->off, r->per);
        size_t base = 0;
        for (int k = 0; k < r->nreg; ++k) {
            r->rbase[k] = base;
            base += (r->per[k] + 15) / 16 * 16;
        }
        r->entry_bytes = base;
    }
    out = r;
    return 0;
}

int enqueue(Server *s, std::shared_ptr<Request> &r, unsigned long long *id)
{
    const size_t dfoot = Arena::round(r->x_bytes) + Arena::round(r->y_bytes) + Arena::round(r->ws_bytes);
    const size_t hfoot = (r->x_pinned ? 0 : Arena::round(r->x_bytes)) + (r->y_pinned ? 0 : Arena::round(r->y_bytes));
    if (s->stop) return -1;
    if (dfoot > s->darena.size() || hfoot > s->harena.size()) {
        s->n_reject++;
        return -4;  // can never be admitted under this server's budget
    }
    r->id = s->next_id++;
    r->t_submit = s->now();
    s->all[r->id] = r;
    s->queue.push_back(r);
    s->n_submit++;
    if (id) *id = r->id;
    s->cv_sched.notify_one();
    return 0;
}

}  // namespace
}  // namespace fpcc

extern "C" int fpcc_server_poll(fpcc_server *h, unsigned long long id, fpcc_request_status *st)
{
    if (!h) return -1;
    Server *s = &h->impl;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return -1;
    fpcc::fill_status(*it->second, st);
    return 0;
}

extern "C" int fpcc_server_wait(fpcc_server *h, unsigned long long id, int timeout_ms, fpcc_request_status *st)
{
    if (!h) return -1;
    Server *s = &h->impl;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return -1;
    std::shared_ptr<fpcc::Request> r = it->second;
    auto done = [&] { return fpcc::final_state(r->state); };
    if (timeout_ms < 0)
        s->cv_wait.wait(lk, done);
    else if (!s->cv_wait.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
        fpcc::fill_status(*r, st);
        return 1;  // timed out, request still pending
    }
    fpcc::fill_status(*r, st);
    return 0;
}

extern "C" int fpcc_server_cancel(fpcc_server *h, unsigned long long id)
{
    if (!h) return -1;
    Server *s = &h->impl;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return -1;
    fpcc::Request &r = *it->second;
    if (fpcc::final_state(r.state)) return 1;
    if (r.state == FPCC_REQ_QUEUED) {
        for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
            if ((*q)->id == id) {
                s->queue.erase(q);
                break;
            }
        r.state = FPCC_REQ_CANCELLED;
        r.t_done = s->now();
        s->n_cancel++;
        s->cv_wait.notify_all();
        s->cv_sched.notify_one();
        return 0;
    }
    r.cancel = true;  // running: its result is dropped when the stream reaches it
    return 0;
}

extern "C" int fpcc_server_forget(fpcc_server *h, unsigned long long id)
{
    if (!h) return -1;
    Server *s = &h->impl;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end() || !fpcc::final_state(it->second->state)) return -1;
    s->all.erase(it);
    return 0;
}

extern "C" int fpcc_server_stats(fpcc_server *h, fpcc_serving_stats *out)
{
    if (!h || !out) return -1;
    Server *s = &h->impl;
    std::vector<double> lat;  // copied under the lock, sorted outside it
    std::unique_lock<std::mutex> lk(s->mu);
    lat = s->lat_ms;
    memset(out, 0, sizeof(*out));
    out->submitted = s->n_submit;
    out->completed = s->n_done;
    out->failed = s->n_fail;
    out->cancelled = s->n_cancel;
    out->rejected = s->n_reject;
    out->cache_hits = s->n_hit;
    out->cache_misses = s->n_miss;
    out->cache_used = s->carena.used();
    out->batched_requests = s->n_batched;
    out->batch_launches = s->n_batch_launch;
    out->queued = s->queue.size();
    out->inflight = s->inflight;
    out->inflight_peak = s->inflight_peak;
    out->max_ms = s->lat_max;
    out->mean_ms = s->n_done ? s->lat_sum / (double)s->n_done : 0;
    out->pair_correlations = (double)s->pair_corr_done;
    out->busy_s = s->first_start < 0 ? 0 : s->last_done - s->first_start;
    out->device_bytes = s->darena.size();
    out->device_used = s->darena.used();
    out->device_peak = s->dev_peak;
    lk.unlock();
    std::sort(lat.begin(), lat.end());
    out->p50_ms = fpcc::percentile(lat, 0.50);
    out->p99_ms = fpcc::percentile(lat, 0.99);
    return 0;
}

extern "C" int fpcc_server_destroy(fpcc_server *h)
{
    if (!h) return -1;
    Server *s = &h->impl;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &r : s->queue) {
            r->state = FPCC_REQ_CANCELLED;
            s->n_cancel++;
        }
        s->queue.clear();
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->sched.join();  // no more launches after this
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_wait.wait(lk, [&] { return s->inflight == 0; });
    }
    s->cv_done.notify_all();
    s->compl_.join();
    (void)hipSetDevice(s->device);
    for (auto &st : s->streams) (void)hipStreamDestroy(st);
    s->cache.clear();
    if (s->cmem) (void)hipFree(s->cmem);
    (void)hipHostFree(s->hmem);
    (void)hipFree(s->dmem);
    delete h;
    return 0;
}

extern "C" void *fpcc_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (!bytes || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

extern "C" void fpcc_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}
