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
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include "../../include/fastpcc.h"
#include "arena.hpp"
#include "engine.hpp"

namespace fpcc {

namespace {

using clock_t_ = std::chrono::steady_clock;

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
};

struct Server;

struct DoneMsg {
    Server *srv;
    std::shared_ptr<Request> req;
};

struct Server {
    fpcc_server_config cfg{};
    int device = 0;
    void *dmem = nullptr, *hmem = nullptr;
    Arena darena, harena;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;
    clock_t_::time_point t0;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_wait;
    std::deque<std::shared_ptr<Request>> queue;
    std::deque<std::shared_ptr<Request>> finished;  // handed over by stream host functions
    std::unordered_map<unsigned long long, std::shared_ptr<Request>> all;
    unsigned long long next_id = 1;
    int inflight = 0, inflight_peak = 0;
    size_t dev_peak = 0;
    bool stop = false;
    unsigned long long admitted = 0;
    // statistics
    unsigned long long n_submit = 0, n_done = 0, n_fail = 0, n_cancel = 0, n_reject = 0;
    std::vector<double> lat_ms;
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

// Launches one admitted request on its stream (scheduler thread, lock not held).
int issue(Server *s, Request &r)
{
    (void)hipSetDevice(s->device);
    hipStream_t st = s->streams[r.stream];
    char *dbase = (char *)s->dmem, *hbase = (char *)s->hmem;
    float *dx = (float *)(dbase + r.off_x), *dy = (float *)(dbase + r.off_y);
    const fpcc_job &j = r.job;
    const size_t row = (size_t)j.n_windows * j.win_len * sizeof(float);
    if (s->cfg.fail_every > 0 && r.id % (unsigned long long)s->cfg.fail_every == 0) return -1;  // injected
    if (r.x_pinned) {
        if (hipMemcpy2DAsync(dx, row, r.x, j.station_stride * sizeof(float), row, j.n_stations,
                             hipMemcpyHostToDevice, st) != hipSuccess)
            return -1;
    } else {
        char *hx = hbase + r.off_sx;
        for (unsigned i = 0; i < j.n_stations; ++i) memcpy(hx + i * row, r.x + (size_t)i * j.station_stride, row);
        if (hipMemcpyAsync(dx, hx, r.x_bytes, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    }
    int rc = fpcc_allpairs_device(&r.dev_job, dx, dy, dbase + r.off_ws, r.ws_bytes, st);
    if (rc) return rc;
    void *yh = r.y_pinned ? (void *)r.y : (void *)(hbase + r.off_sy);
    if (hipMemcpyAsync(yh, dy, r.y_bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    return 0;
}

void release(Server *s, Request &r)
{
    s->darena.release(r.off_x, r.x_bytes);
    s->darena.release(r.off_y, r.y_bytes);
    s->darena.release(r.off_ws, r.ws_bytes);
    if (!r.x_pinned) s->harena.release(r.off_sx, r.x_bytes);
    if (!r.y_pinned) s->harena.release(r.off_sy, r.y_bytes);
    s->stream_load[r.stream]--;
    s->inflight--;
}

// Reserves the request's blocks if they all fit (lock held).
bool try_reserve(Server *s, Request &r)
{
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
        Request &r = *pick;
        int best = 0;
        for (int i = 1; i < (int)s->streams.size(); ++i)
            if (s->stream_load[i] < s->stream_load[best]) best = i;
        r.stream = best;
        s->stream_load[best]++;
        s->inflight++;
        s->inflight_peak = std::max(s->inflight_peak, s->inflight);
        s->dev_peak = std::max(s->dev_peak, s->darena.used());
        r.state = FPCC_REQ_RUNNING;
        r.t_start = s->now();
        if (s->first_start < 0) s->first_start = r.t_start;
        s->admitted++;
        lk.unlock();
        int rc = issue(s, r);
        DoneMsg *m = new DoneMsg{s, pick};
        hipStream_t st = s->streams[r.stream];
        if (rc) r.rc = rc;
        // completion after everything already queued on the stream (also on failure:
        // earlier copies or kernels may still be reading the request's blocks)
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
            s->lat_ms.push_back((r.t_done - r.t_submit) * 1e3);
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

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
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
    s->darena.reset(s->cfg.device_bytes);
    s->harena.reset(s->cfg.pinned_bytes);
    s->streams.resize(s->cfg.n_streams);
    s->stream_load.assign(s->cfg.n_streams, 0);
    for (auto &st : s->streams) {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (auto &t : s->streams)
                if (t) (void)hipStreamDestroy(t);
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

extern "C" int fpcc_server_submit(fpcc_server *h, const fpcc_request *req, unsigned long long *id)
{
    if (!h || !req || !req->x || !req->y) return -1;
    Server *s = &h->impl;
    const fpcc_job &j = req->job;
    const size_t ylen = fpcc_output_len(&j);
    if (!ylen) {
        // resolve's own code: -3 for lags beyond the window, else -1
        const unsigned N = j.win_len;
        if (N >= 2 && ((unsigned)abs(j.lag1) >= N || (unsigned)abs(j.lag2) >= N)) return -3;
        return -1;
    }
    auto r = std::make_shared<fpcc::Request>();
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
    const size_t dfoot = fpcc::Arena::round(r->x_bytes) + fpcc::Arena::round(r->y_bytes) +
                         fpcc::Arena::round(r->ws_bytes);
    const size_t hfoot = (r->x_pinned ? 0 : fpcc::Arena::round(r->x_bytes)) +
                         (r->y_pinned ? 0 : fpcc::Arena::round(r->y_bytes));
    std::lock_guard<std::mutex> lk(s->mu);
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
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = s->n_submit;
    out->completed = s->n_done;
    out->failed = s->n_fail;
    out->cancelled = s->n_cancel;
    out->rejected = s->n_reject;
    out->queued = s->queue.size();
    out->inflight = s->inflight;
    out->inflight_peak = s->inflight_peak;
    out->p50_ms = fpcc::percentile(s->lat_ms, 0.50);
    out->p99_ms = fpcc::percentile(s->lat_ms, 0.99);
    out->max_ms = s->lat_ms.empty() ? 0 : *std::max_element(s->lat_ms.begin(), s->lat_ms.end());
    double sum = 0;
    for (double v : s->lat_ms) sum += v;
    out->mean_ms = s->lat_ms.empty() ? 0 : sum / s->lat_ms.size();
    out->pair_correlations = (double)s->pair_corr_done;
    out->busy_s = s->first_start < 0 ? 0 : s->last_done - s->first_start;
    out->device_bytes = s->darena.size();
    out->device_used = s->darena.used();
    out->device_peak = s->dev_peak;
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
