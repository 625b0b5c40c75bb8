// C ABI of libfastpcc (declared in include/fastpcc.h).
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "../../include/fastpcc.h"
#include "engine.hpp"
#include "kernels.hpp"

using namespace fpcc;

namespace {

struct LagWin {
    int L, lag0;
    unsigned M;
};

// Lag window of pcc2_set/ccgn_set/cc1b_set (src/FFTapps.c:790-804)
int lag_window(int Lag1, int Lag2, unsigned N, LagWin &lw)
{
    if (Lag2 >= Lag1) {
        lw.L = Lag2 - Lag1 + 1;
        lw.lag0 = Lag1;
    } else {
        lw.L = Lag1 - Lag2 + 1;
        lw.lag0 = Lag2;
    }
    unsigned ua1 = (unsigned)abs(Lag1), ua2 = (unsigned)abs(Lag2);
    if (ua1 >= N || ua2 >= N) return -3;
    lw.M = ua1 > ua2 ? ua1 : ua2;
    return 0;
}

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

int validate_job(const fpcc_job *job, LagWin &lw, unsigned &Nz)
{
    if (!job || job->n_stations < 2 || job->n_windows < 1 || job->win_len < 2) return -1;
    if (job->station_stride < (unsigned long long)job->n_windows * job->win_len) return -1;
    if (job->lag1 > job->lag2) return -1;
    int r = lag_window(job->lag1, job->lag2, job->win_len, lw);
    if (r) return r;
    Nz = nz_rule(job->win_len, lw.M);
    if (job->kind != FPCC_PCC2) return -1;
    if (job->stack < FPCC_STACK_NONE || job->stack > FPCC_STACK_GRAM) return -1;
    return 0;
}

// Workspace layout: [F: S*W*Nz spectra][U: (Nz/2+1)*S*2W][G: P*ld], 256-B aligned.
struct WsLayout {
    size_t F, U, G, total;
};
WsLayout ws_layout(const fpcc_job *job, unsigned Nz)
{
    WsLayout l{};
    const size_t S = job->n_stations, W = job->n_windows, P = S * (S - 1) / 2;
    l.F = 0;
    size_t f_bytes = align_up(sizeof(cf32) * S * W * Nz, 256);
    l.total = f_bytes;
    if (job->stack == FPCC_STACK_GRAM) {
        l.U = l.total;
        l.total += align_up(sizeof(cf32) * (Nz / 2 + 1) * S * 2 * W, 256);
        l.G = l.total;
        l.total += align_up(sizeof(cf32) * P * (size_t)gram_ld(Nz), 256);
    }
    return l;
}

}  // namespace

extern "C" size_t fpcc_output_len(const fpcc_job *job)
{
    LagWin lw;
    unsigned Nz;
    if (validate_job(job, lw, Nz)) return 0;
    size_t P = (size_t)job->n_stations * (job->n_stations - 1) / 2;
    return P * (job->stack == FPCC_STACK_NONE ? job->n_windows : 1) * (size_t)lw.L;
}

extern "C" size_t fpcc_workspace_size(const fpcc_job *job)
{
    LagWin lw;
    unsigned Nz;
    if (validate_job(job, lw, Nz)) return 0;
    return ws_layout(job, Nz).total;
}

extern "C" int fpcc_station_stage_device(const fpcc_job *job, const float *d_x, void *d_ws, size_t ws_bytes,
                                         void *stream)
{
    LagWin lw;
    unsigned Nz;
    int r = validate_job(job, lw, Nz);
    if (r) return r;
    if (!d_x || !d_ws || ws_bytes < fpcc_workspace_size(job)) return -1;
    return launch_station_pcc(d_x, (long long)job->station_stride, job->n_stations, job->n_windows, job->win_len,
                              Nz, (cf32 *)d_ws, (hipStream_t)stream);
}

extern "C" int fpcc_pair_stage_device(const fpcc_job *job, void *d_ws, size_t ws_bytes, float *d_y, void *stream)
{
    LagWin lw;
    unsigned Nz;
    int r = validate_job(job, lw, Nz);
    if (r) return r;
    if (!d_y || !d_ws || ws_bytes < fpcc_workspace_size(job)) return -1;
    const int S = job->n_stations, P = S * (S - 1) / 2, W = job->n_windows, N = job->win_len;
    hipStream_t st = (hipStream_t)stream;
    if (job->stack == FPCC_STACK_GRAM) {
        WsLayout l = ws_layout(job, Nz);
        char *ws = (char *)d_ws;
        cf32 *F = (cf32 *)(ws + l.F), *U = (cf32 *)(ws + l.U), *G = (cf32 *)(ws + l.G);
        int ntiles = 0;
        const int2 *tiles = gram_tiles(S, ntiles);
        if (!tiles) return -1;
        if ((r = launch_spectra_to_u(F, S, W, Nz, U, st))) return r;
        if ((r = launch_gram_f32(U, S, W, Nz, tiles, ntiles, G, st))) return r;
        return launch_pair_c2r(G, P, Nz, N, W, lw.lag0, lw.L, d_y, st);
    }
    return launch_pair_pcc2((const cf32 *)d_ws, S, W, N, Nz, lw.lag0, lw.L, job->stack, 0, P, d_y, st);
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
    size_t ylen = fpcc_output_len(job);
    if (!ylen || !h_x || !h_y) {
        LagWin lw;
        unsigned Nz;
        int r = job ? validate_job(job, lw, Nz) : -1;
        return r ? r : -1;
    }
    size_t ws = fpcc_workspace_size(job);
    size_t xlen = (size_t)(job->n_stations - 1) * job->station_stride + (size_t)job->n_windows * job->win_len;
    DeviceBuf dx, dy, dws;
    FPCC_CHECK(hipMalloc(&dx.p, sizeof(float) * xlen));
    FPCC_CHECK(hipMalloc(&dy.p, sizeof(float) * ylen));
    FPCC_CHECK(hipMalloc(&dws.p, ws));
    FPCC_CHECK(hipMemcpy(dx.p, h_x, sizeof(float) * xlen, hipMemcpyHostToDevice));
    int r = fpcc_allpairs_device(job, (const float *)dx.p, (float *)dy.p, dws.p, ws, nullptr);
    if (r) return r;
    FPCC_CHECK(hipMemcpy(h_y, dy.p, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    return 0;
}

// ---- reference-compatible entries (src/FFTapps.h) ----

extern "C" int pcc2_set(float **const y, float **const x1, float **const x2, const unsigned int N,
                        const unsigned int Tr, const int Lag1, const int Lag2)
{
    LagWin lw;
    int r = lag_window(Lag1, Lag2, N, lw);
    if (r) return r;
    if (Tr == 0) return 0;
    if (!y || !x1 || !x2) return -1;
    const unsigned Nz = nz_rule(N, lw.M);
    // stations 0 (= x1 windows) and 1 (= x2 windows), one pair, per-window output
    const size_t xlen = (size_t)2 * Tr * N, ylen = (size_t)Tr * lw.L;
    HostBuf hx, hy;
    DeviceBuf dx, dy, dF;
    FPCC_CHECK(hipHostMalloc(&hx.p, sizeof(float) * xlen, hipHostMallocDefault));
    FPCC_CHECK(hipHostMalloc(&hy.p, sizeof(float) * ylen, hipHostMallocDefault));
    for (unsigned tr = 0; tr < Tr; ++tr) {
        if (!x1[tr] || !x2[tr] || !y[tr]) return -1;
        memcpy((float *)hx.p + (size_t)tr * N, x1[tr], sizeof(float) * N);
        memcpy((float *)hx.p + (size_t)(Tr + tr) * N, x2[tr], sizeof(float) * N);
    }
    FPCC_CHECK(hipMalloc(&dx.p, sizeof(float) * xlen));
    FPCC_CHECK(hipMalloc(&dy.p, sizeof(float) * ylen));
    FPCC_CHECK(hipMalloc(&dF.p, sizeof(cf32) * 2 * Tr * (size_t)Nz));
    FPCC_CHECK(hipMemcpy(dx.p, hx.p, sizeof(float) * xlen, hipMemcpyHostToDevice));
    r = launch_station_pcc((const float *)dx.p, (long long)Tr * N, 2, Tr, N, Nz, (cf32 *)dF.p, nullptr);
    if (r) return r;
    r = launch_pair_pcc2((const cf32 *)dF.p, 2, Tr, N, Nz, lw.lag0, lw.L, 0, 0, 1, (float *)dy.p, nullptr);
    if (r) return r;
    FPCC_CHECK(hipMemcpy(hy.p, dy.p, sizeof(float) * ylen, hipMemcpyDeviceToHost));
    for (unsigned tr = 0; tr < Tr; ++tr) memcpy(y[tr], (float *)hy.p + (size_t)tr * lw.L, sizeof(float) * lw.L);
    return 0;
}
