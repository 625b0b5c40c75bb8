// HIP-graph replay of one job's whole pipeline (no counterpart in the reference).
// fpcc_graph_create runs the job once eagerly (building every plan, twiddle table,
// tile list and filter bank, so nothing allocates or synchronises later), then
// captures fpcc_allpairs_device on the caller's stream (thread-local capture
// mode: other threads' HIP calls are not affected) and instantiates the graph.
// The pipeline clears buffers with fill kernels, never with memset nodes.  Each
// fpcc_graph_launch recomputes y from the current contents of d_x.
#include <cstdio>
#include "../../include/fastpcc.h"
#include "engine.hpp"
#include "profile.hpp"

struct fpcc_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
};

extern "C" int fpcc_graph_create(const fpcc_job *job, const float *d_x, float *d_y, void *d_ws, size_t ws_bytes,
                                 void *stream, fpcc_graph **graph)
{
    if (!graph) return -1;
    *graph = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if (!st) return -1;  // the default stream cannot be captured
    int r = fpcc_allpairs_device(job, d_x, d_y, d_ws, ws_bytes, st);
    if (r) return r;
    FPCC_CHECK(hipStreamSynchronize(st));
    const bool prof = fpcc::prof_on();
    fpcc_profile_enable(0);  // no timing events inside the graph
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        r = fpcc_allpairs_device(job, d_x, d_y, d_ws, ws_bytes, st);
        e = hipStreamEndCapture(st, &g);
    }
    if (prof) fpcc_profile_enable(1);
    if (r || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        fprintf(stderr, "fastpcc: graph capture failed (%d, %s)\n", r, hipGetErrorString(e));
        return r ? r : -1;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return -1;
    }
    *graph = new fpcc_graph{g, ex};
    return 0;
}

extern "C" int fpcc_graph_launch(fpcc_graph *graph, void *stream)
{
    if (!graph) return -1;
    FPCC_CHECK(hipGraphLaunch(graph->e, (hipStream_t)stream));
    return 0;
}

extern "C" int fpcc_graph_destroy(fpcc_graph *graph)
{
    if (!graph) return -1;
    if (graph->e) (void)hipGraphExecDestroy(graph->e);
    if (graph->g) (void)hipGraphDestroy(graph->g);
    delete graph;
    return 0;
}
