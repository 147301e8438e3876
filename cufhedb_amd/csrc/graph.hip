// graph.hip -- HIP graph capture of pipelines of _dev calls on a context.
//
// The reference drives every query row through host calls, one
// cudaStreamSynchronize per bootstrap (src/cuHEDB/fft_gpu/cufft_gpu.cuh:262).
// Here a whole pipeline of batched entry points (a comparison chain, a query
// filter) can be recorded once from the context stream and replayed as one
// graph launch, with the inputs and outputs bound to the device buffers used
// during capture.  What a captured call may not do is enforced: scratch
// growth fails with FHE_ERR_ARG while a capture is open (run the pipeline
// once at its size first), and the entry points a pipeline uses enqueue
// kernels only (no memset or copy nodes, no host synchronisation).
#include "ctx.hpp"

namespace fhe {
thread_local bool g_capturing = false;
}

using namespace fhe;

struct fhe_graph {
    fhe_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int fhe_graph_begin(fhe_ctx *ctx)
{
    FHE_CHECK_ARG(ctx, "fhe_graph_begin: null context");
    FHE_CHECK_ARG(!g_capturing, "fhe_graph_begin: a capture is already open on this thread");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
            "hipStreamBeginCapture");
    g_capturing = true;
    return FHE_OK;
}

int fhe_graph_end(fhe_ctx *ctx, fhe_graph **out)
{
    FHE_CHECK_ARG(ctx && out, "fhe_graph_end: null argument");
    FHE_CHECK_ARG(g_capturing, "fhe_graph_end: no capture open on this thread");
    g_capturing = false;
    *out = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(ctx->stream, &g), "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    *out = new fhe_graph{ctx, g, x};
    return FHE_OK;
}

int fhe_graph_launch(fhe_graph *g, void *stream)
{
    FHE_CHECK_ARG(g, "fhe_graph_launch: null graph");
    HIP_TRY(hipSetDevice(g->ctx->device), "hipSetDevice");
    HIP_TRY(hipGraphLaunch(g->exec, pick(g->ctx, stream)), "hipGraphLaunch");
    return FHE_OK;
}

int fhe_graph_destroy(fhe_graph *g)
{
    if (!g) return FHE_OK;
    (void)hipSetDevice(g->ctx->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return FHE_OK;
}

}  // extern "C"
