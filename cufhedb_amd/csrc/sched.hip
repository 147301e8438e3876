// sched.hip -- coalescing scheduler for a stream of small bootstrap requests.
//
// The mixed-parameter workload is a stream of independent requests: lvl1
// LUT bootstraps (a batch of lvl0 ciphertexts with one test vector) and lvl2
// comparisons (a batch against one threshold ciphertext).  The reference
// serves such a stream with one host thread per CUDA stream and one
// ciphertext per call (src/cuHEDB/gatebootstrap_gpu.cuh:17-25,
// test/HE3_Query/TPCH_Q6.cu:179-193).  A request of a few thousand rows
// fills only part of the 256 CUs (the lvl1 kernel holds 8192 ciphertexts at
// once), so here a flush merges consecutive requests of one kind into
// launches of up to max_rows rows: their inputs are gathered into device
// staging buffers, the per-request test vector / threshold is replicated per
// row (the kernels take one per ciphertext), one batched call runs, and the
// results are scattered back.  A request of more than max_rows rows runs
// alone, straight on its own buffers.  lvl1 launches go to the context's
// stream and lvl2 comparisons to a second stream of the scheduler (default
// priority), so a lvl1 batch and a lvl2 comparison share the CUs (each
// kind in submission order; the second stream starts after the work
// already on the context stream, and the context stream continues after
// both).  The two kinds use disjoint staging and scratch: a comparison's
// chain scratch is the context's, the lvl1 bootstrap only reads the keys,
// whose lazily built layouts are built before the streams fork.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.hpp"

namespace fhe {
namespace {

template <class T>
__global__ void k_replicate(T *__restrict__ dst, const T *__restrict__ src, int W, size_t total)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < total) dst[i] = src[i % size_t(W)];
}

template <class T>
hipError_t replicate(T *dst, const T *src, int W, int rows, hipStream_t s)
{
    const size_t total = size_t(W) * rows;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_replicate<T>, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, dst,
                       src, W, total);
    return hipGetLastError();
}

struct Req {
    int kind;
    const void *in;   // lvl0 TLWEs (PBS1) / lvl2 TLWEs a (CMP2)
    const void *aux;  // test vector (PBS1) / threshold TLWE b (CMP2)
    uint32_t *out;    // lvl1 TLWEs
    int B;
    int op, result_type;
    uint32_t bits;
    uint64_t ticket;
    hipEvent_t start = nullptr, done = nullptr;  // start: the flush that serves it began
    bool flushed = false;
    bool failed = false;  // its flush stopped on an error before its result was complete
};

#define SCHED_TRY(expr)                    \
    do {                                   \
        if (int _rc = (expr)) return _rc;  \
    } while (0)

constexpr int W1 = Lvl1::N + 1;
constexpr int W2 = Lvl2::N + 1;

}  // namespace
}  // namespace fhe

struct fhe_sched {
    fhe_ctx *ctx = nullptr;
    int max_rows = 0;
    uint64_t next_ticket = 1;
    std::vector<fhe::Req> reqs;  // submitted, not yet waited for
    // staging, allocated on first use of each kind
    uint32_t *in1 = nullptr, *tv1 = nullptr, *out1 = nullptr;
    uint64_t *a2 = nullptr, *b2 = nullptr;
    uint32_t *outc = nullptr;
    hipStream_t s2 = nullptr;  // lvl2 comparisons
    hipEvent_t fork = nullptr, join = nullptr;
};

using namespace fhe;

namespace {

int ensure_staging(fhe_sched *s, int kind)
{
    const size_t R = size_t(s->max_rows);
    if (kind == FHE_SCHED_PBS1 && !s->in1) {
        HIP_TRY(hipMalloc(&s->in1, R * (s->ctx->p.n0 + 1) * 4), "scheduler staging");
        HIP_TRY(hipMalloc(&s->tv1, R * Lvl1::N * 4), "scheduler staging");
        HIP_TRY(hipMalloc(&s->out1, R * W1 * 4), "scheduler staging");
    }
    if (kind == FHE_SCHED_CMP2 && !s->a2) {
        HIP_TRY(hipMalloc(&s->a2, R * W2 * 8), "scheduler staging");
        HIP_TRY(hipMalloc(&s->b2, R * W2 * 8), "scheduler staging");
        HIP_TRY(hipMalloc(&s->outc, R * W1 * 4), "scheduler staging");
    }
    return FHE_OK;
}

bool same_group(const Req &a, const Req &b)
{
    return a.kind == b.kind &&
           (a.kind == FHE_SCHED_PBS1 ||
            (a.op == b.op && a.bits == b.bits && a.result_type == b.result_type));
}

// one batched call over rows [0, rows) of the staging buffers
int run_group(fhe_sched *s, const std::vector<Req *> &g, int rows, hipStream_t st)
{
    fhe_ctx *ctx = s->ctx;
    const int kind = g[0]->kind;
    SCHED_TRY(ensure_staging(s, kind));
    const size_t w0 = size_t(ctx->p.n0) + 1;
    size_t off = 0;
    for (Req *r : g) {
        if (kind == FHE_SCHED_PBS1) {
            HIP_TRY(hipMemcpyAsync(s->in1 + off * w0, r->in, size_t(r->B) * w0 * 4,
                                   hipMemcpyDeviceToDevice, st),
                    "scheduler gather");
            HIP_TRY(replicate(s->tv1 + off * Lvl1::N, static_cast<const uint32_t *>(r->aux),
                              Lvl1::N, r->B, st),
                    "scheduler test vectors");
        } else {
            HIP_TRY(hipMemcpyAsync(s->a2 + off * W2, r->in, size_t(r->B) * W2 * 8,
                                   hipMemcpyDeviceToDevice, st),
                    "scheduler gather");
            HIP_TRY(replicate(s->b2 + off * W2, static_cast<const uint64_t *>(r->aux), W2, r->B,
                              st),
                    "scheduler thresholds");
        }
        off += size_t(r->B);
    }
    uint32_t *res;
    if (kind == FHE_SCHED_PBS1) {
        SCHED_TRY(fhe_pbs_lvl01_dev(ctx, s->in1, s->tv1, Lvl1::N, s->out1, rows, st));
        res = s->out1;
    } else {
        const Req &r0 = *g[0];
        SCHED_TRY(fhe_compare_dev(ctx, 2, r0.op, s->a2, s->b2, 0, s->outc, r0.bits,
                                   r0.result_type, rows, st));
        res = s->outc;
    }
    off = 0;
    for (Req *r : g) {
        HIP_TRY(hipMemcpyAsync(r->out, res + off * W1, size_t(r->B) * W1 * 4,
                               hipMemcpyDeviceToDevice, st),
                "scheduler scatter");
        HIP_TRY(hipEventRecord(r->done, st), "hipEventRecord");
        off += size_t(r->B);
    }
    return FHE_OK;
}

// a request too large for the staging buffers runs on its own buffers
int run_alone(fhe_sched *s, Req &r, hipStream_t st)
{
    fhe_ctx *ctx = s->ctx;
    if (r.kind == FHE_SCHED_PBS1)
        SCHED_TRY(fhe_pbs_lvl01_dev(ctx, static_cast<const uint32_t *>(r.in),
                                     static_cast<const uint32_t *>(r.aux), 0, r.out, r.B, st));
    else
        SCHED_TRY(fhe_compare_dev(ctx, 2, r.op, r.in, r.aux, 1, r.out, r.bits, r.result_type,
                                   r.B, st));
    HIP_TRY(hipEventRecord(r.done, st), "hipEventRecord");
    return FHE_OK;
}

// serves every request not yet flushed, each kind in order on its stream
int serve_pending(fhe_sched *s, hipStream_t st1, hipStream_t st2)
{
    auto stream_of = [&](int kind) { return kind == FHE_SCHED_PBS1 ? st1 : st2; };
    std::vector<Req *> group;
    int rows = 0;
    auto close_group = [&]() -> int {
        if (group.empty()) return FHE_OK;
        const int rc = run_group(s, group, rows, stream_of(group[0]->kind));
        group.clear();
        rows = 0;
        return rc;
    };
    for (Req &r : s->reqs) {
        if (r.flushed) continue;
        r.flushed = true;
        if (r.B > s->max_rows) {
            SCHED_TRY(close_group());
            SCHED_TRY(run_alone(s, r, stream_of(r.kind)));
            continue;
        }
        if (!group.empty() && (!same_group(*group[0], r) || rows + r.B > s->max_rows))
            SCHED_TRY(close_group());
        group.push_back(&r);
        rows += r.B;
    }
    return close_group();
}

int flush_pending(fhe_sched *s)
{
    const hipStream_t st1 = s->ctx->stream, st2 = s->s2;
    for (Req &r : s->reqs)
        if (!r.flushed) HIP_TRY(hipEventRecord(r.start, st1), "hipEventRecord");
    SCHED_TRY(prepare_pbs1(s->ctx, st1));
    HIP_TRY(hipEventRecord(s->fork, st1), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(st2, s->fork, 0), "hipStreamWaitEvent");
    const int rc = serve_pending(s, st1, st2);
    // the context stream continues after the comparisons, also after an error
    HIP_TRY(hipEventRecord(s->join, st2), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(st1, s->join, 0), "hipStreamWaitEvent");
    return rc;
}

int submit(fhe_sched *s, Req r, uint64_t *ticket)
{
    FHE_CHECK_ARG(s && ticket && r.in && r.aux && r.out && r.B >= 0, "fhe_sched_submit: bad argument");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    HIP_TRY(hipEventCreate(&r.start), "hipEventCreate");
    if (hipEventCreate(&r.done) != hipSuccess) {
        (void)hipEventDestroy(r.start);
        set_error("fhe_sched_submit: hipEventCreate failed");
        return FHE_ERR_DEVICE;
    }
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    s->reqs.push_back(r);
    return FHE_OK;
}

}  // namespace

extern "C" {

int fhe_sched_create(fhe_ctx *ctx, int max_rows, fhe_sched **out)
{
    FHE_CHECK_ARG(ctx && out && max_rows > 0, "fhe_sched_create: bad argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto *s = new fhe_sched;
    s->ctx = ctx;
    s->max_rows = max_rows;
    hipError_t e = hipStreamCreateWithFlags(&s->s2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (s->join) (void)hipEventDestroy(s->join);
        if (s->fork) (void)hipEventDestroy(s->fork);
        if (s->s2) (void)hipStreamDestroy(s->s2);
        delete s;
        return hip_fail(e, "fhe_sched_create");
    }
    *out = s;
    return FHE_OK;
}

int fhe_sched_destroy(fhe_sched *s)
{
    if (!s) return FHE_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipStreamSynchronize(s->s2);
    for (Req &r : s->reqs) {
        if (r.start) (void)hipEventDestroy(r.start);
        if (r.done) (void)hipEventDestroy(r.done);
    }
    void *bufs[] = {s->in1, s->tv1, s->out1, s->a2, s->b2, s->outc};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    (void)hipEventDestroy(s->fork);
    (void)hipEventDestroy(s->join);
    (void)hipStreamDestroy(s->s2);
    delete s;
    return FHE_OK;
}

int fhe_sched_submit_pbs1(fhe_sched *s, const uint32_t *in_dev, const uint32_t *tv_dev,
                          uint32_t *out_dev, int B, uint64_t *ticket)
{
    Req r{};
    r.kind = FHE_SCHED_PBS1;
    r.in = in_dev;
    r.aux = tv_dev;
    r.out = out_dev;
    r.B = B;
    return submit(s, r, ticket);
}

int fhe_sched_submit_cmp2(fhe_sched *s, int op, const uint64_t *a_dev, const uint64_t *b_dev,
                          uint32_t *out_dev, uint32_t plain_bits, int result_type, int B,
                          uint64_t *ticket)
{
    FHE_CHECK_ARG(op >= FHE_CMP_GT && op <= FHE_CMP_EQ, "fhe_sched_submit_cmp2: unknown comparison %d",
                  op);
    FHE_CHECK_ARG(plain_bits >= 1, "fhe_sched_submit_cmp2: plain_bits >= 1");
    if (plain_bits + 1 > uint32_t(kMaxMsbBits2)) {
        set_error("fhe_sched_submit_cmp2: plain_bits %u above %d: lvl2 comparisons of wider "
                  "values return wrong results (bootstrap noise reaches the decision margin)",
                  plain_bits, kMaxMsbBits2 - 1);
        return FHE_ERR_UNSUPPORTED;
    }
    Req r{};
    r.kind = FHE_SCHED_CMP2;
    r.in = a_dev;
    r.aux = b_dev;
    r.out = out_dev;
    r.B = B;
    r.op = op;
    r.bits = plain_bits;
    r.result_type = result_type ? FHE_LOGIC : FHE_ARITHMETIC;
    return submit(s, r, ticket);
}

int fhe_sched_flush(fhe_sched *s)
{
    FHE_CHECK_ARG(s, "fhe_sched_flush: null scheduler");
    HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
    std::vector<uint64_t> mine;
    for (const Req &r : s->reqs)
        if (!r.flushed) mine.push_back(r.ticket);
    const int rc = flush_pending(s);
    if (rc != FHE_OK) {
        // results of this flush are incomplete: its requests fail on wait
        for (Req &r : s->reqs)
            for (uint64_t t : mine)
                if (r.ticket == t) {
                    r.flushed = true;
                    r.failed = true;
                }
    }
    return rc;
}

int fhe_sched_wait(fhe_sched *s, uint64_t ticket, float *ms)
{
    FHE_CHECK_ARG(s, "fhe_sched_wait: null scheduler");
    for (size_t i = 0; i < s->reqs.size(); i++) {
        Req &r = s->reqs[i];
        if (r.ticket != ticket) continue;
        if (!r.flushed) {
            set_error("fhe_sched_wait: request %llu was not flushed", (unsigned long long)ticket);
            return FHE_ERR_ARG;
        }
        if (r.failed) {
            set_error("fhe_sched_wait: the flush serving request %llu failed",
                      (unsigned long long)ticket);
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.done);
            s->reqs.erase(s->reqs.begin() + long(i));
            return FHE_ERR_DEVICE;
        }
        HIP_TRY(hipEventSynchronize(r.done), "hipEventSynchronize");
        if (ms) HIP_TRY(hipEventElapsedTime(ms, r.start, r.done), "hipEventElapsedTime");
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.done);
        s->reqs.erase(s->reqs.begin() + long(i));
        return FHE_OK;
    }
    set_error("fhe_sched_wait: unknown ticket %llu", (unsigned long long)ticket);
    return FHE_ERR_ARG;
}

}  // extern "C"
