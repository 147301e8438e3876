// sched.hip -- coalescing scheduler for a stream of small bootstrap requests.
//
// The mixed-parameter workload is a stream of independent requests: lvl1
// LUT bootstraps (a batch of lvl0 ciphertexts with one test vector) and lvl2
// comparisons (a batch against one threshold ciphertext).  The reference
// serves such a stream with one host thread per CUDA stream and one
// ciphertext per call (src/cuHEDB/gatebootstrap_gpu.cuh:17-25,
// test/HE3_Query/TPCH_Q6.cu:179-193).  A request of a few thousand rows
// fills only part of the 256 CUs (the lvl1 kernel holds 8192 ciphertexts at
// once).  Here a flush serves every pending lvl1 request with one persistent
// launch (k_blind_rotate1_jobs): a table of jobs (each request's input,
// test vector and output buffers) is handed to resident waves that take the
// next ciphertext of any request from a device cursor, with the key
// resident, no inputs gathered and no test vectors replicated; the kernel
// stamps each request's completion on the device clock.  lvl2 comparisons
// are merged: consecutive requests of one kind (and op / width / result
// type) into launches of up to max_rows rows, their inputs gathered into
// device staging buffers, the per-request threshold replicated per row,
// one batched call, the results scattered back; a comparison of more than
// max_rows rows runs alone on its own buffers.  lvl1 launches go to the context's
// stream and lvl2 comparisons to a second stream of the scheduler (default
// priority), so a lvl1 batch and a lvl2 comparison share the CUs (each
// kind in submission order; the second stream starts after the work
// already on the context stream, and the context stream continues after
// both).  The two kinds use disjoint staging and scratch: a comparison's
// chain scratch is the context's, the lvl1 bootstrap only reads the keys,
// whose lazily built layouts are built before the streams fork.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

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
    int flush = -1;       // lvl1: the flush record holding its completion stamp
    int job = -1;         //       and its job index there
};

// Device memory of one flush's persistent lvl1 launch: the job table, the
// cursor and per-job completion counters, the flush's start time and each
// job's completion time (device clock); kept until the scheduler ends.
struct FlushRec {
    void *mem = nullptr;
    Pbs1Job *host = nullptr;  // pinned host copy of the job table (async upload)
    int njobs = 0;
    Pbs1Job *jobs() const { return static_cast<Pbs1Job *>(mem); }
    int *cursor() const { return reinterpret_cast<int *>(jobs() + njobs); }
    int *done() const { return cursor() + 1; }
    uint64_t *stamps() const  // [0] = start, [1 + j] = job j's completion
    {
        const size_t off = (size_t(njobs) * sizeof(Pbs1Job) + size_t(njobs + 1) * sizeof(int) + 7) & ~size_t(7);
        return reinterpret_cast<uint64_t *>(static_cast<char *>(mem) + off);
    }
    static size_t bytes(int n)
    {
        return ((size_t(n) * sizeof(Pbs1Job) + size_t(n + 1) * sizeof(int) + 7) & ~size_t(7)) +
               size_t(n + 1) * sizeof(uint64_t);
    }
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
    // comparison staging, allocated on first use
    uint64_t *a2 = nullptr, *b2 = nullptr;
    uint32_t *outc = nullptr;
    hipStream_t s2 = nullptr;  // lvl2 comparisons
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<fhe::FlushRec> flushes;
    int clock_khz = 0;  // device clock rate of the completion stamps
};

using namespace fhe;

namespace {

int ensure_staging(fhe_sched *s, int kind)
{
    const size_t R = size_t(s->max_rows);
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

// one batched comparison over rows [0, rows) of the staging buffers
int run_group(fhe_sched *s, const std::vector<Req *> &g, int rows, hipStream_t st)
{
    fhe_ctx *ctx = s->ctx;
    SCHED_TRY(ensure_staging(s, FHE_SCHED_CMP2));
    size_t off = 0;
    for (Req *r : g) {
        HIP_TRY(hipMemcpyAsync(s->a2 + off * W2, r->in, size_t(r->B) * W2 * 8,
                               hipMemcpyDeviceToDevice, st),
                "scheduler gather");
        HIP_TRY(replicate(s->b2 + off * W2, static_cast<const uint64_t *>(r->aux), W2, r->B, st),
                "scheduler thresholds");
        off += size_t(r->B);
    }
    const Req &r0 = *g[0];
    SCHED_TRY(fhe_compare_dev(ctx, 2, r0.op, s->a2, s->b2, 0, s->outc, r0.bits, r0.result_type,
                               rows, st));
    uint32_t *res = s->outc;
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

// a comparison too large for the staging buffers runs on its own buffers
int run_alone(fhe_sched *s, Req &r, hipStream_t st)
{
    SCHED_TRY(fhe_compare_dev(s->ctx, 2, r.op, r.in, r.aux, 1, r.out, r.bits, r.result_type, r.B,
                               st));
    HIP_TRY(hipEventRecord(r.done, st), "hipEventRecord");
    return FHE_OK;
}

// every pending lvl1 request through one persistent launch on st1
int serve_pbs1(fhe_sched *s, const std::vector<Req *> &reqs, hipStream_t st1)
{
    fhe_ctx *ctx = s->ctx;
    if (reqs.empty()) return FHE_OK;
    if (!ctx->bk1) {
        set_error("fhe_sched_flush: lvl01 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    std::vector<Pbs1Job> jobs;
    int total = 0;
    for (Req *r : reqs) {
        if (r->B == 0) continue;
        jobs.push_back(Pbs1Job{static_cast<const uint32_t *>(r->in),
                               static_cast<const uint32_t *>(r->aux), r->out, total, r->B});
        r->job = int(jobs.size()) - 1;
        total += r->B;
    }
    FlushRec f;
    f.njobs = int(jobs.size());
    HIP_TRY(hipMalloc(&f.mem, FlushRec::bytes(f.njobs)), "scheduler job table");
    s->flushes.push_back(f);
    FlushRec &fr = s->flushes.back();
    const int fi = int(s->flushes.size()) - 1;
    for (Req *r : reqs) r->flush = fi;
    HIP_TRY(launch_clock_stamp(fr.stamps(), st1), "scheduler start stamp");
    if (fr.njobs > 0) {
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&fr.host), jobs.size() * sizeof(Pbs1Job),
                              hipHostMallocDefault),
                "scheduler job table");
        std::copy(jobs.begin(), jobs.end(), fr.host);
        HIP_TRY(hipMemcpyAsync(fr.jobs(), fr.host, jobs.size() * sizeof(Pbs1Job),
                               hipMemcpyHostToDevice, st1),
                "scheduler job table");
        HIP_TRY(launch_blind_rotate1_jobs(fr.jobs(), fr.njobs, total, int(ctx->p.n0), ctx->bk1,
                                          fr.cursor(), fr.done(), fr.stamps() + 1, st1),
                "scheduler lvl1 bootstraps");
    }
    for (Req *r : reqs) HIP_TRY(hipEventRecord(r->done, st1), "hipEventRecord");
    return FHE_OK;
}

// serves every request not yet flushed, each kind in order on its stream
int serve_pending(fhe_sched *s, hipStream_t st1, hipStream_t st2)
{
    std::vector<Req *> pbs1, group;
    int rows = 0;
    auto close_group = [&]() -> int {
        if (group.empty()) return FHE_OK;
        const int rc = run_group(s, group, rows, st2);
        group.clear();
        rows = 0;
        return rc;
    };
    for (Req &r : s->reqs) {
        if (r.flushed) continue;
        r.flushed = true;
        if (r.kind == FHE_SCHED_PBS1) {
            pbs1.push_back(&r);
            continue;
        }
        if (r.B > s->max_rows) {
            SCHED_TRY(close_group());
            SCHED_TRY(run_alone(s, r, st2));
            continue;
        }
        if (!group.empty() && (!same_group(*group[0], r) || rows + r.B > s->max_rows))
            SCHED_TRY(close_group());
        group.push_back(&r);
        rows += r.B;
    }
    SCHED_TRY(serve_pbs1(s, pbs1, st1));
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
    hipError_t e = hipDeviceGetAttribute(&s->clock_khz, hipDeviceAttributeWallClockRate,
                                         ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->s2, hipStreamNonBlocking);
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
    void *bufs[] = {s->a2, s->b2, s->outc};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (FlushRec &f : s->flushes) {
        if (f.mem) (void)hipFree(f.mem);
        if (f.host) (void)hipHostFree(f.host);
    }
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
        if (ms && r.kind == FHE_SCHED_PBS1) {
            // the flush's start and this request's completion, device clock
            uint64_t t[2] = {0, 0};
            const FlushRec &f = s->flushes[size_t(r.flush)];
            HIP_TRY(hipMemcpy(&t[0], f.stamps(), sizeof(uint64_t), hipMemcpyDeviceToHost),
                    "scheduler stamps");
            if (r.job >= 0)
                HIP_TRY(hipMemcpy(&t[1], f.stamps() + 1 + r.job, sizeof(uint64_t),
                                  hipMemcpyDeviceToHost),
                        "scheduler stamps");
            else
                t[1] = t[0];  // an empty request is done when its flush starts
            *ms = float(double(t[1] - t[0]) / double(s->clock_khz));
        } else if (ms) {
            HIP_TRY(hipEventElapsedTime(ms, r.start, r.done), "hipEventElapsedTime");
        }
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.done);
        s->reqs.erase(s->reqs.begin() + long(i));
        return FHE_OK;
    }
    set_error("fhe_sched_wait: unknown ticket %llu", (unsigned long long)ticket);
    return FHE_ERR_ARG;
}

}  // extern "C"
