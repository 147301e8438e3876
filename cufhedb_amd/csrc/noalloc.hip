// noalloc.hip -- a lvl1 gate bootstrap entry point that allocates nothing.
//
// fhe_gate_lvl1_dev works in the context's grow-on-demand scratch: the gate's
// linear combination (B lvl1 TLWEs), the key-switched lvl0 batch (B lvl0
// TLWEs) and the test vector, allocated with hipMalloc the first time a batch
// of that size runs.  Callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc) use
// fhe_gate_lvl1_noalloc_dev instead: the same kernels on the same inputs, with
// those three buffers carved out of one file-scope device array, `scratch`,
// of fixed size (batches up to kNoAllocMaxBatch ciphertexts; 16.8 MB that
// every GPU the library runs on holds until the process ends).
//
// The array's name is a host-side stand-in in host code, not a device
// address: its address on each device comes from hipGetSymbolAddress, taken
// once per device when a context is created on it (fhe_ctx_create calls
// noalloc_scratch_init with the device current) and kept in g_scratch.  Every
// call on a device shares that device's array and leaves its contents
// behind: every word a call reads is written earlier in the same call (the
// key switch zero-fills its output with a kernel before accumulating into
// it), and calls that overlap on one GPU (two streams, two host threads)
// race on the array.
#include <mutex>

#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

constexpr int kNoAllocMaxBatch = 2048;
// layout, in 32-bit words: test vector | B lvl1 TLWEs | B lvl0 TLWEs (each
// region starting on a 256-byte boundary)
constexpr size_t kTvWords = Lvl1::N;
constexpr size_t kTmp1Off = kTvWords;
constexpr size_t kTmp0Off = (kTmp1Off + size_t(kNoAllocMaxBatch) * (Lvl1::N + 1) + 63) / 64 * 64;
constexpr size_t kScratchWords = kTmp0Off + size_t(kNoAllocMaxBatch) * (kMaxN0 + 1);

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t *g_scratch[kMaxDevices] = {};

}  // namespace

namespace fhe {

// Records the device address of `device`'s copy of `scratch`; the device
// must be current.  Idempotent.
int noalloc_scratch_init(int device)
{
    FHE_CHECK_ARG(device >= 0 && device < kMaxDevices, "no-alloc scratch: device %d beyond %d",
                  device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return FHE_OK;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "no-alloc scratch address");
    g_scratch[device] = static_cast<uint32_t *>(p);
    return FHE_OK;
}

}  // namespace fhe

extern "C" {

int fhe_gate_lvl1_noalloc_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                              uint32_t *out_dev, int result_type, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && a_dev && b_dev && out_dev && B >= 0,
                  "fhe_gate_lvl1_noalloc_dev: bad argument");
    FHE_CHECK_ARG(B <= kNoAllocMaxBatch,
                  "fhe_gate_lvl1_noalloc_dev: batch of %d above the fixed scratch's %d "
                  "ciphertexts (use fhe_gate_lvl1_dev)", B, kNoAllocMaxBatch);
    // the PBS kernel's key order must already exist: building it allocates
    if (ctx->bk1 && ((ctx->variant == FHE_PBS_FP64 && pbs1_waves(ctx, B) == 2 &&
                      !ctx->bk1w2_valid) ||
                     (ctx->variant == FHE_PBS_MFMA_SPLIT && !ctx->bk1m_valid))) {
        set_error("fhe_gate_lvl1_noalloc_dev: the lvl1 key's kernel layout is missing (load "
                  "the key again)");
        return FHE_ERR_ARG;
    }
    uint32_t *base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    }
    if (!base) {
        set_error("fhe_gate_lvl1_noalloc_dev: no scratch address for device %d", ctx->device);
        return FHE_ERR_DEVICE;
    }
    const GateScratch sc{base + kTmp1Off, base + kTmp0Off, base};
    return gate_lvl1(ctx, gate, a_dev, b_dev, out_dev, result_type, B, pick(ctx, stream), &sc);
}

}  // extern "C"
