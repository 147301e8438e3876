// autotune.hip -- tile shapes of the batched kernels picked by measurement.
//
// The reference has nothing to tune: it launches one ciphertext per call
// (src/cuHEDB/fft_gpu/cufft_gpu.cuh:250-271).  Batched, the lvl2 -> lvl1
// key switch has a real trade-off: a workgroup switching C ciphertexts reads
// each selected key row once for all C of them (less L2 / Infinity Cache
// traffic per ciphertext), while a small batch split into few tiles leaves
// compute units idle.  fhe_autotune times C = 8, 16 and 32 on the context's
// own key at each requested batch size and keeps the fastest per size in the
// context; every candidate computes the same integers, so the choice only
// changes speed.  fhe_tune_get / fhe_tune_set read and restore the table (a
// caller caches it per device and parameter set).
#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

constexpr int kCand[] = {8, 16, 32};
constexpr int kNumCand = 3;

// arbitrary 64-bit words (a hash of the index): the key switch's cost does
// not depend on the values, only on the digits touching every key row
__global__ void k_fill_words(uint64_t *p, size_t n, uint64_t seed)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    p[i] = z ^ (z >> 31);
}

int check_table(const int *batches, const int *tiles, int n)
{
    FHE_CHECK_ARG(n >= 0 && n <= fhe_ctx::kTuneMax, "tune table: at most %d entries (got %d)",
                  fhe_ctx::kTuneMax, n);
    for (int i = 0; i < n; i++) {
        FHE_CHECK_ARG(batches[i] >= 1 && (i == 0 || batches[i] > batches[i - 1]),
                      "tune table: batch sizes must be positive and ascending");
        FHE_CHECK_ARG(tiles[i] == 8 || tiles[i] == 16 || tiles[i] == 32,
                      "tune table: tile must be 8, 16 or 32 (got %d)", tiles[i]);
    }
    return FHE_OK;
}

}  // namespace

extern "C" {

int fhe_tune_set(fhe_ctx *ctx, int knob, const int *batches, const int *tiles, int n)
{
    FHE_CHECK_ARG(ctx && (n == 0 || (batches && tiles)), "fhe_tune_set: null argument");
    FHE_CHECK_ARG(knob == FHE_TUNE_IKS21_TILE, "fhe_tune_set: unknown knob %d", knob);
    if (int rc = check_table(batches, tiles, n)) return rc;
    for (int i = 0; i < n; i++) {
        ctx->tune21_batch[i] = batches[i];
        ctx->tune21_tile[i] = tiles[i];
    }
    ctx->tune21_n = n;
    return FHE_OK;
}

int fhe_tune_get(fhe_ctx *ctx, int knob, int *batches, int *tiles, int cap, int *n)
{
    FHE_CHECK_ARG(ctx && n, "fhe_tune_get: null argument");
    FHE_CHECK_ARG(knob == FHE_TUNE_IKS21_TILE, "fhe_tune_get: unknown knob %d", knob);
    *n = ctx->tune21_n;
    FHE_CHECK_ARG(cap >= ctx->tune21_n && (ctx->tune21_n == 0 || (batches && tiles)),
                  "fhe_tune_get: room for %d entries needed", ctx->tune21_n);
    for (int i = 0; i < ctx->tune21_n; i++) {
        batches[i] = ctx->tune21_batch[i];
        tiles[i] = ctx->tune21_tile[i];
    }
    return FHE_OK;
}

int fhe_autotune(fhe_ctx *ctx, int knob, const int *batches, int n, int *choice, float *ms)
{
    FHE_CHECK_ARG(ctx && batches, "fhe_autotune: null argument");
    FHE_CHECK_ARG(knob == FHE_TUNE_IKS21_TILE, "fhe_autotune: unknown knob %d", knob);
    FHE_CHECK_ARG(n >= 1 && n <= fhe_ctx::kTuneMax, "fhe_autotune: 1 to %d batch sizes",
                  fhe_ctx::kTuneMax);
    for (int i = 0; i < n; i++)
        FHE_CHECK_ARG(batches[i] >= 1 && batches[i] <= (1 << 20) &&
                          (i == 0 || batches[i] > batches[i - 1]),
                      "fhe_autotune: batch sizes must be ascending, in [1, 2^20]");
    const uint32_t *ksk = ctx->ksk[ksk_slot(21)];
    if (!ksk) {
        set_error("fhe_autotune: KSK 21 not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = ctx->stream;
    const int bmax = batches[n - 1];
    const size_t in_words = size_t(bmax) * (Lvl2::N + 1);
    uint64_t *in = nullptr;
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&in, in_words * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&out, size_t(bmax) * (Lvl1::N + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_words, dim3(unsigned((in_words + 255) / 256)), dim3(256), 0, s,
                           in, in_words, 0x5EEDull);
        e = hipGetLastError();
    }
    std::vector<int> tiles(n);
    constexpr int kReps = 3;
    for (int bi = 0; bi < n && e == hipSuccess; bi++) {
        float best = 0.f;
        for (int c = 0; c < kNumCand && e == hipSuccess; c++) {
            e = launch_iks(21, in, ksk, out, Lvl1::N, batches[bi], s, kCand[c]);  // warm-up
            if (e == hipSuccess) e = hipEventRecord(e0, s);
            for (int r = 0; r < kReps && e == hipSuccess; r++)
                e = launch_iks(21, in, ksk, out, Lvl1::N, batches[bi], s, kCand[c]);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            t /= kReps;
            if (ms) ms[bi * kNumCand + c] = t;
            if (c == 0 || t < best) {
                best = t;
                tiles[bi] = kCand[c];
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    HIP_TRY(e, "fhe_autotune");
    if (choice) std::copy(tiles.begin(), tiles.end(), choice);
    return fhe_tune_set(ctx, knob, batches, tiles.data(), n);
}

}  // extern "C"
