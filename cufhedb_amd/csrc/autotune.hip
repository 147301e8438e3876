// autotune.hip -- tile shapes of the batched kernels picked by measurement.
//
// The reference has nothing to tune: it launches one ciphertext per call
// (src/cuHEDB/fft_gpu/cufft_gpu.cuh:250-271).  Batched, the lvl2 -> lvl1
// key switch has a real trade-off: a workgroup switching C ciphertexts reads
// each selected key row once for all C of them (less L2 / Infinity Cache
// traffic per ciphertext), while a small batch split into few tiles leaves
// compute units idle; the FP64 lvl1 blind rotation can give a ciphertext one
// wave or two (two finish a small batch sooner, one uses the GPU best once
// the batch fills it).  fhe_autotune times every candidate of a knob on the
// context's own keys at each requested batch size and keeps the fastest per
// size in the context.  The key-switch candidates compute the same integers;
// the blind-rotation candidates decrypt identically (different FP64
// operation orders).  fhe_tune_get / fhe_tune_set read and restore the table (a
// caller caches it per device and parameter set).
#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

struct Knob {
    int ncand;
    int cand[4];
};

bool knob_of(int knob, Knob &k)
{
    switch (knob) {
    case FHE_TUNE_IKS21_TILE: k = {4, {8, 16, 32, 64}}; return true;
    case FHE_TUNE_PBS1_WAVES: k = {2, {1, 2, 0, 0}}; return true;
    case FHE_TUNE_IKS_SPLIT:
    case FHE_TUNE_IKS20_SPLIT: k = {4, {1, 2, 4, 8}}; return true;
    default: return false;
    }
}

// arbitrary 64-bit words (a hash of the index): the timed kernels' cost does
// not depend on the values (a lvl0 mask word of 0 skips a CMUX step with
// probability 2^-11 per word, the same for every candidate)
__global__ void k_fill_words(uint64_t *p, size_t n, uint64_t seed)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    p[i] = z ^ (z >> 31);
}

int check_table(const Knob &k, const int *batches, const int *vals, int n)
{
    FHE_CHECK_ARG(n >= 0 && n <= fhe_ctx::kTuneMax, "tune table: at most %d entries (got %d)",
                  fhe_ctx::kTuneMax, n);
    for (int i = 0; i < n; i++) {
        FHE_CHECK_ARG(batches[i] >= 1 && (i == 0 || batches[i] > batches[i - 1]),
                      "tune table: batch sizes must be positive and ascending");
        bool ok = false;
        for (int c = 0; c < k.ncand; c++) ok = ok || vals[i] == k.cand[c];
        FHE_CHECK_ARG(ok, "tune table: %d is not a candidate of this knob", vals[i]);
    }
    return FHE_OK;
}

}  // namespace

extern "C" {

int fhe_tune_set(fhe_ctx *ctx, int knob, const int *batches, const int *vals, int n)
{
    FHE_CHECK_ARG(ctx && (n == 0 || (batches && vals)), "fhe_tune_set: null argument");
    Knob k;
    FHE_CHECK_ARG(knob_of(knob, k), "fhe_tune_set: unknown knob %d", knob);
    if (int rc = check_table(k, batches, vals, n)) return rc;
    for (int i = 0; i < n; i++) {
        ctx->tune_batch[knob][i] = batches[i];
        ctx->tune_val[knob][i] = vals[i];
    }
    ctx->tune_n[knob] = n;
    return FHE_OK;
}

int fhe_tune_get(fhe_ctx *ctx, int knob, int *batches, int *vals, int cap, int *n)
{
    FHE_CHECK_ARG(ctx && n, "fhe_tune_get: null argument");
    Knob k;
    FHE_CHECK_ARG(knob_of(knob, k), "fhe_tune_get: unknown knob %d", knob);
    const int m = ctx->tune_n[knob];
    *n = m;
    FHE_CHECK_ARG(cap >= m && (m == 0 || (batches && vals)), "fhe_tune_get: room for %d entries "
                  "needed", m);
    for (int i = 0; i < m; i++) {
        batches[i] = ctx->tune_batch[knob][i];
        vals[i] = ctx->tune_val[knob][i];
    }
    return FHE_OK;
}

int fhe_autotune(fhe_ctx *ctx, int knob, const int *batches, int n, int *choice, float *ms)
{
    FHE_CHECK_ARG(ctx && batches, "fhe_autotune: null argument");
    Knob k;
    FHE_CHECK_ARG(knob_of(knob, k), "fhe_autotune: unknown knob %d", knob);
    FHE_CHECK_ARG(n >= 1 && n <= fhe_ctx::kTuneMax, "fhe_autotune: 1 to %d batch sizes",
                  fhe_ctx::kTuneMax);
    for (int i = 0; i < n; i++)
        FHE_CHECK_ARG(batches[i] >= 1 && batches[i] <= (1 << 20) &&
                          (i == 0 || batches[i] > batches[i - 1]),
                      "fhe_autotune: batch sizes must be ascending, in [1, 2^20]");
    // the kernel each knob shapes, and the key it needs: KSK 21, 10 or 20;
    // BK lvl01
    const int iks_kind = knob == FHE_TUNE_IKS21_TILE   ? 21
                         : knob == FHE_TUNE_IKS_SPLIT   ? 10
                         : knob == FHE_TUNE_IKS20_SPLIT ? 20
                                                        : 0;
    const bool iks = iks_kind != 0;
    if (iks ? !ctx->ksk[ksk_slot(iks_kind)] : !ctx->bk1) {
        if (iks)
            set_error("fhe_autotune: KSK %d not loaded", iks_kind);
        else
            set_error("fhe_autotune: BK lvl01 not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = ctx->stream;
    const int bmax = batches[n - 1];
    const int n0 = int(ctx->p.n0);
    // inputs (64-bit words): lvl2 TLWEs, lvl1 TLWEs, or lvl0 TLWEs plus one
    // test vector
    const size_t in_words = iks_kind == 10 ? (size_t(bmax) * (Lvl1::N + 1) + 1) / 2
                            : iks          ? size_t(bmax) * (Lvl2::N + 1)
                                           : (size_t(bmax) * (n0 + 1) + Lvl1::N + 1) / 2;
    uint64_t *in = nullptr;
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&in, in_words * sizeof(uint64_t));
    const int tn = iks_kind == 21 ? Lvl1::N : iks ? n0 : Lvl1::N;  // output TLWE width - 1
    if (e == hipSuccess) e = hipMalloc(&out, size_t(bmax) * (tn + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_words, dim3(unsigned((in_words + 255) / 256)), dim3(256), 0, s,
                           in, in_words, 0x5EEDull);
        e = hipGetLastError();
    }
    int saved_b[fhe_ctx::kTuneMax], saved_v[fhe_ctx::kTuneMax];
    const int saved = ctx->tune_n[knob];
    std::copy(ctx->tune_batch[knob], ctx->tune_batch[knob] + fhe_ctx::kTuneMax, saved_b);
    std::copy(ctx->tune_val[knob], ctx->tune_val[knob] + fhe_ctx::kTuneMax, saved_v);
    std::vector<int> picks(n);
    constexpr int kReps = 3;
    for (int bi = 0; bi < n && e == hipSuccess; bi++) {
        const int B = batches[bi];
        float best = 0.f;
        for (int c = 0; c < k.ncand && e == hipSuccess; c++) {
            auto run = [&]() -> hipError_t {
                if (iks)
                    return launch_iks(iks_kind, in, ctx->ksk[ksk_slot(iks_kind)], out, tn, B, s,
                                      k.cand[c]);
                // the lvl1 blind rotation with this candidate forced
                ctx->tune_batch[knob][0] = 1;
                ctx->tune_val[knob][0] = k.cand[c];
                ctx->tune_n[knob] = 1;
                const uint32_t *x = reinterpret_cast<const uint32_t *>(in);
                const int rc = launch_pbs1(ctx, x, x + size_t(bmax) * (n0 + 1), 0, out, true, B, s);
                return rc == FHE_OK ? hipSuccess : hipErrorUnknown;
            };
            e = run();  // warm-up (also builds a candidate's key layout)
            if (e == hipSuccess) e = hipEventRecord(e0, s);
            for (int r = 0; r < kReps && e == hipSuccess; r++) e = run();
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            t /= kReps;
            if (ms) ms[bi * k.ncand + c] = t;
            if (c == 0 || t < best) {
                best = t;
                picks[bi] = k.cand[c];
            }
        }
    }
    // a failed run leaves the previous table
    std::copy(saved_b, saved_b + fhe_ctx::kTuneMax, ctx->tune_batch[knob]);
    std::copy(saved_v, saved_v + fhe_ctx::kTuneMax, ctx->tune_val[knob]);
    ctx->tune_n[knob] = saved;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    HIP_TRY(e, "fhe_autotune");
    if (choice) std::copy(picks.begin(), picks.end(), choice);
    return fhe_tune_set(ctx, knob, batches, picks.data(), n);
}

}  // extern "C"
