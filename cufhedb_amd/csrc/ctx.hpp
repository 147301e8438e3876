// ctx.hpp -- the device context behind the extern "C" boundary and the
// helpers every entry-point file shares (error mapping, scratch growth).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "fft64.hpp"
#include "params.hpp"

struct fhe_ctx {
    fhe_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    fhe::cd *bk1 = nullptr;  // lvl01 key, device FD layout
    fhe::cd *bk2 = nullptr;  // lvl02 key, device FD layout
    fhe::cd *rot2 = nullptr; // w2^e, e < 4096: the lvl2 rotation in the frequency domain
    uint32_t *ksk[3] = {nullptr, nullptr, nullptr};  // kinds 10, 20, 21
    fhe::cd *trgsw1 = nullptr;  // external-product TRGSW in device layout
    fhe::cd *trgsw2 = nullptr;
    int variant = FHE_PBS_FP64;
    // split-MFMA variant: the key's three integer limbs, transformed, FP32,
    // in its frequency order (built from bk1 when the variant is selected or
    // a lvl1 key uploaded while it is; bk1m_valid cleared by every lvl1 key
    // upload) and its tables
    float *bk1m = nullptr;  // [i][row][col][limb][512 complex] FP32
    bool bk1m_valid = false;
    // two-waves-per-ciphertext kernel: bk1 permuted into its order (rebuilt
    // on demand, bk1w2_valid cleared by every lvl1 key upload)
    fhe::cd *bk1w2 = nullptr;
    bool bk1w2_valid = false;
    void *mfma_tab = nullptr;
    // grow-on-demand scratch for the bootstrap operators
    uint32_t *tv1 = nullptr;     // one lvl1 test vector
    uint64_t *tv2 = nullptr;     // one lvl2 test vector
    uint32_t *tmp2 = nullptr;    // B x 2049 uint64 (as 2 x uint32 words)
    size_t tmp2_cap = 0;
    uint32_t *tmp1 = nullptr;    // B x 1025
    size_t tmp1_cap = 0;
    uint32_t *tmp0 = nullptr;    // B x (n0 + 1)
    size_t tmp0_cap = 0;
    // scratch of the comparison chains (HomMSB / compare), kept apart from
    // the single-operator scratch above because a chain step feeds one
    // operator's output into the next
    static constexpr int kChainBufs = 7;
    uint32_t *chain[kChainBufs] = {};
    size_t chain_cap[kChainBufs] = {};
    // scratch of the digit-wise comparison: lvl0 digits, bootstrapped
    // digits, their combination, test vectors
    static constexpr int kDigitBufs = 4;
    uint32_t *digit[kDigitBufs] = {};
    size_t digit_cap[kDigitBufs] = {};
    // scratch of the many-table bootstrap: rounded lvl0 batch, accumulators
    uint32_t *many[2] = {};
    size_t many_cap[2] = {};
    // tuned choices per knob (FHE_TUNE_*) and batch size (ascending batch
    // sizes; a batch uses the entry of the largest size <= it)
    static constexpr int kTuneKnobs = 4;
    static constexpr int kTuneMax = 16;
    int tune_batch[kTuneKnobs][kTuneMax] = {};
    int tune_val[kTuneKnobs][kTuneMax] = {};
    int tune_n[kTuneKnobs] = {};
};

namespace fhe {

inline int ksk_slot(int kind) { return kind == 10 ? 0 : kind == 20 ? 1 : kind == 21 ? 2 : -1; }

inline int hip_fail(hipError_t e, const char *what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? FHE_ERR_ALLOC : FHE_ERR_DEVICE;
}

#define HIP_TRY(expr, what)                                 \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return hip_fail(_e, what);    \
    } while (0)

inline hipStream_t pick(fhe_ctx *ctx, void *stream)
{
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

// untuned: batches below this many ciphertexts use two waves per ciphertext
// (one MI355X, n0 = 636: two waves 4.40 / 5.79 / 6.24 ms at 256 / 512 / 768
// ciphertexts, one wave 5.17 / 5.41 / 5.94)
constexpr int kPbs1TwoWaveBelow = 512;

// set on this thread between fhe_graph_begin and fhe_graph_end: scratch
// must not be (re)allocated while the context stream is being captured
extern thread_local bool g_capturing;

inline int grow(uint32_t *&buf, size_t &cap, size_t words)
{
    if (cap >= words) return FHE_OK;
    if (g_capturing) {
        set_error("scratch of %zu words needed during graph capture: run the pipeline once "
                  "at this size before capturing it", words);
        return FHE_ERR_ARG;
    }
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    HIP_TRY(hipMalloc(&buf, words * sizeof(uint32_t)), "scratch allocation");
    cap = words;
    return FHE_OK;
}

// Device key-switching keys are allocated with kKskPadWords spare words at
// the end: the whole-row key-switch kernel reads every row in 64-word chunks
// and may run up to 63 words past the end of the last row (those words are
// discarded).
constexpr size_t kKskPadWords = 64;
inline int alloc_ksk(fhe_ctx *ctx, int kind)
{
    const int slot = ksk_slot(kind);
    if (ctx->ksk[slot]) return FHE_OK;
    const size_t bytes = (fhe_ksk_words(&ctx->p, kind) + kKskPadWords) * sizeof(uint32_t);
    HIP_TRY(hipMalloc(&ctx->ksk[slot], bytes), "KSK allocation");
    return FHE_OK;
}

// tuned value of `knob` for a batch of B, or `dflt` if the knob is untuned
inline int tuned(const fhe_ctx *ctx, int knob, int B, int dflt)
{
    const int n = ctx->tune_n[knob];
    if (n == 0) return dflt;
    int v = ctx->tune_val[knob][0];
    for (int i = 1; i < n && ctx->tune_batch[knob][i] <= B; i++) v = ctx->tune_val[knob][i];
    return v;
}

// workgroup shape of a key switch of B ciphertexts.  Kind 21: ciphertexts
// per workgroup; untuned 8 below 512 ciphertexts, 16 below 2048, else 32
// (one MI355X: 32 ciphertexts 1.9 / 2.7 / 9.2 ms with tiles 8 / 16 / 32,
// 2048 ciphertexts 19.1 / 17.3 / 13.6 ms).  Kinds 10 and 20: workgroups per
// ciphertext over the digit positions; untuned 8 (one MI355X, n0 = 672, split
// 1 / 8: kind 10 16 ciphertexts 0.13 / 0.02 ms, 16384 ciphertexts 15.1 /
// 13.5 ms; kind 20 16384 ciphertexts 31.0 / 28.6 ms).
inline int iks_tile(const fhe_ctx *ctx, int kind, int B)
{
    if (kind == 21) return tuned(ctx, FHE_TUNE_IKS21_TILE, B, B < 256 ? 8 : B < 512 ? 16 : 64);
    return tuned(ctx, kind == 10 ? FHE_TUNE_IKS_SPLIT : FHE_TUNE_IKS20_SPLIT, B, 8);
}

// waves per ciphertext of the FP64 lvl1 blind rotation for a batch of B
inline int pbs1_waves(const fhe_ctx *ctx, int B)
{
    return tuned(ctx, FHE_TUNE_PBS1_WAVES, B, B < kPbs1TwoWaveBelow ? 2 : 1);
}

// lvl1 blind rotation (+ sample extraction) of B ciphertexts with the
// context's variant and tuned shape (capi.hip)
int launch_pbs1(fhe_ctx *ctx, const uint32_t *in, const uint32_t *tv, int tv_stride,
                uint32_t *out, bool se, int B, hipStream_t s);
int prepare_pbs1(fhe_ctx *ctx, hipStream_t s);

// Buffers a lvl1 gate bootstrap works in: B lvl1 TLWEs (the gate's linear
// combination), B lvl0 TLWEs (after the key switch) and one test vector.
struct GateScratch {
    uint32_t *tmp1;
    uint32_t *tmp0;
    uint32_t *tv1;
};

// HomNAND / NOR / XNOR / AND / OR of B pairs of lvl1 TLWEs (capi.hip): in the
// context's grow-on-demand scratch (fixed == nullptr) or in `fixed`
int gate_lvl1(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
              uint32_t *out_dev, int result_type, int B, hipStream_t s, const GateScratch *fixed);

// device address of the allocation-free entry point's fixed scratch on
// `device` (current), taken once per device (noalloc.hip)
int noalloc_scratch_init(int device);

inline int ks_target_n(const fhe_ctx *ctx, int kind) { return kind == 21 ? Lvl1::N : int(ctx->p.n0); }
inline int ks_domain_words(int kind) { return kind == 10 ? Lvl1::N + 1 : Lvl2::N + 1; }

}  // namespace fhe

