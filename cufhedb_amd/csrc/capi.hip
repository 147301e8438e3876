// capi.hip -- the extern "C" boundary: device context, key upload, batched
// entry points, bootstrap operators and memory/timing helpers.
//
// A context corresponds to one cufft::CuFFT_Torus<P> pair of the reference
// (src/cuHEDB/fft_gpu/cufft_gpu.cuh:10-451) but holds both levels, keeps
// every key resident on the device, and takes whole batches per call instead
// of one ciphertext per stream_id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

// The split-MFMA variant's key (three integer limbs per torus word, each
// limb polynomial's transform in FP32, in its own frequency order) and
// tables, derived from the FP64 device key when the variant is selected and
// after each lvl1 upload while it is selected.
int ensure_mfma(fhe_ctx *ctx, hipStream_t s)
{
    if ((!ctx->mfma_tab || !ctx->bk1m_valid) && g_capturing) {
        set_error("split-MFMA key layout needed during graph capture: run one lvl1 bootstrap "
                  "with this variant before capturing");
        return FHE_ERR_ARG;
    }
    if (!ctx->mfma_tab) {
        HIP_TRY(hipMalloc(&ctx->mfma_tab, mfma_tables_bytes()), "MFMA table allocation");
        HIP_TRY(launch_mfma_tables(ctx->mfma_tab, s), "MFMA tables");
    }
    if (!ctx->bk1m_valid) {
        if (!ctx->bk1m)
            HIP_TRY(hipMalloc(&ctx->bk1m, bk1_mfma_floats(int(ctx->p.n0)) * sizeof(float)),
                    "MFMA key allocation");
        HIP_TRY(launch_bk1_to_mfma(ctx->bk1, ctx->bk1m, int(ctx->p.n0), s), "MFMA key layout");
        ctx->bk1m_valid = true;
    }
    return FHE_OK;
}

}  // namespace

namespace fhe {

// the key layouts launch_pbs1 builds on first use (split-MFMA key and
// tables, or the two-wave key order), built now on stream s: afterwards
// launch_pbs1 only reads context memory, so it may run on several streams
int prepare_pbs1(fhe_ctx *ctx, hipStream_t s)
{
    if (!ctx->bk1) return FHE_OK;
    if (ctx->variant == FHE_PBS_MFMA_SPLIT) return ensure_mfma(ctx, s);
    if (!ctx->bk1w2_valid) {
        // fhe_load_bk builds this layout with the key, so only a failed
        // upload leaves it missing; never build it into a captured graph
        if (g_capturing) {
            set_error("two-wave key layout missing during graph capture");
            return FHE_ERR_ARG;
        }
        const int n0 = int(ctx->p.n0);
        if (!ctx->bk1w2)
            HIP_TRY(hipMalloc(&ctx->bk1w2, bk1_device_elems(n0) * sizeof(cd)),
                    "two-wave key allocation");
        HIP_TRY(launch_bk1_to_w2(ctx->bk1, ctx->bk1w2, n0, s), "two-wave key layout");
        ctx->bk1w2_valid = true;
    }
    return FHE_OK;
}

// lvl1 blind rotation (+ sample extraction) with the context's variant
int launch_pbs1(fhe_ctx *ctx, const uint32_t *in, const uint32_t *tv, int tv_stride,
                uint32_t *out, bool se, int B, hipStream_t s)
{
    const int n0 = int(ctx->p.n0);
    if (ctx->variant == FHE_PBS_MFMA_SPLIT) {
        int rc = ensure_mfma(ctx, s);
        if (rc) return rc;
        HIP_TRY(launch_blind_rotate1_mfma(in, n0, ctx->bk1m, ctx->mfma_tab, tv, tv_stride, out, se,
                                          B, s),
                "pbs lvl01 (split MFMA)");
        return FHE_OK;
    }
    if (pbs1_waves(ctx, B) == 2) {
        int rc = prepare_pbs1(ctx, s);
        if (rc) return rc;
        HIP_TRY(launch_blind_rotate1w2(in, n0, ctx->bk1w2, tv, tv_stride, out, se, B, s),
                "pbs lvl01 (two waves per ciphertext)");
        return FHE_OK;
    }
    HIP_TRY(launch_blind_rotate1(in, n0, ctx->bk1, tv, tv_stride, out, se, B, s), "pbs lvl01");
    return FHE_OK;
}

}  // namespace fhe

namespace {

// lvl02 key given in TFHEpp's FD form: back to torus polynomials (TwistFFT,
// FP64, as exact as the FD values allow), then into the blind rotation's
// layout through its own transform
hipError_t bk2_fft_to_q(const double *bk_fft, cd *bk_dev, int n0, hipStream_t s)
{
    const int npoly = n0 * Lvl2::rows * 2;
    uint64_t *polys = nullptr;
    hipError_t e = hipMalloc(&polys, size_t(npoly) * Lvl2::N * sizeof(uint64_t));
    if (e != hipSuccess) return e;
    e = launch_fft2(bk_fft, polys, npoly, s);
    if (e == hipSuccess) e = launch_bk2q_from_int(polys, bk_dev, n0, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(polys);
    return e;
}

}  // namespace

extern "C" {

int fhe_ctx_create(const fhe_params *p, int device, fhe_ctx **out)
{
    FHE_CHECK_ARG(p && out, "fhe_ctx_create: null argument");
    FHE_CHECK_ARG(p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_ctx_create: n0 must be in [1, %d]",
                  kMaxN0);
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    FHE_CHECK_ARG(device >= 0 && device < count, "fhe_ctx_create: device %d of %d", device, count);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    auto *ctx = new fhe_ctx;
    ctx->p = *p;
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipMalloc(&ctx->tv1, Lvl1::N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&ctx->tv2, Lvl2::N * sizeof(uint64_t));
    if (e != hipSuccess) {
        fhe_ctx_destroy(ctx);
        return hip_fail(e, "fhe_ctx_create");
    }
    // this device's fixed scratch of the allocation-free entry point
    if (int rc = noalloc_scratch_init(device)) {
        fhe_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return FHE_OK;
}

int fhe_ctx_destroy(fhe_ctx *ctx)
{
    if (!ctx) return FHE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    void *bufs[] = {ctx->bk1,  ctx->bk2,    ctx->ksk[0], ctx->ksk[1], ctx->ksk[2],
                    ctx->trgsw1, ctx->trgsw2, ctx->tv1,  ctx->tv2,    ctx->tmp1,
                    ctx->tmp0, ctx->tmp2,   ctx->bk1m, ctx->mfma_tab, ctx->bk1w2, ctx->rot2};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (uint32_t *b : ctx->chain)
        if (b) (void)hipFree(b);
    for (uint32_t *b : ctx->digit)
        if (b) (void)hipFree(b);
    for (uint32_t *b : ctx->many)
        if (b) (void)hipFree(b);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FHE_OK;
}

int fhe_ctx_sync(fhe_ctx *ctx)
{
    FHE_CHECK_ARG(ctx, "fhe_ctx_sync: null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return FHE_OK;
}

void *fhe_ctx_stream(fhe_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int fhe_set_pbs_variant(fhe_ctx *ctx, int variant)
{
    FHE_CHECK_ARG(ctx, "fhe_set_pbs_variant: null context");
    FHE_CHECK_ARG(variant == FHE_PBS_FP64 || variant == FHE_PBS_MFMA_SPLIT,
                  "fhe_set_pbs_variant: unknown variant %d", variant);
    FHE_CHECK_ARG(!g_capturing, "fhe_set_pbs_variant: not during a graph capture");
    ctx->variant = variant;
    // the variant's key layout is built now (when a key is loaded), so that
    // no later bootstrap has to allocate or convert
    if (variant == FHE_PBS_MFMA_SPLIT && ctx->bk1) {
        HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
        if (int rc = ensure_mfma(ctx, ctx->stream)) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream), "fhe_set_pbs_variant");
    }
    return FHE_OK;
}

int fhe_load_bk(fhe_ctx *ctx, int level, const void *bk_host, int format)
{
    FHE_CHECK_ARG(ctx && bk_host, "fhe_load_bk: null argument");
    FHE_CHECK_ARG(level == 1 || level == 2, "fhe_load_bk: level must be 1 or 2");
    FHE_CHECK_ARG(format == FHE_BK_INT || format == FHE_BK_FFT, "fhe_load_bk: bad format %d", format);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const int n0 = int(ctx->p.n0);
    const bool l1 = level == 1;
    const size_t elems = l1 ? bk1_device_elems(n0) : bk2q_device_elems(n0);
    cd *&dst = l1 ? ctx->bk1 : ctx->bk2;
    if (!dst) HIP_TRY(hipMalloc(&dst, elems * sizeof(cd)), "BK allocation");
    const int rows = l1 ? Lvl1::rows : Lvl2::rows;
    const int N = l1 ? Lvl1::N : Lvl2::N;
    const size_t word = format == FHE_BK_FFT ? sizeof(double) : (l1 ? 4 : 8);
    const size_t src_bytes = size_t(n0) * rows * 2 * N * word;
    void *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, src_bytes), "BK staging allocation");
    hipError_t e = hipMemcpyAsync(tmp, bk_host, src_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        if (l1)
            e = format == FHE_BK_INT
                    ? launch_bk1_from_int(static_cast<uint32_t *>(tmp), dst, n0, ctx->stream)
                    : launch_bk1_from_fft(static_cast<double *>(tmp), dst, n0, ctx->stream);
        else
            e = format == FHE_BK_INT
                    ? launch_bk2q_from_int(static_cast<uint64_t *>(tmp), dst, n0, ctx->stream)
                    : bk2_fft_to_q(static_cast<double *>(tmp), dst, n0, ctx->stream);
    }
    // the lvl2 blind rotation applies X^a - 1 in the frequency domain with
    // this table of w2^e
    if (e == hipSuccess && !l1 && !ctx->rot2) {
        e = hipMalloc(&ctx->rot2, rot2q_elems() * sizeof(cd));
        if (e == hipSuccess) e = launch_rot2q_table(ctx->rot2, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    HIP_TRY(e, "fhe_load_bk");
    if (l1) {
        // the two-wave kernel's key order is built here, with the key, so
        // that no later bootstrap (possibly inside a graph capture) has to
        ctx->bk1m_valid = ctx->bk1w2_valid = false;
        const int v = ctx->variant;
        ctx->variant = FHE_PBS_FP64;
        int rc = prepare_pbs1(ctx, ctx->stream);
        ctx->variant = v;
        if (rc == FHE_OK && v == FHE_PBS_MFMA_SPLIT) rc = ensure_mfma(ctx, ctx->stream);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(ctx->stream), "fhe_load_bk");
    }
    return FHE_OK;
}

int fhe_load_ksk(fhe_ctx *ctx, int kind, const uint32_t *ksk_host)
{
    FHE_CHECK_ARG(ctx && ksk_host, "fhe_load_ksk: null argument");
    const int slot = ksk_slot(kind);
    FHE_CHECK_ARG(slot >= 0, "fhe_load_ksk: kind must be 10, 20 or 21 (got %d)", kind);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t bytes = fhe_ksk_words(&ctx->p, kind) * sizeof(uint32_t);
    if (int rc = alloc_ksk(ctx, kind)) return rc;
    HIP_TRY(hipMemcpy(ctx->ksk[slot], ksk_host, bytes, hipMemcpyHostToDevice), "KSK upload");
    return FHE_OK;
}

int fhe_bk_loaded(const fhe_ctx *ctx, int level)
{
    if (!ctx) return 0;
    return (level == 1 ? ctx->bk1 : level == 2 ? ctx->bk2 : nullptr) != nullptr;
}

int fhe_ksk_loaded(const fhe_ctx *ctx, int kind)
{
    const int slot = ksk_slot(kind);
    return ctx && slot >= 0 && ctx->ksk[slot] != nullptr;
}

int fhe_blind_rotate_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev,
                               int tv_stride, uint32_t *acc_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && acc_dev, "fhe_blind_rotate_lvl01_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_blind_rotate_lvl01_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl1::N, "tv_stride must be 0 or %d", Lvl1::N);
    if (!ctx->bk1) {
        set_error("fhe_blind_rotate_lvl01_dev: lvl01 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    return launch_pbs1(ctx, in_dev, tv_dev, tv_stride, acc_dev, false, B, pick(ctx, stream));
}

int fhe_pbs_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev, int tv_stride,
                      uint32_t *out_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && out_dev, "fhe_pbs_lvl01_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_pbs_lvl01_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl1::N, "tv_stride must be 0 or %d", Lvl1::N);
    if (!ctx->bk1) {
        set_error("fhe_pbs_lvl01_dev: lvl01 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    return launch_pbs1(ctx, in_dev, tv_dev, tv_stride, out_dev, true, B, pick(ctx, stream));
}

int fhe_pbs_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev, int tv_stride,
                      uint64_t *out_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && out_dev, "fhe_pbs_lvl02_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_pbs_lvl02_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl2::N, "tv_stride must be 0 or %d", Lvl2::N);
    if (!ctx->bk2) {
        set_error("fhe_pbs_lvl02_dev: lvl02 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    HIP_TRY(launch_blind_rotate2q(in_dev, int(ctx->p.n0), ctx->bk2, ctx->rot2, tv_dev, tv_stride, out_dev, true,
                                 B, pick(ctx, stream)),
            "pbs lvl02");
    return FHE_OK;
}

int fhe_blind_rotate_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev,
                               int tv_stride, uint64_t *acc_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && acc_dev, "fhe_blind_rotate_lvl02_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_blind_rotate_lvl02_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl2::N, "tv_stride must be 0 or %d", Lvl2::N);
    if (!ctx->bk2) {
        set_error("fhe_blind_rotate_lvl02_dev: lvl02 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    HIP_TRY(launch_blind_rotate2q(in_dev, int(ctx->p.n0), ctx->bk2, ctx->rot2, tv_dev, tv_stride, acc_dev,
                                 false, B, pick(ctx, stream)),
            "blind rotate lvl02");
    return FHE_OK;
}

int fhe_sample_extract_lvl1_dev(fhe_ctx *ctx, const uint32_t *trlwe_dev, uint32_t *tlwe_dev,
                                int index, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && trlwe_dev && tlwe_dev, "fhe_sample_extract_lvl1_dev: null argument");
    FHE_CHECK_ARG(B >= 0 && index >= 0 && index < Lvl1::N, "fhe_sample_extract_lvl1_dev: bad size");
    HIP_TRY(launch_sample_extract1(trlwe_dev, tlwe_dev, index, B, pick(ctx, stream)),
            "sample extract lvl1");
    return FHE_OK;
}

int fhe_sample_extract_lvl2_dev(fhe_ctx *ctx, const uint64_t *trlwe_dev, uint64_t *tlwe_dev,
                                int index, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && trlwe_dev && tlwe_dev, "fhe_sample_extract_lvl2_dev: null argument");
    FHE_CHECK_ARG(B >= 0 && index >= 0 && index < Lvl2::N, "fhe_sample_extract_lvl2_dev: bad size");
    HIP_TRY(launch_sample_extract2(trlwe_dev, tlwe_dev, index, B, pick(ctx, stream)),
            "sample extract lvl2");
    return FHE_OK;
}

int fhe_iks_dev(fhe_ctx *ctx, int kind, const void *in_dev, uint32_t *out_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev, "fhe_iks_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_iks_dev: negative batch");
    const int slot = ksk_slot(kind);
    FHE_CHECK_ARG(slot >= 0, "fhe_iks_dev: kind must be 10, 20 or 21 (got %d)", kind);
    if (!ctx->ksk[slot]) {
        set_error("fhe_iks_dev: key-switching key %d not loaded", kind);
        return FHE_ERR_NOT_LOADED;
    }
    HIP_TRY(launch_iks(kind, in_dev, ctx->ksk[slot], out_dev, ks_target_n(ctx, kind), B,
                       pick(ctx, stream), iks_tile(ctx, kind, B)),
            "identity key switch");
    return FHE_OK;
}

int fhe_external_product_lvl1_dev(fhe_ctx *ctx, const uint32_t *trlwe_dev,
                                  const double *trgswfft_dev, uint32_t *out_dev, int B,
                                  void *stream)
{
    FHE_CHECK_ARG(ctx && trlwe_dev && trgswfft_dev && out_dev,
                  "fhe_external_product_lvl1_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_external_product_lvl1_dev: negative batch");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->trgsw1)
        HIP_TRY(hipMalloc(&ctx->trgsw1, size_t(Lvl1::rows) * 2 * 512 * sizeof(cd)),
                "TRGSW scratch allocation");
    hipStream_t s = pick(ctx, stream);
    HIP_TRY(launch_trgsw1_from_fft(trgswfft_dev, ctx->trgsw1, s), "TRGSW layout");
    HIP_TRY(launch_extprod1(trlwe_dev, ctx->trgsw1, out_dev, B, s), "external product lvl1");
    return FHE_OK;
}

int fhe_external_product_lvl2_dev(fhe_ctx *ctx, const uint64_t *trlwe_dev,
                                  const double *trgswfft_dev, uint64_t *out_dev, int B,
                                  void *stream)
{
    FHE_CHECK_ARG(ctx && trlwe_dev && trgswfft_dev && out_dev,
                  "fhe_external_product_lvl2_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_external_product_lvl2_dev: negative batch");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->trgsw2)
        HIP_TRY(hipMalloc(&ctx->trgsw2, size_t(Lvl2::rows) * 2 * 1024 * sizeof(cd)),
                "TRGSW scratch allocation");
    hipStream_t s = pick(ctx, stream);
    HIP_TRY(launch_trgsw2_from_fft(trgswfft_dev, ctx->trgsw2, s), "TRGSW layout");
    HIP_TRY(launch_extprod2(trlwe_dev, ctx->trgsw2, out_dev, B, s), "external product lvl2");
    return FHE_OK;
}

int fhe_ifft_lvl1_dev(fhe_ctx *ctx, const uint32_t *poly_dev, double *fd_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && poly_dev && fd_dev && B >= 0, "fhe_ifft_lvl1_dev: bad argument");
    HIP_TRY(launch_ifft1(poly_dev, fd_dev, B, pick(ctx, stream)), "ifft lvl1");
    return FHE_OK;
}

int fhe_fwd_digits_lvl1_mfma_dev(fhe_ctx *ctx, const int32_t *poly_dev, double *fd_dev, int B,
                                 void *stream)
{
    FHE_CHECK_ARG(ctx && poly_dev && fd_dev && B >= 0, "fhe_fwd_digits_lvl1_mfma_dev: bad argument");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    if (!ctx->mfma_tab) {
        HIP_TRY(hipMalloc(&ctx->mfma_tab, mfma_tables_bytes()), "MFMA table allocation");
        HIP_TRY(launch_mfma_tables(ctx->mfma_tab, s), "MFMA tables");
    }
    HIP_TRY(launch_fwd_digits_mfma(poly_dev, ctx->mfma_tab, fd_dev, B, s), "split-MFMA transform");
    return FHE_OK;
}

int fhe_fft_lvl1_dev(fhe_ctx *ctx, const double *fd_dev, uint32_t *poly_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && poly_dev && fd_dev && B >= 0, "fhe_fft_lvl1_dev: bad argument");
    HIP_TRY(launch_fft1(fd_dev, poly_dev, B, pick(ctx, stream)), "fft lvl1");
    return FHE_OK;
}

int fhe_ifft_lvl2_dev(fhe_ctx *ctx, const uint64_t *poly_dev, double *fd_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && poly_dev && fd_dev && B >= 0, "fhe_ifft_lvl2_dev: bad argument");
    HIP_TRY(launch_ifft2q(poly_dev, fd_dev, B, pick(ctx, stream)), "ifft lvl2");
    return FHE_OK;
}

int fhe_fft_lvl2_dev(fhe_ctx *ctx, const double *fd_dev, uint64_t *poly_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && poly_dev && fd_dev && B >= 0, "fhe_fft_lvl2_dev: bad argument");
    HIP_TRY(launch_fft2(fd_dev, poly_dev, B, pick(ctx, stream)), "fft lvl2");
    return FHE_OK;
}

// ---- bootstrap operators --------------------------------------------------

namespace {

// Test vector of a bootstrap operator: constant u, or identity
// 2^scale * (i >> 6) (gpolygen, src/HEDB/comparison/tfhepp_utils.h:113-120).
struct TvSpec {
    bool ide;
    uint64_t value;  // u, or scale_bits
};

// in (lvl `in_level` TLWEs already offset, device) -> IKS to lvl0 -> PBS to
// lvl `out_level` with the test vector -> body += post.  fixed: the lvl0
// batch and test-vector buffers to use instead of the context's
// grow-on-demand scratch (lvl1 output only).
int run_bootstrap(fhe_ctx *ctx, int in_level, const void *in, int out_level, void *out, TvSpec tv,
                  uint64_t post, int B, hipStream_t s, const GateScratch *fixed = nullptr)
{
    const int kind = in_level == 1 ? 10 : 20;
    const int slot = ksk_slot(kind);
    cd *bk = out_level == 1 ? ctx->bk1 : ctx->bk2;
    if (!bk || !ctx->ksk[slot]) {
        set_error("bootstrap operator: lvl0%d BK and KSK %d must be loaded", out_level, kind);
        return FHE_ERR_NOT_LOADED;
    }
    if (B == 0) return FHE_OK;
    int rc = FHE_OK;
    if (!fixed) rc = grow(ctx->tmp0, ctx->tmp0_cap, size_t(B) * (ctx->p.n0 + 1));
    if (rc) return rc;
    uint32_t *tmp0 = fixed ? fixed->tmp0 : ctx->tmp0;
    uint32_t *tv1 = fixed ? fixed->tv1 : ctx->tv1;
    const int n0 = int(ctx->p.n0);
    HIP_TRY(launch_iks(kind, in, ctx->ksk[slot], tmp0, n0, B, s, iks_tile(ctx, kind, B)), "IKS");
    if (out_level == 1) {
        HIP_TRY(tv.ide ? launch_tv_ide<uint32_t>(tv1, uint32_t(tv.value), Lvl1::N, s)
                       : launch_tv_const<uint32_t>(tv1, uint32_t(tv.value), Lvl1::N, s),
                "test vector");
        rc = launch_pbs1(ctx, tmp0, tv1, 0, static_cast<uint32_t *>(out), true, B, s);
        if (rc) return rc;
        if (post)
            HIP_TRY(launch_tlwe_add_body<uint32_t>(static_cast<uint32_t *>(out), Lvl1::N + 1,
                                                   uint32_t(post), B, s),
                    "post offset");
    } else {
        HIP_TRY(tv.ide ? launch_tv_ide<uint64_t>(ctx->tv2, uint32_t(tv.value), Lvl2::N, s)
                       : launch_tv_const<uint64_t>(ctx->tv2, tv.value, Lvl2::N, s),
                "test vector");
        HIP_TRY(launch_blind_rotate2q(tmp0, n0, ctx->bk2, ctx->rot2, ctx->tv2, 0,
                                     static_cast<uint64_t *>(out), true, B, s),
                "pbs lvl02");
        if (post)
            HIP_TRY(launch_tlwe_add_body<uint64_t>(static_cast<uint64_t *>(out), Lvl2::N + 1, post,
                                                   B, s),
                    "post offset");
    }
    return FHE_OK;
}

// in + pre (on the body) into level-matched scratch; returns the scratch
int offset_input(fhe_ctx *ctx, int in_level, const void *in, uint64_t pre, int B, hipStream_t s,
                 const void **staged)
{
    if (in_level == 1) {
        int rc = grow(ctx->tmp1, ctx->tmp1_cap, size_t(B) * (Lvl1::N + 1));
        if (rc) return rc;
        HIP_TRY(launch_tlwe_copy_add_body<uint32_t>(static_cast<const uint32_t *>(in), ctx->tmp1,
                                                    Lvl1::N + 1, uint32_t(pre), B, s),
                "input offset");
        *staged = ctx->tmp1;
    } else {
        int rc = grow(ctx->tmp2, ctx->tmp2_cap, size_t(B) * (Lvl2::N + 1) * 2);
        if (rc) return rc;
        uint64_t *t2 = reinterpret_cast<uint64_t *>(ctx->tmp2);
        HIP_TRY(launch_tlwe_copy_add_body<uint64_t>(static_cast<const uint64_t *>(in), t2,
                                                    Lvl2::N + 1, pre, B, s),
                "input offset");
        *staged = t2;
    }
    return FHE_OK;
}

// c = sa*a + sb*b + off -> IKS lvl10 -> PBS(tv) [-> body += post]
int gate_pipeline(fhe_ctx *ctx, const uint32_t *a, const uint32_t *b, int sa, int sb, uint32_t off,
                  uint32_t u, bool ide, uint32_t scale_bits, uint32_t post, uint32_t *out, int B,
                  hipStream_t s, const GateScratch *fixed = nullptr)
{
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (!fixed) {
        int rc = grow(ctx->tmp1, ctx->tmp1_cap, size_t(B) * (Lvl1::N + 1));
        if (rc) return rc;
    }
    uint32_t *tmp1 = fixed ? fixed->tmp1 : ctx->tmp1;
    HIP_TRY(launch_tlwe1_lincomb(a, b ? b : a, sa, b ? sb : 0, off, tmp1, B, s), "lincomb");
    return run_bootstrap(ctx, 1, tmp1, 1, out, TvSpec{ide, ide ? scale_bits : u}, post, B, s,
                         fixed);
}

}  // namespace

}  // extern "C"

namespace fhe {

int gate_lvl1(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
              uint32_t *out_dev, int result_type, int B, hipStream_t s, const GateScratch *fixed)
{
    const uint32_t mu = Lvl1::mu;
    const bool logic = result_type != 0;
    switch (gate) {
    case FHE_GATE_NAND:  // HomGate<-1,-1,+mu> + GateBootstrapping (TFHEpp gate.cpp:50-65)
        return gate_pipeline(ctx, a_dev, b_dev, -1, -1, mu, mu, false, 0, 0, out_dev, B, s, fixed);
    case FHE_GATE_NOR:
        return gate_pipeline(ctx, a_dev, b_dev, -1, -1, 0u - mu, mu, false, 0, 0, out_dev, B, s,
                             fixed);
    case FHE_GATE_XNOR:
        return gate_pipeline(ctx, a_dev, b_dev, -2, -2, 0u - 2 * mu, mu, false, 0, 0, out_dev, B, s,
                             fixed);
    case FHE_GATE_AND: {  // cuHEDB::HomAND: a + b - mu/2, tv = offset
        const uint32_t u = logic ? mu : mu << 1;
        return gate_pipeline(ctx, a_dev, b_dev, 1, 1, 0u - (mu >> 1), u, false, 0, logic ? 0 : u,
                             out_dev, B, s, fixed);
    }
    case FHE_GATE_OR: {  // cuHEDB::HomOR: a + b + mu/2
        const uint32_t u = logic ? mu : mu << 1;
        return gate_pipeline(ctx, a_dev, b_dev, 1, 1, mu >> 1, u, false, 0, logic ? 0 : u, out_dev,
                             B, s, fixed);
    }
    default:
        set_error("fhe_gate_lvl1_dev: unknown gate %d", gate);
        return FHE_ERR_ARG;
    }
}

}  // namespace fhe

extern "C" {

int fhe_gate_lvl1_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                      uint32_t *out_dev, int result_type, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && a_dev && b_dev && out_dev && B >= 0, "fhe_gate_lvl1_dev: bad argument");
    return gate_lvl1(ctx, gate, a_dev, b_dev, out_dev, result_type, B, pick(ctx, stream), nullptr);
}

int fhe_gate_rescale_lvl1_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                              uint32_t *out_dev, uint32_t scale_bits, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && a_dev && b_dev && out_dev && B >= 0,
                  "fhe_gate_rescale_lvl1_dev: bad argument");
    FHE_CHECK_ARG(gate == FHE_GATE_AND || gate == FHE_GATE_OR,
                  "fhe_gate_rescale_lvl1_dev: gate must be AND or OR (got %d)", gate);
    FHE_CHECK_ARG(scale_bits >= 1 && scale_bits <= 31,
                  "fhe_gate_rescale_lvl1_dev: scale_bits in [1, 31]");
    // HomAND / HomOR's combination a + b -+ mu/2, then log_rescale's test
    // vector m = 2^(scale_bits-1) and post-offset m: true -> 2^scale_bits,
    // false -> 0
    const uint32_t half = Lvl1::mu >> 1;
    const uint32_t m = 1u << (scale_bits - 1);
    return gate_pipeline(ctx, a_dev, b_dev, 1, 1, gate == FHE_GATE_AND ? 0u - half : half, m,
                         false, 0, m, out_dev, B, pick(ctx, stream));
}

int fhe_msb_gate_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev, int result_type,
                          int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_msb_gate_lvl1_dev: bad argument");
    // cuHEDB::MSBGateBootstrapping: +2^26 offset, tv = -u, ARITHMETIC adds u
    const uint32_t u = result_type ? Lvl1::mu : Lvl1::mu << 1;
    return gate_pipeline(ctx, in_dev, nullptr, 1, 0, 1u << 26, 0u - u, false, 0,
                         result_type ? 0 : u, out_dev, B, pick(ctx, stream));
}

int fhe_ide_gate_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev,
                          uint32_t scale_bits, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_ide_gate_lvl1_dev: bad argument");
    FHE_CHECK_ARG(scale_bits < 32, "fhe_ide_gate_lvl1_dev: scale_bits must be < 32");
    // cuHEDB::IdeGateBootstrapping: +2^26 offset, tv = gpolygen(4, scale_bits)
    return gate_pipeline(ctx, in_dev, nullptr, 1, 0, 1u << 26, 0, true, scale_bits, 0, out_dev, B,
                         pick(ctx, stream));
}

int fhe_msb_gate_dev(fhe_ctx *ctx, int in_level, int out_level, const void *in_dev, void *out_dev,
                     int result_type, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_msb_gate_dev: bad argument");
    FHE_CHECK_ARG((in_level == 1 || in_level == 2) && (out_level == 1 || out_level == 2) &&
                      !(in_level == 1 && out_level == 2),
                  "fhe_msb_gate_dev: levels (1,1), (2,1) or (2,2)");
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    // gatebootstrap_gpu.cuh:31-80: offset 2^(w-6) (lvl1 output) / 2^(w-7)
    // (lvl2 output), w = input torus width; tv = -u; ARITHMETIC adds u
    const int w = in_level == 1 ? 32 : 64;
    const uint64_t pre = uint64_t(1) << (w - (out_level == 1 ? 6 : 7));
    const uint64_t mu = out_level == 1 ? uint64_t(Lvl1::mu) : Lvl2::mu;
    const uint64_t u = result_type ? mu : mu << 1;
    const uint64_t mask = out_level == 1 ? 0xFFFFFFFFull : ~0ull;
    const void *staged = nullptr;
    int rc = offset_input(ctx, in_level, in_dev, pre, B, s, &staged);
    if (rc) return rc;
    return run_bootstrap(ctx, in_level, staged, out_level, out_dev, TvSpec{false, (0 - u) & mask},
                         result_type ? 0 : u, B, s);
}

int fhe_ide_gate_dev(fhe_ctx *ctx, int in_level, int out_level, const void *in_dev, void *out_dev,
                     uint32_t scale_bits, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_ide_gate_dev: bad argument");
    FHE_CHECK_ARG((in_level == 1 || in_level == 2) && (out_level == 1 || out_level == 2) &&
                      !(in_level == 1 && out_level == 2),
                  "fhe_ide_gate_dev: levels (1,1), (2,1) or (2,2)");
    // gatebootstrap_gpu.cuh:82-115: lvl2 -> lvl1 uses scale_bits - 32
    uint32_t sb = scale_bits;
    if (in_level == 2 && out_level == 1) {
        FHE_CHECK_ARG(scale_bits >= 32 && scale_bits < 64, "fhe_ide_gate_dev: scale_bits in [32, 64)");
        sb = scale_bits - 32;
    }
    FHE_CHECK_ARG(sb < (out_level == 1 ? 32u : 64u), "fhe_ide_gate_dev: scale_bits too large");
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    const int w = in_level == 1 ? 32 : 64;
    const uint64_t pre = uint64_t(1) << (w - (out_level == 1 ? 6 : 7));
    const void *staged = nullptr;
    int rc = offset_input(ctx, in_level, in_dev, pre, B, s, &staged);
    if (rc) return rc;
    return run_bootstrap(ctx, in_level, staged, out_level, out_dev, TvSpec{true, sb}, 0, B, s);
}

int fhe_log_to_ari_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev, int B,
                            void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_log_to_ari_lvl1_dev: bad argument");
    // HEDB LOG_to_ARI (tfhepp_utils.cpp:18-26): test vector mu_polygen(-2 mu)
    // = +2 mu, body += mu, so true -> 3/8 and false -> -1/8 (1 and 0 at
    // scale 1/2).  The reference's GPU twin (gatebootstrap_gpu.cuh:155-162)
    // passes -2 mu as the constant itself, which inverts the result.
    return gate_pipeline(ctx, in_dev, nullptr, 1, 0, 0, Lvl1::mu << 1, false, 0, Lvl1::mu, out_dev,
                         B, pick(ctx, stream));
}

int fhe_log_rescale_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev,
                             uint32_t scale_bits, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_log_rescale_lvl1_dev: bad argument");
    FHE_CHECK_ARG(scale_bits >= 1 && scale_bits <= 32, "fhe_log_rescale_lvl1_dev: scale_bits in [1, 32]");
    // HEDB log_rescale (tfhepp_utils.cpp:28-35): mu' = 2^(scale_bits-1), test
    // vector mu_polygen(-mu') = +mu', body += mu': true -> 2^scale_bits,
    // false -> 0 (the GPU twin, gatebootstrap_gpu.cuh:164-170, inverts it)
    const uint32_t m = uint32_t(1ull << (scale_bits - 1));
    return gate_pipeline(ctx, in_dev, nullptr, 1, 0, 0, m, false, 0, m, out_dev, B,
                         pick(ctx, stream));
}

// ---- host-pointer forms ---------------------------------------------------

int fhe_pbs_lvl01(fhe_ctx *ctx, const uint32_t *in, const uint32_t *tv, int tv_stride,
                  uint32_t *out, int B)
{
    FHE_CHECK_ARG(ctx && in && tv && out && B >= 0, "fhe_pbs_lvl01: bad argument");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl1::N, "tv_stride must be 0 or %d", Lvl1::N);
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t in_b = size_t(B) * (ctx->p.n0 + 1) * 4, out_b = size_t(B) * (Lvl1::N + 1) * 4;
    const size_t tv_b = (tv_stride ? size_t(B) : 1) * Lvl1::N * 4;
    void *din = nullptr, *dtv = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&din, in_b);
    if (e == hipSuccess) e = hipMalloc(&dtv, tv_b);
    if (e == hipSuccess) e = hipMalloc(&dout, out_b);
    if (e == hipSuccess) e = hipMemcpyAsync(din, in, in_b, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dtv, tv, tv_b, hipMemcpyHostToDevice, ctx->stream);
    int rc = FHE_OK;
    if (e == hipSuccess)
        rc = fhe_pbs_lvl01_dev(ctx, static_cast<uint32_t *>(din), static_cast<uint32_t *>(dtv),
                               tv_stride, static_cast<uint32_t *>(dout), B, nullptr);
    if (e == hipSuccess && rc == FHE_OK)
        e = hipMemcpyAsync(out, dout, out_b, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (din) (void)hipFree(din);
    if (dtv) (void)hipFree(dtv);
    if (dout) (void)hipFree(dout);
    if (rc != FHE_OK) return rc;
    HIP_TRY(e, "fhe_pbs_lvl01");
    return FHE_OK;
}

int fhe_iks(fhe_ctx *ctx, int kind, const void *in, uint32_t *out, int B)
{
    FHE_CHECK_ARG(ctx && in && out && B >= 0, "fhe_iks: bad argument");
    const int slot = ksk_slot(kind);
    FHE_CHECK_ARG(slot >= 0, "fhe_iks: kind must be 10, 20 or 21 (got %d)", kind);
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t in_b = size_t(B) * ks_domain_words(kind) * (kind == 10 ? 4 : 8);
    const size_t out_b = size_t(B) * (ks_target_n(ctx, kind) + 1) * 4;
    void *din = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&din, in_b);
    if (e == hipSuccess) e = hipMalloc(&dout, out_b);
    if (e == hipSuccess) e = hipMemcpyAsync(din, in, in_b, hipMemcpyHostToDevice, ctx->stream);
    int rc = FHE_OK;
    if (e == hipSuccess) rc = fhe_iks_dev(ctx, kind, din, static_cast<uint32_t *>(dout), B, nullptr);
    if (e == hipSuccess && rc == FHE_OK)
        e = hipMemcpyAsync(out, dout, out_b, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (rc != FHE_OK) return rc;
    HIP_TRY(e, "fhe_iks");
    return FHE_OK;
}

// ---- memory / timing helpers ---------------------------------------------

int fhe_malloc(void **ptr, size_t bytes)
{
    FHE_CHECK_ARG(ptr, "fhe_malloc: null pointer");
    *ptr = nullptr;
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
    return FHE_OK;
}

int fhe_free(void *ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr), "hipFree");
    return FHE_OK;
}

int fhe_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    FHE_CHECK_ARG(dst && src, "fhe_memcpy_h2d: null pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return FHE_OK;
}

int fhe_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    FHE_CHECK_ARG(dst && src, "fhe_memcpy_d2h: null pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return FHE_OK;
}

int fhe_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    FHE_CHECK_ARG(dst && src, "fhe_memcpy_d2d: null pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    return FHE_OK;
}

int fhe_memset_dev(void *dst, int value, size_t bytes)
{
    FHE_CHECK_ARG(dst, "fhe_memset_dev: null pointer");
    HIP_TRY(hipMemset(dst, value, bytes), "hipMemset");
    return FHE_OK;
}

int fhe_device_sync(void)
{
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return FHE_OK;
}

int fhe_timer_start(fhe_ctx *ctx, void *stream)
{
    FHE_CHECK_ARG(ctx, "fhe_timer_start: null context");
    HIP_TRY(hipEventRecord(ctx->ev0, pick(ctx, stream)), "hipEventRecord");
    return FHE_OK;
}

int fhe_timer_stop(fhe_ctx *ctx, void *stream, float *ms)
{
    FHE_CHECK_ARG(ctx && ms, "fhe_timer_stop: null argument");
    HIP_TRY(hipEventRecord(ctx->ev1, pick(ctx, stream)), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(ctx->ev1), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
    return FHE_OK;
}

int fhe_device_name(char *buf, int len)
{
    FHE_CHECK_ARG(buf && len > 0, "fhe_device_name: bad buffer");
    hipDeviceProp_t prop;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    snprintf(buf, size_t(len), "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
             prop.multiProcessorCount);
    return FHE_OK;
}

}  // extern "C"
