// manylut.hip -- programmable bootstrap with several lookup tables per
// blind rotation (TFHEpp GateBootstrappingManyLUT<P, num_out>,
// thirdparty/TFHEpp/include/gatebootstrapping.hpp:119-130).
//
// BlindRotate<P, num_out> (gatebootstrapping.hpp:15-43) switches the lvl0
// ciphertext to Z_2N on multiples of 2^bw, bw = bits_needed(num_out - 1)
// (utils.hpp:69-81):
//   bbar = 2N - ((b >> sh) << bw)                     (no rounding)
//   abar = ((a + 2^(sh - 1)) >> sh) << bw              (rounded),
//   sh = 32 - 1 - nbit + bw,
// so the test vector's coefficients k, k + 2^bw, k + 2*2^bw, ... form the
// k-th table, and SampleExtractIndex at k = 0 .. num_out-1 reads them all.
//
// The same switch comes out of the single-table blind rotation when the
// lvl0 words are first rounded to multiples of 2^sh: with a' = ((a +
// 2^(sh-1)) >> sh) << sh, the kernel's (a' + 2^(sh-bw-1)) >> (sh - bw) adds
// less than one unit below bit sh and returns (a' >> sh) << bw exactly, and
// b' = (b >> sh) << sh gives b' >> (sh - bw) = (b >> sh) << bw.  So a
// many-table bootstrap here is one elementwise rounding pass over the lvl0
// batch, the tuned blind rotation kernels unchanged (accumulator out), and
// one extraction pass writing num_out TLWEs per ciphertext.
#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

constexpr int kMaxOut = 64;

int bits_needed(uint32_t v)
{
    int b = 0;
    while (b < 32 && (v >> b)) b++;
    return b;
}

__global__ __launch_bounds__(256) void k_many_round(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, int n0, int sh,
                                                    size_t total)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const uint32_t v = in[gid];
    const bool body = int(gid % size_t(n0 + 1)) == n0;
    out[gid] = body ? (v >> sh) << sh : ((v + (1u << (sh - 1))) >> sh) << sh;
}

// out[b][k] = SampleExtractIndex(acc[b], k) for k < num_out
// (thirdparty/TFHEpp/src/trlwe.cpp:109-118)
template <class T, int N>
__global__ __launch_bounds__(256) void k_extract_many(const T *__restrict__ acc,
                                                      T *__restrict__ out, int num_out,
                                                      size_t total)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int i = int(gid % (N + 1));
    const size_t rec = gid / (N + 1);
    const int k = int(rec % size_t(num_out));
    const T *a = acc + (rec / size_t(num_out)) * 2 * N;
    T v;
    if (i == N)
        v = a[N + k];
    else if (i <= k)
        v = a[k - i];
    else
        v = T(0) - a[N + k - i];
    out[gid] = v;
}

unsigned blocks(size_t n) { return unsigned((n + 255) / 256); }

int many_scratch(fhe_ctx *ctx, int slot, size_t words, uint32_t **p)
{
    int rc = grow(ctx->many[slot], ctx->many_cap[slot], words);
    *p = ctx->many[slot];
    return rc;
}

// rounds the batch into scratch slot 0; returns the shift or a negative code
int round_input(fhe_ctx *ctx, const uint32_t *in_dev, int nbit, int num_out, int B,
                hipStream_t s, const uint32_t **rounded)
{
    const int n0 = int(ctx->p.n0);
    const int sh = 32 - 1 - nbit + bits_needed(uint32_t(num_out - 1));
    uint32_t *r = nullptr;
    const size_t total = size_t(B) * (n0 + 1);
    int rc = many_scratch(ctx, 0, total, &r);
    if (rc) return rc;
    hipLaunchKernelGGL(k_many_round, dim3(blocks(total)), dim3(256), 0, s, in_dev, r, n0, sh,
                       total);
    HIP_TRY(hipGetLastError(), "many-LUT input rounding");
    *rounded = r;
    return FHE_OK;
}

}  // namespace

extern "C" {

int fhe_pbs_many_lut_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev,
                               int tv_stride, int num_out, uint32_t *out_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && out_dev, "fhe_pbs_many_lut_lvl01_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_pbs_many_lut_lvl01_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl1::N, "tv_stride must be 0 or %d", Lvl1::N);
    FHE_CHECK_ARG(num_out >= 1 && num_out <= kMaxOut,
                  "fhe_pbs_many_lut_lvl01_dev: num_out must be in [1, %d] (got %d)", kMaxOut,
                  num_out);
    if (!ctx->bk1) {
        set_error("fhe_pbs_many_lut_lvl01_dev: lvl01 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    const uint32_t *rounded = nullptr;
    int rc = round_input(ctx, in_dev, Lvl1::nbit, num_out, B, s, &rounded);
    if (rc) return rc;
    uint32_t *acc = nullptr;
    rc = many_scratch(ctx, 1, size_t(B) * 2 * Lvl1::N, &acc);
    if (rc) return rc;
    rc = launch_pbs1(ctx, rounded, tv_dev, tv_stride, acc, false, B, s);
    if (rc) return rc;
    const size_t total = size_t(B) * num_out * (Lvl1::N + 1);
    hipLaunchKernelGGL((k_extract_many<uint32_t, Lvl1::N>), dim3(blocks(total)), dim3(256), 0, s,
                       acc, out_dev, num_out, total);
    HIP_TRY(hipGetLastError(), "many-LUT sample extraction lvl1");
    return FHE_OK;
}

int fhe_pbs_many_lut_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev,
                               int tv_stride, int num_out, uint64_t *out_dev, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && tv_dev && out_dev, "fhe_pbs_many_lut_lvl02_dev: null argument");
    FHE_CHECK_ARG(B >= 0, "fhe_pbs_many_lut_lvl02_dev: negative batch");
    FHE_CHECK_ARG(tv_stride == 0 || tv_stride == Lvl2::N, "tv_stride must be 0 or %d", Lvl2::N);
    FHE_CHECK_ARG(num_out >= 1 && num_out <= kMaxOut,
                  "fhe_pbs_many_lut_lvl02_dev: num_out must be in [1, %d] (got %d)", kMaxOut,
                  num_out);
    if (!ctx->bk2) {
        set_error("fhe_pbs_many_lut_lvl02_dev: lvl02 bootstrapping key not loaded");
        return FHE_ERR_NOT_LOADED;
    }
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    const uint32_t *rounded = nullptr;
    int rc = round_input(ctx, in_dev, Lvl2::nbit, num_out, B, s, &rounded);
    if (rc) return rc;
    uint32_t *acc32 = nullptr;
    rc = many_scratch(ctx, 1, size_t(B) * 2 * Lvl2::N * 2, &acc32);
    if (rc) return rc;
    uint64_t *acc = reinterpret_cast<uint64_t *>(acc32);
    HIP_TRY(launch_blind_rotate2q(rounded, int(ctx->p.n0), ctx->bk2, ctx->rot2, tv_dev, tv_stride,
                                  acc, false, B, s),
            "blind rotate lvl02");
    const size_t total = size_t(B) * num_out * (Lvl2::N + 1);
    hipLaunchKernelGGL((k_extract_many<uint64_t, Lvl2::N>), dim3(blocks(total)), dim3(256), 0, s,
                       acc, out_dev, num_out, total);
    HIP_TRY(hipGetLastError(), "many-LUT sample extraction lvl2");
    return FHE_OK;
}

}  // extern "C"
