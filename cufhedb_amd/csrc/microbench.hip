// microbench.hip -- measured ceilings for the roofline denominators: the
// FP64 fused multiply-add rate of the vector ALUs (the unit both blind
// rotations are bound by) and the device-to-device streaming bandwidth (the
// unit the key switches are bound by), each timed with HIP events on the
// context stream.  No reference counterpart.
#include "ctx.hpp"

using namespace fhe;

namespace {

constexpr int kChains = 16;  // independent FMA chains per lane

__global__ __launch_bounds__(256) void k_fma_f64(double *__restrict__ out, int iters, double a,
                                                 double b)
{
    double x[kChains];
    const double t = double(threadIdx.x) * 1e-6;
#pragma unroll
    for (int k = 0; k < kChains; k++) x[k] = t + k;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < kChains; k++) x[k] = __fma_rn(x[k], a, b);
    }
    double s = 0;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += x[k];
    out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

// four independent 16-byte loads in flight per lane; n is a multiple of
// 4 * gridDim.x * blockDim.x
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src,
                                                uint4 *__restrict__ dst, size_t n)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += 4 * stride) {
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = src[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; k++) dst[i + k * stride] = v[k];
    }
}

}  // namespace

extern "C" {

int fhe_measure_peaks(fhe_ctx *ctx, double *fp64_tflops, double *copy_gbs)
{
    FHE_CHECK_ARG(ctx && fp64_tflops && copy_gbs, "fhe_measure_peaks: null argument");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
    hipStream_t s = ctx->stream;
    // FP64 FMA: 8 workgroups of 4 waves per CU, 16 chains per lane
    const int blocks = prop.multiProcessorCount * 8, iters = 1 << 16;
    double *out = nullptr;
    HIP_TRY(hipMalloc(&out, size_t(blocks) * 256 * sizeof(double)), "hipMalloc");
    float ms = 0;
    hipError_t e = hipSuccess;
    for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // first run warms up
        e = hipEventRecord(ctx->ev0, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_fma_f64, dim3(blocks), dim3(256), 0, s, out, iters, 0.999999,
                               1e-7);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    }
    (void)hipFree(out);
    HIP_TRY(e, "FP64 FMA rate");
    *fp64_tflops = 2.0 * kChains * double(iters) * blocks * 256 / (double(ms) * 1e9);
    // streaming copy of 512 MiB: bytes read plus bytes written
    const size_t bytes = size_t(512) << 20, n = bytes / sizeof(uint4);
    const int copy_blocks = 4096;  // n = 2^25 is a multiple of 4 x 4096 x 256
    void *src = nullptr, *dst = nullptr;
    e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, s);
    for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {
        e = hipEventRecord(ctx->ev0, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_copy16, dim3(copy_blocks), dim3(256), 0, s,
                               static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    }
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    HIP_TRY(e, "copy bandwidth");
    *copy_gbs = 2.0 * double(bytes) / (double(ms) * 1e6);
    return FHE_OK;
}

}  // extern "C"
