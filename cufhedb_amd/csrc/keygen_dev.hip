// keygen_dev.hip -- key-switching key generation on the device.
//
// ikskgen (thirdparty/TFHEpp/src/cloudkey.cpp:54-70) is one TLWE encryption
// per (i, j, k): 2.1 M records of n0+1 words for lvl10 (5.6 GB), 4.2 M for
// lvl20 (11.3 GB), 143 k records of 1025 words for lvl21.  The host generator
// (keygen.cpp) draws record r = (i t + j)(2^basebit - 1) + k from the Philox
// stream (seed, tag, r): words 0..tn-1 are the mask and words tn..tn+3 feed
// one Box-Muller normal for the noise.  Here the device draws the masks and
// forms b = <a, s> + m (Philox and the dot product are integer arithmetic,
// identical on host and device) and the host supplies each record's noise
// word, evaluated with the same libm calls as the host generator, so the
// device key is bit-identical to fhe_keygen_ksk's while the host does one
// Philox block per record instead of tn/4 + 1 and uploads 4 bytes per record
// instead of 4 (tn + 1).
//
// Layout: one wave per record; lane l draws Philox blocks l, l+64, ... of the
// mask (4 words each), writes them and accumulates a * s; a wave reduction
// gives <a, s>, lane 0 writes b.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.hpp"
#include "rng.hpp"

namespace fhe {
namespace {

struct KsShape {
    int dn, tn, t, basebit;
    uint32_t tag;
    double alpha;
};

bool ks_shape(const fhe_params &p, int kind, KsShape &g)
{
    switch (kind) {
    case 10: g = {Lvl1::N, int(p.n0), KS10::t, KS10::basebit, TAG_KSK10, p.alpha0}; return true;
    case 20: g = {Lvl2::N, int(p.n0), KS20::t, KS20::basebit, TAG_KSK20, p.alpha0}; return true;
    case 21: g = {Lvl2::N, Lvl1::N, KS21::t, KS21::basebit, TAG_KSK21, p.alpha1}; return true;
    default: return false;
    }
}

__global__ __launch_bounds__(256) void k_ksk_records(uint64_t seed, uint32_t tag, int tn, int t,
                                                     int basebit, const uint32_t *__restrict__ sdom,
                                                     const uint32_t *__restrict__ stgt,
                                                     const uint32_t *__restrict__ noise,
                                                     uint32_t *__restrict__ ksk, uint64_t records)
{
    const uint64_t rec = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (rec >= records) return;
    const int lane = threadIdx.x & 63;
    const Philox ph(seed);
    const uint32_t rlo = uint32_t(rec), rhi = uint32_t(rec >> 32);
    const size_t row = size_t(tn) + 1;
    uint32_t *out = ksk + rec * row;
    uint32_t acc = 0;
    const int blocks = (tn + 3) >> 2;
    for (int bi = lane; bi < blocks; bi += 64) {
        uint32_t w[4];
        ph.block(uint32_t(bi), rlo, rhi, tag, w);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int c = 4 * bi + e;
            if (c < tn) {
                out[c] = w[e];
                acc += w[e] * stgt[c];
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        const uint64_t nd = (uint64_t(1) << basebit) - 1;
        const uint32_t k = uint32_t(rec % nd);
        const uint64_t ij = rec / nd;
        const uint32_t j = uint32_t(ij % uint64_t(t));
        const uint32_t i = uint32_t(ij / uint64_t(t));
        const uint32_t m = sdom[i] * (k + 1) * (1u << (32 - (j + 1) * basebit));
        out[tn] = acc + m + noise[rec];
    }
}

struct NoiseJob {
    uint64_t seed;
    KsShape g;
    uint32_t *noise;
};

void noise_worker(int64_t b, int64_t e, void *arg)
{
    auto *j = static_cast<NoiseJob *>(arg);
    for (int64_t r = b; r < e; r++) {
        Stream s(j->seed, j->g.tag, uint64_t(r));
        s.seek(uint64_t(j->g.tn));
        j->noise[r] = gaussian_t32(s, j->g.alpha);
    }
}

}  // namespace
}  // namespace fhe

using namespace fhe;

extern "C" {

int fhe_keygen_ksk_dev(fhe_ctx *ctx, int kind, uint64_t seed, const void *s_dom,
                       const uint32_t *s_tgt, uint32_t *ksk_dev)
{
    FHE_CHECK_ARG(ctx && s_dom && s_tgt, "fhe_keygen_ksk_dev: null argument");
    KsShape g;
    FHE_CHECK_ARG(ks_shape(ctx->p, kind, g), "fhe_keygen_ksk_dev: kind must be 10, 20 or 21 (got %d)",
                  kind);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const uint64_t records = uint64_t(g.dn) * g.t * ((uint64_t(1) << g.basebit) - 1);
    uint32_t *dst = ksk_dev;
    if (!dst) {
        if (int rc = alloc_ksk(ctx, kind)) return rc;
        dst = ctx->ksk[ksk_slot(kind)];
    }
    // host side: the domain key as 32-bit words (a lvl2 key is 0/1 in 64
    // bits; the host generator truncates the same way) and the noise words
    std::vector<uint32_t> sdom(g.dn);
    for (int i = 0; i < g.dn; i++)
        sdom[i] = kind == 10 ? static_cast<const uint32_t *>(s_dom)[i]
                             : uint32_t(static_cast<const uint64_t *>(s_dom)[i]);
    std::vector<uint32_t> noise(records);
    NoiseJob job{seed, g, noise.data()};
    parallel_for(int64_t(records), host_threads(), noise_worker, &job);

    uint32_t *d_sdom = nullptr, *d_stgt = nullptr, *d_noise = nullptr;
    hipError_t e = hipMalloc(&d_sdom, sdom.size() * 4);
    if (e == hipSuccess) e = hipMalloc(&d_stgt, size_t(g.tn) * 4);
    if (e == hipSuccess) e = hipMalloc(&d_noise, records * 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sdom, sdom.data(), sdom.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_stgt, s_tgt, size_t(g.tn) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_noise, noise.data(), records * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        const unsigned grid = unsigned((records + 3) / 4);
        hipLaunchKernelGGL(k_ksk_records, dim3(grid), dim3(256), 0, ctx->stream, seed, g.tag, g.tn,
                           g.t, g.basebit, d_sdom, d_stgt, d_noise, dst, records);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    if (d_sdom) (void)hipFree(d_sdom);
    if (d_stgt) (void)hipFree(d_stgt);
    if (d_noise) (void)hipFree(d_noise);
    HIP_TRY(e, "fhe_keygen_ksk_dev");
    return FHE_OK;
}

}  // extern "C"
