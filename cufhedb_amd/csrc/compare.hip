// compare.hip -- the comparison layer above the bootstrap operators: HomMSB
// chains, the five comparisons and HomNOT, each batched over B ciphertexts.
//
// Reference:
//   ExtractMSB5/9/10, HomMSB (lvl1)     src/cuHEDB/msb_gpu.cuh:7-58
//   ImExtractMSB5..33, HomMSB (lvl2)    src/cuHEDB/msb_gpu.cuh:60-192
//   greater_than .. equal               src/cuHEDB/HomCompare_gpu.cuh:14-76
//   HomNOT                              src/cuHEDB/gatebootstrap_gpu.cuh:117-121
//   CPU twins                           src/HEDB/comparison/extract_msb.cpp:5-259
// The reference runs these per ciphertext on a host thread, with the key
// switches on the CPU.  Here every step of a chain is one batched device
// operation over the whole batch on the context stream:
//   peel(x, pb):  sh = x << (pb - k);  g = MSB_ARITHMETIC(sh);  sh -= g;
//                 sh = IDE(sh, w - pb);  x -= sh            (pb -> pb - 5)
// with k = 5 on the 32-bit torus and k = 6 on the 64-bit torus, until the
// remaining bits fit one MSB gate.
#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

constexpr int W1 = Lvl1::N + 1;
constexpr int W2 = Lvl2::N + 1;

// chain scratch slots
enum { kX2 = 0, kS2 = 1, kG2 = 2, kD = 3, kS1 = 4, kG1 = 5, kE = 6 };

int chain_buf(fhe_ctx *ctx, int slot, size_t words, uint32_t **p)
{
    int rc = grow(ctx->chain[slot], ctx->chain_cap[slot], words);
    *p = ctx->chain[slot];
    return rc;
}

#define FHE_TRY(expr)              \
    do {                           \
        int _rc = (expr);          \
        if (_rc) return _rc;       \
    } while (0)

// ExtractMSB9/10 first half (msb_gpu.cuh:11-31): out = x - IDE(sh - MSB(sh))
// with sh = x << (pb - 5); out may alias x.
int peel1(fhe_ctx *ctx, const uint32_t *x, uint32_t *out, uint32_t pb, int B, hipStream_t s)
{
    uint32_t *S, *G;
    FHE_TRY(chain_buf(ctx, kS1, size_t(B) * W1, &S));
    FHE_TRY(chain_buf(ctx, kG1, size_t(B) * W1, &G));
    HIP_TRY(launch_tlwe_shl<uint32_t>(x, S, W1, int(pb) - 5, B, s), "chain shift");
    FHE_TRY(fhe_msb_gate_dev(ctx, 1, 1, S, G, FHE_ARITHMETIC, B, s));
    HIP_TRY(launch_tlwe_lincomb<uint32_t>(S, G, 1, -1, 0u, S, W1, false, B, s), "chain sub");
    FHE_TRY(fhe_ide_gate_dev(ctx, 1, 1, S, S, 32 - pb, B, s));
    HIP_TRY(launch_tlwe_lincomb<uint32_t>(x, S, 1, -1, 0u, out, W1, false, B, s), "chain sub");
    return FHE_OK;
}

// lvl2 peeling step (ImExtractMSB14..33, msb_gpu.cuh:70-171): shift by pb - 6,
// MSB and identity bootstrapping lvl2 -> lvl2; out may alias x.
int peel2(fhe_ctx *ctx, const uint64_t *x, uint64_t *out, uint32_t pb, int B, hipStream_t s)
{
    uint32_t *s32, *g32;
    FHE_TRY(chain_buf(ctx, kS2, size_t(B) * W2 * 2, &s32));
    FHE_TRY(chain_buf(ctx, kG2, size_t(B) * W2 * 2, &g32));
    uint64_t *S = reinterpret_cast<uint64_t *>(s32), *G = reinterpret_cast<uint64_t *>(g32);
    HIP_TRY(launch_tlwe_shl<uint64_t>(x, S, W2, int(pb) - 6, B, s), "chain shift");
    FHE_TRY(fhe_msb_gate_dev(ctx, 2, 2, S, G, FHE_ARITHMETIC, B, s));
    HIP_TRY(launch_tlwe_lincomb<uint64_t>(S, G, 1, -1, 0ull, S, W2, false, B, s), "chain sub");
    FHE_TRY(fhe_ide_gate_dev(ctx, 2, 2, S, S, 64 - pb, B, s));
    HIP_TRY(launch_tlwe_lincomb<uint64_t>(x, S, 1, -1, 0ull, out, W2, false, B, s), "chain sub");
    return FHE_OK;
}

// ExtractMSB9 (msb_gpu.cuh:11-31): one peeling step (4 bits) then the MSB
// gate; res and x may alias.
int extract_msb9(fhe_ctx *ctx, const uint32_t *x, uint32_t *res, uint32_t pb, int rt, int B,
                 hipStream_t s)
{
    FHE_TRY(peel1(ctx, x, res, pb, B, s));
    return fhe_msb_gate_dev(ctx, 1, 1, res, res, rt, B, s);
}

// HomMSB on lvl1 (msb_gpu.cuh:51-57): ExtractMSB5 / 9 / 10; res and x may alias.
int hom_msb1(fhe_ctx *ctx, const uint32_t *x, uint32_t *res, uint32_t pb, int rt, int B,
             hipStream_t s)
{
    if (pb <= 5) return fhe_msb_gate_dev(ctx, 1, 1, x, res, rt, B, s);
    if (pb <= 9) return extract_msb9(ctx, x, res, pb, rt, B, s);
    if (pb == 10) {  // ExtractMSB10: peel, then ExtractMSB9 with plain_bits - 4 (msb_gpu.cuh:48)
        FHE_TRY(peel1(ctx, x, res, pb, B, s));
        return extract_msb9(ctx, res, res, pb - 4, rt, B, s);
    }
    set_error("HomMSB lvl1: plain_bits %u out of range (at most 10)", pb);
    return FHE_ERR_ARG;
}

// HomMSB on lvl2 (msb_gpu.cuh:175-192), result lvl1.  ImExtractMSB33 .. 14
// each peel 5 bits at lvl2 and hand on plain_bits - 5; ImExtractMSB14 ends
// with IKS lvl21 and ExtractMSB9; ImExtractMSB9 / 5 switch right away.
int hom_msb2(fhe_ctx *ctx, const uint64_t *x, uint32_t *res, uint32_t pb, int rt, int B,
             hipStream_t s)
{
    if (pb > 33) {
        set_error("HomMSB lvl2: plain_bits %u out of range (at most 33)", pb);
        return FHE_ERR_ARG;
    }
    const uint64_t *cur = x;
    bool peeled = false;
    if (pb > 9) {
        uint32_t *x32;
        FHE_TRY(chain_buf(ctx, kX2, size_t(B) * W2 * 2, &x32));
        uint64_t *X = reinterpret_cast<uint64_t *>(x32);
        while (pb > 9) {
            FHE_TRY(peel2(ctx, cur, X, pb, B, s));
            cur = X;
            pb -= 5;
        }
        peeled = true;
    }
    FHE_TRY(fhe_iks_dev(ctx, 21, cur, res, B, s));
    if (peeled || pb > 5) return extract_msb9(ctx, res, res, pb, rt, B, s);
    return fhe_msb_gate_dev(ctx, 1, 1, res, res, rt, B, s);
}

int hom_msb(fhe_ctx *ctx, int level, const void *in, uint32_t *out, uint32_t pb, int rt, int B,
            hipStream_t s)
{
    return level == 1 ? hom_msb1(ctx, static_cast<const uint32_t *>(in), out, pb, rt, B, s)
                      : hom_msb2(ctx, static_cast<const uint64_t *>(in), out, pb, rt, B, s);
}

// d = sa*a + sb*b at the input level, into chain slot kD
int difference(fhe_ctx *ctx, int level, const void *a, const void *b, bool bcast, int sa, int sb,
               int B, hipStream_t s, const void **d)
{
    uint32_t *D;
    const size_t words = size_t(B) * (level == 1 ? W1 : 2 * W2);
    FHE_TRY(chain_buf(ctx, kD, words, &D));
    if (level == 1) {
        HIP_TRY(launch_tlwe_lincomb<uint32_t>(static_cast<const uint32_t *>(a),
                                              static_cast<const uint32_t *>(b), sa, sb, 0u, D, W1,
                                              bcast, B, s),
                "compare difference");
    } else {
        HIP_TRY(launch_tlwe_lincomb<uint64_t>(static_cast<const uint64_t *>(a),
                                              static_cast<const uint64_t *>(b), sa, sb, 0ull,
                                              reinterpret_cast<uint64_t *>(D), W2, bcast, B, s),
                "compare difference");
    }
    *d = D;
    return FHE_OK;
}

// greater_than / less_than (HomCompare_gpu.cuh:14-24, 40-49): MSB of the
// signed difference sa*a + sb*b with one extra (sign) bit
int cmp_strict(fhe_ctx *ctx, int level, const void *a, const void *b, bool bcast, int sa, int sb,
               uint32_t *out, uint32_t pb, int rt, int B, hipStream_t s)
{
    const void *d;
    FHE_TRY(difference(ctx, level, a, b, bcast, sa, sb, B, s, &d));
    return hom_msb(ctx, level, d, out, pb + 1, rt, B, s);
}

// greater_than_equal / less_than_equal (HomCompare_gpu.cuh:27-37, 52-61):
// NOT of a strict comparison, then LOG_to_ARI for an arithmetic result
int cmp_nonstrict(fhe_ctx *ctx, int level, const void *a, const void *b, bool bcast, int sa,
                  int sb, uint32_t *out, uint32_t pb, int rt, int B, hipStream_t s)
{
    const void *d;
    FHE_TRY(difference(ctx, level, a, b, bcast, sa, sb, B, s, &d));
    FHE_TRY(hom_msb(ctx, level, d, out, pb + 1, FHE_LOGIC, B, s));
    HIP_TRY(launch_tlwe_lincomb<uint32_t>(out, nullptr, -1, 0, 0u, out, W1, false, B, s), "HomNOT");
    if (rt == FHE_ARITHMETIC) return fhe_log_to_ari_lvl1_dev(ctx, out, out, B, s);
    return FHE_OK;
}

}  // namespace

extern "C" {

int fhe_hom_msb_dev(fhe_ctx *ctx, int in_level, const void *in_dev, uint32_t *out_dev,
                    uint32_t plain_bits, int result_type, int B, void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_hom_msb_dev: bad argument");
    FHE_CHECK_ARG(in_level == 1 || in_level == 2, "fhe_hom_msb_dev: level 1 or 2");
    FHE_CHECK_ARG(plain_bits >= 1, "fhe_hom_msb_dev: plain_bits >= 1");
    if (in_level == 2 && plain_bits > kMaxMsbBits2) {
        set_error("fhe_hom_msb_dev: plain_bits %u above %d at level 2: the peeling steps shift "
                  "the lvl2 bootstrap noise onto the decision margin (wrong results from 28 bits on)",
                  plain_bits, kMaxMsbBits2);
        return FHE_ERR_UNSUPPORTED;
    }
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    return hom_msb(ctx, in_level, in_dev, out_dev, plain_bits, result_type ? FHE_LOGIC : FHE_ARITHMETIC,
                   B, pick(ctx, stream));
}

int fhe_compare_dev(fhe_ctx *ctx, int level, int op, const void *a_dev, const void *b_dev,
                    int b_broadcast, uint32_t *out_dev, uint32_t plain_bits, int result_type,
                    int B, void *stream)
{
    FHE_CHECK_ARG(ctx && a_dev && b_dev && out_dev && B >= 0, "fhe_compare_dev: bad argument");
    FHE_CHECK_ARG(level == 1 || level == 2, "fhe_compare_dev: level 1 or 2");
    FHE_CHECK_ARG(plain_bits >= 1 && plain_bits + 1 <= (level == 1 ? 10u : 33u),
                  "fhe_compare_dev: plain_bits + 1 must be at most %d at level %d",
                  level == 1 ? 10 : 33, level);
    if (level == 2 && plain_bits + 1 > kMaxMsbBits2) {
        set_error("fhe_compare_dev: plain_bits %u above %d at level 2: the peeling steps shift "
                  "the lvl2 bootstrap noise onto the decision margin (wrong results from 27 bits on)",
                  plain_bits, kMaxMsbBits2 - 1);
        return FHE_ERR_UNSUPPORTED;
    }
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    const bool bc = b_broadcast != 0;
    const int rt = result_type ? FHE_LOGIC : FHE_ARITHMETIC;
    switch (op) {
    case FHE_CMP_GT:  // a > b  <=>  msb(b - a)
        return cmp_strict(ctx, level, a_dev, b_dev, bc, -1, 1, out_dev, plain_bits, rt, B, s);
    case FHE_CMP_LT:  // a < b  <=>  msb(a - b)
        return cmp_strict(ctx, level, a_dev, b_dev, bc, 1, -1, out_dev, plain_bits, rt, B, s);
    case FHE_CMP_GE:  // a >= b <=>  NOT msb(a - b)
        return cmp_nonstrict(ctx, level, a_dev, b_dev, bc, 1, -1, out_dev, plain_bits, rt, B, s);
    case FHE_CMP_LE:  // a <= b <=>  NOT msb(b - a)
        return cmp_nonstrict(ctx, level, a_dev, b_dev, bc, -1, 1, out_dev, plain_bits, rt, B, s);
    case FHE_CMP_EQ: {  // (a >= b) AND (a <= b), HomCompare_gpu.cuh:64-72
        uint32_t *ge;
        FHE_TRY(chain_buf(ctx, kE, size_t(B) * W1, &ge));
        FHE_TRY(cmp_nonstrict(ctx, level, a_dev, b_dev, bc, 1, -1, ge, plain_bits, FHE_LOGIC, B, s));
        FHE_TRY(cmp_nonstrict(ctx, level, a_dev, b_dev, bc, -1, 1, out_dev, plain_bits, FHE_LOGIC,
                              B, s));
        return fhe_gate_lvl1_dev(ctx, FHE_GATE_AND, ge, out_dev, out_dev, rt, B, s);
    }
    default:
        break;
    }
    set_error("fhe_compare_dev: unknown comparison %d", op);
    return FHE_ERR_ARG;
}

int fhe_hom_not_dev(fhe_ctx *ctx, int level, const void *in_dev, void *out_dev, int B,
                    void *stream)
{
    FHE_CHECK_ARG(ctx && in_dev && out_dev && B >= 0, "fhe_hom_not_dev: bad argument");
    FHE_CHECK_ARG(level == 0 || level == 1 || level == 2, "fhe_hom_not_dev: level 0, 1 or 2");
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    if (level == 2) {
        HIP_TRY(launch_tlwe_lincomb<uint64_t>(static_cast<const uint64_t *>(in_dev), nullptr, -1, 0,
                                              0ull, static_cast<uint64_t *>(out_dev), W2, false, B,
                                              s),
                "HomNOT");
    } else {
        const int w = level == 1 ? W1 : int(ctx->p.n0) + 1;
        HIP_TRY(launch_tlwe_lincomb<uint32_t>(static_cast<const uint32_t *>(in_dev), nullptr, -1, 0,
                                              0u, static_cast<uint32_t *>(out_dev), w, false, B, s),
                "HomNOT");
    }
    return FHE_OK;
}

}  // extern "C"
