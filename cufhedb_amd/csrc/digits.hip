// digits.hip -- digit-wise comparison of radix-encrypted integers with a
// public threshold: one lvl2 bootstrap per digit plus one combining
// bootstrap.
//
// The reference compares whole-value ciphertexts through HomMSB chains
// (src/cuHEDB/HomCompare_gpu.cuh:14-76 over msb_gpu.cuh:175-192: for a 20-bit
// value three 5-bit lvl2 peeling steps, 6 lvl2 + 3 lvl1 bootstraps per row,
// run one after another).  This operator takes the value as `digits`
// separate lvl2 TLWEs of `digit_bits` bits each (digit i holds
// x_i = (x >> (i * digit_bits)) mod 2^digit_bits at v * 2^(63 - digit_bits),
// one padding bit) and the threshold t in the clear:
//
//   1. every digit: IKS lvl20, then a lvl02 bootstrap whose test vector
//      (built from t_i) returns a small signed integer times 2^-digits:
//        ordering (GT, GE, LT, LE):
//          digit 0:  d_0 = [x_0 > t_0]  (GT, LE)  or  [x_0 >= t_0]  (GE, LT)
//          digit i:  d_i = sign(x_i - t_i) * 2^(i-1)            (i >= 1)
//        equality:   d_i = [x_i != t_i]
//      all digits of a batch in one launch (test vector = ciphertext mod
//      digits);
//   2. S = sum_i d_i - 1/2, on the lvl2 ciphertexts (no bootstrap);
//   3. IKS lvl20 and one lvl02 sign bootstrap: S > 0 is x > t (GT, LE),
//      x >= t (GE, LT) or x != t (EQ), and the test vector's sign turns it
//      into the requested result; IKS lvl21 for a lvl1 output.
//
// Why one sign bootstrap decides the whole comparison: with weights
// w_0 = 1 (on a 0/1 digit) and w_i = 2^(i-1) = 1 + w_1 + .. + w_(i-1), the
// most significant differing digit k outweighs everything below it.  If
// x_k > t_k then S + 1/2 >= w_k - (w_1 + .. + w_(k-1)) = 1, if x_k < t_k then
// S + 1/2 <= -w_k + (w_1 + .. + w_(k-1)) + 1 = 0, and if all digits above 0
// agree, S + 1/2 = d_0.  S + 1/2 lies in [-(2^(digits-1) - 1), 2^(digits-1)],
// 2^digits integers, so at scale 2^-digits every phase S * 2^-digits keeps a
// distance of 2^-(digits+1) (2^-6 for five digits) from both decision points
// 0 and 1/2 -- the same margin as the digit bootstraps' 4-bit boxes.
#include "ctx.hpp"
#include "launch.hpp"

using namespace fhe;

namespace {

constexpr int N2 = Lvl2::N;
constexpr int W2 = Lvl2::N + 1;
constexpr int kMaxDigits = 5;
constexpr int kMaxDigitBits = 4;

enum { kLvl0 = 0, kBoot = 1, kSum = 2, kTv = 3 };

__host__ __device__ inline bool ordering_strict(int op)
{
    return op == FHE_CMP_GT || op == FHE_CMP_LE;
}

// Test-vector word j of digit position i.  A digit v sits at phase
// v 2^-(bits+1); after the switch to Z_4096 it is p = v 2^(11-bits) + error,
// so coefficient j holds f(v) for v = round(j / 2^(11-bits)); the top
// half-box (v = 2^bits) is digit 0 seen from below zero, where the
// negacyclic rotation negates the coefficient, so it holds -f(0).
__host__ __device__ inline uint64_t digit_tv_word(int op, int digits, int bits, uint64_t threshold,
                                                  int i, int j)
{
    const int R = 1 << bits;
    const int ti = int((threshold >> (i * bits)) & uint64_t(R - 1));
    const int box = 11 - bits;
    int v = (j + (1 << (box - 1))) >> box;
    const bool wrap = v == R;
    if (wrap) v = 0;
    int64_t f;
    if (op == FHE_CMP_EQ) {
        f = v != ti;
    } else if (i == 0) {
        f = ordering_strict(op) ? v > ti : v >= ti;
    } else {
        f = int64_t((v > ti) - (v < ti)) << (i - 1);
    }
    const uint64_t w = uint64_t(f) << (64 - digits);
    return wrap ? 0 - w : w;
}

// sign bootstrap constant and the body offset added after it
void final_words(int op, int result_type, uint64_t *tv_word, uint64_t *post_add)
{
    const uint64_t u = result_type ? Lvl2::mu : Lvl2::mu << 1;
    const bool true_above = op == FHE_CMP_GT || op == FHE_CMP_GE;
    *tv_word = true_above ? u : 0 - u;
    *post_add = result_type ? 0 : u;
}

uint64_t combine_offset(int digits) { return 0 - (uint64_t(1) << (63 - digits)); }

int check_digit_args(const char *fn, int op, int digits, int digit_bits, uint64_t threshold,
                     int result_type)
{
    FHE_CHECK_ARG(op >= FHE_CMP_GT && op <= FHE_CMP_EQ, "%s: unknown comparison %d", fn, op);
    FHE_CHECK_ARG(digits >= 1 && digits <= kMaxDigits, "%s: digits must be in [1, %d] (got %d)", fn,
                  kMaxDigits, digits);
    FHE_CHECK_ARG(digit_bits >= 1 && digit_bits <= kMaxDigitBits,
                  "%s: digit_bits must be in [1, %d] (got %d)", fn, kMaxDigitBits, digit_bits);
    FHE_CHECK_ARG(threshold >> (digits * digit_bits) == 0,
                  "%s: threshold %llu does not fit %d digits of %d bits", fn,
                  (unsigned long long)threshold, digits, digit_bits);
    FHE_CHECK_ARG(result_type == FHE_LOGIC || result_type == FHE_ARITHMETIC,
                  "%s: result_type must be LOGIC or ARITHMETIC", fn);
    return FHE_OK;
}

__global__ __launch_bounds__(256) void k_digit_tvs(uint64_t *__restrict__ tv, int op, int digits,
                                                   int bits, uint64_t threshold,
                                                   uint64_t final_word)
{
    const int i = blockIdx.x;
    for (int j = threadIdx.x; j < N2; j += 256)
        tv[size_t(i) * N2 + j] =
            i < digits ? digit_tv_word(op, digits, bits, threshold, i, j) : final_word;
}

// out[r] = sum_i in[r * digits + i] (+ body_add on the body word)
__global__ __launch_bounds__(256) void k_digit_sum(const uint64_t *__restrict__ in,
                                                   uint64_t *__restrict__ out, int digits,
                                                   uint64_t body_add, size_t total)
{
    const size_t step = size_t(gridDim.x) * 256;
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += step) {
        const size_t r = e / W2, k = e - r * W2;
        const uint64_t *p = in + r * size_t(digits) * W2 + k;
        uint64_t acc = k == size_t(N2) ? body_add : 0;
        for (int i = 0; i < digits; i++) acc += p[size_t(i) * W2];
        out[e] = acc;
    }
}

int digit_buf(fhe_ctx *ctx, int slot, size_t words, uint32_t **p)
{
    int rc = grow(ctx->digit[slot], ctx->digit_cap[slot], words);
    *p = ctx->digit[slot];
    return rc;
}

#define DIG_TRY(expr)              \
    do {                           \
        int _rc = (expr);          \
        if (_rc) return _rc;       \
    } while (0)

}  // namespace

extern "C" {

int fhe_digit_testvectors(int op, int digits, int digit_bits, uint64_t threshold, int result_type,
                          uint64_t *tv, uint64_t *combine_add, uint64_t *post_add)
{
    FHE_CHECK_ARG(tv && combine_add && post_add, "fhe_digit_testvectors: null argument");
    DIG_TRY(check_digit_args("fhe_digit_testvectors", op, digits, digit_bits, threshold,
                             result_type));
    uint64_t fw;
    final_words(op, result_type, &fw, post_add);
    for (int i = 0; i <= digits; i++)
        for (int j = 0; j < N2; j++)
            tv[size_t(i) * N2 + j] =
                i < digits ? digit_tv_word(op, digits, digit_bits, threshold, i, j) : fw;
    *combine_add = combine_offset(digits);
    return FHE_OK;
}

int fhe_compare_digits_dev(fhe_ctx *ctx, int op, const uint64_t *x_dev, int digits, int digit_bits,
                           uint64_t threshold, int out_level, void *out_dev, int result_type, int B,
                           void *stream)
{
    FHE_CHECK_ARG(ctx && x_dev && out_dev && B >= 0, "fhe_compare_digits_dev: bad argument");
    DIG_TRY(check_digit_args("fhe_compare_digits_dev", op, digits, digit_bits, threshold,
                             result_type));
    FHE_CHECK_ARG(out_level == 1 || out_level == 2, "fhe_compare_digits_dev: out_level 1 or 2");
    FHE_CHECK_ARG(B <= (1 << 30) / digits, "fhe_compare_digits_dev: batch of %d rows too large", B);
    if (!ctx->bk2 || !ctx->ksk[ksk_slot(20)] || (out_level == 1 && !ctx->ksk[ksk_slot(21)])) {
        set_error("fhe_compare_digits_dev: needs BK lvl02, KSK 20%s",
                  out_level == 1 ? " and 21" : "");
        return FHE_ERR_NOT_LOADED;
    }
    if (B == 0) return FHE_OK;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = pick(ctx, stream);
    const int n0 = int(ctx->p.n0);
    const size_t rows = size_t(B) * digits;

    uint32_t *lvl0, *boot32, *sum32, *tv32;
    DIG_TRY(digit_buf(ctx, kLvl0, rows * (n0 + 1), &lvl0));
    DIG_TRY(digit_buf(ctx, kBoot, rows * W2 * 2, &boot32));
    DIG_TRY(digit_buf(ctx, kSum, size_t(B) * W2 * 2, &sum32));
    DIG_TRY(digit_buf(ctx, kTv, size_t(digits + 1) * N2 * 2, &tv32));
    uint64_t *boot = reinterpret_cast<uint64_t *>(boot32);
    uint64_t *sum = reinterpret_cast<uint64_t *>(sum32);
    uint64_t *tv = reinterpret_cast<uint64_t *>(tv32);

    uint64_t fw, post;
    final_words(op, result_type, &fw, &post);
    hipLaunchKernelGGL(k_digit_tvs, dim3(digits + 1), dim3(256), 0, s, tv, op, digits, digit_bits,
                       threshold, fw);
    HIP_TRY(hipGetLastError(), "digit test vectors");

    // 1. every digit through IKS lvl20 and its own bootstrap
    const uint32_t *ksk20 = ctx->ksk[ksk_slot(20)];
    HIP_TRY(launch_iks(20, x_dev, ksk20, lvl0, n0, int(rows), s, iks_tile(ctx, 20, int(rows))),
            "digit key switch");
    HIP_TRY(launch_blind_rotate2q(lvl0, n0, ctx->bk2, ctx->rot2, tv, N2, boot, true, int(rows), s, digits),
            "digit bootstrap");
    // 2. the weighted digits summed, minus half a step
    const size_t total = size_t(B) * W2;
    const unsigned grid = unsigned(std::min<size_t>((total + 255) / 256, 65536));
    hipLaunchKernelGGL(k_digit_sum, dim3(grid), dim3(256), 0, s, boot, sum, digits,
                       combine_offset(digits), total);
    HIP_TRY(hipGetLastError(), "digit combination");
    // 3. one sign bootstrap
    HIP_TRY(launch_iks(20, sum, ksk20, lvl0, n0, B, s, iks_tile(ctx, 20, B)),
            "combine key switch");
    uint64_t *res = out_level == 2 ? static_cast<uint64_t *>(out_dev) : boot;
    HIP_TRY(launch_blind_rotate2q(lvl0, n0, ctx->bk2, ctx->rot2, tv + size_t(digits) * N2, 0, res, true, B, s),
            "combine bootstrap");
    if (post) HIP_TRY(launch_tlwe_add_body<uint64_t>(res, W2, post, B, s), "result offset");
    if (out_level == 1)
        HIP_TRY(launch_iks(21, res, ctx->ksk[ksk_slot(21)], static_cast<uint32_t *>(out_dev),
                           Lvl1::N, B, s, iks_tile(ctx, 21, B)),
                "result key switch");
    return FHE_OK;
}

}  // extern "C"
