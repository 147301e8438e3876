/*
 * tfhe_oracle.c -- CPU restatement of the reference TFHE bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See tfhe_oracle.h for the layout
 * conventions and for the reference function each entry point restates.
 *
 * The FFT here is a plain FP64 iterative radix-2 transform written for this
 * checker; it computes the same mathematical map as spqlios' ifft/fft
 * (TFHEpp/thirdparties/spqlios/spqlios-fft-impl.cpp:204-397 fft_model and
 * :468-640 ifft_model): twist by w^j, DFT of the N/2 complex points
 * z_j = a_j + i*a_{j+N/2}, output in spqlios' bit-reversed order.  Operation
 * order differs from the spqlios asm, so FD values agree to FP64 rounding and
 * torus outputs to a few units in the last place.
 */
#include "tfhe_oracle.h"

#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} cplx;

/* exp(2*pi*i*k/K) with octant reduction so that the table entries are as
 * accurate as libm's sin/cos on [0, pi/4]. */
static cplx cexp_frac(long k, long K)
{
    k %= K;
    if (k < 0) k += K;
    /* reduce to [0, K/8] using symmetries of the unit circle */
    int neg_im = 0, swap = 0, neg_re_after = 0;
    long kk = k;
    if (kk > K / 2) { kk = K - kk; neg_im = 1; }          /* conj */
    if (kk > K / 4) { kk = K / 2 - kk; neg_re_after = 1; } /* pi - x */
    if (kk > K / 8) { kk = K / 4 - kk; swap = 1; }         /* pi/2 - x */
    double ang = 2.0 * M_PI * (double)kk / (double)K;
    double c = cos(ang), s = sin(ang);
    if (swap) { double t = c; c = s; s = t; }
    if (neg_re_after) c = -c;
    if (neg_im) s = -s;
    cplx r = {c, s};
    return r;
}


/* All twiddles used here are 4096-th roots of unity (N <= 2048, so 2N | 4096):
 * exp(2 pi i k / K) = circ[(k mod K) * 4096 / K].  cexp_frac(k*s, K*s) returns
 * the same double as cexp_frac(k, K), so the table changes no value. */
#define ORC_ROOTS 4096
static cplx circ[ORC_ROOTS];
__attribute__((constructor)) static void orc_init_roots(void)
{
    for (int k = 0; k < ORC_ROOTS; k++) circ[k] = cexp_frac(k, ORC_ROOTS);
}
static inline cplx root(long k, long K)
{
    k %= K;
    if (k < 0) k += K;
    return circ[k * (ORC_ROOTS / K)];
}

int orc_fd_rev(int idx, int M)
{
    int r = 0;
    for (int m = M; m > 1; m >>= 1) {
        r = (r << 1) | (idx & 1);
        idx >>= 1;
    }
    return r;
}

/* In-place DFT of size M (power of two), sign = +1 (forward, exp(+2 pi i nm/M))
 * or -1; natural order in and out, unnormalised. */
static void dft_pow2(cplx *x, int M, int sign)
{
    /* bit-reversal permutation */
    for (int i = 0; i < M; i++) {
        int j = orc_fd_rev(i, M);
        if (j > i) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= M; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            cplx w = root((long)sign * k, len);
            for (int s = 0; s < M; s += len) {
                cplx u = x[s + k], v = x[s + k + half];
                cplx t = {v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re};
                x[s + k].re = u.re + t.re;
                x[s + k].im = u.im + t.im;
                x[s + k + half].re = u.re - t.re;
                x[s + k + half].im = u.im - t.im;
            }
        }
    }
}

/* a (N reals) -> spqlios-ordered FD (N doubles). */
static void twist_ifft_real(double *fd, const double *a, int N)
{
    int M = N / 2;
    cplx *z = (cplx *)malloc(sizeof(cplx) * M);
    for (int j = 0; j < M; j++) {
        cplx w = root(j, 2L * N); /* w^j, w = exp(i pi / N) */
        double re = a[j], im = a[j + M];
        z[j].re = re * w.re - im * w.im;
        z[j].im = re * w.im + im * w.re;
    }
    dft_pow2(z, M, +1);
    for (int idx = 0; idx < M; idx++) {
        int m = orc_fd_rev(idx, M);
        fd[idx] = z[m].re;
        fd[idx + M] = z[m].im;
    }
    free(z);
}

/* spqlios-ordered FD -> N reals (scaled by 2/N, untwisted). */
static void twist_fft_real(double *a, const double *fd, int N)
{
    int M = N / 2;
    const double s2n = 2.0 / (double)N;
    cplx *z = (cplx *)malloc(sizeof(cplx) * M);
    for (int m = 0; m < M; m++) {
        int idx = orc_fd_rev(m, M);
        z[m].re = fd[idx] * s2n;
        z[m].im = fd[idx + M] * s2n;
    }
    dft_pow2(z, M, -1);
    for (int j = 0; j < M; j++) {
        cplx w = root(-(long)j, 2L * N);
        a[j] = z[j].re * w.re - z[j].im * w.im;
        a[j + M] = z[j].re * w.im + z[j].im * w.re;
    }
    free(z);
}

void orc_ifft_torus32(double *fd, const uint32_t *a, int N)
{
    double *x = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) x[i] = (double)(int32_t)a[i];
    twist_ifft_real(fd, x, N);
    free(x);
}

void orc_fft_torus32(uint32_t *res, const double *fd, int N)
{
    double *x = (double *)malloc(sizeof(double) * N);
    twist_fft_real(x, fd, N);
    for (int i = 0; i < N; i++) res[i] = (uint32_t)(int64_t)x[i];
    free(x);
}

void orc_ifft_torus64(double *fd, const uint64_t *a, int N)
{
    double *x = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) x[i] = (double)(int64_t)a[i];
    twist_ifft_real(fd, x, N);
    free(x);
}

/* Truncation toward zero of the exact value of x, reduced mod 2^64.
 * Matches execute_direct_torus64's mantissa shift for every |x| >= 1 below
 * 2^116; for |x| < 1 the result is 0. */
static uint64_t d2t64(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int expo = (int)((bits >> 52) & 0x7FF);
    if (expo < 1023) return 0; /* |x| < 1 */
    uint64_t mant = (bits & 0x000FFFFFFFFFFFFFull) | 0x0010000000000000ull;
    int sh = expo - 1075;
    uint64_t v;
    if (sh >= 64)
        v = 0;
    else if (sh >= 0)
        v = mant << sh;
    else
        v = mant >> (-sh);
    return (bits >> 63) ? (uint64_t)0 - v : v;
}

void orc_fft_torus64(uint64_t *res, const double *fd, int N)
{
    double *x = (double *)malloc(sizeof(double) * N);
    twist_fft_real(x, fd, N);
    for (int i = 0; i < N; i++) res[i] = d2t64(x[i]);
    free(x);
}

/* ---- polynomial helpers -------------------------------------------------- */

#define DEF_ROT(T, SUF)                                                              \
    void orc_mul_xai_##SUF(T *res, const T *poly, uint32_t a, int N)                 \
    {                                                                                \
        if (a == 0) {                                                                \
            memcpy(res, poly, sizeof(T) * N);                                        \
        } else if (a < (uint32_t)N) {                                                \
            for (int i = 0; i < (int)a; i++) res[i] = (T)0 - poly[i - a + N];        \
            for (int i = a; i < N; i++) res[i] = poly[i - a];                        \
        } else {                                                                     \
            uint32_t aa = a - N;                                                     \
            for (int i = 0; i < (int)aa; i++) res[i] = poly[i - aa + N];             \
            for (int i = aa; i < N; i++) res[i] = (T)0 - poly[i - aa];               \
        }                                                                            \
    }                                                                                \
    void orc_mul_xai_minus_one_##SUF(T *res, const T *poly, uint32_t a, int N)       \
    {                                                                                \
        if (a < (uint32_t)N) {                                                       \
            for (int i = 0; i < (int)a; i++) res[i] = (T)0 - poly[i - a + N] - poly[i]; \
            for (int i = a; i < N; i++) res[i] = poly[i - a] - poly[i];              \
        } else {                                                                     \
            uint32_t aa = a - N;                                                     \
            for (int i = 0; i < (int)aa; i++) res[i] = poly[i - aa + N] - poly[i];   \
            for (int i = aa; i < N; i++) res[i] = (T)0 - poly[i - aa] - poly[i];     \
        }                                                                            \
    }
DEF_ROT(uint32_t, u32)
DEF_ROT(uint64_t, u64)

/* offsetgen (TFHEpp/src/trgsw.cpp:9-18): sum_{i=1..l} Bg/2 * 2^(w - i*Bgbit) */
#define DEF_DECOMP(T, SUF, W)                                                        \
    void orc_decompose_##SUF(T *dec, const T *poly, int digit, int N, int l, int Bgbit) \
    {                                                                                \
        T offset = 0;                                                                \
        for (int i = 1; i <= l; i++)                                                 \
            offset += ((T)1 << (Bgbit - 1)) * ((T)1 << (W - i * Bgbit));             \
        T roundoffset = (T)1 << (W - l * Bgbit - 1);                                 \
        T mask = ((T)1 << Bgbit) - 1;                                                \
        T halfBg = (T)1 << (Bgbit - 1);                                              \
        int shift = W - (digit + 1) * Bgbit;                                         \
        for (int i = 0; i < N; i++)                                                  \
            dec[i] = (((poly[i] + offset + roundoffset) >> shift) & mask) - halfBg;  \
    }
DEF_DECOMP(uint32_t, u32, 32)
DEF_DECOMP(uint64_t, u64, 64)

#define DEF_NAIVE(T, ST, SUF)                                                        \
    void orc_negacyclic_naive_##SUF(T *res, const T *a, const T *b, int N)           \
    {                                                                                \
        for (int i = 0; i < N; i++) {                                                \
            T r = 0;                                                                 \
            for (int j = 0; j <= i; j++) r += (T)((ST)a[j]) * b[i - j];              \
            for (int j = i + 1; j < N; j++) r -= (T)((ST)a[j]) * b[N + i - j];       \
            res[i] = r;                                                              \
        }                                                                            \
    }
DEF_NAIVE(uint32_t, int32_t, u32)
DEF_NAIVE(uint64_t, int64_t, u64)

/* ---- external product ---------------------------------------------------- */

/* MulInFD / FMAInFD (TFHEpp/include/utils.hpp:83-108), spqlios FD layout. */
static void mul_in_fd(double *res, const double *a, const double *b, int N)
{
    int M = N / 2;
    for (int i = 0; i < M; i++) {
        double aimbim = a[i + M] * b[i + M];
        double arebim = a[i] * b[i + M];
        res[i] = fma(a[i], b[i], -aimbim);
        res[i + M] = fma(a[i + M], b[i], arebim);
    }
}
static void fma_in_fd(double *res, const double *a, const double *b, int N)
{
    int M = N / 2;
    for (int i = 0; i < M; i++) {
        res[i] = fma(a[i], b[i], res[i]);
        res[i + M] = fma(a[i + M], b[i], res[i + M]);
    }
    for (int i = 0; i < M; i++) {
        res[i + M] = fma(a[i], b[i + M], res[i + M]);
        res[i] -= a[i + M] * b[i + M];
    }
}

void orc_trgsw_to_fft_u32(double *trgswfft, const uint32_t *trgsw, int N, int l)
{
    for (int r = 0; r < 2 * l; r++)
        for (int c = 0; c < 2; c++)
            orc_ifft_torus32(trgswfft + ((size_t)r * 2 + c) * N, trgsw + ((size_t)r * 2 + c) * N, N);
}

void orc_trgsw_to_fft_u64(double *trgswfft, const uint64_t *trgsw, int N, int l)
{
    for (int r = 0; r < 2 * l; r++)
        for (int c = 0; c < 2; c++)
            orc_ifft_torus64(trgswfft + ((size_t)r * 2 + c) * N, trgsw + ((size_t)r * 2 + c) * N, N);
}

#define DEF_EXTPROD(T, SUF)                                                          \
    void orc_external_product_##SUF(T *res, const T *trlwe, const double *trgswfft,  \
                                    int N, int l, int Bgbit)                         \
    {                                                                                \
        T *dec = (T *)malloc(sizeof(T) * N);                                         \
        double *decfd = (double *)malloc(sizeof(double) * N);                        \
        double *acc = (double *)malloc(sizeof(double) * 2 * N);                      \
        int first = 1;                                                               \
        for (int k = 0; k < 2; k++) {                                                \
            for (int i = 0; i < l; i++) {                                            \
                orc_decompose_##SUF(dec, trlwe + (size_t)k * N, i, N, l, Bgbit);     \
                orc_ifft_torus##SUF##_dec(decfd, dec, N);                            \
                const double *row = trgswfft + ((size_t)(i + k * l) * 2) * N;        \
                for (int m = 0; m < 2; m++) {                                        \
                    if (first)                                                       \
                        mul_in_fd(acc + (size_t)m * N, decfd, row + (size_t)m * N, N); \
                    else                                                             \
                        fma_in_fd(acc + (size_t)m * N, decfd, row + (size_t)m * N, N); \
                }                                                                    \
                first = 0;                                                           \
            }                                                                        \
        }                                                                            \
        for (int m = 0; m < 2; m++) orc_fft_torus##SUF##_out(res + (size_t)m * N, acc + (size_t)m * N, N); \
        free(dec);                                                                   \
        free(decfd);                                                                 \
        free(acc);                                                                   \
    }
#define orc_ifft_torusu32_dec orc_ifft_torus32
#define orc_ifft_torusu64_dec orc_ifft_torus64
#define orc_fft_torusu32_out orc_fft_torus32
#define orc_fft_torusu64_out orc_fft_torus64
DEF_EXTPROD(uint32_t, u32)
DEF_EXTPROD(uint64_t, u64)

#define DEF_CMUX(T, SUF)                                                             \
    void orc_cmux_rot_##SUF(T *acc, const double *trgswfft, uint32_t a, int N, int l, int Bgbit) \
    {                                                                                \
        T *tmp = (T *)malloc(sizeof(T) * 2 * N);                                     \
        for (int k = 0; k < 2; k++)                                                  \
            orc_mul_xai_minus_one_##SUF(tmp + (size_t)k * N, acc + (size_t)k * N, a, N); \
        orc_external_product_##SUF(tmp, tmp, trgswfft, N, l, Bgbit);                 \
        for (int i = 0; i < 2 * N; i++) acc[i] += tmp[i];                            \
        free(tmp);                                                                   \
    }
DEF_CMUX(uint32_t, u32)
DEF_CMUX(uint64_t, u64)

/* ---- blind rotation ------------------------------------------------------ */
/* The domain (lvl0) torus is 32-bit; nbit = log2 N of the target ring.
 *   bbar = 2N - (b >> (32 - 1 - nbit))                 (no rounding)
 *   abar = (a + 2^(32 - 2 - nbit)) >> (32 - 1 - nbit)  (rounded)
 * (TFHEpp/include/gatebootstrapping.hpp:21-41 with num_out = 1). */
#define DEF_BR(T, SUF)                                                               \
    void orc_blind_rotate_##SUF(T *acc, const uint32_t *tlwe0, int n0, const double *bkfft, \
                                const T *tv, int N, int nbit, int l, int Bgbit)      \
    {                                                                                \
        uint32_t bbar = 2u * N - (tlwe0[n0] >> (32 - 1 - nbit));                     \
        memset(acc, 0, sizeof(T) * N);                                               \
        orc_mul_xai_##SUF(acc + N, tv, bbar, N);                                     \
        const uint32_t roundoffset = 1u << (32 - 2 - nbit);                          \
        const size_t trgswlen = (size_t)2 * l * 2 * N;                               \
        for (int i = 0; i < n0; i++) {                                               \
            uint32_t abar = (tlwe0[i] + roundoffset) >> (32 - 1 - nbit);             \
            if (abar == 0) continue;                                                 \
            orc_cmux_rot_##SUF(acc, bkfft + (size_t)i * trgswlen, abar, N, l, Bgbit); \
        }                                                                            \
    }
DEF_BR(uint32_t, u32)
DEF_BR(uint64_t, u64)

#define DEF_SE(T, SUF)                                                               \
    void orc_sample_extract_##SUF(T *tlwe, const T *trlwe, int index, int N)         \
    {                                                                                \
        for (int i = 0; i <= index; i++) tlwe[i] = trlwe[index - i];                 \
        for (int i = index + 1; i < N; i++) tlwe[i] = (T)0 - trlwe[N + index - i];   \
        tlwe[N] = trlwe[N + index];                                                  \
    }
DEF_SE(uint32_t, u32)
DEF_SE(uint64_t, u64)

#define DEF_PBS(T, SUF)                                                              \
    void orc_pbs_##SUF(T *out, const uint32_t *tlwe0, int n0, const double *bkfft,   \
                       const T *tv, int N, int nbit, int l, int Bgbit)               \
    {                                                                                \
        T *acc = (T *)malloc(sizeof(T) * 2 * N);                                     \
        orc_blind_rotate_##SUF(acc, tlwe0, n0, bkfft, tv, N, nbit, l, Bgbit);        \
        orc_sample_extract_##SUF(out, acc, 0, N);                                    \
        free(acc);                                                                   \
    }
DEF_PBS(uint32_t, u32)
DEF_PBS(uint64_t, u64)

/* GateBootstrappingManyLUT<P, num_out> (TFHEpp/include/gatebootstrapping.hpp:119-130
 * with BlindRotate<P, num_out>, :15-43): bw = bits_needed(num_out - 1)
 * (TFHEpp/include/utils.hpp:69-81, the bit length of num_out - 1), then
 *   bbar = 2N - ((b >> (32 - 1 - nbit + bw)) << bw)
 *   abar = ((a + 2^(32 - 2 - nbit + bw)) >> (32 - 1 - nbit + bw)) << bw
 * and SampleExtractIndex at 0 .. num_out-1 (out: num_out x (N+1) words). */
static int bits_needed_u32(uint32_t v)
{
    int b = 0;
    while (v >> b) b++;
    return b;
}

#define DEF_PBS_MANY(T, SUF)                                                         \
    void orc_pbs_many_##SUF(T *out, const uint32_t *tlwe0, int n0, const double *bkfft, \
                            const T *tv, int N, int nbit, int l, int Bgbit, int num_out) \
    {                                                                                \
        const int bw = bits_needed_u32((uint32_t)num_out - 1);                       \
        const int sh = 32 - 1 - nbit + bw;                                           \
        T *acc = (T *)malloc(sizeof(T) * 2 * N);                                     \
        uint32_t bbar = 2u * N - ((tlwe0[n0] >> sh) << bw);                          \
        memset(acc, 0, sizeof(T) * N);                                               \
        orc_mul_xai_##SUF(acc + N, tv, bbar, N);                                     \
        const uint32_t roundoffset = 1u << (sh - 1);                                 \
        const size_t trgswlen = (size_t)2 * l * 2 * N;                               \
        for (int i = 0; i < n0; i++) {                                               \
            uint32_t abar = ((tlwe0[i] + roundoffset) >> sh) << bw;                  \
            if (abar == 0) continue;                                                 \
            orc_cmux_rot_##SUF(acc, bkfft + (size_t)i * trgswlen, abar, N, l, Bgbit); \
        }                                                                            \
        for (int k = 0; k < num_out; k++)                                            \
            orc_sample_extract_##SUF(out + (size_t)k * (N + 1), acc, k, N);          \
        free(acc);                                                                   \
    }
DEF_PBS_MANY(uint32_t, u32)
DEF_PBS_MANY(uint64_t, u64)

/* ---- identity key switch ------------------------------------------------- */
/* out (target, 32-bit torus, tn+1 words) from in (domain, W-bit torus, dn+1). */
#define DEF_IKS(TD, SUF, W)                                                          \
    void orc_iks_##SUF(uint32_t *out, const TD *in, const uint32_t *ksk, int dn, int tn, \
                       int t, int basebit)                                           \
    {                                                                                \
        const TD prec_offset = (TD)1 << (W - (1 + basebit * t));                     \
        const uint32_t mask = (1u << basebit) - 1;                                   \
        const size_t nd = ((size_t)1 << basebit) - 1;                                \
        const size_t row = (size_t)tn + 1;                                           \
        memset(out, 0, sizeof(uint32_t) * row);                                      \
        if (W == 32)                                                                 \
            out[tn] = (uint32_t)in[dn];                                              \
        else                                                                         \
            out[tn] = (uint32_t)((in[dn] + ((TD)1 << (W > 32 ? W - 33 : 0))) >> (W > 32 ? W - 32 : 0));    \
        for (int i = 0; i < dn; i++) {                                               \
            const TD aibar = in[i] + prec_offset;                                    \
            for (int j = 0; j < t; j++) {                                            \
                const uint32_t aij = (uint32_t)((aibar >> (W - (j + 1) * basebit)) & mask); \
                if (aij == 0) continue;                                              \
                const uint32_t *r = ksk + (((size_t)i * t + j) * nd + (aij - 1)) * row; \
                for (size_t k = 0; k < row; k++) out[k] -= r[k];                     \
            }                                                                        \
        }                                                                            \
    }
DEF_IKS(uint32_t, 32to32, 32)
DEF_IKS(uint64_t, 64to32, 64)

/* ---- readout -------------------------------------------------------------- */
uint32_t orc_phase_u32(const uint32_t *c, const uint32_t *key, int n)
{
    uint32_t p = c[n];
    for (int i = 0; i < n; i++) p -= c[i] * key[i];
    return p;
}
uint64_t orc_phase_u64(const uint64_t *c, const uint64_t *key, int n)
{
    uint64_t p = c[n];
    for (int i = 0; i < n; i++) p -= c[i] * key[i];
    return p;
}
void orc_trlwe_phase_u32(uint32_t *phase, const uint32_t *c, const uint32_t *key, int N)
{
    uint32_t *mul = (uint32_t *)malloc(sizeof(uint32_t) * N);
    /* key coefficients are small (binary): use them as the signed operand */
    orc_negacyclic_naive_u32(mul, key, c, N);
    for (int i = 0; i < N; i++) phase[i] = c[N + i] - mul[i];
    free(mul);
}
void orc_trlwe_phase_u64(uint64_t *phase, const uint64_t *c, const uint64_t *key, int N)
{
    uint64_t *mul = (uint64_t *)malloc(sizeof(uint64_t) * N);
    orc_negacyclic_naive_u64(mul, key, c, N);
    for (int i = 0; i < N; i++) phase[i] = c[N + i] - mul[i];
    free(mul);
}

/* ---- HomNAND -------------------------------------------------------------- */
void orc_hom_nand_lvl1(uint32_t *out, const uint32_t *ca, const uint32_t *cb, const uint32_t *ksk10,
                       int n0, const double *bkfft)
{
    const int N = 1024;
    const uint32_t mu = 1u << 29;
    uint32_t *c = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1));
    uint32_t *c0 = (uint32_t *)malloc(sizeof(uint32_t) * (n0 + 1));
    uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * N);
    for (int i = 0; i <= N; i++) c[i] = (uint32_t)0 - ca[i] - cb[i];
    c[N] += mu;
    orc_iks_32to32(c0, c, ksk10, N, n0, 2, 10);
    for (int i = 0; i < N; i++) tv[i] = mu;
    orc_pbs_u32(out, c0, n0, bkfft, tv, N, 10, 3, 6);
    free(c);
    free(c0);
    free(tv);
}
