// keygen.cpp -- seeded host key generation and encryption.
//
// Distributions follow the reference's TFHEpp:
//   secret keys      lweKey()            thirdparty/TFHEpp/src/key.cpp:8-17,
//                    generate_hamming_secret_key  include/key.hpp:38-49
//   TLWE encryption  tlweSymEncrypt      src/tlwe.cpp:19-31
//   TRLWE zero       trlweSymEncryptZero src/trlwe.cpp:8-24
//   TRGSW            trgswSymEncrypt     src/trgsw.cpp:206-224, hgen trgsw.hpp:33-41
//   bootstrapping    bkgen               src/cloudkey.cpp:7-21
//   key switching    ikskgen             src/cloudkey.cpp:54-70
// The reference draws from an unseedable randen/random_device generator
// (include/utils.hpp:17-22); here every record draws from its own Philox
// stream (rng.hpp) so keys are reproducible from a 64-bit seed and can be
// generated in parallel.  The TRLWE mask product a*s is exact (the reference
// uses its FFT PolyMul for lvl1, which can differ by one unit).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "common.hpp"
#include "params.hpp"
#include "rng.hpp"

namespace fhe {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int host_threads()
{
    if (const char *e = std::getenv("FHE_HOST_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) return v;
    }
    unsigned hc = std::thread::hardware_concurrency();
    return int(std::min(16u, std::max(1u, hc)));
}

void parallel_for(int64_t n, int threads, void (*fn)(int64_t, int64_t, void *), void *arg)
{
    if (n <= 0) return;
    threads = int(std::min<int64_t>(threads, n));
    if (threads <= 1) {
        fn(0, n, arg);
        return;
    }
    std::vector<std::thread> pool;
    int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back(fn, b, e, arg);
    }
    for (auto &th : pool) th.join();
}

// ---- torus-width generic helpers -----------------------------------------
template <class T> inline T draw_uniform(Stream &s);
template <> inline uint32_t draw_uniform<uint32_t>(Stream &s) { return s.u32(); }
template <> inline uint64_t draw_uniform<uint64_t>(Stream &s) { return s.u64(); }
template <class T> inline T draw_gauss(Stream &s, double a);
template <> inline uint32_t draw_gauss<uint32_t>(Stream &s, double a) { return gaussian_t32(s, a); }
template <> inline uint64_t draw_gauss<uint64_t>(Stream &s, double a) { return gaussian_t64(s, a); }

// b = m + e + <a, key>
template <class T>
static void tlwe_encrypt(Stream &s, const T *key, uint32_t n, T m, double alpha, T *out)
{
    T b = 0;
    for (uint32_t i = 0; i < n; i++) {
        out[i] = draw_uniform<T>(s);
        b += out[i] * key[i];
    }
    out[n] = b + m + draw_gauss<T>(s, alpha);
}

// exact negacyclic res += a * key for a binary key of degree N
template <class T>
static void add_mul_binary(T *res, const T *a, const T *key, int N)
{
    for (int i = 0; i < N; i++) {
        if (key[i] == 0) continue;
        // res += X^i * a
        for (int j = 0; j < N - i; j++) res[j + i] += a[j];
        for (int j = N - i; j < N; j++) res[j + i - N] -= a[j];
    }
}

// TRLWE encryption of zero under `key`: a uniform, b = a*s + e.
template <class T>
static void trlwe_encrypt_zero(Stream &s, const T *key, int N, double alpha, T *out)
{
    T *a = out, *b = out + N;
    for (int i = 0; i < N; i++) a[i] = draw_uniform<T>(s);
    for (int i = 0; i < N; i++) b[i] = draw_gauss<T>(s, alpha);
    add_mul_binary<T>(b, a, key, N);
}

// TRGSW of the constant polynomial m: rows r = i + k*l, row (i, k) gets
// m * 2^(W - (i+1) Bgbit) added to coefficient 0 of polynomial k.
template <class T, int N, int L, int B>
static void trgsw_encrypt_const(Stream &s, const T *key, T m, double alpha, T *out)
{
    constexpr int W = int(sizeof(T) * 8);
    for (int r = 0; r < 2 * L; r++) trlwe_encrypt_zero<T>(s, key, N, alpha, out + size_t(r) * 2 * N);
    for (int i = 0; i < L; i++)
        for (int k = 0; k < 2; k++) {
            T h = T(1) << (W - (i + 1) * B);
            out[(size_t(i + k * L) * 2 + k) * N + 0] += m * h;
        }
}

template <class T>
static void secret_key(uint64_t seed, uint32_t tag, uint32_t n, uint32_t hw, T *key)
{
    Stream s(seed, tag, 0);
    std::fill(key, key + n, T(0));
    if (hw == 0) {
        for (uint32_t i = 0; i < n; i++) key[i] = T(s.u32() & 1u);
    } else {
        for (uint32_t i = 0; i < hw; i++) key[s.below(n)] = T(1);
    }
}

}  // namespace fhe

using namespace fhe;

extern "C" {

const char *fhe_last_error(void) { return g_err.c_str(); }

int fhe_params_preset(int preset, fhe_params *out)
{
    FHE_CHECK_ARG(out != nullptr, "fhe_params_preset: null output");
    if (preset == FHE_PRESET_REF672) {
        // thirdparty/TFHEpp/include/params/128bit.hpp:8-10,26,44 and src/key.cpp:11-13
        *out = fhe_params{672, 512, 512, 1024, std::ldexp(1.0, -16), std::ldexp(1.0, -25),
                          std::ldexp(1.0, -52)};
        return FHE_OK;
    }
    if (preset == FHE_PRESET_CFG636) {
        *out = fhe_params{636, 0, 0, 0, std::ldexp(1.0, -15), std::ldexp(1.0, -25),
                          std::ldexp(1.0, -52)};
        return FHE_OK;
    }
    set_error("fhe_params_preset: unknown preset %d", preset);
    return FHE_ERR_ARG;
}

int fhe_keygen_secret(const fhe_params *p, uint64_t seed, uint32_t *s0, uint32_t *s1, uint64_t *s2)
{
    FHE_CHECK_ARG(p && p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_keygen_secret: bad n0");
    if (s0) secret_key<uint32_t>(seed, TAG_SK0, p->n0, p->key0_hw, s0);
    if (s1) secret_key<uint32_t>(seed, TAG_SK1, Lvl1::N, p->key1_hw, s1);
    if (s2) secret_key<uint64_t>(seed, TAG_SK2, Lvl2::N, p->key2_hw, s2);
    return FHE_OK;
}

}  // extern "C"

namespace {
struct BkJob {
    uint64_t seed;
    const uint32_t *s0;
    const void *skey;
    void *bk;
    double alpha;
};
template <class P, uint32_t TAG>
void bk_worker(int64_t b, int64_t e, void *arg)
{
    auto *j = static_cast<BkJob *>(arg);
    using T = typename P::T;
    const size_t len = size_t(P::rows) * 2 * P::N;
    for (int64_t i = b; i < e; i++) {
        Stream s(j->seed, TAG, uint64_t(i));
        trgsw_encrypt_const<T, P::N, P::l, P::Bgbit>(s, static_cast<const T *>(j->skey),
                                                    T(j->s0[i]), j->alpha,
                                                    static_cast<T *>(j->bk) + size_t(i) * len);
    }
}
}  // namespace

extern "C" {

int fhe_keygen_bk_lvl01(const fhe_params *p, uint64_t seed, const uint32_t *s0, const uint32_t *s1,
                        uint32_t *bk)
{
    FHE_CHECK_ARG(p && s0 && s1 && bk, "fhe_keygen_bk_lvl01: null argument");
    FHE_CHECK_ARG(p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_keygen_bk_lvl01: bad n0");
    BkJob j{seed, s0, s1, bk, p->alpha1};
    parallel_for(p->n0, host_threads(), bk_worker<Lvl1, TAG_BK01>, &j);
    return FHE_OK;
}

int fhe_keygen_bk_lvl02(const fhe_params *p, uint64_t seed, const uint32_t *s0, const uint64_t *s2,
                        uint64_t *bk)
{
    FHE_CHECK_ARG(p && s0 && s2 && bk, "fhe_keygen_bk_lvl02: null argument");
    FHE_CHECK_ARG(p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_keygen_bk_lvl02: bad n0");
    BkJob j{seed, s0, s2, bk, p->alpha2};
    parallel_for(p->n0, host_threads(), bk_worker<Lvl2, TAG_BK02>, &j);
    return FHE_OK;
}

}  // extern "C"

namespace {
struct KsGeom {
    int dn, tn, t, basebit;
    uint32_t tag;
    double alpha;
};
bool ks_geom(const fhe_params *p, int kind, KsGeom &g)
{
    switch (kind) {
    case 10: g = {Lvl1::N, int(p->n0), KS10::t, KS10::basebit, TAG_KSK10, p->alpha0}; return true;
    case 20: g = {Lvl2::N, int(p->n0), KS20::t, KS20::basebit, TAG_KSK20, p->alpha0}; return true;
    case 21: g = {Lvl2::N, Lvl1::N, KS21::t, KS21::basebit, TAG_KSK21, p->alpha1}; return true;
    default: return false;
    }
}
struct KsJob {
    KsGeom g;
    uint64_t seed;
    const void *sdom;
    bool dom64;
    const uint32_t *stgt;
    uint32_t *ksk;
};
void ks_worker(int64_t b, int64_t e, void *arg)
{
    auto *j = static_cast<KsJob *>(arg);
    const KsGeom &g = j->g;
    const size_t nd = (size_t(1) << g.basebit) - 1;
    const size_t row = size_t(g.tn) + 1;
    for (int64_t i = b; i < e; i++) {
        uint32_t si = j->dom64 ? uint32_t(static_cast<const uint64_t *>(j->sdom)[i])
                               : static_cast<const uint32_t *>(j->sdom)[i];
        for (int jj = 0; jj < g.t; jj++)
            for (size_t k = 0; k < nd; k++) {
                size_t rec = (size_t(i) * g.t + jj) * nd + k;
                Stream s(j->seed, g.tag, rec);
                uint32_t m = si * uint32_t(k + 1) * (1u << (32 - (jj + 1) * g.basebit));
                tlwe_encrypt<uint32_t>(s, j->stgt, uint32_t(g.tn), m, g.alpha, j->ksk + rec * row);
            }
    }
}
}  // namespace

extern "C" {

size_t fhe_ksk_words(const fhe_params *p, int kind)
{
    KsGeom g;
    if (!p || !ks_geom(p, kind, g)) return 0;
    return size_t(g.dn) * g.t * ((size_t(1) << g.basebit) - 1) * (size_t(g.tn) + 1);
}

int fhe_keygen_ksk(const fhe_params *p, int kind, uint64_t seed, const void *s_dom,
                   const uint32_t *s_tgt, uint32_t *ksk)
{
    FHE_CHECK_ARG(p && s_dom && s_tgt && ksk, "fhe_keygen_ksk: null argument");
    KsGeom g;
    FHE_CHECK_ARG(ks_geom(p, kind, g), "fhe_keygen_ksk: kind must be 10, 20 or 21 (got %d)", kind);
    KsJob j{g, seed, s_dom, kind != 10, s_tgt, ksk};
    parallel_for(g.dn, host_threads(), ks_worker, &j);
    return FHE_OK;
}

int fhe_tlwe_encrypt_u32(uint64_t seed, uint64_t index0, const uint32_t *key, uint32_t n,
                         const uint32_t *m, double alpha, uint32_t *out, int B)
{
    FHE_CHECK_ARG(key && m && out && B >= 0 && n >= 1, "fhe_tlwe_encrypt_u32: bad argument");
    for (int b = 0; b < B; b++) {
        Stream s(seed, TAG_TLWE, index0 + uint64_t(b));
        tlwe_encrypt<uint32_t>(s, key, n, m[b], alpha, out + size_t(b) * (n + 1));
    }
    return FHE_OK;
}

int fhe_tlwe_encrypt_u64(uint64_t seed, uint64_t index0, const uint64_t *key, uint32_t n,
                         const uint64_t *m, double alpha, uint64_t *out, int B)
{
    FHE_CHECK_ARG(key && m && out && B >= 0 && n >= 1, "fhe_tlwe_encrypt_u64: bad argument");
    for (int b = 0; b < B; b++) {
        Stream s(seed, TAG_TLWE, index0 + uint64_t(b));
        tlwe_encrypt<uint64_t>(s, key, n, m[b], alpha, out + size_t(b) * (n + 1));
    }
    return FHE_OK;
}

int fhe_trlwe_encrypt_u32(uint64_t seed, uint64_t index0, const uint32_t *key, const uint32_t *m,
                          double alpha, uint32_t *out, int B)
{
    FHE_CHECK_ARG(key && m && out && B >= 0, "fhe_trlwe_encrypt_u32: bad argument");
    const int N = Lvl1::N;
    for (int b = 0; b < B; b++) {
        Stream s(seed, TAG_TRLWE, index0 + uint64_t(b));
        uint32_t *c = out + size_t(b) * 2 * N;
        trlwe_encrypt_zero<uint32_t>(s, key, N, alpha, c);
        for (int i = 0; i < N; i++) c[N + i] += m[size_t(b) * N + i];
    }
    return FHE_OK;
}

int fhe_trlwe_encrypt_u64(uint64_t seed, uint64_t index0, const uint64_t *key, const uint64_t *m,
                          double alpha, uint64_t *out, int B)
{
    FHE_CHECK_ARG(key && m && out && B >= 0, "fhe_trlwe_encrypt_u64: bad argument");
    const int N = Lvl2::N;
    for (int b = 0; b < B; b++) {
        Stream s(seed, TAG_TRLWE, index0 + uint64_t(b));
        uint64_t *c = out + size_t(b) * 2 * N;
        trlwe_encrypt_zero<uint64_t>(s, key, N, alpha, c);
        for (int i = 0; i < N; i++) c[N + i] += m[size_t(b) * N + i];
    }
    return FHE_OK;
}

int fhe_trgsw_encrypt_u32(uint64_t seed, uint64_t index, const uint32_t *key, int32_t m,
                          double alpha, uint32_t *out)
{
    FHE_CHECK_ARG(key && out, "fhe_trgsw_encrypt_u32: null argument");
    Stream s(seed, TAG_TRGSW, index);
    trgsw_encrypt_const<uint32_t, Lvl1::N, Lvl1::l, Lvl1::Bgbit>(s, key, uint32_t(m), alpha, out);
    return FHE_OK;
}

int fhe_trgsw_encrypt_u64(uint64_t seed, uint64_t index, const uint64_t *key, int64_t m,
                          double alpha, uint64_t *out)
{
    FHE_CHECK_ARG(key && out, "fhe_trgsw_encrypt_u64: null argument");
    Stream s(seed, TAG_TRGSW, index);
    trgsw_encrypt_const<uint64_t, Lvl2::N, Lvl2::l, Lvl2::Bgbit>(s, key, uint64_t(m), alpha, out);
    return FHE_OK;
}

}  // extern "C"
