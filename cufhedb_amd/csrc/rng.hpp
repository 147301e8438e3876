// rng.hpp -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the samplers the
// key generator needs.  Every random quantity is addressed by
// (seed, domain tag, record index, word index), so any record can be
// regenerated independently and in parallel.
#pragma once
#include <cmath>
#include <cstdint>

// Philox is also evaluated by device code (keygen_dev.hip); the host and
// device versions are the same integer arithmetic.
#if defined(__HIPCC__)
#define FHE_HD __host__ __device__
#else
#define FHE_HD
#endif

namespace fhe {

struct Philox {
    uint32_t k0, k1;
    FHE_HD explicit Philox(uint64_t seed) : k0(uint32_t(seed)), k1(uint32_t(seed >> 32)) {}

    FHE_HD static inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo)
    {
        uint64_t p = uint64_t(a) * b;
        hi = uint32_t(p >> 32);
        lo = uint32_t(p);
    }
    // 4 words for counter (c0..c3)
    FHE_HD inline void block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                             uint32_t out[4]) const
    {
        uint32_t key0 = k0, key1 = k1;
        for (int r = 0; r < 10; r++) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c0, hi0, lo0);
            mulhilo(0xCD9E8D57u, c2, hi1, lo1);
            uint32_t n0 = hi1 ^ c1 ^ key0;
            uint32_t n1 = lo1;
            uint32_t n2 = hi0 ^ c3 ^ key1;
            uint32_t n3 = lo0;
            c0 = n0; c1 = n1; c2 = n2; c3 = n3;
            key0 += 0x9E3779B9u;
            key1 += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
};

// Domain tags: one independent stream family per kind of random quantity.
enum RngTag : uint32_t {
    TAG_SK0 = 1, TAG_SK1 = 2, TAG_SK2 = 3,
    TAG_BK01 = 0x10, TAG_BK02 = 0x11,
    TAG_KSK10 = 0x20, TAG_KSK20 = 0x21, TAG_KSK21 = 0x22,
    TAG_TLWE = 0x30, TAG_TRLWE = 0x31, TAG_TRGSW = 0x32,
};

// Sequential word stream for one record: (seed, tag, record) -> words.
class Stream {
  public:
    Stream(uint64_t seed, uint32_t tag, uint64_t record)
        : ph_(seed), rec_lo_(uint32_t(record)), rec_hi_(uint32_t(record >> 32)), tag_(tag) {}

    inline uint32_t u32()
    {
        if (pos_ == 4) {
            ph_.block(ctr_++, rec_lo_, rec_hi_, tag_, buf_);
            pos_ = 0;
        }
        return buf_[pos_++];
    }
    inline uint64_t u64() { uint64_t lo = u32(); return lo | (uint64_t(u32()) << 32); }
    // position the stream at word w of the record (the next u32() returns it)
    inline void seek(uint64_t w)
    {
        ctr_ = uint32_t(w >> 2);
        pos_ = 4;
        for (uint64_t i = 0; i < (w & 3); i++) (void)u32();
    }
    // uniform double in (0, 1], 53-bit
    inline double unit_open0()
    {
        uint64_t v = u64() >> 11;
        return (double(v) + 1.0) * (1.0 / 9007199254740992.0);
    }
    // standard normal (Box-Muller, cosine branch)
    inline double normal()
    {
        double u1 = unit_open0(), u2 = unit_open0();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // uniform integer in [0, n)
    inline uint32_t below(uint32_t n) { return uint32_t((uint64_t(u32()) * n) >> 32); }

  private:
    Philox ph_;
    uint32_t rec_lo_, rec_hi_, tag_;
    uint32_t ctr_ = 0;
    uint32_t buf_[4] = {0, 0, 0, 0};
    int pos_ = 4;
};

// ModularGaussian (thirdparty/TFHEpp/include/utils.hpp:41-66): 32-bit torus
// noise via dtot32 (:35-38), 64-bit torus noise via stdev * z * 2^64.
inline uint32_t gaussian_t32(Stream &s, double stdev)
{
    double d = s.normal() * stdev;
    return uint32_t(int32_t(int64_t((d - double(int64_t(d))) * 4294967296.0)));
}
inline uint64_t gaussian_t64(Stream &s, double stdev)
{
    double v = stdev * s.normal() * 18446744073709551616.0;
    return uint64_t(int64_t(v));
}

}  // namespace fhe
