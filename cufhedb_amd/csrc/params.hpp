// params.hpp -- parameter sets of the TFHE bootstrapping path.
//
// Ring / gadget / key-switch constants follow the reference's compiled set
// (thirdparty/TFHEpp/include/params/128bit.hpp:7-97): lvl1 N=1024 l=3 Bgbit=6,
// lvl2 N=2048 l=4 Bgbit=9, lvl10 t=2 basebit=10, lvl20 t=2 basebit=10,
// lvl21 t=10 basebit=3.  The lvl0 dimension n0, its noise and the secret-key
// distributions are runtime values (fhe_params) so that both the reference set
// (n0 = 672, alpha0 = 2^-16, Hamming-weight keys) and the n0 = 636 set run
// through the same kernels.
#pragma once
#include <cstdint>

namespace fhe {

constexpr int kMaxN0 = 1024;  // kernels accept any n0 in [1, kMaxN0]

// Widest lvl2 HomMSB input (plaintext bits including the comparison's sign
// bit): each 5-bit peeling step shifts the previous identity bootstrap's
// output noise (std 2^36.2 on the 64-bit torus) left by pb - 11, onto the
// 2^57 decision margin of the 6-bit window; from 28 bits on outputs next to
// a slot boundary flip (error rates from ~1e-4 at 28 bits to ~50 % at 33).
constexpr int kMaxMsbBits2 = 27;

struct Lvl1 {
    static constexpr int N = 1024;
    static constexpr int nbit = 10;
    static constexpr int l = 3;
    static constexpr int Bgbit = 6;
    static constexpr int M = N / 2;          // complex FD points
    static constexpr int rows = 2 * l;       // (k+1) l TRGSW rows
    using T = uint32_t;
    static constexpr uint32_t mu = 1u << 29;  // lvl1param::μ
};

struct Lvl2 {
    static constexpr int N = 2048;
    static constexpr int nbit = 11;
    static constexpr int l = 4;
    static constexpr int Bgbit = 9;
    static constexpr int M = N / 2;
    static constexpr int rows = 2 * l;
    using T = uint64_t;
    static constexpr uint64_t mu = 1ull << 61;  // lvl2param::μ
};

// identity key-switch kinds: domain level -> target level
struct KS10 { static constexpr int t = 2, basebit = 10; };
struct KS20 { static constexpr int t = 2, basebit = 10; };
struct KS21 { static constexpr int t = 10, basebit = 3; };

// gadget offset (offsetgen, thirdparty/TFHEpp/src/trgsw.cpp:9-18) plus the
// rounding offset of DecompositionPolynomial (:27-29), folded into one add.
template <class T, int W, int L, int B>
constexpr T decomp_total_offset()
{
    T off = 0;
    for (int i = 1; i <= L; i++) off += (T(1) << (B - 1)) * (T(1) << (W - i * B));
    return off + (T(1) << (W - L * B - 1));
}

}  // namespace fhe
