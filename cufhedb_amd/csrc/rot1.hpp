// rot1.hpp -- LDS layout and column preparation of the lvl1 blind rotation,
// shared by the FP64 kernel (pbs_lvl1.hip) and the split-MFMA kernel
// (pbs_lvl1_mfma.hip).
//
// Reference: PolynomialMulByXaiMinusOne (thirdparty/TFHEpp/include/utils.hpp:110-141)
// followed by the gadget decomposition (thirdparty/TFHEpp/src/trgsw.cpp:9-39),
// as the CMUX applies them (thirdparty/TFHEpp/src/detwfa.cpp:21-30).
#pragma once

#include "fft64.hpp"
#include "params.hpp"

namespace fhe {

// decomposition offset and the flip of every 6-bit field's top bit: a digit
// of (x + offset) ^ flip is one signed 6-bit field extract
constexpr uint32_t kBr1DecOff = decomp_total_offset<uint32_t, 32, Lvl1::l, Lvl1::Bgbit>();
constexpr uint32_t kBr1DigitFlip = (1u << 31) | (1u << 25) | (1u << 19);

// LDS of one blind rotation: the accumulator's two columns with the
// transform tile between them (17 KB).  Between transforms the tile's first
// 16 rows of 64 words extend column 0 and its last 16 rows precede column 1
// (br1_prep_column).  16-byte alignment lets the tile's double2 accesses be
// single 16-byte LDS operations (with 8-byte alignment the compiler splits
// them, and the FP64 kernel spills 26 registers and runs 2.5x slower).
struct alignas(16) Br1Lds {
    uint32_t acc0[Lvl1::N];
    cd tile[kFftLdsElems];
    uint32_t acc1[Lvl1::N];
};
// words from column 0 to column 1
constexpr int kBr1Col1Words = Lvl1::N + kFftLdsElems * 4;
static_assert(kFftLdsElems * 16 >= 2 * Lvl1::N * 4, "the tile holds both columns' extensions");

// (X^abar - 1) acc_c at the lane's 16 coefficients n = lane + 64 k, prepared
// for digit extraction ((w + offset) ^ flip) into u[k].  With r = (lane -
// abar) mod 2048, (X^abar acc)[n] is acc_c[(r + 64 k) mod 1024], negated when
// (r + 64 k) mod 2048 >= 1024.  The column's negation is written as 16 rows
// of 64 words right after column 0 (the tile's first rows) or right before
// column 1 (its last rows); then word k sits at byte 4 (r mod 1024) + 256 k
// from one base, the wrap's sign included, and only the sign from bit 10 of
// r (s = 0 or ~0, the same for the 16 words; flipped for column 1, whose base
// is the negated rows) is left: (v ^ s) + (K - s - p) is +-v - p + K.  The
// tile must be free (no transform in flight); the caller's wave owns lds.
// pn[k] is the column's word lane + 64 k, passed in by a caller that still
// holds it from the accumulator update (br1_prep_column reads it from LDS).
__device__ __forceinline__ void br1_prep_column_pn(Br1Lds &lds, int c, uint32_t abar, int lane,
                                                   const uint32_t pn[16], uint32_t u[16])
{
    constexpr int N1 = Lvl1::N;
    uint32_t *col = lds.acc0 + c * kBr1Col1Words;
    uint32_t *neg = c == 0 ? col + N1 : col - N1;
#pragma unroll
    for (int k = 0; k < 16; k++) neg[lane + 64 * k] = 0u - pn[k];
    wave_lds_sync();
    const uint32_t rr = uint32_t(lane) - abar;
    const uint32_t *src = col - c * N1 + (rr & (N1 - 1));
    const uint32_t s = uint32_t(__builtin_amdgcn_sbfe(int32_t(rr), 10, 1)) ^ (0u - uint32_t(c));
    // K - s held opaque: three operations per word (a subtract, an xor-add,
    // the flip); seen through, K - s - p is reassociated into four
    uint32_t kc = kBr1DecOff - s;
    asm volatile("" : "+v"(kc));
#pragma unroll
    for (int k = 0; k < 16; k++) u[k] = ((src[64 * k] ^ s) + (kc - pn[k])) ^ kBr1DigitFlip;
}

__device__ __forceinline__ void br1_prep_column(Br1Lds &lds, int c, uint32_t abar, int lane,
                                                uint32_t u[16])
{
    const uint32_t *col = lds.acc0 + c * kBr1Col1Words;
    uint32_t pn[16];
#pragma unroll
    for (int k = 0; k < 16; k++) pn[k] = col[lane + 64 * k];
    br1_prep_column_pn(lds, c, abar, lane, pn, u);
}

}  // namespace fhe
