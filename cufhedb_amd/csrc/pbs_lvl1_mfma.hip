// pbs_lvl1_mfma.hip -- lvl1 blind rotation with the forward (digit)
// transforms on the fp16 matrix cores and the bootstrapping key split into
// integer limbs: the "split MFMA" variant.
//
// Same blind rotation, decomposition, rotation and sample extraction as
// pbs_lvl1.hip (reference: thirdparty/TFHEpp/include/gatebootstrapping.hpp:15-43,
// src/detwfa.cpp:21-30, src/trgsw.cpp:9-103; GPU counterpart
// src/cuHEDB/fft_gpu/BlindRotate_gpu.cuh:401-531; the multiply-accumulate it
// replaces: src/cuHEDB/fft_gpu/ExternalProduct_gpu.cuh:67-133,
// thirdparty/TFHEpp/include/utils.hpp:83-108).  What changes is the
// arithmetic of the external product.
//
// Forward transforms.  The six digit polynomials of every CMUX reach the
// frequency domain as
//
//   Z[m2 + 16 m1] = sum_n1 W32^{n1 m1} T(n1, m2) sum_n2 A[m2][n2] z[n1 + 32 n2]
//     A[m2][n2] = exp(i pi n2 (4 m2 + 1) / 32)      (radix-16 stage + twist)
//     T(n1, m2) = exp(i pi n1 (4 m2 + 1) / 1024)    (inter-stage twiddle)
//
// which is the twisted negacyclic map Z[m] = sum_n z_n w^{n(4m+1)} of the FP64
// path, factored as n = n1 + 32 n2, m = m2 + 16 m1.  Both stages are real
// matrix products on v_mfma_f32_*_f16 with FP32 accumulation:
//   stage A  32x32x16: the digits (|d| <= 32, exact in fp16) times the
//            matrix A split in two fp16 limbs (hi + lo, ~22 bits): 4 MFMAs;
//   stage B  16x16x32: the twiddled stage-A output (FP32) and W32, both
//            split hi + lo, products hi*hi + hi*lo + lo*hi: 24 MFMAs.
// The data never leaves the wave: stage A consumes the digits in the lane
// layout the rotation produces (lane L holds z[L + 64 j]); one LDS
// transpose (16 x 34 float2, conflict-free) feeds stage B.  Relative RMS
// error of the transform: 2^-22.7 (FP32 accumulation bounds it).
//
// Key limbs.  That error is far too large for a 32-bit key: the product
// sums reach 2^40.7 (RMS) before the reduction mod 2^32, so a 2^-22.7
// relative error leaves 2^18 per coefficient and 2^27 of output phase
// noise, at the Z_8 decision margin.  The key's torus words are therefore
// split into three signed integer limbs, K = K0 + 2^16 K1 + 2^24 K2 mod
// 2^32 with K0 in [-2^15, 2^15), K1, K2 in [-2^7, 2^7), each limb's
// polynomial transformed once (FP64, at key load) and stored in FP32.  Per
// CMUX and output column, the frequency-domain products with each limb are
// summed over the six rows in FP32 and brought back by one FP64 inverse
// transform per limb; the limb products are exact integers:
//   limbs 1, 2: |product| RMS 2^16.7, error RMS 2^-5.4 -> rounded exactly
//               (0.5 is 21 standard deviations);
//   limb 0:     RMS 2^24.7, error RMS 2^2.6, kept as is (weight 1);
// and the three are recombined with the shifts and reduced mod 2^32 as
// integers.  The external product thus carries ~2^2.6 of arithmetic error
// per coefficient (the FP64 path: below 1), far under the key noise, so the
// output phase noise equals the FP64 path's.
#include "fft64.hpp"
#include "launch.hpp"
#include "params.hpp"
#include "rot1.hpp"

namespace fhe {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int N1 = Lvl1::N;
constexpr int kStrideB = 34;  // float2 row stride of the stage-A -> stage-B transpose

// Per-lane constants of the two stages (built once on the device).
struct MfmaTables {
    half8 a[2][2][64];        // [limb][K step][lane]   stage-A matrix, A operand
    half8 w[2][2][2][64];     // [re/im][limb][tile][lane]  W32, B operand (tile 1 unused)
    float2 t[64][3];          // [lane]: T(n1, 2h), w^{4 n1}, w^{16 n1}  twiddle bases
};

// signed gadget digit d (DecompositionPolynomial, trgsw.cpp:21-39) of a word
// prepared as (x + offset) ^ flip (br1_prep_column), the top bit of every
// 6-bit field flipped: the field read as a signed 6-bit number is the digit
__device__ __forceinline__ int digit1mp(uint32_t u, int d)
{
    return int32_t(u << (Lvl1::Bgbit * d)) >> (32 - Lvl1::Bgbit);
}

// conflict-free slot of frequency m in the 512-entry float2 permutation buffer
__device__ __forceinline__ int perm_slot(int m)
{
    return m ^ (((m >> 3) & 7) << 2) ^ ((m >> 6) & 3);
}

// negate elements 1, 3, 5, 7 (the high half of every dword)
__device__ __forceinline__ half8 neg_odd(half8 x)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, x);
    u ^= 0x80008000u & 0xFFFF0000u;
    return __builtin_bit_cast(half8, u);
}

__device__ __forceinline__ void split2(float x, _Float16 &hi, _Float16 &lo)
{
    hi = _Float16(x);
    lo = _Float16(x - float(hi));
}

__global__ void k_mfma_tables(MfmaTables *tab)
{
    const int lane = threadIdx.x;  // one block of 64
    for (int s = 0; s < 2; s++) {
        const int i = lane & 31;
        const int t = i >> 3, hr = (i >> 2) & 1, q = i & 3;
        const int m2 = 4 * t + 2 * hr + (q >> 1), eo = q & 1;
        for (int e = 0; e < 8; e++) {
            const int n2 = 2 * (4 * s + (e >> 1)) + (lane >> 5), ei = e & 1;
            double sv, cv;
            sincospi(double(n2 * (4 * m2 + 1)) / 32.0, &sv, &cv);
            const double val = eo == 0 ? (ei == 0 ? cv : -sv) : (ei == 0 ? sv : cv);
            const _Float16 hi = _Float16(val);
            tab->a[0][s][lane][e] = hi;
            tab->a[1][s][lane][e] = _Float16(val - double(hi));
        }
    }
    for (int tile = 0; tile < 2; tile++)
        for (int e = 0; e < 8; e++) {
            const int n1 = 8 * (lane >> 4) + e, m1 = 16 * tile + (lane & 15);
            double sv, cv;
            sincospi(double((n1 * m1) & 31) / 16.0, &sv, &cv);
            const _Float16 rh = _Float16(cv), ih = _Float16(sv);
            tab->w[0][0][tile][lane][e] = rh;
            tab->w[0][1][tile][lane][e] = _Float16(cv - double(rh));
            tab->w[1][0][tile][lane][e] = ih;
            tab->w[1][1][tile][lane][e] = _Float16(sv - double(ih));
        }
    // T(n1, m2) = w^{n1 (4 m2 + 1)}, w = exp(i pi / 1024); a lane needs
    // m2 = 4 tt + 2 h + u: T(n1, 2h) times powers of w^{4 n1} (u) and w^{16 n1} (tt)
    const int n1 = lane & 31, h = lane >> 5;
    const int ex[3] = {n1 * (8 * h + 1), 4 * n1, 16 * n1};
    for (int k = 0; k < 3; k++) {
        double sv, cv;
        sincospi(double(ex[k]) / 1024.0, &sv, &cv);
        tab->t[lane][k] = float2{float(cv), float(sv)};
    }
}

// Limb split of a key torus word: K = K0 + 2^16 K1 + 2^24 K2 (mod 2^32),
// K0 in [-2^15, 2^15), K1 and K2 in [-2^7, 2^7) (signed, balanced).
constexpr int kLimbs = 3;
__host__ __device__ constexpr int limb_shift(int b) { return b == 0 ? 0 : b == 1 ? 16 : 24; }
__device__ __forceinline__ void split_limbs(int64_t k, double out[kLimbs])
{
    const int64_t k0 = int64_t(int16_t(uint16_t(k)));
    const int64_t r1 = (k - k0) >> 16;
    const int64_t k1 = int64_t(int8_t(uint8_t(r1)));
    const int64_t r2 = (r1 - k1) >> 8;
    const int64_t k2 = int64_t(int8_t(uint8_t(r2)));
    out[0] = double(k0);
    out[1] = double(k1);
    out[2] = double(k2);
}

// round to the nearest integer, mod 2^32, for |x| < 2^51: adding 1.5 * 2^52
// rounds x to an integer (round-to-nearest-even) whose two's complement
// residue mod 2^32 is the low mantissa word
__device__ __forceinline__ uint32_t round_torus32(double x)
{
    return uint32_t(__double_as_longlong(x + 0x1.8p52));
}

// stage-B output position (float2 index within a 512-point polynomial,
// [k2 < 4][lane][2]) of frequency m: lane l, element k of the forward
// transform holds frequency 4 (l >> 4) + (k & 3) + 16 (l & 15) + 256 (k >> 2)
__device__ __forceinline__ int stage_b_slot(int m)
{
    const int mm = m & 255;
    const int l = (mm >> 4) + 16 * ((mm >> 2) & 3);
    const int k = (mm & 3) + 4 * (m >> 8);
    return (k >> 1) * 128 + l * 2 + (k & 1);
}

// FP64 device-layout key polynomial (lane s, register q <-> frequency
// fd_index1024(s, q), scaled by 1/512) -> its torus words (exact: the FP64
// inverse of an integer polynomial's transform rounds back to it) -> three
// limb polynomials -> FP64 forward transforms, scaled by 1/512, stored in
// FP32 in stage B's output order: [i][row][col][limb][k2 < 4][lane][2].
// One wave per key polynomial.
__global__ __launch_bounds__(64) void k_bk1_limbs_mfma(const cd *__restrict__ src,
                                                        float2 *__restrict__ dst, int npoly)
{
    __shared__ cd tile[kFftLdsElems];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    if (p >= npoly) return;
    const Twiddles1024 tw = make_twiddles1024(lane);
    cd v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = src[size_t(p) * 512 + q * 64 + lane];
    fft512_inv(v, tile, lane, tw);
    double lx[8][kLimbs], ly[8][kLimbs];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        split_limbs(int64_t(rint(v[j].x)), lx[j]);
        split_limbs(int64_t(rint(v[j].y)), ly[j]);
    }
#pragma unroll 1
    for (int b = 0; b < kLimbs; b++) {
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = cd{lx[j][b], ly[j][b]};
        fft512_fwd(v, tile, lane, tw);
        float2 *o = dst + (size_t(p) * kLimbs + b) * 512;
#pragma unroll
        for (int q = 0; q < 8; q++)
            o[stage_b_slot(fd_index1024(lane, q))] =
                float2{float(v[q].x * (1.0 / 512)), float(v[q].y * (1.0 / 512))};
    }
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return float2{__fmaf_rn(a.x, b.x, -a.y * b.y), __fmaf_rn(a.x, b.y, a.y * b.x)};
}

__device__ __forceinline__ float2 fmac2(float2 acc, float2 a, float2 b)
{
    acc.x = __fmaf_rn(a.x, b.x, acc.x);
    acc.x = __fmaf_rn(-a.y, b.y, acc.x);
    acc.y = __fmaf_rn(a.x, b.y, acc.y);
    acc.y = __fmaf_rn(a.y, b.x, acc.y);
    return acc;
}

// Forward split-MFMA transform of one digit polynomial held as stage-A B
// operands (b[s] element e = re/im of z[lane + 64 (4 s + e / 2)]), result as
// two 16x16 output tiles: lane l, tile t, element q holds frequency
// 4 (l >> 4) + q + 16 (l & 15) + 256 t.  buf: 16 x 34 float2 of LDS.
struct MfmaConsts {
    half8 ah[2], al[2];     // stage-A matrix limbs
    float2 tb, t4, t16;     // twiddle bases
};

__device__ __forceinline__ MfmaConsts load_consts(const MfmaTables *__restrict__ tab, int lane)
{
    MfmaConsts mc;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        mc.ah[s] = tab->a[0][s][lane];
        mc.al[s] = tab->a[1][s][lane];
    }
    mc.tb = tab->t[lane][0];
    mc.t4 = tab->t[lane][1];
    mc.t16 = tab->t[lane][2];
    return mc;
}

__device__ __forceinline__ void mfma_fwd(const half8 b[2], const MfmaConsts &mc,
                                         const MfmaTables *__restrict__ tab, float2 *buf, int lane,
                                         f32x4 zr[2], f32x4 zi[2])
{
    const int n1 = lane & 31, h = lane >> 5;
    f32x16 C = {};
    C = __builtin_amdgcn_mfma_f32_32x32x16_f16(mc.ah[0], b[0], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_32x32x16_f16(mc.al[0], b[0], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_32x32x16_f16(mc.ah[1], b[1], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_32x32x16_f16(mc.al[1], b[1], C, 0, 0, 0);
    // inter-stage twiddle, then transpose Y'[m2][n1] through LDS
    {
        float2 tw = mc.tb;
#pragma unroll
        for (int tt = 0; tt < 4; tt++) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const float2 tk = u ? cmulf(tw, mc.t4) : tw;
                const float2 y = cmulf(float2{C[4 * tt + 2 * u], C[4 * tt + 2 * u + 1]}, tk);
                buf[(4 * tt + 2 * h + u) * kStrideB + n1] = y;
            }
            if (tt < 3) tw = cmulf(tw, mc.t16);
        }
    }
    wave_lds_sync();
    half8 yrh, yrl, yih, yil, nyih, nyil;
    {
        const float4 *row =
            reinterpret_cast<const float4 *>(buf + (lane & 15) * kStrideB + 8 * (lane >> 4));
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const float4 x = row[p];
            _Float16 hi, lo;
            split2(x.x, hi, lo); yrh[2 * p] = hi; yrl[2 * p] = lo;
            split2(x.y, hi, lo); yih[2 * p] = hi; yil[2 * p] = lo;
            split2(x.z, hi, lo); yrh[2 * p + 1] = hi; yrl[2 * p + 1] = lo;
            split2(x.w, hi, lo); yih[2 * p + 1] = hi; yil[2 * p + 1] = lo;
        }
    }
    wave_lds_sync();
    nyih = -yih;
    nyil = -yil;
    // W32 columns 16..31 equal columns 0..15 times (-1)^{n1}: tile 1 runs on
    // tile 0's matrix with the odd-n1 (odd-element) inputs negated
    const half8 wrh = tab->w[0][0][0][lane], wrl = tab->w[0][1][0][lane];
    const half8 wih = tab->w[1][0][0][lane], wil = tab->w[1][1][0][lane];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        if (t == 1) {
            yrh = neg_odd(yrh); yrl = neg_odd(yrl);
            yih = neg_odd(yih); yil = neg_odd(yil);
            nyih = neg_odd(nyih); nyil = neg_odd(nyil);
        }
        f32x4 a = {}, c = {};
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrh, wrh, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrh, wrl, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrl, wrh, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(nyih, wih, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(nyih, wil, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(nyil, wih, a, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrh, wih, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrh, wil, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yrl, wih, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yih, wrh, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yih, wrl, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(yil, wrh, c, 0, 0, 0);
        zr[t] = a;
        zi[t] = c;
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void
k_blind_rotate1_mfma(const uint32_t *__restrict__ in, int n0, const float4 *__restrict__ bkm,
                     const MfmaTables *__restrict__ tab, const uint32_t *__restrict__ tv,
                     int tv_stride, uint32_t *__restrict__ out, int sample_extract, int B)
{
    __shared__ Br1Lds lds;  // one wave: all of it is the wave's own
    uint32_t *acc0 = lds.acc0, *acc1 = lds.acc1;
    cd *tile = lds.tile;
    const int lane = threadIdx.x;
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const uint32_t *c0 = in + size_t(ct) * (n0 + 1);
    const uint32_t *tvp = tv + size_t(ct) * tv_stride;
    float2 *buf = reinterpret_cast<float2 *>(tile);  // 576 x 16 B >= 544 / 512 float2 needed

    const Twiddles1024 tw64 = make_twiddles1024(lane);

    const uint32_t bbar = 2u * N1 - (c0[n0] >> (32 - 1 - Lvl1::nbit));
#pragma unroll
    for (int k = 0; k < N1 / 64; k++) {
        const int n = lane + 64 * k;
        acc0[n] = 0;
        uint32_t idx = (uint32_t(n) - bbar) & (2u * N1 - 1);
        uint32_t v = tvp[idx & (N1 - 1)];
        acc1[n] = (idx & N1) ? 0u - v : v;
    }
    wave_lds_sync();

    // key slice of one (row, column, limb): 256 float4
    constexpr int kSlice = 256;
    constexpr uint32_t kRound = 1u << (32 - 2 - Lvl1::nbit);
    for (int i = 0; i < n0; i++) {
        const uint32_t abar = (c0[i] + kRound) >> (32 - 1 - Lvl1::nbit);
        if (abar == 0) continue;
        const float4 *bki = bkm + size_t(i) * Lvl1::rows * 2 * kLimbs * kSlice;
        // the six digit polynomials in the frequency domain (stage-B order:
        // element k = 4 t + q of lane l is frequency 4 (l >> 4) + q + 16 (l & 15)
        // + 256 t), kept in registers for the products with every key limb
        float2 zf[Lvl1::rows][8];
        // (X^abar - 1) acc_c at the lane's 16 coefficients (u[k]: n = lane +
        // 64 k), prepared for digit extraction once per column (at its first
        // row) and kept in registers (br1_prep_column)
        uint32_t u[16];
#pragma unroll
        for (int r = 0; r < Lvl1::rows; r++) {
            const int c = r / Lvl1::l, d = r - c * Lvl1::l;
            if (d == 0) br1_prep_column(lds, c, abar, lane, u);
            // digits as stage-A B operands: step s holds j = 4s..4s+3 (re, im)
            half8 b[2];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                b[j >> 2][2 * (j & 3)] = _Float16(digit1mp(u[j], d));
                b[j >> 2][2 * (j & 3) + 1] = _Float16(digit1mp(u[j + 8], d));
            }
            f32x4 zr[2], zi[2];
            // the stage constants are re-read (L1 / L2) per transform rather
            // than held: registers go to the six transforms' outputs
            const MfmaConsts mc = load_consts(tab, lane);
            mfma_fwd(b, mc, tab, buf, lane, zr, zi);
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) zf[r][4 * t + q] = float2{zr[t][q], zi[t][q]};
        }
        // per output column: for each key limb, the products summed over the
        // rows (FP32), back to the time domain (FP64), rounded and shifted;
        // the limbs' sum mod 2^32 goes into acc
#pragma unroll 1
        for (int c = 0; c < 2; c++) {
#pragma unroll 1
            for (int lb = 0; lb < kLimbs; lb++) {
                float2 fd[8];
#pragma unroll
                for (int k = 0; k < 8; k++) fd[k] = float2{0.f, 0.f};
#pragma unroll
                for (int r = 0; r < Lvl1::rows; r++) {
                    const float4 *kr = bki + size_t((r * 2 + c) * kLimbs + lb) * kSlice + lane;
                    float4 kv[4];
#pragma unroll
                    for (int k2 = 0; k2 < 4; k2++) kv[k2] = kr[k2 * 64];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const float4 p = kv[k >> 1];
                        const float2 key = (k & 1) ? float2{p.z, p.w} : float2{p.x, p.y};
                        fd[k] = fmac2(fd[k], zf[r][k], key);
                    }
                }
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int m = 4 * (lane >> 4) + (k & 3) + 16 * (lane & 15) + 256 * (k >> 2);
                    buf[perm_slot(m)] = fd[k];
                }
                wave_lds_sync();
                cd v[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const float2 f = buf[perm_slot(fd_index1024(lane, q))];
                    v[q] = cd{double(f.x), double(f.y)};
                }
                wave_lds_sync();
                fft512_inv<true, true>(v, tile, lane, tw64);
                // the digits of every row are already in zf, so acc may
                // take each limb's share at once (lane-private words)
                const int sh = limb_shift(lb);
                uint32_t *col = acc0 + c * kBr1Col1Words;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int n = lane + 64 * j;
                    col[n] += round_torus32(v[j].x) << sh;
                    col[n + 512] += round_torus32(v[j].y) << sh;
                }
            }
        }
        wave_lds_sync();
    }

    if (sample_extract) {
        uint32_t *o = out + size_t(ct) * (N1 + 1);
#pragma unroll
        for (int k = 0; k < N1 / 64; k++) {
            const int n = lane + 64 * k;
            o[n] = n == 0 ? acc0[0] : 0u - acc0[N1 - n];
        }
        if (lane == 0) o[N1] = acc1[0];
    } else {
        uint32_t *o = out + size_t(ct) * 2 * N1;
#pragma unroll
        for (int k = 0; k < N1 / 64; k++) {
            o[lane + 64 * k] = acc0[lane + 64 * k];
            o[N1 + lane + 64 * k] = acc1[lane + 64 * k];
        }
    }
}

// The split-MFMA forward transform alone: B small-integer polynomials
// (|x| <= 2048, exact in fp16; gadget digits are |x| <= 32) -> TFHEpp FD
// layout (spqlios order, like k_ifft1), for parity tests of the transform.
__global__ __launch_bounds__(64) void k_fwd_digits_mfma(const int32_t *__restrict__ poly,
                                                         const MfmaTables *__restrict__ tab,
                                                         double *__restrict__ fd, int B)
{
    __shared__ float2 buf[16 * kStrideB];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    if (p >= B) return;
    const MfmaConsts mc = load_consts(tab, lane);
    const int32_t *a = poly + size_t(p) * N1;
    half8 b[2];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        b[j >> 2][2 * (j & 3)] = _Float16(a[lane + 64 * j]);
        b[j >> 2][2 * (j & 3) + 1] = _Float16(a[lane + 64 * j + 512]);
    }
    f32x4 zr[2], zi[2];
    mfma_fwd(b, mc, tab, buf, lane, zr, zi);
    double *o = fd + size_t(p) * N1;
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = 4 * (lane >> 4) + q + 16 * (lane & 15) + 256 * t;
            o[rev9(m)] = zr[t][q];
            o[rev9(m) + 512] = zi[t][q];
        }
}

}  // namespace

hipError_t launch_fwd_digits_mfma(const int32_t *poly, const void *tab, double *fd, int B,
                                  hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fwd_digits_mfma, dim3(B), dim3(64), 0, s, poly,
                       static_cast<const MfmaTables *>(tab), fd, B);
    return hipGetLastError();
}

size_t mfma_tables_bytes() { return sizeof(MfmaTables); }
size_t bk1_mfma_floats(int n0) { return size_t(n0) * Lvl1::rows * 2 * kLimbs * 512 * 2; }

hipError_t launch_mfma_tables(void *tab, hipStream_t s)
{
    hipLaunchKernelGGL(k_mfma_tables, dim3(1), dim3(64), 0, s, static_cast<MfmaTables *>(tab));
    return hipGetLastError();
}

hipError_t launch_bk1_to_mfma(const cd *bk_dev, float *bkm, int n0, hipStream_t s)
{
    const int npoly = n0 * Lvl1::rows * 2;
    hipLaunchKernelGGL(k_bk1_limbs_mfma, dim3(npoly), dim3(64), 0, s, bk_dev,
                       reinterpret_cast<float2 *>(bkm), npoly);
    return hipGetLastError();
}

hipError_t launch_blind_rotate1_mfma(const uint32_t *in, int n0, const float *bkm, const void *tab,
                                     const uint32_t *tv, int tv_stride, uint32_t *out,
                                     bool sample_extract, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_blind_rotate1_mfma, dim3(B), dim3(64), 0, s, in, n0,
                       reinterpret_cast<const float4 *>(bkm), static_cast<const MfmaTables *>(tab),
                       tv, tv_stride, out, int(sample_extract), B);
    return hipGetLastError();
}

}  // namespace fhe
