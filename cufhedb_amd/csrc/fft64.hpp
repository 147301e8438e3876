// fft64.hpp -- FP64 twisted negacyclic FFT of N = 1024 (512 complex points)
// for one wave64 on gfx950.
//
// Mathematical map (same as the reference's TwistIFFT / TwistFFT,
// thirdparty/TFHEpp/include/mulfft.hpp:58-121 over spqlios, and cuFHEDB's
// ifft1024 / fft1024, src/cuHEDB/fft_gpu/fft_gpu.cuh:471-799):
//   forward:  z_n = a_n + i a_{n+512},  Z[m] = sum_n z_n w^{n(4m+1)},
//             w = exp(i pi / 1024), m, n in [0, 512)
//   inverse:  z_n = (1/512) w^{-n} sum_m Z[m] exp(-2 pi i n m / 512)
//
// Layout and schedule (this build's own design):
//   time domain : lane L, register j holds z[L + 64 j]            (natural)
//   freq domain : lane s, register q holds Z[(s>>3) + 8 (s&7) + 64 q]
// 512 = 8 x 8 x 8.  Pass 1 is a radix-8 DFT over the 8 registers of a lane;
// the twist w^n is folded into the constants c_j = w^{64 j} before it and
// the per-lane twiddle w^{L(4k+1)} after it.  Two transposes through an
// 8-row double2 LDS tile regroup the data for pass 2 (radix 8, twiddle
// W64^{L0 p}) and pass 3 (radix 8).  Row strides are chosen for the LDS of
// gfx950, which serves a ds_read/write_b128 16 lanes (256 bytes, all 64
// banks) per cycle: the strided side of transpose 1 has lanes (L0, k1)
// reading row k1, column L0 + 8 l1, so the row stride is 72 double2
// (= 32 banks mod 64) and the two rows one cycle touches fill opposite bank
// halves; the strided side of transpose 2 has lanes (p, k1) reading row p,
// column l0 + 8 k1, so that stride is 65 (4 banks mod 64 per row, the
// k1 parity adding 32).  Row-contiguous sides are conflict-free.  The
// inverse runs the same steps mirrored with conjugate twiddles, so it takes
// the permuted frequency layout and returns natural time order.
// Twiddle powers are generated in registers from three per-lane bases, so
// the transform reads no twiddle tables from memory.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fhe {

struct cd {
    double x, y;
};

__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return {__fma_rn(a.x, b.x, -a.y * b.y), __fma_rn(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
__device__ __forceinline__ cd cmulc(cd a, cd b)
{
    return {__fma_rn(a.x, b.x, a.y * b.y), __fma_rn(a.y, b.x, -a.x * b.y)};
}
// acc += a * b
__device__ __forceinline__ void cmac(cd &acc, cd a, cd b)
{
    acc.x = __fma_rn(a.x, b.x, acc.x);
    acc.x = __fma_rn(-a.y, b.y, acc.x);
    acc.y = __fma_rn(a.x, b.y, acc.y);
    acc.y = __fma_rn(a.y, b.x, acc.y);
}

constexpr double kR2 = 0.70710678118654752440;  // sqrt(1/2)

// LDS ordering among the lanes of one wave: a wave executes its LDS
// operations in order, so a compiler barrier with wavefront-scope fences
// suffices when an LDS area is private to the wave.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x * (S i)
template <int S>
__device__ __forceinline__ cd mul_si(cd a)
{
    return S > 0 ? cd{-a.y, a.x} : cd{a.y, -a.x};
}

// in-register DFT of size 4, X[k] = sum_j a_j (S i)^{jk}
template <int S>
__device__ __forceinline__ void dft4(cd &a0, cd &a1, cd &a2, cd &a3)
{
    cd p0 = cadd(a0, a2), p1 = csub(a0, a2);
    cd q0 = cadd(a1, a3), q1 = mul_si<S>(csub(a1, a3));
    a0 = cadd(p0, q0);
    a1 = cadd(p1, q1);
    a2 = csub(p0, q0);
    a3 = csub(p1, q1);
}

// in-register DFT of size 8, natural order in and out, X[k] = sum_j v_j W8^{S jk}
template <int S>
__device__ __forceinline__ void dft8(cd v[8])
{
    cd a0 = cadd(v[0], v[4]), a1 = cadd(v[1], v[5]), a2 = cadd(v[2], v[6]), a3 = cadd(v[3], v[7]);
    cd b0 = csub(v[0], v[4]);
    cd t1 = csub(v[1], v[5]);
    cd t2 = csub(v[2], v[6]);
    cd t3 = csub(v[3], v[7]);
    // W8^S = (1 + S i)/sqrt2, W8^{2S} = S i, W8^{3S} = (-1 + S i)/sqrt2
    cd b1 = S > 0 ? cd{kR2 * (t1.x - t1.y), kR2 * (t1.x + t1.y)}
                  : cd{kR2 * (t1.x + t1.y), kR2 * (t1.y - t1.x)};
    cd b2 = mul_si<S>(t2);
    cd b3 = S > 0 ? cd{kR2 * (-t3.x - t3.y), kR2 * (t3.x - t3.y)}
                  : cd{kR2 * (t3.y - t3.x), kR2 * (-t3.x - t3.y)};
    dft4<S>(a0, a1, a2, a3);
    dft4<S>(b0, b1, b2, b3);
    v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
    v[1] = b0; v[3] = b1; v[5] = b2; v[7] = b3;
}

// c_j = exp(2 pi i j / 32), j = 0..7 (the w^{64 j} part of the twist)
__device__ __forceinline__ cd twist_const(int j)
{
    constexpr double c1 = 0.98078528040323044913, s1 = 0.19509032201612826785;
    constexpr double c2 = 0.92387953251128675613, s2 = 0.38268343236508977173;
    constexpr double c3 = 0.83146961230254523708, s3 = 0.55557023301960222474;
    switch (j) {
    case 0: return {1.0, 0.0};
    case 1: return {c1, s1};
    case 2: return {c2, s2};
    case 3: return {c3, s3};
    case 4: return {kR2, kR2};
    case 5: return {s3, c3};
    case 6: return {s2, c2};
    default: return {s1, c1};
    }
}

// Per-lane twiddle bases (computed once per kernel):
//   wl  = w^L            (pass-1 twist, w = exp(i pi/1024))
//   w4l = w^{4L}         (pass-1 step: W512^L)
//   w64 = W64^{L & 7}    (pass-2 twiddle base, lane role L0 = lane & 7)
struct Twiddles1024 {
    cd wl, w4l, w64;
};

__device__ __forceinline__ cd unit_root(double num, double den)
{
    // exp(i pi num/den)
    double s, c;
    sincospi(num / den, &s, &c);
    return {c, s};
}

__device__ __forceinline__ Twiddles1024 make_twiddles1024(int lane)
{
    Twiddles1024 t;
    t.wl = unit_root(double(lane), 1024.0);
    t.w4l = unit_root(double(lane), 256.0);
    t.w64 = unit_root(double(lane & 7), 32.0);
    return t;
}

// bases for the plain (untwisted) 512-point DFT
__device__ __forceinline__ Twiddles1024 make_twiddles512_plain(int lane)
{
    Twiddles1024 t;
    t.wl = cd{1.0, 0.0};
    t.w4l = unit_root(double(lane), 256.0);
    t.w64 = unit_root(double(lane & 7), 32.0);
    return t;
}

constexpr int kFftPad = 65;                  // row stride of transpose 2 (double2)
constexpr int kFftPadA = 72;                 // row stride of transpose 1 (double2)
constexpr int kFftLdsElems = 8 * kFftPadA;   // 576 double2 = 9216 bytes

// Ordering around the transposes of a transform tile.  WAVE = true: the tile
// belongs to this wave alone (wave_lds_sync: no wait, a transpose's reads
// queue right behind its writes); WAVE = false: a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void tile_sync()
{
    if (WAVE)
        wave_lds_sync();
    else
        __syncthreads();
}

// Forward transform: v[j] = z[L + 64 j] on entry, Z in permuted layout on exit.
// TWIST = true: twisted negacyclic map above (w = exp(i pi/1024)).
// TWIST = false: plain DFT of size 512, Z[m] = sum_n z_n exp(2 pi i n m/512)
//                (tw.wl must then be 1).
// WAVE: the tile is private to the wave (see tile_sync).
template <bool TWIST = true, bool WAVE = false>
__device__ __forceinline__ void fft512_fwd(cd v[8], cd *lds, int lane, const Twiddles1024 &tw)
{
    // twist constants, pass 1
    if (TWIST) {
#pragma unroll
        for (int j = 1; j < 8; j++) v[j] = cmul(v[j], twist_const(j));
    }
    dft8<1>(v);
    {
        cd t = tw.wl;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (TWIST || k > 0) v[k] = cmul(v[k], t);
            if (k < 7) t = (TWIST || k > 0) ? cmul(t, tw.w4l) : tw.w4l;
        }
    }
    // transpose 1: T[k1][L] -> lane (L0 + 8 k1) register L1
#pragma unroll
    for (int k = 0; k < 8; k++) lds[k * kFftPadA + lane] = v[k];
    tile_sync<WAVE>();
    {
        const int L0 = lane & 7, k1 = lane >> 3;
#pragma unroll
        for (int l1 = 0; l1 < 8; l1++) v[l1] = lds[k1 * kFftPadA + L0 + 8 * l1];
    }
    tile_sync<WAVE>();
    dft8<1>(v);
    {
        cd t = tw.w64;
        cd p = t;
#pragma unroll
        for (int k = 1; k < 8; k++) {
            v[k] = cmul(v[k], p);
            if (k < 7) p = cmul(p, t);
        }
    }
    // transpose 2: T[p][lane] -> lane (p + 8 k1) register L0
#pragma unroll
    for (int k = 0; k < 8; k++) lds[k * kFftPad + lane] = v[k];
    tile_sync<WAVE>();
    {
        const int p = lane & 7, k1 = lane >> 3;
#pragma unroll
        for (int l0 = 0; l0 < 8; l0++) v[l0] = lds[p * kFftPad + l0 + 8 * k1];
    }
    tile_sync<WAVE>();
    dft8<1>(v);
}

// Inverse transform (unnormalised: returns 512 * z): permuted Z in, natural z out.
template <bool TWIST = true, bool WAVE = false>
__device__ __forceinline__ void fft512_inv(cd v[8], cd *lds, int lane, const Twiddles1024 &tw)
{
    dft8<-1>(v);
    {
        const int p = lane & 7, k1 = lane >> 3;
#pragma unroll
        for (int l0 = 0; l0 < 8; l0++) lds[p * kFftPad + l0 + 8 * k1] = v[l0];
    }
    tile_sync<WAVE>();
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = lds[k * kFftPad + lane];
    tile_sync<WAVE>();
    {
        cd t = tw.w64;
        cd p = t;
#pragma unroll
        for (int k = 1; k < 8; k++) {
            v[k] = cmulc(v[k], p);
            if (k < 7) p = cmul(p, t);
        }
    }
    dft8<-1>(v);
    {
        const int L0 = lane & 7, k1 = lane >> 3;
#pragma unroll
        for (int l1 = 0; l1 < 8; l1++) lds[k1 * kFftPadA + L0 + 8 * l1] = v[l1];
    }
    tile_sync<WAVE>();
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = lds[k * kFftPadA + lane];
    tile_sync<WAVE>();
    {
        cd t = tw.wl;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (TWIST || k > 0) v[k] = cmulc(v[k], t);
            if (k < 7) t = (TWIST || k > 0) ? cmul(t, tw.w4l) : tw.w4l;
        }
    }
    dft8<-1>(v);
    if (TWIST) {
#pragma unroll
        for (int j = 1; j < 8; j++) v[j] = cmulc(v[j], twist_const(j));
    }
}

// Two inverse transforms (twisted, as fft512_inv) through one wave-private
// tile, interleaved: each sequence's transpose is written while the other
// one's arithmetic runs, and the second sequence's writes follow the first
// one's reads of the same tile words with no wait (LDS operations of a wave
// are performed in order), so the round trips overlap arithmetic of the
// same wave.  The steps are fft512_inv's, written out again: built from
// helper functions shared with fft512_inv instead, the same instructions
// came out in another order and the blind rotation spilled 12 bytes per
// lane.
__device__ __forceinline__ void fft512_inv2(cd a[8], cd b[8], cd *lds, int lane,
                                            const Twiddles1024 &tw)
{
    const int p = lane & 7, k1 = lane >> 3;
    auto put2 = [&](const cd v[8]) {
#pragma unroll
        for (int l0 = 0; l0 < 8; l0++) lds[p * kFftPad + l0 + 8 * k1] = v[l0];
    };
    auto get2 = [&](cd v[8]) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = lds[k * kFftPad + lane];
    };
    auto put1 = [&](const cd v[8]) {
#pragma unroll
        for (int l1 = 0; l1 < 8; l1++) lds[k1 * kFftPadA + p + 8 * l1] = v[l1];
    };
    auto get1 = [&](cd v[8]) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = lds[k * kFftPadA + lane];
    };
    auto mid = [&](cd v[8]) {
        cd t = tw.w64;
        cd q = t;
#pragma unroll
        for (int k = 1; k < 8; k++) {
            v[k] = cmulc(v[k], q);
            if (k < 7) q = cmul(q, t);
        }
        dft8<-1>(v);
    };
    auto last = [&](cd v[8]) {
        cd t = tw.wl;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k] = cmulc(v[k], t);
            if (k < 7) t = cmul(t, tw.w4l);
        }
        dft8<-1>(v);
#pragma unroll
        for (int j = 1; j < 8; j++) v[j] = cmulc(v[j], twist_const(j));
    };
    dft8<-1>(a);
    put2(a);
    dft8<-1>(b);
    wave_lds_sync();
    get2(a);
    wave_lds_sync();
    put2(b);
    mid(a);
    wave_lds_sync();
    get2(b);
    wave_lds_sync();
    put1(a);
    mid(b);
    wave_lds_sync();
    get1(a);
    wave_lds_sync();
    put1(b);
    last(a);
    wave_lds_sync();
    get1(b);
    wave_lds_sync();
    last(b);
}

constexpr int kLaneTile = 320;  // double2 per wave tile of lanes64_fwd / lanes64_inv

// v[q] *= base^q (CONJ: conj(base)^q), q = 1..3
template <bool CONJ>
__device__ __forceinline__ void twiddle4(cd v[4], cd base)
{
    const cd b2 = cmul(base, base);
    const cd b3 = cmul(b2, base);
    v[1] = CONJ ? cmulc(v[1], base) : cmul(v[1], base);
    v[2] = CONJ ? cmulc(v[2], b2) : cmul(v[2], b2);
    v[3] = CONJ ? cmulc(v[3], b3) : cmul(v[3], b3);
}

// Exchange of one 32-bit word between two registers across lane groups:
// permlane32_swap trades lanes 32-63 of a with lanes 0-31 of b,
// permlane16_swap the odd 16-lane rows of a with the even rows of b
// (gfx950 v_permlane32_swap / v_permlane16_swap; no LDS traffic).
template <int ROWS>
__device__ __forceinline__ void swap_rows(double &a, double &b)
{
    const uint64_t ua = uint64_t(__double_as_longlong(a)), ub = uint64_t(__double_as_longlong(b));
    const uint32_t al = uint32_t(ua), ah = uint32_t(ua >> 32);
    const uint32_t bl = uint32_t(ub), bh = uint32_t(ub >> 32);
    const auto lo = ROWS == 32 ? __builtin_amdgcn_permlane32_swap(al, bl, false, false)
                               : __builtin_amdgcn_permlane16_swap(al, bl, false, false);
    const auto hi = ROWS == 32 ? __builtin_amdgcn_permlane32_swap(ah, bh, false, false)
                               : __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
    a = __longlong_as_double((long long)((uint64_t(hi[0]) << 32) | lo[0]));
    b = __longlong_as_double((long long)((uint64_t(hi[1]) << 32) | lo[1]));
}

// 4 x 4 transpose of the register index with lane bits 4-5: afterwards
// register r, lane (l & 15) + 16 m holds what register m, lane
// (l & 15) + 16 r held.  Register bit 1 trades with lane bit 5 (32-lane
// swap between registers r and r + 2), register bit 0 with lane bit 4
// (16-lane swap between registers r and r + 1).  Its own inverse.
__device__ __forceinline__ void transpose_reg_lane_hi(cd v[4])
{
#pragma unroll
    for (int r = 0; r < 2; r++) {
        swap_rows<32>(v[r].x, v[r + 2].x);
        swap_rows<32>(v[r].y, v[r + 2].y);
    }
#pragma unroll
    for (int r = 0; r < 4; r += 2) {
        swap_rows<16>(v[r].x, v[r + 1].x);
        swap_rows<16>(v[r].y, v[r + 1].y);
    }
}

// The same transpose through the wave's LDS tile (rows of 64 double2).
__device__ __forceinline__ void transpose_reg_lane_hi_lds(cd v[4], cd *tile, int l)
{
#pragma unroll
    for (int m1 = 0; m1 < 4; m1++) tile[m1 * 64 + l] = v[m1];
    wave_lds_sync();
#pragma unroll
    for (int L1 = 0; L1 < 4; L1++) v[L1] = tile[(l >> 4) * 64 + (l & 15) + 16 * L1];
    wave_lds_sync();
}

// 64-point DFT over the lanes of a wave for each of the four register
// sequences, as three radix-4 passes: the first exchange is the register /
// lane-bits-4-5 transpose (PERM: lane swaps; otherwise through the tile),
// the other two go through a wave-private LDS tile of kLaneTile double2
// (row strides 20 and 5, conflict-free for 16-lane b128 accesses).  The
// swaps move VALU work in for LDS traffic out, which pays on short launches
// and costs on long ones: the kernels that use this transform instantiate
// both forms and pick one by batch size (both move the same words, so the
// results are identical).  t64 = W64^(lane & 15),
// t16 = W16^(lane & 3).  Input: register r, lane L holds x_r[L]; output:
// register q, lane l holds X_r[m] for r = l >> 4,
// m = ((l >> 2) & 3) + 4 (l & 3) + 16 q.
template <bool PERM>
__device__ __forceinline__ void lanes64_fwd(cd v[4], cd *tile, int l, cd t64, cd t16)
{
    // T1: (lane L, reg m1) -> lane (L0 = l & 15, m1 = l >> 4), reg L1
    if constexpr (PERM)
        transpose_reg_lane_hi(v);
    else
        transpose_reg_lane_hi_lds(v, tile, l);
    dft4<1>(v[0], v[1], v[2], v[3]);
    twiddle4<false>(v, t64);
    // T2: (lane (L0, m1), reg q1) -> lane (L00 = l & 3, rest = l >> 2 = 4 m1 + q1), reg L01
#pragma unroll
    for (int q1 = 0; q1 < 4; q1++) tile[((l >> 4) * 4 + q1) * 20 + (l & 15)] = v[q1];
    wave_lds_sync();
#pragma unroll
    for (int L01 = 0; L01 < 4; L01++) v[L01] = tile[(l >> 2) * 20 + (l & 3) + 4 * L01];
    wave_lds_sync();
    dft4<1>(v[0], v[1], v[2], v[3]);
    twiddle4<false>(v, t16);
    // T3: (lane (L00, rest), reg q21) -> lane 4 rest + q21, reg L00
#pragma unroll
    for (int q21 = 0; q21 < 4; q21++) tile[((l >> 2) * 4 + q21) * 5 + (l & 3)] = v[q21];
    wave_lds_sync();
#pragma unroll
    for (int L00 = 0; L00 < 4; L00++) v[L00] = tile[l * 5 + L00];
    wave_lds_sync();
    dft4<1>(v[0], v[1], v[2], v[3]);
}

template <bool PERM>
__device__ __forceinline__ void lanes64_inv(cd v[4], cd *tile, int l, cd t64, cd t16)
{
    dft4<-1>(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int L00 = 0; L00 < 4; L00++) tile[l * 5 + L00] = v[L00];
    wave_lds_sync();
#pragma unroll
    for (int q21 = 0; q21 < 4; q21++) v[q21] = tile[((l >> 2) * 4 + q21) * 5 + (l & 3)];
    wave_lds_sync();
    twiddle4<true>(v, t16);
    dft4<-1>(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int L01 = 0; L01 < 4; L01++) tile[(l >> 2) * 20 + (l & 3) + 4 * L01] = v[L01];
    wave_lds_sync();
#pragma unroll
    for (int q1 = 0; q1 < 4; q1++) v[q1] = tile[((l >> 4) * 4 + q1) * 20 + (l & 15)];
    wave_lds_sync();
    twiddle4<true>(v, t64);
    dft4<-1>(v[0], v[1], v[2], v[3]);
    if constexpr (PERM)
        transpose_reg_lane_hi(v);
    else
        transpose_reg_lane_hi_lds(v, tile, l);
}

// frequency index held by (lane s, register q) in the permuted layout
__host__ __device__ constexpr int fd_index1024(int s, int q) { return (s >> 3) + 8 * (s & 7) + 64 * q; }

// bit reversal over 9 bits: spqlios order of the 512 FD points
__host__ __device__ constexpr int rev9(int x)
{
    int r = 0;
    for (int i = 0; i < 9; i++) r |= ((x >> i) & 1) << (8 - i);
    return r;
}

// uint32_t(int64_t(x)) for |x| < 2^51 (TwistFFT truncation,
// fft_processor_spqlios.cpp:121), computed exactly in FP64: t = trunc(x) is
// an integer of magnitude < 2^51, so t + 1.5*2^52 is exact and its low 32
// mantissa bits are t mod 2^32 (two's complement for negative t).
__device__ __forceinline__ uint32_t trunc_torus32(double x)
{
    const double u = trunc(x) + 0x1.8p52;
    return uint32_t(__double_as_longlong(u));
}

}  // namespace fhe
