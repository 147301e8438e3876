// pbs_lvl2q.hip -- lvl2 (N = 2048, 64-bit torus, l = 4, Bgbit = 9) blind
// rotation with four waves per ciphertext.
//
// Reference: thirdparty/TFHEpp/include/gatebootstrapping.hpp:15-43 with
// lvl02param, src/detwfa.cpp:21-30, src/trgsw.cpp:9-103; GPU counterpart
// src/cuHEDB/fft_gpu/BlindRotate_gpu.cuh:534-670.  Laid out for occupancy:
// each wave holds 4 of the 1024 complex points per lane (a two-wave layout
// with 8 per lane needs ~320 registers, one wave per SIMD), so a wave needs
// about half the registers and three waves fit per SIMD (the
// accumulator, 2 x 2048 x 64-bit = 32 KB, stays one copy per ciphertext in
// LDS).  The twisted transform Z[m] = sum_n z_n w2^{n(4m+1)},
// w2 = exp(i pi/2048), factors as
//   step A  (registers, then one exchange between the four waves):
//           u_k[n0] = w2^{n0(4k+1)} sum_p c_p z[n0 + 256 p] i^{pk},
//           c_p = w2^{256 p},  n0 < 256,  k < 4;
//   step C  (wave k): the plain 256-point DFT of u_k, Z[4 m' + k]:
//           a radix-4 pass over the registers, then the 64-point DFT over
//           the lanes as three radix-4 passes: the first exchange by
//           cross-lane register swaps (v_permlane32_swap /
//           v_permlane16_swap); the other two by wave-local LDS transposes
//           (row strides 20 and 5 double2, conflict-free for gfx950's
//           16-lanes-per-cycle b128 accesses).
// Layouts: time domain wave w, lane L, register p holds z[L + 64 w + 256 p];
// frequency domain wave k, lane l, register q holds Z[4 m' + k] with
// m' = (l >> 4) + 4 ((l >> 2) & 3) + 16 (l & 3) + 64 q.  The inverse runs
// the same steps mirrored with conjugate twiddles (unnormalised, x 1024;
// the key is stored pre-scaled by 1/1024, exactly).
// The CMUX ACC += ExtProd((X^a - 1) ACC, BK_i) is computed as
// ACC += (X^a - 1) ExtProd(ACC, BK_i): the gadget digits are those of the
// accumulator itself (each lane decomposes its own words, no rotated LDS
// reads), and X^a - 1 multiplies the product sum in the frequency domain by
// w2^(a (4f+1)) - 1 before the inverse transforms.  Exact arithmetic gives
// the same result; in floating point the decomposition rounding and the
// key noise are multiplied by X^a - 1 instead of entering after it, which
// raises the output phase noise by sqrt(2) (2^-28.4 -> 2^-27.8 of the
// torus on the 64-bit torus, far below every decision margin) for 8 %
// less time per bootstrap.
#include "fft64.hpp"
#include "launch.hpp"
#include "params.hpp"

namespace fhe {

namespace {

constexpr int N2q = Lvl2::N;
constexpr uint64_t kDecOff2q = decomp_total_offset<uint64_t, 64, Lvl2::l, Lvl2::Bgbit>();
constexpr int kTileQ = kLaneTile;    // double2 per wave tile
constexpr double kKeyScale2 = 1.0 / 1024;

// per-lane twiddle bases
struct TwQ {
    cd beta, rho;  // step A: w2^{n0}, w2^{4 n0}, n0 = lane + 64 wave
    cd t256;       // W256^{lane}
    cd t64;        // W64^{lane & 15}
    cd t16;        // W16^{lane & 3}
};

__device__ __forceinline__ TwQ make_twq(int lane, int wave)
{
    TwQ t;
    const int n0 = lane + 64 * wave;
    t.beta = unit_root(double(n0), 2048.0);
    t.rho = unit_root(double(n0), 512.0);
    t.t256 = unit_root(double(lane), 128.0);
    t.t64 = unit_root(double(lane & 15), 32.0);
    t.t16 = unit_root(double(lane & 3), 8.0);
    return t;
}

// c_p = exp(i pi p / 8)
__device__ __forceinline__ cd cq(int p)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
    return p == 1 ? cd{c1, s1} : p == 2 ? cd{kR2, kR2} : cd{s1, c1};
}

// Forward: v[p] = z[lane + 64 wave + 256 p] in, Z[4 m' + wave] (layout above)
// out.  xch: the workgroup's 4 x 256 double2 exchange area, which aliases the
// four wave tiles (tiles = xch + wave * kTileQ).
template <bool PERM = false>
__device__ __forceinline__ void fft1024q_fwd(cd v[4], cd *xch, int lane, int wave, const TwQ &tw)
{
    v[1] = cmul(v[1], cq(1));
    v[2] = cmul(v[2], cq(2));
    v[3] = cmul(v[3], cq(3));
    dft4<1>(v[0], v[1], v[2], v[3]);
    {
        cd g = tw.beta;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cmul(v[k], g);
            if (k < 3) g = cmul(g, tw.rho);
        }
    }
    const int n0 = lane + 64 * wave;
    // the exchange area overlaps the wave tiles: every wave has left its tile
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) xch[k * kTileQ + n0] = v[k];
    __syncthreads();
    // wave k's sequence sits at the start of wave k's own tile, so the wave
    // may go on into its tile without waiting for the others' reads
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = xch[wave * kTileQ + lane + 64 * j];
    // step C: radix-4 over the registers (stride 64), twiddle W256^{L m1}
    dft4<1>(v[0], v[1], v[2], v[3]);
    twiddle4<false>(v, tw.t256);
    lanes64_fwd<PERM>(v, xch + wave * kTileQ, lane, tw.t64, tw.t16);
}

// TAIL_SYNC: wait for every wave's exchange reads before returning (the
// caller's next use of its own tile would overwrite words the others read);
// not needed when the next write into the exchange area is a forward
// transform's, which waits for all waves first.
template <bool PERM = false, bool TAIL_SYNC = true>
__device__ __forceinline__ void fft1024q_inv(cd v[4], cd *xch, int lane, int wave, const TwQ &tw)
{
    lanes64_inv<PERM>(v, xch + wave * kTileQ, lane, tw.t64, tw.t16);
    twiddle4<true>(v, tw.t256);
    dft4<-1>(v[0], v[1], v[2], v[3]);
    // the wave writes its own tile, which no other wave reads or writes
    // between the last barrier (a forward exchange's second, or the previous
    // inverse's tail) and the reads below: no barrier before the write
#pragma unroll
    for (int j = 0; j < 4; j++) xch[wave * kTileQ + lane + 64 * j] = v[j];
    __syncthreads();
    const int n0 = lane + 64 * wave;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = xch[k * kTileQ + n0];
    if (TAIL_SYNC) __syncthreads();
    {
        cd g = tw.beta;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cmulc(v[k], g);
            if (k < 3) g = cmul(g, tw.rho);
        }
    }
    dft4<-1>(v[0], v[1], v[2], v[3]);
    v[1] = cmulc(v[1], cq(1));
    v[2] = cmulc(v[2], cq(2));
    v[3] = cmulc(v[3], cq(3));
}

// natural frequency of (wave, lane, register q) in the output layout
__host__ __device__ constexpr int q_freq(int wave, int l, int q)
{
    return 4 * ((l >> 4) + 4 * ((l >> 2) & 3) + 16 * (l & 3) + 64 * q) + wave;
}

// trunc(x) mod 2^64 (execute_direct_torus64's truncation; |x| < 1 gives 0),
// exact for |x| < 2^83: t = trunc(x) = H 2^32 + L with L in [0, 2^32) (exact
// in FP64), and H mod 2^32 is the low mantissa word of H + 1.5 2^52 (exact
// for |H| < 2^51).  The inverse transform's outputs are sums of 2 x 4 x 2048
// products of a 9-bit digit and a 64-bit key word: RMS ~2^77, so 2^83 is
// ~64 standard deviations away.
__device__ __forceinline__ uint64_t trunc_torus64q(double x)
{
    const double t = trunc(x);
    const double H = floor(t * 0x1p-32);
    const double L = __fma_rn(-H, 0x1p32, t);
    const uint32_t Hm = uint32_t(__double_as_longlong(H + 0x1.8p52));
    return (uint64_t(Hm) << 32) | uint64_t(uint32_t(L));
}

// x * i^k
__device__ __forceinline__ cd mul_ipow2q(cd x, uint32_t k)
{
    switch (k & 3) {
    case 0: return x;
    case 1: return cd{-x.y, x.x};
    case 2: return cd{-x.x, -x.y};
    default: return cd{x.y, -x.x};
    }
}

// rot[e] = w2^e = exp(i pi e / 2048), e < 2N
__global__ __launch_bounds__(256) void k_rot_table2q(cd *__restrict__ rot)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < 2 * N2q) rot[e] = unit_root(double(e), 2048.0);
}

// Gadget digits of x (DecompositionPolynomial, trgsw.cpp:21-39):
//   digit_d = ((x + offset) >> (64 - 9(d+1)) & 511) - 256.
// prep2q adds the offset and flips the top bit of every 9-bit field, so each
// digit is one signed 9-bit field extract of the prepared word: (f ^ 256)
// read as a signed 9-bit number equals f - 256.
constexpr uint64_t kDigitFlip2q = (1ull << 63) | (1ull << 54) | (1ull << 45) | (1ull << 36);
__device__ __forceinline__ uint64_t prep2q(uint64_t x) { return (x + kDecOff2q) ^ kDigitFlip2q; }
__device__ __forceinline__ uint64_t unprep2q(uint64_t u) { return (u ^ kDigitFlip2q) - kDecOff2q; }
__device__ __forceinline__ double digit2qp(uint64_t u, int d)
{
    const int sh = 64 - Lvl2::Bgbit * (d + 1);
    return double(int32_t(__builtin_amdgcn_sbfe(int32_t(uint32_t(u >> sh)), 0, Lvl2::Bgbit)));
}
// Digits 0 to l - 2 lie in the high word (bits 37 and up), so for them only
// the high halves of the accumulator words are read from LDS (32-bit reads,
// one extract each); A/B in one call, 264.3 / 264.6 against 268.3 / 267.3 ms
// per 16384 bootstraps.
static_assert(64 - Lvl2::Bgbit * (Lvl2::l - 1) >= 32, "digits 0 to l - 2 in the high word");
__device__ __forceinline__ double digit2qh(uint64_t u, int d)
{
    const int sh = 32 - Lvl2::Bgbit * (d + 1);
    return double(int32_t(__builtin_amdgcn_sbfe(int32_t(uint32_t(u >> 32)), sh, Lvl2::Bgbit)));
}

constexpr int kQRow = 4 * 4 * 64;  // double2 per (i, row, column): [wave][q][lane]

// The first lane transpose runs as register swaps (the transforms' PERM
// form), which trades LDS traffic for VALU work: A/B in one call, 135.7
// against 144.2 ms per 8192 bootstraps, 268.6 against 289.3 per 16384 and
// 341.0 against 360.4 per 20480 (an earlier build of this kernel, with more
// barriers and the accumulator decomposed per row, had found it faster only
// up to 4096).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_blind_rotate2q(
    const uint32_t *__restrict__ in, int n0, const cd *__restrict__ bk,
    const cd *__restrict__ rot, const uint64_t *__restrict__ tv, int tv_stride, int tv_period,
    uint64_t *__restrict__ out, int sample_extract, int B)
{
    __shared__ uint64_t acc[2 * N2q];
    __shared__ cd xch[4 * kTileQ];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const uint32_t *c0 = in + size_t(ct) * (n0 + 1);
    const uint64_t *tvp = tv + size_t(tv_period > 0 ? ct % tv_period : ct) * tv_stride;
    const TwQ tw = make_twq(lane, wave);

    // the accumulator is kept prepared for digit extraction: LDS holds
    // prep2q(x) = (x + offset) ^ flip, so the l digit rows read their digits
    // straight from it and only the update pays (one xor before and after
    // the addition; the flip touches the high word only)
    const uint32_t bbar = 2u * N2q - (c0[n0] >> (32 - 1 - Lvl2::nbit));
    for (int n = tid; n < N2q; n += 256) {
        acc[n] = prep2q(0);
        uint32_t idx = (uint32_t(n) - bbar) & (2u * N2q - 1);
        uint64_t v = tvp[idx & (N2q - 1)];
        acc[N2q + n] = prep2q((idx & N2q) ? 0ull - v : v);
    }
    __syncthreads();

    constexpr uint32_t kRound = 1u << (32 - 2 - Lvl2::nbit);
    // exponent base of this lane's points: 4 q_freq(wave, lane, q) + 1 = e0 + 1024 q
    const uint32_t e0 = 4u * uint32_t(q_freq(wave, lane, 0)) + 1u;
    for (int i = 0; i < n0; i++) {
        const uint32_t abar = (c0[i] + kRound) >> (32 - 1 - Lvl2::nbit);
        if (abar == 0) continue;
        const cd *bki = bk + size_t(i) * Lvl2::rows * 2 * kQRow + wave * 256 + lane;
        const cd base = rot[(abar * e0) & (2u * N2q - 1)];  // used after the rows
        cd fd0[4], fd1[4];
#pragma unroll
        for (int q = 0; q < 4; q++) fd0[q] = fd1[q] = cd{0.0, 0.0};
#pragma unroll 1
        for (int r = 0; r < Lvl2::rows; r++) {
            const int c = r / Lvl2::l, d = r - c * Lvl2::l;
            const uint64_t *poly = acc + c * N2q;
            cd b0[4], b1[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                b0[q] = bki[(r * 2 + 0) * kQRow + q * 64];
                b1[q] = bki[(r * 2 + 1) * kQRow + q * 64];
            }
            // digits of this lane's own accumulator words
            cd v[4];
            if (d < Lvl2::l - 1) {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int n = lane + 64 * wave + 256 * p;
                    v[p] = cd{digit2qh(poly[n], d), digit2qh(poly[n + 1024], d)};
                }
            } else {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int n = lane + 64 * wave + 256 * p;
                    v[p] = cd{digit2qp(poly[n], d), digit2qp(poly[n + 1024], d)};
                }
            }
            fft1024q_fwd<true>(v, xch, lane, wave, tw);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cmac(fd0[q], v[q], b0[q]);
                cmac(fd1[q], v[q], b1[q]);
            }
        }
        // (X^abar - 1) in the frequency domain: the evaluation at w2^(4f+1)
        // is multiplied by w2^(abar (4f+1)) - 1, and for this lane's points
        // w2^(abar (e0 + 1024 q)) = w2^(abar e0) i^(abar q)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            cd g = mul_ipow2q(base, abar * uint32_t(q));
            g.x -= 1.0;
            fd0[q] = cmul(fd0[q], g);
            fd1[q] = cmul(fd1[q], g);
        }
        // each lane reads and updates only its own accumulator words
        fft1024q_inv<true>(fd0, xch, lane, wave, tw);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int n = lane + 64 * wave + 256 * p;
            acc[n] = prep2q(unprep2q(acc[n]) + trunc_torus64q(fd0[p].x));
            acc[n + 1024] = prep2q(unprep2q(acc[n + 1024]) + trunc_torus64q(fd0[p].y));
        }
        // the next step's first forward exchange waits for all waves before
        // writing into the others' tiles
        fft1024q_inv<true, false>(fd1, xch, lane, wave, tw);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int n = lane + 64 * wave + 256 * p;
            acc[N2q + n] = prep2q(unprep2q(acc[N2q + n]) + trunc_torus64q(fd1[p].x));
            acc[N2q + n + 1024] = prep2q(unprep2q(acc[N2q + n + 1024]) + trunc_torus64q(fd1[p].y));
        }
        // no barrier here: each lane updated only its own accumulator words,
        // and the next forward transform's first barrier follows every wave's
        // exchange reads above
    }
    __syncthreads();  // sample extraction reads other waves' words

    if (sample_extract) {
        uint64_t *o = out + size_t(ct) * (N2q + 1);
        for (int n = tid; n < N2q; n += 256)
            o[n] = n == 0 ? unprep2q(acc[0]) : 0ull - unprep2q(acc[N2q - n]);
        if (tid == 0) o[N2q] = unprep2q(acc[N2q]);
    } else {
        uint64_t *o = out + size_t(ct) * 2 * N2q;
        for (int n = tid; n < 2 * N2q; n += 256) o[n] = unprep2q(acc[n]);
    }
}

// integer torus polynomial -> key layout [wave][q][lane], scaled by 1/1024
__global__ __launch_bounds__(256) void k_poly_to_devfd2q(const uint64_t *__restrict__ poly,
                                                          cd *__restrict__ fd, int npoly)
{
    __shared__ cd xch[4 * kTileQ];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = blockIdx.x;
    if (p >= npoly) return;
    const uint64_t *a = poly + size_t(p) * N2q;
    const TwQ tw = make_twq(lane, wave);
    cd v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int n = lane + 64 * wave + 256 * j;
        v[j] = cd{double(int64_t(a[n])), double(int64_t(a[n + 1024]))};
    }
    fft1024q_fwd(v, xch, lane, wave, tw);
#pragma unroll
    for (int q = 0; q < 4; q++)
        fd[size_t(p) * kQRow + wave * 256 + q * 64 + lane] =
            cd{v[q].x * kKeyScale2, v[q].y * kKeyScale2};
}

// the same transform, spqlios FD order out (parity hook for the layout)
__global__ __launch_bounds__(256) void k_ifft2q(const uint64_t *__restrict__ poly,
                                                 double *__restrict__ fd, int B)
{
    __shared__ cd xch[4 * kTileQ];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = blockIdx.x;
    if (p >= B) return;
    const uint64_t *a = poly + size_t(p) * N2q;
    const TwQ tw = make_twq(lane, wave);
    cd v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int n = lane + 64 * wave + 256 * j;
        v[j] = cd{double(int64_t(a[n])), double(int64_t(a[n + 1024]))};
    }
    fft1024q_fwd(v, xch, lane, wave, tw);
    double *o = fd + size_t(p) * N2q;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int m = q_freq(wave, lane, q), r = 0;
        for (int b = 0; b < 10; b++) r |= ((m >> b) & 1) << (9 - b);
        o[r] = v[q].x;
        o[r + 1024] = v[q].y;
    }
}

}  // namespace

size_t bk2q_device_elems(int n0) { return size_t(n0) * Lvl2::rows * 2 * kQRow; }

hipError_t launch_bk2q_from_int(const uint64_t *bk_int, cd *bk_dev, int n0, hipStream_t s)
{
    const int npoly = n0 * Lvl2::rows * 2;
    hipLaunchKernelGGL(k_poly_to_devfd2q, dim3(npoly), dim3(256), 0, s, bk_int, bk_dev, npoly);
    return hipGetLastError();
}

hipError_t launch_blind_rotate2q(const uint32_t *in, int n0, const cd *bk, const cd *rot,
                                 const uint64_t *tv, int tv_stride, uint64_t *out,
                                 bool sample_extract, int B, hipStream_t s, int tv_period)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_blind_rotate2q, dim3(B), dim3(256), 0, s, in, n0, bk, rot, tv, tv_stride,
                       tv_period, out, int(sample_extract), B);
    return hipGetLastError();
}

size_t rot2q_elems() { return 2 * N2q; }

hipError_t launch_rot2q_table(cd *rot, hipStream_t s)
{
    hipLaunchKernelGGL(k_rot_table2q, dim3((2 * N2q + 255) / 256), dim3(256), 0, s, rot);
    return hipGetLastError();
}

hipError_t launch_ifft2q(const uint64_t *poly, double *fd, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ifft2q, dim3(B), dim3(256), 0, s, poly, fd, B);
    return hipGetLastError();
}

}  // namespace fhe
