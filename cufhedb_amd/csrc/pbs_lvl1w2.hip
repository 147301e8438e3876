// pbs_lvl1w2.hip -- lvl1 blind rotation with two waves per ciphertext, for
// batches too small to fill the GPU with one wave per ciphertext.
//
// Same computation as k_blind_rotate1 (pbs_lvl1.hip; TFHEpp BlindRotate,
// gatebootstrapping.hpp:15-43, CMUX detwfa.cpp:21-30, external product
// trgsw.cpp:78-103), split over two waves: each wave owns 256 of the 512
// complex points of every polynomial (4 per lane), so a ciphertext's CMUX
// step runs on two SIMDs and a small batch finishes in about half the time.
// The accumulator stays in LDS (the rotation reads any coefficient); each
// wave rotates and decomposes its own coefficients, and the transform
// (512 = 4 x 128) exchanges once between the waves:
//   step A (registers): u_k[n0] = w^(n0 (4k+1)) sum_p c_p z[n0 + 128 p] i^(pk),
//           c_p = exp(i pi p / 8), n0 = lane + 64 wave < 128, k < 4;
//   exchange: wave v takes u_k for k = 2v, 2v+1 (all 128 n0);
//   step C: per sequence, radix 2 over n0 = L + 64 h, twiddle W128^(L m1),
//           then the 64-point DFT over the lanes (lanes64_fwd).
// Time domain: wave w, lane L, register p holds z[L + 64 w + 128 p],
// z_n = a_n + i a_(n+512).  Frequency domain: wave w, lane l, register q
// holds Z[f1w(w, l, q)].  The key is the one-wave kernel's device key with
// each (i, row, col) slice permuted into this order (same FD values).
#include "fft64.hpp"
#include "launch.hpp"
#include "params.hpp"

namespace fhe {

namespace {

constexpr int N1 = Lvl1::N;
constexpr int kPolyW = 512;  // complex points per polynomial
constexpr uint32_t kDecOffW = decomp_total_offset<uint32_t, 32, Lvl1::l, Lvl1::Bgbit>();
constexpr uint32_t kDigitFlipW = (1u << 31) | (1u << 25) | (1u << 19);

// natural frequency of (wave, lane, register) in the output layout
__host__ __device__ constexpr int f1w(int w, int l, int q)
{
    return 4 * (((l >> 4) & 1) + 2 * (((l >> 2) & 3) + 4 * (l & 3) + 16 * q)) + 2 * w + (l >> 5);
}

struct TwW {
    cd beta, rho;  // w^n0, w^(4 n0), n0 = lane + 64 wave, w = exp(i pi / 1024)
    cd t128;       // W128^lane
    cd t64, t16;   // lanes64 twiddles
};

__device__ __forceinline__ TwW make_tww(int lane, int wave)
{
    TwW t;
    const int n0 = lane + 64 * wave;
    t.beta = unit_root(double(n0), 1024.0);
    t.rho = unit_root(double(n0), 256.0);
    t.t128 = unit_root(double(lane), 64.0);
    t.t64 = unit_root(double(lane & 15), 32.0);
    t.t16 = unit_root(double(lane & 3), 8.0);
    return t;
}

// exp(i pi p / 8), p = 1..3
__device__ __forceinline__ cd c8w(int p)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
    return p == 1 ? cd{c1, s1} : p == 2 ? cd{kR2, kR2} : cd{s1, c1};
}

// Forward transform.  xch: the workgroup's 2 x kLaneTile double2, the two
// wave tiles; the exchange puts wave v's two sequences (k = 2v, 2v + 1) at
// the start of v's own tile, so after the exchange a wave touches only its
// own tile until the next exchange.
template <bool PERM>
__device__ __forceinline__ void fft512w_fwd(cd v[4], cd *xch, int lane, int wave, const TwW &tw)
{
    v[1] = cmul(v[1], c8w(1));
    v[2] = cmul(v[2], c8w(2));
    v[3] = cmul(v[3], c8w(3));
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
    __syncthreads();  // both waves have left their tiles
#pragma unroll
    for (int k = 0; k < 4; k++) xch[(k >> 1) * kLaneTile + (k & 1) * 128 + n0] = v[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = xch[wave * kLaneTile + (j >> 1) * 128 + lane + 64 * (j & 1)];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const cd a = v[2 * s], b = v[2 * s + 1];
        v[2 * s] = cadd(a, b);
        v[2 * s + 1] = cmul(csub(a, b), tw.t128);
    }
    lanes64_fwd<PERM>(v, xch + wave * kLaneTile, lane, tw.t64, tw.t16);
}

// Inverse (unnormalised: returns 512 z): frequency layout in, time layout out.
// The wave writes its exchange words into its own tile, which the other wave
// does not touch between the last barrier (the forward exchange's second, or
// the previous inverse's tail) and the reads, so no barrier precedes the
// write.  TAIL_SYNC: wait for both waves' exchange reads before returning
// (the caller's next use of its tile would overwrite words the other wave
// reads); not needed when a workgroup barrier follows anyway.
template <bool PERM, bool TAIL_SYNC = true>
__device__ __forceinline__ void fft512w_inv(cd v[4], cd *xch, int lane, int wave, const TwW &tw)
{
    lanes64_inv<PERM>(v, xch + wave * kLaneTile, lane, tw.t64, tw.t16);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const cd a = v[2 * s], b = cmulc(v[2 * s + 1], tw.t128);
        v[2 * s] = cadd(a, b);
        v[2 * s + 1] = csub(a, b);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) xch[wave * kLaneTile + (j >> 1) * 128 + lane + 64 * (j & 1)] = v[j];
    __syncthreads();
    const int n0 = lane + 64 * wave;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = xch[(k >> 1) * kLaneTile + (k & 1) * 128 + n0];
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
    v[1] = cmulc(v[1], c8w(1));
    v[2] = cmulc(v[2], c8w(2));
    v[3] = cmulc(v[3], c8w(3));
}

// (X^a - 1) * poly at coefficient n, a in [0, 2N)
__device__ __forceinline__ uint32_t rot_minus_one_w(const uint32_t *poly, uint32_t a, int n)
{
    const uint32_t idx = (uint32_t(n) - a) & (2u * N1 - 1);
    uint32_t v = poly[idx & (N1 - 1)];
    if (idx & N1) v = 0u - v;
    return v - poly[n];
}

// gadget digits (trgsw.cpp:21-39) as one signed field extract of the word
// prepared with the offset and the flipped top bit of every 6-bit field
__device__ __forceinline__ uint32_t prep_w(uint32_t x) { return (x + kDecOffW) ^ kDigitFlipW; }
__device__ __forceinline__ double digit_w(uint32_t u, int d)
{
    const int sh = 32 - Lvl1::Bgbit * (d + 1);
    return double(int32_t(__builtin_amdgcn_sbfe(int32_t(u), sh, Lvl1::Bgbit)));
}

// three waves per SIMD (168 VGPRs, no spills)
template <bool PERM>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_blind_rotate1w2(
    const uint32_t *__restrict__ in, int n0, const cd *__restrict__ bk,
    const uint32_t *__restrict__ tv, int tv_stride, uint32_t *__restrict__ out, int sample_extract,
    int B)
{
    __shared__ uint32_t acc[2][N1];
    __shared__ cd xch[2 * kLaneTile];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const uint32_t *c0 = in + size_t(ct) * (n0 + 1);
    const uint32_t *tvp = tv + size_t(ct) * tv_stride;
    const TwW tw = make_tww(lane, wave);
    const int nb = lane + 64 * wave;

    // acc = (0, X^bbar tv), bbar = 2N - (b >> 21)   (no rounding)
    {
        const uint32_t bbar = 2u * N1 - (c0[n0] >> (32 - 1 - Lvl1::nbit));
#pragma unroll
        for (int k = 0; k < N1 / 128; k++) {
            const int n = tid + 128 * k;
            const uint32_t idx = (uint32_t(n) - bbar) & (2u * N1 - 1);
            const uint32_t v = tvp[idx & (N1 - 1)];
            acc[0][n] = 0;
            acc[1][n] = (idx & N1) ? 0u - v : v;
        }
    }
    __syncthreads();

    constexpr uint32_t kRound = 1u << (32 - 2 - Lvl1::nbit);
    for (int i = 0; i < n0; i++) {
        const uint32_t abar = (c0[i] + kRound) >> (32 - 1 - Lvl1::nbit);
        if (abar == 0) continue;
        const cd *bki = bk + size_t(i) * (Lvl1::rows * 2 * kPolyW) + wave * 256 + lane;
        cd fd0[4], fd1[4];
#pragma unroll
        for (int q = 0; q < 4; q++) fd0[q] = fd1[q] = cd{0.0, 0.0};
#pragma unroll 1
        for (int c = 0; c < 2; c++) {
            // this wave's coefficients of (X^abar - 1) acc_c, prepared for digits
            uint32_t u[8];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                u[2 * p] = prep_w(rot_minus_one_w(acc[c], abar, nb + 128 * p));
                u[2 * p + 1] = prep_w(rot_minus_one_w(acc[c], abar, nb + 128 * p + 512));
            }
#pragma unroll 1
            for (int d = 0; d < Lvl1::l; d++) {
                const int r = c * Lvl1::l + d;
                // key words requested before the transform
                cd b0[4], b1[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    b0[q] = bki[r * 2 * kPolyW + q * 64];
                    b1[q] = bki[r * 2 * kPolyW + kPolyW + q * 64];
                }
                cd v[4];
#pragma unroll
                for (int p = 0; p < 4; p++) v[p] = cd{digit_w(u[2 * p], d), digit_w(u[2 * p + 1], d)};
                fft512w_fwd<PERM>(v, xch, lane, wave, tw);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    cmac(fd0[q], v[q], b0[q]);
                    cmac(fd1[q], v[q], b1[q]);
                }
            }
        }
        // both waves' rotation reads of acc are complete (the transforms
        // synchronised them); each wave updates only its own coefficients
        fft512w_inv<PERM>(fd0, xch, lane, wave, tw);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            acc[0][nb + 128 * p] += trunc_torus32(fd0[p].x);
            acc[0][nb + 128 * p + 512] += trunc_torus32(fd0[p].y);
        }
        // the step's closing barrier follows this transform's exchange reads
        fft512w_inv<PERM, false>(fd1, xch, lane, wave, tw);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            acc[1][nb + 128 * p] += trunc_torus32(fd1[p].x);
            acc[1][nb + 128 * p + 512] += trunc_torus32(fd1[p].y);
        }
        __syncthreads();
    }

    if (sample_extract) {
        // SampleExtractIndex 0: out[0] = acc0[0], out[n] = -acc0[N - n], out[N] = acc1[0]
        uint32_t *o = out + size_t(ct) * (N1 + 1);
#pragma unroll
        for (int k = 0; k < N1 / 128; k++) {
            const int n = tid + 128 * k;
            o[n] = n == 0 ? acc[0][0] : 0u - acc[0][N1 - n];
        }
        if (tid == 0) o[N1] = acc[1][0];
    } else {
        uint32_t *o = out + size_t(ct) * 2 * N1;
#pragma unroll
        for (int k = 0; k < N1 / 128; k++) {
            o[tid + 128 * k] = acc[0][tid + 128 * k];
            o[N1 + tid + 128 * k] = acc[1][tid + 128 * k];
        }
    }
}

// one-wave device key -> this kernel's order: slice element (w, q, l) holds
// frequency f = f1w(w, l, q), which the one-wave layout keeps at register
// f >> 6 of lane ((f & 7) << 3) | ((f >> 3) & 7)
__global__ __launch_bounds__(256) void k_bk1_to_w2(const cd *__restrict__ src, cd *__restrict__ dst,
                                                  size_t nslices)
{
    const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= nslices * kPolyW) return;
    const size_t slice = e / kPolyW;
    const int k = int(e % kPolyW);
    const int w = k >> 8, q = (k >> 6) & 3, l = k & 63;
    const int f = f1w(w, l, q);
    const int s = ((f & 7) << 3) | ((f >> 3) & 7);
    dst[e] = src[slice * kPolyW + (f >> 6) * 64 + s];
}

}  // namespace

hipError_t launch_bk1_to_w2(const cd *bk1, cd *bk_w2, int n0, hipStream_t s)
{
    const size_t nslices = size_t(n0) * Lvl1::rows * 2;
    const size_t n = nslices * kPolyW;
    hipLaunchKernelGGL(k_bk1_to_w2, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, bk1, bk_w2,
                       nslices);
    return hipGetLastError();
}

hipError_t launch_blind_rotate1w2(const uint32_t *in, int n0, const cd *bk, const uint32_t *tv,
                                  int tv_stride, uint32_t *out, bool sample_extract, int B,
                                  hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    // the lane transform's first transpose as register swaps: measured A/B
    // 4.42 against 4.61 ms at 256 ciphertexts, 6.12 against 5.87 ms at 512
    if (B <= 256)
        hipLaunchKernelGGL(k_blind_rotate1w2<true>, dim3(B), dim3(128), 0, s, in, n0, bk, tv,
                           tv_stride, out, int(sample_extract), B);
    else
        hipLaunchKernelGGL(k_blind_rotate1w2<false>, dim3(B), dim3(128), 0, s, in, n0, bk, tv,
                           tv_stride, out, int(sample_extract), B);
    return hipGetLastError();
}

}  // namespace fhe
