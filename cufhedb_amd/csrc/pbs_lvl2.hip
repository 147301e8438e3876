// pbs_lvl2.hip -- lvl2 (N = 2048, 64-bit torus, l = 4, Bgbit = 9) batched
// external product, TwistFFT and sample extraction for gfx950, FP64 path.
// (The lvl2 blind rotation is k_blind_rotate2q in pbs_lvl2q.hip.)
//
// The twisted transform has 1024 complex points.  A workgroup of two waves
// owns one polynomial pair: one radix-2 decimation-in-frequency step across
// the two waves (through LDS) splits the 1024-point DFT into two plain
// 512-point DFTs, and each wave runs the wave-level FFT of fft64.hpp on its
// half:
//   t_n  = z_n w2^n              (w2 = exp(i pi/2048), z_n = a_n + i a_{n+1024})
//   u0_n = t_n + t_{n+512},  u1_n = (t_n - t_{n+512}) W1024^n,   n < 512
//   Z[2m + h] = DFT512(u_h)[m]
// Wave h, lane L, register j holds point n = L + 64 j + 512 h in time and
// Z[2 fd_index1024(lane, q) + h] in frequency.
//
// Reference semantics: external product / decomposition as in pbs_lvl1.hip,
// instantiated for lvl2param (thirdparty/TFHEpp/include/params/128bit.hpp:35-50);
// torus64 conversions follow execute_reverse_torus64 / execute_direct_torus64
// (thirdparty/TFHEpp/thirdparties/spqlios/fft_processor_spqlios.cpp:88-93,182-221);
// GPU counterpart ExternalProduct<Lvl2> (src/cuHEDB/fft_gpu/ExternalProduct_gpu.cuh:371-473).
#include "fft64.hpp"
#include "launch.hpp"
#include "params.hpp"

namespace fhe {

namespace {

constexpr int N2 = Lvl2::N;
constexpr uint64_t kDecOff2 = decomp_total_offset<uint64_t, 64, Lvl2::l, Lvl2::Bgbit>();
constexpr int kHalfCd = 512;             // complex points per wave
constexpr int kPoly2Cd = 1024;           // complex points per polynomial
constexpr int kRow2Cd = 2 * kPoly2Cd;    // one TRGSW row, both output columns

// exp(2 pi i j / 64), j = 0..7  (w2^{64 j})
__device__ __forceinline__ cd twist2_const(int j)
{
    switch (j) {
    case 0: return {1.0, 0.0};
    case 1: return {0.99518472667219688624, 0.09801714032956060199};
    case 2: return {0.98078528040323044913, 0.19509032201612826785};
    case 3: return {0.95694033573220886494, 0.29028467725446236764};
    case 4: return {0.92387953251128675613, 0.38268343236508977173};
    case 5: return {0.88192126434835502971, 0.47139673682599764856};
    case 6: return {0.83146961230254523708, 0.55557023301960222474};
    default: return {0.77301045336273696081, 0.63439328416364549822};
    }
}

// exp(2 pi i j / 16), j = 0..7  (W1024^{64 j})
__device__ __forceinline__ cd dif_const(int j)
{
    constexpr double c = 0.92387953251128675613, s = 0.38268343236508977173;
    switch (j) {
    case 0: return {1.0, 0.0};
    case 1: return {c, s};
    case 2: return {kR2, kR2};
    case 3: return {s, c};
    case 4: return {0.0, 1.0};
    case 5: return {-s, c};
    case 6: return {-kR2, kR2};
    default: return {-c, s};
    }
}

struct Tw2 {
    Twiddles1024 plain;  // plain 512-point DFT bases
    cd beta;             // w2^{L + 512 h}
    cd gamma;            // W1024^L
};

__device__ __forceinline__ Tw2 make_tw2(int lane, int h)
{
    Tw2 t;
    t.plain = make_twiddles512_plain(lane);
    t.beta = unit_root(double(lane + 512 * h), 2048.0);
    t.gamma = unit_root(double(lane), 512.0);
    return t;
}

// Forward: v[j] = z[L + 64 j + 512 h] on entry -> Z[2 m + h] (permuted m) on exit.
// `xch` is this wave's tile, `pxch` the partner wave's.
__device__ __forceinline__ void fft1024_fwd(cd v[8], cd *xch, const cd *pxch, int lane, int h,
                                            const Tw2 &tw)
{
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = cmul(v[j], cmul(tw.beta, twist2_const(j)));
#pragma unroll
    for (int j = 0; j < 8; j++) xch[j * kFftPad + lane] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const cd o = pxch[j * kFftPad + lane];
        v[j] = h == 0 ? cadd(v[j], o) : cmul(csub(o, v[j]), cmul(tw.gamma, dif_const(j)));
    }
    __syncthreads();
    fft512_fwd<false>(v, xch, lane, tw.plain);
}

// Inverse: Z[2 m + h] (permuted m) in -> 1024 z[L + 64 j + 512 h] out.
__device__ __forceinline__ void fft1024_inv(cd v[8], cd *xch, const cd *pxch, int lane, int h,
                                            const Tw2 &tw)
{
    fft512_inv<false>(v, xch, lane, tw.plain);
#pragma unroll
    for (int j = 0; j < 8; j++) xch[j * kFftPad + lane] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const cd o = pxch[j * kFftPad + lane];
        const cd y0 = h == 0 ? v[j] : o;
        const cd y1 = h == 0 ? o : v[j];
        const cd r = cmulc(y1, cmul(tw.gamma, dif_const(j)));
        v[j] = h == 0 ? cadd(y0, r) : csub(y0, r);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = cmulc(v[j], cmul(tw.beta, twist2_const(j)));
}

// Truncation toward zero of the exact value of x, reduced mod 2^64
// (execute_direct_torus64's mantissa shift; |x| < 1 gives 0), in FP64
// arithmetic that is exact for |x| < 2^85: t = trunc(x) = H 2^32 + L with
// L in [0, 2^32) and H reduced mod 2^32 the same way.
__device__ __forceinline__ uint64_t trunc_torus64(double x)
{
    const double t = trunc(x);
    const double H = floor(t * 0x1p-32);
    const double L = __fma_rn(-H, 0x1p32, t);
    const double Hm = __fma_rn(-floor(H * 0x1p-32), 0x1p32, H);
    return (uint64_t(uint32_t(Hm)) << 32) | uint64_t(uint32_t(L));
}

// Gadget digits of x (DecompositionPolynomial, trgsw.cpp:21-39, lvl2):
//   digit_d = ((x + offset) >> (64 - 9(d+1)) & 511) - 256.
// prep2 adds the offset and flips the top bit of every 9-bit field, so each
// digit is one signed 9-bit field read: the fields at bits 55, 46 and 37 lie
// in the high word, the one at bit 28 straddles both words.
constexpr uint64_t kDigitFlip2 = (1ull << 63) | (1ull << 54) | (1ull << 45) | (1ull << 36);
__device__ __forceinline__ uint64_t prep2(uint64_t x) { return (x + kDecOff2) ^ kDigitFlip2; }
__device__ __forceinline__ double digit2p(uint64_t u, int d)
{
    const uint32_t hi = uint32_t(u >> 32), lo = uint32_t(u);
    const uint32_t top = d < 3 ? hi << (Lvl2::Bgbit * d) : ((hi << 4) | (lo >> 28)) << 23;
    return double(int32_t(top) >> (32 - Lvl2::Bgbit));
}
__device__ __forceinline__ double digit2(uint64_t x, int d) { return digit2p(prep2(x), d); }

// row pointer (both columns) -> this wave's half: [c][h][q][lane]
__device__ __forceinline__ void load_row2(const cd *row, int lane, int h, cd b0[8], cd b1[8])
{
#pragma unroll
    for (int q = 0; q < 8; q++) {
        b0[q] = row[h * kHalfCd + q * 64 + lane];
        b1[q] = row[kPoly2Cd + h * kHalfCd + q * 64 + lane];
    }
}

__global__ __launch_bounds__(128) void k_extprod2(const uint64_t *__restrict__ trlwe,
                                                  const cd *__restrict__ trgsw,
                                                  uint64_t *__restrict__ out, int B)
{
    __shared__ cd tile[2][kFftLdsElems];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const uint64_t *src = trlwe + size_t(ct) * 2 * N2;
    const Tw2 tw = make_tw2(lane, h);
    cd *xch = tile[h];
    const cd *pxch = tile[1 - h];
    cd fd0[8], fd1[8];
#pragma unroll
    for (int q = 0; q < 8; q++) fd0[q] = fd1[q] = cd{0.0, 0.0};
#pragma unroll 1
    for (int r = 0; r < Lvl2::rows; r++) {
        const int c = r / Lvl2::l, d = r - c * Lvl2::l;
        cd b0[8], b1[8];
        load_row2(trgsw + r * kRow2Cd, lane, h, b0, b1);
        cd v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n = lane + 64 * j + 512 * h;
            v[j] = cd{digit2(src[c * N2 + n], d), digit2(src[c * N2 + n + 1024], d)};
        }
        fft1024_fwd(v, xch, pxch, lane, h, tw);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            cmac(fd0[q], v[q], b0[q]);
            cmac(fd1[q], v[q], b1[q]);
        }
    }
    uint64_t *o = out + size_t(ct) * 2 * N2;
    fft1024_inv(fd0, xch, pxch, lane, h, tw);
    fft1024_inv(fd1, xch, pxch, lane, h, tw);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int n = lane + 64 * j + 512 * h;
        o[n] = trunc_torus64(fd0[j].x * (1.0 / 1024));
        o[n + 1024] = trunc_torus64(fd0[j].y * (1.0 / 1024));
        o[N2 + n] = trunc_torus64(fd1[j].x * (1.0 / 1024));
        o[N2 + n + 1024] = trunc_torus64(fd1[j].y * (1.0 / 1024));
    }
}

// bit reversal over 10 bits: spqlios order of the 1024 FD points
__host__ __device__ constexpr int rev10(int x)
{
    int r = 0;
    for (int i = 0; i < 10; i++) r |= ((x >> i) & 1) << (9 - i);
    return r;
}

// frequency index of device position (h, q, lane)
__host__ __device__ constexpr int fd_index2048(int h, int s, int q)
{
    return 2 * fd_index1024(s, q) + h;
}

__global__ void k_spq_to_devfd2(const double *__restrict__ spq, cd *__restrict__ fd, int npoly)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(npoly) * kPoly2Cd) return;
    const size_t p = gid / kPoly2Cd;
    const int e = int(gid % kPoly2Cd);
    const int h = e / kHalfCd, q = (e % kHalfCd) >> 6, s = e & 63;
    const int idx = rev10(fd_index2048(h, s, q));
    fd[gid] = cd{spq[p * N2 + idx], spq[p * N2 + idx + 1024]};
}

__global__ __launch_bounds__(128) void k_fft2(const double *__restrict__ fd,
                                               uint64_t *__restrict__ poly, int B)
{
    __shared__ cd tile[2][kFftLdsElems];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const int p = blockIdx.x;
    if (p >= B) return;
    const double *f = fd + size_t(p) * N2;
    const Tw2 tw = make_tw2(lane, h);
    cd v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int idx = rev10(fd_index2048(h, lane, q));
        v[q] = cd{f[idx], f[idx + 1024]};
    }
    fft1024_inv(v, tile[h], tile[1 - h], lane, h, tw);
    uint64_t *o = poly + size_t(p) * N2;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int n = lane + 64 * j + 512 * h;
        o[n] = trunc_torus64(v[j].x * (1.0 / 1024));
        o[n + 1024] = trunc_torus64(v[j].y * (1.0 / 1024));
    }
}

__global__ void k_sample_extract2(const uint64_t *__restrict__ trlwe, uint64_t *__restrict__ tlwe,
                                  int index, int B)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(B) * (N2 + 1)) return;
    const size_t b = gid / (N2 + 1);
    const int i = int(gid % (N2 + 1));
    const uint64_t *a = trlwe + b * 2 * N2;
    uint64_t v;
    if (i == N2)
        v = a[N2 + index];
    else if (i <= index)
        v = a[index - i];
    else
        v = 0ull - a[N2 + index - i];
    tlwe[gid] = v;
}

inline unsigned grid1d(size_t n, unsigned block) { return unsigned((n + block - 1) / block); }

}  // namespace

hipError_t launch_trgsw2_from_fft(const double *trgswfft, cd *trgsw_dev, hipStream_t s)
{
    const int npoly = Lvl2::rows * 2;
    hipLaunchKernelGGL(k_spq_to_devfd2, dim3(grid1d(size_t(npoly) * kPoly2Cd, 256)), dim3(256), 0,
                       s, trgswfft, trgsw_dev, npoly);
    return hipGetLastError();
}

hipError_t launch_extprod2(const uint64_t *trlwe, const cd *trgsw_dev, uint64_t *out, int B,
                           hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_extprod2, dim3(B), dim3(128), 0, s, trlwe, trgsw_dev, out, B);
    return hipGetLastError();
}

hipError_t launch_fft2(const double *fd, uint64_t *poly, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fft2, dim3(B), dim3(128), 0, s, fd, poly, B);
    return hipGetLastError();
}

hipError_t launch_sample_extract2(const uint64_t *trlwe, uint64_t *tlwe, int index, int B,
                                  hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_extract2, dim3(grid1d(size_t(B) * (N2 + 1), 256)), dim3(256), 0,
                       s, trlwe, tlwe, index, B);
    return hipGetLastError();
}

}  // namespace fhe
