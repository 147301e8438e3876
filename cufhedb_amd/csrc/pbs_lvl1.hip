// pbs_lvl1.hip -- lvl1 (N = 1024, 32-bit torus) kernels for gfx950, FP64 path.
//
// One wave64 owns one ciphertext for the whole blind rotation: its
// accumulator (2 x 1024 words) lives in LDS with the wave's private 9 KB
// transform tile between its two columns (rot1.hpp), ordered by wave-level
// fences only, and the frequency-domain accumulators of the external
// product stay in registers.  A launch bootstraps a whole batch
// (one workgroup per ciphertext); every wave streams BK[i] (96 KB per CMUX)
// from L2 / Infinity Cache, where the 62-66 MB key stays resident across
// the batch.
//
// Reference semantics:
//   BlindRotate            thirdparty/TFHEpp/include/gatebootstrapping.hpp:15-43
//   CMUX with X^a - 1      thirdparty/TFHEpp/src/detwfa.cpp:21-30
//   external product       thirdparty/TFHEpp/src/trgsw.cpp:78-103
//   decomposition          thirdparty/TFHEpp/src/trgsw.cpp:9-39
//   SampleExtractIndex     thirdparty/TFHEpp/src/trlwe.cpp:109-118
//   GPU counterpart        src/cuHEDB/fft_gpu/BlindRotate_gpu.cuh:401-531
// The a-bar == 0 skip of the reference is kept (it is wave-uniform here).
#include <algorithm>

#include "fft64.hpp"
#include "launch.hpp"
#include "params.hpp"
#include "rot1.hpp"

namespace fhe {

namespace {

constexpr int N1 = Lvl1::N;
constexpr uint32_t kDecOff1 = kBr1DecOff;
constexpr int kPolyCd = 512;  // complex points per polynomial
// Device-layout keys hold FD / 512: the inverse transform's 2/N scale
// (fft_processor_spqlios.cpp:97-117) is folded into the key, exactly.
constexpr double kKeyScale1 = 1.0 / 512;

// Gadget digits of x (DecompositionPolynomial, trgsw.cpp:21-39):
//   digit_d = ((x + offset) >> (32 - 6(d+1)) & 63) - 32.
// prep1 adds the offset and flips the top bit of every 6-bit field, so each
// digit is then one signed 6-bit field extract: (f ^ 32) read as a signed
// 6-bit number equals f - 32.
constexpr uint32_t kDigitFlip1 = kBr1DigitFlip;
__device__ __forceinline__ uint32_t prep1(uint32_t x) { return (x + kDecOff1) ^ kDigitFlip1; }
__device__ __forceinline__ double digit1p(uint32_t u, int d)
{
    // signed extract of the field [sh, sh + 6), sh = 32 - 6 (d + 1)
    const int sh = 32 - Lvl1::Bgbit * (d + 1);
    return double(int32_t(__builtin_amdgcn_sbfe(int32_t(u), sh, Lvl1::Bgbit)));
}
__device__ __forceinline__ double digit1(uint32_t x, int d) { return digit1p(prep1(x), d); }

// load one TRGSW row (both output columns) of the device key layout
__device__ __forceinline__ void load_row(const cd *row, int lane, cd b0[8], cd b1[8])
{
#pragma unroll
    for (int q = 0; q < 8; q++) {
        b0[q] = row[q * 64 + lane];
        b1[q] = row[8 * 64 + q * 64 + lane];
    }
}

// The blind rotation of one lvl0 ciphertext c0 by the wave: acc = (0,
// X^bbar tv), n0 CMUX steps, then sample extraction (or the TRLWE) to out.
__device__ __forceinline__ void blind_rotate1_one(const uint32_t *__restrict__ c0, int n0,
                                                  const cd *__restrict__ bk,
                                                  const uint32_t *__restrict__ tvp,
                                                  uint32_t *__restrict__ out, int sample_extract,
                                                  Br1Lds &lds, int lane, const Twiddles1024 &tw)
{
    uint32_t *acc0 = lds.acc0, *acc1 = lds.acc1;
    cd *tile = lds.tile;
    // acc = (0, X^bbar * tv), bbar = 2N - (b >> 21)   (no rounding)
    const uint32_t bbar = 2u * N1 - (c0[n0] >> (32 - 1 - Lvl1::nbit));
#pragma unroll
    for (int k = 0; k < N1 / 64; k++) {
        const int n = lane + 64 * k;
        acc0[n] = 0;
        uint32_t idx = (uint32_t(n) - bbar) & (2u * N1 - 1);
        uint32_t v = tvp[idx & (N1 - 1)];
        acc1[n] = (idx & N1) ? 0u - v : v;
    }
    tile_sync<true>();

    constexpr uint32_t kRound = 1u << (32 - 2 - Lvl1::nbit);
    // column 0 at the lane's 16 words, carried from one step's update to the
    // next step's rotation (acc0 starts at zero); the two columns are written
    // out below instead of looped over, so that these registers are dead
    // during the rows' transforms (in a loop they would stay live through
    // both columns and the kernel would spill)
    uint32_t p0[16];
#pragma unroll
    for (int k = 0; k < 16; k++) p0[k] = 0;
    for (int i = 0; i < n0; i++) {
        const uint32_t abar = (c0[i] + kRound) >> (32 - 1 - Lvl1::nbit);
        if (abar == 0) continue;
        const cd *bki = bk + size_t(i) * (Lvl1::rows * 2 * kPolyCd);
        cd fd0[8], fd1[8];
#pragma unroll
        for (int q = 0; q < 8; q++) fd0[q] = fd1[q] = cd{0.0, 0.0};
        // a column's l rows, from the rotated column prepared for digit
        // extraction (u, in registers for the column's rows)
        auto rows = [&](int c, const uint32_t u[16]) {
#pragma unroll 1
            for (int d = 0; d < Lvl1::l; d++) {
                const int r = c * Lvl1::l + d;
                // issue this row's key loads before the FFT so their latency
                // overlaps the transform
                cd b0[8], b1[8];
                load_row(bki + r * 2 * kPolyCd, lane, b0, b1);
                cd v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) v[j] = cd{digit1p(u[j], d), digit1p(u[j + 8], d)};
                fft512_fwd<true, true>(v, tile, lane, tw);
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    cmac(fd0[q], v[q], b0[q]);
                    cmac(fd1[q], v[q], b1[q]);
                }
            }
        };
        {
            uint32_t u[16];
            br1_prep_column_pn(lds, 0, abar, lane, p0, u);
            rows(0, u);
        }
        {
            uint32_t u[16];
            br1_prep_column(lds, 1, abar, lane, u);
            rows(1, u);
        }
        // every lane's rotation reads are complete (the FFTs synchronised)
        fft512_inv2(fd0, fd1, tile, lane, tw);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n = lane + 64 * j;
            p0[j] = acc0[n] + trunc_torus32(fd0[j].x);
            p0[j + 8] = acc0[n + 512] + trunc_torus32(fd0[j].y);
            acc0[n] = p0[j];
            acc0[n + 512] = p0[j + 8];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int n = lane + 64 * j;
            acc1[n] += trunc_torus32(fd1[j].x);
            acc1[n + 512] += trunc_torus32(fd1[j].y);
        }
        tile_sync<true>();
    }

    if (sample_extract) {
#pragma unroll
        for (int k = 0; k < N1 / 64; k++) {
            const int n = lane + 64 * k;
            out[n] = n == 0 ? acc0[0] : 0u - acc0[N1 - n];
        }
        if (lane == 0) out[N1] = acc1[0];
    } else {
#pragma unroll
        for (int k = 0; k < N1 / 64; k++) {
            out[lane + 64 * k] = acc0[lane + 64 * k];
            out[N1 + lane + 64 * k] = acc1[lane + 64 * k];
        }
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_blind_rotate1(const uint32_t *__restrict__ in, int n0,
                                                      const cd *__restrict__ bk,
                                                      const uint32_t *__restrict__ tv,
                                                      int tv_stride, uint32_t *__restrict__ out,
                                                      int sample_extract, int B)
{
    __shared__ Br1Lds lds;
    const int lane = threadIdx.x;
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const Twiddles1024 tw = make_twiddles1024(lane);
    blind_rotate1_one(in + size_t(ct) * (n0 + 1), n0, bk, tv + size_t(ct) * tv_stride,
                      out + size_t(ct) * (sample_extract ? N1 + 1 : 2 * N1), sample_extract, lds,
                      lane, tw);
}

// Persistent form over a table of jobs (each a batch of lvl0 ciphertexts
// with one test vector and its own output buffer): every wave takes the next
// ciphertext of the concatenated jobs from a device cursor until all are
// done, so any number of requests of any sizes runs as one launch with the
// key resident, and no inputs are gathered or test vectors replicated.  A
// job's last finished ciphertext records the device clock into stamp[job]
// (its completion time).  cursor and done[] must be zero at launch.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_blind_rotate1_jobs(
    const Pbs1Job *__restrict__ jobs, int njobs, int total, int n0, const cd *__restrict__ bk,
    int *__restrict__ cursor, int *__restrict__ done, uint64_t *__restrict__ stamp)
{
    __shared__ Br1Lds lds;
    const int lane = threadIdx.x;
    const Twiddles1024 tw = make_twiddles1024(lane);
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(cursor, 1);
        idx = __builtin_amdgcn_readfirstlane(idx);
        if (idx >= total) break;
        // the job holding ciphertext idx: the last with first <= idx
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].first <= idx) lo = mid;
            else hi = mid - 1;
        }
        const Pbs1Job j = jobs[lo];
        const int ct = idx - j.first;
        blind_rotate1_one(j.in + size_t(ct) * (n0 + 1), n0, bk, j.tv,
                          j.out + size_t(ct) * (N1 + 1), 1, lds, lane, tw);
        // this ciphertext's outputs are written; the job's last one stamps it
        __threadfence();
        if (lane == 0 && atomicAdd(done + lo, 1) == j.count - 1) stamp[lo] = wall_clock64();
        __syncthreads();  // the next ciphertext rewrites the accumulator
    }
}

// Batched external product with one TRGSW (device layout [6][2][8][64]).
__global__ __launch_bounds__(64) void k_extprod1(const uint32_t *__restrict__ trlwe,
                                                 const cd *__restrict__ trgsw,
                                                 uint32_t *__restrict__ out, int B)
{
    __shared__ cd tile[kFftLdsElems];
    const int lane = threadIdx.x;
    const int ct = blockIdx.x;
    if (ct >= B) return;
    const uint32_t *src = trlwe + size_t(ct) * 2 * N1;
    const Twiddles1024 tw = make_twiddles1024(lane);
    cd fd0[8], fd1[8];
#pragma unroll
    for (int q = 0; q < 8; q++) fd0[q] = fd1[q] = cd{0.0, 0.0};
    for (int c = 0; c < 2; c++) {
        uint32_t t[16];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            t[2 * j] = src[c * N1 + lane + 64 * j];
            t[2 * j + 1] = src[c * N1 + lane + 64 * j + 512];
        }
        for (int d = 0; d < Lvl1::l; d++) {
            cd b0[8], b1[8];
            load_row(trgsw + (c * Lvl1::l + d) * 2 * kPolyCd, lane, b0, b1);
            cd v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = cd{digit1(t[2 * j], d), digit1(t[2 * j + 1], d)};
            fft512_fwd(v, tile, lane, tw);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                cmac(fd0[q], v[q], b0[q]);
                cmac(fd1[q], v[q], b1[q]);
            }
        }
    }
    uint32_t *o = out + size_t(ct) * 2 * N1;
    fft512_inv(fd0, tile, lane, tw);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        o[lane + 64 * j] = trunc_torus32(fd0[j].x);
        o[lane + 64 * j + 512] = trunc_torus32(fd0[j].y);
    }
    fft512_inv(fd1, tile, lane, tw);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        o[N1 + lane + 64 * j] = trunc_torus32(fd1[j].x);
        o[N1 + lane + 64 * j + 512] = trunc_torus32(fd1[j].y);
    }
}

// torus polynomial -> device-layout FD (one wave per polynomial)
__global__ __launch_bounds__(64) void k_poly_to_devfd1(const uint32_t *__restrict__ poly,
                                                        cd *__restrict__ fd, int npoly)
{
    __shared__ cd tile[kFftLdsElems];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    if (p >= npoly) return;
    const uint32_t *a = poly + size_t(p) * N1;
    const Twiddles1024 tw = make_twiddles1024(lane);
    cd v[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        v[j] = cd{double(int32_t(a[lane + 64 * j])), double(int32_t(a[lane + 64 * j + 512]))};
    fft512_fwd(v, tile, lane, tw);
#pragma unroll
    for (int q = 0; q < 8; q++)
        fd[size_t(p) * kPolyCd + q * 64 + lane] = cd{v[q].x * kKeyScale1, v[q].y * kKeyScale1};
}

// spqlios-order FD polynomials -> device layout (pure permutation)
__global__ void k_spq_to_devfd1(const double *__restrict__ spq, cd *__restrict__ fd, int npoly)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(npoly) * kPolyCd) return;
    const size_t p = gid / kPolyCd;
    const int e = int(gid % kPolyCd);
    const int q = e >> 6, s = e & 63;
    const int idx = rev9(fd_index1024(s, q));
    fd[gid] = cd{spq[p * N1 + idx] * kKeyScale1, spq[p * N1 + idx + 512] * kKeyScale1};
}

// torus polynomial -> spqlios-order FD (TwistIFFT)
__global__ __launch_bounds__(64) void k_ifft1(const uint32_t *__restrict__ poly,
                                               double *__restrict__ fd, int B)
{
    __shared__ cd tile[kFftLdsElems];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    if (p >= B) return;
    const uint32_t *a = poly + size_t(p) * N1;
    const Twiddles1024 tw = make_twiddles1024(lane);
    cd v[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        v[j] = cd{double(int32_t(a[lane + 64 * j])), double(int32_t(a[lane + 64 * j + 512]))};
    fft512_fwd(v, tile, lane, tw);
    double *o = fd + size_t(p) * N1;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int idx = rev9(fd_index1024(lane, q));
        o[idx] = v[q].x;
        o[idx + 512] = v[q].y;
    }
}

// spqlios-order FD -> torus polynomial (TwistFFT: 2/N scale, truncation)
__global__ __launch_bounds__(64) void k_fft1(const double *__restrict__ fd,
                                              uint32_t *__restrict__ poly, int B)
{
    __shared__ cd tile[kFftLdsElems];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    if (p >= B) return;
    const double *f = fd + size_t(p) * N1;
    const Twiddles1024 tw = make_twiddles1024(lane);
    cd v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int idx = rev9(fd_index1024(lane, q));
        v[q] = cd{f[idx], f[idx + 512]};
    }
    fft512_inv(v, tile, lane, tw);
    uint32_t *o = poly + size_t(p) * N1;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        o[lane + 64 * j] = trunc_torus32(v[j].x * kKeyScale1);
        o[lane + 64 * j + 512] = trunc_torus32(v[j].y * kKeyScale1);
    }
}

__global__ void k_sample_extract1(const uint32_t *__restrict__ trlwe, uint32_t *__restrict__ tlwe,
                                  int index, int B)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(B) * (N1 + 1)) return;
    const size_t b = gid / (N1 + 1);
    const int i = int(gid % (N1 + 1));
    const uint32_t *a = trlwe + b * 2 * N1;
    uint32_t v;
    if (i == N1)
        v = a[N1 + index];
    else if (i <= index)
        v = a[index - i];
    else
        v = 0u - a[N1 + index - i];
    tlwe[gid] = v;
}

inline unsigned grid1d(size_t n, unsigned block) { return unsigned((n + block - 1) / block); }

}  // namespace

size_t bk1_device_elems(int n0) { return size_t(n0) * Lvl1::rows * 2 * kPolyCd; }

hipError_t launch_bk1_from_int(const uint32_t *bk_int, cd *bk_dev, int n0, hipStream_t s)
{
    const int npoly = n0 * Lvl1::rows * 2;
    hipLaunchKernelGGL(k_poly_to_devfd1, dim3(npoly), dim3(64), 0, s, bk_int, bk_dev, npoly);
    return hipGetLastError();
}

hipError_t launch_bk1_from_fft(const double *bk_fft, cd *bk_dev, int n0, hipStream_t s)
{
    const int npoly = n0 * Lvl1::rows * 2;
    hipLaunchKernelGGL(k_spq_to_devfd1, dim3(grid1d(size_t(npoly) * kPolyCd, 256)), dim3(256), 0,
                       s, bk_fft, bk_dev, npoly);
    return hipGetLastError();
}

hipError_t launch_blind_rotate1(const uint32_t *in, int n0, const cd *bk, const uint32_t *tv,
                                int tv_stride, uint32_t *out, bool sample_extract, int B,
                                hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_blind_rotate1, dim3(B), dim3(64), 0, s, in, n0, bk, tv, tv_stride, out,
                       int(sample_extract), B);
    return hipGetLastError();
}

namespace {
__global__ void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_clock_stamp(uint64_t *t) { *t = wall_clock64(); }
}  // namespace

hipError_t launch_clock_stamp(uint64_t *t, hipStream_t s)
{
    hipLaunchKernelGGL(k_clock_stamp, dim3(1), dim3(1), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_blind_rotate1_jobs(const Pbs1Job *jobs, int njobs, int total, int n0,
                                     const cd *bk, int *cursor, int *done, uint64_t *stamp,
                                     hipStream_t s)
{
    if (total <= 0 || njobs <= 0) return hipSuccess;
    // zero-filled by kernels, not hipMemsetAsync, so the sequence stays
    // capturable into a graph
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(1), 0, s, cursor, 1);
    hipLaunchKernelGGL(k_zero_ints, dim3(unsigned((njobs + 255) / 256)), dim3(256), 0, s, done,
                       njobs);
    // as many waves as stay resident (two per SIMD), at most one per ciphertext
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const int waves = std::min(total, cus * 4 * 2);
    hipLaunchKernelGGL(k_blind_rotate1_jobs, dim3(unsigned(waves)), dim3(64), 0, s, jobs, njobs,
                       total, n0, bk, cursor, done, stamp);
    return hipGetLastError();
}

hipError_t launch_trgsw1_from_fft(const double *trgswfft, cd *trgsw_dev, hipStream_t s)
{
    const int npoly = Lvl1::rows * 2;
    hipLaunchKernelGGL(k_spq_to_devfd1, dim3(grid1d(size_t(npoly) * kPolyCd, 256)), dim3(256), 0,
                       s, trgswfft, trgsw_dev, npoly);
    return hipGetLastError();
}

hipError_t launch_extprod1(const uint32_t *trlwe, const cd *trgsw_dev, uint32_t *out, int B,
                           hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_extprod1, dim3(B), dim3(64), 0, s, trlwe, trgsw_dev, out, B);
    return hipGetLastError();
}

hipError_t launch_ifft1(const uint32_t *poly, double *fd, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ifft1, dim3(B), dim3(64), 0, s, poly, fd, B);
    return hipGetLastError();
}

hipError_t launch_fft1(const double *fd, uint32_t *poly, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fft1, dim3(B), dim3(64), 0, s, fd, poly, B);
    return hipGetLastError();
}

hipError_t launch_sample_extract1(const uint32_t *trlwe, uint32_t *tlwe, int index, int B,
                                  hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_extract1, dim3(grid1d(size_t(B) * (N1 + 1), 256)), dim3(256), 0, s,
                       trlwe, tlwe, index, B);
    return hipGetLastError();
}

}  // namespace fhe
