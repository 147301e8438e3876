// launch.hpp -- host launch wrappers of the device kernels (one .hip file per
// kernel family); all take device pointers and a stream and only enqueue.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fhe {

struct cd;

// ---- lvl1 (N = 1024, 32-bit torus), pbs_lvl1.hip ------------------------
// device bootstrapping-key layout: [n0][6 rows][2 cols][8 q][64 lanes] complex
size_t bk1_device_elems(int n0);
hipError_t launch_bk1_from_int(const uint32_t *bk_int, cd *bk_dev, int n0, hipStream_t s);
hipError_t launch_bk1_from_fft(const double *bk_fft, cd *bk_dev, int n0, hipStream_t s);
// out: SE ? B x 1025 TLWE : B x 2048 TRLWE (accumulator)
hipError_t launch_blind_rotate1(const uint32_t *in, int n0, const cd *bk, const uint32_t *tv,
                                int tv_stride, uint32_t *out, bool sample_extract, int B,
                                hipStream_t s);
// One job of the persistent lvl1 bootstrap: count lvl0 TLWEs at in, one
// test vector, count lvl1 TLWEs to out; first = sum of the previous jobs'
// counts.
struct Pbs1Job {
    const uint32_t *in;
    const uint32_t *tv;
    uint32_t *out;
    int first;
    int count;
};
// every ciphertext of njobs jobs (total in all) through one persistent
// launch; cursor (1 int) and done (njobs ints) are zeroed here on stream s;
// stamp[j] receives job j's completion time on the device clock
hipError_t launch_blind_rotate1_jobs(const Pbs1Job *jobs, int njobs, int total, int n0,
                                     const cd *bk, int *cursor, int *done, uint64_t *stamp,
                                     hipStream_t s);
// the device clock (wall_clock64) now, into *t, on stream s
hipError_t launch_clock_stamp(uint64_t *t, hipStream_t s);
hipError_t launch_trgsw1_from_fft(const double *trgswfft, cd *trgsw_dev, hipStream_t s);
hipError_t launch_extprod1(const uint32_t *trlwe, const cd *trgsw_dev, uint32_t *out, int B,
                           hipStream_t s);
hipError_t launch_ifft1(const uint32_t *poly, double *fd, int B, hipStream_t s);
hipError_t launch_fft1(const double *fd, uint32_t *poly, int B, hipStream_t s);
hipError_t launch_sample_extract1(const uint32_t *trlwe, uint32_t *tlwe, int index, int B,
                                  hipStream_t s);

// ---- lvl1, two waves per ciphertext (small batches), pbs_lvl1w2.hip -----
// key: the one-wave device key with each (i, row, col) slice permuted
hipError_t launch_bk1_to_w2(const cd *bk1, cd *bk_w2, int n0, hipStream_t s);
hipError_t launch_blind_rotate1w2(const uint32_t *in, int n0, const cd *bk, const uint32_t *tv,
                                  int tv_stride, uint32_t *out, bool sample_extract, int B,
                                  hipStream_t s);

// ---- lvl1 split-MFMA variant, pbs_lvl1_mfma.hip -------------------------
size_t mfma_tables_bytes();
size_t bk1_mfma_floats(int n0);
hipError_t launch_mfma_tables(void *tab, hipStream_t s);
// FP64 device-layout key -> FP32 key in the variant's frequency order
hipError_t launch_bk1_to_mfma(const cd *bk_dev, float *bkm, int n0, hipStream_t s);
hipError_t launch_fwd_digits_mfma(const int32_t *poly, const void *tab, double *fd, int B,
                                  hipStream_t s);
hipError_t launch_blind_rotate1_mfma(const uint32_t *in, int n0, const float *bkm, const void *tab,
                                     const uint32_t *tv, int tv_stride, uint32_t *out,
                                     bool sample_extract, int B, hipStream_t s);

// ---- lvl2 (N = 2048, 64-bit torus) ----------------------------------------
// four-waves-per-ciphertext blind rotation (pbs_lvl2q.hip) and its key layout
size_t bk2q_device_elems(int n0);
hipError_t launch_bk2q_from_int(const uint64_t *bk_int, cd *bk_dev, int n0, hipStream_t s);
// tv_period > 0: ciphertext c uses test vector c % tv_period (tv_stride apart).
// rot: the table of launch_rot2q_table (w2^e, e < 4096)
hipError_t launch_blind_rotate2q(const uint32_t *in, int n0, const cd *bk, const cd *rot,
                                 const uint64_t *tv, int tv_stride, uint64_t *out,
                                 bool sample_extract, int B, hipStream_t s, int tv_period = 0);
size_t rot2q_elems();
hipError_t launch_rot2q_table(cd *rot, hipStream_t s);
hipError_t launch_ifft2q(const uint64_t *poly, double *fd, int B, hipStream_t s);
// external product, TwistFFT and sample extraction (pbs_lvl2.hip)
hipError_t launch_trgsw2_from_fft(const double *trgswfft, cd *trgsw_dev, hipStream_t s);
hipError_t launch_extprod2(const uint64_t *trlwe, const cd *trgsw_dev, uint64_t *out, int B,
                           hipStream_t s);
hipError_t launch_fft2(const double *fd, uint64_t *poly, int B, hipStream_t s);
hipError_t launch_sample_extract2(const uint64_t *trlwe, uint64_t *tlwe, int index, int B,
                                  hipStream_t s);

// ---- identity key switching and small element-wise ops, iks.hip ---------
// kind 10: uint32 domain (dn = 1024), kind 20/21: uint64 domain (dn = 2048)
// tile: kind 21, ciphertexts per workgroup (8, 16, 32; 0 = 32); kinds 10
// and 20, workgroups per ciphertext over the digit positions (1, 2, 4, 8;
// 0 = 1)
hipError_t launch_iks(int kind, const void *in, const uint32_t *ksk, uint32_t *out, int tn, int B,
                      hipStream_t s, int tile = 0);
// out = sa*a + sb*b (+ offset on the body word), lvl1 TLWEs of 1025 words
hipError_t launch_tlwe1_lincomb(const uint32_t *a, const uint32_t *b, int sa, int sb,
                                uint32_t offset, uint32_t *out, int B, hipStream_t s);
// tlwe[1024] += add for B lvl1 TLWEs
hipError_t launch_tlwe1_add_body(uint32_t *tlwe, uint32_t add, int B, hipStream_t s);
// test vectors: constant u, or gpolygen (scale 2^scale_bits, padding 6)
hipError_t launch_tv1_const(uint32_t *tv, uint32_t u, hipStream_t s);
hipError_t launch_tv1_ide(uint32_t *tv, uint32_t scale_bits, hipStream_t s);

// generic forms for either torus width (W words per TLWE)
template <class T>
hipError_t launch_tlwe_copy_add_body(const T *in, T *out, int W, T add, int B, hipStream_t s);
template <class T>
hipError_t launch_tlwe_add_body(T *tlwe, int W, T add, int B, hipStream_t s);
// out = sa*a + sb*b (+ add on the body); b may be null; b_bcast: b is one TLWE
template <class T>
hipError_t launch_tlwe_lincomb(const T *a, const T *b, int sa, int sb, T add, T *out, int W,
                               bool b_bcast, int B, hipStream_t s);
template <class T>
hipError_t launch_tlwe_shl(const T *in, T *out, int W, int shift, int B, hipStream_t s);
template <class T>
hipError_t launch_tv_const(T *tv, T u, int N, hipStream_t s);
template <class T>
hipError_t launch_tv_ide(T *tv, uint32_t scale_bits, int N, hipStream_t s);

}  // namespace fhe
