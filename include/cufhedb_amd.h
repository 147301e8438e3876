/*
 * cufhedb_amd.h -- C ABI of the MI355X (gfx950) TFHE bootstrapping library.
 *
 * Drop-in boundary for cuFHEDB's GPU hot path.  Every entry point names the
 * reference interface it replaces (paths relative to the reference root):
 *   src/cuHEDB/fft_gpu/cufft_gpu.cuh     class cufft::CuFFT_Torus<P>
 *   src/cuHEDB/gatebootstrap_gpu.cuh     cuHEDB:: bootstrap operators
 *   thirdparty/TFHEpp/...                CPU primitives the operators call
 *
 * Conventions
 *   - All functions return an int status (FHE_OK == 0).  Nothing throws.
 *     fhe_last_error() returns a thread-local message for the last failure.
 *   - Ciphertexts use TFHEpp's aggregate layouts (params.hpp:31-99): a TLWE is
 *     n+1 torus words with b last; a TRLWE is 2 polynomials of N words; a
 *     TRGSW is (k+1)l = 2l TRLWEs.  lvl0/lvl1 words are uint32_t, lvl2 words
 *     are uint64_t.  Batches are contiguous arrays of B such records.
 *   - "_dev" buffers are device pointers (hipMalloc / fhe_malloc); the call is
 *     enqueued on `stream` (a hipStream_t, NULL = the context's stream) and
 *     returns without synchronising.  Other pointers are host memory and the
 *     call is synchronous, like the reference's *_st entry points.
 *   - One context owns the device keys and scratch of one GPU.  Calls that
 *     share a context must come from one host thread at a time.
 */
#ifndef CUFHEDB_AMD_H
#define CUFHEDB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fhe_status {
    FHE_OK = 0,
    FHE_ERR_ARG = 1,         /* bad size / pointer / enum argument       */
    FHE_ERR_NOT_LOADED = 2,  /* the key this call needs was not loaded   */
    FHE_ERR_DEVICE = 3,      /* HIP runtime error (message has details)   */
    FHE_ERR_ALLOC = 4,       /* device or host allocation failed          */
    FHE_ERR_UNSUPPORTED = 5  /* option not available in this build        */
};

/* Runtime parameter set.  Ring/gadget/key-switch constants are fixed to the
 * reference's params/128bit.hpp (lvl1 N=1024,l=3,Bgbit=6; lvl2 N=2048,l=4,
 * Bgbit=9; KS10/20 t=2,basebit=10; KS21 t=10,basebit=3); the lvl0 dimension
 * and the key / noise distributions are runtime values. */
typedef struct fhe_params {
    uint32_t n0;      /* lvl0 LWE dimension (reference: 672)                  */
    uint32_t key0_hw; /* lvl0 key Hamming weight, 0 = uniform binary (512)   */
    uint32_t key1_hw; /* lvl1 key Hamming weight, 0 = uniform binary (512)   */
    uint32_t key2_hw; /* lvl2 key Hamming weight, 0 = uniform binary (1024)  */
    double alpha0;    /* lvl0 noise stdev, also KS10 / KS20 keys (2^-16)     */
    double alpha1;    /* lvl1 noise stdev: BK lvl01, KS21 (2^-25)            */
    double alpha2;    /* lvl2 noise stdev: BK lvl02 (2^-52)                  */
} fhe_params;

enum fhe_preset {
    FHE_PRESET_REF672 = 0, /* thirdparty/TFHEpp/include/params/128bit.hpp + key.cpp:8-17 */
    FHE_PRESET_CFG636 = 1  /* n0 = 636, alpha0 = 2^-15, uniform binary keys          */
};
int fhe_params_preset(int preset, fhe_params *out);

const char *fhe_last_error(void);

/* ------------------------------------------------------------------------ */
/* Host-side key generation and encryption (deterministic from a 64-bit seed).
 * Replace TFHEpp's unseeded lweKey() (src/key.cpp:8-17), bkgen
 * (src/cloudkey.cpp:7-21), ikskgen (src/cloudkey.cpp:54-70) and
 * tlweSymEncrypt (src/tlwe.cpp:19-31); same distributions, seeded RNG
 * (Philox4x32-10) so that tests and benches are reproducible.            */
int fhe_keygen_secret(const fhe_params *p, uint64_t seed, uint32_t *s0 /* n0 */,
                      uint32_t *s1 /* 1024 */, uint64_t *s2 /* 2048 */);
/* Integer bootstrapping keys: TFHEpp BootstrappingKey<lvl01param/lvl02param>
 * layout [n0][2l][2][N] (params.hpp:67-69). */
int fhe_keygen_bk_lvl01(const fhe_params *p, uint64_t seed, const uint32_t *s0, const uint32_t *s1,
                        uint32_t *bk);
int fhe_keygen_bk_lvl02(const fhe_params *p, uint64_t seed, const uint32_t *s0, const uint64_t *s2,
                        uint64_t *bk);
/* Key-switching keys, TFHEpp KeySwitchingKey layout [dn][t][2^basebit-1][tn+1]:
 *   kind 10: lvl1 (dn=1024) -> lvl0 (tn=n0), s_dom = s1 (uint32_t*)
 *   kind 20: lvl2 (dn=2048) -> lvl0 (tn=n0), s_dom = s2 (uint64_t*)
 *   kind 21: lvl2 (dn=2048) -> lvl1 (tn=1024), s_dom = s2 (uint64_t*) */
int fhe_keygen_ksk(const fhe_params *p, int kind, uint64_t seed, const void *s_dom,
                   const uint32_t *s_tgt, uint32_t *ksk);
size_t fhe_ksk_words(const fhe_params *p, int kind);
/* TLWE encryptions of B messages (torus values m[i]) under `key` (n words),
 * noise stdev alpha; ciphertext i draws from RNG stream (seed, index0 + i). */
int fhe_tlwe_encrypt_u32(uint64_t seed, uint64_t index0, const uint32_t *key, uint32_t n,
                         const uint32_t *m, double alpha, uint32_t *out, int B);
int fhe_tlwe_encrypt_u64(uint64_t seed, uint64_t index0, const uint64_t *key, uint32_t n,
                         const uint64_t *m, double alpha, uint64_t *out, int B);
/* TRLWE encryption of B polynomials (trlweSymEncrypt, src/trlwe.cpp:30-37). */
int fhe_trlwe_encrypt_u32(uint64_t seed, uint64_t index0, const uint32_t *key /* N */,
                          const uint32_t *m /* B x N */, double alpha, uint32_t *out, int B);
int fhe_trlwe_encrypt_u64(uint64_t seed, uint64_t index0, const uint64_t *key,
                          const uint64_t *m, double alpha, uint64_t *out, int B);
/* TRGSW of the constant polynomial m (trgswSymEncrypt, src/trgsw.cpp:206-224). */
int fhe_trgsw_encrypt_u32(uint64_t seed, uint64_t index, const uint32_t *key, int32_t m,
                          double alpha, uint32_t *out /* 2l x 2 x N */);
int fhe_trgsw_encrypt_u64(uint64_t seed, uint64_t index, const uint64_t *key, int64_t m,
                          double alpha, uint64_t *out);

/* ------------------------------------------------------------------------ */
/* Device context.  Replaces cufft::CuFFT_Torus<P>::CuFFT_Torus
 * (src/cuHEDB/fft_gpu/cufft_gpu.cuh:135-163) and the global singletons
 * cufftlvl1/cufftlvl2 (src/cuHEDB/gatebootstrap_gpu.cuh:24-25). */
typedef struct fhe_ctx fhe_ctx;
int fhe_ctx_create(const fhe_params *p, int device, fhe_ctx **out);
int fhe_ctx_destroy(fhe_ctx *ctx);
int fhe_ctx_sync(fhe_ctx *ctx);
void *fhe_ctx_stream(fhe_ctx *ctx);

/* Bootstrapping-key upload (CuFFT_Torus::LoadBK, cufft_gpu.cuh:175-184).
 * format FHE_BK_INT:   integer TRGSW key [n0][2l][2][N] (lvl1 uint32_t,
 *                      lvl2 uint64_t); the FFT to the device FD layout runs
 *                      on the GPU.
 * format FHE_BK_FFT:   TFHEpp BootstrappingKeyFFT [n0][2l][2][N] doubles in
 *                      spqlios order (what LoadBK takes); permuted on load. */
enum fhe_bk_format { FHE_BK_INT = 0, FHE_BK_FFT = 1 };
int fhe_load_bk(fhe_ctx *ctx, int level /* 1 or 2 */, const void *bk_host, int format);
/* Key-switching key upload (host TFHEpp layout, kind 10/20/21).  The
 * reference keeps these on the host and switches on the CPU. */
int fhe_load_ksk(fhe_ctx *ctx, int kind, const uint32_t *ksk_host);
/* 1 if the context holds a bootstrapping key of `level` (1, 2) / a
 * key-switching key of `kind` (10, 20, 21), else 0.  No reference
 * counterpart (the reference reads its keys from the caller's EvalKey on
 * every call); the C++ drop-in uses them to upload an EvalKey's keys once. */
int fhe_bk_loaded(const fhe_ctx *ctx, int level);
int fhe_ksk_loaded(const fhe_ctx *ctx, int kind);
/* Key-switching key generated on the device (ikskgen,
 * thirdparty/TFHEpp/src/cloudkey.cpp:54-70), bit-identical to fhe_keygen_ksk
 * for the same seed and keys: the device draws the masks and <a, s>, the
 * host only the noise word of each record.  ksk_dev = NULL writes the
 * context's key slot (as fhe_load_ksk would); otherwise a device buffer of
 * fhe_ksk_words(kind) words.  s_dom / s_tgt are host keys as for
 * fhe_keygen_ksk.  Synchronous. */
int fhe_keygen_ksk_dev(fhe_ctx *ctx, int kind, uint64_t seed, const void *s_dom,
                       const uint32_t *s_tgt, uint32_t *ksk_dev);

/* PBS kernel variant for the lvl1 blind rotation. */
enum fhe_pbs_variant {
    FHE_PBS_FP64 = 0,      /* FP64 negacyclic FFT, LDS-resident (plain method) */
    FHE_PBS_MFMA_SPLIT = 1 /* forward digit transforms as radix-16 x 32 DFT matrix
                              products on fp16 MFMA (hi/lo limbs, FP32
                              accumulation); key split into three integer
                              limbs (16 + 8 + 8 bits), FP32 frequency-domain
                              MAC per limb, one FP64 inverse transform per
                              limb and column, limb products rounded to
                              integers and recombined mod 2^32; lvl1 PBS only */
};
/* Selects the lvl1 blind-rotation variant for every later lvl1 bootstrap of
 * the context (plain PBS, gates, comparisons, scheduler).  The variant's key
 * layout is built here when a key is loaded (and by fhe_load_bk afterwards),
 * so bootstraps never allocate; not allowed during a graph capture. */
int fhe_set_pbs_variant(fhe_ctx *ctx, int variant);

/* ---- batched hot path on device buffers -------------------------------- */
/* Programmable bootstrap lvl0 -> lvl1 = BlindRotate + SampleExtractIndex(0)
 * (TFHEpp GateBootstrappingTLWE2TLWEFFT<lvl01param>,
 * thirdparty/TFHEpp/src/gatebootstrapping.cpp:14-23; GPU kernel
 * GateBootstrappingTLWE2TLWEFFT<Lvl1>, src/cuHEDB/fft_gpu/BlindRotate_gpu.cuh:401-531).
 *   in_dev : B x (n0+1) uint32_t lvl0 TLWEs
 *   tv_dev : test vectors, N uint32_t each; tv_stride = 0 -> one test vector
 *            for the whole batch, tv_stride = N -> one per ciphertext
 *   out_dev: B x 1025 uint32_t lvl1 TLWEs */
int fhe_pbs_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev, int tv_stride,
                      uint32_t *out_dev, int B, void *stream);
/* Blind rotation only (acc, B x 2 x N); BlindRotate, gatebootstrapping.hpp:15-43 */
int fhe_blind_rotate_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev,
                               int tv_stride, uint32_t *acc_dev, int B, void *stream);
/* lvl0 -> lvl2 PBS (GateBootstrappingTLWE2TLWEFFT<lvl02param>;
 * BlindRotate_gpu.cuh:534-670), uint64_t test vectors / outputs. */
int fhe_pbs_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev, int tv_stride,
                      uint64_t *out_dev, int B, void *stream);
/* Blind rotation only, lvl2 accumulator (B x 2 x 2048 uint64_t). */
int fhe_blind_rotate_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev,
                               int tv_stride, uint64_t *acc_dev, int B, void *stream);
/* Programmable bootstrap with num_out lookup tables in one blind rotation
 * (TFHEpp GateBootstrappingManyLUT<lvl01param / lvl02param, num_out>,
 * thirdparty/TFHEpp/include/gatebootstrapping.hpp:119-130): the modulus
 * switch keeps multiples of 2^bw, bw = bits_needed(num_out - 1), so test
 * vector coefficient j belongs to table j mod 2^bw; output k of a
 * ciphertext is SampleExtractIndex(acc, k).  num_out in [1, 64].
 *   out_dev: B x num_out x (N+1) words (num_out TLWEs per ciphertext). */
int fhe_pbs_many_lut_lvl01_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint32_t *tv_dev,
                               int tv_stride, int num_out, uint32_t *out_dev, int B, void *stream);
int fhe_pbs_many_lut_lvl02_dev(fhe_ctx *ctx, const uint32_t *in_dev, const uint64_t *tv_dev,
                               int tv_stride, int num_out, uint64_t *out_dev, int B, void *stream);
/* SampleExtractIndex (thirdparty/TFHEpp/src/trlwe.cpp:109-118;
 * BlindRotate_gpu.cuh:386-397,703-718): B TRLWEs -> B TLWEs. */
int fhe_sample_extract_lvl1_dev(fhe_ctx *ctx, const uint32_t *trlwe_dev, uint32_t *tlwe_dev,
                                int index, int B, void *stream);
int fhe_sample_extract_lvl2_dev(fhe_ctx *ctx, const uint64_t *trlwe_dev, uint64_t *tlwe_dev,
                                int index, int B, void *stream);
/* IdentityKeySwitch (thirdparty/TFHEpp/src/keyswitch.cpp:12-48); the
 * reference runs it on the host before every bootstrap
 * (gatebootstrap_gpu.cuh:42,59,76,90,101,113,133,149,159,167).
 *   kind 10: in B x 1025 uint32_t -> out B x (n0+1) uint32_t
 *   kind 20: in B x 2049 uint64_t -> out B x (n0+1) uint32_t
 *   kind 21: in B x 2049 uint64_t -> out B x 1025 uint32_t */
int fhe_iks_dev(fhe_ctx *ctx, int kind, const void *in_dev, uint32_t *out_dev, int B, void *stream);
/* External product TRGSW (x) TRLWE for a batch of TRLWEs and one TRGSW given
 * in TFHEpp TRGSWFFT layout (spqlios order) on the device
 * (trgswfftExternalProduct, TFHEpp/src/trgsw.cpp:78-103; CuFFT_Torus::
 * ExternalProduct_st, cufft_gpu.cuh:228-248). */
int fhe_external_product_lvl1_dev(fhe_ctx *ctx, const uint32_t *trlwe_dev,
                                  const double *trgswfft_dev, uint32_t *out_dev, int B,
                                  void *stream);
int fhe_external_product_lvl2_dev(fhe_ctx *ctx, const uint64_t *trlwe_dev,
                                  const double *trgswfft_dev, uint64_t *out_dev, int B,
                                  void *stream);
/* Twisted negacyclic transforms in TFHEpp's FD layout (CuFFT_Torus::ifft_st /
 * fft_st, cufft_gpu.cuh:187-212; TwistIFFT / TwistFFT, mulfft.hpp:58-121):
 * B polynomials per call. */
int fhe_ifft_lvl1_dev(fhe_ctx *ctx, const uint32_t *poly_dev, double *fd_dev, int B, void *stream);
int fhe_fft_lvl1_dev(fhe_ctx *ctx, const double *fd_dev, uint32_t *poly_dev, int B, void *stream);
/* TwistIFFT (mulfft.hpp:113-121) of small-integer polynomials (|x| <= 2048,
 * e.g. gadget digits) computed by the split-MFMA variant's transform (fp16
 * MFMA, hi/lo limbs, FP32 accumulation); TFHEpp FD layout out.  Exposes the
 * variant's forward transform for parity tests against TwistIFFT. */
int fhe_fwd_digits_lvl1_mfma_dev(fhe_ctx *ctx, const int32_t *poly_dev, double *fd_dev, int B,
                                 void *stream);
int fhe_ifft_lvl2_dev(fhe_ctx *ctx, const uint64_t *poly_dev, double *fd_dev, int B, void *stream);
int fhe_fft_lvl2_dev(fhe_ctx *ctx, const double *fd_dev, uint64_t *poly_dev, int B, void *stream);

/* ---- bootstrap operators (cuHEDB::, gatebootstrap_gpu.cuh:31-170) --------- */
enum fhe_gate {
    FHE_GATE_NAND = 0, /* TFHEpp HomNAND, gate.cpp:60-65: -a-b+mu          */
    FHE_GATE_AND = 1,  /* cuHEDB::HomAND, gatebootstrap_gpu.cuh:123-137     */
    FHE_GATE_OR = 2,   /* cuHEDB::HomOR, gatebootstrap_gpu.cuh:139-153      */
    FHE_GATE_NOR = 3,  /* TFHEpp HomNOR, gate.cpp:73-78                    */
    FHE_GATE_XNOR = 4  /* TFHEpp HomXNOR, gate.cpp:86-91                   */
};
/* result_type of the operators (HEDB/utils/types.h:4-5) */
enum fhe_result_type { FHE_ARITHMETIC = 0, FHE_LOGIC = 1 };
/* Two-input gate on lvl1 TLWEs: linear combination, IKS lvl10, PBS.
 * result_type: 1 = LOGIC (+-mu), 0 = ARITHMETIC (0 / 2mu), as in
 * HEDB/utils/types.h:4-11; NAND/NOR/XNOR are always LOGIC. */
int fhe_gate_lvl1_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                      uint32_t *out_dev, int result_type, int B, void *stream);
/* The same gate bootstrap on a fixed device scratch: allocates nothing
 * (no hipMalloc / hipFree, no allocating library call).  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * B <= 2048 (FHE_ERR_ARG above, before any launch); the scratch is one
 * 16.8 MB array per GPU, held until the process ends, shared by every call
 * on that GPU: calls that overlap on one GPU (two streams or host threads)
 * race.  The lvl1 key's kernel layout must exist (fhe_load_bk builds it). */
int fhe_gate_lvl1_noalloc_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                              uint32_t *out_dev, int result_type, int B, void *stream);
/* AND / OR of two LOGIC lvl1 TLWEs with the result already at the scale
 * log_rescale gives it: 0 or 2^scale_bits (scale_bits in [1, 31]).  No
 * reference counterpart as one call: it decrypts as
 * log_rescale(HomAND / HomOR(a, b), scale_bits) (tfhepp_utils.cpp:28-35 after
 * gatebootstrap_gpu.cuh:123-153, the pair the query filters run,
 * test/HE3_Query/TPCH_Q1.cu:266-283) with one bootstrap instead of two: the
 * gate's linear combination goes straight into the rescaling bootstrap. */
int fhe_gate_rescale_lvl1_dev(fhe_ctx *ctx, int gate, const uint32_t *a_dev, const uint32_t *b_dev,
                              uint32_t *out_dev, uint32_t scale_bits, int B, void *stream);
/* cuHEDB::MSBGateBootstrapping (lvl1 -> lvl1, gatebootstrap_gpu.cuh:31-46) */
int fhe_msb_gate_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev,
                          int result_type, int B, void *stream);
/* cuHEDB::IdeGateBootstrapping (lvl1 -> lvl1, gatebootstrap_gpu.cuh:82-92) */
int fhe_ide_gate_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev,
                          uint32_t scale_bits, int B, void *stream);

/* Level-generic forms (gatebootstrap_gpu.cuh:31-115): (in_level, out_level)
 * = (1,1), (2,1) through IKS lvl20 + PBS lvl01, or (2,2) through IKS lvl20 +
 * PBS lvl02.  Inputs / outputs are TLWEs of their level (uint32_t for lvl1,
 * uint64_t for lvl2).  For (2,1) identity bootstrapping scale_bits refers to
 * the 64-bit torus, as in the reference (it applies scale_bits - 32). */
int fhe_msb_gate_dev(fhe_ctx *ctx, int in_level, int out_level, const void *in_dev, void *out_dev,
                     int result_type, int B, void *stream);
int fhe_ide_gate_dev(fhe_ctx *ctx, int in_level, int out_level, const void *in_dev, void *out_dev,
                     uint32_t scale_bits, int B, void *stream);
/* cuHEDB::LOG_to_ARI (gatebootstrap_gpu.cuh:155-162) and cuHEDB::log_rescale
 * (:164-170): IKS lvl10, PBS with a constant test vector, body offset.  The
 * test vector follows HEDB's CPU definitions (tfhepp_utils.cpp:18-35,
 * mu_polygen(-mu) = +mu): LOG_to_ARI maps true/false to 3/8 and -1/8 (1 and
 * 0 at scale 1/2), log_rescale to 2^scale_bits and 0.  The GPU twins pass
 * -mu unnegated and so return the inverted bit. */
int fhe_log_to_ari_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev, int B,
                            void *stream);
int fhe_log_rescale_lvl1_dev(fhe_ctx *ctx, const uint32_t *in_dev, uint32_t *out_dev,
                             uint32_t scale_bits, int B, void *stream);

/* ---- comparison layer (cuHEDB msb_gpu.cuh, HomCompare_gpu.cuh) ---------- */
/* HomMSB (msb_gpu.cuh:51-57 lvl1, :175-192 lvl2; CPU twin
 * HEDB/comparison/extract_msb.cpp:241-259): encrypted top bit of a
 * plaintext of `plain_bits` bits held in the top bits of the torus, by 5-bit
 * peeling steps (shift, MSB gate, identity bootstrapping, subtract) down to
 * one MSB gate.  in_level 1: plain_bits <= 10; in_level 2: plain_bits <= 27,
 * through IKS lvl21 once at most 9 bits remain (the reference's
 * ImExtractMSB29/33 exist, but from 28 bits on the peeling steps shift the
 * lvl2 bootstrap noise onto the decision margin and results are wrong at
 * rates from ~1e-4 to ~50 %: FHE_ERR_UNSUPPORTED).  Output: lvl1 TLWEs.
 * Needs BK lvl01 + KSK 10, and for level 2 also BK lvl02 + KSK 20 and 21. */
int fhe_hom_msb_dev(fhe_ctx *ctx, int in_level, const void *in_dev, uint32_t *out_dev,
                    uint32_t plain_bits, int result_type, int B, void *stream);
enum fhe_cmp {
    FHE_CMP_GT = 0, /* cuHEDB::greater_than,       HomCompare_gpu.cuh:14-24 */
    FHE_CMP_GE = 1, /* cuHEDB::greater_than_equal, HomCompare_gpu.cuh:27-37 */
    FHE_CMP_LT = 2, /* cuHEDB::less_than,          HomCompare_gpu.cuh:40-49 */
    FHE_CMP_LE = 3, /* cuHEDB::less_than_equal,    HomCompare_gpu.cuh:52-61 */
    FHE_CMP_EQ = 4  /* cuHEDB::equal,              HomCompare_gpu.cuh:64-72 */
};
/* a op b for B pairs of TLWEs of `level` holding plain_bits-bit values at
 * scale 2^(w - plain_bits - 1) (one padding bit, as in
 * test/HE3_Query/HomComp.cu:75-80).  b_broadcast != 0 compares every a with
 * the single TLWE at b_dev (a shared or public threshold).  Level 1:
 * plain_bits <= 9; level 2: plain_bits <= 26 (FHE_ERR_UNSUPPORTED above, see
 * fhe_hom_msb_dev).  Output: lvl1 TLWEs, LOGIC or ARITHMETIC. */
int fhe_compare_dev(fhe_ctx *ctx, int level, int op, const void *a_dev, const void *b_dev,
                    int b_broadcast, uint32_t *out_dev, uint32_t plain_bits, int result_type,
                    int B, void *stream);
/* HomNOT<P> (gatebootstrap_gpu.cuh:117-121; TFHEpp gate.cpp:30-34): negate
 * every word of B TLWEs of level 0, 1 or 2 (out may equal in). */
int fhe_hom_not_dev(fhe_ctx *ctx, int level, const void *in_dev, void *out_dev, int B,
                    void *stream);
/* Digit-wise comparison with a public threshold: x op threshold for B
 * integers given as radix-2^digit_bits digits.  No single reference entry
 * point: it replaces HomCompare_gpu.cuh:14-76 (a op b through HomMSB chains,
 * b an encrypted or trivially encrypted threshold) for the case of a
 * threshold known in the clear.
 *   x_dev : B x digits lvl2 TLWEs (2049 words each), row-major by value,
 *           digit i (least significant first) of value b at
 *           x_dev + (b * digits + i) * 2049, holding its digit v at
 *           v * 2^(63 - digit_bits) (one padding bit);
 *   digits in [1, 5], digit_bits in [1, 4], threshold < 2^(digits * digit_bits);
 *   out_level 2: B lvl2 TLWEs, 1: B lvl1 TLWEs (through IKS lvl21);
 *   result_type LOGIC (+-mu) or ARITHMETIC (0 / 2mu) as fhe_compare_dev.
 * One lvl02 bootstrap per digit and one combining bootstrap per value.
 * Needs BK lvl02 and KSK 20, and KSK 21 for out_level 1. */
int fhe_compare_digits_dev(fhe_ctx *ctx, int op, const uint64_t *x_dev, int digits, int digit_bits,
                           uint64_t threshold, int out_level, void *out_dev, int result_type, int B,
                           void *stream);
/* The test vectors fhe_compare_digits_dev builds on the device, on the host:
 * tv receives (digits + 1) x 2048 words (one per digit position, then the
 * combining sign bootstrap's); *combine_add is the body offset added to the
 * sum of the bootstrapped digits, *post_add the one added to the result. */
int fhe_digit_testvectors(int op, int digits, int digit_bits, uint64_t threshold, int result_type,
                          uint64_t *tv, uint64_t *combine_add, uint64_t *post_add);

/* ---- tile autotuning ---------------------------------------------------- */
/* No reference counterpart (the reference launches one ciphertext per call).
 * FHE_TUNE_IKS21_TILE: ciphertexts per workgroup of the lvl2 -> lvl1 key
 * switch: 8, 16 or 32 with one output column per lane (more read each key
 * row once for more ciphertexts, fewer give small batches more workgroups),
 * or 64 with one ciphertext per lane (key rows staged in LDS; the fastest
 * form from about 512 ciphertexts up).  fhe_autotune
 * times every candidate on the context's KSK 21 at the n ascending batch
 * sizes and keeps the fastest per size (a batch then uses the entry of the
 * largest size <= it; untuned: 8 below 256 ciphertexts, 16 below 512,
 * else 64); choice (n, may be NULL) receives the picks, ms (n x the
 * knob's candidate count, in candidate order, may be NULL) the mean
 * times.  All candidates compute identical results.  fhe_tune_get /
 * fhe_tune_set read and restore the table (e.g. from a per-device cache). */
enum fhe_tune_knob {
    FHE_TUNE_IKS21_TILE = 0, /* ciphertexts per workgroup: 8, 16, 32, 64      */
    FHE_TUNE_PBS1_WAVES = 1, /* waves per ciphertext of the FP64 lvl1 blind
                                rotation: 1, or 2 (small batches: each
                                ciphertext's CMUX on two SIMDs); untuned: 2
                                below 512 ciphertexts */
    FHE_TUNE_IKS_SPLIT = 2,  /* workgroups per ciphertext of the lvl1 -> lvl0
                                key switch (kind 10; 1, 2, 4, 8: the digit
                                positions split, sums added with integer
                                atomics); untuned: 8; needs KSK 10 */
    FHE_TUNE_IKS20_SPLIT = 3 /* the same for the lvl2 -> lvl0 key switch
                                (kind 20: twice the digit positions, 64-bit
                                inputs), tuned apart; untuned: 8; needs KSK 20 */
};
int fhe_autotune(fhe_ctx *ctx, int knob, const int *batches, int n, int *choice, float *ms);
int fhe_tune_set(fhe_ctx *ctx, int knob, const int *batches, const int *tiles, int n);
int fhe_tune_get(fhe_ctx *ctx, int knob, int *batches, int *tiles, int cap, int *n);

/* ---- HIP graph capture of pipelines ------------------------------------ */
/* Replaces the reference's one-host-call-and-synchronise-per-bootstrap
 * driving of a query (cufft_gpu.cuh:262, TPCH_Q1.cu:255-283): the _dev calls
 * made on the context stream (stream argument NULL) between fhe_graph_begin
 * and fhe_graph_end are recorded, not run, and fhe_graph_launch replays them
 * as one graph on the buffers they were given.  Run the pipeline once at the
 * same sizes before capturing it: scratch that would have to grow during a
 * capture makes the call fail with FHE_ERR_ARG (end the capture with
 * fhe_graph_end all the same).  Host-pointer forms and key loading cannot be
 * captured.  One capture per thread at a time. */
typedef struct fhe_graph fhe_graph;
int fhe_graph_begin(fhe_ctx *ctx);
int fhe_graph_end(fhe_ctx *ctx, fhe_graph **out);
int fhe_graph_launch(fhe_graph *g, void *stream /* NULL = the context stream */);
int fhe_graph_destroy(fhe_graph *g);

/* ---- coalescing scheduler (mixed lvl1 / lvl2 request stream) ------------ */
/* Replaces the reference's serving pattern for a stream of independent
 * requests -- one host thread per CUDA stream, one ciphertext per call
 * (src/cuHEDB/gatebootstrap_gpu.cuh:17-25, test/HE3_Query/TPCH_Q6.cu:179-193)
 * -- for the two request kinds of the mixed workload: lvl1 LUT bootstraps
 * (fhe_pbs_lvl01_dev with one test vector per request) and lvl2 comparisons
 * (fhe_compare_dev level 2 against one threshold ciphertext per request).
 * fhe_sched_flush serves everything submitted so far, in order, merging
 * consecutive requests of one kind (and one op / plain_bits / result_type)
 * into batched launches of up to max_rows rows through device staging
 * buffers; a request of more than max_rows rows runs alone.  Buffers must
 * stay valid until fhe_sched_wait returns for the request.  One host thread
 * per scheduler; the scheduler uses its context's stream. */
typedef struct fhe_sched fhe_sched;
enum fhe_sched_kind { FHE_SCHED_PBS1 = 0, FHE_SCHED_CMP2 = 1 };
int fhe_sched_create(fhe_ctx *ctx, int max_rows, fhe_sched **out);
int fhe_sched_destroy(fhe_sched *s);
/* in_dev: B lvl0 TLWEs; tv_dev: one test vector (N words); out_dev: B lvl1 TLWEs */
int fhe_sched_submit_pbs1(fhe_sched *s, const uint32_t *in_dev, const uint32_t *tv_dev,
                          uint32_t *out_dev, int B, uint64_t *ticket);
/* a_dev: B lvl2 TLWEs; b_dev: one lvl2 TLWE; out_dev: B lvl1 TLWEs (a op b);
 * plain_bits <= 26 (FHE_ERR_UNSUPPORTED above, as fhe_compare_dev) */
int fhe_sched_submit_cmp2(fhe_sched *s, int op, const uint64_t *a_dev, const uint64_t *b_dev,
                          uint32_t *out_dev, uint32_t plain_bits, int result_type, int B,
                          uint64_t *ticket);
int fhe_sched_flush(fhe_sched *s);
/* Blocks until the request is done; ms (may be NULL) = device time from the
 * start of the flush that served it to its completion. */
int fhe_sched_wait(fhe_sched *s, uint64_t ticket, float *ms);

/* ---- key material files ------------------------------------------------- */
/* The reference's key files: TFHEpp cereal PortableBinary archives written by
 * HEDB::generate_*_key and read by HEDB::load_*_key
 * (src/HEDB/utils/serialize.cpp:7-320; file names ./../evalkey/<name>.key),
 * and the raw fastR dumps (test/HE3_Query/fastR.h:6-74).  Layouts:
 *   secret.key      SecretKey::serialize (TFHEpp/include/key.hpp:29-35):
 *                   flag byte, lvl0 key [n0] u32, lvl1 key [1024] u32,
 *                   lvl2 key [2048] u64, lweParams
 *   <member>.key    EvalKey::serialize_<member> (TFHEpp/include/cloudkey.hpp:
 *                   164-228): flag byte, lweParams, presence byte, the array
 * The flag byte is cereal's endianness mark (1 = little-endian); big-endian
 * files are byte-swapped on read, files are written little-endian. */
enum fhe_keyfile {
    FHE_KEYFILE_SECRET = 0,     /* secret.key                                  */
    FHE_KEYFILE_BKLVL01 = 1,    /* bklvl01.key     [n0][6][2][1024] uint32_t   */
    FHE_KEYFILE_BKLVL02 = 2,    /* bklvl02.key     [n0][8][2][2048] uint64_t   */
    FHE_KEYFILE_BKFFTLVL01 = 3, /* bkfftlvl01.key  [n0][6][2][1024] double     */
    FHE_KEYFILE_BKFFTLVL02 = 4, /* bkfftlvl02.key  [n0][8][2][2048] double     */
    FHE_KEYFILE_IKSKLVL10 = 5,  /* iksklvl10.key   [1024][2][1023][n0+1] u32   */
    FHE_KEYFILE_IKSKLVL20 = 6,  /* iksklvl20.key   [2048][2][1023][n0+1] u32   */
    FHE_KEYFILE_IKSKLVL21 = 7   /* iksklvl21.key   [2048][10][7][1025] u32     */
};
/* lweParams as stored (TFHEpp/include/lweParams.hpp:181-190), field order. */
typedef struct fhe_keyfile_params {
    uint32_t n0;  double alpha0;  uint32_t approx_bit0;
    uint32_t nbit1, n1, l1, bgbit1;  double alpha1;  uint32_t approx_bit1;
    uint32_t nbit2, n2, l2, bgbit2;  double alpha2;  uint32_t approx_bit2;
    uint32_t t10, basebit10;  double alpha10;
    uint32_t t20, basebit20;  double alpha20;
    uint32_t t21, basebit21;  double alpha21;
    uint32_t t22, basebit22;  double alpha22;
} fhe_keyfile_params;
/* The parameter block this library writes for p (128bit.hpp values, p's n0
 * and noise). */
int fhe_keyfile_params_of(const fhe_params *p, fhe_keyfile_params *out);
/* Bytes of a file's key array (secret: all three keys). 0 for a bad kind. */
uint64_t fhe_keyfile_payload_bytes(int kind, uint32_t n0);
/* Parse and check a file's header and size; either output may be NULL. */
int fhe_keyfile_info(const char *path, int kind, fhe_keyfile_params *params,
                     uint64_t *payload_bytes);
/* HEDB::load_<member>_key into host memory (kinds 1-7). */
int fhe_keyfile_read(const char *path, int kind, void *dst, uint64_t dst_bytes);
/* HEDB::load_secret_key (serialize.cpp:17-24); params may be NULL. */
int fhe_keyfile_read_secret(const char *path, uint32_t *s0, uint32_t n0_cap, uint32_t *s1,
                            uint64_t *s2, fhe_keyfile_params *params);
/* HEDB::generate_<member>_key's archive step (serialize.cpp:26-313) for a
 * key the caller holds in TFHEpp layout. */
int fhe_keyfile_write(const char *path, int kind, const fhe_params *p, const void *payload);
/* HEDB::generate_secret_key's archive step (serialize.cpp:7-15). */
int fhe_keyfile_write_secret(const char *path, const fhe_params *p, const uint32_t *s0,
                             const uint32_t *s1, const uint64_t *s2);
/* HEDB::load_<member>_key straight into a context: bootstrapping keys go
 * through fhe_load_bk, key-switching keys stream from the file to HBM in
 * 64 MiB pinned chunks.  The file's n0 must equal the context's. */
int fhe_keyfile_load(fhe_ctx *ctx, const char *path, int kind);
/* fastR.h writeToFile / readFromFile (test/HE3_Query/fastR.h:6-74): lvl1
 * key, BootstrappingKeyFFT<lvl01param>, KeySwitchingKey<lvl10param>, raw.
 * fhe_fastr_load loads the two keys into ctx (n0 = ctx's) and returns the
 * lvl1 secret key in s1_out (1024 words) when s1_out is not NULL. */
int fhe_fastr_write(const char *path, uint32_t n0, const uint32_t *s1, const double *bkfft,
                    const uint32_t *iksk10);
int fhe_fastr_load(fhe_ctx *ctx, const char *path, uint32_t *s1_out);

/* ---- host-pointer convenience forms (synchronous, like the *_st calls) -- */
int fhe_pbs_lvl01(fhe_ctx *ctx, const uint32_t *in, const uint32_t *tv, int tv_stride,
                  uint32_t *out, int B);
int fhe_iks(fhe_ctx *ctx, int kind, const void *in, uint32_t *out, int B);

/* ---- device memory and timing helpers ---------------------------------- */
int fhe_malloc(void **ptr, size_t bytes);
int fhe_free(void *ptr);
int fhe_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int fhe_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);
int fhe_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int fhe_memset_dev(void *dst_dev, int value, size_t bytes);
int fhe_device_sync(void);
/* Elapsed milliseconds of the work enqueued on `stream` between
 * fhe_timer_start and fhe_timer_stop (HIP events on that stream). */
int fhe_timer_start(fhe_ctx *ctx, void *stream);
int fhe_timer_stop(fhe_ctx *ctx, void *stream, float *ms);
int fhe_device_name(char *buf, int len);
/* Measured ceilings (no reference counterpart; roofline denominators next
 * to the published ones): FP64 FMA rate of the vector ALUs in TFLOP/s (16
 * independent chains per lane, 8 workgroups of 4 waves per CU) and the
 * streaming copy bandwidth in GB/s (512 MiB, read + write bytes).  Uses the
 * context stream; allocates and frees its own buffers.  Synchronous. */
int fhe_measure_peaks(fhe_ctx *ctx, double *fp64_tflops, double *copy_gbs);

#ifdef __cplusplus
}
#endif
#endif
