/*
 * tfhe_oracle.h -- CPU restatement of the reference TFHE bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  The product library (cufhedb_amd) never calls it.
 *
 * Every function restates one reference function (file:line relative to the
 * reference repository root, "TFHEpp/" = thirdparty/TFHEpp/).  Torus values are
 * passed as uint32_t (lvl0/lvl1, 32-bit torus) or uint64_t (lvl2, 64-bit
 * torus) arrays laid out exactly like TFHEpp's std::array aggregates:
 *   TLWE<P>      : n+1 words, b last                      (TFHEpp/include/params.hpp:34-35)
 *   TRLWE<P>     : (k+1)=2 polynomials of N words          (params.hpp:51-52)
 *   TRGSW<P>     : (k+1)l TRLWEs                           (params.hpp:62-63)
 *   TRGSWFFT<P>  : (k+1)l x (k+1) polynomials of N doubles (params.hpp:64-65)
 *   BootstrappingKeyFFT : n0 x TRGSWFFT                    (params.hpp:72-74)
 *   KeySwitchingKey     : [dn][t][2^basebit-1][tn+1]       (params.hpp:79-83)
 * Frequency-domain (FD) polynomials use the spqlios layout: N doubles,
 * [re(N/2) | im(N/2)], point idx = evaluation at w^(1+4*rev(idx)),
 * w = exp(i*pi/N), rev = bit reversal over log2(N/2) bits.
 *
 * Parity pins (each function is checked by tests/test_oracle_*.py):
 *   - FFT / external product / blind rotate / IKS / SE / decomposition:
 *     against the compiled reference TFHEpp (oracle/_ref) on the same inputs
 *     and against committed golden vectors (tests/golden/);
 *   - FFT: against a brute-force O(N^2) DFT and the exact integer negacyclic
 *     product (TFHEpp/test/polymul.cpp:19-67 bounds);
 *   - decomposition: digit recombination closed form;
 *   - gates / PBS: decryption against the plaintext truth table.
 */
#ifndef TFHE_ORACLE_H
#define TFHE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- twisted negacyclic FFT (spqlios semantics) ------------------------- */
/* TwistIFFT lvl1: TFHEpp/include/mulfft.hpp:113-121 ->
 * fft_processor_spqlios.cpp:41-86 (int32 reinterpretation, twist, DFT). */
void orc_ifft_torus32(double *fd, const uint32_t *a, int N);
/* TwistFFT lvl1: mulfft.hpp:58-66 -> fft_processor_spqlios.cpp:95-122
 * (scale 2/N, inverse DFT, untwist, uint32_t(int64_t(x)) truncation). */
void orc_fft_torus32(uint32_t *res, const double *fd, int N);
/* lvl2: execute_reverse_torus64 (fft_processor_spqlios.cpp:88-93, int64 ->
 * double rounding) and execute_direct_torus64 (:182-221, truncation toward
 * zero of the exact double value, reduced mod 2^64). */
void orc_ifft_torus64(double *fd, const uint64_t *a, int N);
void orc_fft_torus64(uint64_t *res, const double *fd, int N);
/* Bit-reversal index used by the spqlios FD order (M = N/2 points). */
int orc_fd_rev(int idx, int M);

/* ---- polynomial helpers ------------------------------------------------- */
/* PolynomialMulByXai / PolynomialMulByXaiMinusOne: TFHEpp/include/utils.hpp:110-141 */
void orc_mul_xai_u32(uint32_t *res, const uint32_t *poly, uint32_t a, int N);
void orc_mul_xai_minus_one_u32(uint32_t *res, const uint32_t *poly, uint32_t a, int N);
void orc_mul_xai_u64(uint64_t *res, const uint64_t *poly, uint32_t a, int N);
void orc_mul_xai_minus_one_u64(uint64_t *res, const uint64_t *poly, uint32_t a, int N);
/* DecompositionPolynomial: TFHEpp/src/trgsw.cpp:9-39 (offsetgen + round). */
void orc_decompose_u32(uint32_t *dec, const uint32_t *poly, int digit, int N, int l, int Bgbit);
void orc_decompose_u64(uint64_t *dec, const uint64_t *poly, int digit, int N, int l, int Bgbit);
/* Exact negacyclic product mod 2^w, a taken as signed (PolyMulNaieve,
 * TFHEpp/include/mulfft.hpp:172-187). */
void orc_negacyclic_naive_u32(uint32_t *res, const uint32_t *a, const uint32_t *b, int N);
void orc_negacyclic_naive_u64(uint64_t *res, const uint64_t *a, const uint64_t *b, int N);

/* ---- external product / CMUX ------------------------------------------- */
/* ApplyFFT2trgsw: TFHEpp/src/trgsw.cpp:163-170 */
void orc_trgsw_to_fft_u32(double *trgswfft, const uint32_t *trgsw, int N, int l);
void orc_trgsw_to_fft_u64(double *trgswfft, const uint64_t *trgsw, int N, int l);
/* trgswfftExternalProduct: TFHEpp/src/trgsw.cpp:78-103 (k = 1). */
void orc_external_product_u32(uint32_t *res, const uint32_t *trlwe, const double *trgswfft,
                              int N, int l, int Bgbit);
void orc_external_product_u64(uint64_t *res, const uint64_t *trlwe, const double *trgswfft,
                              int N, int l, int Bgbit);
/* CMUXFFTwithPolynomialMulByXaiMinusOne: TFHEpp/src/detwfa.cpp:21-30 */
void orc_cmux_rot_u32(uint32_t *acc, const double *trgswfft, uint32_t a, int N, int l, int Bgbit);
void orc_cmux_rot_u64(uint64_t *acc, const double *trgswfft, uint32_t a, int N, int l, int Bgbit);

/* ---- blind rotation / PBS / sample extract ------------------------------ */
/* BlindRotate (polynomial test vector): TFHEpp/include/gatebootstrapping.hpp:15-43.
 * tlwe0 is a lvl0 TLWE of n0+1 words (32-bit torus). bkfft = n0 TRGSWFFTs. */
void orc_blind_rotate_u32(uint32_t *acc, const uint32_t *tlwe0, int n0, const double *bkfft,
                          const uint32_t *tv, int N, int nbit, int l, int Bgbit);
void orc_blind_rotate_u64(uint64_t *acc, const uint32_t *tlwe0, int n0, const double *bkfft,
                          const uint64_t *tv, int N, int nbit, int l, int Bgbit);
/* SampleExtractIndex: TFHEpp/src/trlwe.cpp:109-118 */
void orc_sample_extract_u32(uint32_t *tlwe, const uint32_t *trlwe, int index, int N);
void orc_sample_extract_u64(uint64_t *tlwe, const uint64_t *trlwe, int index, int N);
/* GateBootstrappingTLWE2TLWEFFT: TFHEpp/src/gatebootstrapping.cpp:14-23 */
void orc_pbs_u32(uint32_t *out, const uint32_t *tlwe0, int n0, const double *bkfft,
                 const uint32_t *tv, int N, int nbit, int l, int Bgbit);
void orc_pbs_u64(uint64_t *out, const uint32_t *tlwe0, int n0, const double *bkfft,
                 const uint64_t *tv, int N, int nbit, int l, int Bgbit);
/* GateBootstrappingManyLUT<P, num_out>: TFHEpp/include/gatebootstrapping.hpp:119-130.
 * out = num_out TLWEs of N+1 words (sample extraction at 0 .. num_out-1). */
void orc_pbs_many_u32(uint32_t *out, const uint32_t *tlwe0, int n0, const double *bkfft,
                      const uint32_t *tv, int N, int nbit, int l, int Bgbit, int num_out);
void orc_pbs_many_u64(uint64_t *out, const uint32_t *tlwe0, int n0, const double *bkfft,
                      const uint64_t *tv, int N, int nbit, int l, int Bgbit, int num_out);

/* ---- identity key switching --------------------------------------------- */
/* IdentityKeySwitch: TFHEpp/src/keyswitch.cpp:12-48, target torus 32-bit. */
void orc_iks_32to32(uint32_t *out, const uint32_t *in, const uint32_t *ksk, int dn, int tn,
                    int t, int basebit);
void orc_iks_64to32(uint32_t *out, const uint64_t *in, const uint32_t *ksk, int dn, int tn,
                    int t, int basebit);

/* ---- encryption readout ------------------------------------------------- */
/* phase = b - <a,s>: tlweSymDecrypt, TFHEpp/src/tlwe.cpp:60-70 */
uint32_t orc_phase_u32(const uint32_t *c, const uint32_t *key, int n);
uint64_t orc_phase_u64(const uint64_t *c, const uint64_t *key, int n);
/* TRLWE phase (trlweSymDecrypt, TFHEpp/src/trlwe.cpp:61-77), exact product. */
void orc_trlwe_phase_u32(uint32_t *phase, const uint32_t *c, const uint32_t *key, int N);
void orc_trlwe_phase_u64(uint64_t *phase, const uint64_t *c, const uint64_t *key, int N);

/* ---- gate (config-1 HomNAND) -------------------------------------------- */
/* HomNAND<lvl1param>: TFHEpp/src/gate.cpp:50-65 + gatebootstrapping.hpp:150-158:
 * c = -ca - cb + mu; IKS lvl10; PBS with test vector mu. */
void orc_hom_nand_lvl1(uint32_t *out, const uint32_t *ca, const uint32_t *cb, const uint32_t *ksk10,
                       int n0, const double *bkfft);

#ifdef __cplusplus
}
#endif
#endif
