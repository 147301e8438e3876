// ref_harness.cpp -- extern "C" entry points into the reference TFHEpp CPU
// library, compiled from the reference's own sources (see oracle/Makefile).
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle (tfhe_oracle.c) and to
// write golden vectors under tests/golden/.  Keys come from the caller (the
// reference's own key generator cannot be seeded: TFHEpp/include/utils.hpp:17-22),
// so every function takes raw arrays laid out as TFHEpp's std::array
// aggregates (TFHEpp/include/params.hpp:31-99) and reinterprets them.
//
// TFHEpp is compiled with its default parameter set params/128bit.hpp, i.e.
// lvl0 n = 672; the n = 636 set is checked through the oracle only.
#include <cstring>
#include <fstream>
#include <memory>
#include <tfhe++.hpp>

#include "HEDB/comparison/tfhepp_utils.h"

using namespace TFHEpp;

namespace {
template <class A>
const A &as(const void *p)
{
    return *reinterpret_cast<const A *>(p);
}
}  // namespace

namespace {
template <class P, uint32_t K>
void many_lut(typename P::targetP::T *out, const uint32_t *tlwe0, const double *bkfft,
              const typename P::targetP::T *tv)
{
    std::array<TLWE<typename P::targetP>, K> r;
    GateBootstrappingManyLUT<P, K>(r, as<TLWE<lvl0param>>(tlwe0),
                                   as<BootstrappingKeyFFT<P>>(bkfft),
                                   as<Polynomial<typename P::targetP>>(tv));
    std::memcpy(out, r.data(), sizeof(r));
}

template <class P>
int many_lut_n(typename P::targetP::T *out, const uint32_t *tlwe0, const double *bkfft,
               const typename P::targetP::T *tv, int num_out)
{
    switch (num_out) {
    case 1: many_lut<P, 1>(out, tlwe0, bkfft, tv); return 0;
    case 2: many_lut<P, 2>(out, tlwe0, bkfft, tv); return 0;
    case 3: many_lut<P, 3>(out, tlwe0, bkfft, tv); return 0;
    case 4: many_lut<P, 4>(out, tlwe0, bkfft, tv); return 0;
    case 8: many_lut<P, 8>(out, tlwe0, bkfft, tv); return 0;
    default: return -1;
    }
}
}  // namespace

extern "C" {

int ref_lvl0_n() { return lvl0param::n; }

void ref_twist_ifft_lvl1(double *fd, const uint32_t *a)
{
    PolynomialInFD<lvl1param> r;
    TwistIFFT<lvl1param>(r, as<Polynomial<lvl1param>>(a));
    std::memcpy(fd, r.data(), sizeof(r));
}

void ref_twist_fft_lvl1(uint32_t *res, const double *fd)
{
    Polynomial<lvl1param> r;
    TwistFFT<lvl1param>(r, as<PolynomialInFD<lvl1param>>(fd));
    std::memcpy(res, r.data(), sizeof(r));
}

void ref_twist_ifft_lvl2(double *fd, const uint64_t *a)
{
    PolynomialInFD<lvl2param> r;
    TwistIFFT<lvl2param>(r, as<Polynomial<lvl2param>>(a));
    std::memcpy(fd, r.data(), sizeof(r));
}

void ref_twist_fft_lvl2(uint64_t *res, const double *fd)
{
    Polynomial<lvl2param> r;
    TwistFFT<lvl2param>(r, as<PolynomialInFD<lvl2param>>(fd));
    std::memcpy(res, r.data(), sizeof(r));
}

void ref_polymul_lvl1(uint32_t *res, const uint32_t *a, const uint32_t *b)
{
    Polynomial<lvl1param> r;
    PolyMul<lvl1param>(r, as<Polynomial<lvl1param>>(a), as<Polynomial<lvl1param>>(b));
    std::memcpy(res, r.data(), sizeof(r));
}

void ref_decompose_lvl1(uint32_t *dec, const uint32_t *poly, int digit)
{
    DecomposedPolynomial<lvl1param> r;
    DecompositionPolynomial<lvl1param>(r, as<Polynomial<lvl1param>>(poly), digit);
    std::memcpy(dec, r.data(), sizeof(r));
}

void ref_decompose_lvl2(uint64_t *dec, const uint64_t *poly, int digit)
{
    DecomposedPolynomial<lvl2param> r;
    DecompositionPolynomial<lvl2param>(r, as<Polynomial<lvl2param>>(poly), digit);
    std::memcpy(dec, r.data(), sizeof(r));
}

void ref_trgsw_to_fft_lvl1(double *out, const uint32_t *trgsw)
{
    auto r = std::make_unique<TRGSWFFT<lvl1param>>();
    *r = ApplyFFT2trgsw<lvl1param>(as<TRGSW<lvl1param>>(trgsw));
    std::memcpy(out, r->data(), sizeof(*r));
}

void ref_trgsw_to_fft_lvl2(double *out, const uint64_t *trgsw)
{
    auto r = std::make_unique<TRGSWFFT<lvl2param>>();
    *r = ApplyFFT2trgsw<lvl2param>(as<TRGSW<lvl2param>>(trgsw));
    std::memcpy(out, r->data(), sizeof(*r));
}

void ref_external_product_lvl1(uint32_t *res, const uint32_t *trlwe, const double *trgswfft)
{
    TRLWE<lvl1param> r;
    trgswfftExternalProduct<lvl1param>(r, as<TRLWE<lvl1param>>(trlwe),
                                       as<TRGSWFFT<lvl1param>>(trgswfft));
    std::memcpy(res, r.data(), sizeof(r));
}

void ref_external_product_lvl2(uint64_t *res, const uint64_t *trlwe, const double *trgswfft)
{
    TRLWE<lvl2param> r;
    trgswfftExternalProduct<lvl2param>(r, as<TRLWE<lvl2param>>(trlwe),
                                       as<TRGSWFFT<lvl2param>>(trgswfft));
    std::memcpy(res, r.data(), sizeof(r));
}

void ref_blind_rotate_lvl01(uint32_t *acc, const uint32_t *tlwe0, const double *bkfft,
                            const uint32_t *tv)
{
    TRLWE<lvl1param> r;
    BlindRotate<lvl01param>(r, as<TLWE<lvl0param>>(tlwe0),
                            as<BootstrappingKeyFFT<lvl01param>>(bkfft),
                            as<Polynomial<lvl1param>>(tv));
    std::memcpy(acc, r.data(), sizeof(r));
}

void ref_pbs_lvl01(uint32_t *out, const uint32_t *tlwe0, const double *bkfft, const uint32_t *tv)
{
    TLWE<lvl1param> r;
    GateBootstrappingTLWE2TLWEFFT<lvl01param>(r, as<TLWE<lvl0param>>(tlwe0),
                                              as<BootstrappingKeyFFT<lvl01param>>(bkfft),
                                              as<Polynomial<lvl1param>>(tv));
    std::memcpy(out, r.data(), sizeof(r));
}

void ref_pbs_lvl02(uint64_t *out, const uint32_t *tlwe0, const double *bkfft, const uint64_t *tv)
{
    TLWE<lvl2param> r;
    GateBootstrappingTLWE2TLWEFFT<lvl02param>(r, as<TLWE<lvl0param>>(tlwe0),
                                              as<BootstrappingKeyFFT<lvl02param>>(bkfft),
                                              as<Polynomial<lvl2param>>(tv));
    std::memcpy(out, r.data(), sizeof(r));
}

/* GateBootstrappingManyLUT<lvl0Xparam, num_out> for num_out in {1, 2, 3, 4, 8}
 * (gatebootstrapping.hpp:119-130); out = num_out TLWEs.  Returns -1 for
 * other counts. */
int ref_pbs_many_lvl01(uint32_t *out, const uint32_t *tlwe0, const double *bkfft,
                       const uint32_t *tv, int num_out)
{
    return many_lut_n<lvl01param>(out, tlwe0, bkfft, tv, num_out);
}

int ref_pbs_many_lvl02(uint64_t *out, const uint32_t *tlwe0, const double *bkfft,
                       const uint64_t *tv, int num_out)
{
    return many_lut_n<lvl02param>(out, tlwe0, bkfft, tv, num_out);
}

void ref_sample_extract_lvl1(uint32_t *tlwe, const uint32_t *trlwe, int index)
{
    TLWE<lvl1param> r;
    SampleExtractIndex<lvl1param>(r, as<TRLWE<lvl1param>>(trlwe), index);
    std::memcpy(tlwe, r.data(), sizeof(r));
}

void ref_sample_extract_lvl2(uint64_t *tlwe, const uint64_t *trlwe, int index)
{
    TLWE<lvl2param> r;
    SampleExtractIndex<lvl2param>(r, as<TRLWE<lvl2param>>(trlwe), index);
    std::memcpy(tlwe, r.data(), sizeof(r));
}

void ref_iks_lvl10(uint32_t *out, const uint32_t *in, const uint32_t *ksk)
{
    TLWE<lvl0param> r;
    IdentityKeySwitch<lvl10param>(r, as<TLWE<lvl1param>>(in),
                                  as<KeySwitchingKey<lvl10param>>(ksk));
    std::memcpy(out, r.data(), sizeof(r));
}

void ref_iks_lvl20(uint32_t *out, const uint64_t *in, const uint32_t *ksk)
{
    TLWE<lvl0param> r;
    IdentityKeySwitch<lvl20param>(r, as<TLWE<lvl2param>>(in),
                                  as<KeySwitchingKey<lvl20param>>(ksk));
    std::memcpy(out, r.data(), sizeof(r));
}

void ref_iks_lvl21(uint32_t *out, const uint64_t *in, const uint32_t *ksk)
{
    TLWE<lvl1param> r;
    IdentityKeySwitch<lvl21param>(r, as<TLWE<lvl2param>>(in),
                                  as<KeySwitchingKey<lvl21param>>(ksk));
    std::memcpy(out, r.data(), sizeof(r));
}

// EvalKey adopts the caller's key arrays for the duration of one call and
// releases them before it is destroyed (it never owns them).
namespace {
struct BorrowedKey {
    EvalKey ek;
    BorrowedKey(const double *bkfft, const uint32_t *ksk10)
    {
        ek.bkfftlvl01.reset(reinterpret_cast<BootstrappingKeyFFT<lvl01param> *>(
            const_cast<double *>(bkfft)));
        ek.iksklvl10.reset(reinterpret_cast<KeySwitchingKey<lvl10param> *>(
            const_cast<uint32_t *>(ksk10)));
    }
    ~BorrowedKey()
    {
        ek.bkfftlvl01.release();
        ek.iksklvl10.release();
    }
};
}  // namespace

void ref_hom_nand_lvl1(uint32_t *out, const uint32_t *ca, const uint32_t *cb,
                       const double *bkfft, const uint32_t *ksk10)
{
    BorrowedKey k(bkfft, ksk10);
    TLWE<lvl1param> r;
    HomNAND<lvl1param>(r, as<TLWE<lvl1param>>(ca), as<TLWE<lvl1param>>(cb), k.ek);
    std::memcpy(out, r.data(), sizeof(r));
}

void ref_msb_gate_lvl1(uint32_t *out, const uint32_t *in, const double *bkfft,
                       const uint32_t *ksk10, int result_type)
{
    BorrowedKey k(bkfft, ksk10);
    TLWE<lvl1param> r;
    MSBGateBootstrapping(r, as<TLWE<lvl1param>>(in), k.ek, result_type != 0);
    std::memcpy(out, r.data(), sizeof(r));
}

void ref_ide_gate_lvl1(uint32_t *out, const uint32_t *in, uint32_t scale_bits,
                       const double *bkfft, const uint32_t *ksk10)
{
    BorrowedKey k(bkfft, ksk10);
    TLWE<lvl1param> r;
    IdeGateBootstrapping(r, as<TLWE<lvl1param>>(in), scale_bits, k.ek);
    std::memcpy(out, r.data(), sizeof(r));
}

}  // extern "C"

// ---- key files ------------------------------------------------------------
// The reference's own archive code: HEDB::generate_*_key / load_*_key
// (src/HEDB/utils/serialize.cpp:7-320) run cereal's PortableBinary archive
// over SecretKey::serialize and EvalKey::serialize_<member>.  `kind` numbers
// the member files as enum fhe_keyfile in include/cufhedb_amd.h does
// (1 bklvl01, 2 bklvl02, 3 bkfftlvl01, 4 bkfftlvl02, 5 iksklvl10,
// 6 iksklvl20, 7 iksklvl21).
namespace {
template <class M>
int member_io(std::unique_ptr<M> &slot, void *data, bool save, bool be, EvalKey &ek,
              void (EvalKey::*ser_out)(cereal::PortableBinaryOutputArchive &),
              void (EvalKey::*ser_in)(cereal::PortableBinaryInputArchive &), const char *path)
{
    if (save) {
        slot.reset(reinterpret_cast<M *>(data));
        {
            std::ofstream ofs{path, std::ios::binary};
            cereal::PortableBinaryOutputArchive ar(
                ofs, be ? cereal::PortableBinaryOutputArchive::Options::BigEndian()
                        : cereal::PortableBinaryOutputArchive::Options::LittleEndian());
            (ek.*ser_out)(ar);
        }
        slot.release();
        return 0;
    }
    {
        std::ifstream ifs{path, std::ios::binary};
        cereal::PortableBinaryInputArchive ar(ifs);
        (ek.*ser_in)(ar);
    }
    if (!slot) return 1;
    std::memcpy(data, slot.get(), sizeof(M));
    return 0;
}

int keyfile_io(const char *path, int kind, void *data, bool save, bool be)
{
    EvalKey ek;
    using O = cereal::PortableBinaryOutputArchive;
    using I = cereal::PortableBinaryInputArchive;
    switch (kind) {
    case 1: return member_io(ek.bklvl01, data, save, be, ek, &EvalKey::serialize_bklvl01<O>,
                             &EvalKey::serialize_bklvl01<I>, path);
    case 2: return member_io(ek.bklvl02, data, save, be, ek, &EvalKey::serialize_bklvl02<O>,
                             &EvalKey::serialize_bklvl02<I>, path);
    case 3: return member_io(ek.bkfftlvl01, data, save, be, ek, &EvalKey::serialize_bkfftlvl01<O>,
                             &EvalKey::serialize_bkfftlvl01<I>, path);
    case 4: return member_io(ek.bkfftlvl02, data, save, be, ek, &EvalKey::serialize_bkfftlvl02<O>,
                             &EvalKey::serialize_bkfftlvl02<I>, path);
    case 5: return member_io(ek.iksklvl10, data, save, be, ek, &EvalKey::serialize_iksklvl10<O>,
                             &EvalKey::serialize_iksklvl10<I>, path);
    case 6: return member_io(ek.iksklvl20, data, save, be, ek, &EvalKey::serialize_iksklvl20<O>,
                             &EvalKey::serialize_iksklvl20<I>, path);
    case 7: return member_io(ek.iksklvl21, data, save, be, ek, &EvalKey::serialize_iksklvl21<O>,
                             &EvalKey::serialize_iksklvl21<I>, path);
    default: return 2;
    }
}
}  // namespace

extern "C" {

// big_endian != 0 writes the archive with cereal's big-endian option
int ref_keyfile_write(const char *path, int kind, const void *data, int big_endian)
{
    return keyfile_io(path, kind, const_cast<void *>(data), true, big_endian != 0);
}

int ref_keyfile_read(const char *path, int kind, void *out)
{
    return keyfile_io(path, kind, out, false, false);
}

int ref_secret_write(const char *path, const uint32_t *s0, const uint32_t *s1, const uint64_t *s2,
                     int big_endian)
{
    auto sk = std::make_unique<SecretKey>();
    std::memcpy(sk->key.lvl0.data(), s0, sizeof(sk->key.lvl0));
    std::memcpy(sk->key.lvl1.data(), s1, sizeof(sk->key.lvl1));
    std::memcpy(sk->key.lvl2.data(), s2, sizeof(sk->key.lvl2));
    std::ofstream ofs{path, std::ios::binary};
    cereal::PortableBinaryOutputArchive ar(
        ofs, big_endian ? cereal::PortableBinaryOutputArchive::Options::BigEndian()
                        : cereal::PortableBinaryOutputArchive::Options::LittleEndian());
    sk->serialize(ar);
    return 0;
}

int ref_secret_read(const char *path, uint32_t *s0, uint32_t *s1, uint64_t *s2, uint32_t *n0)
{
    auto sk = std::make_unique<SecretKey>();
    {
        std::ifstream ifs{path, std::ios::binary};
        cereal::PortableBinaryInputArchive ar(ifs);
        sk->serialize(ar);
    }
    std::memcpy(s0, sk->key.lvl0.data(), sizeof(sk->key.lvl0));
    std::memcpy(s1, sk->key.lvl1.data(), sizeof(sk->key.lvl1));
    std::memcpy(s2, sk->key.lvl2.data(), sizeof(sk->key.lvl2));
    *n0 = sk->params.lvl0.n;
    return 0;
}

}  // extern "C"
