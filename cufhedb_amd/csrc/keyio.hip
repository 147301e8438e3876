// keyio.hip -- the reference's key material files: read, write, and stream
// into a device context.
//
// TFHEpp keys are written by HEDB's generate_*_key and read back by its
// load_*_key (src/HEDB/utils/serialize.cpp:7-320) through cereal's
// PortableBinary archive:
//   * one leading byte, 1 when the stream is little-endian
//     (cereal/archives/portable_binary.hpp:119-126); every arithmetic value
//     follows in that byte order, back to back, with no padding or lengths;
//     a std::array of arithmetic type is one blob (cereal/types/array.hpp:38-56)
//     and nested arrays are their elements in order;
//   * secret.key = SecretKey::serialize (TFHEpp/include/key.hpp:29-35):
//     lvl0 key [n0] u32, lvl1 key [1024] u32, lvl2 key [2048] u64, lweParams;
//   * <member>.key = EvalKey::serialize_<member>
//     (TFHEpp/include/cloudkey.hpp:164-228): lweParams, then the
//     std::unique_ptr member as one byte (1 = present,
//     cereal/types/memory.hpp:346-362) followed by the array.
// lweParams::serialize (TFHEpp/include/lweParams.hpp:181-190) writes 27
// fields (136 bytes; lvl1.Bg and lvl2.Bg are not stored).
// The fastR dumps (test/HE3_Query/fastR.h:6-74) are raw and headerless: lvl1
// key, BootstrappingKeyFFT<lvl01param>, KeySwitchingKey<lvl10param>, host
// byte order.
//
// Large members (key-switching keys, up to 11.3 GB) are streamed from the
// file into device memory through two pinned staging chunks, so the host
// never holds a whole key.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ctx.hpp"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "key files are written little-endian");

namespace fhe {
namespace {

constexpr uint64_t kParamsBytes = 136;

struct Member {
    int elem;        // bytes per element (the byte-swap unit)
    uint64_t count;  // elements
};

// Payload of an EvalKey member file for lvl0 dimension n0.
bool member_of(int kind, uint32_t n0, Member &m)
{
    const uint64_t n = n0;
    switch (kind) {
    case FHE_KEYFILE_BKLVL01: m = {4, n * Lvl1::rows * 2 * Lvl1::N}; return true;
    case FHE_KEYFILE_BKLVL02: m = {8, n * Lvl2::rows * 2 * Lvl2::N}; return true;
    case FHE_KEYFILE_BKFFTLVL01: m = {8, n * Lvl1::rows * 2 * Lvl1::N}; return true;
    case FHE_KEYFILE_BKFFTLVL02: m = {8, n * Lvl2::rows * 2 * Lvl2::N}; return true;
    case FHE_KEYFILE_IKSKLVL10:
        m = {4, uint64_t(Lvl1::N) * KS10::t * ((1u << KS10::basebit) - 1) * (n + 1)};
        return true;
    case FHE_KEYFILE_IKSKLVL20:
        m = {4, uint64_t(Lvl2::N) * KS20::t * ((1u << KS20::basebit) - 1) * (n + 1)};
        return true;
    case FHE_KEYFILE_IKSKLVL21:
        m = {4, uint64_t(Lvl2::N) * KS21::t * ((1u << KS21::basebit) - 1) * (Lvl1::N + 1)};
        return true;
    default: return false;
    }
}

int ksk_kind_of(int kind)
{
    return kind == FHE_KEYFILE_IKSKLVL10 ? 10 : kind == FHE_KEYFILE_IKSKLVL20 ? 20
         : kind == FHE_KEYFILE_IKSKLVL21 ? 21 : 0;
}

void swap_elems(void *p, uint64_t bytes, int elem)
{
    auto *b = static_cast<uint8_t *>(p);
    for (uint64_t i = 0; i + elem <= bytes; i += elem) std::reverse(b + i, b + i + elem);
}

class File {
  public:
    ~File()
    {
        if (f_) std::fclose(f_);
    }
    int open(const char *path, bool write)
    {
        f_ = std::fopen(path, write ? "wb" : "rb");
        if (!f_) {
            set_error("cannot open key file '%s' for %s", path, write ? "writing" : "reading");
            return FHE_ERR_ARG;
        }
        path_ = path;
        if (!write) {
            if (fseeko(f_, 0, SEEK_END) != 0) return io_error("seek");
            size_ = uint64_t(ftello(f_));
            if (fseeko(f_, 0, SEEK_SET) != 0) return io_error("seek");
        }
        return FHE_OK;
    }
    uint64_t size() const { return size_; }
    bool swap() const { return swap_; }
    int read_flag()
    {
        uint8_t le = 0;
        if (int rc = read(&le, 1, 1)) return rc;
        if (le > 1) {
            set_error("key file '%s': first byte %u is not a cereal endianness flag", path_, le);
            return FHE_ERR_ARG;
        }
        swap_ = le == 0;
        return FHE_OK;
    }
    int read(void *dst, uint64_t bytes, int elem)
    {
        if (bytes && std::fread(dst, 1, bytes, f_) != bytes) return io_error("short read");
        if (swap_ && elem > 1) swap_elems(dst, bytes, elem);
        return FHE_OK;
    }
    template <class T>
    int get(T &v)
    {
        return read(&v, sizeof(T), int(sizeof(T)));
    }
    int skip(uint64_t bytes)
    {
        if (fseeko(f_, off_t(bytes), SEEK_CUR) != 0) return io_error("seek");
        return FHE_OK;
    }
    int write(const void *src, uint64_t bytes)
    {
        if (bytes && std::fwrite(src, 1, bytes, f_) != bytes) return io_error("short write");
        return FHE_OK;
    }
    template <class T>
    int put(T v)
    {
        return write(&v, sizeof(T));
    }
    int close()
    {
        FILE *f = f_;
        f_ = nullptr;
        if (f && std::fclose(f) != 0) {
            set_error("key file '%s': close failed", path_);
            return FHE_ERR_ARG;
        }
        return FHE_OK;
    }
    int io_error(const char *what)
    {
        set_error("key file '%s': %s", path_, what);
        return FHE_ERR_ARG;
    }

  private:
    FILE *f_ = nullptr;
    const char *path_ = "";
    uint64_t size_ = 0;
    bool swap_ = false;
};

#define KIO_TRY(expr)                       \
    do {                                    \
        if (int _rc = (expr)) return _rc;   \
    } while (0)

int read_params(File &f, fhe_keyfile_params &q)
{
    KIO_TRY(f.get(q.n0));
    KIO_TRY(f.get(q.alpha0));
    KIO_TRY(f.get(q.approx_bit0));
    KIO_TRY(f.get(q.nbit1));
    KIO_TRY(f.get(q.n1));
    KIO_TRY(f.get(q.l1));
    KIO_TRY(f.get(q.bgbit1));
    KIO_TRY(f.get(q.alpha1));
    KIO_TRY(f.get(q.approx_bit1));
    KIO_TRY(f.get(q.nbit2));
    KIO_TRY(f.get(q.n2));
    KIO_TRY(f.get(q.l2));
    KIO_TRY(f.get(q.bgbit2));
    KIO_TRY(f.get(q.alpha2));
    KIO_TRY(f.get(q.approx_bit2));
    KIO_TRY(f.get(q.t10));
    KIO_TRY(f.get(q.basebit10));
    KIO_TRY(f.get(q.alpha10));
    KIO_TRY(f.get(q.t20));
    KIO_TRY(f.get(q.basebit20));
    KIO_TRY(f.get(q.alpha20));
    KIO_TRY(f.get(q.t21));
    KIO_TRY(f.get(q.basebit21));
    KIO_TRY(f.get(q.alpha21));
    KIO_TRY(f.get(q.t22));
    KIO_TRY(f.get(q.basebit22));
    KIO_TRY(f.get(q.alpha22));
    return FHE_OK;
}

int write_params(File &f, const fhe_keyfile_params &q)
{
    KIO_TRY(f.put(q.n0));
    KIO_TRY(f.put(q.alpha0));
    KIO_TRY(f.put(q.approx_bit0));
    KIO_TRY(f.put(q.nbit1));
    KIO_TRY(f.put(q.n1));
    KIO_TRY(f.put(q.l1));
    KIO_TRY(f.put(q.bgbit1));
    KIO_TRY(f.put(q.alpha1));
    KIO_TRY(f.put(q.approx_bit1));
    KIO_TRY(f.put(q.nbit2));
    KIO_TRY(f.put(q.n2));
    KIO_TRY(f.put(q.l2));
    KIO_TRY(f.put(q.bgbit2));
    KIO_TRY(f.put(q.alpha2));
    KIO_TRY(f.put(q.approx_bit2));
    KIO_TRY(f.put(q.t10));
    KIO_TRY(f.put(q.basebit10));
    KIO_TRY(f.put(q.alpha10));
    KIO_TRY(f.put(q.t20));
    KIO_TRY(f.put(q.basebit20));
    KIO_TRY(f.put(q.alpha20));
    KIO_TRY(f.put(q.t21));
    KIO_TRY(f.put(q.basebit21));
    KIO_TRY(f.put(q.alpha21));
    KIO_TRY(f.put(q.t22));
    KIO_TRY(f.put(q.basebit22));
    KIO_TRY(f.put(q.alpha22));
    return FHE_OK;
}

// The ring, gadget and key-switch shapes this build computes with; a file
// whose parameters differ cannot be used (its arrays would not match).
int check_structure(const fhe_keyfile_params &q, const char *path)
{
    const bool ok = q.approx_bit0 == 32 && q.nbit1 == 10 && q.n1 == Lvl1::N &&
                    q.l1 == uint32_t(Lvl1::l) && q.bgbit1 == uint32_t(Lvl1::Bgbit) &&
                    q.approx_bit1 == 32 && q.nbit2 == 11 && q.n2 == Lvl2::N &&
                    q.l2 == uint32_t(Lvl2::l) && q.bgbit2 == uint32_t(Lvl2::Bgbit) &&
                    q.approx_bit2 == 64 && q.t10 == uint32_t(KS10::t) &&
                    q.basebit10 == uint32_t(KS10::basebit) && q.t20 == uint32_t(KS20::t) &&
                    q.basebit20 == uint32_t(KS20::basebit) && q.t21 == uint32_t(KS21::t) &&
                    q.basebit21 == uint32_t(KS21::basebit);
    if (!ok) {
        set_error("key file '%s': parameter block differs from the 128-bit parameter set "
                  "(lvl1 N=1024 l=3 Bgbit=6, lvl2 N=2048 l=4 Bgbit=9, KS t/basebit 2/10, 2/10, "
                  "10/3)", path);
        return FHE_ERR_UNSUPPORTED;
    }
    if (q.n0 < 1 || q.n0 > uint32_t(kMaxN0)) {
        set_error("key file '%s': lvl0 dimension %u outside [1, %d]", path, q.n0, kMaxN0);
        return FHE_ERR_UNSUPPORTED;
    }
    return FHE_OK;
}

// Opens a member file, parses flag + params + presence byte and checks the
// file size against the payload the parameters imply.
int open_member(File &f, const char *path, int kind, fhe_keyfile_params &q, Member &m)
{
    KIO_TRY(f.open(path, false));
    KIO_TRY(f.read_flag());
    KIO_TRY(read_params(f, q));
    KIO_TRY(check_structure(q, path));
    uint8_t present = 0;
    KIO_TRY(f.read(&present, 1, 1));
    if (present != 1) {
        set_error("key file '%s' holds no key (null member, presence byte %u)", path, present);
        return FHE_ERR_ARG;
    }
    member_of(kind, q.n0, m);
    const uint64_t want = 1 + kParamsBytes + 1 + m.count * m.elem;
    if (f.size() != want) {
        set_error("key file '%s': %llu bytes, expected %llu for this key kind at n0 = %u", path,
                  (unsigned long long)f.size(), (unsigned long long)want, q.n0);
        return FHE_ERR_ARG;
    }
    return FHE_OK;
}

// File -> device memory through two pinned staging chunks (disk read of one
// chunk overlaps the upload of the other).
int stream_to_device(fhe_ctx *ctx, File &f, void *dst_dev, uint64_t bytes, int elem)
{
    constexpr uint64_t kChunk = uint64_t(64) << 20;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    int rc = FHE_OK;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipHostMalloc(&pin[i], kChunk, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
    }
    auto *dst = static_cast<uint8_t *>(dst_dev);
    for (uint64_t off = 0, i = 0; e == hipSuccess && rc == FHE_OK && off < bytes; i ^= 1) {
        const uint64_t n = std::min(kChunk, bytes - off);
        if (pending[i]) e = hipEventSynchronize(done[i]);
        if (e != hipSuccess) break;
        rc = f.read(pin[i], n, elem);
        if (rc != FHE_OK) break;
        e = hipMemcpyAsync(dst + off, pin[i], n, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(done[i], ctx->stream);
        pending[i] = e == hipSuccess;
        off += n;
    }
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    for (int i = 0; i < 2; i++) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (pin[i]) (void)hipHostFree(pin[i]);
    }
    if (rc != FHE_OK) return rc;
    HIP_TRY(e, "key upload");
    return FHE_OK;
}

int load_ksk_stream(fhe_ctx *ctx, File &f, int ks_kind, uint64_t bytes)
{
    const int slot = ksk_slot(ks_kind);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    if (int rc = alloc_ksk(ctx, ks_kind)) return rc;
    return stream_to_device(ctx, f, ctx->ksk[slot], bytes, 4);
}

int load_bk_from_file(fhe_ctx *ctx, File &f, int level, int format, uint64_t bytes, int elem)
{
    std::vector<uint8_t> buf;
    try {
        buf.resize(bytes);
    } catch (...) {
        set_error("host allocation of %llu bytes for the bootstrapping key failed",
                  (unsigned long long)bytes);
        return FHE_ERR_ALLOC;
    }
    KIO_TRY(f.read(buf.data(), bytes, elem));
    return fhe_load_bk(ctx, level, buf.data(), format);
}

uint64_t fastr_bytes(uint32_t n0, uint64_t &bk_off, uint64_t &ksk_off, uint64_t &ksk_bytes)
{
    Member bk, ks;
    member_of(FHE_KEYFILE_BKFFTLVL01, n0, bk);
    member_of(FHE_KEYFILE_IKSKLVL10, n0, ks);
    bk_off = uint64_t(Lvl1::N) * 4;
    ksk_off = bk_off + bk.count * 8;
    ksk_bytes = ks.count * 4;
    return ksk_off + ksk_bytes;
}

}  // namespace
}  // namespace fhe

using namespace fhe;

extern "C" {

int fhe_keyfile_params_of(const fhe_params *p, fhe_keyfile_params *out)
{
    FHE_CHECK_ARG(p && out, "fhe_keyfile_params_of: null argument");
    // TFHEpp/include/params/128bit.hpp:7-97 with this context's n0 and alphas;
    // the key-switch noises follow the reference (lvl10/lvl20 = lvl0 alpha,
    // lvl21 = lvl1 alpha, lvl22 = lvl2 alpha)
    fhe_keyfile_params q{};
    q.n0 = p->n0;
    q.alpha0 = p->alpha0;
    q.approx_bit0 = 32;
    q.nbit1 = 10;
    q.n1 = Lvl1::N;
    q.l1 = Lvl1::l;
    q.bgbit1 = Lvl1::Bgbit;
    q.alpha1 = p->alpha1;
    q.approx_bit1 = 32;
    q.nbit2 = 11;
    q.n2 = Lvl2::N;
    q.l2 = Lvl2::l;
    q.bgbit2 = Lvl2::Bgbit;
    q.alpha2 = p->alpha2;
    q.approx_bit2 = 64;
    q.t10 = KS10::t;
    q.basebit10 = KS10::basebit;
    q.alpha10 = p->alpha0;
    q.t20 = KS20::t;
    q.basebit20 = KS20::basebit;
    q.alpha20 = p->alpha0;
    q.t21 = KS21::t;
    q.basebit21 = KS21::basebit;
    q.alpha21 = p->alpha1;
    q.t22 = 8;
    q.basebit22 = 4;
    q.alpha22 = p->alpha2;
    *out = q;
    return FHE_OK;
}

uint64_t fhe_keyfile_payload_bytes(int kind, uint32_t n0)
{
    Member m;
    if (kind == FHE_KEYFILE_SECRET) return uint64_t(n0) * 4 + Lvl1::N * 4 + Lvl2::N * 8;
    if (!member_of(kind, n0, m)) return 0;
    return m.count * m.elem;
}

int fhe_keyfile_info(const char *path, int kind, fhe_keyfile_params *params,
                     uint64_t *payload_bytes)
{
    FHE_CHECK_ARG(path, "fhe_keyfile_info: null path");
    fhe_keyfile_params q{};
    File f;
    if (kind == FHE_KEYFILE_SECRET) {
        KIO_TRY(f.open(path, false));
        const uint64_t fixed = 1 + Lvl1::N * 4 + Lvl2::N * 8 + kParamsBytes;
        if (f.size() < fixed + 4 || (f.size() - fixed) % 4 != 0) {
            set_error("key file '%s': %llu bytes is not a secret-key archive", path,
                      (unsigned long long)f.size());
            return FHE_ERR_ARG;
        }
        const uint64_t n0 = (f.size() - fixed) / 4;
        KIO_TRY(f.read_flag());
        KIO_TRY(f.skip(n0 * 4 + Lvl1::N * 4 + Lvl2::N * 8));
        KIO_TRY(read_params(f, q));
        KIO_TRY(check_structure(q, path));
        if (q.n0 != n0) {
            set_error("key file '%s': lvl0 key of %llu words but parameter n0 = %u", path,
                      (unsigned long long)n0, q.n0);
            return FHE_ERR_ARG;
        }
        if (payload_bytes) *payload_bytes = fhe_keyfile_payload_bytes(kind, q.n0);
    } else {
        Member m;
        FHE_CHECK_ARG(member_of(kind, 1, m), "fhe_keyfile_info: unknown key file kind %d", kind);
        KIO_TRY(open_member(f, path, kind, q, m));
        if (payload_bytes) *payload_bytes = m.count * m.elem;
    }
    if (params) *params = q;
    return FHE_OK;
}

int fhe_keyfile_read(const char *path, int kind, void *dst, uint64_t dst_bytes)
{
    FHE_CHECK_ARG(path && dst, "fhe_keyfile_read: null argument");
    Member m;
    FHE_CHECK_ARG(kind != FHE_KEYFILE_SECRET && member_of(kind, 1, m),
                  "fhe_keyfile_read: kind %d is not a key member (use fhe_keyfile_read_secret "
                  "for secret keys)", kind);
    File f;
    fhe_keyfile_params q;
    KIO_TRY(open_member(f, path, kind, q, m));
    const uint64_t bytes = m.count * m.elem;
    FHE_CHECK_ARG(dst_bytes >= bytes, "fhe_keyfile_read: buffer of %llu bytes, key needs %llu",
                  (unsigned long long)dst_bytes, (unsigned long long)bytes);
    return f.read(dst, bytes, m.elem);
}

int fhe_keyfile_read_secret(const char *path, uint32_t *s0, uint32_t n0_cap, uint32_t *s1,
                            uint64_t *s2, fhe_keyfile_params *params)
{
    FHE_CHECK_ARG(path && s0 && s1 && s2, "fhe_keyfile_read_secret: null argument");
    fhe_keyfile_params q;
    KIO_TRY(fhe_keyfile_info(path, FHE_KEYFILE_SECRET, &q, nullptr));
    FHE_CHECK_ARG(n0_cap >= q.n0, "fhe_keyfile_read_secret: lvl0 key has %u words, buffer %u",
                  q.n0, n0_cap);
    File f;
    KIO_TRY(f.open(path, false));
    KIO_TRY(f.read_flag());
    KIO_TRY(f.read(s0, uint64_t(q.n0) * 4, 4));
    KIO_TRY(f.read(s1, uint64_t(Lvl1::N) * 4, 4));
    KIO_TRY(f.read(s2, uint64_t(Lvl2::N) * 8, 8));
    if (params) *params = q;
    return FHE_OK;
}

int fhe_keyfile_write(const char *path, int kind, const fhe_params *p, const void *payload)
{
    FHE_CHECK_ARG(path && p && payload, "fhe_keyfile_write: null argument");
    Member m;
    FHE_CHECK_ARG(kind != FHE_KEYFILE_SECRET && member_of(kind, p->n0, m),
                  "fhe_keyfile_write: kind %d is not a key member", kind);
    FHE_CHECK_ARG(p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_keyfile_write: bad n0");
    fhe_keyfile_params q;
    fhe_keyfile_params_of(p, &q);
    File f;
    KIO_TRY(f.open(path, true));
    KIO_TRY(f.put(uint8_t(1)));
    KIO_TRY(write_params(f, q));
    KIO_TRY(f.put(uint8_t(1)));
    KIO_TRY(f.write(payload, m.count * m.elem));
    return f.close();
}

int fhe_keyfile_write_secret(const char *path, const fhe_params *p, const uint32_t *s0,
                             const uint32_t *s1, const uint64_t *s2)
{
    FHE_CHECK_ARG(path && p && s0 && s1 && s2, "fhe_keyfile_write_secret: null argument");
    FHE_CHECK_ARG(p->n0 >= 1 && p->n0 <= uint32_t(kMaxN0), "fhe_keyfile_write_secret: bad n0");
    fhe_keyfile_params q;
    fhe_keyfile_params_of(p, &q);
    File f;
    KIO_TRY(f.open(path, true));
    KIO_TRY(f.put(uint8_t(1)));
    KIO_TRY(f.write(s0, uint64_t(p->n0) * 4));
    KIO_TRY(f.write(s1, uint64_t(Lvl1::N) * 4));
    KIO_TRY(f.write(s2, uint64_t(Lvl2::N) * 8));
    KIO_TRY(write_params(f, q));
    return f.close();
}

int fhe_keyfile_load(fhe_ctx *ctx, const char *path, int kind)
{
    FHE_CHECK_ARG(ctx && path, "fhe_keyfile_load: null argument");
    Member m;
    FHE_CHECK_ARG(kind != FHE_KEYFILE_SECRET && member_of(kind, 1, m),
                  "fhe_keyfile_load: kind %d is not an evaluation key", kind);
    File f;
    fhe_keyfile_params q;
    KIO_TRY(open_member(f, path, kind, q, m));
    if (q.n0 != ctx->p.n0) {
        set_error("key file '%s': n0 = %u, context n0 = %u", path, q.n0, ctx->p.n0);
        return FHE_ERR_ARG;
    }
    const uint64_t bytes = m.count * m.elem;
    switch (kind) {
    case FHE_KEYFILE_BKLVL01: return load_bk_from_file(ctx, f, 1, FHE_BK_INT, bytes, m.elem);
    case FHE_KEYFILE_BKLVL02: return load_bk_from_file(ctx, f, 2, FHE_BK_INT, bytes, m.elem);
    case FHE_KEYFILE_BKFFTLVL01: return load_bk_from_file(ctx, f, 1, FHE_BK_FFT, bytes, m.elem);
    case FHE_KEYFILE_BKFFTLVL02: return load_bk_from_file(ctx, f, 2, FHE_BK_FFT, bytes, m.elem);
    default: return load_ksk_stream(ctx, f, ksk_kind_of(kind), bytes);
    }
}

int fhe_fastr_write(const char *path, uint32_t n0, const uint32_t *s1, const double *bkfft,
                    const uint32_t *iksk10)
{
    FHE_CHECK_ARG(path && s1 && bkfft && iksk10, "fhe_fastr_write: null argument");
    FHE_CHECK_ARG(n0 >= 1 && n0 <= uint32_t(kMaxN0), "fhe_fastr_write: bad n0");
    uint64_t bk_off, ksk_off, ksk_bytes;
    fastr_bytes(n0, bk_off, ksk_off, ksk_bytes);
    File f;
    KIO_TRY(f.open(path, true));
    KIO_TRY(f.write(s1, bk_off));
    KIO_TRY(f.write(bkfft, ksk_off - bk_off));
    KIO_TRY(f.write(iksk10, ksk_bytes));
    return f.close();
}

int fhe_fastr_load(fhe_ctx *ctx, const char *path, uint32_t *s1_out)
{
    FHE_CHECK_ARG(ctx && path, "fhe_fastr_load: null argument");
    uint64_t bk_off, ksk_off, ksk_bytes;
    const uint64_t total = fastr_bytes(ctx->p.n0, bk_off, ksk_off, ksk_bytes);
    File f;
    KIO_TRY(f.open(path, false));
    if (f.size() != total) {
        set_error("fastR file '%s': %llu bytes, expected %llu at n0 = %u", path,
                  (unsigned long long)f.size(), (unsigned long long)total, ctx->p.n0);
        return FHE_ERR_ARG;
    }
    uint32_t s1[Lvl1::N];
    KIO_TRY(f.read(s1, bk_off, 1));
    if (s1_out) std::memcpy(s1_out, s1, sizeof(s1));
    KIO_TRY(load_bk_from_file(ctx, f, 1, FHE_BK_FFT, ksk_off - bk_off, 8));
    return load_ksk_stream(ctx, f, 10, ksk_bytes);
}

}  // extern "C"
