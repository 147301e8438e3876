// Exercises the C++ drop-in (include/cufhedb_amd/cuhedb.hpp) the way a
// cuFHEDB caller would: CuFFT_Torus<P>::LoadBK / *_st calls and the cuHEDB::
// operators.  Exit status 0 = every check passed.  Needs a GPU.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cufhedb_amd/cuhedb.hpp"

using cufhedb_amd::lvl1param;

static uint32_t phase(const uint32_t *c, const uint32_t *key, uint32_t n)
{
    uint32_t p = c[n];
    for (uint32_t i = 0; i < n; i++) p -= c[i] * key[i];
    return p;
}

int main()
{
    fhe_params p;
    cufhedb_amd::check(fhe_params_preset(FHE_PRESET_CFG636, &p), "preset");
    const uint32_t n0 = p.n0;
    std::vector<uint32_t> s0(n0), s1(1024);
    std::vector<uint64_t> s2(2048);
    cufhedb_amd::check(fhe_keygen_secret(&p, 7, s0.data(), s1.data(), s2.data()), "keygen");
    std::vector<uint32_t> bk(size_t(n0) * 6 * 2 * 1024);
    cufhedb_amd::check(fhe_keygen_bk_lvl01(&p, 8, s0.data(), s1.data(), bk.data()), "bk");
    std::vector<uint32_t> ksk(fhe_ksk_words(&p, 10));
    cufhedb_amd::check(fhe_keygen_ksk(&p, 10, 9, s1.data(), s0.data(), ksk.data()), "ksk");

    // the key-switching key travels through the reference's key-file format:
    // written as HEDB::generate_iksklvl10_key would, loaded by
    // HEDB::load_iksklvl10_key; the secret key likewise
    const std::string dir = "/tmp/dropin_evalkey_" + std::to_string(getpid());
    mkdir(dir.c_str(), 0700);
    cufhedb_amd::check(fhe_keyfile_write((dir + "/iksklvl10.key").c_str(), FHE_KEYFILE_IKSKLVL10,
                                         &p, ksk.data()), "write iksk");
    cufhedb_amd::check(fhe_keyfile_write_secret((dir + "/secret.key").c_str(), &p, s0.data(),
                                                s1.data(), s2.data()), "write secret");
    HEDB::SecretKey sk;
    HEDB::load_secret_key(sk, dir);
    int bad = sk.lvl0 != s0 || !std::equal(s1.begin(), s1.end(), sk.lvl1.begin());

    cufhedb_amd::Runtime rt(p, 0);
    HEDB::load_iksklvl10_key(rt, dir);
    std::remove((dir + "/iksklvl10.key").c_str());
    std::remove((dir + "/secret.key").c_str());
    rmdir(dir.c_str());
    cufft::CuFFT_Torus<lvl1param> cufftlvl1(rt);
    cufftlvl1.LoadBKInt(bk.data());

    std::mt19937 gen(3);
    // GateBootstrappingTLWE2TLWEFFT_st with constant test vector u: output
    // +u for phase in [0, 1/2), -u otherwise
    for (int t = 0; t < 8; t++) {
        const int bit = gen() & 1;
        std::vector<uint32_t> c0(n0 + 1);
        const uint32_t m = bit ? (1u << 30) : (3u << 30);
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(11, t, s0.data(), n0, &m, p.alpha0, c0.data(), 1), "enc");
        std::array<uint32_t, 1025> out;
        cufftlvl1.GateBootstrappingTLWE2TLWEFFT_st(out, c0, lvl1param::mu, 0);
        const bool pos = int32_t(phase(out.data(), s1.data(), 1024)) > 0;
        bad += pos != bool(bit);
    }
    // three requests of different sizes and test vectors through the
    // coalescing scheduler (one merged launch); outputs checked per request
    {
        cufhedb_amd::Scheduler sched(rt, 64);
        const int sizes[3] = {5, 17, 2};
        const uint32_t us[3] = {lvl1param::mu, 1u << 28, 3u << 27};
        cufhedb_amd::DeviceBuffer din[3], dtv[3], dout[3];
        std::vector<std::vector<int>> bits(3);
        uint64_t tk[3];
        for (int q = 0; q < 3; q++) {
            const int B = sizes[q];
            std::vector<uint32_t> c(size_t(B) * (n0 + 1)), m(B), tvv(1024, us[q]);
            for (int i = 0; i < B; i++) {
                bits[q].push_back(int(gen() & 1));
                m[i] = bits[q][i] ? (1u << 30) : (3u << 30);
            }
            cufhedb_amd::check(fhe_tlwe_encrypt_u32(40 + q, 0, s0.data(), n0, m.data(), p.alpha0,
                                                    c.data(), B), "enc");
            din[q].reset(c.size() * 4);
            din[q].upload(c.data(), c.size() * 4);
            dtv[q].reset(4096);
            dtv[q].upload(tvv.data(), 4096);
            dout[q].reset(size_t(B) * 1025 * 4);
            tk[q] = sched.submit_pbs1(din[q].as<uint32_t>(), dtv[q].as<uint32_t>(),
                                      dout[q].as<uint32_t>(), B);
        }
        sched.flush();
        for (int q = 0; q < 3; q++) {
            sched.wait(tk[q]);
            std::vector<uint32_t> o(size_t(sizes[q]) * 1025);
            dout[q].download(o.data(), o.size() * 4);
            for (int i = 0; i < sizes[q]; i++) {
                const int32_t ph = int32_t(phase(o.data() + size_t(i) * 1025, s1.data(), 1024));
                const int32_t want = bits[q][i] ? int32_t(us[q]) : -int32_t(us[q]);
                bad += std::abs(int64_t(ph) - want) > (1 << 25);
            }
        }
    }
    // ifft_st / fft_st round trip within one unit (TFHEpp/test/polymul.cpp:19-28)
    std::array<uint32_t, 1024> a, back;
    std::array<double, 1024> fd;
    for (auto &x : a) x = gen();
    cufftlvl1.ifft_st(fd, a, 0);
    cufftlvl1.fft_st(back, fd, 0);
    for (int i = 0; i < 1024; i++) bad += (int32_t(back[i] - a[i]) > 1 || int32_t(a[i] - back[i]) > 1);
    // cuHEDB::HomAND / HomOR (LOGIC) truth tables
    for (int t = 0; t < 8; t++) {
        const int x = t & 1, y = (t >> 1) & 1;
        cuHEDB::TLWELvl1 ca, cb, r;
        const uint32_t mx = x ? lvl1param::mu : 0u - lvl1param::mu;
        const uint32_t my = y ? lvl1param::mu : 0u - lvl1param::mu;
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(12, t, s1.data(), 1024, &mx, p.alpha1, ca.data(), 1), "enc");
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(13, t, s1.data(), 1024, &my, p.alpha1, cb.data(), 1), "enc");
        cuHEDB::HomAND(rt, r, ca, cb, cuHEDB::LOGIC, 0);
        bad += (int32_t(phase(r.data(), s1.data(), 1024)) > 0) != bool(x & y);
        cuHEDB::HomOR(rt, r, ca, cb, cuHEDB::LOGIC, 0);
        bad += (int32_t(phase(r.data(), s1.data(), 1024)) > 0) != bool(x | y);
        cuHEDB::HomNOT<lvl1param>(r, ca);
        bad += (int32_t(phase(r.data(), s1.data(), 1024)) > 0) == bool(x);
    }
    // cuHEDB::greater_than / less_than_equal / equal on 6-bit lvl1 values
    // (HomCompare_gpu.cuh:14-76; values at scale 2^(32 - 6 - 1), as in
    // test/HE3_Query/HomComp.cu:75-80) and LOG_to_ARI
    for (int t = 0; t < 6; t++) {
        const uint32_t bits = 6;
        const uint32_t x = gen() % 64, y = t == 0 ? x : gen() % 64;
        const uint32_t mx = x << (31 - bits), my = y << (31 - bits);
        cuHEDB::TLWELvl1 cx, cy, r;
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(14, t, s1.data(), 1024, &mx, p.alpha1, cx.data(), 1), "enc");
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(15, t, s1.data(), 1024, &my, p.alpha1, cy.data(), 1), "enc");
        cuHEDB::greater_than(rt, cx, cy, r, bits, cuHEDB::LOGIC, 0);
        bad += (int32_t(phase(r.data(), s1.data(), 1024)) > 0) != (x > y);
        cuHEDB::less_than_equal(rt, cx, cy, r, bits, cuHEDB::ARITHMETIC, 0);
        bad += ((phase(r.data(), s1.data(), 1024) + (1u << 30)) >> 31) != uint32_t(x <= y);
        cuHEDB::equal(rt, cx, cy, r, bits, cuHEDB::LOGIC, 0);
        bad += (int32_t(phase(r.data(), s1.data(), 1024)) > 0) != (x == y);
    }
    // CuFFT_Torus::pbs_many_lut_batch (GateBootstrappingManyLUT, 2 tables
    // over Z_8 with a padding bit): table k of ciphertext i gives luts[k][m_i]
    {
        const int B = 16, K = 2;
        uint32_t luts[K][8];
        for (auto &t : luts)
            for (auto &v : t) v = uint32_t(gen() % 16) << 28;
        std::vector<uint32_t> tv(1024);
        for (int j = 0; j < 1024; j++) {
            const int k = j % K, q = (j - k + 64) / 128;
            tv[j] = q == 8 ? 0u - luts[k][0] : luts[k][q];
        }
        std::vector<uint32_t> mm(B), c0(size_t(B) * (n0 + 1)), r(size_t(B) * K * 1025);
        for (auto &v : mm) v = uint32_t(gen() % 8) << 28;
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(20, 0, s0.data(), n0, mm.data(), p.alpha0,
                                                c0.data(), B), "enc0");
        cufftlvl1.pbs_many_lut_batch(r.data(), c0.data(), tv.data(), 0, K, B);
        for (int i = 0; i < B; i++)
            for (int k = 0; k < K; k++) {
                const int32_t d = int32_t(phase(r.data() + (size_t(i) * K + k) * 1025, s1.data(),
                                                1024) - luts[k][mm[i] >> 28]);
                bad += d > (1 << 26) || d < -(1 << 26);
            }
    }
    // cuHEDB::HomAND_rescale_batch: AND of LOGIC bits, decrypting to 0 / 2^29
    {
        const int B = 8;
        std::vector<uint32_t> ma(B), mb(B), ca(size_t(B) * 1025), cb(size_t(B) * 1025), r(ca.size());
        const uint32_t mu = 1u << 29;
        for (int i = 0; i < B; i++) {
            ma[i] = (i & 1) ? mu : 0u - mu;
            mb[i] = (i & 2) ? mu : 0u - mu;
        }
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(21, 0, s1.data(), 1024, ma.data(), p.alpha1,
                                                ca.data(), B), "enc a");
        cufhedb_amd::check(fhe_tlwe_encrypt_u32(22, 0, s1.data(), 1024, mb.data(), p.alpha1,
                                                cb.data(), B), "enc b");
        cuHEDB::HomAND_rescale_batch(rt, r.data(), ca.data(), cb.data(), 29, B);
        for (int i = 0; i < B; i++) {
            const uint32_t want = (i & 1) && (i & 2) ? 1u << 29 : 0u;
            const int32_t d = int32_t(phase(r.data() + size_t(i) * 1025, s1.data(), 1024) - want);
            bad += d > (1 << 26) || d < -(1 << 26);
        }
    }
    // cuHEDB::compare_digits_batch: 12-bit values as three 4-bit lvl2
    // digits against a public threshold (lvl02 key and KSK 20 / 21 made here)
    {
        std::vector<uint64_t> bk2(size_t(n0) * 8 * 2 * 2048);
        cufhedb_amd::check(fhe_keygen_bk_lvl02(&p, 16, s0.data(), s2.data(), bk2.data()), "bk2");
        rt.load_bk(2, bk2.data(), FHE_BK_INT);
        cufhedb_amd::check(fhe_keygen_ksk_dev(rt.ctx(), 20, 17, s2.data(), s0.data(), nullptr),
                           "ksk20");
        cufhedb_amd::check(fhe_keygen_ksk_dev(rt.ctx(), 21, 18, s2.data(), s1.data(), nullptr),
                           "ksk21");
        const int B = 24, D = 3;
        const uint64_t t = 0x5A7;
        std::vector<uint64_t> vals(B), m(size_t(B) * D), x(size_t(B) * D * 2049);
        for (int i = 0; i < B; i++) {
            vals[i] = i < 4 ? t + i - 2 : gen() % 4096;
            for (int d = 0; d < D; d++) m[size_t(i) * D + d] = ((vals[i] >> (4 * d)) & 15) << 59;
        }
        cufhedb_amd::check(fhe_tlwe_encrypt_u64(19, 0, s2.data(), 2048, m.data(), p.alpha2,
                                                x.data(), B * D), "enc2");
        std::vector<uint32_t> r(size_t(B) * 1025);
        cuHEDB::compare_digits_batch(rt, FHE_CMP_LE, x.data(), D, 4, t, r.data(), cuHEDB::LOGIC, B);
        for (int i = 0; i < B; i++)
            bad += (int32_t(phase(r.data() + size_t(i) * 1025, s1.data(), 1024)) > 0) != (vals[i] <= t);
    }
    std::printf("dropin_demo: %s (%d mismatches)\n", bad ? "FAILED" : "ok", bad);
    return bad ? 1 : 0;
}
