// A cuFHEDB caller written against the reference's own signatures: the
// filter of the TPC-H Q1 query (the per-row steps of
// test/HE3_Query/TPCH_Q1.cu:248-283, written anew) over a small seeded
// table, through include/cufhedb_amd/cuhedb.hpp with no call-site change:
// cufftlvl1.LoadBK(*ek.bkfftlvl01), cuHEDB::less_than_equal<Lvl1>(.., ek,
// LOGIC, stream_id), cuHEDB::less_than<Lvl1>, cuHEDB::HomNOT<Lvl1>,
// cuHEDB::HomAND(.., ek, LOGIC, stream_id), cuHEDB::log_rescale(.., ek,
// stream_id) from OpenMP-style worker threads, each with its stream_id.
//
// The prelude below stands in for what the reference's headers give its
// callers (tfhe++.hpp's EvalKey with its member names, HEDB's type names);
// with TFHEpp on the include path TFHEpp::EvalKey is used unchanged.  Keys
// are generated through the C ABI (seeded), the key-switching key on the
// device and copied into the EvalKey, the BK transformed into TFHEpp's FD
// form.  Exit status 0 = every decrypted predicate equals the plaintext one.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "cufhedb_amd/cuhedb.hpp"

using namespace HEDB;

// ---- prelude: TFHEpp::EvalKey's members this path reads (cloudkey.hpp:93-105)
using BKFFT01 = std::array<std::array<std::array<std::array<double, 1024>, 2>, 6>, 672>;
using BKFFT02 = std::array<std::array<std::array<std::array<double, 2048>, 2>, 8>, 672>;
using KSK10 = std::array<std::array<std::array<TLWELvl0, 1023>, 2>, 1024>;
using KSK20 = std::array<std::array<std::array<TLWELvl0, 1023>, 2>, 2048>;
using KSK21 = std::array<std::array<std::array<TLWELvl1, 7>, 10>, 2048>;
struct EvalKey {
    std::unique_ptr<BKFFT01> bkfftlvl01;
    std::unique_ptr<BKFFT02> bkfftlvl02;
    std::unique_ptr<KSK10> iksklvl10;
    std::unique_ptr<KSK20> iksklvl20;
    std::unique_ptr<KSK21> iksklvl21;
};
using TFHEEvalKey = EvalKey;

const uint32_t bin_bits = 1;
const uint32_t date_bits = 9;

static TLWELvl1 encrypt_int(uint32_t v, uint32_t bits, const uint32_t *s1, uint64_t seed)
{
    // tlweSymInt32Encrypt (HEDB/comparison/tfhepp_utils.h:16-30): v * 2^(31 - bits)
    TLWELvl1 c;
    const uint32_t m = v << (31 - bits);
    cufhedb_amd::check(fhe_tlwe_encrypt_u32(seed, 0, s1, 1024, &m, std::pow(2.0, -25), c.data(), 1),
                       "encrypt");
    return c;
}

static uint32_t decrypt_int(const TLWELvl1 &c, double scale, const uint32_t *s1)
{
    // tlweSymInt32Decrypt (tfhepp_utils.h:32-44)
    uint32_t phase = c[1024];
    for (int i = 0; i < 1024; i++) phase -= c[i] * s1[i];
    const uint32_t modulus = uint32_t((1ull << 31) / scale) * 2;
    return uint32_t(std::llround(double(int32_t(phase)) / scale)) % modulus;
}

// ---- the filter, as the reference writes it (TPCH_Q1.cu:248-283)
static void filter(std::vector<std::vector<TLWELvl1>> &cres, std::vector<TLWELvl1> &l_returnflag,
                   std::vector<TLWELvl1> &l_linestatus, std::vector<TLWELvl1> &l_shipdate,
                   TLWELvl1 &date1, TLWELvl1 &cipher1, TFHEEvalKey &ek, int nthreads)
{
    const size_t lineitem_rows = l_returnflag.size();
    cres.assign(4, std::vector<TLWELvl1>(lineitem_rows));
    std::vector<TLWELvl1> pred_cres1(lineitem_rows), pred_cres2(lineitem_rows),
        pred_cres3(lineitem_rows), pred_cres4(lineitem_rows), pred_cres5(lineitem_rows);
    auto row = [&](size_t i, uint32_t stream_id) {
        cuHEDB::less_than_equal<Lvl1>(l_shipdate[i], date1, pred_cres1[i], date_bits, ek, LOGIC, stream_id);
        cuHEDB::less_than<Lvl1>(l_returnflag[i], cipher1, pred_cres2[i], bin_bits, ek, LOGIC, stream_id);
        cuHEDB::less_than<Lvl1>(l_linestatus[i], cipher1, pred_cres3[i], bin_bits, ek, LOGIC, stream_id);
        cuHEDB::HomNOT<Lvl1>(pred_cres4[i], pred_cres2[i]);
        cuHEDB::HomNOT<Lvl1>(pred_cres5[i], pred_cres3[i]);

        cuHEDB::HomAND(cres[0][i], pred_cres1[i], pred_cres2[i], ek, LOGIC, stream_id);
        cuHEDB::HomAND(cres[1][i], cres[0][i], pred_cres5[i], ek, LOGIC, stream_id);  // 01
        cuHEDB::HomAND(cres[0][i], cres[0][i], pred_cres3[i], ek, LOGIC, stream_id);  // 00

        cuHEDB::HomAND(cres[2][i], pred_cres1[i], pred_cres4[i], ek, LOGIC, stream_id);
        cuHEDB::HomAND(cres[3][i], cres[2][i], pred_cres5[i], ek, LOGIC, stream_id);  // 11
        cuHEDB::HomAND(cres[2][i], cres[2][i], pred_cres3[i], ek, LOGIC, stream_id);  // 10

        const uint32_t rlwe_scale_bits = 29;
        for (int g = 0; g < 4; g++) cuHEDB::log_rescale(cres[g][i], cres[g][i], rlwe_scale_bits, ek, stream_id);
    };
    // the reference's `#pragma omp parallel for` with stream_id = omp_get_thread_num()
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < lineitem_rows; i += nthreads) row(i, uint32_t(t));
        });
    for (auto &th : pool) th.join();
}

int main()
{
    fhe_params p;
    cufhedb_amd::check(fhe_params_preset(FHE_PRESET_REF672, &p), "preset");
    std::vector<uint32_t> s0(p.n0), s1(1024);
    std::vector<uint64_t> s2(2048);
    cufhedb_amd::check(fhe_keygen_secret(&p, 21, s0.data(), s1.data(), s2.data()), "secret");

    // the evaluation key, in TFHEpp's forms
    TFHEEvalKey ek;
    {
        std::vector<uint32_t> bk(size_t(p.n0) * 6 * 2 * 1024);
        cufhedb_amd::check(fhe_keygen_bk_lvl01(&p, 22, s0.data(), s1.data(), bk.data()), "bk");
        ek.bkfftlvl01 = std::make_unique<BKFFT01>();
        cufft::CuFFT_Torus<Lvl1> helper(cufhedb_amd::default_runtime());
        std::array<uint32_t, 1024> poly;
        std::array<double, 1024> fd;
        for (size_t q = 0; q < bk.size() / 1024; q++) {
            std::copy(bk.begin() + q * 1024, bk.begin() + (q + 1) * 1024, poly.begin());
            helper.ifft_st(fd, poly, 0);  // TwistIFFT (TFHEpp's FD order)
            (&(*ek.bkfftlvl01)[0][0][0])[q] = fd;
        }
        cufhedb_amd::DeviceBuffer dk(sizeof(KSK10));
        cufhedb_amd::check(fhe_keygen_ksk_dev(cufhedb_amd::default_runtime().ctx(), 10, 23,
                                              s1.data(), s0.data(), dk.as<uint32_t>()),
                           "ksk");
        ek.iksklvl10 = std::make_unique<KSK10>();
        dk.download(ek.iksklvl10.get(), sizeof(KSK10));
    }
    cufftlvl1.LoadBK(*ek.bkfftlvl01);

    // seeded table: returnflag, linestatus in {0, 1}, shipdate < 2^9
    const size_t rows = 12;
    std::mt19937 gen(24);
    std::vector<uint32_t> rf(rows), ls(rows), sd(rows);
    std::vector<TLWELvl1> l_returnflag, l_linestatus, l_shipdate;
    for (size_t i = 0; i < rows; i++) {
        rf[i] = gen() & 1;
        ls[i] = gen() & 1;
        sd[i] = gen() % (1u << date_bits);
        l_returnflag.push_back(encrypt_int(rf[i], bin_bits, s1.data(), 100 + i));
        l_linestatus.push_back(encrypt_int(ls[i], bin_bits, s1.data(), 200 + i));
        l_shipdate.push_back(encrypt_int(sd[i], date_bits, s1.data(), 300 + i));
    }
    const uint32_t plain_date = 300;
    TLWELvl1 date1 = encrypt_int(plain_date, date_bits, s1.data(), 400);
    TLWELvl1 cipher1 = encrypt_int(1, bin_bits, s1.data(), 401);

    std::vector<std::vector<TLWELvl1>> cres;
    filter(cres, l_returnflag, l_linestatus, l_shipdate, date1, cipher1, ek, 4);

    int bad = 0;
    for (size_t i = 0; i < rows; i++) {
        const uint32_t d = sd[i] <= plain_date;
        const uint32_t want[4] = {d & !rf[i] & !ls[i], d & !rf[i] & ls[i], d & rf[i] & !ls[i],
                                  d & rf[i] & ls[i]};
        for (int g = 0; g < 4; g++) {
            const uint32_t got = decrypt_int(cres[g][i], std::pow(2.0, 29), s1.data());
            if (got != want[g]) {
                std::printf("row %zu group %d: got %u want %u\n", i, g, got, want[g]);
                bad++;
            }
        }
    }
    // the per-ciphertext PolyMul_st / MoveBuf / *_test members exist and run
    {
        std::array<std::array<uint32_t, 1024>, 2> trlwe{}, res{};
        trlwe[1][3] = 5;
        cufftlvl1.PolyMul_st(res, trlwe, 0);  // times the all-ones polynomial
        // TwistFFT truncates toward zero: within one unit of the exact product
        for (int j = 0; j < 1024; j++) {
            const int32_t want = j >= 3 ? 5 : -5, got = int32_t(res[1][j]);
            bad += got - want > 1 || want - got > 1;
        }
        std::array<double, 1024> fd;
        double cost = 0;
        cufftlvl1.ifft_test(fd, trlwe[1], 3, cost);
        bad += cost <= 0;
        std::array<std::array<std::array<double, 1024>, 2>, 6> buf;
        cufftlvl1.MoveBuf(buf, 0);
    }
    std::printf("%s: TPC-H Q1 filter, %zu rows, reference signatures\n", bad ? "FAILED" : "ok", rows);
    return bad ? 1 : 0;
}
