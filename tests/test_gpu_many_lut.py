"""Many-table programmable bootstrap (fhe_pbs_many_lut_lvl01/02_dev)
against the oracle's GateBootstrappingManyLUT
(thirdparty/TFHEpp/include/gatebootstrapping.hpp:119-130).

Bars: K = 1 is the single-table PBS bit for bit (same kernels, same
modulus switch); for K > 1 every output is SampleExtractIndex(k) of the
device blind rotation of the rounded lvl0 batch, bit for bit, decrypts to
its table entry, and its phase is within 2^-6 of the torus of the oracle's
(FP64 against FP64 with a different operation order, as for every PBS).
"""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

pytestmark = pytest.mark.gpu


def _round(c0, nbit, K):
    sh = 32 - 1 - nbit + (K - 1).bit_length()
    r = c0.astype(np.uint64)
    a = (((r[:, :-1] + (1 << (sh - 1))) % (1 << 32)) >> sh) << sh
    b = (r[:, -1:] >> sh) << sh
    return np.concatenate([a, b], axis=1).astype(np.uint32)


def test_one_table_is_the_plain_pbs(keys636, ctx636, rng):
    k = keys636
    tv = rng.integers(0, 1 << 32, F.LVL1_N, dtype=np.uint64).astype(np.uint32)
    c0 = rng.integers(0, 1 << 32, (37, k.params.n0 + 1), dtype=np.uint64).astype(np.uint32)
    many = ctx636.pbs_many_lut(1, c0, tv, 1)
    assert np.array_equal(many[:, 0], ctx636.pbs_lvl01(c0, tv))


@pytest.mark.parametrize("K,bits", [(2, 3), (3, 2), (4, 2), (8, 1)])
def test_lvl1_tables_vs_oracle(keys636, ctx636, K, bits):
    k = keys636
    rng = np.random.default_rng(100 * K + bits)
    B = 300
    m = rng.integers(0, 1 << bits, B)
    luts = rng.integers(0, 16, (K, 1 << bits))
    tv = F.many_lut_testvector(luts << 28, bits)
    c0 = k.enc_lvl0((m << (31 - bits)).astype(np.uint32), index0=72000 + 10 * K + bits)
    got = ctx636.pbs_many_lut(1, c0, tv, K)
    assert got.shape == (B, K, F.LVL1_N + 1)
    acc = ctx636.blind_rotate_lvl01(_round(c0, 10, K), tv)
    for kk in range(K):
        assert np.array_equal(got[:, kk], ctx636.sample_extract_lvl1(acc, kk)), kk
    ph = F.tlwe_phase_u32(got.reshape(-1, F.LVL1_N + 1), k.sk.s1).reshape(B, K)
    want = (luts.T[m] << 28).astype(np.uint32)
    assert np.max(np.abs(O.tdiff32(ph, want))) < 1 << 27
    for b in (0, B - 1):
        ref = O.pbs_many32(c0[b], k.bkfft, tv, K)
        rph = F.tlwe_phase_u32(ref, k.sk.s1)
        assert np.max(np.abs(O.tdiff32(ph[b], rph))) <= 1 << 26, b


def test_lvl1_per_ciphertext_tables(keys636, ctx636):
    k = keys636
    rng = np.random.default_rng(9)
    B, K, bits = 20, 2, 3
    m = rng.integers(0, 8, B)
    luts = rng.integers(0, 16, (B, K, 8))
    tvs = np.stack([F.many_lut_testvector(luts[b] << 28, bits) for b in range(B)])
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=73000)
    got = ctx636.pbs_many_lut(1, c0, tvs, K)
    ph = F.tlwe_phase_u32(got.reshape(-1, F.LVL1_N + 1), k.sk.s1).reshape(B, K)
    want = (luts[np.arange(B), :, m] << 28).astype(np.uint32)
    assert np.max(np.abs(O.tdiff32(ph, want))) < 1 << 27


def test_lvl2_tables_vs_oracle(keys636):
    k = keys636
    bk2 = F.keygen_bk_lvl02(k.sk, 636501)
    ctx = F.Context(k.params)
    try:
        ctx.load_bk(2, bk2, F.BK_INT)
        bkfft2 = O.bk_to_fft64(bk2)
        rng = np.random.default_rng(11)
        B, K, bits = 64, 2, 3
        m = rng.integers(0, 8, B)
        luts = rng.integers(0, 16, (K, 8))
        tv = F.many_lut_testvector([[(int(v) << 60) for v in t] for t in luts], bits, F.LVL2_N)
        c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=74000)
        got = ctx.pbs_many_lut(2, c0, tv, K)
        acc = ctx.blind_rotate_lvl02(_round(c0, 11, K), tv)
        for kk in range(K):
            assert np.array_equal(got[:, kk], ctx.sample_extract_lvl2(acc, kk)), kk
        ph = O.tlwe_phase64(got.reshape(-1, F.LVL2_N + 1), k.sk.s2).reshape(B, K)
        want = (luts.T[m].astype(np.uint64) << np.uint64(60))
        assert np.max(np.abs(O.tdiff64(ph, want))) < 1 << 59
        for b in (0, B - 1):
            rph = O.tlwe_phase64(O.pbs_many64(c0[b], bkfft2, tv, K), k.sk.s2)
            assert np.max(np.abs(O.tdiff64(ph[b], rph))) <= 1 << 58, b
    finally:
        ctx.close()


def test_bad_arguments(keys636, ctx636):
    c0 = np.zeros((2, keys636.params.n0 + 1), np.uint32)
    tv = np.zeros(F.LVL1_N, np.uint32)
    for K in (0, 65):
        with pytest.raises(F.FheError):
            ctx636.pbs_many_lut(1, c0, tv, K)


def test_full_batch_8192(keys636, ctx636):
    """At the benchmark's batch (8192 ciphertexts, n0 = 636): two tables of
    Z_8 messages per blind rotation, every output decrypting to its table
    entry, and the one-table form equal to the plain PBS word for word."""
    k = keys636
    rng = np.random.default_rng(8192)
    B = 8192
    m = rng.integers(0, 8, B)
    luts = rng.integers(0, 16, (2, 8))
    tv = F.many_lut_testvector(luts << 28, 3)
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=75000)
    got = ctx636.pbs_many_lut(1, c0, tv, 2)
    ph = F.tlwe_phase_u32(got.reshape(-1, F.LVL1_N + 1), k.sk.s1).reshape(B, 2)
    dec = np.round(O.tdiff32(ph, 0) / (1 << 28)).astype(np.int64) % 16
    assert np.array_equal(dec, luts.T[m])
    one = ctx636.pbs_many_lut(1, c0, tv, 1)
    assert np.array_equal(one[:, 0], ctx636.pbs_lvl01(c0, tv))
