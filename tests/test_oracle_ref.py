"""The oracle against the compiled reference (oracle/_ref) on fresh random
inputs.  Skipped where the reference build is absent (it is built only where
the reference sources were available).  CPU only."""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

pytestmark = pytest.mark.skipif(not O.ref_built(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    # loaded only when one of these tests runs, not at collection
    return O.ref()

N1, N2 = 1024, 2048


def test_fft_lvl1_vs_reference(ref, rng):
    for _ in range(8):
        a = rng.integers(0, 1 << 32, N1, dtype=np.uint64).astype(np.uint32)
        fr = np.zeros(N1)
        ref.ref_twist_ifft_lvl1(O.P(fr), O.P(a))
        assert np.max(np.abs(O.ifft32(a) - fr)) <= 1e-13 * np.max(np.abs(fr))
        br = np.zeros(N1, np.uint32)
        ref.ref_twist_fft_lvl1(O.P(br), O.P(fr))
        assert np.max(np.abs(O.tdiff32(O.fft32(fr), br))) <= 1


def test_fft_lvl2_vs_reference(ref, rng):
    for _ in range(4):
        a = rng.integers(0, 1 << 64, N2, dtype=np.uint64)
        fr = np.zeros(N2)
        ref.ref_twist_ifft_lvl2(O.P(fr), O.P(a))
        assert np.max(np.abs(O.ifft64(a) - fr)) <= 1e-13 * np.max(np.abs(fr))


def test_decomposition_vs_reference(ref, rng):
    x = rng.integers(0, 1 << 32, N1, dtype=np.uint64).astype(np.uint32)
    x2 = rng.integers(0, 1 << 64, N2, dtype=np.uint64)
    for d in range(3):
        r = np.zeros(N1, np.uint32)
        ref.ref_decompose_lvl1(O.P(r), O.P(x), d)
        assert np.array_equal(O.decompose32(x, d), r)
    for d in range(4):
        r = np.zeros(N2, np.uint64)
        ref.ref_decompose_lvl2(O.P(r), O.P(x2), d)
        assert np.array_equal(O.decompose64(x2, d), r)


def test_sample_extract_vs_reference(ref, rng):
    t = rng.integers(0, 1 << 32, (2, N1), dtype=np.uint64).astype(np.uint32)
    for idx in (0, 3, 1023):
        r = np.zeros(N1 + 1, np.uint32)
        ref.ref_sample_extract_lvl1(O.P(r), O.P(t), idx)
        assert np.array_equal(O.sample_extract32(t, idx), r)


def test_external_product_vs_reference(ref, keys672, rng):
    s1 = keys672.sk.s1
    trgsw = F.trgsw_encrypt_u32(91, 1, s1, 1, 2.0 ** -25)
    fr = np.zeros((6, 2, N1))
    ref.ref_trgsw_to_fft_lvl1(O.P(fr), O.P(trgsw))
    assert np.max(np.abs(O.trgsw_to_fft32(trgsw) - fr)) <= 1e-13 * np.max(np.abs(fr))
    for _ in range(3):
        c = rng.integers(0, 1 << 32, (2, N1), dtype=np.uint64).astype(np.uint32)
        r = np.zeros((2, N1), np.uint32)
        ref.ref_external_product_lvl1(O.P(r), O.P(c), O.P(fr))
        assert np.max(np.abs(O.tdiff32(O.external_product32(c, fr), r))) <= 2


def test_iks_lvl10_vs_reference_real_key(ref, keys672, rng):
    k = keys672
    c1 = k.enc_lvl1((rng.integers(0, 8, 3) << 28).astype(np.uint32), index0=77)
    for b in range(3):
        r = np.zeros(673, np.uint32)
        ref.ref_iks_lvl10(O.P(r), O.P(c1[b]), O.P(k.ksk10))
        assert np.array_equal(O.iks32(c1[b], k.ksk10, 672), r)


def test_many_lut_vs_reference(ref, keys672, rng):
    """GateBootstrappingManyLUT<lvl01param, K> for K = 3 and 8 (tables of a
    non-power-of-two count, and the widest stride the harness instantiates)
    against the oracle on fresh ciphertexts: phases within 2^-6 of the torus."""
    k = keys672
    for K, bits in ((3, 2), (8, 1)):
        luts = rng.integers(0, 16, (K, 1 << bits))
        tv = F.many_lut_testvector(luts << 28, bits)
        m = rng.integers(0, 1 << bits, 2)
        c0 = k.enc_lvl0((m << (31 - bits)).astype(np.uint32), index0=900 + K)
        for b in range(2):
            r = np.zeros((K, N1 + 1), np.uint32)
            assert ref.ref_pbs_many_lvl01(O.P(r), O.P(c0[b]), O.P(k.bkfft), O.P(tv), K) == 0
            got = O.pbs_many32(c0[b], k.bkfft, tv, K)
            d = O.tdiff32(F.tlwe_phase_u32(got, k.sk.s1), F.tlwe_phase_u32(r, k.sk.s1))
            assert np.max(np.abs(d)) <= 1 << 26, (K, b)
