"""The oracle against the golden vectors written by the compiled reference
(tests/golden/make_golden.py).  CPU only.

Bars: integer pieces (decomposition, sample extract, IKS) bit-exact; FD
values within FP64 rounding (relative 1e-13); torus outputs of a single
FFT-bearing step within 2 units; PBS / gates decrypt-exact with the phase
within 2^-6 of the reference's.  Two FP64 bootstrapping implementations with
different rounding diverge in their low gadget digits within the first CMUX
steps and then accumulate independent noise: measured over 128 PBS on the
reference parameter set, both have output phase-noise std 2^22.3 and their
per-sample phase gap has std 2^22.8 (= sqrt(2) x noise), so 2^26 is a ~9 sigma
bound that still leaves half of the Z_8 decoding margin (2^27).
"""

import os

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O
from golden.make_golden import sparse_ksk

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N1, N2 = 1024, 2048
PHASE_TOL = 1 << 26  # per-sample phase gap vs the reference (see module docstring)


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_golden_fft_lvl1():
    g = load("fft_lvl1.npz")
    for b in range(g["a"].shape[0]):
        fd = O.ifft32(g["a"][b])
        assert np.max(np.abs(fd - g["fd"][b])) <= 1e-13 * np.max(np.abs(g["fd"][b]))
        assert np.max(np.abs(O.tdiff32(O.fft32(g["fd"][b]), g["back"][b]))) <= 1
        # PolyMul: TwistIFFT x2, MulInFD, TwistFFT (mulfft.hpp:124-140)
        fa, fb = O.ifft32(g["pm_a"][b]), O.ifft32(g["pm_b"][b])
        M = N1 // 2
        prod = np.empty(N1)
        prod[:M] = fa[:M] * fb[:M] - fa[M:] * fb[M:]
        prod[M:] = fa[:M] * fb[M:] + fa[M:] * fb[:M]
        assert np.max(np.abs(O.tdiff32(O.fft32(prod), g["pm"][b]))) <= 1


def test_golden_fft_lvl2():
    g = load("fft_lvl2.npz")
    for b in range(g["a"].shape[0]):
        fd = O.ifft64(g["a"][b])
        assert np.max(np.abs(fd - g["fd"][b])) <= 1e-13 * np.max(np.abs(g["fd"][b]))
        # 64-bit torus: FP64 keeps ~53 bits, so compare at 2^16 granularity
        assert np.max(np.abs(O.tdiff64(O.fft64(g["fd"][b]), g["back"][b]))) <= 1 << 16


def test_golden_decomposition_bit_exact():
    g = load("decomp.npz")
    for b in range(2):
        for d in range(3):
            assert np.array_equal(O.decompose32(g["p1"][b], d), g["d1"][b, d])
        for d in range(4):
            assert np.array_equal(O.decompose64(g["p2"][b], d), g["d2"][b, d])


def _sk672():
    return F.SecretKeys(F.preset(F.PRESET_REF672), 672001)


def test_golden_external_product_lvl1():
    g = load("extprod_lvl1.npz")
    sk = _sk672()
    for mi, m in enumerate(g["ms"]):
        trgsw = F.trgsw_encrypt_u32(int(g["trgsw_seed"]), int(m) + 5, sk.s1, int(m), 2.0 ** -25)
        fd = O.trgsw_to_fft32(trgsw)
        for b in range(g["trlwe"].shape[0]):
            got = O.external_product32(g["trlwe"][b], fd)
            assert np.max(np.abs(O.tdiff32(got, g["out"][mi, b]))) <= 2


def test_golden_external_product_lvl2():
    g = load("extprod_lvl2.npz")
    sk = _sk672()
    for mi, m in enumerate(g["ms"]):
        trgsw = np.zeros((8, 2, N2), np.uint64)
        F.check(F.lib().fhe_trgsw_encrypt_u64(int(g["trgsw_seed"]), int(m) + 5, F._ptr(sk.s2),
                                              int(m), 2.0 ** -52, F._ptr(trgsw)))
        fd = np.zeros((8, 2, N2))
        O.orc().orc_trgsw_to_fft_u64(O.P(fd), O.P(trgsw), N2, 4)
        for b in range(g["trlwe"].shape[0]):
            got = np.zeros((2, N2), np.uint64)
            O.orc().orc_external_product_u64(O.P(got), O.P(np.ascontiguousarray(g["trlwe"][b])),
                                             O.P(fd), N2, 4, 9)
            # uniform-word inputs: products reach 2^80, FP64 resolves ~2^28
            assert np.max(np.abs(O.tdiff64(got, g["out"][mi, b]))) <= 1 << 32


def test_golden_pbs_lvl01_ref672():
    g = load("pbs_lvl01_ref672.npz")
    sk = F.SecretKeys(F.preset(F.PRESET_REF672), int(g["sk_seed"]))
    bkfft = O.bk_to_fft32(F.keygen_bk_lvl01(sk, int(g["bk_seed"])))
    for b in range(4):
        out = O.pbs32(g["c0"][b], bkfft, g["tv"])
        ph = F.tlwe_phase_u32(out, sk.s1)[0]
        assert abs(int(O.tdiff32(ph, g["out_phase"][b]))) <= PHASE_TOL
        assert int(np.round(O.tdiff32(ph, 0) / (1 << 28))) % 16 == g["lut"][g["m"][b]]
    for b in range(2):
        acc = O.blind_rotate32(g["c0"][b], bkfft, g["tv"])
        d = O.tdiff32(O.trlwe_phase32(acc, sk.s1), O.trlwe_phase32(g["acc"][b], sk.s1))
        assert np.max(np.abs(d)) <= PHASE_TOL
        # SampleExtractIndex of the reference accumulator is its PBS output
        assert np.array_equal(O.sample_extract32(g["acc"][b]), g["out"][b])


def test_golden_pbs_many_lut_lvl01_ref672():
    """GateBootstrappingManyLUT<lvl01param, K> of the compiled reference for
    K = 2 and 4 against the oracle: every output within PHASE_TOL of the
    reference's phase and decrypting to its table entry."""
    g = load("pbs_many_lvl01_ref672.npz")
    sk = F.SecretKeys(F.preset(F.PRESET_REF672), int(g["sk_seed"]))
    bkfft = O.bk_to_fft32(F.keygen_bk_lvl01(sk, int(g["bk_seed"])))
    for K in (2, 4):
        m, luts = g[f"m{K}"], g[f"luts{K}"]
        for b in range(3):
            out = O.pbs_many32(g[f"c0_{K}"][b], bkfft, g[f"tv{K}"], K)
            ph = F.tlwe_phase_u32(out, sk.s1)
            for k in range(K):
                assert abs(int(O.tdiff32(ph[k], g[f"phase{K}"][b, k]))) <= PHASE_TOL, (K, b, k)
                assert int(np.round(O.tdiff32(ph[k], 0) / (1 << 28))) % 16 == luts[k][m[b]]


# lvl2 (64-bit torus): output phase-noise std 2^35.65 (2^-64 units) for the
# reference's operation order, which the oracle follows; two FP64
# implementations then differ per sample by ~sqrt(2) x that, 2^36.15, so 2^40
# is a ~14 sigma bound, 2^20 below the Z_8 decision margin (2^59).
PHASE_TOL2 = 1 << 40


def test_golden_pbs_lvl02_ref672():
    """GateBootstrappingTLWE2TLWEFFT<lvl02param> and GateBootstrappingManyLUT
    <lvl02param, 2> of the compiled reference against the oracle on the same
    seeded key and inputs."""
    g = load("pbs_lvl02_ref672.npz")
    sk = F.SecretKeys(F.preset(F.PRESET_REF672), int(g["sk_seed"]))
    bkfft = O.bk_to_fft64(F.keygen_bk_lvl02(sk, int(g["bk_seed"])))
    gaps = []
    for b in range(len(g["m"])):
        ph = O.tlwe_phase64(O.pbs64(g["c0"][b], bkfft, g["tv"]), sk.s2)[0]
        gap = int(O.tdiff64(ph, g["out_phase"][b]))
        gaps.append(gap)
        assert abs(gap) <= PHASE_TOL2, (b, gap)
        assert (int(ph) + (1 << 59)) % (1 << 64) >> 60 == g["lut"][g["m"][b]]
    for b in range(len(g["mm"])):
        out = O.pbs_many64(g["cm"][b], bkfft, g["tvm"], 2)
        ph = O.tlwe_phase64(out, sk.s2)
        for k in range(2):
            gap = int(O.tdiff64(ph[k], g["outm_phase"][b, k]))
            gaps.append(gap)
            assert abs(gap) <= PHASE_TOL2, (b, k, gap)
            assert (int(ph[k]) + (1 << 59)) % (1 << 64) >> 60 == g["luts"][k][g["mm"][b]]
    print(f"lvl02 oracle - reference phase gaps: max 2^{np.log2(max(map(abs, gaps)) + 1):.1f}")


def test_many_lut_is_rounded_single_lut():
    """The identity the device path is built on: the K-table bootstrap's
    accumulator equals the single-table blind rotation of the lvl0 words
    rounded to multiples of 2^sh (sh = 32 - 1 - nbit + bw), and K = 1 is the
    plain PBS.  Checked on the oracle, bit for bit."""
    p = F.preset(F.PRESET_CFG636)
    sk = F.SecretKeys(p, 61)
    bkfft = O.bk_to_fft32(F.keygen_bk_lvl01(sk, 62))
    rng = np.random.default_rng(63)
    c0 = rng.integers(0, 1 << 32, (2, p.n0 + 1), dtype=np.uint64).astype(np.uint32)
    tv = rng.integers(0, 1 << 32, 1024, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(O.pbs_many32(c0[0], bkfft, tv, 1)[0], O.pbs32(c0[0], bkfft, tv))
    for K, bw in ((2, 1), (3, 2), (8, 3)):
        sh = 21 + bw
        r = c0[1].astype(np.uint64)
        a = (((r[:-1] + (1 << (sh - 1))) % (1 << 32)) >> sh) << sh
        b = (r[-1:] >> sh) << sh
        rounded = np.concatenate([a, b]).astype(np.uint32)
        acc = O.blind_rotate32(rounded, bkfft, tv)
        many = O.pbs_many32(c0[1], bkfft, tv, K)
        for k in range(K):
            assert np.array_equal(many[k], O.sample_extract32(acc, k)), (K, k)


def test_golden_iks_bit_exact():
    g = load("iks.npz")
    for b in range(2):
        ksk, mm = sparse_ksk((N1, 2, 1023, 673), int(g["seeds10"][b]), g["in10"][b], 32, 2, 10)
        assert np.array_equal(O.iks32(g["in10"][b], ksk, 672), g["out10"][b])
        del ksk
        mm.close()
        ksk, mm = sparse_ksk((N2, 2, 1023, 673), int(g["seeds20"][b]), g["in2"][b], 64, 2, 10)
        assert np.array_equal(O.iks64(g["in2"][b], ksk, 672, 2, 10), g["out20"][b])
        del ksk
        mm.close()
        ksk, mm = sparse_ksk((N2, 10, 7, N1 + 1), int(g["seeds21"][b]), g["in2"][b], 64, 10, 3)
        assert np.array_equal(O.iks64(g["in2"][b], ksk, N1, 10, 3), g["out21"][b])
        del ksk
        mm.close()


def test_golden_hom_nand_ref672():
    g = load("gates_ref672.npz")
    sk = F.SecretKeys(F.preset(F.PRESET_REF672), int(g["sk_seed"]))
    bkfft = O.bk_to_fft32(F.keygen_bk_lvl01(sk, int(g["bk_seed"])))
    ksk10 = F.keygen_ksk(sk, 10, int(g["ksk_seed"]))
    bits = g["bits"]
    for i in range(6):
        out = O.hom_nand(g["ca"][i], g["cb"][i], ksk10, 672, bkfft)
        ph = F.tlwe_phase_u32(out, sk.s1)[0]
        assert abs(int(O.tdiff32(ph, g["nand_phase"][i]))) <= PHASE_TOL
        assert (int(ph) < (1 << 31) and int(ph) != 0) == (not (bits[i, 0] and bits[i, 1]))
    # reference outputs decrypt to the truth table
    ref_bits = (g["nand_phase"].view(np.int32) > 0)
    assert np.array_equal(ref_bits, ~(bits[:, 0].astype(bool) & bits[:, 1].astype(bool)))
