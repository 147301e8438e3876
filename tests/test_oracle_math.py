"""Pins of the oracle by what the mathematics fixes (CPU only).

Each check is chosen so that a plausible restatement mistake fails it; the
`planted` tests plant such mistakes in copies of oracle outputs and assert
that the check rejects them.

  * TwistIFFT == evaluation of the polynomial at w^(4m+1) (w = exp(i pi/N)),
    stored in spqlios bit-reversed order as [re | im]   (brute force O(N^2))
  * TwistFFT(TwistIFFT(a)) == a within 1 (lvl1) / 2^14 (lvl2) and FFT
    polymul == exact negacyclic product within 1       (TFHEpp/test/polymul.cpp)
  * decomposition: digits in [-Bg/2, Bg/2), sum_d digit_d 2^(W-(d+1)Bgbit)
    within 2^(W - l Bgbit - 1) of x                       (trgsw.cpp:21-39)
  * X^a rotations: X^a X^b = X^(a+b), X^N = -1, (X^a - 1)p = X^a p - p,
    and X^a p equals the exact product with the monomial (utils.hpp:110-141)
  * TRGSW(+-1) (x) TRLWE(m) decrypts to +-m              (externalproduct.cpp)
  * CMUX(TRGSW(b), c1, c0) decrypts to c_b               (cmux.cpp)
  * sample extract phase == coefficient 0 of the TRLWE phase
  * IKS preserves the phase within key-switch noise     (iks.cpp)
  * blind rotation: acc phase == X^(-phi~) tv within noise
"""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

N = 1024


def dft_eval(a):
    """spqlios-ordered FD by direct evaluation at w^(4m+1)."""
    Nn = a.shape[0]
    M = Nn // 2
    x = a.astype(np.int64).astype(np.float64) if a.dtype != np.float64 else a
    k = np.arange(Nn)
    out = np.zeros(Nn)
    for idx in range(M):
        m = O.orc().orc_fd_rev(idx, M)
        z = np.exp(1j * np.pi * (4 * m + 1) * k / Nn)
        v = np.sum(x * z)
        out[idx] = v.real
        out[idx + M] = v.imag
    return out


def fft_matches_dft(fd, a):
    want = dft_eval(a)
    return np.max(np.abs(fd - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("Nn", [16, 64, 256])
def test_ifft_is_evaluation_at_odd_roots(Nn, rng):
    a = (rng.integers(-1000, 1000, Nn)).astype(np.int32).view(np.uint32)
    assert fft_matches_dft(O.ifft32(a, Nn), a.view(np.int32))


def test_ifft_full_size_spot_check(rng):
    a = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    fd = O.ifft32(a)
    x = a.view(np.int32).astype(np.float64)
    k = np.arange(N)
    for idx in (0, 1, 77, 255, 511):
        m = O.orc().orc_fd_rev(idx, N // 2)
        v = np.sum(x * np.exp(1j * np.pi * (4 * m + 1) * k / N))
        assert abs(fd[idx] - v.real) <= 1e-12 * abs(v) + 1e-3
        assert abs(fd[idx + N // 2] - v.imag) <= 1e-12 * abs(v) + 1e-3


def test_planted_fft_mistakes_are_caught(rng):
    Nn = 64
    a = rng.integers(-1000, 1000, Nn).astype(np.int32).view(np.uint32)
    fd = O.ifft32(a, Nn)
    M = Nn // 2
    conj = fd.copy()
    conj[M:] *= -1  # wrong twist / DFT sign
    natural = np.concatenate([fd[[O.orc().orc_fd_rev(i, M) for i in range(M)]],
                              fd[[M + O.orc().orc_fd_rev(i, M) for i in range(M)]]])
    swapped = np.concatenate([fd[M:], fd[:M]])  # re/im halves exchanged
    for bad in (conj, natural, swapped):
        assert not fft_matches_dft(bad, a.view(np.int32))


def test_fft_roundtrip_and_polymul_bounds(rng):
    for _ in range(20):
        a = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
        assert np.max(np.abs(O.tdiff32(O.fft32(O.ifft32(a)), a))) <= 1
        d = (rng.integers(0, 64, N) - 32).astype(np.int64).astype(np.uint32)
        fa, fb = O.ifft32(d), O.ifft32(a)
        M = N // 2
        prod = np.empty(N)
        prod[:M] = fa[:M] * fb[:M] - fa[M:] * fb[M:]
        prod[M:] = fa[:M] * fb[M:] + fa[M:] * fb[:M]
        assert np.max(np.abs(O.tdiff32(O.fft32(prod), O.naive32(d, a)))) <= 1
    for _ in range(5):
        a2 = rng.integers(0, 1 << 64, 2048, dtype=np.uint64)
        assert np.max(np.abs(O.tdiff64(O.fft64(O.ifft64(a2)), a2))) <= 1 << 14


def test_naive_product_small_cases():
    # (1 + X) * X^(N-1) = X^(N-1) + X^N = -1 + X^(N-1)  (mod X^N + 1)
    a = np.zeros(8, np.uint32)
    a[0] = a[1] = 1
    b = np.zeros(8, np.uint32)
    b[7] = 1
    got = O.naive32(a, b).view(np.int32)
    assert list(got) == [-1, 0, 0, 0, 0, 0, 0, 1]


def recombine(digits, W, l, Bgbit):
    tot = 0
    for d in range(l):
        tot += digits[d].astype(object) * (1 << (W - (d + 1) * Bgbit))
    return tot


@pytest.mark.parametrize("W,l,Bgbit", [(32, 3, 6), (64, 4, 9)])
def test_decomposition_closed_form(W, l, Bgbit, rng):
    Nn = 256
    if W == 32:
        x = rng.integers(0, 1 << 32, Nn, dtype=np.uint64).astype(np.uint32)
        digs = [O.decompose32(np.resize(x, Nn), d, l, Bgbit).view(np.int32) for d in range(l)]
    else:
        x = rng.integers(0, 1 << 64, Nn, dtype=np.uint64)
        digs = [O.decompose64(x, d, l, Bgbit).view(np.int64) for d in range(l)]
    half = 1 << (Bgbit - 1)
    for d in range(l):
        assert digs[d].min() >= -half and digs[d].max() < half
    rec = recombine(digs, W, l, Bgbit)
    mod = 1 << W
    for xi, ri in zip(x.tolist(), rec.tolist()):
        err = (int(xi) - int(ri)) % mod
        err = err - mod if err >= mod // 2 else err
        assert abs(err) <= 1 << (W - l * Bgbit - 1)
    # zero decomposes to all-zero digits (the a-bar == 0 skip is neutral)
    z = O.decompose32(np.zeros(8, np.uint32), 0) if W == 32 else O.decompose64(np.zeros(8, np.uint64), 0)
    assert not z.any()


def test_planted_decomposition_mistake_is_caught(rng):
    x = rng.integers(0, 1 << 32, 64, dtype=np.uint64).astype(np.uint32)
    digs = [O.decompose32(x, d).view(np.int32) for d in range(3)]
    digs[1] = -digs[1]  # sign error on one digit
    rec = recombine(digs, 32, 3, 6)
    errs = [((int(a) - int(b)) % (1 << 32)) for a, b in zip(x.tolist(), rec.tolist())]
    errs = [e - (1 << 32) if e >= 1 << 31 else e for e in errs]
    assert max(abs(e) for e in errs) > 1 << 13


def test_rotation_identities(rng):
    p = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    for a, b in [(0, 5), (3, 1021), (1024, 7), (1500, 600), (2047, 1)]:
        ab = O.mul_xai32(O.mul_xai32(p, a), b)
        assert np.array_equal(ab, O.mul_xai32(p, (a + b) % 2048))
        mono = np.zeros(N, np.uint32)
        if a < N:
            mono[a] = 1
        else:
            mono[a - N] = 0xFFFFFFFF  # X^a = -X^(a-N)
        assert np.array_equal(O.mul_xai32(p, a), O.naive32(mono, p))
        assert np.array_equal(O.mul_xai32(p, a, minus_one=True), O.mul_xai32(p, a) - p)
    assert np.array_equal(O.mul_xai32(p, N), (0 - p.astype(np.int64)).astype(np.uint32))


def test_external_product_decrypts(keys672, rng):
    s1 = keys672.sk.s1
    msgs = rng.integers(0, 8, (3, N)).astype(np.int64) << 28
    trlwe = F.trlwe_encrypt_u32(31, 0, s1, msgs.astype(np.uint32), 2.0 ** -25)
    for m in (1, -1, 0):
        fd = O.trgsw_to_fft32(F.trgsw_encrypt_u32(32, m + 9, s1, m, 2.0 ** -25))
        for b in range(3):
            ph = O.trlwe_phase32(O.external_product32(trlwe[b], fd), s1)
            dec = np.round(O.tdiff32(ph, 0) / (1 << 28)).astype(np.int64) % 16
            assert np.array_equal(dec, (m * msgs[b] >> 28) % 16)


def test_cmux_selects(keys672, rng):
    s1 = keys672.sk.s1
    m0 = rng.integers(0, 8, N).astype(np.int64) << 28
    m1 = rng.integers(0, 8, N).astype(np.int64) << 28
    c0 = F.trlwe_encrypt_u32(41, 0, s1, m0.astype(np.uint32), 2.0 ** -25)[0]
    c1 = F.trlwe_encrypt_u32(41, 1, s1, m1.astype(np.uint32), 2.0 ** -25)[0]
    for bit in (0, 1):
        fd = O.trgsw_to_fft32(F.trgsw_encrypt_u32(42, bit, s1, bit, 2.0 ** -25))
        # CMUXFFT (detwfa.cpp:5-14): c0 + TRGSW(b) (x) (c1 - c0)
        res = O.external_product32(c1 - c0, fd) + c0
        dec = np.round(O.tdiff32(O.trlwe_phase32(res, s1), 0) / (1 << 28)).astype(np.int64) % 16
        assert np.array_equal(dec, ((m1 if bit else m0) >> 28) % 16)


def test_sample_extract_phase(keys672, rng):
    s1 = keys672.sk.s1
    c = F.trlwe_encrypt_u32(51, 0, s1, rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32),
                            2.0 ** -25)[0]
    ph = O.trlwe_phase32(c, s1)
    for idx in (0, 1, 500, 1023):
        t = np.zeros(N + 1, np.uint32)
        O.orc().orc_sample_extract_u32(O.P(t), O.P(c), idx, N)
        assert F.tlwe_phase_u32(t, s1)[0] == ph[idx]


def test_iks_preserves_phase(keys636, rng):
    k = keys636
    m = (rng.integers(0, 8, 4) << 28).astype(np.uint32)
    c1 = k.enc_lvl1(m, index0=300)
    for b in range(4):
        c0 = O.iks32(c1[b], k.ksk10, k.n0)
        d = O.tdiff32(F.tlwe_phase_u32(c0, k.sk.s0)[0], m[b])
        # KS noise: sqrt(1024 * 2 * var(alpha0)) ~ 2^22.5, plus the 2^11
        # rounding per digit: far below the 2^27 Z_8 half-box
        assert abs(int(d)) < 1 << 25


def test_blind_rotate_phase_is_rotated_testvector(keys636, rng):
    from test_gpu_lvl1 import z8_testvector
    k = keys636
    tv = z8_testvector(rng.integers(0, 8, 8))
    m = int(rng.integers(0, 8))
    c0 = k.enc_lvl0(np.array([m << 28], np.uint32), index0=400)[0]
    acc = O.blind_rotate32(c0, k.bkfft, tv)
    ph = O.trlwe_phase32(acc, k.sk.s1)
    # phi~ from the exact phase of c0 (b-bar unrounded, a-bar rounded)
    n0 = k.n0
    bbar = 2048 - (int(c0[n0]) >> 21)
    phit = bbar
    for i in range(n0):
        abar = ((int(c0[i]) + (1 << 20)) % (1 << 32)) >> 21
        phit += abar * int(k.sk.s0[i])
    rot = O.mul_xai32(tv, phit % 2048)
    assert np.max(np.abs(O.tdiff32(ph, rot))) < 1 << 25
