"""GPU parity: lvl1 FP64 path (HIP kernels through the C ABI) vs the oracle.

Bars (FP64 FFT-bearing outputs cannot be bit-identical to a different FP64
operation order; the reference's own GPU differs from its CPU by +-1 torus
unit, test/Arc_Query/fft_res.txt):
  * TwistIFFT on the GPU vs oracle: |dFD| <= 1e-13 * max|FD|
  * TwistFFT on the GPU vs oracle: torus difference <= 1 (truncation edge)
  * external product vs oracle: <= 2 torus units per coefficient
  * sample extract: bit-exact
  * PBS / blind rotation: decrypted plaintexts bit-exact vs the LUT, output
    phase within 2^-6 of the oracle's phase on the same keys and inputs (two
    FP64 implementations accumulate independent noise after their first low-digit
    divergence; measured gap std 2^22.8, noise std 2^22.3 -- see
    tests/test_oracle_golden.py and tests/test_gpu_noise.py).
"""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

pytestmark = pytest.mark.gpu

N = 1024
DELTA8 = 1 << 28  # Z_8 with one padding bit
PHASE_TOL = 1 << 26  # 2^-6 of the torus: ~9 sigma of the measured FP64-vs-FP64 phase gap


def z8_testvector(lut):
    """tv[j] = LUT[round(j/128)] * 2^28, negacyclic wrap for the last half box."""
    j = np.arange(N)
    box = (j + 64) >> 7
    vals = np.array(lut, np.int64)[box % 8] * DELTA8
    vals = np.where(box == 8, -vals, vals)
    return (vals % (1 << 32)).astype(np.uint32)


def test_gpu_ifft_matches_oracle(ctx672, rng):
    polys = rng.integers(0, 1 << 32, (16, N), dtype=np.uint64).astype(np.uint32)
    got = ctx672.ifft_lvl1(polys)
    for b in range(polys.shape[0]):
        want = O.ifft32(polys[b])
        err = np.max(np.abs(got[b] - want)) / np.max(np.abs(want))
        assert err <= 1e-13, (b, err)


def test_gpu_fft_matches_oracle_and_roundtrip(ctx672, rng):
    polys = rng.integers(0, 1 << 32, (16, N), dtype=np.uint64).astype(np.uint32)
    fds = np.stack([O.ifft32(p) for p in polys])
    got = ctx672.fft_lvl1(fds)
    for b in range(polys.shape[0]):
        assert np.max(np.abs(O.tdiff32(got[b], O.fft32(fds[b])))) <= 1
        # TFHEpp/test/polymul.cpp:19-28 bound: FFT round trip within 1
        assert np.max(np.abs(O.tdiff32(got[b], polys[b]))) <= 1


def test_gpu_polymul_bound(ctx672, rng):
    """digit poly x torus poly through GPU FFTs vs exact product (polymul.cpp:31-51)."""
    for _ in range(4):
        a = (rng.integers(0, 64, N) - 32).astype(np.int64).astype(np.uint32)
        b = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
        fa, fb = ctx672.ifft_lvl1(a)[0], ctx672.ifft_lvl1(b)[0]
        M = N // 2
        prod = np.empty(N)
        prod[:M] = fa[:M] * fb[:M] - fa[M:] * fb[M:]
        prod[M:] = fa[:M] * fb[M:] + fa[M:] * fb[:M]
        got = ctx672.fft_lvl1(prod)[0]
        assert np.max(np.abs(O.tdiff32(got, O.naive32(a, b)))) <= 1


def test_gpu_external_product(keys672, ctx672, rng):
    s1 = keys672.sk.s1
    B = 8
    msgs = rng.integers(0, 8, (B, N)).astype(np.int64) * DELTA8
    trlwe = F.trlwe_encrypt_u32(11, 0, s1, msgs.astype(np.uint32), 2.0 ** -25)
    for m in (1, 0, -1):
        trgsw = F.trgsw_encrypt_u32(12, m + 5, s1, m, 2.0 ** -25)
        fd = O.trgsw_to_fft32(trgsw)
        got = ctx672.external_product_lvl1(trlwe, fd)
        for b in range(B):
            want = O.external_product32(trlwe[b], fd)
            assert np.max(np.abs(O.tdiff32(got[b], want))) <= 2
            ph = O.trlwe_phase32(got[b], s1)
            dec = np.round(O.tdiff32(ph, 0) / DELTA8).astype(np.int64) % 16
            assert np.array_equal(dec, (m * msgs[b] // DELTA8) % 16)


def test_gpu_sample_extract_bit_exact(ctx672, rng):
    trlwe = rng.integers(0, 1 << 32, (6, 2, N), dtype=np.uint64).astype(np.uint32)
    for idx in (0, 1, 511, 1023):
        got = ctx672.sample_extract_lvl1(trlwe, idx)
        for b in range(trlwe.shape[0]):
            want = np.zeros(N + 1, np.uint32)
            O.orc().orc_sample_extract_u32(O.P(want), O.P(np.ascontiguousarray(trlwe[b])), idx, N)
            assert np.array_equal(got[b], want)


@pytest.mark.parametrize("which", [F.PRESET_REF672, F.PRESET_CFG636])
def test_gpu_pbs_z8_lut(which, rng):
    from conftest import keys_for, ctx_for
    k, ctx = keys_for(which), ctx_for(which)
    B = 64
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector(lut)
    c0 = k.enc_lvl0((m * DELTA8).astype(np.uint32), index0=1000)
    out = ctx.pbs_lvl01(c0, tv)
    ph = F.tlwe_phase_u32(out, k.sk.s1)
    dec = np.round(O.tdiff32(ph, 0) / DELTA8).astype(np.int64) % 16
    assert np.array_equal(dec, lut[m])
    # phase parity against the oracle on the first few ciphertexts
    for b in range(3):
        want = O.pbs32(c0[b], k.bkfft, tv)
        dphi = O.tdiff32(ph[b], F.tlwe_phase_u32(want, k.sk.s1)[0])
        assert abs(int(dphi)) <= PHASE_TOL, (b, int(dphi))


def test_gpu_pbs_per_ciphertext_testvectors(keys636, ctx636, rng):
    B = 16
    m = rng.integers(0, 8, B)
    luts = rng.integers(0, 8, (B, 8))
    tvs = np.stack([z8_testvector(l) for l in luts])
    c0 = keys636.enc_lvl0((m * DELTA8).astype(np.uint32), index0=5000)
    out = ctx636.pbs_lvl01(c0, tvs)
    ph = F.tlwe_phase_u32(out, keys636.sk.s1)
    dec = np.round(O.tdiff32(ph, 0) / DELTA8).astype(np.int64) % 16
    assert np.array_equal(dec, luts[np.arange(B), m])


def test_gpu_blind_rotate_phase(keys672, ctx672, rng):
    B = 2
    m = rng.integers(0, 8, B)
    tv = z8_testvector(rng.integers(0, 8, 8))
    c0 = keys672.enc_lvl0((m * DELTA8).astype(np.uint32), index0=7000)
    acc = ctx672.blind_rotate_lvl01(c0, tv)
    for b in range(B):
        want = O.blind_rotate32(c0[b], keys672.bkfft, tv)
        dphi = O.tdiff32(O.trlwe_phase32(acc[b], keys672.sk.s1),
                         O.trlwe_phase32(want, keys672.sk.s1))
        assert np.max(np.abs(dphi)) <= PHASE_TOL
        # sample extract of the GPU accumulator equals the GPU PBS output
        assert np.array_equal(O.sample_extract32(acc[b]), ctx672.pbs_lvl01(c0[b:b + 1], tv)[0])


def test_gpu_pbs_trivial_and_boundary_masks(keys672, ctx672, rng):
    """Edge inputs of the blind rotation.
    * Trivial ciphertexts (mask 0): every a-bar is 0, every CMUX step is the
      reference's skip, so the output is exactly SampleExtract(0, X^-bbar tv):
      a zero mask and body tv rotated -- bit-exact, independent of the FFT.
    * Mask words on both sides of the a-bar rounding boundary (2^20 - 1 -> 0,
      2^20 -> 1, and the top of the torus): phase against the oracle.
    * An empty batch is a no-op."""
    N = F.LVL1_N
    n0 = keys672.n0
    tv = z8_testvector(rng.integers(0, 8, 8))
    bodies = np.array([0, 1, 1 << 21, (1 << 21) - 1, 0x80000000, 0xFFFFFFFF, 1234 << 21],
                      np.uint32)
    c0 = np.zeros((len(bodies), n0 + 1), np.uint32)
    c0[:, n0] = bodies
    out = ctx672.pbs_lvl01(c0, tv)
    for b, body in enumerate(bodies):
        bbar = (2 * N - (int(body) >> 21)) % (2 * N)
        # coefficient 0 of X^bbar tv (negacyclic): tv[-bbar mod 2N], negated
        # when the index wraps past N
        idx = (0 - bbar) % (2 * N)
        want = tv[idx] if idx < N else (0 - int(tv[idx - N])) % (1 << 32)
        assert np.all(out[b, :N] == 0), b
        assert int(out[b, N]) == int(want), b
    # masks at the a-bar rounding boundary
    m = rng.integers(0, 8, 4)
    c = keys672.enc_lvl0((m * DELTA8).astype(np.uint32), index0=9100)
    c[0, :n0] = (1 << 20) - 1
    c[1, :n0] = 1 << 20
    c[2, : n0 // 2] = 0xFFFFFFFF
    c[3, ::3] = 0
    got = ctx672.pbs_lvl01(c, tv)
    for b in range(4):
        ref = O.pbs32(c[b], keys672.bkfft, tv)
        dphi = O.tdiff32(F.tlwe_phase_u32(got[b:b + 1], keys672.sk.s1)[0],
                         F.tlwe_phase_u32(ref, keys672.sk.s1)[0])
        assert abs(int(dphi)) <= PHASE_TOL, (b, int(dphi))
    # empty batch
    d_in = F.DeviceBuffer(4 * (n0 + 1))
    d_tv = F.DeviceBuffer.from_array(tv)
    d_out = F.DeviceBuffer(4 * (N + 1))
    F.check(F.lib().fhe_pbs_lvl01_dev(ctx672.h, d_in.ptr, d_tv.ptr, 0, d_out.ptr, 0, None))
    ctx672.sync()
    for buf in (d_in, d_tv, d_out):
        buf.free()
