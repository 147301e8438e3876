"""Digit-wise comparison with a public threshold on the GPU
(fhe_compare_digits_dev): one lvl2 bootstrap per 4-bit digit plus one
combining bootstrap.

* The 20-bit workload at n0 = 636 / alpha0 = 2^-15: 2^14 seeded uniform
  values (plus the rows next to the threshold) against a seeded public
  threshold, x <= t, decrypted results bit-exact against the plaintext.
  The HomMSB-chain compare cannot run reliably at this parameter set (its
  6-bit windows leave ~3 sigma of key-switch noise); the digit scheme keeps
  2^-6 of margin at every bootstrap.
* Every operator and result type at the reference's n0 = 672, agreement
  with the HomMSB-chain compare (fhe_compare_dev against a trivially
  encrypted threshold) on the same values.
* Parity with the oracle's pipeline (tests/_oracle.py compare_digits: the
  oracle's key switch and FP64 bootstrap) at a small n0: results decrypt
  to the same bits and their phases agree to the bootstrap's FFT error.
* Edge shapes: 1-bit and 1-digit radixes, thresholds 0 and 2^(digits*bits)-1,
  an empty batch, argument checks.
"""

import numpy as np
import pytest

import cufhedb_amd as F

import _oracle as O

pytestmark = pytest.mark.gpu

OPS = {F.CMP_GT: np.greater, F.CMP_GE: np.greater_equal, F.CMP_LT: np.less,
       F.CMP_LE: np.less_equal, F.CMP_EQ: np.equal}
MU1 = 1 << 29


def _bits_lvl1(out, s1):
    ph = F.tlwe_phase_u32(out, s1).view(np.int32).astype(np.int64)
    return ph > 0, ph


def _near_threshold(rng, t, top, n):
    """Values at and around t: equal, +-1, and t with one digit changed."""
    vals = [t, t - 1, t + 1, 0, top]
    for i in range(5):
        for d in (-1, 1):
            vals.append(t + d * (1 << (4 * i)))
    vals += list(rng.integers(max(0, t - 40), min(top, t + 40) + 1, n))
    return np.array([v for v in vals if 0 <= v <= top], np.uint64)


def test_config3_20bit_le_n636(full636):
    k, ctx = full636
    rng = np.random.default_rng(3)
    top = (1 << 20) - 1
    t = int(rng.integers(0, top + 1))
    x = np.concatenate([rng.integers(0, top + 1, 1 << 14).astype(np.uint64),
                        _near_threshold(rng, t, top, 64)])
    ct = F.encrypt_digits(303, 0, k.sk.s2, x, 5, 4, k.params.alpha2)
    out = F.compare_digits(ctx, F.CMP_LE, ct, 5, 4, t)
    got, ph = _bits_lvl1(out, k.sk.s1)
    want = x <= np.uint64(t)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} wrong of {x.size}, t={t}, x={x[bad[:4]]}"
    # every result sits near +-mu (margin report: worst distance from +-mu)
    dev = np.abs(np.abs(ph) - MU1)
    assert dev.max() < MU1 // 4, dev.max()
    print(f"\n  {x.size} rows, t={t}: 0 errors, "
          f"max |phase - (+-mu)| = 2^{np.log2(dev.max() + 1):.2f}")


@pytest.mark.parametrize("op", list(OPS))
def test_every_operator_ref672(full672, op):
    k, ctx = full672
    rng = np.random.default_rng(40 + op)
    top = (1 << 20) - 1
    t = int(rng.integers(0, top + 1))
    x = np.concatenate([rng.integers(0, top + 1, 1024).astype(np.uint64),
                        _near_threshold(rng, t, top, 128)])
    ct = F.encrypt_digits(404 + op, 0, k.sk.s2, x, 5, 4, k.params.alpha2)
    want = OPS[op](x, np.uint64(t))
    # lvl1 LOGIC
    got, _ = _bits_lvl1(F.compare_digits(ctx, op, ct, 5, 4, t), k.sk.s1)
    assert np.array_equal(got, want)
    # lvl2 ARITHMETIC: true at 1/2, false at 0
    out2 = F.compare_digits(ctx, op, ct, 5, 4, t, out_level=2, result_type=F.ARITHMETIC)
    ph2 = O.tlwe_phase64(out2, k.sk.s2)
    d_true = np.abs(O.tdiff64(ph2, np.uint64(1 << 63)))
    d_false = np.abs(O.tdiff64(ph2, np.uint64(0)))
    assert np.array_equal(d_true < d_false, want)
    assert np.minimum(d_true, d_false).max() < 1 << 58
    # lvl1 ARITHMETIC (through IKS lvl21): 1/2 and 0 on the 32-bit torus
    out1a = F.compare_digits(ctx, op, ct, 5, 4, t, result_type=F.ARITHMETIC)
    ph1 = F.tlwe_phase_u32(out1a, k.sk.s1)
    assert np.array_equal(np.abs(O.tdiff32(ph1, 1 << 31)) < (1 << 30), want)


def test_matches_hom_msb_chain_compare(full672):
    """The reference's comparison (fhe_compare_dev, a HomMSB chain, with the
    threshold as a trivial lvl2 TLWE) and the digit-wise one agree."""
    k, ctx = full672
    rng = np.random.default_rng(77)
    top = (1 << 20) - 1
    t = int(rng.integers(0, top + 1))
    x = np.concatenate([rng.integers(0, top + 1, 256).astype(np.uint64),
                        _near_threshold(rng, t, top, 32)])
    pb = 20
    whole = F.tlwe_encrypt_u64(505, 0, k.sk.s2, x << np.uint64(64 - pb - 1), k.params.alpha2)
    thr = np.zeros(F.LVL2_N + 1, np.uint64)
    thr[-1] = np.uint64(t) << np.uint64(64 - pb - 1)
    chain, _ = _bits_lvl1(F.compare(ctx, 2, F.CMP_LE, whole, thr, pb), k.sk.s1)
    ct = F.encrypt_digits(506, 0, k.sk.s2, x, 5, 4, k.params.alpha2)
    digit, _ = _bits_lvl1(F.compare_digits(ctx, F.CMP_LE, ct, 5, 4, t), k.sk.s1)
    want = x <= np.uint64(t)
    assert np.array_equal(digit, want)
    assert np.array_equal(chain, want)


@pytest.fixture(scope="module")
def small():
    """n0 = 37: the oracle's key switch needs the KSK on the host."""
    p = F.preset(F.PRESET_CFG636)
    p.n0 = 37
    sk = F.SecretKeys(p, 37037)
    bk2 = F.keygen_bk_lvl02(sk, 1)
    ksk20 = F.keygen_ksk(sk, 20, 2)
    ksk21 = F.keygen_ksk(sk, 21, 3)
    ctx = F.Context(p)
    ctx.load_bk(2, bk2, F.BK_INT)
    ctx.load_ksk(20, ksk20)
    ctx.load_ksk(21, ksk21)
    yield p, sk, O.bk_to_fft64(bk2), ksk20, ksk21, ctx
    ctx.close()


@pytest.mark.parametrize("op", [F.CMP_LE, F.CMP_GT, F.CMP_EQ])
def test_parity_with_oracle(small, op):
    p, sk, bkfft2, ksk20, ksk21, ctx = small
    t = 0x8A3C5
    x = np.array([t, t + 1, t - 1, 0x8A3D5, 0x7FFFF, 0xFFFFF, 0, 0x8A2C5], np.uint64)
    ct = F.encrypt_digits(606 + op, 0, sk.s2, x, 5, 4, p.alpha2)
    want = OPS[op](x, np.uint64(t))
    for out_level, ksk in ((2, None), (1, ksk21)):
        got = F.compare_digits(ctx, op, ct, 5, 4, t, out_level=out_level)
        ref = O.compare_digits(ct, op, 5, 4, t, ksk20, bkfft2, p.n0, ksk21=ksk)
        if out_level == 2:
            pg, pr = O.tlwe_phase64(got, sk.s2), O.tlwe_phase64(ref, sk.s2)
            assert np.array_equal(pg.view(np.int64) > 0, want)
            assert np.array_equal(pr.view(np.int64) > 0, want)
            assert np.max(np.abs(O.tdiff64(pg, pr))) < 1 << 46
        else:
            pg, pr = F.tlwe_phase_u32(got, sk.s1), F.tlwe_phase_u32(ref, sk.s1)
            assert np.array_equal(pg.view(np.int32) > 0, want)
            assert np.array_equal(pr.view(np.int32) > 0, want)
            assert np.max(np.abs(O.tdiff32(pg, pr))) < 1 << 18


@pytest.mark.parametrize("digits,bits", [(1, 1), (1, 4), (2, 3), (5, 1), (3, 4)])
def test_small_radixes_exhaustive_thresholds(full672, digits, bits):
    k, ctx = full672
    top = (1 << (digits * bits)) - 1
    x = np.arange(top + 1, dtype=np.uint64)
    ct = F.encrypt_digits(700 + digits * 10 + bits, 0, k.sk.s2, x, digits, bits, k.params.alpha2)
    ts = sorted({0, top, top // 2, 1 % (top + 1)})
    for t in ts:
        for op in (F.CMP_LT, F.CMP_EQ):
            got, _ = _bits_lvl1(F.compare_digits(ctx, op, ct, digits, bits, t), k.sk.s1)
            assert np.array_equal(got, OPS[op](x, np.uint64(t))), (t, op)


def test_empty_batch_and_bad_arguments(full672):
    k, ctx = full672
    empty = np.zeros((0, 5, F.LVL2_N + 1), np.uint64)
    assert F.compare_digits(ctx, F.CMP_LE, empty, 5, 4, 7).shape == (0, F.LVL1_N + 1)
    x = F.encrypt_digits(1, 0, k.sk.s2, np.array([3], np.uint64), 5, 4, k.params.alpha2)
    with pytest.raises(ValueError):
        F.compare_digits(ctx, F.CMP_LE, x, 4, 4, 7)     # shape does not match digits
    # the C entry point checks its arguments before touching the buffers
    dx = F.DeviceBuffer.from_array(x)
    dout = F.DeviceBuffer(8 * (F.LVL2_N + 1))
    lib = F.lib()
    for op, digits, bits, t, level, rt in [(9, 5, 4, 7, 1, 1), (F.CMP_LE, 6, 4, 7, 1, 1),
                                           (F.CMP_LE, 0, 4, 7, 1, 1), (F.CMP_LE, 5, 5, 7, 1, 1),
                                           (F.CMP_LE, 5, 4, 1 << 20, 1, 1),
                                           (F.CMP_LE, 5, 4, 7, 3, 1), (F.CMP_LE, 5, 4, 7, 1, 2)]:
        rc = lib.fhe_compare_digits_dev(ctx.h, op, dx.ptr, digits, bits, t, level, dout.ptr, rt,
                                        1, None)
        assert rc == F.FHE_ERR_ARG, (op, digits, bits, t, level, rt, rc)
    assert lib.fhe_compare_digits_dev(ctx.h, F.CMP_LE, dx.ptr, 5, 4, 7, 1, dout.ptr, 1, -1,
                                      None) == F.FHE_ERR_ARG
    dx.free()
    dout.free()
    # a context without KSK 21 serves lvl2 outputs and refuses lvl1 ones
    bare = F.Context(k.params)
    bare.load_bk(2, F.keygen_bk_lvl02(k.sk, k.seed + 3), F.BK_INT)
    bare.keygen_ksk(20, k.sk, k.seed + 4)
    with pytest.raises(F.FheError):
        F.compare_digits(bare, F.CMP_LE, x, 5, 4, 7)
    out = F.compare_digits(bare, F.CMP_LE, x, 5, 4, 7, out_level=2)
    assert O.tlwe_phase64(out, k.sk.s2).view(np.int64)[0] > 0
    bare.close()


def test_key_switched_digit_noise_n636(full636):
    """The noise the digit bootstraps see: IKS lvl20 of the digit ciphertexts
    at n0 = 636 (predicted std 2^-9: 4096 key words of std 2^-15)."""
    k, ctx = full636
    rng = np.random.default_rng(8)
    x = rng.integers(0, 1 << 20, 512).astype(np.uint64)
    ct = F.encrypt_digits(808, 0, k.sk.s2, x, 5, 4, k.params.alpha2).reshape(-1, F.LVL2_N + 1)
    lvl0 = ctx.iks(20, ct)
    ph = F.tlwe_phase_u32(lvl0, k.sk.s0)
    m = (O.tlwe_phase64(ct, k.sk.s2) >> np.uint64(32)).astype(np.uint32)
    e = O.tdiff32(ph, m).astype(np.float64) / 2.0 ** 32
    sd = e.std()
    print(f"\n  IKS lvl20 noise std 2^{np.log2(sd):.2f}, max |e| 2^{np.log2(np.abs(e).max()):.2f}")
    assert sd < 2.0 ** -8.5
    assert np.abs(e).max() < 2.0 ** -6
