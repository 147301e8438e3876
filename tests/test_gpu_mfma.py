"""GPU: the split-MFMA lvl1 PBS variant (forward digit transforms as fp16
matrix products with hi/lo limbs and FP32 accumulation; the key split into
three integer limbs K = K0 + 2^16 K1 + 2^24 K2, an FP32 frequency-domain MAC
and one FP64 inverse transform per limb, limb products rounded and
recombined mod 2^32).

What is pinned here:
* the forward transform against TwistIFFT (mulfft.hpp:113-121) through the
  oracle: relative RMS error below 2^-22 (FP32 accumulation bounds it);
* one CMUX step (a context with n0 = 1) against the exact integer
  computation (oracle decomposition, rotation and naive negacyclic products
  mod 2^32): the limb products rounded exactly leave only limb 0's
  arithmetic error, a few units per coefficient (the error model puts its
  RMS at 2^2.6); a key kept whole in FP32 would be off by ~2^18 there, a
  wrong limb shift or sign by ~2^16 and more;
* the full blind rotation over a batch of 8192 Z_8 LUT bootstraps at
  n0 = 636: every output decrypts to LUT[m], and the phase-error std equals
  the FP64 path's on the same inputs (the arithmetic error per CMUX is far
  below the key noise);
* the 10^6-bootstrap failure count is in test_gpu_failures.py.
"""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

pytestmark = pytest.mark.gpu
N, DELTA8 = 1024, 1 << 28


def z8_testvector(lut):
    j = np.arange(N)
    box = (j + 64) >> 7
    vals = np.array(lut, np.int64)[box % 8] << 28
    vals = np.where(box == 8, -vals, vals)
    return (vals % (1 << 32)).astype(np.uint32)


def _ctx(params, bk, variant):
    ctx = F.Context(params)
    ctx.load_bk(1, bk, F.BK_INT)
    ctx.set_pbs_variant(variant)
    return ctx


@pytest.fixture(scope="module")
def mctx636(keys636):
    ctx = _ctx(keys636.params, keys636.bk, F.PBS_MFMA_SPLIT)
    yield ctx
    ctx.close()


def test_mfma_forward_transform_vs_twist_ifft(mctx636):
    # the variant's forward transform on gadget-digit polynomials and on
    # wider small integers, against the FP64 TwistIFFT of the oracle
    rng = np.random.default_rng(77)
    for lim in (32, 2048):
        a = rng.integers(-lim, lim, (16, N)).astype(np.int32)
        got = mctx636.fwd_digits_lvl1_mfma(a)
        want = np.stack([O.ifft32(x.astype(np.uint32)) for x in a])
        err = got - want
        rel = np.sqrt(np.mean(err ** 2) / np.mean(want ** 2))
        print(f"split-MFMA transform |x| < {lim}: rms relative error 2^{np.log2(rel):.2f}, "
              f"max abs {np.max(np.abs(err)):.3g} of rms {np.sqrt(np.mean(want ** 2)):.3g}")
        assert rel < 2.0 ** -22


def _one_cmux_exact(c0, bk, tv):
    """acc after one CMUX of the blind rotation (n0 = 1), exactly: the test
    vector rotated by -b-bar, then acc += BK_0 (x) ((X^a-bar - 1) acc) with
    the external product as exact negacyclic products mod 2^32
    (TFHEpp gatebootstrapping.hpp:15-43, detwfa.cpp:21-30, trgsw.cpp:9-103)."""
    a, b = int(c0[0]), int(c0[1])
    bbar = 2 * N - (b >> 21)
    abar = ((a + (1 << 20)) & 0xFFFFFFFF) >> 21
    acc = np.zeros((2, N), np.uint32)
    acc[1] = O.mul_xai32(tv, bbar % (2 * N))
    if abar == 0:
        return acc
    out = acc.copy()
    for c in range(2):
        rot = O.mul_xai32(acc[c], abar, minus_one=True)
        for d in range(3):
            dig = O.decompose32(rot, d)
            for oc in range(2):
                out[oc] += O.naive32(dig, bk[0, 3 * c + d, oc])
    return out


def test_mfma_one_cmux_vs_exact(keys636):
    p = F.preset(F.PRESET_CFG636)
    p.n0 = 1
    sk = F.SecretKeys(p, 4711)
    bk = F.keygen_bk_lvl01(sk, 4712)
    rng = np.random.default_rng(4713)
    B = 24
    c0 = rng.integers(0, 1 << 32, (B, 2), dtype=np.uint64).astype(np.uint32)
    c0[0, 0] = 0  # a-bar = 0: the step is skipped
    tv = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    want = np.stack([_one_cmux_exact(c0[i], bk, tv) for i in range(B)])
    errs = {}
    for name, variant in (("fp64", F.PBS_FP64), ("mfma", F.PBS_MFMA_SPLIT)):
        ctx = _ctx(p, bk, variant)
        got = ctx.blind_rotate_lvl01(c0, tv)
        ctx.close()
        errs[name] = O.tdiff32(got.reshape(B, -1), want.reshape(B, -1)).astype(np.int64)
    for name, e in errs.items():
        print(f"one CMUX, {name} - exact: max |err| {np.max(np.abs(e))}, "
              f"rms {np.sqrt(np.mean(e.astype(np.float64) ** 2)):.2f}")
    assert np.array_equal(errs["mfma"][0], np.zeros(2 * N, np.int64))
    assert np.max(np.abs(errs["fp64"])) <= 2
    # limb 0's FP32 error (RMS ~2^2.6) is all that remains
    assert np.max(np.abs(errs["mfma"])) < 1 << 7
    assert np.sqrt(np.mean(errs["mfma"].astype(np.float64) ** 2)) < 2 ** 4


def test_mfma_pbs_decrypts_with_fp64_noise(keys636, ctx636, mctx636):
    # same 8192 inputs through both variants; error against the exact LUT value
    k = keys636
    B = 8192
    rng = np.random.default_rng(424242)
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector(lut)
    c0 = k.enc_lvl0((m * DELTA8).astype(np.uint32), index0=9200)
    want = (lut[m] * DELTA8).astype(np.int64)
    errs, dec, ph = {}, {}, {}
    for name, ctx in (("fp64", ctx636), ("mfma", mctx636)):
        ph[name] = F.tlwe_phase_u32(ctx.pbs_lvl01(c0, tv), k.sk.s1)
        p = ph[name].view(np.int32).astype(np.int64)
        errs[name] = ((p - want + (1 << 31)) % (1 << 32)) - (1 << 31)
        dec[name] = np.round(p / DELTA8).astype(np.int64) % 16
    sf, sm = np.std(errs["fp64"]), np.std(errs["mfma"])
    gap = O.tdiff32(ph["mfma"], ph["fp64"]).astype(np.float64)
    print(f"phase-error std: fp64 2^{np.log2(sf):.2f}, split-MFMA 2^{np.log2(sm):.2f}; "
          f"max |err| split-MFMA 2^{np.log2(np.max(np.abs(errs['mfma']))):.2f} (margin 2^27); "
          f"split-MFMA - fp64 phase gap std 2^{np.log2(np.std(gap)):.2f}")
    assert np.array_equal(dec["fp64"], lut[m])
    assert np.array_equal(dec["mfma"], lut[m])
    # equal noise: the std of a 8192-sample std is ~0.8 %
    assert abs(np.log2(sm / sf)) < 0.1
    assert abs(np.mean(errs["mfma"])) < 4 * sm / np.sqrt(B)


def test_mfma_variant_selection(keys636):
    ctx = F.Context(keys636.params)
    ctx.set_pbs_variant(F.PBS_MFMA_SPLIT)  # before a key: the layout is built at load
    ctx.load_bk(1, keys636.bk, F.BK_INT)
    rng = np.random.default_rng(5)
    m = rng.integers(0, 8, 64)
    lut = rng.integers(0, 8, 8)
    out = ctx.pbs_lvl01(keys636.enc_lvl0((m * DELTA8).astype(np.uint32)), z8_testvector(lut))
    ph = F.tlwe_phase_u32(out, keys636.sk.s1).view(np.int32).astype(np.int64)
    assert np.array_equal(np.round(ph / DELTA8).astype(np.int64) % 16, lut[m])
    ctx.set_pbs_variant(F.PBS_FP64)
    with pytest.raises(F.FheError):
        ctx.set_pbs_variant(7)
    ctx.close()
