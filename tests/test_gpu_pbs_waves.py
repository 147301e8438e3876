"""The lvl1 blind rotation with two waves per ciphertext (small batches,
csrc/pbs_lvl1w2.hip) against the one-wave kernel and the oracle.

Both shapes compute TFHEpp's BlindRotate (gatebootstrapping.hpp:15-43) in
FP64 with different operation orders, so the bar is the one of every
bootstrap here: decrypt-exact, phase within 2^-6 of the oracle's, and equal
noise statistics; the accumulator (no sample extraction) must sample-extract
to the PBS output bit for bit.  The shape is chosen per batch size
(FHE_TUNE_PBS1_WAVES; untuned: two waves below 512 ciphertexts).
"""

import numpy as np
import pytest

import cufhedb_amd as F

import _oracle as O
from test_gpu_lvl1 import z8_testvector

pytestmark = pytest.mark.gpu
KNOB = F.TUNE_PBS1_WAVES


def _run(ctx, waves, c0, tv):
    ctx.set_tune_table([(1, waves)], knob=KNOB)
    try:
        return ctx.pbs_lvl01(c0, tv)
    finally:
        ctx.set_tune_table([], knob=KNOB)


@pytest.mark.parametrize("B", [1, 5, 64, 700])
def test_two_waves_decrypt_and_match_one_wave(keys636, ctx636, B):
    k = keys636
    rng = np.random.default_rng(B)
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector(lut)
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=70000 + B)
    want = (lut[m] << 28).astype(np.uint32)
    e1 = O.tdiff32(F.tlwe_phase_u32(_run(ctx636, 1, c0, tv), k.sk.s1), want)
    e2 = O.tdiff32(F.tlwe_phase_u32(_run(ctx636, 2, c0, tv), k.sk.s1), want)
    assert np.max(np.abs(e2)) < 1 << 27          # decrypts to LUT[m]
    assert np.max(np.abs(e2 - e1)) < 1 << 26     # same ciphertext up to the noise
    if B >= 64:
        ratio = e2.std() / e1.std()
        assert 0.8 < ratio < 1.25, ratio


def test_two_waves_vs_oracle_and_accumulator(keys636, ctx636):
    k = keys636
    rng = np.random.default_rng(77)
    B = 24
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector(lut)
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=71000)
    ctx636.set_tune_table([(1, 2)], knob=KNOB)
    try:
        out = ctx636.pbs_lvl01(c0, tv)
        acc = ctx636.blind_rotate_lvl01(c0, tv)
        # per-ciphertext test vectors through the same kernel
        tvs = np.stack([z8_testvector(np.roll(lut, b)) for b in range(B)])
        dev_in, dev_tv = F.DeviceBuffer.from_array(c0), F.DeviceBuffer.from_array(tvs)
        dev_out = F.DeviceBuffer(B * 1025 * 4)
        F.check(F.lib().fhe_pbs_lvl01_dev(ctx636.h, dev_in.ptr, dev_tv.ptr, F.LVL1_N,
                                          dev_out.ptr, B, None))
        ctx636.sync()
        out_pc = dev_out.to_array(np.uint32, (B, 1025))
    finally:
        ctx636.set_tune_table([], knob=KNOB)
    assert np.array_equal(ctx636.sample_extract_lvl1(acc, 0), out)
    ph = F.tlwe_phase_u32(out, k.sk.s1)
    for b in (0, 11, 23):
        ref = F.tlwe_phase_u32(O.pbs32(c0[b], k.bkfft, tv)[None], k.sk.s1)[0]
        assert abs(int(O.tdiff32(ph[b], ref))) < 1 << 26
    want_pc = np.array([np.roll(lut, b)[m[b]] for b in range(B)], np.int64) << 28
    e = O.tdiff32(F.tlwe_phase_u32(out_pc, k.sk.s1), want_pc.astype(np.uint32))
    assert np.max(np.abs(e)) < 1 << 27
    for buf in (dev_in, dev_tv, dev_out):
        buf.free()


def test_autotune_waves_table(keys636, ctx636):
    try:
        res = ctx636.autotune(batches=(64, 2048), knob=KNOB)
        print("\n  lvl1 waves per ciphertext:", res)
        table = ctx636.tune_table(knob=KNOB)
        assert [b for b, _ in table] == [64, 2048]
        assert all(w in (1, 2) for _, w in table)
    finally:
        ctx636.set_tune_table([], knob=KNOB)
    with pytest.raises(F.FheError):
        ctx636.set_tune_table([(1, 3)], knob=KNOB)
    assert ctx636.tune_table(knob=KNOB) == []


def test_two_wave_transpose_forms_agree(keys636, ctx636):
    """Two-wave batches up to 256 run the lane transform's first transpose as
    register swaps, larger ones through LDS; both move the same words, so a
    256- and a 257-ciphertext batch agree word for word on their common
    ciphertexts."""
    k = keys636
    rng = np.random.default_rng(257)
    m = rng.integers(0, 8, 257)
    tv = z8_testvector(rng.integers(0, 8, 8))
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=77000)
    ctx636.set_tune_table([(1, 2)], knob=KNOB)
    try:
        big = ctx636.pbs_lvl01(c0, tv)
        small = ctx636.pbs_lvl01(c0[:256], tv)
    finally:
        ctx636.set_tune_table([], knob=KNOB)
    assert np.array_equal(big[:256], small)
