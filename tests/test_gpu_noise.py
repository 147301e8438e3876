"""GPU vs oracle output-noise statistics for the lvl1 PBS (n0 = 636 set).

A per-sample phase comparison between two FP64 implementations is only
meaningful up to the noise (they accumulate independent rounding noise once
a low gadget digit flips), so this checks the distribution: over B outputs
the GPU's phase-error std must match the oracle's (computed on the same keys
and a subsample of the same inputs) within the statistical error, its mean
must be centred, and no output may decrypt wrongly.  A lost bit of FFT
precision, a wrong rounding offset or a biased truncation would move one of
these by many standard errors.
"""

import concurrent.futures as cf
import os

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O
from test_gpu_lvl1 import z8_testvector

pytestmark = pytest.mark.gpu


def test_gpu_pbs_noise_matches_oracle(keys636, ctx636):
    k = keys636
    rng = np.random.default_rng(31337)
    B = 4096
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector(lut)
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=50000)
    out = ctx636.pbs_lvl01(c0, tv)
    want = (lut[m] << 28).astype(np.uint32)
    e_gpu = O.tdiff32(F.tlwe_phase_u32(out, k.sk.s1), want)
    assert np.max(np.abs(e_gpu)) < 1 << 27  # no decryption failure
    S = 256
    threads = min(16, os.cpu_count() or 1)
    bkfft = k.bkfft
    with cf.ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(lambda b: O.pbs32(c0[b], bkfft, tv), range(S)))
    e_orc = O.tdiff32(F.tlwe_phase_u32(np.stack(outs), k.sk.s1), want[:S])
    s_gpu, s_orc = e_gpu.std(), e_orc.std()
    ratio = s_gpu / s_orc
    print(f"phase-error std: gpu 2^{np.log2(s_gpu):.2f} over {B}, "
          f"oracle 2^{np.log2(s_orc):.2f} over {S}, ratio {ratio:.3f}, "
          f"gpu mean {e_gpu.mean():.0f}")
    assert 0.85 < ratio < 1.17  # ~3 standard errors of a 256-sample std
    assert abs(e_gpu.mean()) < 4 * s_gpu / np.sqrt(B)
    # per-sample phase gap GPU - oracle on the same inputs and keys: two
    # independent noise draws once a low digit differs, so std ~ sqrt(2) x
    gap = O.tdiff32(F.tlwe_phase_u32(out[:S], k.sk.s1), F.tlwe_phase_u32(np.stack(outs), k.sk.s1))
    print(f"GPU - oracle phase gap over {S}: std 2^{np.log2(gap.std()):.2f}, "
          f"|gap| histogram (2^-32 units) {log2_hist(gap)}")
    assert np.max(np.abs(gap)) < 1 << 26


def log2_hist(e, lo=18, hi=27):
    """Counts of |e| per octave: [0, 2^lo), [2^lo, 2^(lo+1)), ..., [2^hi, inf)."""
    a = np.abs(np.asarray(e, np.int64))
    edges = [0] + [1 << j for j in range(lo, hi + 1)] + [1 << 62]
    counts = np.histogram(a, bins=edges)[0]
    labels = [f"<2^{lo}"] + [f"2^{j}" for j in range(lo, hi)] + [f">=2^{hi}"]
    return {lab: int(c) for lab, c in zip(labels, counts)}


def z8_testvector2(lut, N=2048):
    j = np.arange(N)
    box = (j + N // 16) // (N // 8)
    vals = np.array(lut, np.uint64)[box % 8] << np.uint64(60)
    return np.where(box == 8, np.uint64(0) - vals, vals).astype(np.uint64)


def test_gpu_pbs_lvl02_noise_vs_oracle(keys672):
    """lvl2 PBS (reference parameter set): the GPU applies X^a - 1 in the
    frequency domain after the key product, ACC += (X^a - 1) ExtProd(ACC),
    where the reference (and the oracle) compute ACC += ExtProd((X^a - 1)
    ACC).  The two are equal in exact arithmetic; in FP64 the decomposition
    rounding error and the key noise get multiplied by X^a - 1 (norm
    sqrt(2)), so the GPU's output noise std is sqrt(2) times the oracle's.
    That ratio is what is asserted: a lost bit of FFT precision or a wrong
    rotation would move it far off, and outputs must all decrypt."""
    k = keys672
    bk2 = F.keygen_bk_lvl02(k.sk, 672500)
    ctx = F.Context(k.params)
    ctx.load_bk(2, bk2, F.BK_INT)
    rng = np.random.default_rng(7272)
    B, S = 2048, 128
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector2(lut)
    c0 = k.enc_lvl0((m << 28).astype(np.uint32), index0=70000)
    want = lut[m].astype(np.uint64) << np.uint64(60)
    e_gpu = O.tdiff64(O.tlwe_phase64(ctx.pbs_lvl02(c0, tv), k.sk.s2), want).astype(np.float64)
    ctx.close()
    assert np.max(np.abs(e_gpu)) < 2.0 ** 59  # no decryption failure
    bkfft2 = O.bk_to_fft64(bk2)
    threads = min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(threads) as ex:
        outs = np.stack(list(ex.map(lambda b: O.pbs64(c0[b], bkfft2, tv), range(S))))
    e_orc = O.tdiff64(O.tlwe_phase64(outs, k.sk.s2), want[:S]).astype(np.float64)
    s_gpu, s_orc = e_gpu.std(), e_orc.std()
    ratio = s_gpu / s_orc
    print(f"lvl2 phase-error std (2^-64 units): gpu 2^{np.log2(s_gpu):.2f} over {B}, "
          f"oracle 2^{np.log2(s_orc):.2f} over {S}, ratio {ratio:.3f} (expected sqrt(2)); "
          f"max |gpu err| 2^{np.log2(np.max(np.abs(e_gpu))):.1f} against the 2^59 margin")
    assert 1.15 < ratio < 1.75  # sqrt(2) within ~3 standard errors of a 128-sample std
    assert abs(e_gpu.mean()) < 4 * s_gpu / np.sqrt(B)
