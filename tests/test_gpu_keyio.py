"""Key material on the GPU: device key-switching key generation and loading
the reference's key files into a context.

* fhe_keygen_ksk_dev (ikskgen, TFHEpp/src/cloudkey.cpp:54-70) must equal the
  host generator word for word (kinds 10, 20 at an n0 that is not a multiple
  of 4, so the noise words straddle Philox blocks; kind 21).
* fhe_keyfile_load (HEDB::load_*_key, src/HEDB/utils/serialize.cpp) and
  fhe_fastr_load (test/HE3_Query/fastR.h:40-72) must leave a context that
  computes bit-identically to one loaded from the same arrays directly.  The
  file formats themselves are checked against the reference's own cereal
  code in test_keyio.py.
"""

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import keyio as K

import _oracle as O

pytestmark = pytest.mark.gpu

N0 = 37


@pytest.fixture(scope="module")
def small():
    p = F.preset(F.PRESET_CFG636)
    p.n0 = N0
    sk = F.SecretKeys(p, 37001)
    return p, sk


@pytest.mark.parametrize("kind", [10, 20, 21])
def test_keygen_ksk_dev_bit_exact(kind, small):
    p, sk = small
    ctx = F.Context(p)
    host = F.keygen_ksk(sk, kind, 4242 + kind)
    dev = F.DeviceBuffer(host.nbytes)
    ctx.keygen_ksk(kind, sk, 4242 + kind, dev)
    got = dev.to_array(np.uint32, host.shape)
    dev.free()
    bad = np.flatnonzero(got.reshape(-1) != host.reshape(-1))
    assert bad.size == 0, f"{bad.size} words differ, first at {bad[:4]}"
    # into the context slot: key switching with it equals the uploaded key's
    ctx.keygen_ksk(kind, sk, 4242 + kind)
    ref = F.Context(p)
    ref.load_ksk(kind, host)
    rng = np.random.default_rng(kind)
    if kind == 10:
        x = rng.integers(0, 2 ** 32, (33, F.LVL1_N + 1), dtype=np.uint64).astype(np.uint32)
    else:
        x = rng.integers(0, 2 ** 63, (33, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2)
    assert np.array_equal(ctx.iks(kind, x), ref.iks(kind, x))
    ctx.close()
    ref.close()


def _pbs_inputs(p, sk, B=64):
    rng = np.random.default_rng(5)
    m = rng.integers(1, 8, B)  # phases inside (0, 1/2): the constant LUT gives +mu
    c0 = F.tlwe_encrypt_u32(777, 0, sk.s0, (m << 28).astype(np.uint32), p.alpha0)
    tv = np.full(F.LVL1_N, F.LVL1_MU, np.uint32)
    return c0, tv


def test_keyfile_load_matches_direct_upload(small, tmp_path):
    p, sk = small
    bk1 = F.keygen_bk_lvl01(sk, 11)
    bk2 = F.keygen_bk_lvl02(sk, 12)
    ksk = {10: F.keygen_ksk(sk, 10, 13), 20: F.keygen_ksk(sk, 20, 14),
           21: F.keygen_ksk(sk, 21, 15)}
    files = {K.BKLVL01: bk1, K.BKLVL02: bk2, K.IKSKLVL10: ksk[10], K.IKSKLVL20: ksk[20],
             K.IKSKLVL21: ksk[21]}
    K.save_evalkey_dir(str(tmp_path), p, sk, files)
    q, s0, s1, s2 = K.read_secret_key(str(tmp_path / "secret.key"))
    assert q.n0 == N0 and np.array_equal(s0, sk.s0) and np.array_equal(s2, sk.s2)

    a = F.Context(p)
    loaded = K.load_evalkey_dir(a, str(tmp_path))
    assert sorted(loaded) == sorted(files)
    b = F.Context(p)
    b.load_bk(1, bk1, F.BK_INT)
    b.load_bk(2, bk2, F.BK_INT)
    for kind, k in ksk.items():
        b.load_ksk(kind, k)

    c0, tv = _pbs_inputs(p, sk)
    assert np.array_equal(a.pbs_lvl01(c0, tv), b.pbs_lvl01(c0, tv))
    rng = np.random.default_rng(9)
    x1 = rng.integers(0, 2 ** 32, (40, F.LVL1_N + 1), dtype=np.uint64).astype(np.uint32)
    x2 = rng.integers(0, 2 ** 63, (40, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2)
    assert np.array_equal(a.iks(10, x1), b.iks(10, x1))
    assert np.array_equal(a.iks(20, x2), b.iks(20, x2))
    assert np.array_equal(a.iks(21, x2), b.iks(21, x2))
    # lvl02 PBS through both keys
    tv2 = np.full(F.LVL2_N, 1 << 61, np.uint64)
    assert np.array_equal(a.pbs_lvl02(c0[:32], tv2), b.pbs_lvl02(c0[:32], tv2))

    # FFT-form lvl1 key file (TFHEpp BootstrappingKeyFFT): the FD values are
    # used as given (as LoadBK does), so the result equals an upload of the
    # same FD array, and decrypts like the integer key's
    bkfft = O.bk_to_fft32(bk1)
    K.write_key(str(tmp_path / "bkfftlvl01.key"), K.BKFFTLVL01, p, bkfft)
    K.load_key(a, str(tmp_path / "bkfftlvl01.key"), K.BKFFTLVL01)
    b.load_bk(1, bkfft, F.BK_FFT)
    out = a.pbs_lvl01(c0, tv)
    assert np.array_equal(out, b.pbs_lvl01(c0, tv))
    ph = F.tlwe_phase_u32(out, sk.s1).view(np.int32).astype(np.int64)
    assert np.all(np.abs(ph - F.LVL1_MU) < 1 << 26)

    # a file for another n0 is refused
    p636 = F.preset(F.PRESET_CFG636)
    K.write_key(str(tmp_path / "other.key"), K.BKLVL01, p636, np.zeros((636, 6, 2, 1024),
                                                                         np.uint32))
    with pytest.raises(F.FheError):
        K.load_key(a, str(tmp_path / "other.key"), K.BKLVL01)
    a.close()
    b.close()


def test_fastr_load(small, tmp_path):
    p, sk = small
    bk = F.keygen_bk_lvl01(sk, 21)
    ksk10 = F.keygen_ksk(sk, 10, 22)
    path = str(tmp_path / "key.bin")
    K.write_fastr(path, sk.s1, O.bk_to_fft32(bk), ksk10, N0)
    a = F.Context(p)
    s1 = K.load_fastr(a, path)
    assert np.array_equal(s1, sk.s1)
    b = F.Context(p)
    b.load_bk(1, O.bk_to_fft32(bk), F.BK_FFT)
    b.load_ksk(10, ksk10)
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2, (2, 48))
    mu, neg_mu = np.uint32(F.LVL1_MU), np.uint32((1 << 32) - F.LVL1_MU)
    ca = F.tlwe_encrypt_u32(31, 0, sk.s1, np.where(bits[0] == 1, mu, neg_mu).astype(np.uint32),
                            p.alpha1)
    cb = F.tlwe_encrypt_u32(32, 0, sk.s1, np.where(bits[1] == 1, mu, neg_mu).astype(np.uint32),
                            p.alpha1)
    ga = a.gate(F.GATE_NAND, ca, cb)
    assert np.array_equal(ga, b.gate(F.GATE_NAND, ca, cb))
    ph = F.tlwe_phase_u32(ga, sk.s1).view(np.int32)
    assert np.array_equal(ph > 0, ~((bits[0] == 1) & (bits[1] == 1)))
    with open(path, "ab") as f:
        f.write(b"\0")
    with pytest.raises(F.FheError):
        K.load_fastr(a, path)
    a.close()
    b.close()
