"""Key material files (host side, no GPU).

The reference's archive code (HEDB::generate_*_key / load_*_key over cereal,
src/HEDB/utils/serialize.cpp:7-320) is compiled into oracle/_ref from the
reference sources; these tests check, at the reference's parameter set
(n0 = 672), that the files this library writes are byte-identical to the
ones the reference writes for the same keys, that each side reads the
other's files back to the same arrays, and that big-endian archives (cereal's
BigEndian option) read back to the same arrays.  Malformed files are
rejected.  The fastR raw layout (test/HE3_Query/fastR.h:6-72) is checked at
a small n0 against its section offsets.
"""

import ctypes
import os

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import keyio as K

import _oracle as O

ref_only = pytest.mark.skipif(not O.ref_built(), reason="oracle/_ref not built (no reference sources)")


@pytest.fixture(scope="module")
def p672():
    return F.preset(F.PRESET_REF672)


@pytest.fixture(scope="module")
def sk672(p672):
    return F.SecretKeys(p672, 672001)


def _member(kind, n0, seed):
    dtype, shape = K.member_shape(kind, n0)
    rng = np.random.default_rng(seed)
    if dtype == np.float64:
        return rng.standard_normal(shape) * 2.0 ** 20
    return rng.integers(0, np.iinfo(dtype).max, size=shape, dtype=dtype, endpoint=True)


def _bytes(path):
    with open(path, "rb") as f:
        return f.read()


@ref_only
def test_secret_key_file_matches_reference(tmp_path, p672, sk672):
    ours, theirs = str(tmp_path / "ours.key"), str(tmp_path / "ref.key")
    K.write_secret_key(ours, p672, sk672)
    assert O.ref().ref_secret_write(theirs.encode(), O.P(sk672.s0), O.P(sk672.s1),
                                    O.P(sk672.s2), 0) == 0
    assert _bytes(ours) == _bytes(theirs)
    # 1 flag byte + keys + 136-byte lweParams
    assert os.path.getsize(ours) == 1 + 672 * 4 + 1024 * 4 + 2048 * 8 + 136
    q, s0, s1, s2 = K.read_secret_key(theirs)
    assert (q.n0, q.n1, q.l1, q.bgbit1, q.n2, q.l2, q.bgbit2) == (672, 1024, 3, 6, 2048, 4, 9)
    assert q.alpha0 == 2.0 ** -16 and q.alpha1 == 2.0 ** -25 and q.alpha2 == 2.0 ** -52
    assert (q.t10, q.basebit10, q.t21, q.basebit21, q.t22, q.basebit22) == (2, 10, 10, 3, 8, 4)
    assert np.array_equal(s0, sk672.s0) and np.array_equal(s1, sk672.s1)
    assert np.array_equal(s2, sk672.s2)
    r0, r1, r2 = np.zeros(672, np.uint32), np.zeros(1024, np.uint32), np.zeros(2048, np.uint64)
    n0 = ctypes.c_uint32()
    assert O.ref().ref_secret_read(ours.encode(), O.P(r0), O.P(r1), O.P(r2),
                                   ctypes.byref(n0)) == 0
    assert n0.value == 672
    assert np.array_equal(r0, sk672.s0) and np.array_equal(r1, sk672.s1)
    assert np.array_equal(r2, sk672.s2)


@ref_only
def test_secret_key_big_endian_archive(tmp_path, sk672):
    be = str(tmp_path / "be.key")
    O.ref().ref_secret_write(be.encode(), O.P(sk672.s0), O.P(sk672.s1), O.P(sk672.s2), 1)
    assert _bytes(be)[0] == 0
    q, s0, s1, s2 = K.read_secret_key(be)
    assert q.n0 == 672 and q.alpha1 == 2.0 ** -25
    assert np.array_equal(s0, sk672.s0) and np.array_equal(s2, sk672.s2)


@ref_only
@pytest.mark.parametrize("kind", [K.BKLVL01, K.BKFFTLVL01, K.BKLVL02, K.BKFFTLVL02, K.IKSKLVL21])
def test_member_file_matches_reference(kind, tmp_path, p672):
    key = _member(kind, 672, 9000 + kind)
    ours, theirs = str(tmp_path / "ours.key"), str(tmp_path / "ref.key")
    K.write_key(ours, kind, p672, key)
    assert O.ref().ref_keyfile_write(theirs.encode(), kind, O.P(key), 0) == 0
    a, b = _bytes(ours), _bytes(theirs)
    assert len(a) == 1 + 136 + 1 + key.nbytes
    assert a == b
    del a, b
    assert np.array_equal(K.read_key(theirs, kind), key)
    back = np.zeros_like(key)
    assert O.ref().ref_keyfile_read(ours.encode(), kind, O.P(back)) == 0
    assert np.array_equal(back, key)


@ref_only
@pytest.mark.parametrize("kind", [K.BKLVL02, K.BKFFTLVL01])
def test_member_big_endian_archive(kind, tmp_path):
    key = _member(kind, 672, 9100 + kind)
    be = str(tmp_path / "be.key")
    assert O.ref().ref_keyfile_write(be.encode(), kind, O.P(key), 1) == 0
    assert _bytes(be)[0] == 0
    assert np.array_equal(K.read_key(be, kind), key)


def test_keyfile_params_block(p672):
    q = K.params_of(p672)
    assert (q.n0, q.approx_bit0, q.nbit1, q.approx_bit1, q.nbit2, q.approx_bit2) == \
        (672, 32, 10, 32, 11, 64)
    assert q.alpha10 == q.alpha0 == p672.alpha0 and q.alpha21 == p672.alpha1
    assert q.alpha22 == p672.alpha2
    n = F.lib().fhe_keyfile_payload_bytes
    assert n(K.IKSKLVL10, 672) == 5_640_019_968          # 1024*2*1023*673*4
    assert n(K.IKSKLVL20, 672) == 11_280_039_936
    assert n(K.IKSKLVL21, 672) == 587_776_000
    assert n(K.BKFFTLVL01, 672) == 66_060_288
    assert n(K.BKLVL02, 672) == 176_160_768
    assert n(99, 672) == 0


def test_malformed_files_rejected(tmp_path):
    p = F.preset(F.PRESET_CFG636)
    key = _member(K.BKLVL01, 636, 5)
    good = str(tmp_path / "good.key")
    K.write_key(good, K.BKLVL01, p, key)
    raw = bytearray(_bytes(good))
    assert np.array_equal(K.read_key(good, K.BKLVL01), key)

    def bad(mutate, name):
        b = bytearray(raw)
        b = mutate(b)
        path = str(tmp_path / name)
        with open(path, "wb") as f:
            f.write(bytes(b))
        with pytest.raises(F.FheError):
            K.read_key(path, K.BKLVL01)

    bad(lambda b: b[:-4], "short.key")                        # truncated payload
    bad(lambda b: b + b"\0", "long.key")                      # trailing bytes
    bad(lambda b: bytes([7]) + b[1:], "flag.key")             # not an endianness flag
    bad(lambda b: b[:137] + b"\0" + b[138:], "null.key")      # null unique_ptr member
    # lvl1 l (offset 1 + 16 + 8) changed from 3 to 4
    bad(lambda b: b[:25] + (4).to_bytes(4, "little") + b[29:], "l1.key")
    with pytest.raises(F.FheError):
        K.read_key(good, K.BKLVL02)                          # wrong kind for the size
    with pytest.raises(F.FheError):
        K.read_key(str(tmp_path / "missing.key"), K.BKLVL01)
    with pytest.raises(F.FheError):
        K.info(good, 42)


def test_evalkey_dir_layout(tmp_path):
    p = F.preset(F.PRESET_CFG636)
    sk = F.SecretKeys(p, 3)
    K.save_evalkey_dir(str(tmp_path), p, sk, {K.BKLVL01: _member(K.BKLVL01, 636, 1)})
    assert sorted(os.listdir(tmp_path)) == ["bklvl01.key", "secret.key"]
    q, s0, _, _ = K.read_secret_key(str(tmp_path / "secret.key"))
    assert q.n0 == 636 and q.alpha0 == 2.0 ** -15 and np.array_equal(s0, sk.s0)


def test_fastr_layout_small_n0(tmp_path):
    """writeToFile (fastR.h:6-38) writes sk, then every innermost array of
    bkfft and isk in order, with no header: the file is the three arrays'
    bytes back to back."""
    n0 = 16
    rng = np.random.default_rng(7)
    s1 = rng.integers(0, 2, 1024, dtype=np.uint32)
    bk = rng.standard_normal((n0, 6, 2, 1024))
    ksk = rng.integers(0, 2 ** 32, (1024, 2, 1023, n0 + 1), dtype=np.uint64).astype(np.uint32)
    path = str(tmp_path / "key.bin")
    K.write_fastr(path, s1, bk, ksk, n0)
    raw = _bytes(path)
    assert len(raw) == 4096 + bk.nbytes + ksk.nbytes
    assert raw[:4096] == s1.tobytes()
    assert raw[4096:4096 + bk.nbytes] == bk.tobytes()
    assert raw[4096 + bk.nbytes:] == ksk.tobytes()
    with pytest.raises(ValueError):
        K.write_fastr(path, s1, bk, ksk, n0 + 1)
