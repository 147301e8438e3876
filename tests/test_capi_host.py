"""The C ABI library on the CPU: it loads, exports every function declared in
include/cufhedb_amd.h, validates arguments, and its seeded key generation is
deterministic and decrypts correctly.  No device compute is issued here.
"""

import ctypes
import os
import re

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                   "cufhedb_amd.h")


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fhe_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 40
    L = ctypes.CDLL(F.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(F._SIGS)


def test_presets_match_reference_parameters():
    r = F.preset(F.PRESET_REF672)
    # thirdparty/TFHEpp/include/params/128bit.hpp:8-10,26,44; src/key.cpp:11-13
    assert (r.n0, r.key0_hw, r.key1_hw, r.key2_hw) == (672, 512, 512, 1024)
    assert (r.alpha0, r.alpha1, r.alpha2) == (2.0 ** -16, 2.0 ** -25, 2.0 ** -52)
    c = F.preset(F.PRESET_CFG636)
    assert (c.n0, c.key0_hw, c.alpha0) == (636, 0, 2.0 ** -15)


def test_argument_errors_are_reported():
    p = F.Params()
    rc = F.lib().fhe_params_preset(7, ctypes.byref(p))
    assert rc == F.FHE_ERR_ARG and b"unknown preset" in F.lib().fhe_last_error()
    good = F.preset(F.PRESET_REF672)
    assert F.lib().fhe_keygen_ksk(ctypes.byref(good), 11, 1, None, None, None) == F.FHE_ERR_ARG
    bad = F.preset(F.PRESET_REF672)
    bad.n0 = 5000
    s0 = np.zeros(8, np.uint32)
    assert F.lib().fhe_keygen_secret(ctypes.byref(bad), 1, F._ptr(s0), None, None) == F.FHE_ERR_ARG
    with pytest.raises(F.FheError):
        F.check(F.lib().fhe_params_preset(-1, ctypes.byref(p)))


def test_ksk_sizes_match_reference_layout():
    p = F.preset(F.PRESET_REF672)
    # KeySwitchingKey<P> = [dn][t][2^basebit-1][tn+1] (params.hpp:79-83)
    assert F.lib().fhe_ksk_words(ctypes.byref(p), 10) * 4 == 5_640_019_968
    assert F.lib().fhe_ksk_words(ctypes.byref(p), 20) * 4 == 11_280_039_936
    assert F.lib().fhe_ksk_words(ctypes.byref(p), 21) * 4 == 587_776_000


def test_context_without_device_fails_cleanly():
    p = F.preset(F.PRESET_CFG636)
    h = ctypes.c_void_p()
    rc = F.lib().fhe_ctx_create(ctypes.byref(p), 0, ctypes.byref(h))
    if rc == F.FHE_OK:
        F.lib().fhe_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == F.FHE_ERR_DEVICE and not h.value
    assert F.lib().fhe_last_error()


def test_keygen_is_deterministic():
    p = F.preset(F.PRESET_CFG636)
    a, b, c = F.SecretKeys(p, 1), F.SecretKeys(p, 1), F.SecretKeys(p, 2)
    assert np.array_equal(a.s0, b.s0) and np.array_equal(a.s2, b.s2)
    assert not np.array_equal(a.s1, c.s1)
    assert np.array_equal(F.keygen_bk_lvl01(a, 5), F.keygen_bk_lvl01(b, 5))
    # uniform binary key for the n0 = 636 set, Hamming-weight keys for 672
    assert set(np.unique(a.s0)) <= {0, 1} and 250 < a.s0.sum() < 390
    r = F.SecretKeys(F.preset(F.PRESET_REF672), 1)
    assert r.s0.sum() <= 512 and r.s2.sum() <= 1024


def test_bootstrapping_key_rows_decrypt(keys672):
    k = keys672
    h = [1 << (32 - 6 * (d + 1)) for d in range(3)]
    for i in (0, 1, 300, k.n0 - 1):
        for d in range(3):
            # row (d, k=1): phase[0] = s0[i] h_d + e
            ph_b = O.trlwe_phase32(k.bk[i, d + 3], k.sk.s1)
            want = np.zeros(1024, np.int64)
            want[0] = int(k.sk.s0[i]) * h[d]
            assert np.max(np.abs(O.tdiff32(ph_b, want.astype(np.uint32)))) < 1 << 12
            # row (d, k=0): phase = e - s0[i] h_d s1(X)
            ph_a = O.trlwe_phase32(k.bk[i, d], k.sk.s1)
            want = (-(int(k.sk.s0[i]) * h[d]) * k.sk.s1.astype(np.int64)) % (1 << 32)
            assert np.max(np.abs(O.tdiff32(ph_a, want.astype(np.uint32)))) < 1 << 12


def test_tlwe_encryption_decrypts(keys636, rng):
    m = rng.integers(0, 1 << 32, 16, dtype=np.uint64).astype(np.uint32)
    c = keys636.enc_lvl0(m)
    e = O.tdiff32(F.tlwe_phase_u32(c, keys636.sk.s0), m)
    assert np.max(np.abs(e)) < 6 * 2.0 ** -15 * 2 ** 32
    assert np.array_equal(keys636.enc_lvl0(m), c)  # deterministic per (seed, index)


def test_scheduler_and_keyfile_argument_checks():
    """Argument errors of the scheduler and key-file entry points are
    reported before any device work (no GPU needed)."""
    L = F.lib()
    out = ctypes.c_void_p()
    assert L.fhe_sched_create(None, 16, ctypes.byref(out)) == F.FHE_ERR_ARG
    assert L.fhe_sched_flush(None) == F.FHE_ERR_ARG
    assert L.fhe_sched_wait(None, 1, None) == F.FHE_ERR_ARG
    t = ctypes.c_uint64()
    assert L.fhe_sched_submit_cmp2(None, 9, None, None, None, 20, 1, 4, ctypes.byref(t)) == \
        F.FHE_ERR_ARG
    assert L.fhe_keyfile_load(None, b"x", 1) == F.FHE_ERR_ARG
    assert L.fhe_fastr_load(None, b"x", None) == F.FHE_ERR_ARG
    assert L.fhe_keygen_ksk_dev(None, 10, 1, None, None, None) == F.FHE_ERR_ARG
