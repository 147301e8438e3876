"""Key material files of the reference, read and written by the native
library (``csrc/keyio.hip``).

* TFHEpp cereal archives as HEDB writes and reads them
  (src/HEDB/utils/serialize.cpp:7-320): ``secret.key`` and one file per
  evaluation-key member (``bklvl01.key``, ``bkfftlvl01.key``,
  ``iksklvl10.key``, ...), by default in ``./../evalkey/`` as there.
* fastR dumps (test/HE3_Query/fastR.h:6-74): lvl1 key, BootstrappingKeyFFT
  and lvl10 key-switching key, raw.

``load_*`` into a Context go straight to device memory (key-switching keys
stream from the file in 64 MiB pinned chunks); ``read_*`` return host numpy
arrays in TFHEpp layout.
"""

import ctypes
import os

import numpy as np

from . import LVL1_N, LVL2_N, Params, _ptr, check, lib

SECRET, BKLVL01, BKLVL02, BKFFTLVL01, BKFFTLVL02, IKSKLVL10, IKSKLVL20, IKSKLVL21 = range(8)

FILE_NAMES = {
    SECRET: "secret.key",
    BKLVL01: "bklvl01.key",
    BKLVL02: "bklvl02.key",
    BKFFTLVL01: "bkfftlvl01.key",
    BKFFTLVL02: "bkfftlvl02.key",
    IKSKLVL10: "iksklvl10.key",
    IKSKLVL20: "iksklvl20.key",
    IKSKLVL21: "iksklvl21.key",
}

# default location used by HEDB::generate_*_key / load_*_key
EVALKEY_DIR = os.path.join("..", "evalkey")


class KeyfileParams(ctypes.Structure):
    """lweParams as stored in every archive (TFHEpp/include/lweParams.hpp:181-190)."""
    _fields_ = [
        ("n0", ctypes.c_uint32), ("alpha0", ctypes.c_double), ("approx_bit0", ctypes.c_uint32),
        ("nbit1", ctypes.c_uint32), ("n1", ctypes.c_uint32), ("l1", ctypes.c_uint32),
        ("bgbit1", ctypes.c_uint32), ("alpha1", ctypes.c_double),
        ("approx_bit1", ctypes.c_uint32),
        ("nbit2", ctypes.c_uint32), ("n2", ctypes.c_uint32), ("l2", ctypes.c_uint32),
        ("bgbit2", ctypes.c_uint32), ("alpha2", ctypes.c_double),
        ("approx_bit2", ctypes.c_uint32),
        ("t10", ctypes.c_uint32), ("basebit10", ctypes.c_uint32), ("alpha10", ctypes.c_double),
        ("t20", ctypes.c_uint32), ("basebit20", ctypes.c_uint32), ("alpha20", ctypes.c_double),
        ("t21", ctypes.c_uint32), ("basebit21", ctypes.c_uint32), ("alpha21", ctypes.c_double),
        ("t22", ctypes.c_uint32), ("basebit22", ctypes.c_uint32), ("alpha22", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _b(path):
    return os.fsencode(path)


def member_shape(kind, n0):
    """numpy dtype and TFHEpp shape of a member file's array."""
    return {
        BKLVL01: (np.uint32, (n0, 6, 2, LVL1_N)),
        BKLVL02: (np.uint64, (n0, 8, 2, LVL2_N)),
        BKFFTLVL01: (np.float64, (n0, 6, 2, LVL1_N)),
        BKFFTLVL02: (np.float64, (n0, 8, 2, LVL2_N)),
        IKSKLVL10: (np.uint32, (LVL1_N, 2, 1023, n0 + 1)),
        IKSKLVL20: (np.uint32, (LVL2_N, 2, 1023, n0 + 1)),
        IKSKLVL21: (np.uint32, (LVL2_N, 10, 7, LVL1_N + 1)),
    }[kind]


def params_of(params):
    q = KeyfileParams()
    check(lib().fhe_keyfile_params_of(ctypes.byref(params), ctypes.byref(q)))
    return q


def info(path, kind):
    """(KeyfileParams, payload bytes) of a key file; raises on a malformed file."""
    q = KeyfileParams()
    n = ctypes.c_uint64()
    check(lib().fhe_keyfile_info(_b(path), kind, ctypes.byref(q), ctypes.byref(n)))
    return q, n.value


def write_secret_key(path, params, sk):
    """HEDB::generate_secret_key's archive (serialize.cpp:7-15) of a SecretKeys."""
    check(lib().fhe_keyfile_write_secret(_b(path), ctypes.byref(params), _ptr(sk.s0),
                                         _ptr(sk.s1), _ptr(sk.s2)))


def read_secret_key(path):
    """HEDB::load_secret_key (serialize.cpp:17-24) -> (KeyfileParams, s0, s1, s2)."""
    q, _ = info(path, SECRET)
    s0 = np.zeros(q.n0, np.uint32)
    s1 = np.zeros(LVL1_N, np.uint32)
    s2 = np.zeros(LVL2_N, np.uint64)
    check(lib().fhe_keyfile_read_secret(_b(path), _ptr(s0), q.n0, _ptr(s1), _ptr(s2), None))
    return q, s0, s1, s2


def write_key(path, kind, params, key):
    """HEDB::generate_<member>_key's archive of a key held in TFHEpp layout."""
    dtype, shape = member_shape(kind, params.n0)
    key = np.ascontiguousarray(key, dtype)
    if key.size != int(np.prod(shape)):
        raise ValueError(f"key of {key.size} elements, file kind {kind} needs {shape}")
    check(lib().fhe_keyfile_write(_b(path), kind, ctypes.byref(params), _ptr(key)))


def read_key(path, kind):
    """HEDB::load_<member>_key into a host array of TFHEpp shape."""
    q, nbytes = info(path, kind)
    dtype, shape = member_shape(kind, q.n0)
    out = np.empty(shape, dtype)
    check(lib().fhe_keyfile_read(_b(path), kind, _ptr(out), nbytes))
    return out


def load_key(ctx, path, kind):
    """HEDB::load_<member>_key straight into a device context."""
    check(lib().fhe_keyfile_load(ctx.h, _b(path), kind))


def save_evalkey_dir(dirpath, params, sk, keys):
    """Writes secret.key and the given members ({kind: array}) under dirpath."""
    os.makedirs(dirpath, exist_ok=True)
    write_secret_key(os.path.join(dirpath, FILE_NAMES[SECRET]), params, sk)
    for kind, arr in keys.items():
        write_key(os.path.join(dirpath, FILE_NAMES[kind]), kind, params, arr)


def load_evalkey_dir(ctx, dirpath=EVALKEY_DIR, kinds=None):
    """Loads every evaluation-key file present under dirpath (or the listed
    kinds) into ctx; an FFT-form bootstrapping key is preferred over the
    integer one of the same level.  Returns the kinds loaded."""
    present = [k for k in FILE_NAMES if k != SECRET
               and os.path.exists(os.path.join(dirpath, FILE_NAMES[k]))]
    if kinds is None:
        kinds = [k for k in present
                 if not (k == BKLVL01 and BKFFTLVL01 in present)
                 and not (k == BKLVL02 and BKFFTLVL02 in present)]
    for k in kinds:
        load_key(ctx, os.path.join(dirpath, FILE_NAMES[k]), k)
    return kinds


def write_fastr(path, s1, bkfft, iksk10, n0):
    """fastR.h writeToFile (fastR.h:6-38)."""
    s1 = np.ascontiguousarray(s1, np.uint32)
    bkfft = np.ascontiguousarray(bkfft, np.float64)
    iksk10 = np.ascontiguousarray(iksk10, np.uint32)
    if s1.size != LVL1_N or bkfft.size != n0 * 6 * 2 * LVL1_N or \
            iksk10.size != LVL1_N * 2 * 1023 * (n0 + 1):
        raise ValueError("fastR arrays do not match n0")
    check(lib().fhe_fastr_write(_b(path), n0, _ptr(s1), _ptr(bkfft), _ptr(iksk10)))


def load_fastr(ctx, path):
    """fastR.h readFromFile (fastR.h:40-72) into ctx; returns the lvl1 secret key."""
    s1 = np.zeros(LVL1_N, np.uint32)
    check(lib().fhe_fastr_load(ctx.h, _b(path), _ptr(s1)))
    return s1


__all__ = ["SECRET", "BKLVL01", "BKLVL02", "BKFFTLVL01", "BKFFTLVL02", "IKSKLVL10",
           "IKSKLVL20", "IKSKLVL21", "FILE_NAMES", "EVALKEY_DIR", "KeyfileParams", "Params",
           "member_shape", "params_of", "info", "write_secret_key", "read_secret_key",
           "write_key", "read_key", "load_key", "save_evalkey_dir", "load_evalkey_dir",
           "write_fastr", "load_fastr"]
