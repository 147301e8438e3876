import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU check (reference cross-check at full size)")


import cufhedb_amd as F  # noqa: E402


class Keys:
    """Seeded key material for one parameter preset (cached per session)."""

    def __init__(self, which, seed):
        self.params = F.preset(which)
        self.n0 = self.params.n0
        self.seed = seed
        self.sk = F.SecretKeys(self.params, seed)
        self.bk = F.keygen_bk_lvl01(self.sk, seed + 1)
        self._bkfft = None
        self._ksk10 = None

    @property
    def bkfft(self):
        """FD key in TFHEpp BootstrappingKeyFFT layout (oracle ApplyFFT2trgsw)."""
        if self._bkfft is None:
            import _oracle as O
            self._bkfft = O.bk_to_fft32(self.bk)
        return self._bkfft

    @property
    def ksk10(self):
        if self._ksk10 is None:
            self._ksk10 = F.keygen_ksk(self.sk, 10, self.seed + 2)
        return self._ksk10

    def enc_lvl0(self, m, index0=0, seed=None):
        return F.tlwe_encrypt_u32(self.seed + 100 if seed is None else seed, index0, self.sk.s0,
                                  m, self.params.alpha0)

    def enc_lvl1(self, m, index0=0, seed=None):
        return F.tlwe_encrypt_u32(self.seed + 200 if seed is None else seed, index0, self.sk.s1,
                                  m, self.params.alpha1)


_KEYS = {}


def keys_for(which):
    if which not in _KEYS:
        _KEYS[which] = Keys(which, {F.PRESET_REF672: 672001, F.PRESET_CFG636: 636001}[which])
    return _KEYS[which]


@pytest.fixture(scope="session")
def keys672():
    return keys_for(F.PRESET_REF672)


@pytest.fixture(scope="session")
def keys636():
    return keys_for(F.PRESET_CFG636)


_CTX = {}


def ctx_for(which):
    if which not in _CTX:
        k = keys_for(which)
        ctx = F.Context(k.params)
        ctx.load_bk(1, k.bk, F.BK_INT)
        _CTX[which] = ctx
    return _CTX[which]


@pytest.fixture(scope="session")
def ctx672():
    return ctx_for(F.PRESET_REF672)


@pytest.fixture(scope="session")
def ctx636():
    return ctx_for(F.PRESET_CFG636)


_FULL = {}


def full_ctx_for(which):
    """Context holding every key the operator layer uses: BK lvl01 and lvl02,
    KSK 10, 20 and 21 (20 and 21 generated on the device, bit-identical to the
    host generator's keys for the same seeds)."""
    if which not in _FULL:
        k = keys_for(which)
        ctx = F.Context(k.params)
        ctx.load_bk(1, k.bk, F.BK_INT)
        ctx.load_bk(2, F.keygen_bk_lvl02(k.sk, k.seed + 3), F.BK_INT)
        ctx.load_ksk(10, k.ksk10)
        for kind, off in ((20, 4), (21, 5)):
            ctx.keygen_ksk(kind, k.sk, k.seed + off)
        _FULL[which] = ctx
    return keys_for(which), _FULL[which]


@pytest.fixture(scope="session")
def full636():
    return full_ctx_for(F.PRESET_CFG636)


@pytest.fixture(scope="session")
def full672():
    """The reference's own parameter set (n0 = 672, alpha0 = 2^-16).  The
    lvl2 comparison chains run here: with n0 = 636 / alpha0 = 2^-15 the
    key-switched lvl0 noise (IKS lvl20) sits about 3 sigma from the 6-bit
    windows' decision boundary, a per-row failure rate near 1e-4."""
    return full_ctx_for(F.PRESET_REF672)


@pytest.fixture
def rng():
    return np.random.default_rng(20261202)
