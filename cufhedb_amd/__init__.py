"""Python host interface to the MI355X TFHE bootstrapping library.

This package is a thin ctypes layer over ``libcufhedb_amd.so`` (HIP kernels
for gfx950 behind the extern "C" ABI declared in ``include/cufhedb_amd.h``).
It mirrors the reference's operator surface for the bootstrapping hot path:

* ``Context.load_bk`` / ``Context.load_ksk``   -- cufft::CuFFT_Torus::LoadBK
  (src/cuHEDB/fft_gpu/cufft_gpu.cuh:175-184), keys resident on the device;
* ``Context.pbs_lvl01``                         -- GateBootstrappingTLWE2TLWEFFT
  (cufft_gpu.cuh:250-271), batched;
* ``Context.iks``                               -- TFHEpp IdentityKeySwitch
  (thirdparty/TFHEpp/src/keyswitch.cpp:12-48), batched on the GPU;
* ``Context.gate`` / ``msb_gate`` / ``ide_gate``  -- cuHEDB:: operators
  (src/cuHEDB/gatebootstrap_gpu.cuh:31-170).

Arrays are numpy arrays in TFHEpp's aggregate layouts (a TLWE is n+1 torus
words, b last).  The library must be built (``python -c "import
__graft_entry__ as g; g.build()"`` or ``make``); importing this package
without it raises, so no computation can silently fall back to the CPU.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcufhedb_amd.so")

FHE_OK = 0
FHE_ERR_ARG = 1
FHE_ERR_NOT_LOADED = 2
FHE_ERR_DEVICE = 3
FHE_ERR_ALLOC = 4
FHE_ERR_UNSUPPORTED = 5

PRESET_REF672 = 0
PRESET_CFG636 = 1
BK_INT = 0
BK_FFT = 1
PBS_FP64 = 0
PBS_MFMA_SPLIT = 1

TUNE_IKS21_TILE = 0
TUNE_PBS1_WAVES = 1
TUNE_IKS_SPLIT = 2
TUNE_IKS20_SPLIT = 3

GATE_NAND, GATE_AND, GATE_OR, GATE_NOR, GATE_XNOR = range(5)
LOGIC = 1
ARITHMETIC = 0

LVL1_N = 1024
LVL2_N = 2048
LVL1_MU = 1 << 29


class FheError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fhe status {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("n0", ctypes.c_uint32),
        ("key0_hw", ctypes.c_uint32),
        ("key1_hw", ctypes.c_uint32),
        ("key2_hw", ctypes.c_uint32),
        ("alpha0", ctypes.c_double),
        ("alpha1", ctypes.c_double),
        ("alpha2", ctypes.c_double),
    ]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "fhe_params_preset": (_i, [_i, _vp]),
    "fhe_last_error": (ctypes.c_char_p, []),
    "fhe_keygen_secret": (_i, [_vp, _u64, _vp, _vp, _vp]),
    "fhe_keygen_bk_lvl01": (_i, [_vp, _u64, _vp, _vp, _vp]),
    "fhe_keygen_bk_lvl02": (_i, [_vp, _u64, _vp, _vp, _vp]),
    "fhe_keygen_ksk": (_i, [_vp, _i, _u64, _vp, _vp, _vp]),
    "fhe_ksk_words": (_sz, [_vp, _i]),
    "fhe_tlwe_encrypt_u32": (_i, [_u64, _u64, _vp, _u32, _vp, _d, _vp, _i]),
    "fhe_tlwe_encrypt_u64": (_i, [_u64, _u64, _vp, _u32, _vp, _d, _vp, _i]),
    "fhe_trlwe_encrypt_u32": (_i, [_u64, _u64, _vp, _vp, _d, _vp, _i]),
    "fhe_trlwe_encrypt_u64": (_i, [_u64, _u64, _vp, _vp, _d, _vp, _i]),
    "fhe_trgsw_encrypt_u32": (_i, [_u64, _u64, _vp, ctypes.c_int32, _d, _vp]),
    "fhe_trgsw_encrypt_u64": (_i, [_u64, _u64, _vp, ctypes.c_int64, _d, _vp]),
    "fhe_ctx_create": (_i, [_vp, _i, _vp]),
    "fhe_ctx_destroy": (_i, [_vp]),
    "fhe_ctx_sync": (_i, [_vp]),
    "fhe_ctx_stream": (_vp, [_vp]),
    "fhe_load_bk": (_i, [_vp, _i, _vp, _i]),
    "fhe_load_ksk": (_i, [_vp, _i, _vp]),
    "fhe_bk_loaded": (_i, [_vp, _i]),
    "fhe_ksk_loaded": (_i, [_vp, _i]),
    "fhe_keygen_ksk_dev": (_i, [_vp, _i, _u64, _vp, _vp, _vp]),
    "fhe_set_pbs_variant": (_i, [_vp, _i]),
    "fhe_pbs_lvl01_dev": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp]),
    "fhe_blind_rotate_lvl01_dev": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp]),
    "fhe_pbs_lvl02_dev": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp]),
    "fhe_pbs_many_lut_lvl01_dev": (_i, [_vp, _vp, _vp, _i, _i, _vp, _i, _vp]),
    "fhe_pbs_many_lut_lvl02_dev": (_i, [_vp, _vp, _vp, _i, _i, _vp, _i, _vp]),
    "fhe_blind_rotate_lvl02_dev": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp]),
    "fhe_sample_extract_lvl1_dev": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "fhe_sample_extract_lvl2_dev": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "fhe_iks_dev": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "fhe_external_product_lvl1_dev": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fhe_external_product_lvl2_dev": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fhe_ifft_lvl1_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_fft_lvl1_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_ifft_lvl2_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_fft_lvl2_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_measure_peaks": (_i, [_vp, _vp, _vp]),
    "fhe_gate_lvl1_dev": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp]),
    "fhe_gate_lvl1_noalloc_dev": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp]),
    "fhe_gate_rescale_lvl1_dev": (_i, [_vp, _i, _vp, _vp, _vp, _u32, _i, _vp]),
    "fhe_msb_gate_lvl1_dev": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "fhe_ide_gate_lvl1_dev": (_i, [_vp, _vp, _vp, _u32, _i, _vp]),
    "fhe_msb_gate_dev": (_i, [_vp, _i, _i, _vp, _vp, _i, _i, _vp]),
    "fhe_ide_gate_dev": (_i, [_vp, _i, _i, _vp, _vp, _u32, _i, _vp]),
    "fhe_log_to_ari_lvl1_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_log_rescale_lvl1_dev": (_i, [_vp, _vp, _vp, _u32, _i, _vp]),
    "fhe_hom_msb_dev": (_i, [_vp, _i, _vp, _vp, _u32, _i, _i, _vp]),
    "fhe_fwd_digits_lvl1_mfma_dev": (_i, [_vp, _vp, _vp, _i, _vp]),
    "fhe_compare_dev": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _u32, _i, _i, _vp]),
    "fhe_hom_not_dev": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "fhe_compare_digits_dev": (_i, [_vp, _i, _vp, _i, _i, _u64, _i, _vp, _i, _i, _vp]),
    "fhe_autotune": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "fhe_tune_set": (_i, [_vp, _i, _vp, _vp, _i]),
    "fhe_tune_get": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "fhe_graph_begin": (_i, [_vp]),
    "fhe_graph_end": (_i, [_vp, _vp]),
    "fhe_graph_launch": (_i, [_vp, _vp]),
    "fhe_graph_destroy": (_i, [_vp]),
    "fhe_digit_testvectors": (_i, [_i, _i, _i, _u64, _i, _vp, _vp, _vp]),
    "fhe_keyfile_params_of": (_i, [_vp, _vp]),
    "fhe_keyfile_payload_bytes": (_u64, [_i, _u32]),
    "fhe_keyfile_info": (_i, [ctypes.c_char_p, _i, _vp, _vp]),
    "fhe_keyfile_read": (_i, [ctypes.c_char_p, _i, _vp, _u64]),
    "fhe_keyfile_read_secret": (_i, [ctypes.c_char_p, _vp, _u32, _vp, _vp, _vp]),
    "fhe_keyfile_write": (_i, [ctypes.c_char_p, _i, _vp, _vp]),
    "fhe_keyfile_write_secret": (_i, [ctypes.c_char_p, _vp, _vp, _vp, _vp]),
    "fhe_keyfile_load": (_i, [_vp, ctypes.c_char_p, _i]),
    "fhe_fastr_write": (_i, [ctypes.c_char_p, _u32, _vp, _vp, _vp]),
    "fhe_fastr_load": (_i, [_vp, ctypes.c_char_p, _vp]),
    "fhe_sched_create": (_i, [_vp, _i, _vp]),
    "fhe_sched_destroy": (_i, [_vp]),
    "fhe_sched_submit_pbs1": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fhe_sched_submit_cmp2": (_i, [_vp, _i, _vp, _vp, _vp, _u32, _i, _i, _vp]),
    "fhe_sched_flush": (_i, [_vp]),
    "fhe_sched_wait": (_i, [_vp, _u64, _vp]),
    "fhe_pbs_lvl01": (_i, [_vp, _vp, _vp, _i, _vp, _i]),
    "fhe_iks": (_i, [_vp, _i, _vp, _vp, _i]),
    "fhe_malloc": (_i, [_vp, _sz]),
    "fhe_free": (_i, [_vp]),
    "fhe_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "fhe_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "fhe_memcpy_d2d": (_i, [_vp, _vp, _sz]),
    "fhe_memset_dev": (_i, [_vp, _i, _sz]),
    "fhe_device_sync": (_i, []),
    "fhe_timer_start": (_i, [_vp, _vp]),
    "fhe_timer_stop": (_i, [_vp, _vp, _vp]),
    "fhe_device_name": (_i, [_vp, _i]),
}

_lib = None


def lib():
    """Load libcufhedb_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -j8` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    """Pointer argument that keeps `a` alive for the duration of the call."""
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def check(rc):
    if rc != FHE_OK:
        raise FheError(rc, lib().fhe_last_error().decode(errors="replace"))


def preset(which):
    p = Params()
    check(lib().fhe_params_preset(which, ctypes.byref(p)))
    return p


# ---------------------------------------------------------------- key material
class SecretKeys:
    """lvl0 / lvl1 / lvl2 secret keys (TFHEpp lweKey, key.cpp:8-17), seeded."""

    def __init__(self, params, seed):
        self.params = params
        self.seed = seed
        self.s0 = np.zeros(params.n0, np.uint32)
        self.s1 = np.zeros(LVL1_N, np.uint32)
        self.s2 = np.zeros(LVL2_N, np.uint64)
        check(lib().fhe_keygen_secret(ctypes.byref(params), seed, _ptr(self.s0), _ptr(self.s1),
                                      _ptr(self.s2)))


def keygen_bk_lvl01(sk, seed):
    """Integer BK [n0][6][2][1024] uint32 (TFHEpp BootstrappingKey<lvl01param>)."""
    bk = np.zeros((sk.params.n0, 6, 2, LVL1_N), np.uint32)
    check(lib().fhe_keygen_bk_lvl01(ctypes.byref(sk.params), seed, _ptr(sk.s0), _ptr(sk.s1),
                                    _ptr(bk)))
    return bk


def keygen_bk_lvl02(sk, seed):
    bk = np.zeros((sk.params.n0, 8, 2, LVL2_N), np.uint64)
    check(lib().fhe_keygen_bk_lvl02(ctypes.byref(sk.params), seed, _ptr(sk.s0), _ptr(sk.s2),
                                    _ptr(bk)))
    return bk


def ksk_shape(params, kind):
    if kind == 10:
        return (LVL1_N, 2, 1023, params.n0 + 1)
    if kind == 20:
        return (LVL2_N, 2, 1023, params.n0 + 1)
    if kind == 21:
        return (LVL2_N, 10, 7, LVL1_N + 1)
    raise ValueError(kind)


def keygen_ksk(sk, kind, seed):
    """KeySwitchingKey [dn][t][2^basebit-1][tn+1] uint32 (TFHEpp ikskgen)."""
    ksk = np.empty(ksk_shape(sk.params, kind), np.uint32)
    sdom = sk.s1 if kind == 10 else sk.s2
    stgt = sk.s1 if kind == 21 else sk.s0
    check(lib().fhe_keygen_ksk(ctypes.byref(sk.params), kind, seed, _ptr(sdom), _ptr(stgt),
                               _ptr(ksk)))
    return ksk


def tlwe_encrypt_u32(seed, index0, key, m, alpha):
    m = np.ascontiguousarray(m, np.uint32)
    n = key.shape[0]
    out = np.zeros((m.shape[0], n + 1), np.uint32)
    check(lib().fhe_tlwe_encrypt_u32(seed, index0, _ptr(key), n, _ptr(m), alpha, _ptr(out),
                                     m.shape[0]))
    return out


def tlwe_encrypt_u64(seed, index0, key, m, alpha):
    m = np.ascontiguousarray(m, np.uint64)
    n = key.shape[0]
    out = np.zeros((m.shape[0], n + 1), np.uint64)
    check(lib().fhe_tlwe_encrypt_u64(seed, index0, _ptr(key), n, _ptr(m), alpha, _ptr(out),
                                     m.shape[0]))
    return out


def trlwe_encrypt_u32(seed, index0, key, m, alpha):
    m = np.ascontiguousarray(m, np.uint32).reshape(-1, LVL1_N)
    out = np.zeros((m.shape[0], 2, LVL1_N), np.uint32)
    check(lib().fhe_trlwe_encrypt_u32(seed, index0, _ptr(key), _ptr(m), alpha, _ptr(out),
                                      m.shape[0]))
    return out


def trgsw_encrypt_u32(seed, index, key, m, alpha):
    out = np.zeros((6, 2, LVL1_N), np.uint32)
    check(lib().fhe_trgsw_encrypt_u32(seed, index, _ptr(key), int(m), alpha, _ptr(out)))
    return out


def tlwe_phase_u32(c, key):
    """b - <a, s> mod 2^32 for a batch of TLWEs (numpy, exact)."""
    c = np.atleast_2d(c).astype(np.uint64)
    n = key.shape[0]
    dot = (c[:, :n] * key.astype(np.uint64)[None, :]).sum(axis=1, dtype=np.uint64)
    return ((c[:, n] - dot) & 0xFFFFFFFF).astype(np.uint32)


def torus32_signed(x):
    return np.asarray(x, np.uint32).view(np.int32).astype(np.int64)


# ---------------------------------------------------------------- device side
class Scheduler:
    """Coalescing scheduler over a context (fhe_sched_*): requests are
    queued by submit_*, served by flush() in merged launches of up to
    max_rows rows, and waited for by ticket."""

    PBS1, CMP2 = 0, 1

    def __init__(self, ctx, max_rows=16384):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().fhe_sched_create(ctx.h, int(max_rows), ctypes.byref(self.h)))

    def submit_pbs1(self, in_buf, tv_buf, out_buf, B):
        t = ctypes.c_uint64()
        check(lib().fhe_sched_submit_pbs1(self.h, in_buf.ptr, tv_buf.ptr, out_buf.ptr, int(B),
                                          ctypes.byref(t)))
        return t.value

    def submit_cmp2(self, op, a_buf, b_buf, out_buf, bits, result_type, B):
        t = ctypes.c_uint64()
        check(lib().fhe_sched_submit_cmp2(self.h, int(op), a_buf.ptr, b_buf.ptr, out_buf.ptr,
                                          int(bits), int(result_type), int(B), ctypes.byref(t)))
        return t.value

    def flush(self):
        check(lib().fhe_sched_flush(self.h))

    def wait(self, ticket):
        """Milliseconds from the start of the serving flush to completion."""
        ms = ctypes.c_float()
        check(lib().fhe_sched_wait(self.h, ctypes.c_uint64(ticket), ctypes.byref(ms)))
        return float(ms.value)

    def close(self):
        if self.h:
            lib().fhe_sched_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device allocation owned by Python (fhe_malloc / fhe_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().fhe_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        check(lib().fhe_memcpy_h2d(buf.ptr, _ptr(a), a.nbytes))
        return buf

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().fhe_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().fhe_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU, resident keys (fhe_ctx)."""

    def __init__(self, params, device=0):
        self.params = params
        h = ctypes.c_void_p()
        check(lib().fhe_ctx_create(ctypes.byref(params), device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().fhe_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().fhe_ctx_stream(self.h)

    def sync(self):
        check(lib().fhe_ctx_sync(self.h))

    def load_bk(self, level, bk, fmt=BK_INT):
        bk = np.ascontiguousarray(bk)
        check(lib().fhe_load_bk(self.h, level, _ptr(bk), fmt))

    def load_ksk(self, kind, ksk):
        ksk = np.ascontiguousarray(ksk, np.uint32)
        check(lib().fhe_load_ksk(self.h, kind, _ptr(ksk)))

    def keygen_ksk(self, kind, sk, seed, out_dev=None):
        """Key-switching key generated on the device, bit-identical to
        keygen_ksk(sk, kind, seed); into the context's key slot, or into the
        DeviceBuffer out_dev."""
        sdom = sk.s1 if kind == 10 else sk.s2
        stgt = sk.s1 if kind == 21 else sk.s0
        check(lib().fhe_keygen_ksk_dev(self.h, kind, seed, _ptr(sdom), _ptr(stgt),
                                       out_dev.ptr if out_dev is not None else None))

    # ---- autotuning (fhe_autotune / fhe_tune_get / fhe_tune_set)
    TUNE_CANDIDATES = {TUNE_IKS21_TILE: (8, 16, 32, 64), TUNE_PBS1_WAVES: (1, 2),
                       TUNE_IKS_SPLIT: (1, 2, 4, 8), TUNE_IKS20_SPLIT: (1, 2, 4, 8)}

    def autotune(self, batches=(64, 256, 1024, 4096, 16384), knob=None):
        """Time every candidate of `knob` (TUNE_IKS21_TILE: ciphertexts per
        workgroup of the lvl2 -> lvl1 key switch, needs KSK 21;
        TUNE_PBS1_WAVES: waves per ciphertext of the lvl1 blind rotation,
        needs BK lvl01; TUNE_IKS_SPLIT / TUNE_IKS20_SPLIT: workgroups per
        ciphertext of the lvl1 -> lvl0 / lvl2 -> lvl0 key switch, needs KSK
        10 / 20) at each batch size and keep the fastest; returns
        {batch: (choice, {candidate: ms})}."""
        knob = TUNE_IKS21_TILE if knob is None else knob
        cand = self.TUNE_CANDIDATES[knob]
        b = np.ascontiguousarray(sorted(batches), np.int32)
        choice = np.zeros(b.size, np.int32)
        ms = np.zeros((b.size, len(cand)), np.float32)
        check(lib().fhe_autotune(self.h, knob, _ptr(b), int(b.size), _ptr(choice), _ptr(ms)))
        return {int(x): (int(c), {t: float(v) for t, v in zip(cand, row)})
                for x, c, row in zip(b, choice, ms)}

    def tune_table(self, knob=None):
        """[(batch, choice), ...] of `knob` ([] = untuned: built-in rule)."""
        knob = TUNE_IKS21_TILE if knob is None else knob
        b = np.zeros(16, np.int32)
        t = np.zeros(16, np.int32)
        n = ctypes.c_int()
        check(lib().fhe_tune_get(self.h, knob, _ptr(b), _ptr(t), 16, ctypes.byref(n)))
        return [(int(b[i]), int(t[i])) for i in range(n.value)]

    def set_tune_table(self, table, knob=None):
        knob = TUNE_IKS21_TILE if knob is None else knob
        b = np.ascontiguousarray([x for x, _ in table], np.int32)
        t = np.ascontiguousarray([y for _, y in table], np.int32)
        check(lib().fhe_tune_set(self.h, knob, _ptr(b), _ptr(t), int(b.size)))

    def set_pbs_variant(self, variant):
        check(lib().fhe_set_pbs_variant(self.h, variant))

    # ---- per-device tuning cache (a JSON file of fhe_tune_get tables)
    def _tune_key(self):
        p = self.params
        return f"{device_name()}|n0={p.n0}"

    def save_tuning(self, path):
        """Merge this context's tuned tables into the JSON cache at `path`,
        under the device's name and the parameter set (n0): a later process
        on the same kind of device restores them with load_tuning instead of
        timing the candidates again."""
        import json
        import os
        cache = {}
        if os.path.exists(path):
            with open(path) as f:
                cache = json.load(f)
        cache[self._tune_key()] = {str(k): self.tune_table(knob=k) for k in self.TUNE_CANDIDATES}
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)

    def load_tuning(self, path):
        """Restore the tables saved for this device and parameter set; returns
        False (tables unchanged) when the cache has none."""
        import json
        import os
        if not os.path.exists(path):
            return False
        with open(path) as f:
            entry = json.load(f).get(self._tune_key())
        if entry is None:
            return False
        for k, table in entry.items():
            if int(k) in self.TUNE_CANDIDATES:
                self.set_tune_table([(int(b), int(v)) for b, v in table], knob=int(k))
        return True

    # ---- host-array convenience forms (synchronous)
    def pbs_lvl01(self, tlwe0, tv):
        tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
        tv = np.ascontiguousarray(tv, np.uint32)
        B = tlwe0.shape[0]
        stride = 0 if tv.ndim == 1 else LVL1_N
        out = np.zeros((B, LVL1_N + 1), np.uint32)
        check(lib().fhe_pbs_lvl01(self.h, _ptr(tlwe0), _ptr(tv), stride, _ptr(out), B))
        return out

    def iks(self, kind, tlwe):
        tlwe = np.ascontiguousarray(tlwe)
        B = tlwe.shape[0]
        tn = LVL1_N if kind == 21 else self.params.n0
        out = np.zeros((B, tn + 1), np.uint32)
        check(lib().fhe_iks(self.h, kind, _ptr(tlwe), _ptr(out), B))
        return out

    def _run_dev(self, fn, inputs, out_shape, out_dtype, *extra, stride_first=None):
        """Copy inputs to the device, run fn(ctx, *dev_in, [stride,] dev_out, *extra,
        stream=None), copy the output back."""
        dins = [DeviceBuffer.from_array(a) for a in inputs]
        dout = DeviceBuffer(int(np.prod(out_shape)) * np.dtype(out_dtype).itemsize)
        mid = [] if stride_first is None else [stride_first]
        check(fn(self.h, *[d.ptr for d in dins], *mid, dout.ptr, *extra, None))
        self.sync()
        out = dout.to_array(out_dtype, out_shape)
        for d in dins:
            d.free()
        dout.free()
        return out

    def blind_rotate_lvl01(self, tlwe0, tv):
        tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
        tv = np.ascontiguousarray(tv, np.uint32)
        B = tlwe0.shape[0]
        stride = 0 if tv.ndim == 1 else LVL1_N
        dins = [DeviceBuffer.from_array(tlwe0), DeviceBuffer.from_array(tv)]
        dout = DeviceBuffer(B * 2 * LVL1_N * 4)
        check(lib().fhe_blind_rotate_lvl01_dev(self.h, dins[0].ptr, dins[1].ptr, stride, dout.ptr,
                                               B, None))
        self.sync()
        return dout.to_array(np.uint32, (B, 2, LVL1_N))

    def sample_extract_lvl1(self, trlwe, index=0):
        trlwe = np.ascontiguousarray(trlwe, np.uint32)
        B = trlwe.shape[0]
        return self._run_dev(lib().fhe_sample_extract_lvl1_dev, [trlwe], (B, LVL1_N + 1),
                             np.uint32, index, B)

    def external_product_lvl1(self, trlwe, trgswfft):
        trlwe = np.ascontiguousarray(trlwe, np.uint32)
        B = trlwe.shape[0]
        return self._run_dev(lib().fhe_external_product_lvl1_dev,
                             [trlwe, np.ascontiguousarray(trgswfft, np.float64)],
                             (B, 2, LVL1_N), np.uint32, B)

    def fwd_digits_lvl1_mfma(self, poly):
        """Split-MFMA forward transform of small-integer polynomials (B x 1024
        int32, |x| <= 2048) -> B x 1024 FD in TFHEpp order."""
        poly = np.ascontiguousarray(poly, np.int32)
        B = poly.shape[0]
        return self._run_dev(lib().fhe_fwd_digits_lvl1_mfma_dev, [poly], (B, LVL1_N), np.float64,
                             B)

    def pbs_lvl02(self, tlwe0, tv):
        tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
        tv = np.ascontiguousarray(tv, np.uint64)
        B = tlwe0.shape[0]
        stride = 0 if tv.ndim == 1 else LVL2_N
        return self._run_dev(lib().fhe_pbs_lvl02_dev, [tlwe0, tv], (B, LVL2_N + 1), np.uint64,
                             B, stride_first=stride)

    def pbs_many_lut(self, level, tlwe0, tv, num_out):
        """GateBootstrappingManyLUT: num_out lookup tables from one blind
        rotation (fhe_pbs_many_lut_lvl0X_dev).  tv: one test vector (N words)
        or one per ciphertext (B x N), tables interleaved as
        many_lut_testvector builds them.  Returns B x num_out x (N+1)."""
        N, dt, fn = ((LVL1_N, np.uint32, lib().fhe_pbs_many_lut_lvl01_dev) if level == 1 else
                     (LVL2_N, np.uint64, lib().fhe_pbs_many_lut_lvl02_dev))
        tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
        tv = np.ascontiguousarray(tv, dt)
        B = tlwe0.shape[0]
        stride = 0 if tv.ndim == 1 else N
        dins = [DeviceBuffer.from_array(tlwe0), DeviceBuffer.from_array(tv)]
        dout = DeviceBuffer(max(B, 1) * num_out * (N + 1) * np.dtype(dt).itemsize)
        try:
            check(fn(self.h, dins[0].ptr, dins[1].ptr, stride, num_out, dout.ptr, B, None))
            self.sync()
            return dout.to_array(dt, (B, num_out, N + 1))
        finally:
            for d in dins + [dout]:
                d.free()

    def blind_rotate_lvl02(self, tlwe0, tv):
        tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
        tv = np.ascontiguousarray(tv, np.uint64)
        B = tlwe0.shape[0]
        stride = 0 if tv.ndim == 1 else LVL2_N
        return self._run_dev(lib().fhe_blind_rotate_lvl02_dev, [tlwe0, tv], (B, 2, LVL2_N),
                             np.uint64, B, stride_first=stride)

    def sample_extract_lvl2(self, trlwe, index=0):
        trlwe = np.ascontiguousarray(trlwe, np.uint64)
        B = trlwe.shape[0]
        return self._run_dev(lib().fhe_sample_extract_lvl2_dev, [trlwe], (B, LVL2_N + 1),
                             np.uint64, index, B)

    def external_product_lvl2(self, trlwe, trgswfft):
        trlwe = np.ascontiguousarray(trlwe, np.uint64)
        B = trlwe.shape[0]
        return self._run_dev(lib().fhe_external_product_lvl2_dev,
                             [trlwe, np.ascontiguousarray(trgswfft, np.float64)],
                             (B, 2, LVL2_N), np.uint64, B)

    def ifft_lvl2(self, poly):
        poly = np.ascontiguousarray(poly, np.uint64).reshape(-1, LVL2_N)
        B = poly.shape[0]
        return self._run_dev(lib().fhe_ifft_lvl2_dev, [poly], (B, LVL2_N), np.float64, B)

    def fft_lvl2(self, fd):
        fd = np.ascontiguousarray(fd, np.float64).reshape(-1, LVL2_N)
        B = fd.shape[0]
        return self._run_dev(lib().fhe_fft_lvl2_dev, [fd], (B, LVL2_N), np.uint64, B)

    def ifft_lvl1(self, poly):
        poly = np.ascontiguousarray(poly, np.uint32).reshape(-1, LVL1_N)
        B = poly.shape[0]
        return self._run_dev(lib().fhe_ifft_lvl1_dev, [poly], (B, LVL1_N), np.float64, B)

    def fft_lvl1(self, fd):
        fd = np.ascontiguousarray(fd, np.float64).reshape(-1, LVL1_N)
        B = fd.shape[0]
        return self._run_dev(lib().fhe_fft_lvl1_dev, [fd], (B, LVL1_N), np.uint32, B)

    def gate(self, gate, a, b, result_type=LOGIC):
        a = np.ascontiguousarray(a, np.uint32)
        b = np.ascontiguousarray(b, np.uint32)
        B = a.shape[0]
        da, db = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
        dout = DeviceBuffer(a.nbytes)
        check(lib().fhe_gate_lvl1_dev(self.h, gate, da.ptr, db.ptr, dout.ptr, result_type, B,
                                      None))
        self.sync()
        return dout.to_array(np.uint32, a.shape)

    def msb_gate(self, a, result_type=LOGIC):
        a = np.ascontiguousarray(a, np.uint32)
        return self._run_dev(lib().fhe_msb_gate_lvl1_dev, [a], a.shape, np.uint32, result_type,
                             a.shape[0])

    def ide_gate(self, a, scale_bits):
        a = np.ascontiguousarray(a, np.uint32)
        return self._run_dev(lib().fhe_ide_gate_lvl1_dev, [a], a.shape, np.uint32, scale_bits,
                             a.shape[0])


def _level_op(ctx, fn, in_level, out_level, a, arg):
    a = np.ascontiguousarray(a, np.uint32 if in_level == 1 else np.uint64)
    B = a.shape[0]
    N = LVL1_N if out_level == 1 else LVL2_N
    dt = np.uint32 if out_level == 1 else np.uint64
    din = DeviceBuffer.from_array(a)
    dout = DeviceBuffer(B * (N + 1) * np.dtype(dt).itemsize)
    check(fn(ctx.h, in_level, out_level, din.ptr, dout.ptr, arg, B, None))
    ctx.sync()
    return dout.to_array(dt, (B, N + 1))


def msb_gate_levels(ctx, in_level, out_level, a, result_type=LOGIC):
    """cuHEDB::MSBGateBootstrapping for (1,1), (2,1) or (2,2)."""
    return _level_op(ctx, lib().fhe_msb_gate_dev, in_level, out_level, a, result_type)


def ide_gate_levels(ctx, in_level, out_level, a, scale_bits):
    """cuHEDB::IdeGateBootstrapping for (1,1), (2,1) or (2,2)."""
    return _level_op(ctx, lib().fhe_ide_gate_dev, in_level, out_level, a, scale_bits)


def log_to_ari(ctx, a):
    a = np.ascontiguousarray(a, np.uint32)
    return ctx._run_dev(lib().fhe_log_to_ari_lvl1_dev, [a], a.shape, np.uint32, a.shape[0])


def log_rescale(ctx, a, scale_bits):
    a = np.ascontiguousarray(a, np.uint32)
    return ctx._run_dev(lib().fhe_log_rescale_lvl1_dev, [a], a.shape, np.uint32, scale_bits,
                        a.shape[0])


# comparison operators (fhe_compare_dev), HomCompare_gpu.cuh:14-76
CMP_GT, CMP_GE, CMP_LT, CMP_LE, CMP_EQ = 0, 1, 2, 3, 4


def hom_msb(ctx, level, a, plain_bits, result_type=LOGIC):
    """HomMSB (msb_gpu.cuh:51-57, 175-192): lvl1 or lvl2 TLWEs -> lvl1 TLWEs."""
    a = np.ascontiguousarray(a, np.uint32 if level == 1 else np.uint64)
    B = a.shape[0]
    din = DeviceBuffer.from_array(a)
    dout = DeviceBuffer(B * (LVL1_N + 1) * 4)
    check(lib().fhe_hom_msb_dev(ctx.h, level, din.ptr, dout.ptr, plain_bits, result_type, B,
                                None))
    ctx.sync()
    return dout.to_array(np.uint32, (B, LVL1_N + 1))


def compare(ctx, level, op, a, b, plain_bits, result_type=LOGIC):
    """a op b (CMP_*) for B pairs of lvl1/lvl2 TLWEs; b may be one TLWE
    (broadcast to every a).  Returns lvl1 TLWEs."""
    dt = np.uint32 if level == 1 else np.uint64
    a = np.ascontiguousarray(a, dt)
    b = np.ascontiguousarray(b, dt)
    B = a.shape[0]
    bcast = 1 if b.ndim == 1 else 0
    da, db = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
    dout = DeviceBuffer(B * (LVL1_N + 1) * 4)
    check(lib().fhe_compare_dev(ctx.h, level, op, da.ptr, db.ptr, bcast, dout.ptr, plain_bits,
                                result_type, B, None))
    ctx.sync()
    return dout.to_array(np.uint32, (B, LVL1_N + 1))


class Graph:
    """A pipeline of _dev calls on ``ctx``'s stream captured into a HIP graph
    (fhe_graph_*).  ``fn()`` is called once while capturing, so its calls are
    recorded, not run; ``launch()`` replays them on the same device buffers.
    Run ``fn`` once eagerly at the same sizes first (scratch cannot be
    allocated during a capture)."""

    def __init__(self, ctx, fn):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().fhe_graph_begin(ctx.h))
        try:
            fn()
        except BaseException:
            g = ctypes.c_void_p()
            if lib().fhe_graph_end(ctx.h, ctypes.byref(g)) == 0:
                lib().fhe_graph_destroy(g)
            raise
        check(lib().fhe_graph_end(ctx.h, ctypes.byref(self.h)))

    def launch(self):
        check(lib().fhe_graph_launch(self.h, None))

    def destroy(self):
        if self.h:
            lib().fhe_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def many_lut_testvector(luts, plain_bits, N=LVL1_N):
    """Test vector of a many-table bootstrap (pbs_many_lut) for messages
    m * 2^(w - plain_bits - 1) (one padding bit; w = 32 for N = 1024, 64 for
    N = 2048): luts is num_out x 2^plain_bits torus words, table k answering
    luts[k][m].  Coefficient j belongs to table j mod 2^bw (bw = bit length of
    num_out - 1, the stride of BlindRotate<P, num_out>'s modulus switch,
    thirdparty/TFHEpp/include/gatebootstrapping.hpp:21-41); its box is
    (base + box/2) // box for base = j - (j mod 2^bw) and box = N / 2^plain_bits,
    the last half box wrapping negacyclically to -luts[k][0]."""
    dt = np.uint32 if N == LVL1_N else np.uint64
    w = 32 if N == LVL1_N else 64
    luts = [[int(v) for v in t] for t in luts]
    K = len(luts)
    stride = 1 << (K - 1).bit_length()
    box = N >> plain_bits
    tv = np.zeros(N, dt)
    for j in range(N):
        k = j % stride
        if k >= K:
            continue
        q = (j - k + box // 2) // box
        v = -luts[k][0] if q == 1 << plain_bits else luts[k][q]
        tv[j] = v % (1 << w)
    return tv


def encrypt_digits(seed, index0, s2, values, digits, digit_bits, alpha):
    """B integers as B x digits lvl2 TLWEs under the lvl2 key s2, least
    significant digit first, digit v at v * 2^(63 - digit_bits) (the input
    layout of compare_digits).  Noise words come from (seed, index0 + row)."""
    v = np.ascontiguousarray(values, np.uint64).reshape(-1)
    shifts = np.arange(digits, dtype=np.uint64) * np.uint64(digit_bits)
    d = (v[:, None] >> shifts[None, :]) & np.uint64((1 << digit_bits) - 1)
    m = d << np.uint64(63 - digit_bits)
    ct = tlwe_encrypt_u64(seed, index0, s2, m.reshape(-1), alpha)
    return ct.reshape(v.shape[0], digits, LVL2_N + 1)


def compare_digits(ctx, op, x, digits, digit_bits, threshold, out_level=1, result_type=LOGIC):
    """x op threshold (CMP_*) for B radix-encrypted integers x (B x digits x
    2049 lvl2 words, as encrypt_digits) against a public threshold: one lvl2
    bootstrap per digit plus one combining bootstrap (fhe_compare_digits_dev).
    Returns lvl1 (out_level 1) or lvl2 TLWEs."""
    x = np.ascontiguousarray(x, np.uint64)
    B = x.shape[0]
    if x.shape != (B, digits, LVL2_N + 1):
        raise ValueError(f"x must be B x {digits} x {LVL2_N + 1} lvl2 words, got {x.shape}")
    dx = DeviceBuffer.from_array(x) if x.size else DeviceBuffer(8)
    N, dt = (LVL1_N, np.uint32) if out_level == 1 else (LVL2_N, np.uint64)
    dout = DeviceBuffer(max(B, 1) * (N + 1) * np.dtype(dt).itemsize)
    check(lib().fhe_compare_digits_dev(ctx.h, op, dx.ptr, digits, digit_bits, int(threshold),
                                       out_level, dout.ptr, result_type, B, None))
    ctx.sync()
    return dout.to_array(dt, (B, N + 1))


def digit_testvectors(op, digits, digit_bits, threshold, result_type=LOGIC):
    """The test vectors of compare_digits built on the host:
    ((digits + 1) x 2048 words, combine body offset, result body offset)."""
    tv = np.zeros((digits + 1, LVL2_N), np.uint64)
    add, post = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().fhe_digit_testvectors(op, digits, digit_bits, int(threshold), result_type,
                                      _ptr(tv), ctypes.byref(add), ctypes.byref(post)))
    return tv, add.value, post.value


def hom_not(ctx, level, a):
    """HomNOT<P>: every word negated (level 0, 1 or 2)."""
    a = np.ascontiguousarray(a, np.uint64 if level == 2 else np.uint32)
    din = DeviceBuffer.from_array(a)
    dout = DeviceBuffer(a.nbytes)
    check(lib().fhe_hom_not_dev(ctx.h, level, din.ptr, dout.ptr, a.shape[0], None))
    ctx.sync()
    return dout.to_array(a.dtype, a.shape)


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().fhe_device_name(buf, 256))
    return buf.value.decode()
