"""ctypes bindings to the checker libraries (test infrastructure only).

* ``orc``: oracle/liboracle.so, the C restatement of the reference path.
* ``ref``: oracle/_ref/libtfhepp_ref.so, the reference's own TFHEpp compiled
  from its sources (present only where the reference sources were available
  at build time; tests that need it skip otherwise).
"""

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtfhepp_ref.so")

_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32


def P(a):
    """Pointer argument that keeps `a` alive for the duration of the call."""
    return a.ctypes.data_as(ctypes.c_void_p)


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_phase_u32.restype = ctypes.c_uint32
        L.orc_phase_u64.restype = ctypes.c_uint64
        L.orc_fd_rev.restype = _i
        _orc = L
    return _orc


def ref_built():
    """Whether the compiled reference exists (without loading it)."""
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def ref_n0():
    r = ref()
    return r.ref_lvl0_n() if r is not None else None


# ---------------------------------------------------------------- wrappers
def ifft32(a, N=1024):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.zeros(N)
    orc().orc_ifft_torus32(P(out), P(a), N)
    return out


def fft32(fd, N=1024):
    fd = np.ascontiguousarray(fd, np.float64)
    out = np.zeros(N, np.uint32)
    orc().orc_fft_torus32(P(out), P(fd), N)
    return out


def ifft64(a, N=2048):
    a = np.ascontiguousarray(a, np.uint64)
    out = np.zeros(N)
    orc().orc_ifft_torus64(P(out), P(a), N)
    return out


def fft64(fd, N=2048):
    fd = np.ascontiguousarray(fd, np.float64)
    out = np.zeros(N, np.uint64)
    orc().orc_fft_torus64(P(out), P(fd), N)
    return out


def naive32(a, b):
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    out = np.zeros_like(b)
    orc().orc_negacyclic_naive_u32(P(out), P(a), P(b), b.shape[0])
    return out


def naive64(a, b):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.zeros_like(b)
    orc().orc_negacyclic_naive_u64(P(out), P(a), P(b), b.shape[0])
    return out


def decompose32(poly, digit, l=3, Bgbit=6):
    poly = np.ascontiguousarray(poly, np.uint32)
    out = np.zeros_like(poly)
    orc().orc_decompose_u32(P(out), P(poly), digit, poly.shape[0], l, Bgbit)
    return out


def decompose64(poly, digit, l=4, Bgbit=9):
    poly = np.ascontiguousarray(poly, np.uint64)
    out = np.zeros_like(poly)
    orc().orc_decompose_u64(P(out), P(poly), digit, poly.shape[0], l, Bgbit)
    return out


def mul_xai32(poly, a, minus_one=False):
    poly = np.ascontiguousarray(poly, np.uint32)
    out = np.zeros_like(poly)
    f = orc().orc_mul_xai_minus_one_u32 if minus_one else orc().orc_mul_xai_u32
    f(P(out), P(poly), _u32(a), poly.shape[0])
    return out


def trgsw_to_fft32(trgsw, l=3):
    trgsw = np.ascontiguousarray(trgsw, np.uint32)
    N = trgsw.shape[-1]
    out = np.zeros(trgsw.shape, np.float64)
    orc().orc_trgsw_to_fft_u32(P(out), P(trgsw), N, l)
    return out


def bk_to_fft32(bk):
    """[n0][6][2][N] uint32 -> FD doubles, same layout (ApplyFFT2trgsw per entry)."""
    out = np.zeros(bk.shape, np.float64)
    for i in range(bk.shape[0]):
        out[i] = trgsw_to_fft32(bk[i])
    return out


def external_product32(trlwe, trgswfft, l=3, Bgbit=6):
    trlwe = np.ascontiguousarray(trlwe, np.uint32)
    fd = np.ascontiguousarray(trgswfft, np.float64)
    out = np.zeros_like(trlwe)
    orc().orc_external_product_u32(P(out), P(trlwe), P(fd), trlwe.shape[-1], l, Bgbit)
    return out


def blind_rotate32(tlwe0, bkfft, tv, nbit=10, l=3, Bgbit=6):
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    N = tv.shape[0]
    out = np.zeros((2, N), np.uint32)
    tv = np.ascontiguousarray(tv, np.uint32)
    orc().orc_blind_rotate_u32(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit,
                               l, Bgbit)
    return out


def pbs32(tlwe0, bkfft, tv, nbit=10, l=3, Bgbit=6):
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    N = tv.shape[0]
    out = np.zeros(N + 1, np.uint32)
    tv = np.ascontiguousarray(tv, np.uint32)
    orc().orc_pbs_u32(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit, l, Bgbit)
    return out


def pbs_many32(tlwe0, bkfft, tv, num_out, nbit=10, l=3, Bgbit=6):
    """GateBootstrappingManyLUT<lvl01param, num_out>: num_out x (N+1) words."""
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    N = tv.shape[0]
    out = np.zeros((num_out, N + 1), np.uint32)
    tv = np.ascontiguousarray(tv, np.uint32)
    orc().orc_pbs_many_u32(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit, l,
                           Bgbit, num_out)
    return out


def pbs_many64(tlwe0, bkfft, tv, num_out, nbit=11, l=4, Bgbit=9):
    """GateBootstrappingManyLUT<lvl02param, num_out>: num_out x (N+1) words."""
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    N = tv.shape[0]
    out = np.zeros((num_out, N + 1), np.uint64)
    tv = np.ascontiguousarray(tv, np.uint64)
    orc().orc_pbs_many_u64(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit, l,
                           Bgbit, num_out)
    return out


def sample_extract32(trlwe, index=0):
    trlwe = np.ascontiguousarray(trlwe, np.uint32)
    N = trlwe.shape[-1]
    out = np.zeros(N + 1, np.uint32)
    orc().orc_sample_extract_u32(P(out), P(trlwe), index, N)
    return out


def iks32(tlwe, ksk, tn, t=2, basebit=10):
    tlwe = np.ascontiguousarray(tlwe, np.uint32)
    out = np.zeros(tn + 1, np.uint32)
    orc().orc_iks_32to32(P(out), P(tlwe), P(ksk), tlwe.shape[0] - 1, tn, t, basebit)
    return out


def iks64(tlwe, ksk, tn, t, basebit):
    tlwe = np.ascontiguousarray(tlwe, np.uint64)
    out = np.zeros(tn + 1, np.uint32)
    orc().orc_iks_64to32(P(out), P(tlwe), P(ksk), tlwe.shape[0] - 1, tn, t, basebit)
    return out


def trlwe_phase32(c, key):
    c = np.ascontiguousarray(c, np.uint32).reshape(2, -1)
    key = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(c.shape[1], np.uint32)
    orc().orc_trlwe_phase_u32(P(out), P(c), P(key), c.shape[1])
    return out


def hom_nand(ca, cb, ksk10, n0, bkfft):
    out = np.zeros(1025, np.uint32)
    ca = np.ascontiguousarray(ca, np.uint32)
    cb = np.ascontiguousarray(cb, np.uint32)
    orc().orc_hom_nand_lvl1(P(out), P(ca), P(cb), P(ksk10), n0, P(bkfft))
    return out


def tdiff32(a, b):
    """signed difference a - b on the 32-bit torus (int64 array)."""
    d = (np.asarray(a, np.uint32).astype(np.int64) - np.asarray(b, np.uint32).astype(np.int64))
    return (d + (1 << 31)) % (1 << 32) - (1 << 31)


def tdiff64(a, b):
    return (np.asarray(a, np.uint64) - np.asarray(b, np.uint64)).view(np.int64)


# ---------------------------------------------------------------- lvl2 (64-bit)
def trgsw_to_fft64(trgsw, l=4):
    trgsw = np.ascontiguousarray(trgsw, np.uint64)
    out = np.zeros(trgsw.shape, np.float64)
    orc().orc_trgsw_to_fft_u64(P(out), P(trgsw), trgsw.shape[-1], l)
    return out


def bk_to_fft64(bk):
    out = np.zeros(bk.shape, np.float64)
    for i in range(bk.shape[0]):
        out[i] = trgsw_to_fft64(bk[i])
    return out


def external_product64(trlwe, trgswfft, l=4, Bgbit=9):
    trlwe = np.ascontiguousarray(trlwe, np.uint64)
    fd = np.ascontiguousarray(trgswfft, np.float64)
    out = np.zeros_like(trlwe)
    orc().orc_external_product_u64(P(out), P(trlwe), P(fd), trlwe.shape[-1], l, Bgbit)
    return out


def pbs64(tlwe0, bkfft, tv, nbit=11, l=4, Bgbit=9):
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    tv = np.ascontiguousarray(tv, np.uint64)
    N = tv.shape[0]
    out = np.zeros(N + 1, np.uint64)
    orc().orc_pbs_u64(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit, l, Bgbit)
    return out


def blind_rotate64(tlwe0, bkfft, tv, nbit=11, l=4, Bgbit=9):
    tlwe0 = np.ascontiguousarray(tlwe0, np.uint32)
    tv = np.ascontiguousarray(tv, np.uint64)
    N = tv.shape[0]
    out = np.zeros((2, N), np.uint64)
    orc().orc_blind_rotate_u64(P(out), P(tlwe0), tlwe0.shape[0] - 1, P(bkfft), P(tv), N, nbit,
                               l, Bgbit)
    return out


def sample_extract64(trlwe, index=0):
    trlwe = np.ascontiguousarray(trlwe, np.uint64)
    N = trlwe.shape[-1]
    out = np.zeros(N + 1, np.uint64)
    orc().orc_sample_extract_u64(P(out), P(trlwe), index, N)
    return out


def trlwe_phase64(c, key):
    c = np.ascontiguousarray(c, np.uint64).reshape(2, -1)
    key = np.ascontiguousarray(key, np.uint64)
    out = np.zeros(c.shape[1], np.uint64)
    orc().orc_trlwe_phase_u64(P(out), P(c), P(key), c.shape[1])
    return out


def tlwe_phase64(c, key):
    c = np.atleast_2d(np.asarray(c, np.uint64))
    n = key.shape[0]
    dot = (c[:, :n] * np.asarray(key, np.uint64)[None, :]).sum(axis=1, dtype=np.uint64)
    return c[:, n] - dot


# ------------------------------------------- digit-wise comparison (lvl2)
# No reference counterpart: the operator replaces the HomMSB-chain compare
# (src/cuHEDB/HomCompare_gpu.cuh:14-76) for a threshold known in the clear.
# This restatement is pinned by the mathematics: decisions against Python's
# own comparison operators over every (x, t) pair of small radixes
# (test_digits.py) and over encrypted rows (test_gpu_digits.py).
CMP_GT, CMP_GE, CMP_LT, CMP_LE, CMP_EQ = 0, 1, 2, 3, 4


def digit_terms(op, digits, bits, t):
    """terms[i][v]: the integer digit position i contributes for digit value
    v when the threshold is t -- the lexicographic comparison unrolled: the
    least significant digit as a 0/1 comparison, each higher digit as a
    sign weighted by 2^(i-1), which exceeds everything below it; for
    equality, 1 per differing digit."""
    R = 1 << bits
    terms = []
    for i in range(digits):
        ti = (t >> (i * bits)) % R
        row = []
        for v in range(R):
            if op == CMP_EQ:
                row.append(int(v != ti))
            elif i == 0:
                row.append(int(v > ti) if op in (CMP_GT, CMP_LE) else int(v >= ti))
            else:
                row.append(((v > ti) - (v < ti)) << (i - 1))
        terms.append(row)
    return terms


def digit_testvectors(op, digits, bits, t, result_type=1):
    """(digits + 1) x 2048 test vectors: digit v's box is the coefficients
    j whose phase j / 4096 rounds to v / 2^(bits+1); the box that rounds to
    1/2 belongs to digit 0 approached from below zero and holds the negated
    value (negacyclic wrap).  Row `digits` is the sign bootstrap's constant."""
    N = 2048
    R = 1 << bits
    j = np.arange(N)
    v = (j * 2 * R + N) // (2 * N)          # floor(j / box + 1/2), box = N / R
    terms = digit_terms(op, digits, bits, t)
    tv = np.zeros((digits + 1, N), np.uint64)
    for i in range(digits):
        f = np.array([terms[i][x % R] for x in v], dtype=np.int64)
        f = np.where(v == R, -f, f)
        tv[i] = (f.astype(np.uint64) << np.uint64(64 - digits))
    mu = 1 << 61
    u = mu if result_type else mu << 1
    tv[digits] = u if op in (CMP_GT, CMP_GE) else (1 << 64) - u
    combine_add = (1 << 64) - (1 << (63 - digits))
    post_add = 0 if result_type else u
    return tv, combine_add, post_add


def compare_digits(x, op, digits, bits, t, ksk20, bkfft2, n0, result_type=1, ksk21=None):
    """x: B x digits lvl2 TLWEs -> B lvl2 (ksk21 None) or lvl1 TLWEs, through
    the oracle's key switch and bootstrap."""
    tv, add, post = digit_testvectors(op, digits, bits, t, result_type)
    out = []
    for row in np.asarray(x, np.uint64):
        acc = np.zeros(2049, np.uint64)
        for i in range(digits):
            acc += pbs64(iks64(row[i], ksk20, n0, 2, 10), bkfft2, tv[i])
        acc[2048:] += np.uint64(add)
        res = pbs64(iks64(acc, ksk20, n0, 2, 10), bkfft2, tv[digits])
        res[2048:] += np.uint64(post)
        out.append(res if ksk21 is None else iks64(res, ksk21, 1024, 10, 3))
    return np.array(out)
