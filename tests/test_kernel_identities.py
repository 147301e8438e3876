"""Integer identities the HIP kernels rely on, checked exhaustively or on
seeded random words on the CPU (no GPU): a wrong constant or index in the
kernels' closed forms fails here before any device run.

  * select-free rotation of the lvl1 blind rotation (pbs_lvl1.hip,
    pbs_lvl1_mfma.hip): with t = 4 (n - a) mod 8192, the word of X^a p at
    coefficient n is p[(t & 4092) / 4], negated when bit 12 of t is set, and
    ((v ^ s) + (K - s - p_n)) ^ F, s = 0 or ~0 from that bit, equals the
    prepared word ((X^a - 1) p)_n + K) ^ F of the plain rotation
    (utils.hpp:110-141, trgsw.cpp:21-39);
  * digit extraction from the prepared word: the signed 6-bit field d equals
    DecompositionPolynomial's digit d (trgsw.cpp:21-39).
"""

import numpy as np

N = 1024
BGBIT, L = 6, 3
# offset of DecompositionPolynomial (offsetgen + rounding half, trgsw.cpp:9-29)
K = sum((1 << (BGBIT - 1)) << (32 - (d + 1) * BGBIT) for d in range(L)) + (1 << (32 - L * BGBIT - 1))
FLIP = (1 << 31) | (1 << 25) | (1 << 19)
M32 = (1 << 32) - 1


def rot_minus_one(p, a):
    """((X^a - 1) p) mod 2^32 by the definition (negacyclic, N = 1024)."""
    n = np.arange(N)
    idx = (n - a) % (2 * N)
    v = p[idx % N].astype(np.int64)
    v = np.where(idx >= N, -v, v)
    return ((v - p.astype(np.int64)) % (1 << 32)).astype(np.uint64)


def select_free(p, a):
    """The kernels' form: t = 4 (n - a) mod 8192, s from bit 12 of t."""
    n = np.arange(N, dtype=np.int64)
    t = (4 * (n - a)) % 8192
    v = p[(t & 4092) >> 2].astype(np.uint64)
    s = np.where((t >> 12) & 1, np.uint64(M32), np.uint64(0))
    q = (np.uint64(K) - s - p.astype(np.uint64)) & np.uint64(M32)
    return (((v ^ s) + q) & np.uint64(M32)) ^ np.uint64(FLIP)


def test_offset_closed_form():
    # offsetgen (trgsw.cpp:9-18) plus the rounding half: 2181562368 + 8192
    assert K == 2181562368 + 8192


def test_select_free_rotation_every_shift():
    rng = np.random.default_rng(2026)
    p = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    p[:4] = [0, 1, M32, 1 << 31]  # edge words
    for a in range(2 * N):  # every rotation
        want = ((rot_minus_one(p, a) + np.uint64(K)) & np.uint64(M32)) ^ np.uint64(FLIP)
        assert np.array_equal(select_free(p, a), want), a


def test_digit_fields_equal_decomposition():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.integers(0, 1 << 32, 4096, dtype=np.uint64),
                        np.array([0, 1, M32, 1 << 31, (1 << 31) - 1], np.uint64)])
    u = ((x + np.uint64(K)) & np.uint64(M32)) ^ np.uint64(FLIP)
    for d in range(L):
        sh = 32 - BGBIT * (d + 1)
        field = ((u >> np.uint64(sh)) & np.uint64(63)).astype(np.int64)
        signed = np.where(field >= 32, field - 64, field)
        # DecompositionPolynomial: ((x + offset) >> sh & 63) - 32
        ref = (((x + np.uint64(K)) & np.uint64(M32)) >> np.uint64(sh) & np.uint64(63)).astype(np.int64) - 32
        assert np.array_equal(signed, ref), d
    # and the digits recombine to x within the rounding of the last digit
    rec = sum((((((x + np.uint64(K)) & np.uint64(M32)) >> np.uint64(32 - BGBIT * (d + 1)))
                & np.uint64(63)).astype(np.int64) - 32) << (32 - BGBIT * (d + 1)) for d in range(L))
    gap = (rec - x.astype(np.int64)) % (1 << 32)
    gap = np.where(gap >= 1 << 31, gap - (1 << 32), gap)
    assert np.max(np.abs(gap)) <= 1 << (32 - L * BGBIT - 1)


def test_planted_wrong_sign_bit_is_caught():
    # taking the sign from bit 11 instead of bit 12 of t must fail the check
    rng = np.random.default_rng(3)
    p = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    a = 700
    n = np.arange(N, dtype=np.int64)
    t = (4 * (n - a)) % 8192
    v = p[(t & 4092) >> 2].astype(np.uint64)
    s = np.where((t >> 11) & 1, np.uint64(M32), np.uint64(0))
    q = (np.uint64(K) - s - p.astype(np.uint64)) & np.uint64(M32)
    bad = (((v ^ s) + q) & np.uint64(M32)) ^ np.uint64(FLIP)
    want = ((rot_minus_one(p, a) + np.uint64(K)) & np.uint64(M32)) ^ np.uint64(FLIP)
    assert not np.array_equal(bad, want)
