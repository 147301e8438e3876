"""Integer identities the HIP kernels rely on, checked exhaustively or on
seeded random words on the CPU (no GPU): a wrong constant or index in the
kernels' closed forms fails here before any device run.

  * rotation from the extended columns of the lvl1 blind rotations
    (rot1.hpp, used by pbs_lvl1.hip and pbs_lvl1_mfma.hip): LDS holds
    column 0, the 2304-word transform tile, then
    column 1; column c's negation is written as 16 rows of 64 words after
    column 0 / before column 1, and with r = (L - a) mod 2^32 the lane's 16
    words are read at base + 64 k (base: column start, minus 1024 words for
    column 1, plus r mod 1024), s = 0 or ~0 from bit 10 of r, inverted for
    column 1; ((v ^ s) + (K - s - p_n)) ^ F equals the prepared word
    ((X^a - 1) p)_n + K) ^ F of the plain rotation (utils.hpp:110-141,
    trgsw.cpp:21-39);
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


def test_offset_closed_form():
    # offsetgen (trgsw.cpp:9-18) plus the rounding half: 2181562368 + 8192
    assert K == 2181562368 + 8192


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


# LDS words of the lvl1 blind rotations: column 0, tile (576 double2), column 1
TILE_WORDS = 576 * 4
COL1 = N + TILE_WORDS


def extended_rotation(p, a, c, flip_col1=True):
    """The kernel's column preparation for column c, lane by lane, on a model
    of its LDS: returns the prepared words in coefficient order n = L + 64 k."""
    rng = np.random.default_rng(a + 7 * c)
    lds = rng.integers(0, 1 << 32, COL1 + N, dtype=np.uint64)  # stale tile / other column
    col = c * COL1
    lds[col:col + N] = p
    neg = col + N if c == 0 else col - N
    lds[neg:neg + N] = (-p.astype(np.int64)) % (1 << 32)
    out = np.zeros(N, np.uint64)
    for lane in range(64):
        rr = (lane - a) % (1 << 32)
        base = col - c * N + (rr & (N - 1))
        s = M32 if (rr >> 10) & 1 else 0
        if c == 1 and flip_col1:
            s ^= M32
        kc = (K - s) & M32
        k = np.arange(16)
        idx = base + 64 * k
        assert idx.min() >= 0 and idx.max() < COL1 + N
        # column 0's reads stay below column 1, column 1's above column 0
        assert (c == 0 and idx.max() < COL1) or (c == 1 and idx.min() >= N)
        v = lds[idx]
        pn = p[lane + 64 * k].astype(np.uint64)
        out[lane + 64 * k] = (((v ^ np.uint64(s)) + ((np.uint64(kc) - pn) & np.uint64(M32)))
                              & np.uint64(M32)) ^ np.uint64(FLIP)
    return out


def test_extended_column_rotation_every_shift():
    rng = np.random.default_rng(2027)
    for c in (0, 1):
        p = rng.integers(0, 1 << 32, N, dtype=np.uint64)
        p[:4] = [0, 1, M32, 1 << 31]
        for a in range(0, 2 * N, 3 if c else 1):  # every shift for column 0, a third for column 1
            want = ((rot_minus_one(p, a) + np.uint64(K)) & np.uint64(M32)) ^ np.uint64(FLIP)
            assert np.array_equal(extended_rotation(p, a, c), want), (c, a)


def test_planted_unflipped_column1_sign_is_caught():
    rng = np.random.default_rng(5)
    p = rng.integers(0, 1 << 32, N, dtype=np.uint64)
    want = ((rot_minus_one(p, 300) + np.uint64(K)) & np.uint64(M32)) ^ np.uint64(FLIP)
    assert not np.array_equal(extended_rotation(p, 300, 1, flip_col1=False), want)
