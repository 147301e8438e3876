"""Digit-wise comparison with a public threshold: host logic (no GPU).

fhe_digit_testvectors builds the test vectors the device operator uses.  It
is checked against the oracle's construction word for word, and -- what
pins both -- against the mathematics: over every pair (x, t) of small
radixes, evaluating each digit's test vector where the digit's phase lands,
summing, and applying the sign bootstrap must give Python's own x op t, with
the phase at least 2^-(digits+1) from both decision points.  The test
vectors' boxes are checked over every rotation of the negacyclic extension.
"""

import numpy as np
import pytest

import cufhedb_amd as F

import _oracle as O

N = 2048
OPS = {F.CMP_GT: np.greater, F.CMP_GE: np.greater_equal, F.CMP_LT: np.less,
       F.CMP_LE: np.less_equal, F.CMP_EQ: np.equal}
SHAPES = [(1, 1), (1, 4), (2, 4), (3, 3), (5, 2), (5, 4)]


def _ext(tv_row, p):
    """Constant coefficient of X^-p * tv (negacyclic): the value a blind
    rotation by p reads."""
    p = np.asarray(p) % (2 * N)
    return np.where(p < N, tv_row[p % N], (0 - tv_row[p % N].astype(np.uint64)))


@pytest.mark.parametrize("digits,bits", SHAPES)
@pytest.mark.parametrize("op", list(OPS))
def test_testvectors_match_oracle(op, digits, bits):
    rng = np.random.default_rng(op * 100 + digits * 10 + bits)
    top = (1 << (digits * bits)) - 1
    for t in {0, top, int(rng.integers(0, top + 1))}:
        for rt in (F.LOGIC, F.ARITHMETIC):
            tv, add, post = F.digit_testvectors(op, digits, bits, t, rt)
            otv, oadd, opost = O.digit_testvectors(op, digits, bits, t, rt)
            assert np.array_equal(tv, otv) and add == oadd and post == opost, (t, rt)


def _decide(op, digits, bits, t, xs, rt):
    """Run the operator's arithmetic on noiseless phases for every x in xs:
    digit i of x read from its test vector at the rotation its phase gives,
    the words summed with the combine offset, then the sign bootstrap."""
    tv, add, post = F.digit_testvectors(op, digits, bits, t, rt)
    box = N >> bits
    acc = np.full(xs.shape, add, np.uint64)
    for i in range(digits):
        v = (xs >> (i * bits)) & ((1 << bits) - 1)
        acc += _ext(tv[i], v * box)
    s = acc.view(np.int64)
    fw = int(tv[digits][0])
    res = np.where(s >= 0, np.uint64(fw), np.uint64((-fw) % (1 << 64))) + np.uint64(post)
    return s, res


@pytest.mark.parametrize("digits,bits", [(2, 4), (3, 3), (5, 2), (4, 2), (2, 2), (1, 4)])
@pytest.mark.parametrize("op", list(OPS))
def test_decision_exhaustive(op, digits, bits):
    R = 1 << (digits * bits)
    xs = np.arange(R, dtype=np.int64)
    margin = 1 << (63 - digits)            # 2^-(digits+1) on the 64-bit torus
    for t in range(R):
        for rt in (F.LOGIC, F.ARITHMETIC):
            s, res = _decide(op, digits, bits, t, xs, rt)
            if rt == F.LOGIC:
                got = res.view(np.int64) > 0
                mu = np.uint64(1 << 61)
                assert np.all((res == mu) | (res == np.uint64((1 << 64) - (1 << 61))))
            else:
                got = res == np.uint64(1 << 63)
                assert np.all((res == 0) | got)
            want = OPS[op](xs, t)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, (t, rt, xs[bad[:4]])
        # distance of every phase to 0 and to 1/2
        a = np.abs(s.astype(np.float64))
        assert a.min() >= margin and (2.0 ** 63 - a).min() >= margin, t


@pytest.mark.parametrize("digits,bits", SHAPES)
def test_digit_boxes_every_rotation(digits, bits):
    """A digit's phase may drift by less than half a box either way (across
    zero for digit 0) and still read its own value."""
    box = N >> bits
    R = 1 << bits
    t = int(np.random.default_rng(digits * bits).integers(0, 1 << (digits * bits)))
    for op in OPS:
        tv, _, _ = F.digit_testvectors(op, digits, bits, t, F.LOGIC)
        terms = O.digit_terms(op, digits, bits, t)
        for i in range(digits):
            for v in range(R):
                ps = v * box + np.arange(-(box // 2), box // 2)
                want = np.uint64((terms[i][v] << (64 - digits)) % (1 << 64))
                assert np.all(_ext(tv[i], ps) == want), (op, i, v)


def test_combine_range_fits_the_torus():
    """S + 1/2 over every digit pattern spans 2^digits integers, the widest
    range one sign bootstrap separates with a margin of half a step."""
    for digits in range(1, 6):
        for op in (F.CMP_LE, F.CMP_GE):
            terms = O.digit_terms(op, digits, 2, 0b10 * sum(4 ** i for i in range(digits)))
            lo = sum(min(r) for r in terms)
            hi = sum(max(r) for r in terms)
            assert hi - lo + 1 <= 1 << digits, (digits, op, lo, hi)


@pytest.mark.parametrize("args", [
    (7, 5, 4, 0), (F.CMP_LE, 0, 4, 0), (F.CMP_LE, 6, 4, 0), (F.CMP_LE, 5, 0, 0),
    (F.CMP_LE, 5, 5, 0), (F.CMP_LE, 5, 4, 1 << 20), (F.CMP_LE, 2, 3, 64)])
def test_bad_arguments_rejected(args):
    with pytest.raises(F.FheError):
        F.digit_testvectors(*args)
    with pytest.raises(F.FheError):
        F.digit_testvectors(F.CMP_LE, 5, 4, 0, 7)


def test_encrypt_digits_layout():
    p = F.preset(F.PRESET_CFG636)
    sk = F.SecretKeys(p, 99)
    vals = np.array([0, 1, 0xFFFFF, 0x12345, 0xA0B0C], np.uint64)
    x = F.encrypt_digits(7, 0, sk.s2, vals, 5, 4, p.alpha2)
    assert x.shape == (5, 5, N + 1)
    ph = O.tlwe_phase64(x.reshape(-1, N + 1), sk.s2).reshape(5, 5)
    dec = ((ph + np.uint64(1 << 58)) >> np.uint64(59)) & np.uint64(31)
    for b, v in enumerate(vals):
        assert [int(d) for d in dec[b]] == [(int(v) >> (4 * i)) & 15 for i in range(5)]
