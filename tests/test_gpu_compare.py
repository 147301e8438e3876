"""GPU: the comparison layer -- HomMSB chains (src/cuHEDB/msb_gpu.cuh:7-193)
and greater_than .. equal (src/cuHEDB/HomCompare_gpu.cuh:14-76) -- batched,
decrypted and checked bit-exactly against the plaintext comparison.

Encoding as in test/HE3_Query/HomComp.cu:75-80: a `bits`-bit value v at
scale 2^(w - bits - 1) (one padding bit).  LOGIC results decode as
phase > 0, ARITHMETIC results as round(phase / 2^31) mod 2.
"""

import numpy as np
import pytest

import cufhedb_amd as F
import _oracle as O

pytestmark = pytest.mark.gpu

OPS = {
    F.CMP_GT: np.greater,
    F.CMP_GE: np.greater_equal,
    F.CMP_LT: np.less,
    F.CMP_LE: np.less_equal,
    F.CMP_EQ: np.equal,
}


def enc(k, level, v, bits, seed):
    v = np.asarray(v, np.uint64)
    if level == 1:
        m = (v << np.uint64(32 - bits - 1)).astype(np.uint32)
        return F.tlwe_encrypt_u32(seed, 0, k.sk.s1, m, k.params.alpha1)
    m = v << np.uint64(64 - bits - 1)
    return F.tlwe_encrypt_u64(seed, 0, k.sk.s2, m, k.params.alpha2)


def dec(k, c, result_type):
    ph = O.tdiff32(F.tlwe_phase_u32(c, k.sk.s1), 0)
    if result_type == F.LOGIC:
        return (ph > 0).astype(np.int64)
    return np.round(ph / 2.0 ** 31).astype(np.int64) % 2


def operands(rng, bits, n):
    hi = 1 << bits
    a = rng.integers(0, hi, n)
    b = rng.integers(0, hi, n)
    # edge cases: equal pairs, extremes, neighbours
    edge_a = [0, hi - 1, 0, hi - 1, 5 % hi, 7 % hi, hi // 2, hi // 2]
    edge_b = [0, hi - 1, hi - 1, 0, 5 % hi, 6 % hi, hi // 2 - 1, hi // 2 + 1]
    a[: len(edge_a)] = edge_a
    b[: len(edge_b)] = edge_b
    a[len(edge_a): len(edge_a) + 4] = b[len(edge_a): len(edge_a) + 4]
    return a, b


@pytest.mark.parametrize("bits", [4, 8, 9])
def test_compare_lvl1(full636, rng, bits):
    k, ctx = full636
    a, b = operands(rng, bits, 24)
    ca, cb = enc(k, 1, a, bits, 7100 + bits), enc(k, 1, b, bits, 7200 + bits)
    for op, ref in OPS.items():
        for rt in (F.LOGIC, F.ARITHMETIC):
            out = F.compare(ctx, 1, op, ca, cb, bits, rt)
            got = dec(k, out, rt)
            assert np.array_equal(got, ref(a, b).astype(np.int64)), (bits, op, rt, a, b, got)


# Plain widths stay at or below 26 bits (27 with the sign): a peeling step
# shifts the previous identity bootstrap's output noise left by pb - 6 - 5,
# and the lvl2 PBS output noise (std ~2^36.2 on the 64-bit torus, set by
# FP64 rounding of 64-bit products as in the reference's FP64 transforms,
# times sqrt(2) from applying X^a - 1 in the frequency domain)
# then reaches the 2^57 decision margin of the 6-bit window from about
# pb = 28 on; windows next to a slot boundary (plaintexts 0 and -1) are the
# first to flip.  The reference's ImExtractMSB29/33 share this limit.
@pytest.mark.parametrize("bits", [12, 20, 26])
def test_compare_lvl2(full672, rng, bits):
    k, ctx = full672
    a, b = operands(rng, bits, 16)
    ca, cb = enc(k, 2, a, bits, 7300 + bits), enc(k, 2, b, bits, 7400 + bits)
    for op in (F.CMP_GE, F.CMP_LT, F.CMP_EQ):
        out = F.compare(ctx, 2, op, ca, cb, bits, F.LOGIC)
        assert np.array_equal(dec(k, out, F.LOGIC), OPS[op](a, b).astype(np.int64)), (bits, op)


def test_compare_lvl2_public_threshold_broadcast(full672, rng):
    # the 20-bit range query x <= t with one encrypted threshold shared by all rows
    k, ctx = full672
    bits = 20
    x = rng.integers(0, 1 << bits, 40)
    t = int(rng.integers(0, 1 << bits))
    x[:3] = [t, t - 1, t + 1]
    cx = enc(k, 2, x, bits, 7501)
    ct = enc(k, 2, [t], bits, 7502)[0]
    for op in (F.CMP_LE, F.CMP_GT):
        out = F.compare(ctx, 2, op, cx, ct, bits, F.ARITHMETIC)
        assert np.array_equal(dec(k, out, F.ARITHMETIC), OPS[op](x, t).astype(np.int64)), op


@pytest.mark.parametrize("level,pb", [(1, 5), (1, 7), (1, 10), (2, 6), (2, 14), (2, 24), (2, 27)])
def test_hom_msb(full636, full672, rng, level, pb):
    # signed pb-bit plaintexts in the top pb bits; HomMSB = sign bit
    k, ctx = full636 if level == 1 else full672
    w = 32 if level == 1 else 64
    v = rng.integers(0, 1 << pb, 16, dtype=np.uint64)
    v[:4] = [0, (1 << (pb - 1)) - 1, 1 << (pb - 1), (1 << pb) - 1]
    m = v << np.uint64(w - pb)
    c = (F.tlwe_encrypt_u32(7600 + pb, 0, k.sk.s1, m.astype(np.uint32), k.params.alpha1)
         if level == 1 else F.tlwe_encrypt_u64(7600 + pb, 0, k.sk.s2, m, k.params.alpha2))
    want = ((v >> np.uint64(pb - 1)) & np.uint64(1)).astype(np.int64)
    for rt in (F.LOGIC, F.ARITHMETIC):
        out = F.hom_msb(ctx, level, c, pb, rt)
        assert np.array_equal(dec(k, out, rt), want), (level, pb, rt)


def test_hom_not_bit_exact(full636, rng):
    k, ctx = full636
    for level, dt, w in ((1, np.uint32, 1025), (2, np.uint64, 2049), (0, np.uint32, k.n0 + 1)):
        a = rng.integers(0, np.iinfo(dt).max, (5, w), dtype=np.uint64).astype(dt)
        assert np.array_equal(F.hom_not(ctx, level, a), (0 - a.astype(np.uint64)).astype(dt))


def test_compare_rejects_bad_arguments(full636):
    k, ctx = full636
    c = np.zeros((2, 1025), np.uint32)
    with pytest.raises(F.FheError):
        F.compare(ctx, 1, F.CMP_LT, c, c, 10)      # plain_bits + 1 > 10 at lvl1
    with pytest.raises(F.FheError):
        F.compare(ctx, 1, 9, c, c, 4)              # unknown op
    with pytest.raises(F.FheError):
        F.hom_msb(ctx, 1, c, 11)


def test_lvl2_width_limit_rejected(full672):
    # level-2 comparisons of more than 26 bits (HomMSB inputs of more than
    # 27) return wrong results at these parameters: refused before any launch
    k, ctx = full672
    c = np.zeros((2, 2049), np.uint64)
    with pytest.raises(F.FheError) as e:
        F.compare(ctx, 2, F.CMP_LT, c, c, 27)
    assert e.value.code == F.FHE_ERR_UNSUPPORTED
    with pytest.raises(F.FheError) as e:
        F.hom_msb(ctx, 2, c, 28)
    assert e.value.code == F.FHE_ERR_UNSUPPORTED
