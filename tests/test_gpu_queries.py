"""GPU: encrypted query pipelines on the batched operators, checked
bit-exactly against the plaintext query.

* range filter x <= t over a column of 20-bit lvl2 values with one shared
  encrypted threshold (cuHEDB::less_than_equal per row in the reference);
* TPC-H Q1 filter (test/HE3_Query/TPCH_Q1.cu:215-295): group predicates under
  shipdate <= date, log_rescale to 2^29, decrypt, then SUM(quantity) and
  COUNT per (returnflag, linestatus) group as in TPCH_Q1.cu:176-200.
"""

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import queries as Q

pytestmark = pytest.mark.gpu


def enc1(k, v, bits, seed):
    m = (np.asarray(v, np.uint64) << np.uint64(32 - bits - 1)).astype(np.uint32)
    return F.tlwe_encrypt_u32(seed, 0, k.sk.s1, m, k.params.alpha1)


def test_range_filter_lvl2(full672, rng):
    k, ctx = full672
    rows, bits = 1024, 20
    x = rng.integers(0, 1 << bits, rows)
    t = int(rng.integers(1, (1 << bits) - 1))
    x[:4] = [t, t - 1, t + 1, 0]
    cx = F.tlwe_encrypt_u64(8101, 0, k.sk.s2, x.astype(np.uint64) << np.uint64(64 - bits - 1),
                            k.params.alpha2)
    ct = F.tlwe_encrypt_u64(8102, 0, k.sk.s2, np.array([t], np.uint64) << np.uint64(64 - bits - 1),
                            k.params.alpha2)
    dx, dt = F.DeviceBuffer.from_array(cx), F.DeviceBuffer.from_array(ct)
    out = Q.range_filter(ctx, dx, dt, rows, bits)
    ctx.sync()
    ph = F.tlwe_phase_u32(out.to_array(np.uint32, (rows, 1025)), k.sk.s1).view(np.int32)
    assert np.array_equal((ph > 0).astype(np.int64), (x <= t).astype(np.int64))


def test_tpch_q1_filter_and_aggregate(full636, rng):
    k, ctx = full636
    rows, date_bits, qty_bits = 512, 9, 8
    rf = rng.integers(0, 2, rows)
    ls = rng.integers(0, 2, rows)
    date = rng.integers(0, 1 << date_bits, rows)
    qty = rng.integers(0, 1 << qty_bits, rows)
    thr = int(rng.integers(100, 400))
    date[:3] = [thr, thr + 1, thr - 1]
    cols = [F.DeviceBuffer.from_array(enc1(k, v, b, s)) for v, b, s in
            ((date, date_bits, 8201), (rf, 1, 8202), (ls, 1, 8203))]
    dthr = F.DeviceBuffer.from_array(enc1(k, [thr], date_bits, 8204))
    one = F.DeviceBuffer.from_array(enc1(k, [1], 1, 8205))
    s = Q.tpch_q1_filter(ctx, cols[0], cols[1], cols[2], dthr, one, rows)
    ctx.sync()
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    ok = date <= thr
    for g, (a, b) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        want = ok & (rf == a) & (ls == b)
        assert np.array_equal(pred[g], want.astype(np.int64)), g
        assert int(np.sum(qty * pred[g])) == int(np.sum(qty[want]))
        assert int(np.sum(pred[g])) == int(np.sum(want))
    s.free()
