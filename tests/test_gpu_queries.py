"""GPU: encrypted query pipelines on the batched operators, checked
bit-exactly against the plaintext query.

* range filter x <= t over a column of 20-bit lvl2 values with one shared
  encrypted threshold (cuHEDB::less_than_equal per row in the reference);
* TPC-H Q1 filter (test/HE3_Query/TPCH_Q1.cu:215-295): group predicates under
  shipdate <= date, log_rescale to 2^29, decrypt, then SUM(quantity) and
  COUNT per (returnflag, linestatus) group as in TPCH_Q1.cu:176-200;
* group-by filter over the wider synthetic table (12-bit dates against a
  public cut-off, compared digit-wise on lvl2; six (cat1, cat2) groups):
  predicates, then SUM(quantity), SUM(price * (100 - discount)) and COUNT
  per group, against the plaintext SQL.
"""

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import queries as Q

import _oracle as O

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


@pytest.mark.parametrize("fuse", [True, False])
def test_tpch_q1_filter_and_aggregate(full636, rng, fuse):
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
    s = Q.tpch_q1_filter(ctx, cols[0], cols[1], cols[2], dthr, one, rows, fuse_rescale=fuse)
    ctx.sync()
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    ok = date <= thr
    for g, (a, b) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        want = ok & (rf == a) & (ls == b)
        assert np.array_equal(pred[g], want.astype(np.int64)), g
        assert int(np.sum(qty * pred[g])) == int(np.sum(qty[want]))
        assert int(np.sum(pred[g])) == int(np.sum(want))
    s.free()


def synthetic_table(rng, rows):
    """The filter-aggregate table: quantity [1, 50], price [900, 105000],
    discount [0, 10], tax [0, 8], cat1 {0, 1, 2}, cat2 {0, 1}, date [0, 2556]."""
    return {"quantity": rng.integers(1, 51, rows), "price": rng.integers(900, 105001, rows),
            "discount": rng.integers(0, 11, rows), "tax": rng.integers(0, 9, rows),
            "cat1": rng.integers(0, 3, rows), "cat2": rng.integers(0, 2, rows),
            "date": rng.integers(0, 2557, rows)}


def plaintext_sql(t, cutoff):
    """SELECT cat1, cat2, SUM(quantity), SUM(price * (100 - discount)), COUNT(*)
    WHERE date <= cutoff GROUP BY cat1, cat2 -- in Python integers, row by row."""
    agg = {g: [0, 0, 0] for g in Q.GROUPS}
    for i in range(len(t["date"])):
        if int(t["date"][i]) <= cutoff:
            a = agg[(int(t["cat1"][i]), int(t["cat2"][i]))]
            a[0] += int(t["quantity"][i])
            a[1] += int(t["price"][i]) * (100 - int(t["discount"][i]))
            a[2] += 1
    return np.array([agg[g] for g in Q.GROUPS], np.int64)


def test_groupby_filter_and_aggregates(full636):
    k, ctx = full636
    rng = np.random.default_rng(8300)
    rows, cutoff = 600, 2400
    t = synthetic_table(rng, rows)
    t["date"][:6] = [cutoff, cutoff + 1, cutoff - 1, 0, 2556, 4095]
    dd = F.DeviceBuffer.from_array(F.encrypt_digits(8301, 0, k.sk.s2, t["date"], 3, 4,
                                                    k.params.alpha2))
    c1 = F.DeviceBuffer.from_array(enc1(k, t["cat1"], 2, 8302))
    c2 = F.DeviceBuffer.from_array(enc1(k, t["cat2"], 1, 8303))
    consts = [F.DeviceBuffer.from_array(enc1(k, [v], b, 8304 + i))
              for i, (v, b) in enumerate(((1, 2), (2, 2), (1, 1)))]
    s = Q.groupby_filter(ctx, dd, c1, c2, *consts, rows, cutoff)
    ctx.sync()
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    ok = t["date"] <= cutoff
    for g, (a, b) in enumerate(Q.GROUPS):
        want = ok & (t["cat1"] == a) & (t["cat2"] == b)
        assert np.array_equal(pred[g], want.astype(np.int64)), (a, b)
    got = Q.group_aggregates(pred, t["quantity"], t["price"], t["discount"])
    assert np.array_equal(got, plaintext_sql(t, cutoff))
    # every row lands in at most one group, and exactly one when date <= cutoff
    assert np.array_equal(np.sum(pred, axis=0), ok.astype(np.int64))
    s.free()
    for b in [dd, c1, c2] + consts:
        b.free()


def test_gate_rescale_matches_gate_then_rescale(full636):
    """fhe_gate_rescale_lvl1_dev decrypts as log_rescale(AND / OR) for every
    input pair, with fresh-bootstrap noise (one bootstrap)."""
    k, ctx = full636
    rows = 256
    rng = np.random.default_rng(8400)
    a, b = rng.integers(0, 2, rows), rng.integers(0, 2, rows)
    a[:4], b[:4] = [0, 0, 1, 1], [0, 1, 0, 1]
    mu = 1 << 29
    enc = [F.DeviceBuffer.from_array(F.tlwe_encrypt_u32(
        8401 + i, 0, k.sk.s1, np.where(v == 1, mu, (1 << 32) - mu).astype(np.uint32),
        k.params.alpha1)) for i, v in enumerate((a, b))]
    out, mid = F.DeviceBuffer(rows * 1025 * 4), F.DeviceBuffer(rows * 1025 * 4)
    L = F.lib()
    for gate, want in ((F.GATE_AND, a & b), (F.GATE_OR, a | b)):
        for sb in (29, 31, 1):
            F.check(L.fhe_gate_rescale_lvl1_dev(ctx.h, gate, enc[0].ptr, enc[1].ptr, out.ptr, sb,
                                                 rows, None))
            F.check(L.fhe_gate_lvl1_dev(ctx.h, gate, enc[0].ptr, enc[1].ptr, mid.ptr, F.LOGIC,
                                        rows, None))
            F.check(L.fhe_log_rescale_lvl1_dev(ctx.h, mid.ptr, mid.ptr, sb, rows, None))
            ctx.sync()
            ph = F.tlwe_phase_u32(out.to_array(np.uint32, (rows, 1025)), k.sk.s1)
            ph2 = F.tlwe_phase_u32(mid.to_array(np.uint32, (rows, 1025)), k.sk.s1)
            wv = (want.astype(np.uint64) << np.uint64(sb)) % (1 << 32)
            e = O.tdiff32(ph, wv.astype(np.uint32))
            assert np.max(np.abs(e)) < 1 << 26, (gate, sb)
            assert np.max(np.abs(O.tdiff32(ph2, wv.astype(np.uint32)))) < 1 << 26
    with pytest.raises(F.FheError):
        F.check(L.fhe_gate_rescale_lvl1_dev(ctx.h, F.GATE_NAND, enc[0].ptr, enc[1].ptr, out.ptr,
                                            29, rows, None))
    with pytest.raises(F.FheError):
        F.check(L.fhe_gate_rescale_lvl1_dev(ctx.h, F.GATE_AND, enc[0].ptr, enc[1].ptr, out.ptr,
                                            32, rows, None))
    for buf in enc + [out, mid]:
        buf.free()


def test_groupby_full_table_2e14_rows(full636):
    """The filter-aggregate query at its full size, 2^14 rows: every group
    predicate and every aggregate equal to the plaintext query's."""
    k, ctx = full636
    rng = np.random.default_rng(8500)
    rows, cutoff = 1 << 14, 2400
    t = synthetic_table(rng, rows)
    t["date"][:3] = [cutoff, cutoff + 1, 2556]
    dd = F.DeviceBuffer.from_array(F.encrypt_digits(8501, 0, k.sk.s2, t["date"], 3, 4,
                                                    k.params.alpha2))
    c1 = F.DeviceBuffer.from_array(enc1(k, t["cat1"], 2, 8502))
    c2 = F.DeviceBuffer.from_array(enc1(k, t["cat2"], 1, 8503))
    consts = [F.DeviceBuffer.from_array(enc1(k, [v], b, 8504 + i))
              for i, (v, b) in enumerate(((1, 2), (2, 2), (1, 1)))]
    s = Q.groupby_filter(ctx, dd, c1, c2, *consts, rows, cutoff)
    ctx.sync()
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    ok = t["date"] <= cutoff
    want = []
    for g, (a, b) in enumerate(Q.GROUPS):
        m = ok & (t["cat1"] == a) & (t["cat2"] == b)
        assert np.array_equal(pred[g], m.astype(np.int64)), (a, b)
        want.append((int(t["quantity"][m].sum()),
                     int((t["price"][m] * (100 - t["discount"][m])).sum()), int(m.sum())))
    got = Q.group_aggregates(pred, t["quantity"], t["price"], t["discount"])
    assert np.array_equal(got, np.array(want, np.int64))
    s.free()
    for b in [dd, c1, c2] + consts:
        b.free()
