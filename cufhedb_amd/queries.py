"""Encrypted query pipelines on the batched operator layer.

Every step is one batched call into the HIP library on device-resident
ciphertexts (no host round trip between steps); the ciphertexts of a column
are contiguous TLWE records.

* ``range_filter``: ``x <= t`` for every row of a column of ``bits``-bit
  values against one shared encrypted threshold (cuHEDB::less_than_equal,
  src/cuHEDB/HomCompare_gpu.cuh:52-61, as driven per row by
  test/HE3_Query/TPCH_Q6.cu).
* ``tpch_q1_filter``: the four (returnflag, linestatus) group predicates of
  TPC-H Q1 under ``shipdate <= date`` (test/HE3_Query/TPCH_Q1.cu:215-295,
  Filter_Cipher_d), followed by log_rescale to scale 2^29 so each predicate
  decrypts to 0 / 1 (TPCH_Q1.cu:277-283).
* ``groupby_filter``: the same pipeline shape for a wider table: a 12-bit
  date column against a public cut-off, compared digit-wise on the 64-bit
  torus (``fhe_compare_digits_dev``), and six (cat1 in {0,1,2}, cat2 in
  {0,1}) group predicates.  ``group_aggregates`` then sums the plaintext
  measure columns under the decrypted predicates, as the reference's
  ``Query`` does after ``Filter_Cipher_d`` (TPCH_Q1.cu:176-205); the
  reference's other aggregation route, repacking the LWE predicates into a
  SEAL CKKS ciphertext (TPCH_Q1.cu:374-450), is not part of this library.
"""

import numpy as np

from . import CMP_LE, CMP_LT, GATE_AND, LOGIC, LVL1_N, LVL2_N, DeviceBuffer, check, lib

W1 = LVL1_N + 1
W2 = LVL2_N + 1


def _buf(rows, level=1):
    return DeviceBuffer(rows * (W1 * 4 if level == 1 else W2 * 8))


def range_filter(ctx, x_dev, t_dev, rows, bits, level=2, out_dev=None, result_type=LOGIC):
    """x <= t for `rows` TLWEs at x_dev against the single TLWE at t_dev.
    Returns the lvl1 result buffer (rows x 1025 uint32)."""
    out = out_dev or _buf(rows)
    check(lib().fhe_compare_dev(ctx.h, level, CMP_LE, x_dev.ptr, t_dev.ptr, 1, out.ptr, bits,
                                result_type, rows, None))
    return out


class Q1Scratch:
    """Device buffers of one tpch_q1_filter call (reusable across calls)."""

    def __init__(self, rows):
        self.rows = rows
        self.p = [_buf(rows) for _ in range(5)]
        self.out = [_buf(rows) for _ in range(4)]

    def free(self):
        for b in self.p + self.out:
            b.free()


def _last_and(L, h, x, y, out, rows, rescale_bits, fuse):
    """A group predicate's final AND, fused with its log_rescale (one
    bootstrap, fhe_gate_rescale_lvl1_dev) or, with fuse False, the LOGIC AND
    alone (the caller rescales afterwards, as the reference's driver does)."""
    if fuse:
        check(L.fhe_gate_rescale_lvl1_dev(h, GATE_AND, x, y, out, rescale_bits, rows, None))
    else:
        check(L.fhe_gate_lvl1_dev(h, GATE_AND, x, y, out, LOGIC, rows, None))


def tpch_q1_filter(ctx, shipdate, returnflag, linestatus, date_thr, one, rows, date_bits=9,
                   bin_bits=1, scratch=None, rescale_bits=29, fuse_rescale=True):
    """Group predicates of TPC-H Q1 on lvl1 columns (device buffers):
    out[0..3] = [shipdate <= date] AND (returnflag, linestatus) ==
    (0,0), (0,1), (1,0), (1,1), each log_rescaled to 2^rescale_bits.
    date_thr and one are single lvl1 TLWEs (the date constant and an
    encryption of 1 at the bin scale).  fuse_rescale: each group's last AND
    and its log_rescale as one bootstrap (False: two, the reference's
    sequence)."""
    s = scratch or Q1Scratch(rows)
    L = lib()
    h = ctx.h
    p1, p2, p3, p4, p5 = (b.ptr for b in s.p)
    o = [b.ptr for b in s.out]
    # TPCH_Q1.cu:260-264
    check(L.fhe_compare_dev(h, 1, CMP_LE, shipdate.ptr, date_thr.ptr, 1, p1, date_bits, LOGIC,
                            rows, None))
    check(L.fhe_compare_dev(h, 1, CMP_LT, returnflag.ptr, one.ptr, 1, p2, bin_bits, LOGIC, rows,
                            None))
    check(L.fhe_compare_dev(h, 1, CMP_LT, linestatus.ptr, one.ptr, 1, p3, bin_bits, LOGIC, rows,
                            None))
    check(L.fhe_hom_not_dev(h, 1, p2, p4, rows, None))
    check(L.fhe_hom_not_dev(h, 1, p3, p5, rows, None))
    # TPCH_Q1.cu:266-272
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, p1, p2, o[0], LOGIC, rows, None))
    _last_and(L, h, o[0], p5, o[1], rows, rescale_bits, fuse_rescale)   # (0, 1)
    _last_and(L, h, o[0], p3, o[0], rows, rescale_bits, fuse_rescale)   # (0, 0)
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, p1, p4, o[2], LOGIC, rows, None))
    _last_and(L, h, o[2], p5, o[3], rows, rescale_bits, fuse_rescale)   # (1, 1)
    _last_and(L, h, o[2], p3, o[2], rows, rescale_bits, fuse_rescale)   # (1, 0)
    if not fuse_rescale:
        # TPCH_Q1.cu:277-283
        for k in range(4):
            check(L.fhe_log_rescale_lvl1_dev(h, o[k], o[k], rescale_bits, rows, None))
    return s


#: (cat1, cat2) of groupby_filter's outputs, in order
GROUPS = tuple((a, b) for a in range(3) for b in range(2))


class GroupByScratch:
    """Device buffers of one groupby_filter call (reusable across calls)."""

    def __init__(self, rows):
        self.rows = rows
        self.p = [_buf(rows) for _ in range(11)]
        self.out = [_buf(rows) for _ in range(len(GROUPS))]

    def free(self):
        for b in self.p + self.out:
            b.free()


def groupby_filter(ctx, date_digits, cat1, cat2, one2, two2, one1, rows, date_cutoff,
                   digits=3, digit_bits=4, scratch=None, rescale_bits=29, fuse_rescale=True):
    """Group predicates [date <= date_cutoff] AND cat1 == a AND cat2 == b for
    the six (a, b) of GROUPS, each log_rescaled to 2^rescale_bits.

    date_digits: rows x digits lvl2 TLWEs (radix 2^digit_bits, least
    significant digit first, as encrypt_digits writes them); the cut-off is a
    public constant of the query.  cat1 (values 0..2, 2-bit scale 2^29) and
    cat2 (0..1, 1-bit scale 2^30) are lvl1 TLWE columns; one2 / two2 are
    single lvl1 encryptions of 1 and 2 at cat1's scale, one1 of 1 at cat2's.
    Category equality comes from the strict comparisons of the reference's
    filters (cuHEDB::less_than, HomCompare_gpu.cuh:40-49): cat1 == 0 is
    cat1 < 1, cat1 == 1 is (cat1 < 2) AND NOT (cat1 < 1), cat1 == 2 is
    NOT (cat1 < 2); the ANDs are cuHEDB::HomAND (gatebootstrap_gpu.cuh:123-137).
    fuse_rescale: each group's last AND and its log_rescale as one bootstrap.
    """
    s = scratch or GroupByScratch(rows)
    L = lib()
    h = ctx.h
    d, lt1, lt2, c0, nlt1, nlt2, c1, eq1, da0, da1, da2 = (b.ptr for b in s.p)
    o = [b.ptr for b in s.out]
    check(L.fhe_compare_digits_dev(h, CMP_LE, date_digits.ptr, digits, digit_bits,
                                   int(date_cutoff), 1, d, LOGIC, rows, None))
    check(L.fhe_compare_dev(h, 1, CMP_LT, cat1.ptr, one2.ptr, 1, lt1, 2, LOGIC, rows, None))
    check(L.fhe_compare_dev(h, 1, CMP_LT, cat1.ptr, two2.ptr, 1, lt2, 2, LOGIC, rows, None))
    check(L.fhe_compare_dev(h, 1, CMP_LT, cat2.ptr, one1.ptr, 1, c0, 1, LOGIC, rows, None))
    check(L.fhe_hom_not_dev(h, 1, lt1, nlt1, rows, None))
    check(L.fhe_hom_not_dev(h, 1, lt2, nlt2, rows, None))
    check(L.fhe_hom_not_dev(h, 1, c0, c1, rows, None))
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, lt2, nlt1, eq1, LOGIC, rows, None))
    for da, cat in ((da0, lt1), (da1, eq1), (da2, nlt2)):
        check(L.fhe_gate_lvl1_dev(h, GATE_AND, d, cat, da, LOGIC, rows, None))
    for g, (a, b) in enumerate(GROUPS):
        _last_and(L, h, (da0, da1, da2)[a], (c0, c1)[b], o[g], rows, rescale_bits, fuse_rescale)
    if not fuse_rescale:
        for g in range(len(GROUPS)):
            check(L.fhe_log_rescale_lvl1_dev(h, o[g], o[g], rescale_bits, rows, None))
    return s


def group_aggregates(pred, quantity, price, discount):
    """SUM(quantity), SUM(price * (100 - discount)) and COUNT(*) per group
    over the rows each decrypted 0/1 predicate selects (the client-side
    aggregation of TPCH_Q1.cu:176-205, widened to the revenue measure).
    Returns a len(pred) x 3 int64 array."""
    q = np.asarray(quantity, np.int64)
    rev = np.asarray(price, np.int64) * (100 - np.asarray(discount, np.int64))
    out = np.zeros((len(pred), 3), np.int64)
    for g, p in enumerate(pred):
        p = np.asarray(p, np.int64)
        out[g] = (int(np.dot(p, q)), int(np.dot(p, rev)), int(p.sum()))
    return out


def decrypt_predicates(bufs, rows, s1, scale_bits=29):
    """tlweSymInt32Decrypt at scale 2^scale_bits of each lvl1 buffer -> 0/1 arrays."""
    from . import tlwe_phase_u32
    res = []
    for b in bufs:
        c = b.to_array(np.uint32, (rows, W1))
        ph = tlwe_phase_u32(c, s1).astype(np.uint64)
        res.append(((ph + (np.uint64(1) << np.uint64(scale_bits - 1))) >> np.uint64(scale_bits))
                   .astype(np.int64) % (1 << (32 - scale_bits)))
    return res


__all__ = ["range_filter", "tpch_q1_filter", "Q1Scratch", "groupby_filter", "GroupByScratch",
           "GROUPS", "group_aggregates", "decrypt_predicates"]
