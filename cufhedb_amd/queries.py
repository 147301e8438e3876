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


def tpch_q1_filter(ctx, shipdate, returnflag, linestatus, date_thr, one, rows, date_bits=9,
                   bin_bits=1, scratch=None, rescale_bits=29):
    """Group predicates of TPC-H Q1 on lvl1 columns (device buffers):
    out[0..3] = [shipdate <= date] AND (returnflag, linestatus) ==
    (0,0), (0,1), (1,0), (1,1), each log_rescaled to 2^rescale_bits.
    date_thr and one are single lvl1 TLWEs (the date constant and an
    encryption of 1 at the bin scale)."""
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
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, o[0], p5, o[1], LOGIC, rows, None))   # (0, 1)
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, o[0], p3, o[0], LOGIC, rows, None))   # (0, 0)
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, p1, p4, o[2], LOGIC, rows, None))
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, o[2], p5, o[3], LOGIC, rows, None))   # (1, 1)
    check(L.fhe_gate_lvl1_dev(h, GATE_AND, o[2], p3, o[2], LOGIC, rows, None))   # (1, 0)
    # TPCH_Q1.cu:277-283
    for k in range(4):
        check(L.fhe_log_rescale_lvl1_dev(h, o[k], o[k], rescale_bits, rows, None))
    return s


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


__all__ = ["range_filter", "tpch_q1_filter", "Q1Scratch", "decrypt_predicates"]
