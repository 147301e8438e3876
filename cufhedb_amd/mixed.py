"""Mixed-parameter batch stream: lvl1 lookup-table bootstraps and lvl2
20-bit comparisons interleaved, with uneven batch sizes.

The reference serves such a stream with one host thread per CUDA stream and
one ciphertext per call (src/cuHEDB/gatebootstrap_gpu.cuh:17-25,
test/HE3_Query/TPCH_Q6.cu:179-193).  Two ways are measured here:

* ``run``: every batch is one batched call on the context stream (a lvl1
  batch is one fhe_pbs_lvl01_dev launch, a lvl2 batch one fhe_compare_dev
  chain), timed with HIP events on its own;
* ``run_coalesced``: the whole stream is submitted to the scheduler
  (fhe_sched_*): every lvl1 batch runs in one persistent launch over a job
  table (the waves take the next ciphertext of any batch, the key stays
  resident, each batch's completion is stamped on the device clock), the
  comparisons are merged into launches of up to ``max_rows`` rows on a
  second stream; a batch's latency is the time from the start of the flush
  to its completion.

Results are decrypted and checked after the timed region in both.

Plan (seeded): batch sizes clip(round(lognormal(ln 2048, sigma)), 256, 65536),
a fraction `frac_lvl1` of the batches are Z_8 LUT PBS (lvl0 -> lvl1), the rest
are x <= t range compares of 20-bit lvl2 values: ``compare="chain"`` against
one encrypted threshold per batch (fhe_compare_dev, the reference's HomMSB
chain), ``compare="digits"`` against one public threshold per batch with the
values as five 4-bit lvl2 digits (fhe_compare_digits_dev; one call per
batch, ``run`` only).
"""

import ctypes

import numpy as np

from . import (CMP_LE, LOGIC, LVL1_N, DeviceBuffer, Scheduler, check, encrypt_digits, lib,
               tlwe_encrypt_u32, tlwe_encrypt_u64, tlwe_phase_u32)

DIGITS, DIGIT_BITS = 5, 4


def z8_testvector(lut, N=LVL1_N):
    j = np.arange(N)
    box = (j + 64) >> 7
    vals = np.array(lut, np.int64)[box % 8] << 28
    vals = np.where(box == 8, -vals, vals)
    return (vals % (1 << 32)).astype(np.uint32)


def make_plan(seed, nbatches=200, sigma=1.0, frac_lvl1=0.7, max_size=65536):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(2048), sigma, nbatches)), 256, max_size)
    kinds = np.where(rng.random(nbatches) < frac_lvl1, "pbs1", "cmp2")
    return [(str(k), int(s)) for k, s in zip(kinds, sizes)]


class _Batch:
    """Device inputs / outputs and the plaintext expectation of one batch."""

    def __init__(self, bi, kind, B, sk, params, rng, bits, compare="chain"):
        self.kind, self.B, self.compare = kind, B, compare
        if kind == "pbs1":
            m = rng.integers(0, 8, B)
            lut = rng.integers(0, 8, 8)
            c0 = tlwe_encrypt_u32(10_000 + bi, 0, sk.s0, (m << 28).astype(np.uint32),
                                  params.alpha0)
            self.inp = DeviceBuffer.from_array(c0)
            self.aux = DeviceBuffer.from_array(z8_testvector(lut))
            self.expect = lut[m]
        elif compare == "digits":
            assert bits <= DIGITS * DIGIT_BITS
            x = rng.integers(0, 1 << bits, B)
            self.threshold = int(rng.integers(0, 1 << bits))
            self.inp = DeviceBuffer.from_array(encrypt_digits(20_000 + bi, 0, sk.s2, x, DIGITS,
                                                              DIGIT_BITS, params.alpha2))
            self.aux = None
            self.expect = x <= self.threshold
        else:
            x = rng.integers(0, 1 << bits, B)
            t = int(rng.integers(0, 1 << bits))
            sc = np.uint64(64 - bits - 1)
            self.inp = DeviceBuffer.from_array(tlwe_encrypt_u64(20_000 + bi, 0, sk.s2,
                                                                x.astype(np.uint64) << sc,
                                                                params.alpha2))
            self.aux = DeviceBuffer.from_array(tlwe_encrypt_u64(30_000 + bi, 0, sk.s2,
                                                                np.array([t], np.uint64) << sc,
                                                                params.alpha2))
            self.expect = x <= t
        self.out = DeviceBuffer(B * (LVL1_N + 1) * 4)

    def result(self):
        return self.out.to_array(np.uint32, (self.B, LVL1_N + 1))

    def errors(self, sk):
        ph = tlwe_phase_u32(self.result(), sk.s1).view(np.int32).astype(np.int64)
        if self.kind == "pbs1":
            return int(np.sum(np.round(ph / (1 << 28)).astype(np.int64) % 16 != self.expect))
        return int(np.sum((ph > 0) != self.expect))

    def phase_errors(self, sk):
        """Signed distance (torus units of 2^-32) of every output phase from
        the exact encoding of the expected result: LUT value * 2^28 for the
        lvl1 PBS (decision margin 2^27), +-2^29 for the comparisons (margin
        2^29)."""
        ph = tlwe_phase_u32(self.result(), sk.s1).astype(np.int64)
        if self.kind == "pbs1":
            want = self.expect.astype(np.int64) << 28
        else:
            want = np.where(self.expect, 1 << 29, -(1 << 29))
        return (ph - want + (1 << 31)) % (1 << 32) - (1 << 31)

    def record(self, sk, ms):
        e = self.phase_errors(sk).astype(np.float64)
        return {"kind": self.kind, "size": self.B, "ms": float(ms), "errors": self.errors(sk),
                "bootstraps": self.bootstraps, "err_max": float(np.max(np.abs(e))),
                "err_ss": float(np.sum(e * e))}

    @property
    def bootstraps(self):
        # a 21-bit lvl2 HomMSB chain is 6 lvl2 + 3 lvl1 bootstraps per row,
        # the digit-wise comparison 6 lvl2 (one per digit, one combining)
        if self.kind == "pbs1":
            return self.B
        return (DIGITS + 1) * self.B if self.compare == "digits" else 9 * self.B

    def free(self):
        for b in (self.inp, self.aux, self.out):
            if b is not None:
                b.free()


def run(ctx, sk, params, plan, seed=5, bits=20, compare="chain"):
    """Each batch one call, timed alone; returns per-batch records and a summary."""
    L = lib()
    rng = np.random.default_rng(seed)
    records = []
    for bi, (kind, B) in enumerate(plan):
        b = _Batch(bi, kind, B, sk, params, rng, bits, compare)
        ctx.sync()
        check(L.fhe_timer_start(ctx.h, None))
        if kind == "pbs1":
            check(L.fhe_pbs_lvl01_dev(ctx.h, b.inp.ptr, b.aux.ptr, 0, b.out.ptr, B, None))
        elif compare == "digits":
            check(L.fhe_compare_digits_dev(ctx.h, CMP_LE, b.inp.ptr, DIGITS, DIGIT_BITS,
                                           b.threshold, 1, b.out.ptr, LOGIC, B, None))
        else:
            check(L.fhe_compare_dev(ctx.h, 2, CMP_LE, b.inp.ptr, b.aux.ptr, 1, b.out.ptr, bits,
                                    LOGIC, B, None))
        ms = ctypes.c_float()
        check(L.fhe_timer_stop(ctx.h, None, ctypes.byref(ms)))
        records.append(b.record(sk, ms.value))
        b.free()
    return records, summarize(records)


def run_coalesced(ctx, sk, params, plan, seed=5, bits=20, max_rows=16384):
    """The whole plan through the coalescing scheduler in one flush (inputs
    prepared on the device first); returns per-batch records and a summary
    whose device_ms is the flush's span."""
    rng = np.random.default_rng(seed)
    batches = [_Batch(bi, kind, B, sk, params, rng, bits) for bi, (kind, B) in enumerate(plan)]
    sched = Scheduler(ctx, max_rows)
    ctx.sync()
    tickets = []
    for b in batches:
        if b.kind == "pbs1":
            tickets.append(sched.submit_pbs1(b.inp, b.aux, b.out, b.B))
        else:
            tickets.append(sched.submit_cmp2(CMP_LE, b.inp, b.aux, b.out, bits, LOGIC, b.B))
    sched.flush()
    lat = [sched.wait(t) for t in tickets]
    sched.close()
    records = [b.record(sk, ms) for b, ms in zip(batches, lat)]
    for b in batches:
        b.free()
    out = summarize(records, span_ms=max(lat) if lat else 0.0)
    out["max_rows"] = max_rows
    return records, out


def summarize(records, span_ms=None):
    """Per-kind rows/s and batch latency percentiles.  span_ms: wall span of
    a coalesced run (batches overlap in time); None = batches ran one after
    another, so the span is the sum of their times."""
    lat = np.array([r["ms"] for r in records])
    total_ms = float(lat.sum()) if span_ms is None else float(span_ms)
    out = {
        "batches": len(records),
        "device_ms": total_ms,
        "errors": int(sum(r["errors"] for r in records)),
        "p50_batch_ms": float(np.percentile(lat, 50)),
        "p99_batch_ms": float(np.percentile(lat, 99)),
    }
    for kind in ("pbs1", "cmp2"):
        rs = [r for r in records if r["kind"] == kind]
        if not rs:
            continue
        rows = sum(r["size"] for r in rs)
        entry = {"batches": len(rs), "rows": rows,
                 "p50_batch_ms": float(np.percentile([r["ms"] for r in rs], 50)),
                 "p99_batch_ms": float(np.percentile([r["ms"] for r in rs], 99))}
        if all("err_max" in r for r in rs):
            # output phase error against the exact encoding, over every row
            entry["phase_err_max_log2"] = float(np.log2(max(r["err_max"] for r in rs) + 1))
            entry["phase_err_std_log2"] = float(
                0.5 * np.log2(sum(r["err_ss"] for r in rs) / rows + 1))
            entry["decision_margin_log2"] = 27.0 if kind == "pbs1" else 29.0
        if span_ms is None:
            ms = sum(r["ms"] for r in rs)
            entry["device_ms"] = ms
            entry["rows_per_s"] = rows / (ms / 1e3)
        out[kind] = entry
    out["bootstraps_per_s"] = sum(r["bootstraps"] for r in records) / (total_ms / 1e3)
    out["rows_per_s"] = sum(r["size"] for r in records) / (total_ms / 1e3)
    return out


__all__ = ["make_plan", "run", "run_coalesced", "summarize", "z8_testvector"]
