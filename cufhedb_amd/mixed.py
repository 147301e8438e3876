"""Mixed-parameter batch stream: lvl1 lookup-table bootstraps and lvl2
20-bit comparisons interleaved, with uneven batch sizes.

The reference serves such a stream with one host thread per CUDA stream and
one ciphertext per call (src/cuHEDB/gatebootstrap_gpu.cuh:17-25,
test/HE3_Query/TPCH_Q6.cu:179-193).  Here every batch is one batched call on
the context stream (a lvl1 batch is one fhe_pbs_lvl01_dev launch, a lvl2
batch one fhe_compare_dev chain), so the device sees few, large launches and
no host round trip inside a batch.  Each batch is timed with HIP events on
the stream; results are decrypted and checked after the timed call.

Plan (seeded): batch sizes clip(round(lognormal(ln 2048, sigma)), 256, 65536),
a fraction `frac_lvl1` of the batches are Z_8 LUT PBS (lvl0 -> lvl1), the rest
are x <= t range compares of 20-bit lvl2 values against one shared
encrypted threshold.
"""

import ctypes

import numpy as np

from . import (CMP_LE, LOGIC, LVL1_N, DeviceBuffer, check, lib, tlwe_encrypt_u32,
               tlwe_encrypt_u64, tlwe_phase_u32)


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


def run(ctx, sk, params, plan, seed=5, bits=20):
    """Runs the plan; returns per-batch records and a summary."""
    L = lib()
    rng = np.random.default_rng(seed)
    records = []
    for bi, (kind, B) in enumerate(plan):
        if kind == "pbs1":
            m = rng.integers(0, 8, B)
            lut = rng.integers(0, 8, 8)
            c0 = tlwe_encrypt_u32(10_000 + bi, 0, sk.s0, (m << 28).astype(np.uint32),
                                  params.alpha0)
            d_in = DeviceBuffer.from_array(c0)
            d_tv = DeviceBuffer.from_array(z8_testvector(lut))
            d_out = DeviceBuffer(B * (LVL1_N + 1) * 4)
            call = (lambda: check(L.fhe_pbs_lvl01_dev(ctx.h, d_in.ptr, d_tv.ptr, 0, d_out.ptr,
                                                       B, None)))
            bufs = [d_in, d_tv, d_out]
        else:
            x = rng.integers(0, 1 << bits, B)
            t = int(rng.integers(0, 1 << bits))
            sc = np.uint64(64 - bits - 1)
            d_in = DeviceBuffer.from_array(tlwe_encrypt_u64(20_000 + bi, 0, sk.s2,
                                                            x.astype(np.uint64) << sc,
                                                            params.alpha2))
            d_t = DeviceBuffer.from_array(tlwe_encrypt_u64(30_000 + bi, 0, sk.s2,
                                                           np.array([t], np.uint64) << sc,
                                                           params.alpha2))
            d_out = DeviceBuffer(B * (LVL1_N + 1) * 4)
            call = (lambda: check(L.fhe_compare_dev(ctx.h, 2, CMP_LE, d_in.ptr, d_t.ptr, 1,
                                                     d_out.ptr, bits, LOGIC, B, None)))
            bufs = [d_in, d_t, d_out]
        ctx.sync()
        check(L.fhe_timer_start(ctx.h, None))
        call()
        ms = ctypes.c_float()
        check(L.fhe_timer_stop(ctx.h, None, ctypes.byref(ms)))
        out = d_out.to_array(np.uint32, (B, LVL1_N + 1))
        ph = tlwe_phase_u32(out, sk.s1).view(np.int32).astype(np.int64)
        if kind == "pbs1":
            errors = int(np.sum(np.round(ph / (1 << 28)).astype(np.int64) % 16 != lut[m]))
            pbs = B
        else:
            errors = int(np.sum((ph > 0) != (x <= t)))
            pbs = 9 * B  # 6 lvl2 + 3 lvl1 bootstraps per row (21-bit HomMSB chain)
        records.append({"kind": kind, "size": B, "ms": float(ms.value), "errors": errors,
                        "bootstraps": pbs})
        for b in bufs:
            b.free()
    return records, summarize(records)


def summarize(records):
    lat = np.array([r["ms"] for r in records])
    total_ms = float(lat.sum())
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
        ms = sum(r["ms"] for r in rs)
        rows = sum(r["size"] for r in rs)
        out[kind] = {"batches": len(rs), "rows": rows, "device_ms": ms,
                     "rows_per_s": rows / (ms / 1e3),
                     "p50_batch_ms": float(np.percentile([r["ms"] for r in rs], 50)),
                     "p99_batch_ms": float(np.percentile([r["ms"] for r in rs], 99))}
    out["bootstraps_per_s"] = sum(r["bootstraps"] for r in records) / (total_ms / 1e3)
    return out


__all__ = ["make_plan", "run", "summarize", "z8_testvector"]
