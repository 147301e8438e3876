"""GPU: decryption-failure count over 10^6 lvl1 programmable bootstraps.

Both lvl1 variants run it: the FP64 path and the split-MFMA path (fp16
matrix-core forward transforms, key in three integer limbs).

Workload: the batched Z_8 lookup-table PBS of the benchmark (config preset
n0 = 636, alpha0 = 2^-15, batch 8192), 123 batches = 1,007,616 PBS with fresh
plaintexts, encryption seeds and a fresh 8-entry LUT per batch.  Each output
is decrypted (round(phase / 2^28) mod 16) and compared with LUT[m].  The
reference parameters target a failure probability far below 10^-6 per
bootstrap (TFHEpp/include/params/128bit.hpp), so the expected count is 0.
The test also reports the output phase-error distribution (phase minus the
exact LUT value), whose std matches the oracle's (test_gpu_noise.py).
"""

import numpy as np
import pytest

import cufhedb_amd as F

pytestmark = pytest.mark.gpu
N, DELTA8 = 1024, 1 << 28


def z8_testvector(lut):
    j = np.arange(N)
    box = (j + 64) >> 7
    vals = np.array(lut, np.int64)[box % 8] << 28
    vals = np.where(box == 8, -vals, vals)
    return (vals % (1 << 32)).astype(np.uint32)


@pytest.fixture(params=["fp64", "mfma_split"])
def variant_ctx(request, keys636, ctx636):
    if request.param == "fp64":
        yield ctx636
        return
    ctx = F.Context(keys636.params)
    ctx.load_bk(1, keys636.bk, F.BK_INT)
    ctx.set_pbs_variant(F.PBS_MFMA_SPLIT)
    yield ctx
    ctx.close()


def test_no_decryption_failures_in_1e6_pbs(keys636, variant_ctx):
    k, ctx = keys636, variant_ctx
    B, batches = 8192, 123
    rng = np.random.default_rng(1000003)
    d_in = F.DeviceBuffer(B * (k.n0 + 1) * 4)
    d_tv = F.DeviceBuffer(N * 4)
    d_out = F.DeviceBuffer(B * (N + 1) * 4)
    L = F.lib()
    failures, total = 0, 0
    err_max, err_sq = 0, 0.0
    edges = [0] + [1 << j for j in range(18, 28)] + [1 << 62]
    hist = np.zeros(len(edges) - 1, np.int64)
    for b in range(batches):
        m = rng.integers(0, 8, B)
        lut = rng.integers(0, 8, 8)
        c0 = F.tlwe_encrypt_u32(3_000_000 + b, 0, k.sk.s0, (m * DELTA8).astype(np.uint32),
                                k.params.alpha0)
        F.check(L.fhe_memcpy_h2d(d_in.ptr, F._ptr(c0), c0.nbytes))
        tv = z8_testvector(lut)
        F.check(L.fhe_memcpy_h2d(d_tv.ptr, F._ptr(tv), tv.nbytes))
        F.check(L.fhe_pbs_lvl01_dev(ctx.h, d_in.ptr, d_tv.ptr, 0, d_out.ptr, B, None))
        ctx.sync()
        out = d_out.to_array(np.uint32, (B, N + 1))
        ph = F.tlwe_phase_u32(out, k.sk.s1).view(np.int32).astype(np.int64)
        want = lut[m] * DELTA8
        err = (ph - want + (1 << 31)) % (1 << 32) - (1 << 31)
        dec = np.round(ph / DELTA8).astype(np.int64) % 16
        failures += int(np.sum(dec != lut[m]))
        total += B
        err_max = max(err_max, int(np.max(np.abs(err))))
        err_sq += float(np.sum(err.astype(np.float64) ** 2))
        hist += np.histogram(np.abs(err), bins=edges)[0]
    std = (err_sq / total) ** 0.5
    labels = ["<2^18"] + [f"2^{j}" for j in range(18, 27)] + [">=2^27"]
    print(f"failures {failures} of {total} PBS; phase error std 2^{np.log2(std):.2f}, "
          f"max |err| 2^{np.log2(max(err_max, 1)):.2f} (decision margin 2^27); |err| per "
          f"octave (2^-32 units): {dict(zip(labels, (int(c) for c in hist)))}")
    assert hist[-1] == failures
    assert total >= 1_000_000
    assert failures == 0
