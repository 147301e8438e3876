"""GPU: the mixed-parameter batch stream (lvl1 Z_8 LUT PBS and lvl2 20-bit
range compares, uneven batch sizes), every output checked against the
plaintext; a small seeded plan of the full workload's shape."""

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import mixed

@pytest.mark.gpu
def test_mixed_stream_small_plan(full672):
    k, ctx = full672
    plan = mixed.make_plan(seed=555, nbatches=10, sigma=0.8, frac_lvl1=0.6, max_size=2048)
    assert {kind for kind, _ in plan} <= {"pbs1", "cmp2"}
    records, summary = mixed.run(ctx, k.sk, k.params, plan, seed=556)
    print("mixed stream:", summary)
    assert summary["errors"] == 0
    assert len(records) == len(plan)
    assert all(r["ms"] > 0 for r in records)
    for kind in ("pbs1", "cmp2"):
        if kind in summary:
            e = summary[kind]
            # every output phase within its decision margin, with room to spare
            assert e["phase_err_max_log2"] < e["decision_margin_log2"] - 1, (kind, e)
            assert e["phase_err_std_log2"] < e["phase_err_max_log2"]


@pytest.mark.gpu
def test_mixed_stream_digit_compares_n636(full636):
    """The same stream at n0 = 636 with the 20-bit comparisons digit-wise
    against public thresholds (fhe_compare_digits_dev): every row right."""
    k, ctx = full636
    plan = mixed.make_plan(seed=559, nbatches=8, sigma=0.8, frac_lvl1=0.5, max_size=1024)
    records, summary = mixed.run(ctx, k.sk, k.params, plan, seed=560, compare="digits")
    print("mixed stream, digit-wise compares:", summary)
    assert summary["errors"] == 0
    cmp = [r for r in records if r["kind"] == "cmp2"]
    assert cmp and all(r["bootstraps"] == 6 * r["size"] for r in cmp)


def test_mixed_plan_shape():
    plan = mixed.make_plan(seed=1, nbatches=200)
    sizes = np.array([s for _, s in plan])
    assert sizes.min() >= 256 and sizes.max() <= 65536
    frac = np.mean([k == "pbs1" for k, _ in plan])
    assert 0.55 < frac < 0.85


@pytest.mark.gpu
def test_scheduler_matches_direct_calls(full672):
    """Requests served by the coalescing scheduler (fhe_sched_*) give the
    same ciphertexts, bit for bit, as one direct call per request: each
    ciphertext's bootstrap is independent of its batch.  lvl1 requests all go
    through one persistent launch over a job table (the one-wave blind
    rotation, so the direct calls are made with one wave per ciphertext);
    comparisons are merged, split at a change of comparison, served alone
    above max_rows.  Single-row requests included."""
    k, ctx = full672
    rng = np.random.default_rng(77)
    spec = [("pbs1", 37, None), ("pbs1", 500, None), ("cmp2", 20, F.CMP_LE),
            ("cmp2", 300, F.CMP_LE), ("cmp2", 40, F.CMP_GT), ("pbs1", 1, None),
            ("pbs1", 700, None), ("pbs1", 1500, None), ("cmp2", 1100, F.CMP_LE),
            ("pbs1", 64, None)]
    batches = [mixed._Batch(i, kind, B, k.sk, k.params, rng, 20) for i, (kind, B, _) in
               enumerate(spec)]
    sched = F.Scheduler(ctx, max_rows=1024)
    tickets = []
    for b, (_, _, op) in zip(batches, spec):
        if b.kind == "pbs1":
            tickets.append(sched.submit_pbs1(b.inp, b.aux, b.out, b.B))
        else:
            tickets.append(sched.submit_cmp2(op, b.inp, b.aux, b.out, 20, F.LOGIC, b.B))
    sched.flush()
    # the context stream continues after both kinds: a negation enqueued on
    # it right after the flush (HomNOT, no bootstrap) reads the last
    # comparison's finished result
    last = batches[8]
    snap = F.DeviceBuffer(last.B * (F.LVL1_N + 1) * 4)
    F.check(F.lib().fhe_hom_not_dev(ctx.h, 1, last.out.ptr, snap.ptr, last.B, None))
    lat = [sched.wait(t) for t in tickets]
    assert all(ms > 0 for ms in lat)
    ctx.sync()
    neg = (np.uint64(1 << 32) - last.result().astype(np.uint64)) % np.uint64(1 << 32)
    assert np.array_equal(snap.to_array(np.uint32, (last.B, F.LVL1_N + 1)),
                          neg.astype(np.uint32))
    snap.free()
    with pytest.raises(F.FheError):
        sched.wait(tickets[0])  # already waited for
    sched.close()
    L = F.lib()
    for b, (_, _, op) in zip(batches, spec):
        got = b.result()
        direct = F.DeviceBuffer(b.B * (F.LVL1_N + 1) * 4)
        if b.kind == "pbs1":
            try:
                ctx.set_tune_table([(1, 1)], knob=F.TUNE_PBS1_WAVES)
                F.check(L.fhe_pbs_lvl01_dev(ctx.h, b.inp.ptr, b.aux.ptr, 0, direct.ptr, b.B,
                                            None))
            finally:
                ctx.set_tune_table([], knob=F.TUNE_PBS1_WAVES)
        else:
            F.check(L.fhe_compare_dev(ctx.h, 2, op, b.inp.ptr, b.aux.ptr, 1, direct.ptr, 20,
                                      F.LOGIC, b.B, None))
        ctx.sync()
        want = direct.to_array(np.uint32, (b.B, F.LVL1_N + 1))
        direct.free()
        assert np.array_equal(got, want), (b.kind, b.B, op)
        if op != F.CMP_GT:
            assert b.errors(k.sk) == 0
        b.free()


@pytest.mark.gpu
def test_mixed_stream_coalesced(full672):
    k, ctx = full672
    plan = mixed.make_plan(seed=557, nbatches=12, sigma=0.8, frac_lvl1=0.6, max_size=2048)
    records, summary = mixed.run_coalesced(ctx, k.sk, k.params, plan, seed=558, max_rows=4096)
    print("coalesced mixed stream:", summary)
    assert summary["errors"] == 0 and len(records) == len(plan)


@pytest.mark.gpu
def test_scheduler_failed_flush_fails_its_requests(keys636):
    """A flush that stops on an error (here: no bootstrapping key loaded)
    returns it, and waiting on its requests reports the failure instead of
    returning incomplete results."""
    ctx = F.Context(keys636.params)  # no keys
    sched = F.Scheduler(ctx, max_rows=64)
    n = 8
    inp = F.DeviceBuffer(n * (keys636.params.n0 + 1) * 4)
    tv = F.DeviceBuffer(F.LVL1_N * 4)
    out = F.DeviceBuffer(n * (F.LVL1_N + 1) * 4)
    F.check(F.lib().fhe_memset_dev(inp.ptr, 0, inp.nbytes))
    F.check(F.lib().fhe_memset_dev(tv.ptr, 0, tv.nbytes))
    t = sched.submit_pbs1(inp, tv, out, n)
    with pytest.raises(F.FheError):
        sched.flush()
    with pytest.raises(F.FheError):
        sched.wait(t)
    sched.close()
    for b in (inp, tv, out):
        b.free()
    ctx.close()


def test_summary_error_bounds():
    """summarize: per-kind maximum and RMS phase error (log2) over every row
    of every batch, next to the decision margin."""
    recs = [{"kind": "pbs1", "size": 4, "ms": 1.0, "errors": 0, "bootstraps": 4,
             "err_max": 2.0 ** 20, "err_ss": 4 * 2.0 ** 36},
            {"kind": "pbs1", "size": 12, "ms": 3.0, "errors": 0, "bootstraps": 12,
             "err_max": 2.0 ** 22, "err_ss": 12 * 2.0 ** 36},
            {"kind": "cmp2", "size": 8, "ms": 2.0, "errors": 0, "bootstraps": 72,
             "err_max": 2.0 ** 24 - 1, "err_ss": 8 * 2.0 ** 40}]
    out = mixed.summarize(recs)
    assert abs(out["pbs1"]["phase_err_max_log2"] - 22) < 1e-6
    assert abs(out["pbs1"]["phase_err_std_log2"] - 18) < 1e-6
    assert out["pbs1"]["decision_margin_log2"] == 27.0
    assert abs(out["cmp2"]["phase_err_max_log2"] - 24) < 1e-6
    assert abs(out["cmp2"]["phase_err_std_log2"] - 20) < 1e-6
    assert out["cmp2"]["decision_margin_log2"] == 29.0
    assert out["errors"] == 0 and out["batches"] == 3


@pytest.mark.gpu
def test_mixed_stream_full_size_coalesced(full672):
    """The mixed-parameter workload at its own size: 200 batches, sizes
    lognormal around 2048 clipped to [256, 65536], 70 % lvl1 Z_8 LUT
    bootstraps and 30 % lvl2 20-bit comparisons (HomMSB chain against one
    encrypted threshold per batch), served by the coalescing scheduler in one
    flush.  Every output decrypts to its plaintext result, and every output's
    phase error stays more than a factor 2 inside its decision margin."""
    k, ctx = full672
    plan = mixed.make_plan(2026, 200)
    sizes = [s for _, s in plan]
    print(f"plan: {len(plan)} batches, {sum(sizes)} rows, largest {max(sizes)}", flush=True)
    records, summary = mixed.run_coalesced(ctx, k.sk, k.params, plan)
    print("mixed stream, coalesced:", summary)
    assert len(records) == 200
    assert summary["errors"] == 0
    for kind in ("pbs1", "cmp2"):
        e = summary[kind]
        assert e["phase_err_max_log2"] < e["decision_margin_log2"] - 1, (kind, e)
