"""HIP graph capture of query pipelines (fhe_graph_*).

A pipeline of _dev calls recorded from the context stream and replayed as
one graph must give, on every replay, exactly the ciphertexts the eager
calls give (the kernels are deterministic), also with other eager work
between the replays and after the input buffers were rewritten.  A capture
of a pipeline whose scratch is not yet allocated must fail cleanly and leave
the context usable.
"""

import numpy as np
import pytest

import cufhedb_amd as F
from cufhedb_amd import queries as Q

pytestmark = pytest.mark.gpu

W1 = F.LVL1_N + 1


def _enc1(k, v, bits, seed):
    m = (np.asarray(v, np.uint64) << np.uint64(32 - bits - 1)).astype(np.uint32)
    return F.tlwe_encrypt_u32(seed, 0, k.sk.s1, m, k.params.alpha1)


def _outs(s, rows):
    return [b.to_array(np.uint32, (rows, W1)) for b in s.out]


def test_tpch_q1_filter_graph_replays(full636):
    k, ctx = full636
    rows = 96
    rng = np.random.default_rng(61)
    date = rng.integers(0, 512, rows)
    rf = rng.integers(0, 2, rows)
    ls = rng.integers(0, 2, rows)
    thr = 300
    cols = [F.DeviceBuffer.from_array(_enc1(k, date, 9, 1)),
            F.DeviceBuffer.from_array(_enc1(k, rf, 1, 2)),
            F.DeviceBuffer.from_array(_enc1(k, ls, 1, 3))]
    dthr = F.DeviceBuffer.from_array(_enc1(k, [thr], 9, 4))
    one = F.DeviceBuffer.from_array(_enc1(k, [1], 1, 5))
    s = Q.Q1Scratch(rows)

    def run():
        Q.tpch_q1_filter(ctx, cols[0], cols[1], cols[2], dthr, one, rows, scratch=s)

    run()
    ctx.sync()
    eager = _outs(s, rows)
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    for g, (a, b) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        assert np.array_equal(pred[g], (date <= thr) & (rf == a) & (ls == b))

    graph = F.Graph(ctx, run)
    # other eager work on the same context between the replays
    other = F.DeviceBuffer.from_array(_enc1(k, rng.integers(0, 512, rows), 9, 6))
    scratch_out = F.DeviceBuffer(rows * W1 * 4)
    for rep in range(3):
        for b in s.out:
            F.check(F.lib().fhe_memset_dev(b.ptr, 0, rows * W1 * 4))
        graph.launch()
        ctx.sync()
        got = _outs(s, rows)
        assert all(np.array_equal(x, y) for x, y in zip(got, eager)), rep
        Q.range_filter(ctx, other, dthr, rows, 9, level=1, out_dev=scratch_out)
        ctx.sync()

    # new inputs in the same buffers: the replay follows them
    date2 = rng.integers(0, 512, rows)
    c2 = _enc1(k, date2, 9, 7)
    F.check(F.lib().fhe_memcpy_h2d(cols[0].ptr, F._ptr(c2), c2.nbytes))
    graph.launch()
    ctx.sync()
    replay = _outs(s, rows)
    run()
    ctx.sync()
    assert all(np.array_equal(x, y) for x, y in zip(replay, _outs(s, rows)))
    pred = Q.decrypt_predicates(s.out, rows, k.sk.s1)
    assert np.array_equal(pred[0], (date2 <= thr) & (rf == 0) & (ls == 0))
    graph.destroy()
    s.free()
    for b in cols + [dthr, one, other, scratch_out]:
        b.free()


def test_digit_compare_graph_replays(full672):
    k, ctx = full672
    rows = 64
    rng = np.random.default_rng(62)
    x = rng.integers(0, 1 << 20, rows).astype(np.uint64)
    t = int(rng.integers(0, 1 << 20))
    dx = F.DeviceBuffer.from_array(F.encrypt_digits(8, 0, k.sk.s2, x, 5, 4, k.params.alpha2))
    out = F.DeviceBuffer(rows * W1 * 4)
    L = F.lib()

    def run():
        F.check(L.fhe_compare_digits_dev(ctx.h, F.CMP_LE, dx.ptr, 5, 4, t, 1, out.ptr, F.LOGIC,
                                         rows, None))

    run()
    ctx.sync()
    eager = out.to_array(np.uint32, (rows, W1))
    ph = F.tlwe_phase_u32(eager, k.sk.s1).view(np.int32)
    assert np.array_equal(ph > 0, x <= t)
    graph = F.Graph(ctx, run)
    for rep in range(3):
        F.check(L.fhe_memset_dev(out.ptr, 0, rows * W1 * 4))
        graph.launch()
        ctx.sync()
        assert np.array_equal(out.to_array(np.uint32, (rows, W1)), eager), rep
    graph.destroy()
    dx.free()
    out.free()


def test_capture_needing_scratch_fails_cleanly(keys636):
    k = keys636
    ctx = F.Context(k.params)
    ctx.load_bk(1, k.bk, F.BK_INT)
    ctx.load_ksk(10, k.ksk10)
    rows = 8
    a = F.DeviceBuffer.from_array(_enc1(k, [1] * rows, 1, 9))
    o = F.DeviceBuffer(rows * W1 * 4)
    L = F.lib()

    def gate():
        F.check(L.fhe_gate_lvl1_dev(ctx.h, F.GATE_AND, a.ptr, a.ptr, o.ptr, F.LOGIC, rows, None))

    with pytest.raises(F.FheError):
        F.Graph(ctx, gate)          # the gate's scratch is not allocated yet
    gate()                          # eager use afterwards works and sizes it
    ctx.sync()
    ph = F.tlwe_phase_u32(o.to_array(np.uint32, (rows, W1)), k.sk.s1).view(np.int32)
    assert np.all(ph > 0)
    graph = F.Graph(ctx, gate)      # now capturable
    F.check(L.fhe_memset_dev(o.ptr, 0, rows * W1 * 4))
    graph.launch()
    ctx.sync()
    ph = F.tlwe_phase_u32(o.to_array(np.uint32, (rows, W1)), k.sk.s1).view(np.int32)
    assert np.all(ph > 0)
    graph.destroy()
    a.free()
    o.free()
    ctx.close()


def test_small_batch_capture_after_large_eager_run(keys636):
    # A context that has only bootstrapped large batches (one wave per
    # ciphertext) captures a small batch, which runs the two-wave kernel and
    # its own key order.  That order is built when the key is loaded, so the
    # capture records no layout work and an eager small batch before the
    # first replay reads a complete key.
    k = keys636
    ctx = F.Context(k.params)
    try:
        ctx.load_bk(1, k.bk, F.BK_INT)
        rng = np.random.default_rng(62)
        lut = rng.integers(0, 8, 8)
        j = np.arange(F.LVL1_N)
        box = (j + 64) >> 7
        tv = np.where(box == 8, -lut[box % 8], lut[box % 8]).astype(np.int64) << 28
        d_tv = F.DeviceBuffer.from_array((tv % (1 << 32)).astype(np.uint32))
        L = F.lib()

        def batch(B, seed):
            m = rng.integers(0, 8, B)
            c0 = k.enc_lvl0((m << 28).astype(np.uint32), seed=seed)
            return m, F.DeviceBuffer.from_array(c0), F.DeviceBuffer(B * W1 * 4)

        mb, inb, outb = batch(4096, 63)
        F.check(L.fhe_pbs_lvl01_dev(ctx.h, inb.ptr, d_tv.ptr, 0, outb.ptr, 4096, None))
        ctx.sync()
        ms, ins, outs = batch(64, 64)

        def small():
            F.check(L.fhe_pbs_lvl01_dev(ctx.h, ins.ptr, d_tv.ptr, 0, outs.ptr, 64, None))

        graph = F.Graph(ctx, small)
        eager_out = F.DeviceBuffer(64 * W1 * 4)
        F.check(L.fhe_pbs_lvl01_dev(ctx.h, ins.ptr, d_tv.ptr, 0, eager_out.ptr, 64, None))
        ctx.sync()
        eager = eager_out.to_array(np.uint32, (64, W1))
        for rep in range(3):
            graph.launch()
            ctx.sync()
            assert np.array_equal(outs.to_array(np.uint32, (64, W1)), eager), rep
        for m, o in ((ms, eager), (mb, outb.to_array(np.uint32, (4096, W1)))):
            ph = F.tlwe_phase_u32(o, k.sk.s1).view(np.int32).astype(np.int64)
            assert np.array_equal(np.round(ph / (1 << 28)).astype(np.int64) % 16, lut[m])
        graph.destroy()
    finally:
        ctx.close()
