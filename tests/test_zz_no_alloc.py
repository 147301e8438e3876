"""fhe_gate_lvl1_noalloc_dev: the lvl1 gate bootstrap on the library's fixed
device scratch, against fhe_gate_lvl1_dev (grow-on-demand scratch).

Both run the same kernels (lincomb, key switch 10, test vector, PBS) on the
same inputs, so the outputs must be equal word for word.  Two calls in a
row, the second on a smaller batch with another gate, check that nothing of
the first call's scratch contents leaks into the second; the decrypted bits
are checked against the plaintext gate.  Batches above the fixed maximum
are refused before any launch.
"""

import ctypes

import numpy as np
import pytest

import cufhedb_amd as F

pytestmark = pytest.mark.gpu
W1 = F.LVL1_N + 1
MU = 1 << 29


def _fn():
    fn = F.lib().fhe_gate_lvl1_noalloc_dev
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return fn


@pytest.fixture(scope="module")
def gctx(keys636):
    k = keys636
    ctx = F.Context(k.params)
    ctx.load_bk(1, k.bk, F.BK_INT)
    ctx.keygen_ksk(10, k.sk, k.seed + 2)
    yield k, ctx
    ctx.close()


def _bits(k, B, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2, B)
    m = np.where(x == 1, MU, -MU).astype(np.int64) % (1 << 32)
    return x, F.DeviceBuffer.from_array(k.enc_lvl1(m.astype(np.uint32), seed=seed))


PLAIN = {F.GATE_NAND: lambda a, b: 1 - (a & b), F.GATE_AND: lambda a, b: a & b,
         F.GATE_OR: lambda a, b: a | b, F.GATE_NOR: lambda a, b: 1 - (a | b),
         F.GATE_XNOR: lambda a, b: 1 - (a ^ b)}


def test_noalloc_gate_equals_allocating_gate(gctx):
    k, ctx = gctx
    L = F.lib()
    noalloc = _fn()
    # the larger batch first, then a smaller one with another gate: the
    # second call reads only what it wrote itself
    for B, gate, seed in ((1536, F.GATE_NAND, 501), (300, F.GATE_AND, 502),
                          (2048, F.GATE_XNOR, 503), (7, F.GATE_OR, 504)):
        xa, da = _bits(k, B, seed)
        xb, db = _bits(k, B, seed + 1000)
        ref, got = F.DeviceBuffer(B * W1 * 4), F.DeviceBuffer(B * W1 * 4)
        print(f"noalloc gate {gate}, batch {B}", flush=True)
        F.check(L.fhe_gate_lvl1_dev(ctx.h, gate, da.ptr, db.ptr, ref.ptr, F.LOGIC, B, None))
        F.check(noalloc(ctx.h, gate, da.ptr, db.ptr, got.ptr, F.LOGIC, B, None))
        ctx.sync()
        r = ref.to_array(np.uint32, (B, W1))
        g = got.to_array(np.uint32, (B, W1))
        assert np.array_equal(g, r), (B, gate)
        ph = F.tlwe_phase_u32(g, k.sk.s1).view(np.int32)
        assert np.array_equal((ph > 0).astype(np.int64), PLAIN[gate](xa, xb)), (B, gate)
        for buf in (da, db, ref, got):
            buf.free()


def test_noalloc_gate_refuses_oversize_batch(gctx):
    k, ctx = gctx
    B = 2049
    buf = F.DeviceBuffer(B * W1 * 4)
    rc = _fn()(ctx.h, F.GATE_NAND, buf.ptr, buf.ptr, buf.ptr, F.LOGIC, B, None)
    assert rc == F.FHE_ERR_ARG
    assert b"2048" in F.lib().fhe_last_error()
    buf.free()
