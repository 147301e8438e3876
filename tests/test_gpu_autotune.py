"""Tile autotuning of the lvl2 -> lvl1 key switch (fhe_autotune, fhe_tune_*).

Every tile (8, 16, 32 ciphertexts per workgroup with a column per lane, 64
with a ciphertext per lane) must give the oracle's key switch word for word
(IdentityKeySwitch<lvl21param>, TFHEpp/src/keyswitch.cpp:12-48) at batch
sizes that leave partial tiles, and the two forms must agree word for word
at 4,096 and 6,000 ciphertexts;
autotuning must leave results unchanged and record a valid table; bad
tables are refused.
"""

import numpy as np
import pytest

import cufhedb_amd as F

import _oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ksk21(full636):
    k, _ = full636
    # the context's KSK 21 was generated on the device from this seed; the
    # host generator gives the same key word for word
    return F.keygen_ksk(k.sk, 21, k.seed + 5)


@pytest.mark.parametrize("tile", [8, 16, 32, 64])
def test_every_tile_matches_oracle(full636, ksk21, tile):
    k, ctx = full636
    rng = np.random.default_rng(tile)
    try:
        ctx.set_tune_table([(1, tile)])
        for B in (1, 7, 33, 100, 130):
            x = rng.integers(0, 2 ** 63, (B, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2) \
                + rng.integers(0, 2, (B, F.LVL2_N + 1), dtype=np.uint64)
            got = ctx.iks(21, x)
            for b in sorted({0, B // 2, B - 1}):
                assert np.array_equal(got[b], O.iks64(x[b], ksk21, 1024, 10, 3)), (B, b)
    finally:
        ctx.set_tune_table([])


@pytest.mark.parametrize("B", [4096, 6000])
def test_forms_agree_at_full_size(full636, ksk21, B):
    """4,096 and 6,000 lvl2 ciphertexts of random words (16 and 32 output
    columns per wave in the ciphertext-per-lane form; 6,000 leaves a partial
    tile): the ciphertext-per-lane form (tile 64) equals the column-per-lane
    form (tile 32) word for word, and three rows equal the oracle."""
    _, ctx = full636
    rng = np.random.default_rng(B)
    x = rng.integers(0, 2 ** 63, (B, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2) \
        + rng.integers(0, 2, (B, F.LVL2_N + 1), dtype=np.uint64)
    try:
        ctx.set_tune_table([(1, 32)])
        cols = ctx.iks(21, x)
        ctx.set_tune_table([(1, 64)])
        lanes = ctx.iks(21, x)
    finally:
        ctx.set_tune_table([])
    assert np.array_equal(lanes, cols)
    for b in (0, 2049, B - 1):
        assert np.array_equal(lanes[b], O.iks64(x[b], ksk21, 1024, 10, 3)), b


def test_autotune_keeps_results(full636):
    k, ctx = full636
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2 ** 63, (300, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2)
    before = ctx.iks(21, x)
    try:
        res = ctx.autotune(batches=(32, 256, 2048))
        print("\n  lvl21 key-switch tiles:", {b: (t, {c: round(v, 3) for c, v in ms.items()})
                                             for b, (t, ms) in res.items()})
        table = ctx.tune_table()
        assert [b for b, _ in table] == [32, 256, 2048]
        assert all(t in (8, 16, 32, 64) for _, t in table)
        assert all(all(v > 0 for v in ms.values()) for _, ms in res.values())
        assert np.array_equal(ctx.iks(21, x), before)
        assert np.array_equal(ctx.iks(21, x[:5]), before[:5])   # below the first size
    finally:
        ctx.set_tune_table([])
    assert ctx.tune_table() == []


def test_bad_tables_refused(full636):
    _, ctx = full636
    for table in ([(1, 12)], [(64, 8), (32, 16)], [(0, 8)], [(1, 8)] * 17):
        with pytest.raises(F.FheError):
            ctx.set_tune_table(table)
    with pytest.raises(F.FheError):
        ctx.autotune(batches=())
    assert ctx.tune_table() == []


@pytest.mark.parametrize("split", [2, 4, 8])
def test_iks_split_bit_exact(full636, split):
    """Key switches 10 and 20 with the digit positions split over 2, 4 or 8
    workgroups per ciphertext (integer atomics) equal the unsplit kernel
    (itself pinned to the oracle) word for word; kind 10 also against the
    oracle directly."""
    k, ctx = full636
    rng = np.random.default_rng(100 + split)
    knobs = (F.TUNE_IKS_SPLIT, F.TUNE_IKS20_SPLIT)
    x1 = rng.integers(0, 2 ** 32, (37, F.LVL1_N + 1), dtype=np.uint64).astype(np.uint32)
    x2 = rng.integers(0, 2 ** 63, (37, F.LVL2_N + 1), dtype=np.uint64) * np.uint64(2)
    try:
        for knob in knobs:
            ctx.set_tune_table([(1, 1)], knob=knob)
        ref10, ref20 = ctx.iks(10, x1), ctx.iks(20, x2)
        for knob in knobs:
            ctx.set_tune_table([(1, split)], knob=knob)
        got10, got20 = ctx.iks(10, x1), ctx.iks(20, x2)
    finally:
        for knob in knobs:
            ctx.set_tune_table([], knob=knob)
    assert np.array_equal(got10, ref10)
    assert np.array_equal(got20, ref20)
    for b in (0, 36):
        assert np.array_equal(got10[b], O.iks32(x1[b], k.ksk10, k.n0))


@pytest.mark.parametrize("knob", [F.TUNE_IKS_SPLIT, F.TUNE_IKS20_SPLIT])
def test_autotune_split_table(full636, knob):
    # kinds 10 and 20 are tuned apart, each on its own key; tuning one
    # leaves the other's table alone
    _, ctx = full636
    other = F.TUNE_IKS20_SPLIT if knob == F.TUNE_IKS_SPLIT else F.TUNE_IKS_SPLIT
    try:
        ctx.set_tune_table([(1, 2)], knob=other)
        res = ctx.autotune(batches=(16, 256, 4096), knob=knob)
        print(f"\n  key-switch split, knob {knob}:", res)
        assert [b for b, _ in ctx.tune_table(knob=knob)] == [16, 256, 4096]
        assert all(s in (1, 2, 4, 8) for _, s in ctx.tune_table(knob=knob))
        assert ctx.tune_table(knob=other) == [(1, 2)]
    finally:
        ctx.set_tune_table([], knob=knob)
        ctx.set_tune_table([], knob=other)


def test_tuning_cache_round_trip(full636, tmp_path):
    """A tuned table saved to the per-device JSON cache comes back into a
    fresh context of the same parameter set; another parameter set finds no
    entry and keeps its built-in rule."""
    k, ctx = full636
    path = str(tmp_path / "tune.json")
    tables = {F.TUNE_IKS21_TILE: [(1, 8), (512, 16), (2048, 32)],
              F.TUNE_PBS1_WAVES: [(1, 2), (1024, 1)],
              F.TUNE_IKS_SPLIT: [(1, 8), (256, 2)],
              F.TUNE_IKS20_SPLIT: [(1, 4)]}
    try:
        for knob, t in tables.items():
            ctx.set_tune_table(t, knob=knob)
        ctx.save_tuning(path)
    finally:
        for knob in tables:
            ctx.set_tune_table([], knob=knob)
    fresh = F.Context(k.params)
    try:
        assert fresh.load_tuning(path)
        for knob, t in tables.items():
            assert fresh.tune_table(knob=knob) == t
    finally:
        fresh.close()
    other = F.Context(F.preset(F.PRESET_REF672))
    try:
        assert not other.load_tuning(path)
        assert other.tune_table(knob=F.TUNE_IKS21_TILE) == []
    finally:
        other.close()
