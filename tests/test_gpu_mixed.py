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


def test_mixed_plan_shape():
    plan = mixed.make_plan(seed=1, nbatches=200)
    sizes = np.array([s for _, s in plan])
    assert sizes.min() >= 256 and sizes.max() <= 65536
    frac = np.mean([k == "pbs1" for k, _ in plan])
    assert 0.55 < frac < 0.85
