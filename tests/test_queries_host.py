"""Host side of the query layer: the client-side group aggregation under
decrypted predicates against a row-by-row plaintext SQL evaluation."""

import numpy as np

from cufhedb_amd import queries as Q


def test_group_aggregates_match_row_by_row_sql():
    rng = np.random.default_rng(5)
    rows = 2000
    qty = rng.integers(1, 51, rows)
    price = rng.integers(900, 105001, rows)
    disc = rng.integers(0, 11, rows)
    cat1, cat2 = rng.integers(0, 3, rows), rng.integers(0, 2, rows)
    keep = rng.integers(0, 2, rows).astype(bool)
    pred = [(keep & (cat1 == a) & (cat2 == b)).astype(np.int64) for a, b in Q.GROUPS]
    want = []
    for a, b in Q.GROUPS:
        s_q = s_r = n = 0
        for i in range(rows):
            if keep[i] and cat1[i] == a and cat2[i] == b:
                s_q += int(qty[i])
                s_r += int(price[i]) * (100 - int(disc[i]))
                n += 1
        want.append((s_q, s_r, n))
    assert np.array_equal(Q.group_aggregates(pred, qty, price, disc), np.array(want, np.int64))


def test_group_aggregates_full_table_does_not_overflow():
    rows = 1 << 14
    ones = [np.ones(rows, np.int64)]
    got = Q.group_aggregates(ones, np.full(rows, 50), np.full(rows, 105000), np.zeros(rows))
    assert got.tolist() == [[50 * rows, 105000 * 100 * rows, rows]]
