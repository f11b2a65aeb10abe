"""The high-cardinality GROUP BY at its full size: 100M rows, about 4.85M
groups (SELECT k, COUNT(*), SUM(v1), MIN(v2), MAX(v2), AVG(v3) GROUP BY k; the
reference runs it on its per-thread-map path, physical_planner.cpp:7613-7931).
The result is checked against an exact sort-based aggregation in numpy
(COUNT, SUM, MIN, MAX per group, bit for bit) and against properties that hold
at any size: every key once, the counts summing to the rows, the sum of
AVG x COUNT equal to the column's sum."""
import numpy as np
import pytest

from slothdb_amd import datagen, gbagg as G

pytestmark = pytest.mark.gpu


def test_highcard_100m_exact():
    q, cols = datagen.highcard_workload()
    n = len(cols[0].values)
    assert n == 100_000_000
    print("[fullsize] 100M rows: GPU aggregate", flush=True)
    res, info = G.run(q, cols, device=True)
    k, cnt, s1, mn, mx, avg = (np.asarray(c) for c in res.columns[:6])
    print(f"[fullsize] {len(k)} groups, strategy {info['strategy']}; numpy sort-based aggregation", flush=True)
    kv = cols[0].values
    order = np.argsort(kv, kind="stable")
    ks = kv[order]
    starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    ek = ks[starts]
    ecnt = np.diff(np.r_[starts, n])
    es1 = np.add.reduceat(cols[1].values[order], starts)  # int64, wraps like the GPU's int64 sums
    v2s = cols[2].values[order]
    emn = np.minimum.reduceat(v2s, starts)
    emx = np.maximum.reduceat(v2s, starts)
    del order, ks, v2s
    # result rows in key order
    ro = np.argsort(k)
    assert len(k) == len(ek) and np.array_equal(k[ro], ek), "keys: every key exactly once"
    assert np.array_equal(cnt[ro], ecnt) and int(cnt.sum()) == n
    assert np.array_equal(s1[ro], es1)
    assert np.array_equal(mn[ro], emn) and np.array_equal(mx[ro], emx)
    v3 = cols[3].values
    tot = float(np.sum(avg * cnt.astype(np.float64)))
    assert abs(tot - float(np.sum(v3))) <= 1e-9 * float(np.sum(np.abs(v3)))
