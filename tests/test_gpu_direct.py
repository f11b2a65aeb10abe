"""The direct-mapped strategy (IntGroupAgg's dense slot_of[key],
src/execution/int_group_agg.cpp:69-277) against the oracle: LDS-resident and
global dense tables, NULL / SENTINEL / filtered keys, dictionary keys, TopN,
several batches, and the fallback when the key range is too wide."""
import numpy as np
import pytest

from slothdb_amd import datagen, gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, assert_topn_tie_aware, oracle_rows

pytestmark = pytest.mark.gpu


def _exact(q, cols, dicts=None):
    c, v, t, ov = O.groupby_exact(q, cols)
    assert not ov
    return oracle_rows(c, v, t, dicts), t


def _check(q, cols, params=None, dicts=None):
    print(f"[direct] strategy={q.strategy} groups={q.group_cols} rows={len(cols[0])} params={params}", flush=True)
    res, info = G.run(q, cols, device=True, params=params)
    exp, types = _exact(q, cols, dicts)
    assert_same_groups(res.rows(), exp, len(q.group_cols), types)
    return res, info


@pytest.mark.parametrize("strategy", [G.DIRECT, G.AUTO])
def test_plain_1m_direct(strategy):
    # the default plan's rule (selector off) picks DIRECT for 1000 dense keys
    q, cols = datagen.plain_workload(n=1_000_000)
    q.strategy = strategy
    _, info = _check(q, cols, params={"selector": 0})
    assert info["strategy"] == G.DIRECT and info["lds_slots"] == 1000  # keys 0..999 in LDS
    # with the built-in selector the plan may differ, the result may not
    _, info = _check(q, cols)
    print(f"[direct] with the selector: strategy {info['strategy']}, selector used {info['selector_used']}", flush=True)


def test_direct_global_table():
    # 300k distinct dense keys: too many for a workgroup's LDS, one global record each
    rng = np.random.default_rng(1)
    n = 2_000_000
    k = (rng.integers(0, 300_000, n) - 150_000).astype(np.int32)
    v = rng.integers(-10**9, 10**9, n)
    q = G.Query([G.INTEGER, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1), (G.MIN, 1), (G.MAX, 1)],
                nullable=[False, False], strategy=G.DIRECT)
    _, info = _check(q, [G.Column(G.INTEGER, k), G.Column(G.BIGINT, v)])
    assert info["strategy"] == G.DIRECT and info["lds_slots"] == 0


def test_direct_nulls_sentinel_and_filter():
    rng = np.random.default_rng(2)
    n = 500_000
    k = rng.integers(-5, 100, n).astype(np.int64)
    k[rng.integers(0, n, 300)] = np.iinfo(np.int64).min  # the table's empty-slot word, as a real key
    kvalid = (rng.random(n) > 0.05).astype(np.uint8)
    x = rng.integers(0, 100, n).astype(np.int32)
    d = rng.normal(0, 1e3, n)
    cols = [G.Column(G.BIGINT, k, kvalid), G.Column(G.INTEGER, x), G.Column(G.DOUBLE, d)]
    for drop in (False, True):
        q = G.Query([G.BIGINT, G.INTEGER, G.DOUBLE], [0],
                    [(G.COUNT_STAR, None), (G.SUM, 1), (G.AVG, 2), (G.MAX, 2)],
                    predicates=[(1, G.LT, 70)], nullable=[True, False, False], drop_null_keys=drop,
                    strategy=G.DIRECT)
        _, info = _check(q, cols)
        assert info["strategy"] == G.DIRECT


def test_direct_dictionary_key_topn():
    q, cols = datagen.composite_workload(n=1_000_000, dict_size=3000, topn=None)
    dic = cols[1].dictionary
    # GROUP BY b (dictionary codes) WHERE c < 30, ORDER BY COUNT(*) DESC LIMIT 7
    q1 = G.Query(q.column_types, [1], [(G.COUNT_STAR, None), (G.SUM, 2)], predicates=q.predicates,
                 nullable=[False] * 3, strategy=G.DIRECT)
    res, info = _check(q1, cols, dicts={0: dic})
    assert info["strategy"] == G.DIRECT
    exp, types = _exact(q1, cols, {0: dic})
    q1.topn = (1, False, 7)
    res2, info2 = G.run(q1, cols, device=True)
    assert info2["strategy"] == G.DIRECT and res2.num_rows == 7
    assert_topn_tie_aware(res2.rows(), exp, order_col=1, k=7, types=types)


def test_direct_host_batches():
    q, cols = datagen.plain_workload(n=600_000)
    q.strategy = G.DIRECT
    op = G.GpuHashAggregate(q)
    op.Init()
    for s in range(0, 600_000, 122_880):
        op.consume([G.Column(c.type, c.values[s:s + 122_880]) for c in cols])
    op.finalize()
    exp, types = _exact(q, cols)
    assert_same_groups(op.result().rows(), exp, 1, types)
    assert op.plan_info()["strategy"] == G.DIRECT
    op.close()


def test_direct_too_wide_falls_back():
    q, cols = datagen.highcard_workload(n=400_000, nkeys=100_000)  # sparse 64-bit ids
    q.strategy = G.DIRECT
    _, info = _check(q, cols)
    assert info["strategy"] != G.DIRECT
