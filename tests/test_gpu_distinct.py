"""COUNT(DISTINCT x) (the generic path's per-group distinct_set,
physical_planner.cpp:12325-12330) through the C ABI against the oracle:
grouped and scalar, next to COUNT/SUM/AVG/MIN/MAX of other columns, INTEGER,
BIGINT and VARCHAR arguments and keys, NULL arguments, filters, TopN, several
batches, and a high-cardinality input whose first phase runs the radix path."""
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


def _check(q, cols, dicts=None, device=True):
    print(f"[distinct] groups={q.group_cols} aggs={q.aggs} rows={len(cols[0])}", flush=True)
    res, info = G.run(q, cols, device=device)
    exp, types = _exact(q, cols, dicts)
    assert_same_groups(res.rows(), exp, len(q.group_cols), types)
    return res, info


def _table(n, seed):
    rng = np.random.default_rng(seed)
    words = G.Dictionary([f"s{j:04d}" for j in range(700)] + ["s0001"])  # one repeated string
    return [G.Column(G.INTEGER, rng.integers(0, 400, n).astype(np.int32)),
            G.Column(G.BIGINT, rng.integers(-2**40, 2**40, 5000)[rng.integers(0, 5000, n)]),
            G.Column(G.VARCHAR, rng.integers(0, 701, n).astype(np.uint32), dictionary=words),
            G.Column(G.INTEGER, rng.integers(-1000, 1000, n).astype(np.int32)),
            G.Column(G.DOUBLE, rng.normal(50, 10, n), (rng.random(n) > 0.2).astype(np.uint8))], words


@pytest.mark.parametrize("x", [0, 1, 2, 3])
def test_count_distinct_with_other_aggregates(x):
    cols, words = _table(400_000, 3)
    types = [c.type for c in cols]
    nullable = [False, False, False, False, True]
    aggs = [(G.COUNT_DISTINCT, x), (G.COUNT_STAR, None), (G.SUM, 3), (G.AVG, 4), (G.MIN, 4), (G.MAX, 3),
            (G.COUNT, 4)]
    for k in (0, 2):  # INTEGER key, VARCHAR key
        if k == x:
            continue
        q = G.Query(types, [k], aggs, nullable=nullable)
        _check(q, cols, {0: words} if k == 2 else None)


def test_scalar_and_nullable_argument():
    cols, words = _table(300_000, 4)
    rng = np.random.default_rng(5)
    xn = G.Column(G.BIGINT, cols[1].values, (rng.random(len(cols[1])) > 0.3).astype(np.uint8))
    cols = cols + [xn]
    types = [c.type for c in cols]
    nullable = [False, False, False, False, True, True]
    _check(G.Query(types, [], [(G.COUNT_DISTINCT, 2), (G.COUNT_STAR, None), (G.AVG, 4)], nullable=nullable), cols)
    _check(G.Query(types, [], [(G.COUNT_DISTINCT, 5)], nullable=nullable), cols)
    _check(G.Query(types, [0], [(G.COUNT_DISTINCT, 5), (G.COUNT_DISTINCT, 5)], nullable=nullable), cols)
    # empty input: one row, COUNT(DISTINCT) = 0
    empty = [G.Column(c.type, c.values[:0], None if c.validity is None else c.validity[:0],
                      dictionary=c.dictionary) for c in cols]
    res, _ = G.run(G.Query(types, [], [(G.COUNT_DISTINCT, 1), (G.SUM, 3)], nullable=nullable), empty, device=False)
    assert res.rows() == [(0, None)]


def test_filter_topn_and_batches():
    cols, words = _table(600_000, 6)
    types = [c.type for c in cols]
    nullable = [False, False, False, False, True]
    q = G.Query(types, [0], [(G.COUNT_DISTINCT, 2), (G.SUM, 3)], predicates=[(3, G.LT, 300)],
                nullable=nullable)
    exp, t = _exact(q, cols)
    q.topn = (1, False, 10)  # ORDER BY COUNT(DISTINCT b) DESC LIMIT 10
    res, _ = G.run(q, cols, device=True)
    assert res.num_rows == 10
    assert_topn_tie_aware(res.rows(), exp, order_col=1, k=10, types=t)
    # host row groups
    q.topn = None
    op = G.GpuHashAggregate(q)
    op.Init()
    for s in range(0, 600_000, 122_880):
        op.consume([G.Column(c.type, c.values[s:s + 122_880], None if c.validity is None else c.validity[s:s + 122_880],
                             dictionary=c.dictionary) for c in cols])
    op.finalize()
    assert_same_groups(op.result().rows(), exp, 1, t)
    op.close()


def test_high_cardinality_first_phase():
    # 3M rows: COUNT(DISTINCT v2) per sparse BIGINT key (first phase over ~2.9M (k, v2) pairs)
    q, cols = datagen.highcard_workload(n=3_000_000, nkeys=1_000_000)
    rng = np.random.default_rng(8)
    cols[2] = G.Column(G.INTEGER, rng.integers(0, 3, len(cols[2])).astype(np.int32))
    q2 = G.Query(q.column_types, [0], [(G.COUNT_DISTINCT, 2), (G.COUNT_STAR, None), (G.SUM, 1), (G.AVG, 3)],
                 nullable=[False] * 4)
    _check(q2, cols)


def test_groups_whose_argument_is_always_null():
    # every k divisible by 7 has a NULL x on all its rows: COUNT(DISTINCT x)
    # is 0 there, and the other aggregates still count those rows
    rng = np.random.default_rng(12)
    n = 200_000
    k = rng.integers(0, 1000, n).astype(np.int32)
    x = rng.integers(0, 50, n)
    valid = ((rng.random(n) > 0.3) & (k % 7 != 0)).astype(np.uint8)
    cols = [G.Column(G.INTEGER, k), G.Column(G.BIGINT, x, valid), G.Column(G.BIGINT, rng.integers(-100, 100, n))]
    for aggs in ([(G.COUNT_DISTINCT, 1)], [(G.COUNT_DISTINCT, 1), (G.COUNT_STAR, None), (G.SUM, 2), (G.COUNT, 1)]):
        for gc in ([0], []):
            q = G.Query([G.INTEGER, G.BIGINT, G.BIGINT], gc, aggs, nullable=[False, True, False])
            res, _ = _check(q, cols, device=bool(gc))
            if gc:
                zero = {r[0] for r in res.rows() if r[1] == 0}
                assert zero == {v for v in range(0, 1000, 7)}


def _norm(rows, nk):
    """Keys as SQL groups them, and sortable: -0.0 as 0.0, NaN as +inf (the
    tables hold no infinity)."""
    def f(x):
        if isinstance(x, float):
            return float("inf") if x != x else (0.0 if x == 0.0 else x)
        return x
    return [tuple(f(x) for x in r[:nk]) + tuple(r[nk:]) for r in rows]


@pytest.mark.parametrize("keys", [[], [0], [5], [0, 2], [2, 6, 0], [6, 5, 0, 2]])
def test_double_argument_and_multi_column_keys(keys):
    # DOUBLE arguments and keys holding -0.0, 0.0 and NaNs of two bit patterns
    # (distinct by value), two VARCHAR key columns with their own dictionaries,
    # nullable keys whose NULLs form groups, and the other aggregates beside
    n = 300_000
    cols, words = _table(n, 40)
    rng = np.random.default_rng(41)
    pool = np.concatenate([np.array([0.0, -0.0, np.nan, -np.nan]), rng.normal(0, 100, 3000)])
    cols.append(G.Column(G.DOUBLE, rng.choice(pool, n), (rng.random(n) > 0.1).astype(np.uint8)))
    other = G.Dictionary([f"t{j}" for j in range(40)])
    cols.append(G.Column(G.VARCHAR, rng.integers(0, 40, n).astype(np.uint32), (rng.random(n) > 0.05).astype(np.uint8),
                         dictionary=other))
    types = [c.type for c in cols]
    dmap = {2: words, 6: other}
    for x in (5, 4, 2):
        if x in keys:
            continue
        for drop in (False, True):
            q = G.Query(types, keys, [(G.COUNT_DISTINCT, x), (G.COUNT_STAR, None), (G.SUM, 3), (G.AVG, 4),
                                      (G.MAX, 1)], drop_null_keys=drop)
            dicts = {i: dmap[c] for i, c in enumerate(keys) if c in dmap}
            print(f"[distinct] keys={keys} x={x} drop={drop}", flush=True)
            res, _ = G.run(q, cols, device=True)
            exp, t = _exact(q, cols, dicts)
            assert_same_groups(_norm(res.rows(), len(keys)), _norm(exp, len(keys)), len(keys), t)


# ---- several DISTINCT argument columns in one query (a distinct set per
# aggregate, physical_planner.cpp:6076, 12325-12349) ----------------------

def test_two_distinct_arguments_next_to_other_aggregates():
    cols, words = _table(400_000, 11)
    rng = np.random.default_rng(12)
    xn = G.Column(G.BIGINT, cols[1].values, (rng.random(len(cols[1])) > 0.25).astype(np.uint8))
    cols = cols + [xn]
    types = [c.type for c in cols]
    nullable = [False, False, False, False, True, True]
    # phase 2 loads the key, the three arguments and four partial states
    # (SUM(3), AVG(4) as SUM and COUNT, COUNT(*)): the 8-column plan limit
    aggs = [(G.COUNT_DISTINCT, 1), (G.SUM, 3), (G.COUNT_DISTINCT, 2), (G.AVG, 4),
            (G.COUNT_DISTINCT, 5), (G.COUNT_DISTINCT, 1), (G.COUNT_STAR, None)]
    for k in ([0], [2], []):  # INTEGER key, VARCHAR key, scalar
        q = G.Query(types, k, aggs, nullable=nullable)
        _check(q, cols, {g: words for g, c in enumerate(k) if c == 2})
    # two key columns (packed: both loaded) and three (an interned tuple code)
    aggs2 = [(G.COUNT_DISTINCT, 1), (G.COUNT_DISTINCT, 5), (G.SUM, 3), (G.COUNT_STAR, None)]
    for k in ([0, 2], [0, 2, 3]):
        q = G.Query(types, k, aggs2, nullable=nullable)
        _check(q, cols, {g: words for g, c in enumerate(k) if c == 2})


def test_distinct_int_string_and_double_arguments_with_filter_and_nulls():
    rng = np.random.default_rng(13)
    n = 300_000
    cols, words = _table(n, 14)
    kv = (rng.random(n) > 0.05).astype(np.uint8)
    cols[0] = G.Column(G.INTEGER, cols[0].values % 50, kv)
    types = [c.type for c in cols]
    nullable = [True, False, False, False, True]
    aggs = [(G.COUNT_DISTINCT, 2), (G.COUNT_DISTINCT, 4), (G.COUNT_DISTINCT, 3), (G.SUM, 1)]
    for drop in (False, True):
        q = G.Query(types, [0], aggs, nullable=nullable, predicates=[(3, G.GT, -500)], drop_null_keys=drop)
        _check(q, cols)


def test_distinct_arguments_beyond_the_column_limit_are_rejected():
    cols, words = _table(1000, 16)
    types = [c.type for c in cols]
    aggs = [(G.COUNT_DISTINCT, 1), (G.COUNT_DISTINCT, 2), (G.COUNT_DISTINCT, 4), (G.SUM, 3), (G.AVG, 4),
            (G.MIN, 4), (G.MAX, 3)]
    q = G.Query(types, [0], aggs, nullable=[False, False, False, False, True])
    with pytest.raises(G.GbaggError, match="at most 8"):
        G.run(q, cols)


def test_two_distinct_arguments_topn_and_host_batches():
    cols, words = _table(500_000, 15)
    types = [c.type for c in cols]
    aggs = [(G.COUNT_DISTINCT, 3), (G.COUNT_DISTINCT, 1), (G.COUNT_STAR, None)]
    q = G.Query(types, [0], aggs, nullable=[False, False, False, False, True])
    exp, etypes = _exact(q, cols)
    qt = G.Query(types, [0], aggs, nullable=[False, False, False, False, True], topn=(2, False, 10))
    res, info = G.run(qt, cols, device=True)
    assert_topn_tie_aware(res.rows(), exp, order_col=2, k=10, types=etypes)
    op = G.GpuHashAggregate(q)
    op.Init()
    RG = 122_880
    for s in range(0, len(cols[0]), RG):
        op.consume([G.Column(c.type, c.values[s:s + RG],
                             None if c.validity is None else c.validity[s:s + RG], dictionary=c.dictionary)
                    for c in cols])
    op.finalize()
    assert_same_groups(op.result().rows(), exp, 1, etypes)
    op.close()
