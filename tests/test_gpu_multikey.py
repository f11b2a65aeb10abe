"""Group keys beyond one word and 16 bytes (the reference's generic paths
group by any number of columns of any type: the row-key loop of
physical_planner.cpp:12291-12516 and RunCountGroupAggregate's binary keys,
count_group_agg.cpp:32-227): three to eight columns, DOUBLE columns, and two
nullable columns whose NULLs form groups. The operator interns each row's
packed key tuple on the GPU; results are checked against the oracle."""
import numpy as np
import pytest

from slothdb_amd import gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, assert_topn_tie_aware, oracle_rows
from tests.test_gpu_strings import _split, _as_shuffled_dict, _words

pytestmark = pytest.mark.gpu


def _expected(q, cols, dicts):
    c, v, t, ov = O.groupby_exact(q, cols)
    assert not ov
    return oracle_rows(c, v, t, dicts), t


def _run(q, batches, device, params=None):
    op = G.GpuHashAggregate(q, params)
    op.Init()
    keep = []
    for cols in batches:
        if device:
            dc = [G.DeviceColumn(c) for c in cols]
            keep.append(dc)
            op.consume_device(dc)
        else:
            op.consume(cols)
    op.finalize()
    rows = op.result().rows()
    info = op.plan_info()
    op.close()
    return rows, info


def _norm(rows, nk):
    """Keys as SQL groups them, and sortable: -0.0 as 0.0, NaN as +inf (the
    tables hold no infinity)."""
    def f(x):
        if isinstance(x, float):
            return float("inf") if x != x else (0.0 if x == 0.0 else x)
        return x
    return [tuple(f(x) for x in r[:nk]) + tuple(r[nk:]) for r in rows]


def _table(n, seed, words):
    rng = np.random.default_rng(seed)
    dic = G.Dictionary(words, unique=True)
    return [G.Column(G.INTEGER, rng.integers(-20, 20, n).astype(np.int32), (rng.random(n) > 0.05).astype(np.uint8)),
            G.Column(G.BIGINT, rng.integers(0, 30, n) * 1_000_000_007),
            G.Column(G.DOUBLE, rng.choice(np.array([0.0, -0.0, 0.5, np.nan, 1e300, -3.25]), n),
                     (rng.random(n) > 0.05).astype(np.uint8)),
            G.Column(G.VARCHAR, rng.integers(0, len(words), n).astype(np.uint32),
                     (rng.random(n) > 0.05).astype(np.uint8), dictionary=dic),
            G.Column(G.BIGINT, rng.integers(-10**9, 10**9, n)),
            G.Column(G.DOUBLE, rng.normal(10, 3, n))], dic  # AVG argument: no sums that cancel to near 0


AGGS = [(G.COUNT_STAR, None), (G.SUM, 4), (G.MIN, 4), (G.MAX, 4), (G.AVG, 5), (G.COUNT, 2)]
TYPES = [G.INTEGER, G.BIGINT, G.DOUBLE, G.VARCHAR, G.BIGINT, G.DOUBLE]
NULLABLE = [True, False, True, True, False, False]


@pytest.mark.parametrize("strategy", [G.AUTO, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX])
def test_four_column_keys_with_nulls(strategy):
    words = _words(300, 31)
    cols, dic = _table(400_000, 32, words)
    q = G.Query(TYPES, [0, 1, 2, 3], AGGS, nullable=NULLABLE, strategy=strategy)
    exp, t = _expected(q, cols, {3: dic})
    for device in (True, False):
        print(f"[multikey] 4 columns strategy={strategy} device={device}", flush=True)
        got, _ = _run(q, _split(cols, 3), device)
        assert_same_groups(_norm(got, 4), _norm(exp, 4), 4, t)


def test_drop_null_keys_strings_and_mixed_forms():
    # the VARCHAR column arrives as strings in some batches and as codes
    # (each batch its own dictionary) in others; rows with a NULL key dropped
    words = _words(2_000, 33)
    cols, dic = _table(300_000, 34, words)
    q = G.Query(TYPES, [3, 0, 1], AGGS, nullable=NULLABLE, drop_null_keys=True)
    exp, t = _expected(q, cols, {0: dic})
    parts = _split(cols, 4)
    for i, p in enumerate(parts):
        codes, valid = p[3].values, p[3].validity
        strs = [dic.get(int(c)) if valid[r] else None for r, c in enumerate(codes)]
        raw = G.Column.strings(strs)
        p[3] = raw if i % 2 == 0 else _as_shuffled_dict(raw, i)
    for device in (True, False):
        print(f"[multikey] strings + codes device={device}", flush=True)
        got, _ = _run(q, parts, device)
        assert_same_groups(got, exp, 3, t)


def test_double_key_and_nullable_pairs():
    words = _words(50, 35)
    cols, dic = _table(200_000, 36, words)
    for gcols in ([2], [5], [0, 2], [0, 1]):
        q = G.Query(TYPES, gcols, AGGS, nullable=NULLABLE)
        print(f"[multikey] group columns {gcols}", flush=True)
        got, _ = _run(q, [cols], True)
        exp, t = _expected(q, cols, None)
        assert_same_groups(_norm(got, len(gcols)), _norm(exp, len(gcols)), len(gcols), t)


def test_eight_columns_filter_and_topn():
    rng = np.random.default_rng(37)
    n = 500_000
    cols = [G.Column(G.INTEGER, rng.integers(0, 3, n).astype(np.int32)) for _ in range(7)]
    cols.append(G.Column(G.BIGINT, rng.integers(0, 5, n)))
    cols.append(G.Column(G.BIGINT, rng.integers(-100, 100, n)))
    types = [G.INTEGER] * 7 + [G.BIGINT, G.BIGINT]
    q = G.Query(types, list(range(8)), [(G.COUNT_STAR, None), (G.SUM, 8)], predicates=[(8, G.GT, -50)],
                nullable=[False] * 9)
    exp, t = _expected(q, cols, None)
    got, _ = _run(q, [cols], True)
    assert_same_groups(got, exp, 8, t)
    q.topn = (8, False, 10)  # ORDER BY COUNT(*) DESC LIMIT 10
    got, _ = _run(q, _split(cols, 2), True)
    assert len(got) == 10
    assert_topn_tie_aware(got, exp, order_col=8, k=10, types=t)


def test_tuple_hash_collisions_forced():
    # one stored hash bit for the strings and the tuples: all distinct tuples
    # share a hash (each round settles one); groups must not change
    rng = np.random.default_rng(39)
    n = 5_000
    words = _words(5, 38)
    dic = G.Dictionary(words, unique=True)
    cols = [G.Column(G.INTEGER, rng.integers(0, 3, n).astype(np.int32), (rng.random(n) > 0.1).astype(np.uint8)),
            G.Column(G.VARCHAR, rng.integers(0, 5, n).astype(np.uint32), (rng.random(n) > 0.1).astype(np.uint8),
                     dictionary=dic),
            G.Column(G.BIGINT, rng.integers(-1000, 1000, n))]
    q = G.Query([G.INTEGER, G.VARCHAR, G.BIGINT], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)],
                nullable=[True, True, False])
    exp, t = _expected(q, cols, {1: dic})
    got, _ = _run(q, _split(cols, 2), True, params={"str_hash_bits": 1})
    assert_same_groups(got, exp, 2, t)
    assert len(got) == 24  # (3 values + NULL) x (5 strings + NULL)
