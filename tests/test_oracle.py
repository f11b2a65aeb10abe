"""The CPU oracle pinned against the reference's own tests, the reference's
compiled aggregators (golden fixtures), independent implementations (pandas,
brute force) and closed forms. Runs without a GPU."""
import json
import os
from collections import defaultdict

import numpy as np
import pytest

from oracle import oracle as O
from slothdb_amd import datagen, gbagg as G
from tests._util import assert_same_groups, oracle_rows

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def exact_rows(q, cols, dicts=None):
    c, v, t, ov = O.groupby_exact(q, cols)
    return oracle_rows(c, v, t, dicts), t, ov


# ---- hash -----------------------------------------------------------------

def closed_form_hash(x):
    p = (x & (2**64 - 1)) * 0x9E3779B97F4A7C15
    return (p & (2**64 - 1)) ^ (p >> 64)


def test_hash_golden_vectors():
    g = load("hash_u64.json")
    for x, h in g["i64"]:
        assert O.hash_u64(x) == h
        assert closed_form_hash(x) == h
    for x, h in g["u64"]:
        assert O.hash_u64(x) == h
    # int32 keys are sign-extended to 64 bits before hashing
    for x, h in g["i32"]:
        assert O.hash_u64(x) == h


def test_hash_matches_reference_build_when_present():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(5)
    for x in rng.integers(-2**63, 2**63 - 1, 2000, dtype=np.int64).tolist():
        assert R.ref_hash_i64(x) == O.hash_u64(x)


# ---- the reference's known-answer tests ------------------------------------

def emp():
    depts = G.Dictionary(["Engineering", "Sales", "Marketing"])
    return depts, [G.Column(G.VARCHAR, np.array([0, 0, 1, 1, 2], np.uint32), dictionary=depts),
                   G.Column(G.INTEGER, np.array([100, 120, 80, 90, 95], np.int32))]


def test_known_answers_emp():
    depts, cols = emp()
    aggs = [(G.COUNT_STAR, None), (G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 1)]
    rows, _, _ = exact_rows(G.Query([G.VARCHAR, G.INTEGER], [], aggs, nullable=[0, 0]), cols)
    assert rows == [(5, 485, 97.0, 80, 120)]  # test_aggregation.cpp:16-56
    rows, _, _ = exact_rows(G.Query([G.VARCHAR, G.INTEGER], [0], aggs, nullable=[0, 0]), cols, {0: depts})
    assert sorted(rows) == [("Engineering", 2, 220, 110.0, 100, 120), ("Marketing", 1, 95, 95.0, 95, 95),
                            ("Sales", 2, 170, 85.0, 80, 90)]  # test_aggregation.cpp:58-120


def test_known_answer_empty_and_series():
    q = G.Query([G.INTEGER], [], [(G.COUNT_STAR, None), (G.SUM, 0)], nullable=[0])
    rows, _, _ = exact_rows(q, [G.Column(G.INTEGER, np.zeros(0, np.int32))])
    assert rows == [(0, None)]  # test_aggregation.cpp:122-130
    q = G.Query([G.INTEGER], [], [(G.SUM, 0)], nullable=[0])
    rows, _, _ = exact_rows(q, [G.Column(G.INTEGER, np.arange(1, 10001, dtype=np.int32))])
    assert rows == [(50005000,)]  # test_parallel.cpp:105-123


def test_known_answer_gpu_engine_hash_aggregate():
    q = G.Query([G.INTEGER, G.INTEGER], [0], [(G.SUM, 1), (G.COUNT_STAR, None)], nullable=[0, 0])
    cols = [G.Column(G.INTEGER, [1, 2, 1, 2, 1]), G.Column(G.INTEGER, [10, 20, 30, 40, 50])]
    rows, _, _ = exact_rows(q, cols)
    assert sorted(rows) == [(1, 90, 3), (2, 60, 2)]  # test_performance.cpp:153-171


def test_closed_form_arithmetic_series_per_group():
    # key = i % m, value = i: SUM per key is an arithmetic series
    n, m = 1_000_003, 97
    i = np.arange(n, dtype=np.int64)
    q = G.Query([G.BIGINT, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1), (G.MIN, 1), (G.MAX, 1)],
                nullable=[0, 0])
    rows, _, _ = exact_rows(q, [G.Column(G.BIGINT, i % m), G.Column(G.BIGINT, i)])
    for k, cnt, s, mn, mx in rows:
        c = (n - 1 - k) // m + 1
        assert cnt == c and mn == k and mx == k + (c - 1) * m
        assert s == c * k + m * c * (c - 1) // 2


# ---- brute force on tiny inputs --------------------------------------------

_NAN = float("nan")


def _canon(v):
    """One object for every NaN key, so that NaN keys form one group (Python
    compares tuple and dict keys by identity before equality)."""
    return _NAN if isinstance(v, float) and v != v else v


def brute(q, cols):
    """Direct restatement of SQL GROUP BY semantics in pure Python."""
    n = len(cols[0])
    groups = defaultdict(list)

    def val(c, r):
        col = cols[c]
        if col.validity is not None and not col.validity[r]:
            return None
        x = col.values[r].item()
        if col.type == G.VARCHAR:
            return col.dictionary.get(x)
        return x

    for r in range(n):
        ok = True
        for (c, op, lit) in q.predicates:
            x = val(c, r)
            if x is None:
                ok = False
                break
            if op == G.IN:
                ok = x in lit
            elif op in (G.LIKE, G.NOT_LIKE):
                ok = (lit in x) == (op == G.LIKE)
            else:
                if cols[c].type in (G.INTEGER, G.BIGINT) and isinstance(lit, float):
                    lit = int(lit)  # ival = (int64)dval
                ok = {G.LT: x < lit, G.LE: x <= lit, G.GT: x > lit, G.GE: x >= lit,
                      G.EQ: x == lit, G.NE: x != lit}[op]
            if not ok:
                break
        if not ok:
            continue
        key = tuple(_canon(val(c, r)) for c in q.group_cols)
        if q.drop_null_keys and any(k is None for k in key):
            continue
        groups[key].append(r)
    out = []
    for key, rs in groups.items():
        row = list(key)
        for kind, c in q.aggs:
            if kind == G.COUNT_STAR:
                row.append(len(rs))
                continue
            xs = [val(c, r) for r in rs if val(c, r) is not None]
            if kind == G.COUNT:
                row.append(len(xs))
            elif kind == G.COUNT_DISTINCT:
                # SQL value equality: -0.0 == 0.0 (so in one set), every NaN one value
                row.append(len({"nan" if isinstance(v, float) and v != v else v for v in xs}))
            elif not xs:
                row.append(None)
            elif kind == G.SUM:
                row.append(sum(xs) if cols[c].type != G.DOUBLE else float(np.sum(np.array(xs, np.longdouble))))
            elif kind == G.AVG:
                row.append(float(sum(xs)) / len(xs) if cols[c].type != G.DOUBLE
                           else float(np.sum(np.array(xs, np.longdouble)) / len(xs)))
            elif kind == G.MIN:
                row.append(min(xs))
            else:
                row.append(max(xs))
        out.append(tuple(row))
    return out


def tiny_table(seed, n=400):
    rng = np.random.default_rng(seed)
    words = G.Dictionary(["a", "bb", "ab", "ba", "", "abc", "bb"])  # duplicate "bb"
    cols = [G.Column(G.INTEGER, rng.integers(-3, 4, n), (rng.random(n) > 0.1).astype(np.uint8)),
            G.Column(G.BIGINT, rng.integers(-2**50, 2**50, n), (rng.random(n) > 0.2).astype(np.uint8)),
            G.Column(G.DOUBLE, rng.normal(0, 10, n), (rng.random(n) > 0.2).astype(np.uint8)),
            G.Column(G.VARCHAR, rng.integers(0, 7, n).astype(np.uint32), (rng.random(n) > 0.1).astype(np.uint8),
                     dictionary=words),
            G.Column(G.INTEGER, rng.integers(0, 5, n))]
    return cols


@pytest.mark.parametrize("seed", range(6))
def test_exact_vs_brute_force(seed):
    cols = tiny_table(seed)
    types = [c.type for c in cols]
    aggs = [(G.COUNT_STAR, None), (G.COUNT, 1), (G.SUM, 1), (G.MIN, 2), (G.MAX, 0), (G.AVG, 2), (G.SUM, 2),
            (G.AVG, 0)]
    cases = [
        G.Query(types, [0], aggs),
        G.Query(types, [3], aggs),
        G.Query(types, [4, 3], aggs, drop_null_keys=True),
        G.Query(types, [0], aggs, predicates=[(4, G.LT, 3), (2, G.GT, -5.5)]),
        G.Query(types, [4], aggs, predicates=[(3, G.LIKE, "b")]),
        G.Query(types, [4], aggs, predicates=[(3, G.NE, "bb"), (0, G.IN, [-1, 2, 3])]),
        G.Query(types, [0], aggs, predicates=[(4, G.LE, 2.7)]),  # int column, double literal
        G.Query(types, [], aggs),
    ]
    for q in cases:
        dicts = {i: cols[c].dictionary for i, c in enumerate(q.group_cols) if cols[c].type == G.VARCHAR}
        got, t, _ = exact_rows(q, cols, dicts)
        exp = brute(q, cols)
        if not q.group_cols and not exp:
            exp = [tuple([0 if k in (G.COUNT_STAR, G.COUNT) else None for k, _ in q.aggs])]
        assert_same_groups(got, exp, len(q.group_cols), t)


@pytest.mark.parametrize("seed", range(4))
def test_count_distinct_vs_brute_force(seed):
    # distinct non-NULL values per group (physical_planner.cpp:12325-12330);
    # VARCHAR values by content (the dictionary repeats "bb")
    cols = tiny_table(seed)
    types = [c.type for c in cols]
    for x in (0, 1, 3, 4):
        for q in (G.Query(types, [4], [(G.COUNT_DISTINCT, x), (G.COUNT_STAR, None), (G.SUM, 1)]),
                  G.Query(types, [3], [(G.COUNT_DISTINCT, x)], drop_null_keys=True),
                  G.Query(types, [], [(G.COUNT_DISTINCT, x), (G.MAX, 2)])):
            dicts = {i: cols[c].dictionary for i, c in enumerate(q.group_cols) if cols[c].type == G.VARCHAR}
            got, t, _ = exact_rows(q, cols, dicts)
            assert_same_groups(got, brute(q, cols), len(q.group_cols), t)


@pytest.mark.parametrize("seed", range(3))
def test_count_distinct_double_and_multi_column_keys_vs_brute_force(seed):
    # DOUBLE arguments and keys with -0.0, 0.0 and NaNs of two bit patterns;
    # zero to three group columns, nullable keys kept or dropped
    cols = tiny_table(seed)
    rng = np.random.default_rng(100 + seed)
    n = len(cols[0])
    pool = np.array([0.0, -0.0, 1.5, -1.5, np.nan, -np.nan, 3.0, 1e-300])
    cols.append(G.Column(G.DOUBLE, rng.choice(pool, n), (rng.random(n) > 0.15).astype(np.uint8)))
    types = [c.type for c in cols]
    for x in (5, 2, 0, 3):
        for keys, drop in (([], False), ([4, 0], False), ([0, 3, 5], True), ([5], False), ([4, 3, 0], False)):
            if x in keys:
                continue
            q = G.Query(types, keys, [(G.COUNT_DISTINCT, x), (G.COUNT_STAR, None), (G.MIN, 1)],
                        drop_null_keys=drop)
            dicts = {i: cols[c].dictionary for i, c in enumerate(q.group_cols) if cols[c].type == G.VARCHAR}
            got, t, _ = exact_rows(q, cols, dicts)
            exp = brute(q, cols)
            if not keys and not exp:
                exp = [(0, 0, None)]
            # keys as SQL groups them, and sortable: -0.0 as 0.0, NaN as +inf (absent here)
            norm = lambda rows: [tuple(float("inf") if isinstance(v, float) and v != v else (0.0 if v == 0.0 and
                                       isinstance(v, float) else v) for v in r[:len(keys)]) + tuple(r[len(keys):])
                                 for r in rows]
            assert_same_groups(norm(got), norm(exp), len(keys), t)


def test_count_distinct_vs_pandas_nunique():
    import pandas as pd
    rng = np.random.default_rng(11)
    n = 200_000
    k = rng.integers(0, 300, n).astype(np.int32)
    x = rng.integers(-5000, 5000, n)
    xv = (rng.random(n) > 0.1).astype(np.uint8)
    q = G.Query([G.INTEGER, G.BIGINT], [0], [(G.COUNT_DISTINCT, 1)], nullable=[False, True])
    got, _, _ = exact_rows(q, [G.Column(G.INTEGER, k), G.Column(G.BIGINT, x, xv)])
    exp = pd.DataFrame({"k": k, "x": np.where(xv == 1, x, np.nan)}).groupby("k")["x"].nunique()
    assert dict(got) == {int(a): int(b) for a, b in exp.items()}


def test_planted_mistake_is_caught():
    """The comparison helper flags an oracle result with one SUM off by one."""
    cols = tiny_table(3)
    q = G.Query([c.type for c in cols], [0], [(G.SUM, 1)])
    c, v, t, _ = O.groupby_exact(q, cols)
    good = oracle_rows(c, v, t)
    i = next(r for r in range(len(good)) if good[r][1] is not None)
    c[1][i] += 1
    bad = oracle_rows(c, v, t)
    with pytest.raises(AssertionError):
        assert_same_groups(bad, good, 1, t)


def test_neumaier_sum():
    x = np.array([1e16, 1.0, -1e16, 1.0] * 1000)
    q = G.Query([G.INTEGER, G.DOUBLE], [0], [(G.SUM, 1), (G.AVG, 1)], nullable=[0, 0])
    rows, _, _ = exact_rows(q, [G.Column(G.INTEGER, np.zeros(len(x), np.int32)), G.Column(G.DOUBLE, x)])
    assert rows == [(0, 2000.0, 0.5)]  # naive left-to-right summation gives 0


def _special_values_table(seed, n=3000, groups=40):
    """Doubles with NaN (both signs), +-Inf, zeros of both signs, subnormals
    and wide magnitudes, grouped so that some groups hold only NaN, only
    infinities, both infinities, or no special at all."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, groups, n).astype(np.int32)
    x = rng.normal(0, 1, n) * 10.0 ** rng.integers(-30, 30, n)
    kind = rng.integers(0, 20, n)
    neg_nan = np.frombuffer(np.uint64(0xFFF8000000000001).tobytes(), np.float64)[0]
    x[kind == 0] = np.nan
    x[kind == 1] = neg_nan
    x[kind == 2] = np.inf
    x[kind == 3] = -np.inf
    x[kind == 4] = -0.0
    x[kind == 5] = 5e-324
    x[k == 0] = np.nan                       # only NaN
    x[k == 1] = np.where(np.arange((k == 1).sum()) % 2, neg_nan, np.nan)
    x[k == 2] = np.inf                       # only +Inf
    x[(k == 3) & (kind < 10)] = np.inf       # +Inf and finite
    x[(k == 4) & (kind < 10)] = -np.inf      # -Inf and finite
    xs = x[k == 5]
    xs[:2] = [np.inf, -np.inf]               # both infinities
    x[k == 5] = xs
    x[(k >= 6) & (k < 12)] = rng.normal(0, 1, ((k >= 6) & (k < 12)).sum())  # no specials
    return k, x


def _brute_sum(vals):
    """IEEE result of the exact sum: NaN for a NaN or +Inf with -Inf, the
    infinity otherwise, else the correctly rounded exact sum (math.fsum)."""
    import math
    if any(math.isnan(v) for v in vals) or (math.inf in vals and -math.inf in vals):
        return math.nan
    if math.inf in vals:
        return math.inf
    if -math.inf in vals:
        return -math.inf
    return math.fsum(vals)


def _brute_minmax(vals, mn):
    """MIN/MAX skipping NaN, NaN when every value is NaN."""
    import math
    num = [v for v in vals if not math.isnan(v)]
    if not num:
        return math.nan
    return min(num) if mn else max(num)


def test_double_specials_and_nan_min_max_vs_brute_force():
    import math
    k, x = _special_values_table(5)
    q = G.Query([G.INTEGER, G.DOUBLE], [0], [(G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 1), (G.COUNT, 1)],
                nullable=[0, 0])
    rows, _, _ = exact_rows(q, [G.Column(G.INTEGER, k), G.Column(G.DOUBLE, x)])
    by = defaultdict(list)
    for kk, xx in zip(k.tolist(), x.tolist()):
        by[kk].append(xx)

    def same(a, b):
        return (math.isnan(a) and math.isnan(b)) or a == b or abs(a - b) <= 1e-12 * abs(b)

    assert len(rows) == len(by)
    for key, sm, avg, mn, mx, cnt in rows:
        vals = by[key]
        assert cnt == len(vals)
        exp = _brute_sum(vals)
        assert same(sm, exp), (key, sm, exp)
        assert same(avg, exp / len(vals)), (key, avg)
        assert same(mn, _brute_minmax(vals, True)) and same(mx, _brute_minmax(vals, False)), (key, mn, mx)
    got = {r[0]: r for r in rows}
    assert math.isnan(got[0][1]) and math.isnan(got[0][3]) and math.isnan(got[1][4])
    assert got[2][1] == math.inf and math.isnan(got[5][1])


def test_planted_nan_policy_mistake_is_caught():
    """A MIN that lets a NaN win (the reference's first-value behaviour) is
    not what the brute-force rule accepts."""
    import math
    vals = [math.nan, 3.0, -1.0]
    assert _brute_minmax(vals, True) == -1.0
    assert math.isnan(vals[0]) and _brute_minmax(vals, False) == 3.0


def test_int64_sum_beyond_2_53_is_exact():
    v = np.array([2**53 + 1, 2**53 + 1, -3], np.int64)
    q = G.Query([G.INTEGER, G.BIGINT], [0], [(G.SUM, 1)], nullable=[0, 0])
    rows, _, ov = exact_rows(q, [G.Column(G.INTEGER, np.zeros(3, np.int32)), G.Column(G.BIGINT, v)])
    assert rows == [(0, 2**54 - 1)] and not ov


# ---- the reference's compiled aggregators (golden fixtures) ----------------

def test_inline_row_agg_golden():
    from tests.golden.make_golden import inline_columns
    g = load("inline_row_agg.json")
    a, b, v0, v1, m0, m1 = inline_columns(**g["params"])
    cols = [G.Column(G.INTEGER, a), G.Column(G.INTEGER, b), G.Column(G.BIGINT, v0, m0), G.Column(G.BIGINT, v1, m1)]
    q = G.Query([G.INTEGER, G.INTEGER, G.BIGINT, G.BIGINT], [0, 1],
                [(G.COUNT_STAR, None), (G.SUM, 2), (G.SUM, 3), (G.COUNT, 2), (G.COUNT, 3)],
                nullable=[0, 0, 1, 1])
    rows, _, _ = exact_rows(q, cols)
    # InlineRowAgg reports SUM 0 where the oracle reports NULL (no non-NULL input)
    got = sorted((r[0], r[1], r[2], r[3] or 0, r[4] or 0, r[5], r[6]) for r in rows)
    assert got == [tuple(r) for r in g["rows"]]


def test_radix_count_golden():
    g = load("radix_count.json")
    p = g["params"]
    k = datagen.highcard(p["n"], p["nkeys"], p["s"], seed=p["seed"], columns=("k",))["k"]
    q = G.Query([G.BIGINT], [0], [(G.COUNT_STAR, None)], nullable=[0])
    c, v, t, _ = O.groupby_exact(q, [G.Column(G.BIGINT, k)])
    keys, counts = c[0], c[1]
    assert len(keys) == g["groups"] and int(counts.sum()) == g["rows"]
    assert int(np.bitwise_xor.reduce((keys * np.int64(1000003)) ^ counts)) == g["xor_checksum"]
    assert sorted(counts.tolist(), reverse=True)[:50] == g["top50_counts"]
    order = np.lexsort((keys, counts))
    assert [[int(keys[i]), int(counts[i])] for i in order[-200:]] == g["all_pairs_sorted_head"]
    # the restated RadixCountAgg gives the same groups and top-k counts
    pk, pc = O.port_radix_count(k, 8, 0)
    assert sorted(zip(pk.tolist(), pc.tolist())) == sorted(zip(keys.tolist(), counts.tolist()))
    _, tc = O.port_radix_count(k, 8, 50)
    assert tc.tolist() == g["top50_counts"]


def test_plain_1m_golden_pandas():
    g = load("plain_1m.json")
    q, cols = datagen.plain_workload(1_000_000)
    rows, t, _ = exact_rows(q, cols)
    assert sorted(rows) == [tuple(r) for r in g["rows"]]


def test_port_equals_exact_where_double_is_exact():
    for q, cols in (datagen.plain_workload(300_000), datagen.highcard_workload(300_000, 50_000)):
        exp, t, _ = exact_rows(q, cols)
        c, v, t2, _ = O.port_groupby(q, cols)
        got = oracle_rows(c, v, t2)
        assert_same_groups(got, exp, 1, t, rel=1e-9)


def test_port_topn_count_desc():
    q, cols = datagen.highcard_workload(200_000, 20_000)
    q.topn = (1, False, 10)
    c, v, t, _ = O.port_groupby(q, cols)
    got = sorted(c[1].tolist(), reverse=True)
    q.topn = None
    allc, _, _, _ = O.groupby_exact(q, cols)
    assert got == sorted(allc[1].tolist(), reverse=True)[:10]


def test_oracle_topn_tie_break():
    q = G.Query([G.INTEGER], [0], [(G.COUNT_STAR, None)], nullable=[0])
    k = np.array([5, 5, 3, 3, 9, 1, 1, 7], np.int32)
    res = O.OrcResult()
    import ctypes as C
    qd, keep = q.desc()
    views, keep2 = G._views([G.Column(G.INTEGER, k)], False)
    O.lib().orc_groupby_exact(C.byref(qd), views, len(k), 1, C.byref(res))
    idx = np.zeros(4, np.uint64)
    m = O.lib().orc_topn(C.byref(res), 1, 1, 0, 4, idx.ctypes.data)
    c, v, t = O._to_table(res)
    O.lib().orc_result_free(C.byref(res))
    assert [(int(c[0][i]), int(c[1][i])) for i in idx[:m]] == [(1, 2), (3, 2), (5, 2), (7, 1)]



def _dkey(x):
    """SQL grouping of a double: -0.0 with 0.0, every NaN together."""
    if x != x:
        return "nan"
    return 0.0 if x == 0.0 else x


@pytest.mark.parametrize("drop", [False, True])
def test_multi_column_and_double_keys_vs_brute_force(drop):
    # GROUP BY (INTEGER, BIGINT, DOUBLE, VARCHAR) with NULLs, -0.0 and NaN
    # keys; brute force groups Python tuples
    rng = np.random.default_rng(21)
    n = 4000
    a = rng.integers(-3, 3, n).astype(np.int32)
    b = rng.integers(0, 4, n) * (1 << 40)
    d = rng.choice(np.array([0.0, -0.0, 1.5, np.nan, -2.25]), n)
    words = G.Dictionary(["x", "y", "z", "x"])  # duplicate entry: same string, one group
    s = rng.integers(0, 4, n).astype(np.uint32)
    v = rng.integers(-1000, 1000, n)
    va, vd, vs = (rng.random(n) > 0.1).astype(np.uint8), (rng.random(n) > 0.1).astype(np.uint8), \
        (rng.random(n) > 0.1).astype(np.uint8)
    cols = [G.Column(G.INTEGER, a, va), G.Column(G.BIGINT, b), G.Column(G.DOUBLE, d, vd),
            G.Column(G.VARCHAR, s, vs, dictionary=words), G.Column(G.BIGINT, v)]
    q = G.Query([c.type for c in cols], [0, 1, 2, 3], [(G.COUNT_STAR, None), (G.SUM, 4), (G.MIN, 4)],
                nullable=[True, False, True, True, False], drop_null_keys=drop)
    c_, v_, t, ov = O.groupby_exact(q, cols)
    got = {}
    for row in oracle_rows(c_, v_, t, {3: words}):
        key = (row[0], row[1], None if row[2] is None else _dkey(row[2]), row[3])
        assert key not in got
        got[key] = row[4:]
    exp = defaultdict(lambda: [0, 0, None])
    for r in range(n):
        key = (int(a[r]) if va[r] else None, int(b[r]), _dkey(float(d[r])) if vd[r] else None,
               words.get(int(s[r])) if vs[r] else None)
        if drop and None in key:
            continue
        e = exp[key]
        e[0] += 1
        e[1] += int(v[r])
        e[2] = int(v[r]) if e[2] is None else min(e[2], int(v[r]))
    assert got == {k: tuple(e) for k, e in exp.items()}


def test_varchar_min_max_vs_brute_force():
    # MIN/MAX of strings in byte order (prefixes first, non-ASCII after ASCII),
    # a dictionary with a repeated entry, NULL arguments, an all-NULL group
    rng = np.random.default_rng(23)
    words = ["b", "a", "ab", "", "zz", "é", "Z", "a"]
    dic = G.Dictionary(words)
    n = 3000
    k = rng.integers(0, 40, n).astype(np.int32)
    s = rng.integers(0, len(words), n).astype(np.uint32)
    valid = ((rng.random(n) > 0.2) & (k != 7)).astype(np.uint8)
    cols = [G.Column(G.INTEGER, k), G.Column(G.VARCHAR, s, valid, dictionary=dic)]
    q = G.Query([G.INTEGER, G.VARCHAR], [0], [(G.MIN, 1), (G.MAX, 1), (G.COUNT, 1)], nullable=[False, True])
    c_, v_, t, ov = O.groupby_exact(q, cols)
    got = {r[0]: r[1:] for r in oracle_rows(c_, v_, t, {1: dic, 2: dic})}
    exp = {}
    for key in range(40):
        vals = [words[s[r]].encode() for r in range(n) if k[r] == key and valid[r]]
        exp[key] = (min(vals).decode(), max(vals).decode(), len(vals)) if vals else (None, None, 0)
    assert got == exp
