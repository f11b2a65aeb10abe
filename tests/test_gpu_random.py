"""Seeded random queries against the oracle: random schemas (INTEGER, BIGINT,
DOUBLE, VARCHAR as dictionary codes or as strings), 0-4 group columns of any
of those types, 1-4 aggregates, 0-2 predicates, NULLs, drop_null_keys,
forced strategies, and several host or device batches. Each case prints its
number and seed before it runs."""
import numpy as np
import pytest

from slothdb_amd import gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, oracle_rows
from tests.test_gpu_strings import _as_dict, _split

pytestmark = pytest.mark.gpu

WORDS = [f"w{i:03d}" + "z" * (i % 7) for i in range(400)] + ["", "a", "ab"]


def _column(rng, t, n, null_p, raw):
    valid = (rng.random(n) >= null_p).astype(np.uint8) if null_p else None
    if t == G.INTEGER:
        hi = int(rng.choice([3, 50, 100_000]))
        return G.Column(t, rng.integers(-hi, hi, n).astype(np.int32), valid)
    if t == G.BIGINT:  # |value| <= 1e9: sums of 200K rows stay far inside int64
        hi, mult = [(5, 1), (1000, 1_000_003), (10**9, 1)][int(rng.integers(0, 3))]
        return G.Column(t, rng.integers(-hi, hi, n) * mult, valid)
    if t == G.DOUBLE:
        if rng.random() < 0.5:  # few values, -0.0 among them
            return G.Column(t, rng.choice(np.array([0.0, -0.0, 1.5, 2.25, 1e10]), n), valid)
        return G.Column(t, rng.uniform(1, 100, n), valid)
    m = int(rng.choice([3, 40, len(WORDS)]))
    codes = rng.integers(0, m, n).astype(np.uint32)
    if raw:
        ok = np.ones(n, bool) if valid is None else valid.astype(bool)
        return G.Column.strings([WORDS[c] if ok[i] else None for i, c in enumerate(codes)])
    return G.Column(t, codes, valid, dictionary=G.Dictionary(WORDS[:m], unique=True))


def _case(seed):
    rng = np.random.default_rng(seed)
    n = int(np.exp(rng.uniform(np.log(1), np.log(200_000))))
    ncol = int(rng.integers(3, 7))
    types = [int(rng.choice([G.INTEGER, G.BIGINT, G.DOUBLE, G.VARCHAR])) for _ in range(ncol)]
    types[-1] = G.BIGINT  # one numeric column for SUM/MIN/MAX
    raw = rng.random() < 0.5
    nulls = [float(rng.choice([0.0, 0.0, 0.05, 0.3])) for _ in range(ncol)]
    cols = [_column(rng, t, n, p, raw) for t, p in zip(types, nulls)]
    ng = int(rng.integers(0, min(4, ncol) + 1))
    gcols = [int(c) for c in rng.permutation(ncol)[:ng]]
    numeric = [c for c in range(ncol) if types[c] != G.VARCHAR]
    aggs = [(G.COUNT_STAR, None)]
    for _ in range(int(rng.integers(0, 4))):
        kind = int(rng.choice([G.COUNT, G.SUM, G.AVG, G.MIN, G.MAX]))
        col = int(rng.choice(range(ncol) if kind in (G.COUNT, G.MIN, G.MAX) else numeric))
        aggs.append((kind, col))
    preds = []
    for _ in range(int(rng.integers(0, 3))):
        c = int(rng.integers(0, ncol))
        if types[c] == G.VARCHAR:
            preds.append((c, int(rng.choice([G.EQ, G.NE, G.LIKE, G.NOT_LIKE])), str(rng.choice(["w01", "a", "z", ""]))))
        elif types[c] == G.DOUBLE:
            preds.append((c, int(rng.choice([G.LT, G.GE])), float(rng.choice([1.0, 50.0]))))
        else:
            preds.append((c, int(rng.choice([G.LT, G.GT, G.NE])), int(rng.integers(-3, 4))))
    strategy = int(rng.choice([G.AUTO, G.AUTO, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX]))
    q = G.Query(types, gcols, aggs, predicates=preds, nullable=[p > 0 for p in nulls],
                drop_null_keys=bool(rng.random() < 0.3), strategy=strategy)
    return q, cols, int(rng.integers(1, 4)), bool(rng.random() < 0.5)


def _dicts(q, dcols):
    """Dictionaries of the output columns holding VARCHAR codes in the oracle's
    table: VARCHAR group columns and MIN/MAX over VARCHAR."""
    d = {g: dcols[c].dictionary for g, c in enumerate(q.group_cols) if dcols[c].dictionary is not None}
    for i, (kind, col) in enumerate(q.aggs):
        if kind in (G.MIN, G.MAX) and dcols[col].dictionary is not None:
            d[len(q.group_cols) + i] = dcols[col].dictionary
    return d


def _norm(rows, nk):
    def f(x):
        return (0.0 if x == 0.0 else x) if isinstance(x, float) else x
    return [tuple(f(x) for x in r[:nk]) + tuple(r[nk:]) for r in rows]


def test_random_queries_against_oracle():
    ran = 0
    for case in range(150):
        seed = 1000 + case
        q, cols, nbatch, device = _case(seed)
        print(f"[random] case {case} seed {seed}: rows {len(cols[0])} types {q.column_types} groups {q.group_cols} "
              f"aggs {q.aggs} preds {q.predicates} drop {q.drop_null_keys} strategy {q.strategy} "
              f"batches {nbatch} device {device}", flush=True)
        op = G.GpuHashAggregate(q)
        try:
            op.Init()
        except G.GbaggError as e:  # plans outside the operator's limits (e.g. more than 8 input columns used)
            print(f"[random] case {case}: plan refused: {e}", flush=True)
            op.close()
            continue
        keep = []
        for b in _split(cols, nbatch):
            if device:
                dc = [G.DeviceColumn(c) for c in b]
                keep.append(dc)
                op.consume_device(dc)
            else:
                op.consume(b)
        op.finalize()
        got = op.result().rows()
        op.close()
        dcols = [_as_dict(c) if c.raw is not None else c for c in cols]
        dicts = _dicts(q, dcols)
        c_, v_, t, ov = O.groupby_exact(q, dcols)
        assert not ov
        exp = oracle_rows(c_, v_, t, dicts)
        nk = len(q.group_cols)
        assert_same_groups(_norm(got, nk), _norm(exp, nk), nk, t)
        ran += 1
    assert ran >= 120


def test_random_topn_and_count_distinct():
    from tests._util import assert_topn_tie_aware
    ran = 0
    for case in range(80):
        seed = 5000 + case
        q, cols, nbatch, device = _case(seed)
        rng = np.random.default_rng(seed + 1)
        types = q.column_types
        nk = len(q.group_cols)
        if case % 2 == 0:
            # ORDER BY COUNT(*) or COUNT(c) (never NULL), either direction
            counts = [nk + i for i, (kind, _) in enumerate(q.aggs) if kind in (G.COUNT_STAR, G.COUNT)]
            q.topn = (int(rng.choice(counts)), bool(rng.random() < 0.3), int(rng.choice([1, 5, 50])))
        else:
            # COUNT(DISTINCT x) of any type next to COUNT(*) and COUNT(x), by
            # zero to three group columns of any type, half of them with a
            # second COUNT(DISTINCT) argument
            x = int(rng.integers(0, len(types)))
            rest = [c for c in range(len(types)) if c != x]
            keys = [int(c) for c in rng.permutation(rest)[:int(rng.integers(0, min(3, len(rest)) + 1))]]
            q.group_cols = keys
            nk = len(keys)
            q.aggs = [(G.COUNT_DISTINCT, x), (G.COUNT_STAR, None), (G.COUNT, x)][:int(rng.integers(1, 4))]
            q.drop_null_keys = bool(rng.random() < 0.5)
            others = [c for c in rest if c not in keys]
            if others and rng.random() < 0.5:  # a second argument column: a distinct set of its own
                q.aggs.append((G.COUNT_DISTINCT, int(rng.choice(others))))
        print(f"[random] case {case} seed {seed}: rows {len(cols[0])} types {types} groups {q.group_cols} "
              f"aggs {q.aggs} preds {q.predicates} drop {q.drop_null_keys} topn {q.topn} "
              f"batches {nbatch} device {device}", flush=True)
        op = G.GpuHashAggregate(q)
        try:
            op.Init()
        except G.GbaggError as e:
            print(f"[random] case {case}: plan refused: {e}", flush=True)
            op.close()
            continue
        keep = []
        for b in _split(cols, nbatch):
            if device:
                dc = [G.DeviceColumn(c) for c in b]
                keep.append(dc)
                op.consume_device(dc)
            else:
                op.consume(b)
        op.finalize()
        got = op.result().rows()
        op.close()
        topn = q.topn
        q.topn = None
        dcols = [_as_dict(c) if c.raw is not None else c for c in cols]
        dicts = _dicts(q, dcols)
        c_, v_, t, ov = O.groupby_exact(q, dcols)
        assert not ov
        exp = _norm(oracle_rows(c_, v_, t, dicts), nk)
        got = _norm(got, nk)
        if topn:
            col, asc, k = topn
            assert len(got) == min(k, len(exp))
            if got:
                assert_topn_tie_aware(got, exp, order_col=col, k=min(k, len(exp)), types=t, descending=not asc)
        else:
            assert_same_groups(got, exp, nk, t)
        ran += 1
    assert ran >= 60


def test_random_params_and_row_limit():
    # the same random queries under other knobs: DIRECT forced, generic row
    # kernels only, 4-bit string/tuple hashes (shared hashes everywhere), and
    # a bare LIMIT (any `row_limit` groups: a subset of the oracle's groups)
    ran = 0
    for case in range(60):
        seed = 9000 + case
        q, cols, nbatch, device = _case(seed)
        rng = np.random.default_rng(seed + 7)
        params = {}
        knob = case % 4
        if knob == 0:
            q.strategy = G.DIRECT
        elif knob == 1:
            params["static_shapes"] = 0
        elif knob == 2:
            params["str_hash_bits"] = 4
        else:
            q.row_limit = int(rng.choice([1, 7, 100]))
        print(f"[random] case {case} seed {seed}: rows {len(cols[0])} types {q.column_types} groups {q.group_cols} "
              f"aggs {q.aggs} preds {q.predicates} drop {q.drop_null_keys} strategy {q.strategy} params {params} "
              f"limit {q.row_limit} batches {nbatch} device {device}", flush=True)
        op = G.GpuHashAggregate(q, params)
        try:
            op.Init()
        except G.GbaggError as e:
            print(f"[random] case {case}: plan refused: {e}", flush=True)
            op.close()
            continue
        keep = []
        for b in _split(cols, nbatch):
            if device:
                dc = [G.DeviceColumn(c) for c in b]
                keep.append(dc)
                op.consume_device(dc)
            else:
                op.consume(b)
        op.finalize()
        got = op.result().rows()
        op.close()
        limit = q.row_limit
        q.row_limit = 0
        dcols = [_as_dict(c) if c.raw is not None else c for c in cols]
        dicts = _dicts(q, dcols)
        c_, v_, t, ov = O.groupby_exact(q, dcols)
        assert not ov
        nk = len(q.group_cols)
        exp = _norm(oracle_rows(c_, v_, t, dicts), nk)
        got = _norm(got, nk)
        if limit:
            assert len(got) == min(limit, len(exp))
            by_key = {r[:nk]: r for r in exp}
            assert_same_groups(got, [by_key[r[:nk]] for r in got], nk, t)
        else:
            assert_same_groups(got, exp, nk, t)
        ran += 1
    assert ran >= 45


def test_random_large_queries():
    # fewer, larger cases: 1-1.5M rows, so RADIX partitions, heavy hitters and
    # the two interning stages run at size
    for case in range(6):
        seed = 12000 + case
        q, cols, nbatch, device = _case(seed)
        rng = np.random.default_rng(seed)
        n = int(rng.integers(1_000_000, 1_500_000))
        cols = [_column(rng, t, n, 0.05 if q.nullable[i] else 0.0, bool(case % 2)) for i, t in enumerate(q.column_types)]
        print(f"[random] large case {case} seed {seed}: rows {n} types {q.column_types} groups {q.group_cols} "
              f"aggs {q.aggs} preds {q.predicates} drop {q.drop_null_keys} strategy {q.strategy}", flush=True)
        op = G.GpuHashAggregate(q)
        op.Init()
        keep = []
        for b in _split(cols, nbatch):
            dc = [G.DeviceColumn(c) for c in b]
            keep.append(dc)
            op.consume_device(dc)
        op.finalize()
        got = op.result().rows()
        op.close()
        dcols = [_as_dict(c) if c.raw is not None else c for c in cols]
        dicts = _dicts(q, dcols)
        c_, v_, t, ov = O.groupby_exact(q, dcols)
        assert not ov
        nk = len(q.group_cols)
        assert_same_groups(_norm(got, nk), _norm(oracle_rows(c_, v_, t, dicts), nk), nk, t)
