"""Double SUM/AVG are exact fixed-point sums on the GPU: bit-identical across
runs, strategies, batch splits and row orders; the reference's cancellation
case (test_oracle's 1e16, 1, -1e16, 1) exact under every strategy; IEEE
specials and MIN/MAX over NaN as the oracle; a value beyond the sampled bound
re-runs with the exact bound; a group whose small inputs were rounded re-runs
with IEEE atomics. Every result is also compared with the oracle (1e-12)."""
import numpy as np
import pytest

from oracle import oracle as O
from slothdb_amd import datagen, gbagg as G
from tests._util import assert_same_groups, oracle_rows
from tests.test_oracle import _special_values_table

pytestmark = pytest.mark.gpu

STRATEGIES = [G.AUTO, G.DIRECT, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX]


def _exact(q, cols):
    c, v, t, ov = O.groupby_exact(q, cols)
    assert not ov
    return oracle_rows(c, v, t), t


def _run(q, cols, strategy=G.AUTO, batches=1, device=True):
    q.strategy = strategy
    op = G.GpuHashAggregate(q)
    op.Init()
    n = len(cols[0].values)
    cuts = np.linspace(0, n, batches + 1).astype(np.int64)
    for i in range(batches):
        part = [G.Column(c.type, c.values[cuts[i]:cuts[i + 1]],
                         None if c.validity is None else c.validity[cuts[i]:cuts[i + 1]]) for c in cols]
        if device:
            op.consume_device([G.DeviceColumn(c) for c in part])
        else:
            op.consume(part)
    op.finalize()
    res, info = op.result(), op.plan_info()
    op.close()
    q.strategy = G.AUTO
    return res, info


def _bits(res, nkeys=1):
    """Result columns as raw bytes, rows in key order."""
    order = np.lexsort([np.asarray(res.columns[k]) for k in reversed(range(nkeys))])
    return [np.asarray(c)[order].tobytes() for c in res.columns]


def test_reference_cancellation_every_strategy():
    # the oracle test's 1e16, 1, -1e16, 1 pattern (left-to-right double sums
    # give 0), in one group next to 300 ordinary groups whose values (quarters)
    # sit on the scale's grid; groups of much smaller rounded values next to
    # 1e16 take the IEEE fallback (test_rounded_small_group_falls_back...)
    rng = np.random.default_rng(1)
    x = np.concatenate([np.array([1e16, 1.0, -1e16, 1.0] * 1000), rng.integers(-4000, 4000, 60_000) / 4.0])
    k = np.concatenate([np.zeros(4000, np.int32), rng.integers(1, 301, 60_000).astype(np.int32)])
    q = G.Query([G.INTEGER, G.DOUBLE], [0], [(G.SUM, 1), (G.AVG, 1), (G.COUNT_STAR, None)], nullable=[0, 0])
    cols = [G.Column(G.INTEGER, k), G.Column(G.DOUBLE, x)]
    exp, types = _exact(q, cols)
    for st in STRATEGIES:
        print(f"[fsum] cancellation, strategy {st}", flush=True)
        res, info = _run(q, cols, st)
        rows = {r[0]: r for r in res.rows()}
        assert rows[0][1] == 2000.0 and rows[0][2] == 0.5, rows[0]
        assert info["sum_ieee"] == 0
        assert_same_groups(res.rows(), exp, 1, types)


def test_double_sums_bit_identical_across_runs_strategies_and_row_orders():
    q, cols = datagen.highcard_workload(n=2_000_000, nkeys=500_000)
    exp, types = _exact(q, cols)
    print("[fsum] reference run (AUTO)", flush=True)
    ref, info = _run(q, cols)
    assert info["sum_ieee"] == 0
    assert_same_groups(ref.rows(), exp, 1, types)
    want = _bits(ref)
    perm = np.random.default_rng(9).permutation(len(cols[0].values))
    shuffled = [G.Column(c.type, c.values[perm]) for c in cols]
    cases = [("again", cols, G.AUTO, 1), ("radix", cols, G.RADIX, 1), ("lds_hash", cols, G.LDS_HASH, 1),
             ("global_hash", cols, G.GLOBAL_HASH, 1), ("7 batches", cols, G.AUTO, 7),
             ("rows permuted", shuffled, G.AUTO, 1), ("rows permuted, radix, 3 batches", shuffled, G.RADIX, 3)]
    for name, cc, st, nb in cases:
        print(f"[fsum] {name}", flush=True)
        res, info = _run(q, cc, st, nb)
        assert info["sum_ieee"] == 0
        assert _bits(res) == want, f"{name}: results differ bitwise"


def test_specials_and_nan_min_max_match_oracle():
    for seed, n in ((5, 3000), (6, 400_000)):
        k, x = _special_values_table(seed, n=n, groups=40 if n < 10_000 else 5000)
        q = G.Query([G.INTEGER, G.DOUBLE], [0],
                    [(G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 1), (G.COUNT, 1)], nullable=[0, 0])
        cols = [G.Column(G.INTEGER, k), G.Column(G.DOUBLE, x)]
        exp, types = _exact(q, cols)
        for st in STRATEGIES:
            print(f"[fsum] specials n={n} strategy {st}", flush=True)
            res, _ = _run(q, cols, st)
            assert_same_groups(res.rows(), exp, 1, types)


def test_value_beyond_sampled_bound_reruns_exactly():
    # one large value on a row the strided sample skips: the sampled bound is
    # exceeded, the query re-runs with the exact bound, results stay exact
    n = 1_000_000
    rng = np.random.default_rng(2)
    k = rng.integers(0, 5000, n).astype(np.int32)
    x = rng.normal(100, 25, n)
    x[1] = 1e6
    q = G.Query([G.INTEGER, G.DOUBLE], [0], [(G.SUM, 1), (G.AVG, 1)], nullable=[0, 0])
    cols = [G.Column(G.INTEGER, k), G.Column(G.DOUBLE, x)]
    exp, types = _exact(q, cols)
    for st in (G.AUTO, G.RADIX, G.LDS_HASH):
        res, info = _run(q, cols, st)
        print(f"[fsum] bound rerun strategy {st}: {info['sum_reruns']} reruns, ieee {info['sum_ieee']}", flush=True)
        assert info["sum_reruns"] == 1 and info["sum_ieee"] == 0
        assert_same_groups(res.rows(), exp, 1, types)


def test_rounded_small_group_falls_back_to_ieee_atomics():
    # group 0 holds values near 1e15, group 1 values near 1e-6: the shared
    # scale rounds group 1's inputs to nothing, its sum is flagged, and the
    # query re-runs with IEEE double atomics
    rng = np.random.default_rng(3)
    n = 200_000
    k = (rng.random(n) < 0.5).astype(np.int32)
    x = np.where(k == 0, rng.normal(1e15, 1e13, n), rng.normal(1e-6, 1e-7, n))
    q = G.Query([G.INTEGER, G.DOUBLE], [0], [(G.SUM, 1), (G.AVG, 1)], nullable=[0, 0])
    cols = [G.Column(G.INTEGER, k), G.Column(G.DOUBLE, x)]
    exp, types = _exact(q, cols)
    res, info = _run(q, cols)
    print(f"[fsum] hazard: {info['sum_reruns']} reruns, ieee {info['sum_ieee']}", flush=True)
    assert info["sum_ieee"] == 1
    assert_same_groups(res.rows(), exp, 1, types)


def test_topn_by_double_sum_and_avg():
    q, cols = datagen.highcard_workload(n=1_000_000, nkeys=200_000)
    full = G.Query(q.column_types, q.group_cols, q.aggs, nullable=[False] * 4)
    exp, types = _exact(full, cols)
    for col, asc in ((5, False), (5, True)):
        qt = G.Query(q.column_types, q.group_cols, q.aggs, nullable=[False] * 4, topn=(col, asc, 10))
        res, info = _run(qt, cols)
        print(f"[fsum] TopN on AVG asc={asc}: strategy {info['strategy']}", flush=True)
        rows = res.rows()
        vals = sorted((r[col] for r in exp), reverse=not asc)[:10]
        got = sorted((r[col] for r in rows), reverse=not asc)
        assert np.allclose(got, vals, rtol=1e-12, atol=0)
        by_key = {r[0]: r for r in exp}
        assert_same_groups(rows, [by_key[r[0]] for r in rows], 1, types)
