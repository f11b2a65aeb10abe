"""GPU parity: the HIP operator (through the C ABI) against the CPU oracle on
the same seeded inputs. Integer results bit-exact; DOUBLE SUM/AVG within
1e-12 relative of the Kahan-Neumaier oracle."""
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


def _check(q, cols, device=True, dicts=None, params=None):
    print(f"[parity] groups={q.group_cols} aggs={q.aggs} rows={len(cols[0])}", flush=True)
    res, info = G.run(q, cols, device=device, params=params)
    exp, types = _exact(q, cols, dicts)
    worst = assert_same_groups(res.rows(), exp, len(q.group_cols), types)
    return res, info, worst


# ---- the reference's own known-answer tests ------------------------------

def _emp():
    # test_aggregation.cpp:7-14: (dept VARCHAR, salary INTEGER)
    depts = G.Dictionary(["Engineering", "Sales", "Marketing"])
    code = np.array([0, 0, 1, 1, 2], np.uint32)
    salary = np.array([100, 120, 80, 90, 95], np.int32)
    return depts, [G.Column(G.VARCHAR, code, dictionary=depts), G.Column(G.INTEGER, salary)]


def test_emp_scalar_aggregates():
    _, cols = _emp()
    q = G.Query([G.VARCHAR, G.INTEGER], [],
                [(G.COUNT_STAR, None), (G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 1)],
                nullable=[False, False])
    res, _ = G.run(q, cols)
    assert res.rows() == [(5, 485, 97.0, 80, 120)]  # test_aggregation.cpp:16-56


def test_emp_group_by_dept():
    _, cols = _emp()
    q = G.Query([G.VARCHAR, G.INTEGER], [0],
                [(G.COUNT_STAR, None), (G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 1)],
                nullable=[False, False])
    res, _ = G.run(q, cols)
    got = {r[0]: r[1:] for r in res.rows()}
    # test_aggregation.cpp:58-120
    assert got == {"Engineering": (2, 220, 110.0, 100, 120), "Sales": (2, 170, 85.0, 80, 90),
                   "Marketing": (1, 95, 95.0, 95, 95)}


def test_emp_order_by_count_desc_tiebreak():
    _, cols = _emp()
    q = G.Query([G.VARCHAR, G.INTEGER], [0], [(G.COUNT_STAR, None)], nullable=[False, False],
                topn=(1, False, 3))
    res, _ = G.run(q, cols)
    # test_aggregation.cpp:196-212: cnt DESC, dept ASC
    assert res.rows() == [("Engineering", 2), ("Sales", 2), ("Marketing", 1)]


def test_empty_table_count():
    q = G.Query([G.INTEGER], [], [(G.COUNT_STAR, None), (G.SUM, 0)], nullable=[False])
    res, _ = G.run(q, [G.Column(G.INTEGER, np.zeros(0, np.int32))])
    assert res.rows() == [(0, None)]  # test_aggregation.cpp:122-130 + SUM of empty = NULL


def test_parallel_sum_1_to_10000():
    x = np.arange(1, 10001, dtype=np.int32)
    q = G.Query([G.INTEGER], [], [(G.SUM, 0)], nullable=[False])
    res, _ = G.run(q, [G.Column(G.INTEGER, x)])
    assert res.rows() == [(50005000,)]  # test_parallel.cpp:105-123


def test_parallel_50k_five_depts():
    i = np.arange(50000)
    depts = G.Dictionary(["Eng", "Sales", "Marketing", "HR", "Finance"])
    q = G.Query([G.INTEGER, G.VARCHAR, G.INTEGER], [1], [(G.COUNT_STAR, None)], nullable=[False] * 3)
    cols = [G.Column(G.INTEGER, i), G.Column(G.VARCHAR, (i % 5).astype(np.uint32), dictionary=depts),
            G.Column(G.INTEGER, 50000 + (i % 100) * 1000)]
    res, _ = G.run(q, cols, device=False)
    assert sorted(res.rows()) == sorted((d, 10000) for d in ["Eng", "Sales", "Marketing", "HR", "Finance"])


def test_gpu_engine_hash_aggregate():
    k, s, c = G.hash_aggregate_i32([1, 2, 1, 2, 1], [10, 20, 30, 40, 50])
    got = {int(a): (int(b), int(d)) for a, b, d in zip(k, s, c)}
    assert got == {1: (90, 3), 2: (60, 2)}  # test_performance.cpp:153-171


# ---- workloads at oracle-friendly sizes ----------------------------------

def test_plain_1m():
    q, cols = datagen.plain_workload(n=1_000_000)
    res, info, _ = _check(q, cols)
    assert res.num_rows == 1000


def test_plain_host_batches_rowgroups():
    q, cols = datagen.plain_workload(n=500_000)
    op = G.GpuHashAggregate(q)
    op.Init()
    RG = 122_880  # ROW_GROUP_SIZE
    for s in range(0, 500_000, RG):
        op.consume([G.Column(c.type, c.values[s:s + RG]) for c in cols])
    op.finalize()
    exp, types = _exact(q, cols)
    assert_same_groups(op.result().rows(), exp, 1, types)
    op.close()


@pytest.mark.parametrize("shape,strategy", [("highcard", G.AUTO), ("highcard", G.RADIX), ("highcard", G.LDS_HASH),
                                            ("highcard", G.GLOBAL_HASH), ("plain", G.DIRECT), ("plain", G.AUTO)])
def test_many_uneven_batches(shape, strategy):
    # batches of 0 to ~90,000 rows, empty and one-row ones among them, every
    # pass of every strategy one launch over all of them: workgroups whose
    # rows span several batches, and batches spread over several workgroups;
    # device batches, then the same rows as host batches
    n = 1_500_000
    if shape == "highcard":
        q, cols = datagen.highcard_workload(n=n, nkeys=300_000)
    else:
        q, cols = datagen.plain_workload(n=n)
    q.strategy = strategy
    rng = np.random.default_rng(11 + strategy)
    cuts = sorted(set(rng.integers(0, n, 40).tolist()) | {0, 1, 2, 2, 700_000, 700_001, n, n})
    cuts += [n] * 3  # empty batches at the end too
    bounds = list(zip(cuts[:-1], cuts[1:]))
    print(f"[parity] {shape} strategy {strategy}: {len(bounds)} batches, sizes {[b - a for a, b in bounds][:12]}...",
          flush=True)
    exp, types = _exact(q, cols)
    for device in (True, False):
        op = G.GpuHashAggregate(q)
        op.Init()
        for a, b in bounds:
            part = [G.Column(c.type, c.values[a:b]) for c in cols]
            if device:
                op.consume_device([G.DeviceColumn(c) for c in part])
            else:
                op.consume(part)
        op.finalize()
        info = op.plan_info()
        assert info["rows_consumed"] == n
        assert_same_groups(op.result().rows(), exp, 1, types)
        op.close()


@pytest.mark.parametrize("n,nkeys", [(2_000_000, 1_000_000), (4_000_000, 10_000_000)])
def test_highcard(n, nkeys):
    q, cols = datagen.highcard_workload(n=n, nkeys=nkeys)
    res, info, worst = _check(q, cols)
    print(f"highcard n={n} groups={res.num_rows} worst_rel={worst:.2e} info={info}")


@pytest.mark.parametrize("strategy", [G.GLOBAL_HASH, G.LDS_HASH])
def test_strategies_agree(strategy):
    q, cols = datagen.highcard_workload(n=1_000_000, nkeys=100_000)
    q.strategy = strategy
    _check(q, cols)


def test_composite_filter_topn():
    q, cols = datagen.composite_workload(n=2_000_000, dict_size=100_000, topn=10)
    dic = cols[1].dictionary
    res, info = G.run(q, cols, device=True)
    exp, types = _exact(datagen.composite_query(topn=None), cols, {1: dic})
    got = res.rows()
    assert len(got) == 10
    assert_topn_tie_aware(got, exp, order_col=2, k=10, types=types)


def test_composite_full_result():
    q, cols = datagen.composite_workload(n=1_000_000, dict_size=50_000, topn=None)
    _check(q, cols, dicts={1: cols[1].dictionary})


def test_nullable_values_and_keys():
    rng = np.random.default_rng(7)
    n = 300_000
    k = rng.integers(-50, 50, n).astype(np.int64)
    kv = (rng.random(n) > 0.05).astype(np.uint8)
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vv = (rng.random(n) > 0.3).astype(np.uint8)
    d = rng.normal(5, 3, n)
    dv = (rng.random(n) > 0.5).astype(np.uint8)
    cols = [G.Column(G.BIGINT, k, kv), G.Column(G.BIGINT, v, vv), G.Column(G.DOUBLE, d, dv)]
    q = G.Query([G.BIGINT, G.BIGINT, G.DOUBLE], [0],
                [(G.COUNT_STAR, None), (G.COUNT, 1), (G.SUM, 1), (G.AVG, 1), (G.MIN, 1), (G.MAX, 2),
                 (G.SUM, 2), (G.AVG, 2), (G.MIN, 2)])
    res, info, _ = _check(q, cols)
    assert any(r[0] is None for r in res.rows())  # the NULL-key group
    q.drop_null_keys = True
    _check(q, cols)


def test_predicates_int_double_in_like():
    rng = np.random.default_rng(11)
    n = 200_000
    a = rng.integers(0, 1000, n).astype(np.int32)
    b = rng.integers(-5, 5, n).astype(np.int64)
    x = rng.normal(0, 1, n)
    words = G.Dictionary(["apple", "banana", "cherry", "grape", "pineapple", ""])
    s = rng.integers(0, 6, n).astype(np.uint32)
    cols = [G.Column(G.INTEGER, a), G.Column(G.BIGINT, b), G.Column(G.DOUBLE, x),
            G.Column(G.VARCHAR, s, dictionary=words)]
    for preds in ([(0, G.GE, 100), (0, G.LT, 900)], [(1, G.IN, [-3, 0, 4])], [(2, G.GT, 0.25)],
                  [(3, G.LIKE, "apple")], [(3, G.NE, "")], [(3, G.NOT_LIKE, "an"), (1, G.NE, 0)]):
        q = G.Query([G.INTEGER, G.BIGINT, G.DOUBLE, G.VARCHAR], [1],
                    [(G.COUNT_STAR, None), (G.SUM, 0), (G.MAX, 2)], predicates=preds,
                    nullable=[False] * 4)
        _check(q, cols)


def test_sentinel_and_extreme_keys():
    keys = np.array([-2**63, 2**63 - 1, 0, -1, -2**63, 5, 2**63 - 1] * 1000, np.int64)
    vals = np.arange(len(keys), dtype=np.int64)
    q = G.Query([G.BIGINT, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1), (G.MIN, 1)],
                nullable=[False, False])
    _check(q, [G.Column(G.BIGINT, keys), G.Column(G.BIGINT, vals)])


def test_topn_count_desc_matches_exact():
    q, cols = datagen.highcard_workload(n=1_000_000, nkeys=200_000)
    q.topn = (1, False, 10)
    res, info = G.run(q, cols, device=True)
    exp, types = _exact(datagen.highcard_query(), cols)
    assert_topn_tie_aware(res.rows(), exp, order_col=1, k=10, types=types)


# ---- radix-partitioned strategy ------------------------------------------

@pytest.mark.parametrize("strategy", [G.AUTO, G.RADIX])
def test_radix_highcard(strategy):
    q, cols = datagen.highcard_workload(n=3_000_000, nkeys=2_000_000)
    q.strategy = strategy
    res, info, worst = _check(q, cols, params={"selector": 0})  # the default plan's rule
    assert info["strategy"] == G.RADIX
    assert info["heavy_keys"] > 0 and info["cold_rows"] < 3_000_000
    print(f"radix info {info} worst_rel={worst:.2e}")


@pytest.mark.parametrize("params", [
    {"heavy_max": 0},                                  # no heavy hitters: the hot key is cold
    {"part_bits": 0, "lds_kb": 8, "fallback_cap": 1024},  # tiny LDS: most rows take the fallback table, which grows
    {"part_bits": 13},                                 # maximal fanout: many empty partitions
])
def test_radix_params(params):
    q, cols = datagen.highcard_workload(n=500_000, nkeys=200_000)
    q.strategy = G.RADIX
    print(f"[parity] radix params {params}", flush=True)
    res, info = G.run(q, cols, device=True, params=params)
    exp, types = _exact(q, cols)
    assert_same_groups(res.rows(), exp, 1, types)
    print(info)


@pytest.mark.parametrize("part_bits,msplit", [(4, 64), (6, 64), (6, 0), (9, 8192), (13, 8192), (0, 1)])
def test_radix_scatter_offsets_by_multisplit(part_bits, msplit):
    # forced fanouts with the scatter's wave multisplit on: 1, 16 and 64
    # partitions (lanes of a wave share partitions: the multisplit's groups),
    # 512 and 8192; and 64 with it off (the default); Zipf keys make some
    # partitions far larger than others
    q, cols = datagen.highcard_workload(n=600_000, nkeys=2_000)
    q.strategy = G.RADIX
    params = {"part_bits": part_bits, "msplit_fanout": msplit, "selector": 0}
    print(f"[parity] radix multisplit {params}", flush=True)
    res, info = G.run(q, cols, device=True, params=params)
    assert info["radix_bits"] == part_bits
    exp, types = _exact(q, cols)
    assert_same_groups(res.rows(), exp, len(q.group_cols), types)


def test_radix_nullable_keys_and_inputs():
    rng = np.random.default_rng(21)
    n = 400_000
    k = rng.integers(-2**40, 2**40, n // 4).astype(np.int64)[rng.integers(0, n // 4, n)]
    cols = [G.Column(G.BIGINT, k, (rng.random(n) > 0.02).astype(np.uint8)),
            G.Column(G.INTEGER, rng.integers(-1000, 1000, n), (rng.random(n) > 0.3).astype(np.uint8)),
            G.Column(G.DOUBLE, rng.normal(0, 1e6, n), (rng.random(n) > 0.6).astype(np.uint8))]
    for drop in (False, True):
        q = G.Query([G.BIGINT, G.INTEGER, G.DOUBLE], [0],
                    [(G.COUNT_STAR, None), (G.COUNT, 1), (G.SUM, 1), (G.MIN, 1), (G.MAX, 2), (G.AVG, 2)],
                    strategy=G.RADIX, drop_null_keys=drop)
        _check(q, cols)


def test_radix_host_batches():
    q, cols = datagen.highcard_workload(n=600_000, nkeys=300_000)
    q.strategy = G.RADIX
    op = G.GpuHashAggregate(q)
    op.Init()
    RG = 122_880
    for s in range(0, 600_000, RG):
        op.consume([G.Column(c.type, c.values[s:s + RG]) for c in cols])
    op.finalize()
    exp, types = _exact(q, cols)
    assert_same_groups(op.result().rows(), exp, 1, types)
    assert op.plan_info()["strategy"] == G.RADIX
    op.close()


def test_radix_composite_topn_and_full():
    q, cols = datagen.composite_workload(n=3_000_000, dict_size=200_000, topn=10)
    dic = cols[1].dictionary
    exp, types = _exact(datagen.composite_query(topn=None), cols, {1: dic})
    res, info = G.run(q, cols, device=True, params={"selector": 0})  # the default plan's rule
    assert info["strategy"] == G.RADIX and res.num_rows == 10
    assert_topn_tie_aware(res.rows(), exp, order_col=2, k=10, types=types)
    q2 = datagen.composite_query(topn=None)
    res2, info2 = G.run(q2, cols, device=True)
    assert_same_groups(res2.rows(), exp, 2, types)


# The radix count pass launched before the plan is read back (default: inputs
# of 4M rows or more), forced here at smaller sizes with radix_ahead_rows = 0.
@pytest.mark.parametrize("case", ["highcard_auto", "highcard_radix", "highcard_selector", "lowcard_auto",
                                  "composite_topn", "nullable_radix"])
def test_radix_count_ahead_of_plan(case):
    params = {"radix_ahead_rows": 0, "selector": 0}
    print(f"[parity] count ahead: {case}", flush=True)
    if case.startswith("highcard"):
        q, cols = datagen.highcard_workload(n=1_500_000, nkeys=600_000)
        if case == "highcard_radix":
            q.strategy = G.RADIX
        if case == "highcard_selector":
            params.pop("selector")  # the built-in selector may then pick another path
        res, info, _ = _check(q, cols, params=params)
        if case != "highcard_selector":
            assert info["strategy"] == G.RADIX
            # the count's marks stay with the radix plan that used them
            assert info["ms_count"] > 0
    elif case == "lowcard_auto":
        # the device gate says low cardinality: the count's workgroups return
        # at once and the low-cardinality plan runs
        q, cols = datagen.plain_workload(n=1_000_000)
        res, info, _ = _check(q, cols, params=params)
        assert info["strategy"] in (G.DIRECT, G.LDS_HASH)
    elif case == "composite_topn":
        q, cols = datagen.composite_workload(n=2_000_000, dict_size=100_000, topn=10)
        dic = cols[1].dictionary
        res, info = G.run(q, cols, device=True, params=params)
        exp, types = _exact(datagen.composite_query(topn=None), cols, {1: dic})
        assert info["strategy"] == G.RADIX and res.num_rows == 10
        assert_topn_tie_aware(res.rows(), exp, order_col=2, k=10, types=types)
    else:
        rng = np.random.default_rng(23)
        n = 800_000
        k = rng.integers(0, 400_000, n).astype(np.int64)
        kv = (rng.random(n) > 0.02).astype(np.uint8)
        v = rng.integers(-10**9, 10**9, n).astype(np.int64)
        vv = (rng.random(n) > 0.2).astype(np.uint8)
        cols = [G.Column(G.BIGINT, k, kv), G.Column(G.BIGINT, v, vv)]
        q = G.Query([G.BIGINT, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1), (G.MIN, 1), (G.AVG, 1)])
        q.strategy = G.RADIX
        res, info, _ = _check(q, cols, params=params)
        assert any(r[0] is None for r in res.rows())  # the NULL-key group


@pytest.mark.parametrize("k", [10, 32])
def test_radix_topn_all_partitions_produce(k):
    # 8192 partitions each push K candidates: the candidate buffer must hold
    # every producer's list (a group dropped here would change the top K)
    q, cols = datagen.composite_workload(n=2_000_000, dict_size=100_000, topn=k)
    q.strategy = G.RADIX
    res, info = G.run(q, cols, device=True, params={"part_bits": 13})
    assert info["radix_bits"] == 13 and res.num_rows == k
    exp, types = _exact(datagen.composite_query(topn=None), cols, {1: cols[1].dictionary})
    assert_topn_tie_aware(res.rows(), exp, order_col=2, k=k, types=types)


def test_radix_topn_sum_ascending():
    q, cols = datagen.highcard_workload(n=800_000, nkeys=400_000)
    q.strategy = G.RADIX
    q.topn = (2, True, 25)  # ORDER BY SUM(v1) ASC LIMIT 25
    res, info = G.run(q, cols, device=True)
    exp, types = _exact(datagen.highcard_query(), cols)
    assert_topn_tie_aware(res.rows(), exp, order_col=2, k=25, types=types, descending=False)


@pytest.mark.parametrize("static", [1, 0])
def test_static_and_dynamic_kernels_agree(static):
    """Compiled query shapes and the generic row kernels give the same result."""
    for (q, cols) in (datagen.plain_workload(300_000), datagen.highcard_workload(600_000, 200_000),
                      datagen.composite_workload(800_000, dict_size=30_000, topn=None)):
        for strategy in (G.LDS_HASH, G.RADIX):
            q.strategy = strategy
            res, info = G.run(q, cols, device=True, params={"static_shapes": static})
            if static:
                assert info["kernel_shape"] != "dynamic"
            else:
                assert info["kernel_shape"] == "dynamic"
            dicts = {i: cols[c].dictionary for i, c in enumerate(q.group_cols) if cols[c].type == G.VARCHAR}
            exp, types = _exact(q, cols, dicts)
            assert_same_groups(res.rows(), exp, len(q.group_cols), types)
