"""16-byte group keys (two columns, at least one BIGINT; the reference's
InlineRowAggBigKey / RadixMultiAggBigKey, inline_row_agg.cpp:269-517,
radix_multi_agg.cpp:218-225) on every strategy against the oracle, including
keys built to share a key word under the first seed, which must rerun and
still return exact groups."""
import numpy as np
import pytest

from slothdb_amd import datagen, gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, assert_topn_tie_aware, oracle_rows
from tests.test_widekey_mix import wide_mix

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _exact(q, cols, dicts=None):
    c, v, t, ov = O.groupby_exact(q, cols)
    assert not ov
    return oracle_rows(c, v, t, dicts), t


def _mix_arr(x):
    x = x.astype(np.uint64)
    z = x + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _wide_cols(n, distinct, s, seed):
    ranks = datagen.zipf_ranks(n, distinct, s, seed=seed)
    a = _mix_arr(ranks).view(np.int64)
    b = _mix_arr(ranks ^ np.uint64(0x5555)).view(np.int64)
    rng = np.random.default_rng(seed)
    v1 = rng.integers(-10**9, 10**9, n)
    v2 = rng.integers(-2**31, 2**31, n).astype(np.int32)
    v3 = rng.normal(100, 25, n)
    return [G.Column(G.BIGINT, a), G.Column(G.BIGINT, b), G.Column(G.BIGINT, v1),
            G.Column(G.INTEGER, v2), G.Column(G.DOUBLE, v3)]


def _wide_query(strategy=G.AUTO, topn=None):
    return G.Query([G.BIGINT, G.BIGINT, G.BIGINT, G.INTEGER, G.DOUBLE], [0, 1],
                   [(G.COUNT_STAR, None), (G.SUM, 2), (G.MIN, 3), (G.MAX, 3), (G.AVG, 4)],
                   nullable=[False] * 5, strategy=strategy, topn=topn)


@pytest.mark.parametrize("strategy", [G.AUTO, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX])
def test_two_bigint_keys_all_strategies(strategy):
    cols = _wide_cols(1_500_000, 600_000, 1.1, 5)
    q = _wide_query(strategy)
    print(f"[wide] strategy={strategy}", flush=True)
    res, info = G.run(q, cols, device=True)
    exp, types = _exact(q, cols)
    assert_same_groups(res.rows(), exp, 2, types)
    if strategy != G.AUTO:
        assert info["strategy"] == strategy
    if strategy == G.RADIX:
        assert info["heavy_keys"] > 0


def test_wide_key_orders_and_types():
    # (INTEGER, BIGINT), (BIGINT, INTEGER), (VARCHAR, BIGINT), (BIGINT, VARCHAR)
    rng = np.random.default_rng(9)
    n = 300_000
    i32 = rng.integers(-50, 50, n).astype(np.int32)
    i64 = rng.integers(-2**62, 2**62, 3000)[rng.integers(0, 3000, n)]
    words = [f"w{j:05d}" for j in range(400)]
    dic = G.Dictionary(words)
    code = rng.integers(0, len(words), n).astype(np.uint32)
    val = rng.integers(-1000, 1000, n)
    for types, cols, dicts in [
        ([G.INTEGER, G.BIGINT], [G.Column(G.INTEGER, i32), G.Column(G.BIGINT, i64)], None),
        ([G.BIGINT, G.INTEGER], [G.Column(G.BIGINT, i64), G.Column(G.INTEGER, i32)], None),
        ([G.VARCHAR, G.BIGINT], [G.Column(G.VARCHAR, code, dictionary=dic), G.Column(G.BIGINT, i64)], {0: dic}),
        ([G.BIGINT, G.VARCHAR], [G.Column(G.BIGINT, i64), G.Column(G.VARCHAR, code, dictionary=dic)], {1: dic}),
    ]:
        for strategy in (G.LDS_HASH, G.RADIX):
            q = G.Query(types + [G.BIGINT], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2), (G.MIN, 2)],
                        nullable=[False] * 3, strategy=strategy)
            print(f"[wide] types={types} strategy={strategy}", flush=True)
            res, _ = G.run(q, cols + [G.Column(G.BIGINT, val)], device=True)
            exp, t = _exact(q, cols + [G.Column(G.BIGINT, val)], dicts)
            assert_same_groups(res.rows(), exp, 2, t)


@pytest.mark.parametrize("strategy", [G.LDS_HASH, G.GLOBAL_HASH, G.RADIX])
def test_shared_key_words_rerun(strategy):
    # pairs (a, b) and (a ^ mix(b) ^ mix(b'), b') share the key word a ^ mix(b)
    # under the first seed: the emit must see it and the rerun must separate them
    seed0 = 77
    rng = np.random.default_rng(4)
    m = 2000
    a = rng.integers(-2**63, 2**63 - 1, m, dtype=np.int64)
    b = rng.integers(-2**63, 2**63 - 1, m, dtype=np.int64)
    b2 = rng.integers(-2**63, 2**63 - 1, 50, dtype=np.int64)
    a2 = np.array([((int(a[i]) & M64) ^ wide_mix(int(b[i]) & M64, seed0) ^ wide_mix(int(b2[i]) & M64, seed0))
                   for i in range(50)], np.uint64).view(np.int64)
    ka = np.concatenate([a, a2])
    kb = np.concatenate([b, b2])
    idx = rng.integers(0, len(ka), 200_000)
    cols = [G.Column(G.BIGINT, ka[idx]), G.Column(G.BIGINT, kb[idx]),
            G.Column(G.BIGINT, rng.integers(-10**6, 10**6, len(idx)))]
    q = G.Query([G.BIGINT] * 3, [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)], nullable=[False] * 3,
                strategy=strategy)
    res, info = G.run(q, cols, device=True, params={"wide_seed": seed0})
    exp, types = _exact(q, cols)
    assert_same_groups(res.rows(), exp, 2, types)
    assert info["retries"] >= 1
    # the next seed alone separates them: no rerun
    res2, info2 = G.run(q, cols, device=True, params={"wide_seed": seed0 + 1})
    assert_same_groups(res2.rows(), exp, 2, types)
    assert info2["retries"] == 0


def test_wide_key_topn_count_desc():
    cols = _wide_cols(2_000_000, 1_000_000, 0.9, 11)
    q = _wide_query(G.RADIX, topn=(2, False, 10))
    res, info = G.run(q, cols, device=True)
    exp, types = _exact(_wide_query(), cols)
    assert res.num_rows == 10
    assert_topn_tie_aware(res.rows(), exp, order_col=2, k=10, types=types)
