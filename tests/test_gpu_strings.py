"""VARCHAR columns given as strings (GBAGG_STR_RAW: offsets + bytes, the
flattened string_t vectors of a DataChunk), interned on the GPU, against the
oracle run on the same strings dictionary-encoded on the host: group keys
alone and with an INTEGER key, NULL strings, VARCHAR predicates, TopN,
COUNT(DISTINCT), several host and device batches, and hash collisions forced
by keeping only a few hash bits (the reference's RadixHashCountStr groups on
its 64-bit string hash, radix_count_agg.cpp:390-761; here groups stay exact
whatever the hash)."""
import numpy as np
import pytest

from slothdb_amd import datagen, gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, assert_topn_tie_aware, oracle_rows

pytestmark = pytest.mark.gpu


def _words(m, seed, maxlen=40):
    """m distinct strings: the empty string, prefixes of one another, strings
    that differ only in their last byte, non-ASCII text, random lengths."""
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz0123456789", np.uint8)
    base = ["", "a", "ab", "abc", "abcd", "abcde", "abcdefgh", "abcdefghi", "x" * 33, "x" * 32,
            "café", "cafe", "日本語", "tail0", "tail1", "tail2"]
    seen = set(base)
    out = list(base)
    while len(out) < m:
        n = int(rng.integers(1, maxlen + 1))
        w = bytes(alpha[rng.integers(0, len(alpha), n)]).decode()
        if w not in seen:
            seen.add(w)
            out.append(w)
    return out[:m]


def _strings(words, idx, null_p=0.0, seed=0):
    rng = np.random.default_rng(seed + 1)
    vals = [words[i] for i in idx]
    if null_p:
        for r in np.nonzero(rng.random(len(idx)) < null_p)[0]:
            vals[r] = None
    return G.Column.strings(vals)


def _as_dict(col):
    """The same strings dictionary-encoded on the host (the oracle's input)."""
    n = len(col)
    valid = np.ones(n, bool) if col.validity is None else col.validity.astype(bool)
    strs = [col.string(i) if valid[i] else None for i in range(n)]
    uniq = sorted({s for s in strs if s is not None})
    pos = {s: i for i, s in enumerate(uniq)}
    codes = np.array([pos[s] if s is not None else 0 for s in strs], np.uint32)
    d = G.Dictionary(uniq or [""], unique=True)
    return G.Column(G.VARCHAR, codes, col.validity, dictionary=d)


def _as_shuffled_dict(col, seed):
    """The strings of a batch as codes into a dictionary of its own (shuffled,
    one entry repeated, so not unique): a batch from another row group."""
    n = len(col)
    valid = np.ones(n, bool) if col.validity is None else col.validity.astype(bool)
    strs = [col.string(i) if valid[i] else None for i in range(n)]
    uniq = sorted({x for x in strs if x is not None})
    np.random.default_rng(seed).shuffle(uniq)
    entries = uniq + uniq[:1]
    pos = {x: i for i, x in enumerate(uniq)}
    codes = np.array([pos[x] if x is not None else 7 for x in strs], np.uint32)
    return G.Column(G.VARCHAR, codes, col.validity, dictionary=G.Dictionary(entries or [""]))


def _expected(q, cols):
    dcols = [_as_dict(c) if c.raw is not None else c for c in cols]
    dicts = {g: dcols[c].dictionary for g, c in enumerate(q.group_cols) if dcols[c].dictionary is not None}
    c, v, t, ov = O.groupby_exact(q, dcols)
    assert not ov
    return oracle_rows(c, v, t, dicts), t


def _run(q, batches, device, params=None):
    op = G.GpuHashAggregate(q, params)
    op.Init()
    dev = []
    for cols in batches:
        if device:
            dc = [G.DeviceColumn(c) for c in cols]
            dev.append(dc)
            op.consume_device(dc)
        else:
            op.consume(cols)
    op.finalize()
    rows = op.result().rows()
    info = op.plan_info()
    op.close()
    return rows, info


def _split(cols, parts):
    n = len(cols[0])
    cuts = np.linspace(0, n, parts + 1).astype(int)
    out = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        piece = []
        for c in cols:
            if c.raw is not None:
                off = c.values[a:b + 1].astype(np.int64)
                piece.append(G.Column(G.VARCHAR, (off - off[0]).astype(np.uint32),
                                      None if c.validity is None else c.validity[a:b],
                                      raw=c.raw[off[0]:off[-1]] if off[-1] > off[0] else np.zeros(1, np.uint8)))
            else:
                piece.append(G.Column(c.type, c.values[a:b], None if c.validity is None else c.validity[a:b],
                                      dictionary=c.dictionary))
        out.append(piece)
    return out


@pytest.mark.parametrize("device", [True, False])
def test_string_key_aggregates(device):
    # GROUP BY s: COUNT(*), SUM(v), MIN(v), MAX(v), AVG(d); 400K rows over 20K strings
    n, m = 400_000, 20_000
    words = _words(m, 1)
    idx = datagen.zipf_ranks(n, m, 0.9, seed=2).astype(np.int64) % m
    rng = np.random.default_rng(3)
    cols = [_strings(words, idx), G.Column(G.BIGINT, rng.integers(-10**9, 10**9, n)),
            G.Column(G.DOUBLE, rng.normal(10, 3, n))]
    q = G.Query([G.VARCHAR, G.BIGINT, G.DOUBLE], [0],
                [(G.COUNT_STAR, None), (G.SUM, 1), (G.MIN, 1), (G.MAX, 1), (G.AVG, 2)], nullable=[False] * 3)
    print(f"[strings] key alone device={device}", flush=True)
    got, info = _run(q, [cols], device)
    exp, t = _expected(q, cols)
    assert_same_groups(got, exp, 1, t)
    assert len(got) == len(set(words[i] for i in idx))


@pytest.mark.parametrize("strategy", [G.AUTO, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX])
def test_int_and_string_key_with_nulls_and_batches(strategy):
    # GROUP BY a INTEGER, s VARCHAR (rows with a NULL string dropped, as the
    # reference's radix paths do), 4 batches
    n, m = 300_000, 5_000
    words = _words(m, 4)
    rng = np.random.default_rng(5)
    idx = rng.integers(0, m, n)
    cols = [G.Column(G.INTEGER, rng.integers(0, 50, n).astype(np.int32)), _strings(words, idx, null_p=0.05, seed=6),
            G.Column(G.INTEGER, rng.integers(0, 100, n).astype(np.int32))]
    q = G.Query([G.INTEGER, G.VARCHAR, G.INTEGER], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2), (G.COUNT, 1)],
                nullable=[False, True, False], strategy=strategy, drop_null_keys=True)
    exp, t = _expected(q, cols)
    for device in (True, False):
        print(f"[strings] composite strategy={strategy} device={device}", flush=True)
        got, _ = _run(q, _split(cols, 4), device)
        assert_same_groups(got, exp, 2, t)
    # batches 1 and 3 as dictionary codes (each its own dictionary): the
    # column is interned across both forms
    parts = _split(cols, 4)
    for i in (1, 3):
        parts[i][1] = _as_shuffled_dict(parts[i][1], i)
    print(f"[strings] composite strategy={strategy} mixed forms", flush=True)
    got, _ = _run(q, parts, True)
    assert_same_groups(got, exp, 2, t)


def test_forced_hash_collisions():
    # keep 1 or 8 hash bits: distinct strings share stored hashes, rows probe
    # on past other strings' codes; groups must not change
    words = _words(3_000, 7)
    rng = np.random.default_rng(8)
    for bits, n, m in ((1, 4_000, 60), (8, 200_000, 3_000)):
        idx = rng.integers(0, m, n)
        cols = [_strings(words, idx), G.Column(G.INTEGER, rng.integers(-5, 5, n).astype(np.int32))]
        q = G.Query([G.VARCHAR, G.INTEGER], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[False, False])
        print(f"[strings] hash bits {bits} rows {n} strings {m}", flush=True)
        got, _ = _run(q, _split(cols, 3), True, params={"str_hash_bits": bits})
        exp, t = _expected(q, cols)
        assert_same_groups(got, exp, 1, t)
        assert len(got) == len(set(idx.tolist()))


@pytest.mark.parametrize("codes", [1, 16])
def test_tables_grow_from_small(codes):
    # first tables of 1 or 16 codes and slots: the dictionary and the slot table
    # fill, interning restarts with larger ones (rows back at their home slots)
    # until every string has its code; NULL strings form their own group
    words = _words(3_000, 21)
    rng = np.random.default_rng(22)
    n = 150_000
    idx = datagen.zipf_ranks(n, len(words), 0.8, seed=23).astype(np.int64) % len(words)
    cols = [_strings(words, idx, null_p=0.05, seed=24), G.Column(G.BIGINT, rng.integers(-10**6, 10**6, n)),
            G.Column(G.INTEGER, rng.integers(0, 7, n).astype(np.int32))]
    q = G.Query([G.VARCHAR, G.BIGINT, G.INTEGER], [2, 0], [(G.COUNT_STAR, None), (G.SUM, 1)],
                nullable=[True, False, False])
    print(f"[strings] first tables of {codes} codes", flush=True)
    got, _ = _run(q, _split(cols, 2), True, params={"str_codes": codes})
    exp, t = _expected(q, cols)
    assert_same_groups(got, exp, 2, t)


def test_long_strings():
    # strings of 65,535 bytes and more: the row key's length field saturates,
    # so with one hash bit kept, strings that differ only in length share keys
    # and the code's own length must separate them
    base = "x" * 65_534
    words = [base, base + "x", base + "xx", base + "xy", "x" * 70_000, "x" * 70_001, "", "x", "y" * 40]
    rng = np.random.default_rng(31)
    n = 400
    idx = rng.integers(0, len(words), n)
    cols = [_strings(words, idx), G.Column(G.BIGINT, rng.integers(-100, 100, n))]
    q = G.Query([G.VARCHAR, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[False, False])
    exp, t = _expected(q, cols)
    for bits in (64, 1):
        print(f"[strings] long strings, hash bits {bits}", flush=True)
        got, _ = _run(q, _split(cols, 2), True, params={"str_hash_bits": bits})
        assert_same_groups(got, exp, 1, t)
        assert len(got) == len(set(idx.tolist()))


def test_string_predicates():
    # WHERE s = / <> / LIKE '%x%' / NOT LIKE over a string column, GROUP BY k
    n, m = 200_000, 2_000
    words = _words(m, 9)
    rng = np.random.default_rng(10)
    idx = rng.integers(0, m, n)
    cols = [G.Column(G.INTEGER, rng.integers(0, 30, n).astype(np.int32)), _strings(words, idx, null_p=0.02, seed=11),
            G.Column(G.BIGINT, rng.integers(0, 1000, n))]
    for pred in [(1, G.EQ, "abcd"), (1, G.NE, "abcd"), (1, G.LIKE, "ab"), (1, G.NOT_LIKE, "a"), (1, G.LIKE, "")]:
        q = G.Query([G.INTEGER, G.VARCHAR, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 2)],
                    predicates=[pred], nullable=[False, True, False])
        print(f"[strings] predicate {pred}", flush=True)
        got, _ = _run(q, _split(cols, 2), True)
        exp, t = _expected(q, cols)
        assert_same_groups(got, exp, 1, t)


def test_topn_and_count_distinct():
    # C3's shape with the strings given as such: WHERE c < 30 GROUP BY a, s
    # ORDER BY COUNT(*) DESC LIMIT 10; and COUNT(DISTINCT s) GROUP BY a
    n, m = 1_000_000, 100_000
    words = _words(m, 12, maxlen=32)
    rng = np.random.default_rng(13)
    idx = datagen.zipf_ranks(n, m, 0.9, seed=14).astype(np.int64) % m
    cols = [G.Column(G.INTEGER, rng.integers(0, 256, n).astype(np.int32)), _strings(words, idx),
            G.Column(G.INTEGER, rng.integers(0, 100, n).astype(np.int32))]
    types = [G.INTEGER, G.VARCHAR, G.INTEGER]
    q = G.Query(types, [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)], predicates=[(2, G.LT, 30)],
                nullable=[False] * 3)
    exp, t = _expected(q, cols)
    q.topn = (2, False, 10)
    print("[strings] topn", flush=True)
    got, _ = _run(q, [cols], True)
    assert len(got) == 10
    assert_topn_tie_aware(got, exp, order_col=2, k=10, types=t)
    qd = G.Query(types, [0], [(G.COUNT_DISTINCT, 1), (G.COUNT_STAR, None)], nullable=[False] * 3)
    print("[strings] count distinct", flush=True)
    got, _ = _run(qd, _split(cols, 3), True)
    exp, t = _expected(qd, cols)
    assert_same_groups(got, exp, 1, t)


def test_empty_null_and_invalid_batches():
    q = G.Query([G.VARCHAR, G.INTEGER], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[True, False])
    # only NULL strings: one NULL group
    cols = [G.Column.strings([None] * 5), G.Column(G.INTEGER, np.arange(5, dtype=np.int32))]
    got, _ = _run(q, [cols], True)
    assert got == [(None, 5, 10)]
    # an empty batch next to a non-empty one; empty strings are a group of their own
    e = [G.Column.strings([]), G.Column(G.INTEGER, np.zeros(0, np.int32))]
    cols = [G.Column.strings(["", "x", "", None]), G.Column(G.INTEGER, np.array([1, 2, 3, 4], np.int32))]
    got, _ = _run(q, [e, cols, e], False)
    assert sorted(got, key=lambda r: (r[0] is not None, r[0] or "")) == [(None, 1, 4), ("", 2, 4), ("x", 1, 2)]
    # strings in one batch, dictionary codes in the next: one set of groups;
    # a code outside its dictionary drops its row, as for dictionary keys
    d = G.Dictionary(["x", "y"], unique=True)
    coded = [G.Column(G.VARCHAR, np.array([0, 1, 0, 5], np.uint32), np.array([1, 1, 0, 1], np.uint8), dictionary=d),
             G.Column(G.INTEGER, np.array([10, 20, 30, 40], np.int32))]
    for device in (False, True):
        got, _ = _run(q, [cols, coded], device)
        assert sorted(got, key=lambda r: (r[0] is not None, r[0] or "")) == \
            [(None, 2, 34), ("", 2, 4), ("x", 2, 12), ("y", 1, 20)]
    # descending offsets: refused on the host (host batch) and by the hash kernel (device batch)
    bad = G.Column(G.VARCHAR, np.array([0, 3, 1, 4], np.uint32), raw=np.frombuffer(b"abcd", np.uint8))
    vals = G.Column(G.INTEGER, np.zeros(3, np.int32))
    q2 = G.Query([G.VARCHAR, G.INTEGER], [0], [(G.COUNT_STAR, None)], nullable=[False, False])
    for device in (False, True):
        with pytest.raises(G.GbaggError, match="ascending"):
            _run(q2, [[bad, vals]], device)
    # the flag on a non-VARCHAR column the query reads
    op = G.GpuHashAggregate(G.Query([G.VARCHAR, G.INTEGER], [0], [(G.SUM, 1)], nullable=[False, False]))
    op.Init()
    with pytest.raises(G.GbaggError, match="for VARCHAR columns"):
        op.consume([G.Column.strings(["a", "b", "c"]), G.Column(G.INTEGER, np.zeros(4, np.uint32).view(np.int32)[:3],
                                                                raw=np.zeros(1, np.uint8))])
    op.close()


def test_varchar_min_max():
    # MIN/MAX of strings (byte order: prefixes first, non-ASCII after ASCII),
    # the column given as strings in some batches and as codes with their own
    # dictionaries in others; NULL strings; a group whose strings are all NULL
    words = _words(3_000, 41)
    rng = np.random.default_rng(42)
    n = 300_000
    k = rng.integers(0, 500, n).astype(np.int32)
    idx = rng.integers(0, len(words), n)
    s = _strings(words, idx, null_p=0.1, seed=43)
    valid = s.validity.copy()
    valid[k == 13] = 0
    s = G.Column.strings([s.string(i) if valid[i] else None for i in range(n)])
    cols = [G.Column(G.INTEGER, k), s, G.Column(G.BIGINT, rng.integers(-50, 50, n))]
    dcols = [cols[0], _as_dict(cols[1]), cols[2]]
    dic = dcols[1].dictionary
    for gcols, aggs in (([0], [(G.MIN, 1), (G.MAX, 1), (G.COUNT, 1), (G.SUM, 2)]),
                        ([], [(G.MIN, 1), (G.MAX, 1)]),
                        ([0, 2], [(G.MAX, 1), (G.COUNT_STAR, None)])):
        q = G.Query([G.INTEGER, G.VARCHAR, G.BIGINT], gcols, aggs, nullable=[False, True, False])
        c_, v_, t, ov = O.groupby_exact(q, dcols)
        outs = {len(gcols) + i: dic for i, (kind, col) in enumerate(aggs) if kind in (G.MIN, G.MAX)}
        exp = oracle_rows(c_, v_, t, outs)
        parts = _split(cols, 4)
        for i in (1, 3):
            parts[i][1] = _as_shuffled_dict(parts[i][1], i)
        for device in (True, False):
            print(f"[strings] MIN/MAX groups {gcols} aggs {aggs} device {device}", flush=True)
            got, _ = _run(q, parts, device)
            assert_same_groups(got, exp, len(gcols), t)
    got, _ = _run(G.Query([G.INTEGER, G.VARCHAR, G.BIGINT], [0], [(G.MIN, 1)], nullable=[False, True, False]),
                  [cols], True)
    assert dict(got)[13] is None
