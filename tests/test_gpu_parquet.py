"""Parquet row-group decode on the GPU (gbagg_parquet_*) against pyarrow's own
reading of the same files (pyarrow writes them: uncompressed, dictionary and
PLAIN pages, data pages V1 and V2, NULLs, several pages and row groups, string
columns whose dictionary overflows into PLAIN pages), and the decoded row
groups fed straight into the operator against the oracle."""
import io

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

from slothdb_amd import gbagg as G
from oracle import oracle as O
from tests._util import assert_same_groups, oracle_rows

pytestmark = pytest.mark.gpu


def _table(n, seed):
    rng = np.random.default_rng(seed)
    words = np.array([f"w{j:04d}" for j in range(500)])
    i64 = rng.integers(-2**40, 2**40, 20_000)[rng.integers(0, 20_000, n)]
    return pa.table({
        "a": pa.array(rng.integers(0, 65536, n).astype(np.int32)),
        "b": pa.array(words[rng.integers(0, 500, n)], mask=rng.random(n) < 0.05),
        "c": pa.array(rng.integers(0, 100, n).astype(np.int32)),
        "k": pa.array(i64, mask=rng.random(n) < 0.1),
        "d": pa.array(rng.normal(100, 25, n), mask=rng.random(n) < 0.2),
        "u": pa.array(rng.integers(-2**62, 2**62, n)),  # all distinct: PLAIN after the dictionary limit
        # many distinct strings: dictionary pages, then PLAIN pages once the dictionary is full
        "s": pa.array([f"s{v:07d}" + "x" * int(v % 23) for v in rng.integers(0, 10**6, n)],
                      mask=rng.random(n) < 0.03),
    }, schema=pa.schema([pa.field("a", pa.int32(), nullable=False), ("b", pa.string()),
                         pa.field("c", pa.int32(), nullable=False), ("k", pa.int64()), ("d", pa.float64()),
                         pa.field("u", pa.int64(), nullable=False), ("s", pa.string())]))


def _write(t, **kw):
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="NONE", **kw)
    return buf.getvalue()


def _expected(col, tcol):
    arr = tcol.combine_chunks() if hasattr(tcol, "combine_chunks") else tcol
    valid = ~np.asarray(arr.is_null().to_numpy(zero_copy_only=False))
    if pa.types.is_string(arr.type):
        return [s if v else None for s, v in zip(arr.to_pylist(), valid)]
    vals = np.asarray(arr.fill_null(0).to_numpy(zero_copy_only=False))
    return vals, valid


def _check_row_group(cols, t, row, n, rg):
    for c, dc in enumerate(cols):
        exp = _expected(c, t.column(c).slice(row, n))
        if dc.type == G.VARCHAR:
            assert dc.strings() == exp, f"column {c} row group {rg}"
        else:
            vals, valid = dc.to_host()
            ev, evalid = exp
            if valid is None:
                assert evalid.all()
                valid = np.ones(n, bool)
            assert np.array_equal(valid.astype(bool), evalid), f"validity of column {c} row group {rg}"
            assert np.array_equal(vals[evalid], ev[evalid].astype(vals.dtype)), f"column {c} row group {rg}"


def _check_file(image, t, rg_rows, resident=False):
    f = G.ParquetFile(image)
    if resident:
        f.upload()
    assert f.num_columns == t.num_columns and f.num_rows == t.num_rows
    assert [c[0] for c in f.columns] == t.column_names
    row = 0
    for rg in range(f.num_row_groups):
        n = f.row_group_rows(rg)
        print(f"[parquet] row group {rg} rows {n}", flush=True)
        _check_row_group(f.decode(rg), t, row, n, rg)
        row += n
    assert row == t.num_rows
    # every row group at once: one concatenated batch, dictionaries merged
    cols = f.decode(0, last=f.num_row_groups)
    for c, dc in enumerate(cols):
        exp = _expected(c, t.column(c))
        if dc.type == G.VARCHAR:
            if dc.dictionary is not None:
                assert len(set(dc.dictionary.get(i) for i in range(len(dc.dictionary)))) == len(dc.dictionary)
            assert dc.strings() == exp, f"column {c}, all row groups"
        else:
            vals, valid = dc.to_host()
            ev, evalid = exp
            valid = np.ones(t.num_rows, bool) if valid is None else valid.astype(bool)
            assert np.array_equal(valid, evalid) and np.array_equal(vals[evalid], ev[evalid].astype(vals.dtype))
    f.close()


# 122,880-row row groups: every call reads few column chunks, and the host
# reads the index run headers; 10,000-row ones: the whole-file call reads 210
# chunks, and the device reads them (k_pq_streams)
@pytest.mark.parametrize("version,resident,rg_rows", [("1.0", False, 122_880), ("2.0", False, 122_880),
                                                      ("1.0", True, 122_880), ("1.0", True, 10_000),
                                                      ("2.0", False, 10_000)])
def test_decode_matches_pyarrow(version, resident, rg_rows):
    t = _table(300_000, 1)
    image = _write(t, row_group_size=rg_rows, data_page_size=32 * 1024, data_page_version=version,
                   dictionary_pagesize_limit=64 * 1024, use_dictionary=["a", "b", "k", "d", "u", "s"])
    _check_file(image, t, rg_rows, resident)
    # s overflows its dictionary: delivered as strings; b stays dictionary codes
    f = G.ParquetFile(image)
    cols = f.decode(0)
    assert cols[6].raw is not None and cols[1].raw is None
    f.close()


def test_plain_only_and_single_row_group():
    t = _table(50_000, 2)
    image = _write(t, row_group_size=1_000_000, use_dictionary=False)
    _check_file(image, t, 1_000_000)
    f = G.ParquetFile(image)
    cols = f.decode(0)
    assert cols[1].raw is not None and cols[6].raw is not None
    f.close()


def test_compressed_chunks_are_rejected():
    t = _table(1000, 3)
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="SNAPPY")
    f = G.ParquetFile(buf.getvalue())
    with pytest.raises(G.GbaggError, match="compressed"):
        f.decode(0)
    f.close()


def _host_columns(t, types, nullable):
    """pyarrow's reading of the table as the oracle's input columns."""
    cols = []
    for c in range(t.num_columns):
        arr = t.column(c).combine_chunks()
        valid = (~np.asarray(arr.is_null().to_numpy(zero_copy_only=False))).astype(np.uint8)
        if pa.types.is_string(arr.type):
            d = arr.dictionary_encode()
            dic = G.Dictionary([str(x) for x in d.dictionary.to_pylist()])
            codes = np.asarray(d.indices.fill_null(0).to_numpy(zero_copy_only=False)).astype(np.uint32)
            cols.append(G.Column(G.VARCHAR, codes, valid, dictionary=dic))
        else:
            vals = np.asarray(arr.fill_null(0).to_numpy(zero_copy_only=False))
            cols.append(G.Column(types[c], vals, valid if nullable[c] else None))
    return cols


def test_parquet_row_groups_into_the_operator():
    # WHERE c < 30 GROUP BY a, b: COUNT(*), SUM(c), AVG(d), MAX(k); every
    # decoded row group is one consume_device batch of one operator
    t = _table(400_000, 4)
    image = _write(t, row_group_size=122_880, data_page_size=64 * 1024)
    f = G.ParquetFile(image)
    types = [c[1] for c in f.columns]
    nullable = [c[2] for c in f.columns]
    q = G.Query(types, [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2), (G.AVG, 4), (G.MAX, 3)],
                predicates=[(2, G.LT, 30)], nullable=nullable, drop_null_keys=True)
    cols = _host_columns(t, types, nullable)
    c_, v_, ty, ov = O.groupby_exact(q, cols)
    exp = oracle_rows(c_, v_, ty, {1: cols[1].dictionary})
    for whole in (False, True):  # one batch per row group, or one for the file
        op = G.GpuHashAggregate(q)
        op.Init()
        if whole:
            op.consume_device(f.decode(0, last=f.num_row_groups))
        else:
            for rg in range(f.num_row_groups):
                op.consume_device(f.decode(rg))
        op.finalize()
        assert_same_groups(op.result().rows(), exp, 2, ty)
        op.close()
    # GROUP BY s, whose row groups arrive as strings (PLAIN pages after the
    # dictionary overflows): interned by the operator
    q = G.Query(types, [6], [(G.COUNT_STAR, None), (G.SUM, 2), (G.MAX, 3)], nullable=nullable)
    c_, v_, ty, ov = O.groupby_exact(q, cols)
    exp = oracle_rows(c_, v_, ty, {0: cols[6].dictionary})
    for whole in (False, True):
        op = G.GpuHashAggregate(q)
        op.Init()
        batches = [f.decode(0, last=f.num_row_groups)] if whole else [f.decode(rg) for rg in range(f.num_row_groups)]
        assert any(b[6].raw is not None for b in batches)
        for b in batches:
            op.consume_device(b)
        op.finalize()
        assert_same_groups(op.result().rows(), exp, 1, ty)
        op.close()
    f.close()


def test_random_files_match_pyarrow():
    # random flat tables: column types, NULL rates, cardinalities (bit widths
    # of dictionary indices from 0 to 20), data page sizes, row group sizes,
    # page versions and dictionary on/off per column
    for case in range(24):
        rng = np.random.default_rng(300 + case)
        n = int(rng.integers(1, 60_000))
        fields, arrays, use_dict = [], {}, []
        for j in range(int(rng.integers(1, 6))):
            kind = int(rng.integers(0, 4))
            card = int(rng.choice([1, 2, 3, 200, 70_000]))
            null_p = float(rng.choice([0.0, 0.0, 0.01, 0.5, 1.0]))
            mask = rng.random(n) < null_p if null_p else None
            name = f"c{j}"
            if kind == 0:
                a = pa.array(rng.integers(-card, card + 1, n).astype(np.int32), mask=mask, type=pa.int32())
            elif kind == 1:
                a = pa.array(rng.integers(0, card, n) * 7_777_777_777, mask=mask, type=pa.int64())
            elif kind == 2:
                a = pa.array(rng.integers(0, card, n) / 8.0, mask=mask, type=pa.float64())
            else:
                a = pa.array([f"v{v}" + "q" * int(v % 13) for v in rng.integers(0, card, n)], mask=mask,
                             type=pa.string())
            nullable = mask is not None or rng.random() < 0.5
            fields.append(pa.field(name, a.type, nullable=nullable))
            arrays[name] = a
            if rng.random() < 0.7:
                use_dict.append(name)
        t = pa.table(arrays, schema=pa.schema(fields))
        kw = dict(row_group_size=int(rng.choice([1_000, 30_000, 1_000_000])),
                  data_page_size=int(rng.choice([512, 8 * 1024, 1 << 20])),
                  data_page_version=str(rng.choice(["1.0", "2.0"])),
                  dictionary_pagesize_limit=int(rng.choice([256, 64 * 1024, 1 << 20])),
                  use_dictionary=use_dict)
        print(f"[parquet] random case {case}: rows {n} schema {t.schema.types} {kw}", flush=True)
        image = _write(t, **kw)
        _check_file(image, t, kw["row_group_size"], resident=bool(case % 2))


@pytest.mark.parametrize("resident,flush_rows", [(True, 1), (False, 100_000), (True, 1 << 40)])
def test_queued_row_groups_match_pyarrow(resident, flush_rows):
    # gbagg_parquet_decode_async: every row group queued, launched every
    # flush_rows rows (1: each on its own; 2^40: all at the sync), then read
    t = _table(300_000, 5)
    image = _write(t, row_group_size=40_000, data_page_size=32 * 1024, data_page_version="2.0",
                   dictionary_pagesize_limit=64 * 1024, use_dictionary=["a", "b", "k", "d", "u", "s"])
    f = G.ParquetFile(image)
    if resident:
        f.upload()
    f.set_flush_rows(flush_rows)
    queued = [f.decode_async(rg) for rg in range(f.num_row_groups)]
    f.sync()
    row = 0
    for rg, cols in enumerate(queued):
        n = f.row_group_rows(rg)
        assert all(c.n == n for c in cols)
        print(f"[parquet] queued row group {rg} rows {n}", flush=True)
        _check_row_group(cols, t, row, n, rg)
        row += n
    assert row == t.num_rows
    # a synchronous decode after queued ones, and queued ones dropped by reset
    _check_row_group(f.decode(1), t, 40_000, 40_000, 1)
    f.decode_async(0)
    f.reset()
    f.sync()
    f.close()


def test_queued_row_groups_into_the_operator():
    t = _table(400_000, 6)
    image = _write(t, row_group_size=122_880, data_page_size=64 * 1024)
    f = G.ParquetFile(image)
    f.upload()
    types = [c[1] for c in f.columns]
    nullable = [c[2] for c in f.columns]
    q = G.Query(types, [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2), (G.AVG, 4), (G.MAX, 3)],
                predicates=[(2, G.LT, 30)], nullable=nullable, drop_null_keys=True)
    cols = _host_columns(t, types, nullable)
    c_, v_, ty, ov = O.groupby_exact(q, cols)
    exp = oracle_rows(c_, v_, ty, {1: cols[1].dictionary})
    op = G.GpuHashAggregate(q)
    op.Init()
    for rg in range(f.num_row_groups):
        op.consume_device(f.decode_async(rg))  # recorded; read at finalize
    f.sync()
    op.finalize()
    assert_same_groups(op.result().rows(), exp, 2, ty)
    op.close()
    f.close()


def _dictionary_page_num_values(image, offset):
    """Byte range of the num_values varint of the dictionary page header at
    `offset` (Thrift compact PageHeader: field 7 is the DictionaryPageHeader,
    whose field 1 is num_values)."""
    p = offset

    def varint():
        nonlocal p
        v, sh = 0, 0
        while True:
            b = image[p]
            p += 1
            v |= (b & 0x7F) << sh
            sh += 7
            if not b & 0x80:
                return v

    fid = 0
    while True:
        h = image[p]
        p += 1
        assert h, "no dictionary page header"
        fid = fid + (h >> 4) if h >> 4 else None
        assert fid is not None, "long-form field header"
        if (h & 15) == 5:
            varint()
        elif (h & 15) == 12 and fid == 7:
            h2 = image[p]
            p += 1
            assert h2 == 0x15, "num_values is the first field"
            start = p
            n = varint() >> 1
            return start, p, n
        else:
            raise AssertionError(f"unexpected field type {h & 15}")


def test_queued_kernel_errors_reported_by_sync():
    # a dictionary page that claims fewer entries than the data pages index:
    # the host accepts the page (its values fit), the kernels find indices
    # outside the dictionary, and the queued decode reports it at the sync
    rng = np.random.default_rng(7)
    t = pa.table({"x": pa.array(rng.integers(0, 100, 50_000).astype(np.int32))},
                 schema=pa.schema([pa.field("x", pa.int32(), nullable=False)]))
    image = bytearray(_write(t, row_group_size=50_000, use_dictionary=True))
    meta = pq.ParquetFile(io.BytesIO(bytes(image))).metadata
    off = meta.row_group(0).column(0).dictionary_page_offset
    start, end, n = _dictionary_page_num_values(image, off)
    assert n == 100 and end - start == 2  # zigzag 200: two bytes
    image[start:end] = bytes([0x8C, 0x01])  # 70 entries, same two bytes
    for queued in (True, False):
        f = G.ParquetFile(bytes(image))
        if queued:
            f.decode_async(0)
            with pytest.raises(G.GbaggError, match="outside the dictionary"):
                f.sync()
            f.sync()  # reported once
        else:
            with pytest.raises(G.GbaggError, match="outside the dictionary"):
                f.decode(0)
        f.close()
