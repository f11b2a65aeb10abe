"""Parquet footer parsing (gbagg_parquet_open / _shape / _column; no GPU
needed) against pyarrow's view of the same files, and malformed images."""
import io

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

from slothdb_amd import gbagg as G


def test_footer_matches_pyarrow():
    n = 10_000
    t = pa.table({"i": pa.array(np.arange(n, dtype=np.int32)),
                  "s": pa.array([f"x{j % 7}" for j in range(n)]),
                  "l": pa.array(np.arange(n, dtype=np.int64), mask=np.arange(n) % 3 == 0),
                  "d": pa.array(np.ones(n))},
                 schema=pa.schema([pa.field("i", pa.int32(), nullable=False), ("s", pa.string()),
                                   ("l", pa.int64()), pa.field("d", pa.float64(), nullable=False)]))
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="NONE", row_group_size=3000)
    f = G.ParquetFile(buf.getvalue())
    meta = pq.ParquetFile(io.BytesIO(buf.getvalue())).metadata
    assert (f.num_row_groups, f.num_columns, f.num_rows) == (meta.num_row_groups, 4, n)
    assert [f.row_group_rows(i) for i in range(f.num_row_groups)] == \
        [meta.row_group(i).num_rows for i in range(meta.num_row_groups)]
    assert f.columns == [("i", G.INTEGER, False), ("s", G.VARCHAR, True), ("l", G.BIGINT, True),
                         ("d", G.DOUBLE, False)]
    f.close()


@pytest.mark.parametrize("image", [b"", b"PAR1", b"PAR1" + b"\0" * 20 + b"PAR1",
                                   b"NOPE" + b"\0" * 20 + b"\x04\0\0\0PAR1"])
def test_malformed_images_are_rejected(image):
    with pytest.raises(G.GbaggError):
        G.ParquetFile(image)


def test_truncated_footer_is_rejected():
    t = pa.table({"i": pa.array(np.arange(100, dtype=np.int32))})
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="NONE")
    img = bytearray(buf.getvalue())
    flen = int.from_bytes(img[-8:-4], "little")
    img[-8:-4] = (flen // 2).to_bytes(4, "little")  # footer cut in half
    with pytest.raises(G.GbaggError):
        G.ParquetFile(bytes(img))


def test_corrupted_footers_fail_cleanly():
    # random byte changes and cuts inside the footer (Thrift compact protocol:
    # field headers, varints, list sizes, string lengths): every image either
    # opens with a consistent shape or raises, and never reads out of bounds
    n = 5_000
    t = pa.table({"i": pa.array(np.arange(n, dtype=np.int32)),
                  "s": pa.array([f"x{j % 7}" for j in range(n)]),
                  "l": pa.array(np.arange(n, dtype=np.int64), mask=np.arange(n) % 3 == 0)})
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="NONE", row_group_size=1000)
    img = buf.getvalue()
    flen = int.from_bytes(img[-8:-4], "little")
    f0 = len(img) - 8 - flen
    rng = np.random.default_rng(77)
    opened = failed = 0
    for trial in range(3000):
        b = bytearray(img)
        for _ in range(int(rng.integers(1, 4))):
            pos = f0 + int(rng.integers(0, flen))
            b[pos] = int(rng.integers(0, 256)) if rng.random() < 0.7 else b[pos] ^ (1 << int(rng.integers(0, 8)))
        if trial % 10 == 0:  # also cut the footer short, keeping its length field
            cut = int(rng.integers(1, flen))
            b = b[:f0 + flen - cut] + b[-8:]
            b[-8:-4] = (flen - cut).to_bytes(4, "little")
        try:
            f = G.ParquetFile(bytes(b))
        except G.GbaggError:
            failed += 1
            continue
        assert len(f.columns) == f.num_columns
        for rg in range(f.num_row_groups):
            f.row_group_rows(rg)
        f.close()
        opened += 1
    assert failed > 0 and opened + failed == 3000
