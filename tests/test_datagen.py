"""The seeded generators are deterministic, range-independent and follow the
distributions the workloads name. Runs without a GPU."""
import numpy as np

from slothdb_amd import datagen


def test_deterministic_and_row_offset_consistent():
    k1, v1 = datagen.plain(10_000, seed=5)
    k2, v2 = datagen.plain(10_000, seed=5)
    assert (k1 == k2).all() and (v1 == v2).all()
    k3, v3 = datagen.plain(4_000, seed=5, row0=6_000)
    assert (k3 == k1[6000:]).all() and (v3 == v1[6000:]).all()
    a = datagen.highcard(50_000, 1000, seed=9)
    b = datagen.highcard(20_000, 1000, seed=9, row0=30_000)
    for c in ("k", "v1", "v2", "v3"):
        assert np.array_equal(a[c][30_000:], b[c])


def test_plain_ranges_and_uniformity():
    k, v = datagen.plain(1_000_000)
    assert k.min() == 0 and k.max() == 999
    assert v.min() >= -1_000_000 and v.max() <= 1_000_000
    counts = np.bincount(k, minlength=1000)
    chi2 = ((counts - 1000.0) ** 2 / 1000.0).sum()
    assert chi2 < 1200  # 999 dof: mean 999, sd ~45


def test_highcard_distribution():
    n, nkeys = 2_000_000, 100_000
    d = datagen.highcard(n, nkeys)
    exp = datagen.zipf_expected_distinct(1.1, nkeys, n)
    got = len(np.unique(d["k"]))
    assert abs(got - exp) < 5 * np.sqrt(exp)
    # top key share of bounded Zipf(1.1, N): 1 / H(N, 1.1)
    H = (np.arange(1, nkeys + 1, dtype=np.float64) ** -1.1).sum()
    _, c = np.unique(d["k"], return_counts=True)
    assert abs(c.max() / n - 1 / H) < 0.002
    assert d["v1"].min() >= -10**9 and d["v1"].max() <= 10**9
    assert abs(d["v3"].mean() - 100) < 0.1 and abs(d["v3"].std() - 25) < 0.1
    assert d["v2"].min() < -2**30 and d["v2"].max() > 2**30


def test_highcard_full_size_expected_groups():
    # E[distinct] of 100M Zipf(1.1) draws over 10M ranks, the high-cardinality workload
    assert abs(datagen.zipf_expected_distinct(1.1, 10_000_000, 100_000_000) - 4_854_583) < 5


def test_dictionary_and_composite():
    d = datagen.dictionary(20_000)
    strs = [d.get(i) for i in range(len(d))]
    assert len(set(strs)) == len(strs)
    lens = np.diff(d.offsets)
    assert lens.min() == 4 and lens.max() == 32
    assert all(ch.isdigit() or 'a' <= ch <= 'z' for s in strs[:500] for ch in s)
    a, b, c = datagen.composite(500_000, dict_size=20_000)
    assert a.min() >= 0 and a.max() < 65536 and b.max() < 20_000 and c.max() < 100
    assert abs((c < 30).mean() - 0.30) < 0.005


def test_query_i32():
    k, v = datagen.query_i32(100_000, 5000, 0.7, seed=3)
    assert len(np.unique(k)) <= 5000
    k2, _ = datagen.query_i32(100_000, 5000, 0.7, seed=3)
    assert (k == k2).all()


def test_strings_from_codes():
    # every row's string is its code's dictionary entry; offsets ascend from 0
    q, cols = datagen.composite_workload(n=50_000, dict_size=2_000)
    s = datagen.strings_from_codes(cols[1])
    d = cols[1].dictionary
    assert len(s) == 50_000 and s.values[0] == 0 and (np.diff(s.values.astype(np.int64)) >= 4).all()
    for i in range(0, 50_000, 113):
        assert s.string(i) == d.get(int(cols[1].values[i]))
