"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference tree, after `make -C oracle`):

  hash_u64.json          ankerl::unordered_dense hash of int64/int32/uint64
                         keys, from the reference's vendored header
                         (include/third_party/unordered_dense.h) compiled into
                         oracle/_ref
  radix_count.json       the reference RadixCountAgg (radix_count_agg.cpp:17-106)
                         on a seeded Zipf key column: group count, all
                         (key, count) pairs of the top 50, and a checksum of
                         all pairs
  inline_row_agg.json    the reference InlineRowAgg<2> (inline_row_agg.cpp:22-266)
                         on packed (int32, int32) keys with two nullable int64
                         inputs: every group's count, sums and non-NULL counts
  plain_1m.json          the plain aggregate workload (1M rows, 1000 groups)
                         aggregated by pandas, an implementation independent of
                         the oracle

Inputs are regenerated from slothdb_amd.datagen with the seeds recorded in
each file, so the fixtures hold only inputs' parameters and outputs.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from slothdb_amd import datagen  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def hash_inputs():
    xs = [0, 1, 2, 3, -1, -2, 42, 1000, 123456789, -987654321, 2**31 - 1, -2**31, 2**32, 2**53 + 1,
          2**62, 2**63 - 1, -2**63, 0x9E3779B97F4A7C15 - 2**64, 0x5555555555555555, -0x5555555555555556]
    rng = np.random.default_rng(2024)
    xs += [int(x) for x in rng.integers(-2**63, 2**63 - 1, 44, dtype=np.int64)]
    return xs


def radix_inputs():
    return dict(n=300_000, nkeys=50_000, s=1.1, seed=7)


def inline_inputs():
    return dict(n=200_000, a_card=40, b_card=60, seed=99)


def inline_columns(n, a_card, b_card, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(-a_card // 2, a_card // 2, n).astype(np.int32)
    b = rng.integers(0, b_card, n).astype(np.int32)
    v0 = rng.integers(-10**9, 10**9, n).astype(np.int64)
    v1 = rng.integers(-10**15, 10**15, n).astype(np.int64)
    m0 = (rng.random(n) > 0.1).astype(np.uint8)
    m1 = (rng.random(n) > 0.4).astype(np.uint8)
    return a, b, v0, v1, m0, m1


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libslothref.so not built: needs the reference tree")

    xs = hash_inputs()
    hashes = {"i64": [[x, R.ref_hash_i64(x)] for x in xs],
              "i32": [[x, R.ref_hash_i32(x)] for x in xs if -2**31 <= x < 2**31],
              "u64": [[x & (2**64 - 1), R.ref_hash_u64(x & (2**64 - 1))] for x in xs]}
    with open(os.path.join(HERE, "hash_u64.json"), "w") as f:
        json.dump({"source": "include/third_party/unordered_dense.h:158-191,266-267,388-394", **hashes}, f)

    p = radix_inputs()
    keys = datagen.highcard(p["n"], p["nkeys"], p["s"], seed=p["seed"], columns=("k",))["k"]
    cap = p["n"]
    ok, oc, nn = np.zeros(cap, np.int64), np.zeros(cap, np.int64), C.c_uint64()
    R.ref_radix_count(keys.ctypes.data, len(keys), 8, 0, ok.ctypes.data, oc.ctypes.data, cap, C.byref(nn))
    allk, allc = ok[:nn.value].copy(), oc[:nn.value].copy()
    R.ref_radix_count(keys.ctypes.data, len(keys), 8, 50, ok.ctypes.data, oc.ctypes.data, cap, C.byref(nn))
    order = np.lexsort((allk, allc))
    chk = int(np.bitwise_xor.reduce((allk * np.int64(1000003)) ^ allc)) if len(allk) else 0
    with open(os.path.join(HERE, "radix_count.json"), "w") as f:
        json.dump({"source": "src/execution/radix_count_agg.cpp:17-106", "params": p,
                   "groups": int(len(allk)), "rows": int(allc.sum()), "xor_checksum": chk,
                   "top50_counts": oc[:nn.value].tolist(),
                   "all_pairs_sorted_head": [[int(allk[i]), int(allc[i])] for i in order[-200:]]}, f)

    p = inline_inputs()
    a, b, v0, v1, m0, m1 = inline_columns(**p)
    packed = ((a.astype(np.int64).view(np.uint64) & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | \
             (b.astype(np.int64).view(np.uint64) & np.uint64(0xFFFFFFFF))
    vals = np.concatenate([v0, v1])
    valid = np.concatenate([m0, m1])
    cap = p["n"]
    okey, ocs = np.zeros(cap, np.uint64), np.zeros(cap, np.int64)
    osum, ocnt = np.zeros(cap * 2, np.int64), np.zeros(cap * 2, np.int64)
    R.ref_inline_row_agg(2, packed.ctypes.data, vals.ctypes.data, valid.ctypes.data, p["n"], 8, 0,
                         okey.ctypes.data, ocs.ctypes.data, osum.ctypes.data, ocnt.ctypes.data, cap,
                         C.byref(nn))
    m = nn.value
    rows = []
    for i in range(m):
        k = int(okey[i])
        ka = (k >> 32) & 0xFFFFFFFF
        kb = k & 0xFFFFFFFF
        ka = ka - 2**32 if ka >= 2**31 else ka
        kb = kb - 2**32 if kb >= 2**31 else kb
        rows.append([ka, kb, int(ocs[i]), int(osum[2 * i]), int(osum[2 * i + 1]), int(ocnt[2 * i]),
                     int(ocnt[2 * i + 1])])
    rows.sort()
    with open(os.path.join(HERE, "inline_row_agg.json"), "w") as f:
        json.dump({"source": "src/execution/inline_row_agg.cpp:22-266", "params": p,
                   "columns": ["a", "b", "count_star", "sum0", "sum1", "cnt0", "cnt1"], "rows": rows}, f)

    import pandas as pd
    k, v = datagen.plain(1_000_000)
    df = pd.DataFrame({"k": k, "v": v})
    g = df.groupby("k")["v"].agg(["count", "sum", "min", "max"]).reset_index()
    with open(os.path.join(HERE, "plain_1m.json"), "w") as f:
        json.dump({"source": "pandas groupby of slothdb_amd.datagen.plain(1_000_000), seed 42",
                   "columns": ["k", "count", "sum", "min", "max"],
                   "rows": g.astype("int64").values.tolist()}, f)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
