"""gbagg_hash_aggregate_i32_noalloc, the allocation-free variant of
gbagg_hash_aggregate_i32 (GROUP BY int32 key, SUM and COUNT of int32 values),
called through the C ABI on device buffers: two calls in a row, the second
with a smaller input and other keys, each equal to the allocating entry
point's result and to the oracle, bit for bit; the row limit and argument
errors; an empty input."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from slothdb_amd import gbagg as G

pytestmark = pytest.mark.gpu

MAX_ROWS = 524288


class _Dev:
    """Device buffers allocated by the test (the entry point allocates none)."""

    def __init__(self):
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        G._check(G.lib().gbagg_device_alloc(max(nbytes, 1), C.byref(p)))
        self.ptrs.append(p.value)
        return p.value

    def upload(self, arr):
        p = self.alloc(arr.nbytes)
        if arr.nbytes:
            G._check(G.lib().gbagg_memcpy_h2d(p, arr.ctypes.data, arr.nbytes))
        return p

    def download(self, p, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            G._check(G.lib().gbagg_memcpy_d2h(out.ctypes.data, p, out.nbytes))
        return out

    def free(self):
        for p in self.ptrs:
            G.lib().gbagg_device_free(p)
        self.ptrs = []


def _noalloc(dev, keys, vals, capacity):
    dk, dv = dev.upload(keys), dev.upload(vals)
    ok, os_, oc = dev.alloc(capacity * 4), dev.alloc(capacity * 8), dev.alloc(capacity * 8)
    og = dev.alloc(8)
    print(f"[noalloc] call: {len(keys)} rows, capacity {capacity}", flush=True)
    st = G.lib().gbagg_hash_aggregate_i32_noalloc(dk, dv, len(keys), ok, os_, oc, capacity, og, None)
    assert st == G.OK, G.lib().gbagg_last_error().decode()
    G._check(G.lib().gbagg_device_synchronize())
    n = int(dev.download(og, np.uint64, 1)[0])
    m = min(n, capacity)
    return n, dev.download(ok, np.int32, m), dev.download(os_, np.int64, m), dev.download(oc, np.int64, m)


def _sorted(k, s, c):
    o = np.argsort(k, kind="stable")
    return k[o], s[o], c[o]


def _oracle(keys, vals):
    q = G.Query([G.INTEGER, G.INTEGER], [0], [(G.SUM, 1), (G.COUNT_STAR, None)], nullable=[False, False])
    cols, _, _, ov = O.groupby_exact(q, [G.Column(G.INTEGER, keys), G.Column(G.INTEGER, vals)])
    assert not ov
    return _sorted(cols[0].astype(np.int32), cols[1].astype(np.int64), cols[2].astype(np.int64))


def _check_call(dev, keys, vals):
    n, k, s, c = _noalloc(dev, keys, vals, MAX_ROWS)
    got = _sorted(k, s, c)
    ref = _sorted(*G.hash_aggregate_i32(keys, vals))  # the allocating entry point
    exp = _oracle(keys, vals)
    assert n == len(ref[0]) == len(exp[0])
    for a, b, e in zip(got, ref, exp):
        assert np.array_equal(a, b) and np.array_equal(a, e)
    return n


def test_noalloc_two_calls_match_allocating_entry_point_and_oracle():
    rng = np.random.default_rng(4)
    dev = _Dev()
    try:
        # the plain aggregate's shape at the entry point's row limit: 1000 keys
        n1 = MAX_ROWS
        k1 = rng.integers(0, 1000, n1).astype(np.int32)
        v1 = rng.integers(-2**31, 2**31, n1).astype(np.int32)
        assert _check_call(dev, k1, v1) == 1000
        # a smaller input over other keys (many of them singletons): nothing of
        # the first call's groups may remain in the shared scratch table
        n2 = 70_001
        k2 = rng.integers(-200_000, -100_000, n2).astype(np.int32)
        v2 = rng.integers(-1000, 1000, n2).astype(np.int32)
        g2 = _check_call(dev, k2, v2)
        assert g2 > 40_000
        # an empty input: no groups
        n, k, _, _ = _noalloc(dev, k2[:0], v2[:0], 16)
        assert n == 0 and len(k) == 0
    finally:
        dev.free()


def test_noalloc_capacity_and_limits():
    rng = np.random.default_rng(5)
    dev = _Dev()
    try:
        k = rng.integers(0, 5000, 100_000).astype(np.int32)
        v = np.ones(100_000, np.int32)
        # capacity below the group count: the count is reported, capacity rows written
        n, kk, ss, cc = _noalloc(dev, k, v, 100)
        assert n == len(np.unique(k)) and len(kk) == 100
        exp = dict(zip(*np.unique(k, return_counts=True)))
        assert all(exp[int(a)] == int(b) == int(c) for a, b, c in zip(kk, ss, cc))
        # beyond the row limit: rejected before anything runs
        og = dev.alloc(8)
        st = G.lib().gbagg_hash_aggregate_i32_noalloc(og, og, MAX_ROWS + 1, og, og, og, 1, og, None)
        assert st == G.INVALID and "524288" in G.lib().gbagg_last_error().decode()
        st = G.lib().gbagg_hash_aggregate_i32_noalloc(None, None, 10, og, og, og, 1, og, None)
        assert st == G.INVALID
    finally:
        dev.free()
