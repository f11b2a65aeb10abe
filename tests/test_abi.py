"""The C-ABI library loads without a GPU, exports every symbol the public
header declares, and validates plans on the host (no device calls)."""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

from slothdb_amd import gbagg as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slothdb_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbagg_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    names = declared_functions()
    assert len(names) >= 18
    assert set(names) == set(G._SIGS), set(names) ^ set(G._SIGS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    lib = G.lib()  # loads and binds every signature
    for n in declared_functions():
        assert hasattr(lib, n)


def test_abi_version_and_host_hash():
    assert G.lib().gbagg_abi_version() == 1
    with open(os.path.join(ROOT, "tests", "golden", "hash_u64.json")) as f:
        gold = json.load(f)
    for x, h in gold["u64"]:
        assert G.lib().gbagg_hash_u64(x) == h


def _plan(q):
    d, keep = q.desc()
    h = C.c_void_p()
    st = G.lib().gbagg_plan_create(C.byref(d), C.byref(h))
    msg = G.lib().gbagg_last_error().decode()
    if st == G.OK:
        G.lib().gbagg_destroy(h)
    return st, msg


def test_multi_column_key_plans_create_without_device():
    # more than two columns, DOUBLE columns, nullable pairs that keep NULL groups
    assert _plan(G.Query([G.INTEGER, G.BIGINT, G.DOUBLE, G.VARCHAR, G.BIGINT], [0, 1, 2, 3],
                         [(G.COUNT_STAR, None), (G.SUM, 4)]))[0] == G.OK
    assert _plan(G.Query([G.DOUBLE], [0], [(G.COUNT_STAR, None)]))[0] == G.OK
    assert _plan(G.Query([G.BIGINT, G.BIGINT], [0, 1], [(G.COUNT_STAR, None)]))[0] == G.OK
    assert _plan(G.Query([G.INTEGER] * 9, list(range(8)), [(G.MAX, 8)], topn=(8, False, 5)))[0] == G.OK


def test_valid_plans_create_without_device():
    assert _plan(G.Query([G.INTEGER, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1)]))[0] == G.OK
    assert _plan(G.Query([G.INTEGER, G.VARCHAR, G.INTEGER], [0, 1], [(G.SUM, 2)],
                         predicates=[(2, G.LT, 30)], nullable=[0, 0, 0], topn=(2, False, 10)))[0] == G.OK
    assert _plan(G.Query([G.DOUBLE], [], [(G.AVG, 0), (G.MIN, 0)]))[0] == G.OK
    # COUNT(DISTINCT) with other aggregates, grouped and scalar
    assert _plan(G.Query([G.INTEGER, G.VARCHAR, G.DOUBLE], [0],
                         [(G.COUNT_DISTINCT, 1), (G.AVG, 2), (G.COUNT_STAR, None)], nullable=[0, 0, 1]))[0] == G.OK
    assert _plan(G.Query([G.BIGINT], [], [(G.COUNT_DISTINCT, 0)]))[0] == G.OK
    # MIN/MAX over VARCHAR (codes relabelled in byte order)
    assert _plan(G.Query([G.INTEGER, G.VARCHAR], [0], [(G.MIN, 1), (G.MAX, 1)]))[0] == G.OK
    # a nullable argument next to other aggregates (its NULL rows form a phase-1 group)
    assert _plan(G.Query([G.INTEGER] * 3, [0], [(G.COUNT_DISTINCT, 1), (G.SUM, 2)], nullable=[0, 1, 0]))[0] == G.OK
    # COUNT(DISTINCT) over several argument columns, grouped, by two columns and scalar
    for g in ([0], [0, 1], []):
        assert _plan(G.Query([G.INTEGER, G.VARCHAR, G.BIGINT, G.DOUBLE], g,
                             [(G.COUNT_DISTINCT, 2), (G.COUNT_DISTINCT, 3), (G.SUM, 2)],
                             nullable=[0, 0, 0, 1]))[0] == G.OK
    # 16-byte keys: two BIGINT columns, BIGINT with a VARCHAR or INTEGER column
    for types in ([G.BIGINT, G.BIGINT], [G.VARCHAR, G.BIGINT], [G.BIGINT, G.INTEGER]):
        assert _plan(G.Query(types + [G.BIGINT], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)],
                             nullable=[0, 0, 0]))[0] == G.OK


@pytest.mark.parametrize("q,needle", [
    (G.Query([G.INTEGER] * 10, list(range(9)), [(G.COUNT_STAR, None)], nullable=[0] * 10), "at most 8 group"),
    (G.Query([G.INTEGER] * 16, [0, 1, 2], [(G.COUNT_STAR, None)], nullable=[0] * 16), "at most 15 columns"),
    (G.Query([G.INTEGER] * 3, [0, 1, 2], [(G.COUNT_STAR, None)], topn=(1, False, 3)), "TopN on a group column"),
    (G.Query([G.INTEGER, G.VARCHAR], [0], [(G.SUM, 1)]), "numeric"),
    (G.Query([G.INTEGER, G.VARCHAR, G.INTEGER], [0], [(G.MIN, 1), (G.COUNT_DISTINCT, 2)]), "MIN/MAX over VARCHAR"),
    (G.Query([G.INTEGER], [0], [(G.COUNT_STAR, None)], predicates=[(0, G.LIKE, "x")]), "LIKE"),
    (G.Query([G.INTEGER], [0], [(G.SUM, 3)]), "out of range"),
    (G.Query([G.INTEGER] * 10, [0], [(G.SUM, c) for c in range(1, 10)]), "8 distinct"),
    (G.Query([G.INTEGER], [], []), "nothing"),
    (G.Query([G.INTEGER] * 9, list(range(8)), [(G.COUNT_DISTINCT, 8)], nullable=[0] * 9), "at most 7 group"),
    (G.Query([G.INTEGER], [0], [(G.COUNT_STAR, None)], topn=(5, False, 3)), "topn_col"),
])
def test_invalid_plans_are_rejected(q, needle):
    st, msg = _plan(q)
    assert st == G.INVALID
    assert needle in msg


def test_null_arguments():
    assert G.lib().gbagg_plan_create(None, None) == G.INVALID
    assert G.lib().gbagg_finalize(None) == G.INVALID
    assert G.lib().gbagg_reset(None) == G.INVALID
    G.lib().gbagg_destroy(None)
