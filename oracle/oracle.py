"""ctypes wrapper of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE
ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker or as the timed CPU baseline -- never by the operator.

Also wraps oracle/_ref/libslothref.so (the reference's own RadixCountAgg,
InlineRowAgg and ankerl hash compiled from /root/reference) when it was built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libslothref.so")

MAX_OUT = 24  # ORC_MAX_OUT: 8 group columns + 16 aggregates
INTEGER, BIGINT, DOUBLE, VARCHAR = 5, 6, 12, 15


class OrcResult(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32),
                ("types", C.c_int32 * MAX_OUT), ("cells", C.c_void_p * MAX_OUT),
                ("valid", C.c_void_p * MAX_OUT), ("overflow", C.c_int32)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_hash_u64.restype = C.c_uint64
        L.orc_hash_u64.argtypes = [C.c_uint64]
        L.orc_groupby_exact.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.POINTER(OrcResult)]
        L.orc_port_groupby.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.POINTER(OrcResult)]
        L.orc_port_radix_count.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int,
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_uint64)]
        L.orc_topn.restype = C.c_uint64
        L.orc_topn.argtypes = [C.POINTER(OrcResult), C.c_uint32, C.c_int32, C.c_int32, C.c_uint64, C.c_void_p]
        L.orc_result_free.argtypes = [C.POINTER(OrcResult)]
        L.orc_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def ref():
    """The reference's own aggregators (None if oracle/_ref was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        L.ref_hash_u64.restype = C.c_uint64
        L.ref_hash_u64.argtypes = [C.c_uint64]
        L.ref_hash_i64.restype = C.c_uint64
        L.ref_hash_i64.argtypes = [C.c_int64]
        L.ref_hash_i32.restype = C.c_uint64
        L.ref_hash_i32.argtypes = [C.c_int32]
        L.ref_radix_count.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.POINTER(C.c_uint64)]
        L.ref_inline_row_agg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int,
                                         C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint64, C.POINTER(C.c_uint64)]
        _ref = L
    return _ref


def _to_table(r):
    """OrcResult -> (columns, validity, types) as numpy arrays."""
    n = r.num_rows
    cols, valid, types = [], [], []
    for c in range(r.num_columns):
        t = r.types[c]
        cells = np.ctypeslib.as_array(C.cast(r.cells[c], C.POINTER(C.c_int64)), (n,)).copy() if n \
            else np.zeros(0, np.int64)
        v = np.ctypeslib.as_array(C.cast(r.valid[c], C.POINTER(C.c_uint8)), (n,)).astype(bool) if n \
            else np.zeros(0, bool)
        if t == DOUBLE:
            cells = cells.view(np.float64)
        elif t == INTEGER:
            cells = cells.astype(np.int32)
        elif t == VARCHAR:
            cells = cells.astype(np.int64)
        cols.append(cells)
        valid.append(v)
        types.append(t)
    return cols, valid, types


def _call(fn, query, columns, threads):
    from slothdb_amd import gbagg as G
    q, keep = query.desc()
    views, keep2 = G._views(columns, device=False)
    res = OrcResult()
    n = len(columns[0]) if columns else 0
    rc = fn(C.byref(q), views, n, threads, C.byref(res))
    if rc:
        raise RuntimeError(f"oracle returned {rc}")
    out = _to_table(res)
    overflow = res.overflow
    lib().orc_result_free(C.byref(res))
    return out + (overflow,)


def groupby_exact(query, columns, threads=8):
    """Exact GROUP BY -> (columns, validity, types, overflow). VARCHAR keys
    come back as canonical dictionary codes (first entry with those bytes)."""
    return _call(lib().orc_groupby_exact, query, columns, threads)


def port_groupby(query, columns, threads=8):
    """The reference's per-thread-map CPU path, restated."""
    return _call(lib().orc_port_groupby, query, columns, threads)


def port_radix_count(keys, threads=8, k=0):
    keys = np.ascontiguousarray(keys, np.int64)
    pk, pc, n = C.c_void_p(), C.c_void_p(), C.c_uint64()
    lib().orc_port_radix_count(keys.ctypes.data, len(keys), threads, k, C.byref(pk), C.byref(pc), C.byref(n))
    m = n.value
    ks = np.ctypeslib.as_array(C.cast(pk, C.POINTER(C.c_int64)), (m,)).copy() if m else np.zeros(0, np.int64)
    cs = np.ctypeslib.as_array(C.cast(pc, C.POINTER(C.c_int64)), (m,)).copy() if m else np.zeros(0, np.int64)
    lib().orc_free(pk)
    lib().orc_free(pc)
    return ks, cs


def hash_u64(x):
    return lib().orc_hash_u64(int(x) & 0xFFFFFFFFFFFFFFFF)
