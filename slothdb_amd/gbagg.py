"""Python binding of the MI355X GROUP BY operator (include/slothdb_gpu.h).

The native library does all the work; this module only marshals numpy arrays
into the C ABI. It mirrors the reference's operator surface
(PhysicalOperator: SetTopNHint / SetRowLimit / Init / GetData,
include/slothdb/execution/physical_operator.hpp:23-65, and the fused-path
RGConsumerFn batches of physical_planner.cpp:1060-1070) so callers drive it
the way the engine drives PhysicalHashAggregate.

Loading fails loudly when libslothdb_gpu.so is missing: there is no CPU
fallback on this path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libslothdb_gpu.so")

# gbagg_type (LogicalTypeId values)
INTEGER, BIGINT, DOUBLE, VARCHAR = 5, 6, 12, 15
# gbagg_agg_kind
COUNT_STAR, COUNT, SUM, AVG, MIN, MAX, COUNT_DISTINCT = 0, 1, 2, 3, 4, 5, 6
# gbagg_cmp_op
LT, LE, GT, GE, EQ, NE, IN, LIKE, NOT_LIKE = range(9)
# gbagg_strategy
AUTO, DIRECT, LDS_HASH, GLOBAL_HASH, RADIX = range(5)
DICT_UNIQUE = 1
STR_RAW = 2

OK, ERROR, INVALID = 0, 1, 2
VECTOR_SIZE = 2048  # rows per DataChunk (include/slothdb/common/constants.hpp:16)

_NP = {INTEGER: np.int32, BIGINT: np.int64, DOUBLE: np.float64, VARCHAR: np.uint32}


class ColumnView(C.Structure):
    _fields_ = [("type", C.c_int), ("data", C.c_void_p), ("validity", C.c_void_p),
                ("dict_offsets", C.c_void_p), ("dict_bytes", C.c_void_p),
                ("dict_size", C.c_uint32), ("flags", C.c_uint32)]


class Agg(C.Structure):
    _fields_ = [("kind", C.c_int), ("column", C.c_int32)]


class Predicate(C.Structure):
    _fields_ = [("column", C.c_int32), ("op", C.c_int), ("ival", C.c_int64), ("dval", C.c_double),
                ("in_set", C.c_void_p), ("in_count", C.c_uint32), ("sval", C.c_char_p),
                ("slen", C.c_uint32)]


class QueryDesc(C.Structure):
    _fields_ = [("num_columns", C.c_uint32), ("column_types", C.c_void_p),
                ("column_nullable", C.c_void_p), ("num_group_cols", C.c_uint32),
                ("group_cols", C.c_void_p), ("num_aggs", C.c_uint32), ("aggs", C.c_void_p),
                ("num_predicates", C.c_uint32), ("predicates", C.c_void_p),
                ("topn_col", C.c_int32), ("topn_ascending", C.c_int32), ("topn_limit", C.c_uint64),
                ("row_limit", C.c_uint64), ("expected_rows", C.c_uint64),
                ("drop_null_keys", C.c_int32), ("strategy", C.c_int)]


class ResultColumn(C.Structure):
    _fields_ = [("type", C.c_int), ("data", C.c_void_p), ("validity", C.c_void_p),
                ("str_offsets", C.c_void_p), ("str_bytes", C.c_void_p)]


class ResultView(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32),
                ("columns", C.POINTER(ResultColumn))]


class PlanInfo(C.Structure):
    _fields_ = [("strategy", C.c_int), ("estimated_groups", C.c_uint64), ("groups", C.c_uint64),
                ("table_capacity", C.c_uint64), ("lds_slots", C.c_uint32),
                ("radix_bits", C.c_uint32), ("workgroups", C.c_uint32), ("retries", C.c_uint32),
                ("rows_consumed", C.c_uint64), ("heavy_keys", C.c_uint64), ("cold_rows", C.c_uint64),
                ("sample_rows", C.c_uint64), ("sample_distinct", C.c_uint64), ("sample_f1", C.c_uint64),
                ("ms_aggregate", C.c_double), ("ms_total", C.c_double), ("ms_sample", C.c_double),
                ("ms_count", C.c_double), ("ms_partition", C.c_double), ("ms_emit", C.c_double),
                ("kernel_shape", C.c_char * 32), ("selector_distinct", C.c_double),
                ("selector_used", C.c_int32), ("selector_path", C.c_int32),
                ("ms_strings", C.c_double), ("strings_distinct", C.c_uint64),
                ("sum_reruns", C.c_uint32), ("sum_ieee", C.c_int32)]


# Every symbol include/slothdb_gpu.h declares, with its ctypes signature.
_SIGS = {
    "gbagg_abi_version": (C.c_int32, []),
    "gbagg_last_error": (C.c_char_p, []),
    "gbagg_device_name": (C.c_int, [C.c_char_p, C.c_uint32]),
    "gbagg_plan_create": (C.c_int, [C.POINTER(QueryDesc), C.POINTER(C.c_void_p)]),
    "gbagg_consume": (C.c_int, [C.c_void_p, C.POINTER(ColumnView), C.c_uint64]),
    "gbagg_consume_device": (C.c_int, [C.c_void_p, C.POINTER(ColumnView), C.c_uint64]),
    "gbagg_finalize": (C.c_int, [C.c_void_p]),
    "gbagg_result": (C.c_int, [C.c_void_p, C.POINTER(ResultView)]),
    "gbagg_reset": (C.c_int, [C.c_void_p]),
    "gbagg_plan_info_get": (C.c_int, [C.c_void_p, C.POINTER(PlanInfo)]),
    "gbagg_destroy": (None, [C.c_void_p]),
    "gbagg_set_param": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "gbagg_hash_aggregate_i32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_uint64,
                                           C.POINTER(C.c_uint64)]),
    "gbagg_hash_aggregate_i32_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "gbagg_batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "gbagg_batch_set_tune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gbagg_batch_set_param": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "gbagg_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "gbagg_batch_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "gbagg_batch_result": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "gbagg_batch_info_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gbagg_batch_destroy": (None, [C.c_void_p]),
    "gbagg_selector_set": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "gbagg_plan_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gbagg_selector_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                       C.c_void_p, C.c_uint32, C.c_void_p]),
    "gbagg_selector_infer": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]),
    "gbagg_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "gbagg_device_free": (C.c_int, [C.c_void_p]),
    "gbagg_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "gbagg_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "gbagg_device_synchronize": (C.c_int, []),
    "gbagg_hash_u64": (C.c_uint64, [C.c_uint64]),
    "gbagg_parquet_open": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gbagg_parquet_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint64)]),
    "gbagg_parquet_column": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_uint8),
                                       C.POINTER(C.c_char_p)]),
    "gbagg_parquet_row_group_rows": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "gbagg_parquet_decode": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                       C.POINTER(C.c_uint64)]),
    "gbagg_parquet_decode_range": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                             C.POINTER(C.c_uint64)]),
    "gbagg_parquet_decode_async": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                             C.POINTER(C.c_uint64)]),
    "gbagg_parquet_sync": (C.c_int, [C.c_void_p]),
    "gbagg_parquet_set_flush_rows": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gbagg_parquet_upload": (C.c_int, [C.c_void_p]),
    "gbagg_parquet_reset": (C.c_int, [C.c_void_p]),
    "gbagg_parquet_close": (None, [C.c_void_p]),
}

_lib = None


class GbaggError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gbagg status {status}: {msg}")
        self.status = status


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m slothdb_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(st):
    if st != OK:
        raise GbaggError(st, lib().gbagg_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data


class Dictionary:
    """A VARCHAR dictionary as offsets + bytes (ParquetColumnData::str_dict_values)."""

    def __init__(self, strings=None, offsets=None, data=None, unique=False):
        if strings is not None:
            enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
            offsets = np.zeros(len(enc) + 1, np.uint32)
            offsets[1:] = np.cumsum([len(s) for s in enc])
            data = np.frombuffer(b"".join(enc) or b"\0", np.uint8).copy()
        self.offsets = np.ascontiguousarray(offsets, np.uint32)
        self.data = np.ascontiguousarray(data, np.uint8)
        self.unique = unique

    def __len__(self):
        return len(self.offsets) - 1

    def get(self, i):
        return bytes(self.data[self.offsets[i]:self.offsets[i + 1]]).decode()


class Column:
    """One input column of a batch: typed values, optional validity bytes
    (0 = NULL) and, for VARCHAR, uint32 codes into a Dictionary, or the
    strings themselves (Column.strings)."""

    def __init__(self, type_, values, validity=None, dictionary=None, raw=None):
        self.type = type_
        self.values = np.ascontiguousarray(values, _NP[type_])
        self.validity = None if validity is None else np.ascontiguousarray(validity, np.uint8)
        self.dictionary = dictionary
        self.raw = None if raw is None else np.ascontiguousarray(raw, np.uint8)

    @classmethod
    def strings(cls, strings, offsets=None, data=None, validity=None):
        """VARCHAR column given as its strings (GBAGG_STR_RAW): `values` holds
        the n + 1 uint32 offsets into the bytes `raw`; None entries are NULL."""
        if strings is not None:
            enc = [b"" if x is None else (x.encode() if isinstance(x, str) else bytes(x)) for x in strings]
            offsets = np.zeros(len(enc) + 1, np.uint32)
            offsets[1:] = np.cumsum([len(x) for x in enc])
            data = np.frombuffer(b"".join(enc) or b"\0", np.uint8).copy()
            if validity is None and any(x is None for x in strings):
                validity = np.array([x is not None for x in strings], np.uint8)
        return cls(VARCHAR, offsets, validity, raw=data)

    def __len__(self):
        return len(self.values) - 1 if self.raw is not None else len(self.values)

    def string(self, i):
        return bytes(self.raw[self.values[i]:self.values[i + 1]]).decode()


class DeviceColumn:
    """A column resident in HBM (allocated through the library)."""

    def __init__(self, col):
        self.type = col.type
        self.n = len(col)
        self.dictionary = col.dictionary
        self._ptrs = []
        self.data = self._upload(col.values)
        self.validity = None if col.validity is None else self._upload(col.validity)
        self.raw = None if col.raw is None else self._upload(col.raw)

    def _upload(self, arr):
        p = C.c_void_p()
        _check(lib().gbagg_device_alloc(max(arr.nbytes, 1), C.byref(p)))
        self._ptrs.append(p.value)
        if arr.nbytes:
            _check(lib().gbagg_memcpy_h2d(p.value, arr.ctypes.data, arr.nbytes))
        return p.value

    def free(self):
        for p in self._ptrs:
            lib().gbagg_device_free(p)
        self._ptrs = []

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _views(columns, device):
    arr = (ColumnView * len(columns))()
    keep = []
    for i, c in enumerate(columns):
        v = arr[i]
        v.type = c.type
        if device:
            v.data = c.data
            v.validity = c.validity
        else:
            v.data = _ptr(c.values)
            v.validity = _ptr(c.validity)
        if getattr(c, "raw", None) is not None:
            v.dict_bytes = c.raw if device else c.raw.ctypes.data
            v.flags = STR_RAW
        d = c.dictionary
        if d is not None:
            v.dict_offsets = d.offsets.ctypes.data
            v.dict_bytes = d.data.ctypes.data
            v.dict_size = len(d)
            v.flags = DICT_UNIQUE if d.unique else 0
            keep.append(d)
    return arr, keep


class Query:
    """A planned GROUP BY: input schema, group columns, aggregates
    [(kind, column)], compiled predicates [(column, op, literal)], and the
    TopN hint (output column, ascending, limit)."""

    def __init__(self, column_types, group_cols, aggs, predicates=(), nullable=None,
                 topn=None, row_limit=0, drop_null_keys=False, strategy=AUTO, expected_rows=0):
        self.column_types = list(column_types)
        self.group_cols = list(group_cols)
        self.aggs = [tuple(a) for a in aggs]
        self.predicates = list(predicates)
        self.nullable = nullable
        self.topn = topn
        self.row_limit = row_limit
        self.drop_null_keys = drop_null_keys
        self.strategy = strategy
        self.expected_rows = expected_rows

    def desc(self):
        """A QueryDesc plus the buffers it points into (keep both alive)."""
        keep = []
        types = np.array(self.column_types, np.int32)
        keep.append(types)
        q = QueryDesc()
        q.num_columns = len(types)
        q.column_types = types.ctypes.data
        if self.nullable is not None:
            nl = np.array([1 if x else 0 for x in self.nullable], np.uint8)
            keep.append(nl)
            q.column_nullable = nl.ctypes.data
        gc = np.array(self.group_cols or [0], np.uint32)
        keep.append(gc)
        q.num_group_cols = len(self.group_cols)
        q.group_cols = gc.ctypes.data
        aggs = (Agg * max(1, len(self.aggs)))()
        for i, (k, c) in enumerate(self.aggs):
            aggs[i].kind = k
            aggs[i].column = -1 if c is None else c
        keep.append(aggs)
        q.num_aggs = len(self.aggs)
        q.aggs = C.addressof(aggs)
        preds = (Predicate * max(1, len(self.predicates)))()
        for i, p in enumerate(self.predicates):
            col, op, lit = p[0], p[1], p[2]
            P = preds[i]
            P.column = col
            P.op = op
            if op == IN:
                s = np.array(sorted(set(int(x) for x in lit)), np.int64)
                keep.append(s)
                P.in_set = s.ctypes.data
                P.in_count = len(s)
            elif isinstance(lit, (str, bytes)):
                b = lit.encode() if isinstance(lit, str) else lit
                buf = C.create_string_buffer(b, len(b) + 1)
                keep.append(buf)
                P.sval = C.cast(buf, C.c_char_p)
                P.slen = len(b)
            elif isinstance(lit, float):
                # TryCompileSimplePredicateImpl: dval = literal, ival = (int64)dval
                P.dval = lit
                P.ival = int(lit)
            else:
                P.ival = int(lit)
                P.dval = float(lit)
        keep.append(preds)
        q.num_predicates = len(self.predicates)
        q.predicates = C.addressof(preds)
        if self.topn is None:
            q.topn_col = -1
        else:
            q.topn_col, q.topn_ascending, q.topn_limit = self.topn[0], int(self.topn[1]), self.topn[2]
        q.row_limit = self.row_limit
        q.expected_rows = self.expected_rows
        q.drop_null_keys = int(self.drop_null_keys)
        q.strategy = self.strategy
        return q, keep


class Result:
    """Columnar result: per output column a numpy array, a bool validity
    array, and for VARCHAR keys a list of str."""

    def __init__(self, columns, validity, types):
        self.columns = columns
        self.validity = validity
        self.types = types

    @property
    def num_rows(self):
        return len(self.validity[0]) if self.validity else 0

    def rows(self):
        """List of tuples with None for NULL (GetValue semantics)."""
        out = []
        for r in range(self.num_rows):
            row = []
            for c, col in enumerate(self.columns):
                if not self.validity[c][r]:
                    row.append(None)
                else:
                    v = col[r]
                    row.append(v if isinstance(v, str) else v.item())
            out.append(tuple(row))
        return out


def _result_from_view(rv):
    """Copies a gbagg_result_view into a Result."""
    n = rv.num_rows
    cols, valid, types = [], [], []
    for c in range(rv.num_columns):
        rc = rv.columns[c]
        types.append(rc.type)
        v = np.ctypeslib.as_array(C.cast(rc.validity, C.POINTER(C.c_uint8)), (n,)).astype(bool) \
            if n else np.zeros(0, bool)
        if rc.type == VARCHAR:
            off = np.ctypeslib.as_array(C.cast(rc.str_offsets, C.POINTER(C.c_uint32)), (n + 1,)) \
                if n else np.zeros(1, np.uint32)
            total = int(off[-1]) if n else 0
            raw = C.string_at(rc.str_bytes, total) if total else b""
            col = [raw[off[i]:off[i + 1]].decode() for i in range(n)]
        else:
            dt = _NP[rc.type]
            col = np.ctypeslib.as_array(C.cast(rc.data, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                        (n,)).copy() if n else np.zeros(0, dt)
        cols.append(col)
        valid.append(v)
    return Result(cols, valid, types)


class BatchItem(C.Structure):
    _fields_ = [("query", C.c_void_p), ("columns", C.c_void_p), ("num_rows", C.c_uint64)]


class BatchTune(C.Structure):
    _fields_ = [("tile_rows", C.c_uint32), ("lds_slots", C.c_uint32), ("block", C.c_uint32),
                ("wgs_per_cu", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [("queries", C.c_uint32), ("batched", C.c_uint32), ("single", C.c_uint32),
                ("graph_captures", C.c_uint32), ("graph_replays", C.c_uint32),
                ("ms_device", C.c_double), ("ms_single", C.c_double), ("tune", BatchTune),
                ("selector_queries", C.c_uint32)]


class GpuBatch:
    """Many GROUP BY queries in one submission (gbagg_batch_*): each entry is
    (Query, [DeviceColumn]); the columns must outlive the batch's runs."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().gbagg_batch_create(C.byref(h)))
        self._h = h
        self._keep = []
        self._items = None

    def set_queries(self, entries):
        keep, items = [], (BatchItem * max(1, len(entries)))()
        for i, (q, dcols) in enumerate(entries):
            desc, k = q.desc()
            views, kv = _views(dcols, device=True)
            keep += [desc, k, views, kv, dcols]
            items[i].query = C.addressof(desc)
            items[i].columns = C.addressof(views)
            items[i].num_rows = dcols[0].n if dcols else 0
        self._keep, self._items, self._n = keep, items, len(entries)

    def set_tune(self, tile_rows, lds_slots, block, wgs_per_cu):
        t = BatchTune(tile_rows, lds_slots, block, wgs_per_cu)
        _check(lib().gbagg_batch_set_tune(self._h, C.byref(t)))

    def set_param(self, name, value):
        _check(lib().gbagg_batch_set_param(self._h, name.encode(), int(value)))

    def autotune(self):
        _check(lib().gbagg_batch_autotune(self._h, self._items, self._n))

    def run(self):
        _check(lib().gbagg_batch_run(self._h, self._items, self._n))

    def result(self, i):
        rv = ResultView()
        _check(lib().gbagg_batch_result(self._h, i, C.byref(rv)))
        return _result_from_view(rv)

    def info(self):
        bi = BatchInfo()
        _check(lib().gbagg_batch_info_get(self._h, C.byref(bi)))
        d = {k: getattr(bi, k) for k, _ in BatchInfo._fields_ if k != "tune"}
        d["tune"] = {k: getattr(bi.tune, k) for k, _ in BatchTune._fields_}
        return d

    def close(self):
        if self._h:
            lib().gbagg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuHashAggregate:
    """The operator: plan once, consume batches, finalize, read the result."""

    def __init__(self, query, params=None, selector=None):
        """selector: None, "default" (slothdb_amd/selector_weights.json) or a
        parameter vector of the learned strategy selector."""
        self.query = query
        self.params = dict(params or {})
        self.selector = selector
        self._h = None

    # PhysicalOperator surface
    def SetTopNHint(self, col_idx, ascending, limit):
        self._require_unplanned()
        self.query.topn = (col_idx, ascending, limit)

    def SetRowLimit(self, n):
        self._require_unplanned()
        self.query.row_limit = n

    def Init(self):
        if self._h is None:
            q, keep = self.query.desc()
            h = C.c_void_p()
            _check(lib().gbagg_plan_create(C.byref(q), C.byref(h)))
            self._h = h
            for k, v in self.params.items():
                _check(lib().gbagg_set_param(self._h, k.encode(), int(v)))
            if self.selector is not None:
                from . import selector as S
                p = S.load()[0] if isinstance(self.selector, str) and self.selector == "default" else self.selector
                S.attach(self, p)
        else:
            _check(lib().gbagg_reset(self._h))
        self._emit_pos = 0
        self._result = None

    def _require_unplanned(self):
        if self._h is not None:
            raise RuntimeError("hints must be set before Init()")

    def consume(self, columns):
        """Host batch (RGWork semantics: buffers only needed during the call)."""
        n = len(columns[0]) if columns else 0
        views, keep = _views(columns, device=False)
        _check(lib().gbagg_consume(self._h, views, n))

    def consume_device(self, columns):
        """HBM-resident batch of DeviceColumn; must outlive finalize()."""
        n = columns[0].n if columns else 0
        views, keep = _views(columns, device=True)
        _check(lib().gbagg_consume_device(self._h, views, n))
        self._device_keep = getattr(self, "_device_keep", []) + [columns, keep]

    def finalize(self):
        _check(lib().gbagg_finalize(self._h))

    def result(self):
        if self._result is not None:
            return self._result
        rv = ResultView()
        _check(lib().gbagg_result(self._h, C.byref(rv)))
        self._result = _result_from_view(rv)
        return self._result

    def GetData(self):
        """Next chunk of at most VECTOR_SIZE rows as a list of tuples, or None
        at the end (PhysicalHashAggregate::GetData, physical_planner.cpp:6040-6060)."""
        res = self.result()
        if self._emit_pos >= res.num_rows:
            return None
        rows = res.rows()[self._emit_pos:self._emit_pos + VECTOR_SIZE]
        self._emit_pos += len(rows)
        return rows

    def plan_info(self):
        pi = PlanInfo()
        _check(lib().gbagg_plan_info_get(self._h, C.byref(pi)))
        out = {f: getattr(pi, f) for f, _ in PlanInfo._fields_}
        out["kernel_shape"] = out["kernel_shape"].decode()
        return out

    def close(self):
        if self._h is not None:
            lib().gbagg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(query, columns, device=False, params=None):
    """Plan, consume one batch, finalize and return the Result."""
    op = GpuHashAggregate(query, params)
    op.Init()
    if device:
        dcols = [DeviceColumn(c) for c in columns]
        op.consume_device(dcols)
    else:
        op.consume(columns)
    op.finalize()
    res = op.result()
    info = op.plan_info()
    op.close()
    return res, info


def hash_aggregate_i32(keys, values):
    """GPUEngine::HashAggregate (gpu_engine.hpp:79-85): keys, sums, counts."""
    keys = np.ascontiguousarray(keys, np.int32)
    values = np.ascontiguousarray(values, np.int32)
    n = len(keys)
    cap = max(n, 1)
    ok, os_, oc = np.zeros(cap, np.int32), np.zeros(cap, np.int64), np.zeros(cap, np.int64)
    g = C.c_uint64()
    _check(lib().gbagg_hash_aggregate_i32(keys.ctypes.data, values.ctypes.data, n, ok.ctypes.data,
                                          os_.ctypes.data, oc.ctypes.data, cap, C.byref(g)))
    m = g.value
    return ok[:m], os_[:m], oc[:m]


class DecodedColumn:
    """A column decoded into HBM by ParquetFile.decode (valid until the next
    decode on that file); usable wherever a DeviceColumn is."""

    def __init__(self, view, n):
        self.type = view.type
        self.n = n
        self.data = view.data
        self.validity = view.validity
        self.dictionary = None
        self.raw = None  # BYTE_ARRAY with PLAIN pages: data = offsets, raw = bytes (device)
        if view.type == VARCHAR and view.flags & STR_RAW:
            self.raw = view.dict_bytes
        elif view.type == VARCHAR:
            k = view.dict_size
            off = np.ctypeslib.as_array(C.cast(view.dict_offsets, C.POINTER(C.c_uint32)), shape=(k + 1,)).copy()
            nb = max(int(off[-1]), 1)
            data = np.ctypeslib.as_array(C.cast(view.dict_bytes, C.POINTER(C.c_uint8)), shape=(nb,)).copy()
            self.dictionary = Dictionary(offsets=off, data=data)

    def strings(self):
        """A string column's rows as str (None for NULL), copied back to the host."""
        valid = np.ones(self.n, bool)
        if self.validity:
            v = np.empty(self.n, np.uint8)
            if self.n:
                _check(lib().gbagg_memcpy_d2h(v.ctypes.data, self.validity, self.n))
            valid = v.astype(bool)
        if self.raw is None:
            vals, _ = self.to_host()
            return [self.dictionary.get(int(x)) if ok else None for x, ok in zip(vals, valid)]
        off = np.empty(self.n + 1, np.uint32)
        _check(lib().gbagg_memcpy_d2h(off.ctypes.data, self.data, off.nbytes))
        data = np.empty(max(1, int(off[-1])), np.uint8)
        if off[-1]:
            _check(lib().gbagg_memcpy_d2h(data.ctypes.data, self.raw, int(off[-1])))
        return [bytes(data[off[i]:off[i + 1]]).decode() if valid[i] else None for i in range(self.n)]

    def to_host(self):
        """(values, validity) copied back to host numpy arrays."""
        vals = np.empty(self.n, _NP[self.type])
        if self.n:
            _check(lib().gbagg_memcpy_d2h(vals.ctypes.data, self.data, vals.nbytes))
        valid = None
        if self.validity:
            valid = np.empty(self.n, np.uint8)
            if self.n:
                _check(lib().gbagg_memcpy_d2h(valid.ctypes.data, self.validity, self.n))
        return vals, valid


class ParquetFile:
    """A Parquet file image decoded on the GPU one row group at a time
    (gbagg_parquet_*; the reference's ParquetReader::ReadColumnInto,
    include/slothdb/storage/parquet.hpp:263-265)."""

    def __init__(self, image):
        self._buf = np.frombuffer(bytes(image), np.uint8)
        h = C.c_void_p()
        _check(lib().gbagg_parquet_open(self._buf.ctypes.data, self._buf.nbytes, C.byref(h)))
        self._h = h.value
        try:
            nrg, ncol, nrows = C.c_uint32(), C.c_uint32(), C.c_uint64()
            _check(lib().gbagg_parquet_shape(self._h, C.byref(nrg), C.byref(ncol), C.byref(nrows)))
            self.num_row_groups, self.num_columns, self.num_rows = nrg.value, ncol.value, nrows.value
            self.columns = []
            for c in range(self.num_columns):
                t, nl, name = C.c_int(), C.c_uint8(), C.c_char_p()
                _check(lib().gbagg_parquet_column(self._h, c, C.byref(t), C.byref(nl), C.byref(name)))
                # column names are the file's bytes: not necessarily UTF-8
                self.columns.append((name.value.decode("utf-8", "replace"), t.value, bool(nl.value)))
        except BaseException:
            self.close()
            raise

    def upload(self):
        """Keep the whole file in HBM: decodes then read their chunks there."""
        _check(lib().gbagg_parquet_upload(self._h))

    def reset(self):
        """Forget every decoded column; later decodes reuse its device memory."""
        _check(lib().gbagg_parquet_reset(self._h))

    def row_group_rows(self, rg):
        n = C.c_uint64()
        _check(lib().gbagg_parquet_row_group_rows(self._h, rg, C.byref(n)))
        return n.value

    def decode(self, rg, cols=None, last=None):
        """DecodedColumns of row group rg, or of row groups [rg, last)
        concatenated (all columns by default)."""
        cols = list(range(self.num_columns)) if cols is None else list(cols)
        idx = np.array(cols or [0], np.uint32)
        views = (ColumnView * max(1, len(cols)))()
        n = C.c_uint64()
        last = rg + 1 if last is None else last
        _check(lib().gbagg_parquet_decode_range(self._h, rg, last, idx.ctypes.data, len(cols), C.addressof(views),
                                                C.byref(n)))
        return [DecodedColumn(views[i], n.value) for i in range(len(cols))]

    def decode_async(self, rg, cols=None):
        """Queued decode of row group rg (gbagg_parquet_decode_async): its
        DecodedColumns, whose device values are there after sync()."""
        cols = list(range(self.num_columns)) if cols is None else list(cols)
        idx = np.array(cols or [0], np.uint32)
        views = (ColumnView * max(1, len(cols)))()
        n = C.c_uint64()
        _check(lib().gbagg_parquet_decode_async(self._h, rg, idx.ctypes.data, len(cols), C.addressof(views),
                                                C.byref(n)))
        return [DecodedColumn(views[i], n.value) for i in range(len(cols))]

    def sync(self):
        """Runs the queued decodes and waits for them; raises what their kernels found."""
        _check(lib().gbagg_parquet_sync(self._h))

    def set_flush_rows(self, rows):
        _check(lib().gbagg_parquet_set_flush_rows(self._h, rows))

    def close(self):
        if self._h:
            lib().gbagg_parquet_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
