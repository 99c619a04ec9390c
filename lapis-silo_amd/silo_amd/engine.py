"""ctypes binding of the engine's C ABI (include/rdb.h).

`Engine` owns one GPU context; `DeviceTable` uploads a `TableSpec` and answers filter, count,
Mutations and Aggregated queries on the device.  There is no CPU fallback: without the built
library or a GPU these calls raise.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .table import AlignedSeqColumnData, FinalizedSeqColumnData, TableSpec, alphabet_chars

RDB_OK = 0
_STATUS = {1: "RDB_EINVAL", 2: "RDB_ENOMEM", 3: "RDB_EDEVICE", 4: "RDB_EQUERY", 5: "RDB_ECOMPILE"}

_vp, _u32, _u64, _i32, _f64, _cp, _int = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                                          ctypes.c_double, ctypes.c_char_p, ctypes.c_int)

# (name, restype, argtypes) of every entry point declared in include/rdb.h
SIGNATURES = [
    ("rdb_open", _int, [_int, ctypes.POINTER(_vp)]),
    ("rdb_close", None, [_vp]),
    ("rdb_last_error", _cp, [_vp]),
    ("rdb_table_create", _int, [_vp, _vp, _u32, ctypes.POINTER(_vp)]),
    ("rdb_table_free", None, [_vp]),
    ("rdb_table_add_sequence_column", _int, [_vp, _cp, _int, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("rdb_table_add_sequence_column_finalized", _int, [_vp, _cp, _int, _vp, _vp, _u32, _u64, _vp, _vp, _vp, _vp, _vp, _vp,
                                                        _vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    ("rdb_archive_scan_containers", _int, [_vp, _u64, _u64, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _vp]),
    ("rdb_archive_scan_row_bitmaps", _int, [_vp, _u64, _u64, _u64, _vp, _vp, _vp, _vp]),
    ("rdb_archive_bitmap_runs", _int, [_vp, _u64, _vp, _vp, _u64, _vp, _vp, _vp, _u64]),
    ("rdb_table_add_sequence_column_finalized_runs", _int, [_vp, _cp, _int, _vp, _vp, _u32, _u64, _vp, _vp, _vp, _vp, _vp,
                                                             _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("rdb_seq_builder_create", _int, [_vp, _cp, _int, _vp, _u32, ctypes.POINTER(_vp)]),
    ("rdb_seq_builder_append", _int, [_vp, _vp, _vp, _vp, _vp, _u64]),
    ("rdb_seq_builder_finish", _int, [_vp]),
    ("rdb_seq_builder_free", None, [_vp]),
    ("rdb_table_column_digest", _int, [_vp, _cp, _vp]),
    ("rdb_table_add_int_column", _int, [_vp, _cp, _int, _vp, _vp]),
    ("rdb_table_add_dict_column", _int, [_vp, _cp, _vp, _vp, _u32, _cp, _vp, _vp]),
    ("rdb_measure_read_bandwidth", _int, [_vp, _u64, _int, ctypes.POINTER(ctypes.c_double)]),
    ("rdb_table_add_lineage_column", _int, [_vp, _cp, _vp, _vp, _u32, _cp, _vp, _vp, _vp]),
    ("rdb_table_add_bitmap", _int, [_vp, _vp, _u64, _vp]),
    ("rdb_table_finalize", _int, [_vp]),
    ("rdb_table_local_reference", _int, [_vp, _cp, _vp]),
    ("rdb_table_stats", _int, [_vp, _cp, _vp, _vp, _vp]),
    ("rdb_table_build_profile", _int, [_vp, _cp, _vp, _u32]),
    ("rdb_filter", _int, [_vp, _cp, ctypes.POINTER(_vp), _vp]),
    ("rdb_bitmap_cardinality", _int, [_vp, _vp]),
    ("rdb_bitmap_rows", _int, [_vp, _vp, _u64, _vp]),
    ("rdb_bitmap_free", None, [_vp]),
    ("rdb_count", _int, [_vp, _cp, _vp]),
    ("rdb_mutation_counts", _int, [_vp, _vp, _cp, _vp]),
    ("rdb_mutations", _int, [_vp, _vp, _cp, _f64, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_mutations_query", _int, [_vp, _cp, _cp, _f64, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_mutations_query_noalloc", _int, [_vp, _cp, _cp, _f64, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_mutation_rows_ndjson", _int, [_int, _cp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, _vp]),
    ("rdb_mutation_rows_arrow_ipc", _int, [_int, _cp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp, _u64, _vp]),
    ("rdb_mutations_ndjson", _int, [_vp, _vp, _cp, _f64, _u32, _vp, _u64, _vp]),
    ("rdb_mutations_arrow_ipc", _int, [_vp, _vp, _cp, _f64, _u32, _vp, _u64, _vp]),
    ("rdb_aggregate", _int, [_vp, _vp, _cp, _vp, _vp, _u64, _vp]),
    ("rdb_profile", _int, [_vp, _int]),
    ("rdb_profile_read", _int, [_vp, _vp, _vp, _vp, _vp]),
    ("rdb_profile_read_scan", _int, [_vp, _vp, _vp, _vp]),
    ("rdb_mutations_traffic", _int, [_vp, _vp, _cp, _vp]),
    ("rdb_batch_submit", _int, [_vp, _vp, _u32, ctypes.POINTER(_vp)]),
    ("rdb_batch_wait", _int, [_vp]),
    ("rdb_batch_poll", _int, [_vp, _vp]),
    ("rdb_batch_lane_bytes", _int, [_vp, _u32, _u32, _vp]),
    ("rdb_device_memory", _int, [_vp, _vp, _vp]),
    ("rdb_batch_count", _int, [_vp, _u32, _vp]),
    ("rdb_batch_mutations", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_batch_aggregate", _int, [_vp, _u32, _vp, _vp, _u64, _vp]),
    ("rdb_batch_summary", _int, [_vp, _u32, _vp, _vp, _vp]),
    ("rdb_batch_mutations_all", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_batch_free", None, [_vp]),
    ("rdb_mlp_create", _int, [_vp, _u32, _u32, _u32, _u64, ctypes.POINTER(_vp)]),
    ("rdb_mlp_free", None, [_vp]),
    ("rdb_mlp_parameter_count", _int, [_vp, _vp]),
    ("rdb_mlp_get_parameters", _int, [_vp, _vp, _u64]),
    ("rdb_mlp_set_parameters", _int, [_vp, _vp, _u64]),
    ("rdb_mlp_forward", _int, [_vp, _vp, _u32, _vp]),
    ("rdb_mlp_gradients", _int, [_vp, _vp, _vp, _u32, _vp, _vp]),
    ("rdb_mlp_train", _int, [_vp, _vp, _vp, _u64, _u32, _u32, ctypes.c_float, _u64, _vp]),
    ("rdb_plan_samples", _int, [_vp, _vp, _u32, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_plan_estimates", _int, [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _u64, _vp]),
    ("rdb_set_planner", _int, [_vp, _int, _vp]),
    ("rdb_filter_explain", _int, [_vp, _cp, _vp, _u64, _vp]),
    ("rdb_set_tuning", _int, [_vp, _cp, ctypes.c_int64]),
    ("rdb_get_tuning", _int, [_vp, _cp, _vp]),
    ("rdb_diff_sequence", _int, [_int, _cp, _u32, _cp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _u64, _vp]),
]

class RdbQuery(ctypes.Structure):
    """struct rdb_query of include/rdb.h."""
    _fields_ = [("filter", ctypes.c_char_p), ("action", ctypes.c_int), ("column", ctypes.c_char_p),
                ("min_proportion", ctypes.c_double), ("dims", ctypes.c_char_p)]


ACTION_COUNT, ACTION_MUTATIONS, ACTION_AGGREGATE = 0, 1, 2

_bound = False


def lib() -> ctypes.CDLL:
    global _bound
    L = _native.load(_native.ENGINE_LIB)
    if not _bound:
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{_STATUS.get(code, code)}: {message}")
        self.code = code
        self.status = _STATUS.get(code, str(code))


def _check(rc: int, ctx) -> None:
    if rc != RDB_OK:
        raise EngineError(rc, lib().rdb_last_error(ctx).decode(errors="replace"))


def diff_sequence(sequence: str, reference: str, alphabet: int = 0):
    """The engine's restatement of extractCoverageAndMutationsFromSequence (offset 0)."""
    n = len(sequence)
    start, end = _u32(), _u32()
    miss = np.zeros(n + 1, np.uint32)
    mpos = np.zeros(n + 1, np.uint32)
    msym = np.zeros(n + 1, np.uint8)
    nm, nmut = _u64(), _u64()
    rc = lib().rdb_diff_sequence(alphabet, sequence.encode(), n, reference.encode(), ctypes.byref(start),
                                 ctypes.byref(end), _p(miss), n + 1, ctypes.byref(nm), _p(mpos), _p(msym), n + 1,
                                 ctypes.byref(nmut))
    _check(rc, None)
    return (start.value, end.value, miss[: nm.value].tolist(),
            list(zip(mpos[: nmut.value].tolist(), msym[: nmut.value].tolist())))


FIELDS_ALL = 127  # rdb_mutation_field: every output column of the Mutations action


def _sink_call(fn, head: list, ctx=None) -> bytes:
    """Calls a sink entry point twice: for the length, then into a buffer of that size."""
    n = _u64()
    rc = fn(*head, None, 0, ctypes.byref(n))
    if rc not in (0, 2):  # RDB_ENOMEM reports the length
        _check(rc, ctx)
    buf = ctypes.create_string_buffer(max(n.value, 1))
    _check(fn(*head, buf, n.value, ctypes.byref(n)), ctx)
    return buf.raw[: n.value]


def _row_arrays(rows: list):
    n = len(rows)
    pos = np.asarray([r[0] for r in rows], np.uint32)
    fr = np.asarray([r[1] for r in rows], np.uint8)
    to = np.asarray([r[2] for r in rows], np.uint8)
    cnt = np.asarray([r[3] for r in rows], np.uint32)
    cov = np.asarray([r[4] for r in rows], np.uint32)
    prop = np.asarray([r[5] for r in rows], np.float64)
    return n, (pos, fr, to, cnt, cov, prop)


def mutation_rows_ndjson(rows: list, sequence_name: str, alphabet: int = 0, fields: int = FIELDS_ALL) -> str:
    """Mutations rows (position, from, to, count, coverage, proportion) as the reference's NDJSON."""
    n, arrs = _row_arrays(rows)
    out = _sink_call(lib().rdb_mutation_rows_ndjson, [alphabet, sequence_name.encode(), fields] +
                     [_p(a) for a in arrs] + [n])
    return out.decode()


def mutation_rows_arrow_ipc(rows: list, sequence_name: str, alphabet: int = 0, fields: int = FIELDS_ALL) -> bytes:
    """Mutations rows as an Arrow IPC stream (schema + one record batch)."""
    n, arrs = _row_arrays(rows)
    return _sink_call(lib().rdb_mutation_rows_arrow_ipc, [alphabet, sequence_name.encode(), fields] +
                      [_p(a) for a in arrs] + [n])


class Engine:
    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().rdb_open(device, ctypes.byref(h)), None)
        self._h = h
        self._tables = []

    @property
    def handle(self):
        return self._h

    def table(self, spec: TableSpec) -> "DeviceTable":
        t = DeviceTable(self, spec)
        self._tables.append(t)
        return t

    def profile(self, enable: bool) -> None:
        """HIP-event timing of the Mutations counting kernel and the filter kernel (resets totals)."""
        _check(lib().rdb_profile(self._h, int(enable)), self._h)

    def profile_read(self) -> dict:
        mm, mn, fm, fn = _f64(), _u64(), _f64(), _u64()
        _check(lib().rdb_profile_read(self._h, ctypes.byref(mm), ctypes.byref(mn), ctypes.byref(fm), ctypes.byref(fn)),
               self._h)
        sm, sn, dq = _f64(), _u64(), _u64()
        _check(lib().rdb_profile_read_scan(self._h, ctypes.byref(sm), ctypes.byref(sn), ctypes.byref(dq)), self._h)
        return {"mutations_ms": mm.value, "mutations_launches": mn.value, "filter_ms": fm.value,
                "filter_launches": fn.value, "scan_ms": sm.value, "scan_launches": sn.value,
                "scan_dense_queries": dq.value}

    def read_bandwidth(self, n_bytes: int, reps: int = 20) -> float:
        """GB/s of a plain streaming read of n_bytes (rounded down to 16) on this GPU."""
        g = ctypes.c_double()
        _check(lib().rdb_measure_read_bandwidth(self._h, int(n_bytes) // 16 * 16, reps, ctypes.byref(g)), self._h)
        return g.value

    def device_memory(self) -> tuple:
        """(free, total) bytes of the GPU."""
        f, t = _u64(), _u64()
        _check(lib().rdb_device_memory(self._h, ctypes.byref(f), ctypes.byref(t)), self._h)
        return f.value, t.value

    def set_tuning(self, key: str, value: int) -> None:
        """Kernel knobs: "sparse_max_rows" (-1 = dense only), "mutation_groups" (0 = automatic)."""
        _check(lib().rdb_set_tuning(self._h, key.encode(), int(value)), self._h)

    def get_tuning(self, key: str) -> int:
        v = ctypes.c_int64()
        _check(lib().rdb_get_tuning(self._h, key.encode(), ctypes.byref(v)), self._h)
        return v.value

    def mlp(self, n_in: int = 32, n_hidden: int = 128, n_hidden_layers: int = 3, seed: int = 0) -> "Mlp":
        return Mlp(self, n_in, n_hidden, n_hidden_layers, seed)

    def set_planner(self, mode: str, mlp: "Mlp | None" = None) -> None:
        """mode: "reference" (default), "independence" or "learned" (with the estimator `mlp`)."""
        m = {"reference": 0, "independence": 1, "learned": 2}[mode]
        _check(lib().rdb_set_planner(self._h, m, None if mlp is None else mlp.handle), self._h)

    def close(self) -> None:
        for t in self._tables:
            t.close()
        self._tables = []
        if self._h:
            lib().rdb_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Bitmap:
    """A filter result resident on the device (the engine's CopyOnWriteBitmap)."""

    def __init__(self, table: "DeviceTable", handle, cardinality: int):
        self.table = table
        self._h = handle
        self.cardinality = cardinality

    @property
    def handle(self):
        return self._h

    def rows(self) -> np.ndarray:
        out = np.zeros(max(self.cardinality, 1), np.uint32)
        n = _u64()
        _check(lib().rdb_bitmap_rows(self._h, _p(out), len(out), ctypes.byref(n)), self.table.engine.handle)
        return out[: n.value]

    def free(self) -> None:
        if self._h:
            lib().rdb_bitmap_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class DeviceTable:
    def __init__(self, engine: Engine, spec: TableSpec):
        self.engine = engine
        L = lib()
        ctx = engine.handle
        sizes = np.ascontiguousarray(spec.chunk_sizes, dtype=np.uint32)
        h = _vp()
        _check(L.rdb_table_create(ctx, _p(sizes), len(sizes), ctypes.byref(h)), ctx)
        self._h = h
        self.seq_lengths = {}
        self.alphabets = {}
        self.dict_names = {}
        try:
            for c in spec.seq_columns:
                if isinstance(c, FinalizedSeqColumnData):
                    self._add_finalized(c)
                    continue
                if isinstance(c, AlignedSeqColumnData):
                    self._add_aligned(c)
                    continue
                arrs = [np.ascontiguousarray(x) for x in (
                    np.asarray(c.reference, np.uint8), np.asarray(c.is_null, np.uint8), np.asarray(c.cov_start, np.uint32),
                    np.asarray(c.cov_end, np.uint32), np.asarray(c.mut_ptr, np.uint64), np.asarray(c.mut_pos, np.uint32),
                    np.asarray(c.mut_sym, np.uint8), np.asarray(c.nrun_ptr, np.uint64), np.asarray(c.nrun_start, np.uint32),
                    np.asarray(c.nrun_end, np.uint32))]
                _check(L.rdb_table_add_sequence_column(h, c.name.encode(), c.alphabet, _p(arrs[0]), len(c.reference),
                                                       *[_p(a) for a in arrs[1:]]), ctx)
                self.seq_lengths[c.name] = len(c.reference)
                self.alphabets[c.name] = c.alphabet
            for c in spec.int_columns:
                v = np.ascontiguousarray(c.values, dtype=np.int32)
                nl = np.ascontiguousarray(c.is_null, dtype=np.uint8)
                _check(L.rdb_table_add_int_column(h, c.name.encode(), int(c.is_date), _p(v), _p(nl)), ctx)
            for c in spec.dict_columns:
                ids = np.ascontiguousarray(c.ids, dtype=np.uint32)
                nl = np.ascontiguousarray(c.is_null, dtype=np.uint8)
                blobs = [s.encode() for s in c.names]
                offs = np.zeros(len(blobs) + 1, dtype=np.uint32)
                offs[1:] = np.cumsum([len(b) for b in blobs])
                parent = None if c.lineage_parent is None else np.ascontiguousarray(c.lineage_parent, dtype=np.int32)
                alias = None if c.lineage_alias is None else np.ascontiguousarray(c.lineage_alias, dtype=np.int32)
                _check(L.rdb_table_add_lineage_column(h, c.name.encode(), _p(ids), _p(nl), len(c.names),
                                                      b"".join(blobs), _p(offs),
                                                      None if parent is None else _p(parent),
                                                      None if alias is None else _p(alias)), ctx)
                self.dict_names[c.name] = list(c.names)
            for b in spec.bitmaps:
                arr = np.ascontiguousarray(b, dtype=np.uint32)
                bid = _u32()
                _check(L.rdb_table_add_bitmap(h, _p(arr), len(arr), ctypes.byref(bid)), ctx)
            _check(L.rdb_table_finalize(h), ctx)
        except Exception:
            self.close()
            raise
        self.n_rows = spec.n_rows

    def _add_finalized(self, c: "FinalizedSeqColumnData") -> None:
        runs = c.missing_run_ptr is not None
        tail = ((c.missing_run_ptr, np.uint64), (c.missing_run_start, np.uint32), (c.missing_run_end, np.uint32)) if runs \
            else ((c.missing_ptr, np.uint64), (c.missing_pos, np.uint32))
        a = [np.ascontiguousarray(np.asarray(x, dt)) for x, dt in (
            (c.reference, np.uint8), (c.local_reference, np.uint8), (c.key_position, np.uint32),
            (c.key_chunk, np.uint16), (c.key_symbol, np.uint8), (c.typecode, np.uint8), (c.cardinality, np.uint32),
            (c.payload_offset, np.uint64), (c.payload, np.uint8), (c.is_null, np.uint8), (c.cov_start, np.uint32),
            (c.cov_end, np.uint32)) + tail]
        fn = lib().rdb_table_add_sequence_column_finalized_runs if runs else lib().rdb_table_add_sequence_column_finalized
        _check(fn(self._h, c.name.encode(), c.alphabet, _p(a[0]), _p(a[1]), len(c.reference), len(a[2]), _p(a[2]),
                  _p(a[3]), _p(a[4]), _p(a[5]), _p(a[6]), _p(a[7]), _p(a[8]), len(a[8]), _p(a[9]), _p(a[10]), _p(a[11]),
                  *[_p(x) for x in a[12:]]), self.engine.handle)
        self.seq_lengths[c.name] = len(c.reference)
        self.alphabets[c.name] = c.alphabet

    def _add_aligned(self, c: "AlignedSeqColumnData") -> None:
        L, ctx = lib(), self.engine.handle
        ref = np.ascontiguousarray(c.reference, np.uint8)
        b = _vp()
        _check(L.rdb_seq_builder_create(self._h, c.name.encode(), c.alphabet, _p(ref), len(ref), ctypes.byref(b)), ctx)
        try:
            batches = c.batches() if callable(c.batches) else c.batches
            for chars, row_ptr, offsets, is_null in batches:
                chars = np.ascontiguousarray(chars, np.uint8)
                row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
                offs = None if offsets is None else np.ascontiguousarray(offsets, np.uint32)
                nl = None if is_null is None else np.ascontiguousarray(is_null, np.uint8)
                _check(L.rdb_seq_builder_append(b, _p(chars), _p(row_ptr), None if offs is None else _p(offs),
                                                None if nl is None else _p(nl), len(row_ptr) - 1), ctx)
            _check(L.rdb_seq_builder_finish(b), ctx)
        finally:
            L.rdb_seq_builder_free(b)
        self.seq_lengths[c.name] = len(ref)
        self.alphabets[c.name] = c.alphabet

    def digest(self, column: str) -> int:
        """64-bit digest of a sequence column's device layout (equal layouts, equal digests)."""
        d = _u64()
        _check(lib().rdb_table_column_digest(self._h, column.encode(), ctypes.byref(d)), self._ctx())
        return int(d.value)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            lib().rdb_table_free(self._h)
            self._h = None

    def _ctx(self):
        return self.engine.handle

    def bitmap(self, expr: str) -> Bitmap:
        h = _vp()
        card = _u64()
        _check(lib().rdb_filter(self._h, expr.encode(), ctypes.byref(h), ctypes.byref(card)), self._ctx())
        return Bitmap(self, h, card.value)

    def filter(self, expr: str) -> np.ndarray:
        with self.bitmap(expr) as b:
            return b.rows()

    def count(self, expr: str) -> int:
        n = _u64()
        _check(lib().rdb_count(self._h, expr.encode(), ctypes.byref(n)), self._ctx())
        return n.value

    def explain(self, expr: str) -> str:
        """The compiled operator tree of the filter as text (Operator::toString)."""
        n = _u64()
        lib().rdb_filter_explain(self._h, expr.encode(), None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(lib().rdb_filter_explain(self._h, expr.encode(), buf, n.value + 1, ctypes.byref(n)), self._ctx())
        return buf.value.decode()

    def local_reference(self, column: str) -> np.ndarray:
        out = np.zeros(self.seq_lengths[column], np.uint8)
        _check(lib().rdb_table_local_reference(self._h, column.encode(), _p(out)), self._ctx())
        return out

    def stats(self, column: str | None = None) -> dict:
        db, nc, pb = _u64(), _u64(), _u64()
        _check(lib().rdb_table_stats(self._h, None if column is None else column.encode(), ctypes.byref(db),
                                     ctypes.byref(nc), ctypes.byref(pb)), self._ctx())
        return {"device_bytes": db.value, "containers": nc.value, "payload_bytes": pb.value}

    BUILD_PHASES = ("checks", "adaptation", "containers", "tiles", "row_arrays", "layout_checks", "upload")

    def build_profile(self, column: str) -> dict:
        """Wall time (ms) of the phases of a sequence column's index build."""
        ms = np.zeros(len(self.BUILD_PHASES), np.float64)
        _check(lib().rdb_table_build_profile(self._h, column.encode(), _p(ms), len(ms)), self._ctx())
        return dict(zip(self.BUILD_PHASES, ms.tolist()))

    def _with_bitmap(self, filt, fn):
        if isinstance(filt, Bitmap):
            return fn(filt)
        with self.bitmap(filt) as b:
            return fn(b)

    def mutations_traffic(self, filt, column: str) -> int:
        """Algorithmic bytes of one Mutations evaluation under the filter."""
        n = _u64()

        def run(b):
            _check(lib().rdb_mutations_traffic(self._h, b.handle, column.encode(), ctypes.byref(n)), self._ctx())
            return n.value
        return self._with_bitmap(filt, run)

    def mutation_counts(self, filt, column: str) -> np.ndarray:
        n_sym = len(alphabet_chars(self.alphabets[column]))
        out = np.zeros((n_sym, self.seq_lengths[column]), np.uint32)

        def run(b):
            _check(lib().rdb_mutation_counts(self._h, b.handle, column.encode(), _p(out)), self._ctx())
            return out
        return self._with_bitmap(filt, run)

    def _row_buffers(self, column: str):
        cap = 32 * self.seq_lengths.get(column, 0) + 16  # an unknown column fails in the call itself
        key = (column, cap)
        bufs = getattr(self, "_bufs", {})
        if key not in bufs:
            bufs[key] = (np.zeros(cap, np.uint32), np.zeros(cap, np.uint8), np.zeros(cap, np.uint8),
                         np.zeros(cap, np.uint32), np.zeros(cap, np.uint32), np.zeros(cap, np.float64))
            self._bufs = bufs
        return cap, bufs[key]

    def mutations_query(self, expr: str, column: str, min_proportion: float, arrays: bool = False):
        """Filter + Mutations in one call (one device round trip).  Rows as (position, from, to,
        count, coverage, proportion) tuples, or with `arrays` as one MutationArrays of columns."""
        cap, bufs = self._row_buffers(column)
        pos, fr, to, cnt, cov, prop = bufs
        n = _u64()
        _check(lib().rdb_mutations_query(self._h, expr.encode(), column.encode(), float(min_proportion), _p(pos), _p(fr),
                                         _p(to), _p(cnt), _p(cov), _p(prop), cap, ctypes.byref(n)), self._ctx())
        return _mutation_rows(bufs, n.value, arrays)

    def mutations_ndjson(self, filt, column: str, min_proportion: float, fields: int = FIELDS_ALL) -> str:
        def run(b):
            return _sink_call(lib().rdb_mutations_ndjson, [self._h, b.handle, column.encode(), float(min_proportion),
                                                           fields], self._ctx()).decode()
        return self._with_bitmap(filt, run)

    def mutations_arrow_ipc(self, filt, column: str, min_proportion: float, fields: int = FIELDS_ALL) -> bytes:
        def run(b):
            return _sink_call(lib().rdb_mutations_arrow_ipc, [self._h, b.handle, column.encode(), float(min_proportion),
                                                              fields], self._ctx())
        return self._with_bitmap(filt, run)

    def mutations(self, filt, column: str, min_proportion: float) -> list:
        cap, (pos, fr, to, cnt, cov, prop) = self._row_buffers(column)
        n = _u64()

        def run(b):
            _check(lib().rdb_mutations(self._h, b.handle, column.encode(), float(min_proportion), _p(pos), _p(fr), _p(to),
                                       _p(cnt), _p(cov), _p(prop), cap, ctypes.byref(n)), self._ctx())
            k = n.value
            return list(zip(pos[:k].tolist(), fr[:k].tolist(), to[:k].tolist(), cnt[:k].tolist(), cov[:k].tolist(),
                            prop[:k].tolist()))
        return self._with_bitmap(filt, run)

    def aggregate(self, filt, dims: list) -> list:
        cap = 1 << 20
        nd = max(len(dims), 1)
        ids = np.zeros(cap * nd, np.uint32)
        counts = np.zeros(cap, np.uint64)
        n = _u64()

        def run(b):
            _check(lib().rdb_aggregate(self._h, b.handle, ",".join(dims).encode(), _p(ids), _p(counts), cap,
                                       ctypes.byref(n)), self._ctx())
            out = []
            for i in range(n.value):
                key = tuple(None if int(x) == 0xFFFFFFFF else int(x) for x in ids[i * len(dims):(i + 1) * len(dims)])
                out.append((key, int(counts[i])))
            return out
        return self._with_bitmap(filt, run)

    def lane_bytes(self, max_queries: int, row_capacity: int) -> int:
        n = _u64()
        _check(lib().rdb_batch_lane_bytes(self._h, max_queries, row_capacity, ctypes.byref(n)), self._ctx())
        return n.value

    # ---- batched path --------------------------------------------------------------------------
    def batch_submit(self, queries: list) -> "Batch":
        """Submits dicts {"filter", "action": "count"|"mutations"|"aggregate", "column",
        "min_proportion", "dims": [...]} as one batch; returns without waiting."""
        arr = (RdbQuery * max(len(queries), 1))()
        keep = []
        for i, q in enumerate(queries):
            act = {"count": ACTION_COUNT, "mutations": ACTION_MUTATIONS, "aggregate": ACTION_AGGREGATE}[q["action"]]
            f = q["filter"].encode()
            col = q.get("column", "").encode()
            dims = ",".join(q.get("dims", [])).encode()
            keep += [f, col, dims]
            arr[i] = RdbQuery(f, act, col, float(q.get("min_proportion", 0.0)), dims)
        h = _vp()
        _check(lib().rdb_batch_submit(self._h, arr, len(queries), ctypes.byref(h)), self._ctx())
        return Batch(self, h, queries)

    def batch(self, queries: list, arrays: bool = False) -> list:
        """Runs a batch and returns per query: its cardinality, Mutations rows (columns with
        `arrays`) or groups (or the EngineError of a query that failed)."""
        with self.batch_submit(queries) as b:
            return b.results(arrays)


class MutationArrays(tuple):
    """Mutations rows as columns: position (u32), from and to symbol (u8), count and coverage (u32),
    proportion (f64), each a numpy array of the row count."""

    def __new__(cls, cols):
        return super().__new__(cls, cols)

    def rows(self) -> list:
        return list(zip(*(c.tolist() for c in self)))


def _mutation_rows(bufs, k: int, arrays: bool):
    if arrays:
        return MutationArrays(tuple(b[:k].copy() for b in bufs))
    pos, fr, to, cnt, cov, prop = bufs
    return list(zip(pos[:k].tolist(), fr[:k].tolist(), to[:k].tolist(), cnt[:k].tolist(), cov[:k].tolist(),
                    prop[:k].tolist()))


class Batch:
    def __init__(self, table: DeviceTable, handle, queries: list):
        self.table = table
        self._h = handle
        self.queries = queries

    def wait(self) -> None:
        _check(lib().rdb_batch_wait(self._h), self.table._ctx())

    def poll(self) -> bool:
        """True once every job of the batch has finished (collects finished jobs, never waits)."""
        d = ctypes.c_int()
        _check(lib().rdb_batch_poll(self._h, ctypes.byref(d)), self.table._ctx())
        return bool(d.value)

    def count(self, q: int) -> int:
        n = _u64()
        _check(lib().rdb_batch_count(self._h, q, ctypes.byref(n)), self.table._ctx())
        return n.value

    def mutations(self, q: int, arrays: bool = False):
        cap, bufs = self.table._row_buffers(self.queries[q]["column"])
        pos, fr, to, cnt, cov, prop = bufs
        n = _u64()
        _check(lib().rdb_batch_mutations(self._h, q, _p(pos), _p(fr), _p(to), _p(cnt), _p(cov), _p(prop), cap,
                                         ctypes.byref(n)), self.table._ctx())
        return _mutation_rows(bufs, n.value, arrays)

    def aggregate(self, q: int) -> list:
        dims = self.queries[q].get("dims", [])
        cap = 1 << 20
        ids = np.zeros(cap * max(len(dims), 1), np.uint32)
        counts = np.zeros(cap, np.uint64)
        n = _u64()
        _check(lib().rdb_batch_aggregate(self._h, q, _p(ids), _p(counts), cap, ctypes.byref(n)), self.table._ctx())
        out = []
        for i in range(n.value):
            key = tuple(None if int(x) == 0xFFFFFFFF else int(x) for x in ids[i * len(dims):(i + 1) * len(dims)])
            out.append((key, int(counts[i])))
        return out

    def results(self, arrays: bool = False) -> list:
        """Every query's result in submission order (Mutations rows as in DeviceTable.mutations_query,
        as columns with `arrays`: then every Mutations query's rows come back in one call)."""
        if arrays:
            return self._results_columns()
        out = []
        for i, q in enumerate(self.queries):
            try:
                if q["action"] == "mutations":
                    out.append(self.mutations(i, arrays))
                elif q["action"] == "aggregate":
                    out.append(self.aggregate(i))
                else:
                    out.append(self.count(i))
            except EngineError as e:
                out.append(e)
        return out

    def _results_columns(self) -> list:
        n = len(self.queries)
        status = np.zeros(n, np.int32)
        card = np.zeros(n, np.uint64)
        size = np.zeros(n, np.uint64)
        ctx = self.table._ctx()
        _check(lib().rdb_batch_summary(self._h, n, _p(status), _p(card), _p(size)), ctx)
        is_mut = np.array([q["action"] == "mutations" for q in self.queries], bool)
        total = int(size[is_mut & (status == RDB_OK)].sum())
        offsets = np.zeros(n + 1, np.uint64)
        cols = (np.empty(total, np.uint32), np.empty(total, np.uint8), np.empty(total, np.uint8),
                np.empty(total, np.uint32), np.empty(total, np.uint32), np.empty(total, np.float64))
        got = _u64()
        _check(lib().rdb_batch_mutations_all(self._h, n, _p(offsets), *[_p(c) for c in cols], total, ctypes.byref(got)),
               ctx)
        out = []
        for i, q in enumerate(self.queries):
            if status[i] != RDB_OK:
                try:
                    self.count(i)  # raises the query's own error
                except EngineError as e:
                    out.append(e)
                continue
            if q["action"] == "mutations":
                a, b = int(offsets[i]), int(offsets[i + 1])
                out.append(MutationArrays(tuple(c[a:b] for c in cols)))
            elif q["action"] == "aggregate":
                out.append(self.aggregate(i))
            else:
                out.append(int(card[i]))
        return out

    def free(self) -> None:
        if self._h:
            lib().rdb_batch_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class Mlp:
    """The planner's cardinality estimator (BF16 parameters, FP32 accumulation, HIP kernels)."""

    def __init__(self, engine: Engine, n_in: int, n_hidden: int, n_hidden_layers: int, seed: int):
        self.engine = engine
        self.n_in = n_in
        h = _vp()
        _check(lib().rdb_mlp_create(engine.handle, n_in, n_hidden, n_hidden_layers, seed, ctypes.byref(h)), engine.handle)
        self._h = h
        n = _u64()
        _check(lib().rdb_mlp_parameter_count(h, ctypes.byref(n)), engine.handle)
        self.n_params = n.value

    @property
    def handle(self):
        return self._h

    def parameters(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(lib().rdb_mlp_get_parameters(self._h, _p(p), self.n_params), self.engine.handle)
        return p

    def set_parameters(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().rdb_mlp_set_parameters(self._h, _p(p), len(p)), self.engine.handle)

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(len(x), np.float32)
        _check(lib().rdb_mlp_forward(self._h, _p(x), len(x), _p(y)), self.engine.handle)
        return y

    def gradients(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(lib().rdb_mlp_gradients(self._h, _p(x), _p(y), len(x), ctypes.byref(loss), _p(g)), self.engine.handle)
        return loss.value, g

    def train(self, x: np.ndarray, y: np.ndarray, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().rdb_mlp_train(self._h, _p(x), _p(y), len(x), epochs, batch, ctypes.c_float(lr), seed,
                                   ctypes.byref(loss)), self.engine.handle)
        return loss.value

    def free(self) -> None:
        if self._h:
            lib().rdb_mlp_free(self._h)
            self._h = None


def _cstrings(items):
    arr = (ctypes.c_char_p * max(len(items), 1))()
    enc = [s.encode() for s in items]
    for i, s in enumerate(enc):
        arr[i] = s
    return arr, enc


def plan_samples(table: DeviceTable, filters: list):
    """Training rows of the planner: (features [n, 32], targets [n], tree index [n])."""
    arr, keep = _cstrings(filters)
    n = _u64()
    rc = lib().rdb_plan_samples(table.handle, arr, len(filters), None, None, None, 0, ctypes.byref(n))
    if rc not in (RDB_OK, 2):
        _check(rc, table._ctx())
    cap = n.value
    f = np.zeros((cap, 32), np.float32)
    y = np.zeros(cap, np.float32)
    tr = np.zeros(cap, np.uint32)
    _check(lib().rdb_plan_samples(table.handle, arr, len(filters), _p(f), _p(y), _p(tr), cap, ctypes.byref(n)),
           table._ctx())
    return f, y, tr


def plan_estimates(table: DeviceTable, filters: list, mlp: "Mlp | None" = None, exact: bool = True):
    """Per node of every filter: learned estimate, independence estimate, exact rows, tree index."""
    arr, keep = _cstrings(filters)
    n = _u64()
    rc = lib().rdb_plan_estimates(table.handle, None if mlp is None else mlp.handle, arr, len(filters), None, None,
                                  None, None, 0, ctypes.byref(n))
    if rc not in (RDB_OK, 2):
        _check(rc, table._ctx())
    cap = n.value
    est = np.zeros(cap, np.float32)
    ind = np.zeros(cap, np.float32)
    ex = np.zeros(cap, np.uint64)
    tr = np.zeros(cap, np.uint32)
    _check(lib().rdb_plan_estimates(table.handle, None if mlp is None else mlp.handle, arr, len(filters), _p(est), _p(ind),
                                    _p(ex) if exact else None, _p(tr), cap, ctypes.byref(n)), table._ctx())
    return est, ind, (ex if exact else None), tr
