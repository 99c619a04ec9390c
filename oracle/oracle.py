"""ORACLE — test infrastructure only.  ctypes wrapper around `oracle/liboracle.so`.

Only tests, `__graft_entry__.smoke()` and the CPU-baseline leg of `bench.py` use this module, and
only as the checker (or as the timed CPU restatement); the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32, f64, cp, ip = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                                          ctypes.c_double, ctypes.c_char_p, ctypes.c_int)
        L.or_table_create.restype = vp
        L.or_table_create.argtypes = [vp, u32]
        L.or_table_free.argtypes = [vp]
        L.or_last_error.restype = cp
        L.or_last_error.argtypes = [vp]
        L.or_add_sequence_column.restype = ip
        L.or_add_sequence_column.argtypes = [vp, cp, ip, vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.or_add_int_column.restype = ip
        L.or_add_int_column.argtypes = [vp, cp, ip, vp, vp]
        L.or_add_dict_column.restype = ip
        L.or_add_dict_column.argtypes = [vp, cp, vp, vp, u32, cp, vp, vp]
        L.or_add_bitmap.restype = ip
        L.or_add_bitmap.argtypes = [vp, vp, u64]
        L.or_local_reference.restype = ip
        L.or_local_reference.argtypes = [vp, cp, vp]
        L.or_export_sequence_column.restype = ip
        L.or_export_sequence_column.argtypes = [vp, cp, vp, vp, vp] + [vp] * 13
        L.or_index_stats.restype = ip
        L.or_index_stats.argtypes = [vp, cp, vp, vp]
        L.or_filter.restype = ip
        L.or_filter.argtypes = [vp, cp, vp, u64, vp]
        L.or_filter_root_type.restype = ip
        L.or_filter_root_type.argtypes = [vp, cp, vp]
        L.or_mutation_counts.restype = ip
        L.or_mutation_counts.argtypes = [vp, cp, cp, vp]
        L.or_mutations.restype = ip
        L.or_mutations.argtypes = [vp, cp, cp, f64, vp, vp, vp, vp, vp, vp, u64, vp]
        L.or_aggregate.restype = ip
        L.or_aggregate.argtypes = [vp, cp, cp, vp, vp, u64, vp]
        L.or_diff_sequence.restype = ip
        L.or_diff_sequence.argtypes = [ip, cp, u32, cp, vp, vp, vp, u64, vp, vp, vp, u64, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


class OracleTable:
    """The oracle's in-memory table built from a `TableSpec` (same ingest arrays as the engine)."""

    def __init__(self, spec):
        L = lib()
        sizes = np.ascontiguousarray(spec.chunk_sizes, dtype=np.uint32)
        self._h = L.or_table_create(_p(sizes), len(sizes))
        if not self._h:
            raise OracleError("invalid row layout")
        self.seq_lengths = {}
        self.alphabets = {}
        self.references = {}
        self.dict_names = {}
        for c in spec.seq_columns:
            arrs = [np.ascontiguousarray(x) for x in (c.reference, c.is_null, c.cov_start, c.cov_end, c.mut_ptr,
                                                      c.mut_pos, c.mut_sym, c.nrun_ptr, c.nrun_start, c.nrun_end)]
            rc = L.or_add_sequence_column(self._h, c.name.encode(), c.alphabet, _p(arrs[0]), len(c.reference),
                                          *[_p(a) for a in arrs[1:]])
            self._check(rc)
            self.seq_lengths[c.name] = len(c.reference)
            self.alphabets[c.name] = c.alphabet
            self.references[c.name] = np.asarray(c.reference, np.uint8)
        for c in spec.int_columns:
            v = np.ascontiguousarray(c.values, dtype=np.int32)
            n = np.ascontiguousarray(c.is_null, dtype=np.uint8)
            self._check(L.or_add_int_column(self._h, c.name.encode(), int(c.is_date), _p(v), _p(n)))
        for c in spec.dict_columns:
            ids = np.ascontiguousarray(c.ids, dtype=np.uint32)
            n = np.ascontiguousarray(c.is_null, dtype=np.uint8)
            blobs = [s.encode() for s in c.names]
            offs = np.zeros(len(blobs) + 1, dtype=np.uint32)
            offs[1:] = np.cumsum([len(b) for b in blobs])
            if getattr(c, "lineage_alias", None) is not None and np.any(np.asarray(c.lineage_alias) >= 0):
                raise OracleError("lineage aliases are checked against oracle/brute.py, not this oracle")
            parent = None if c.lineage_parent is None else np.ascontiguousarray(c.lineage_parent, dtype=np.int32)
            rc = L.or_add_dict_column(self._h, c.name.encode(), _p(ids), _p(n), len(c.names), b"".join(blobs),
                                      _p(offs), None if parent is None else _p(parent))
            self._check(rc)
            self.dict_names[c.name] = list(c.names)
        for b in spec.bitmaps:
            arr = np.ascontiguousarray(b, dtype=np.uint32)
            L.or_add_bitmap(self._h, _p(arr), len(arr))
        self.n_rows = spec.n_rows

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise OracleError(lib().or_last_error(self._h).decode())

    def close(self) -> None:
        if self._h:
            lib().or_table_free(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def local_reference(self, column: str) -> np.ndarray:
        out = np.zeros(self.seq_lengths[column], dtype=np.uint8)
        self._check(lib().or_local_reference(self._h, column.encode(), _p(out)))
        return out

    def export_finalized(self, column: str):
        """The column's finalized storage (the state a finalized reference SequenceColumn holds) as
        silo_amd.table.FinalizedSeqColumnData."""
        from silo_amd.table import FinalizedSeqColumnData

        nc, nb, nm = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        c = column.encode()
        self._check(lib().or_export_sequence_column(self._h, c, ctypes.byref(nc), ctypes.byref(nb), ctypes.byref(nm),
                                                    *([None] * 13)))
        k, b, m, n = nc.value, nb.value, nm.value, self.n_rows
        out = dict(local_reference=np.zeros(self.seq_lengths[column], np.uint8), key_position=np.zeros(k, np.uint32),
                   key_chunk=np.zeros(k, np.uint16), key_symbol=np.zeros(k, np.uint8), typecode=np.zeros(k, np.uint8),
                   cardinality=np.zeros(k, np.uint32), payload_offset=np.zeros(k, np.uint64),
                   payload=np.zeros(max(b, 1), np.uint8), is_null=np.zeros(n, np.uint8),
                   cov_start=np.zeros(n, np.uint32), cov_end=np.zeros(n, np.uint32),
                   missing_ptr=np.zeros(n + 1, np.uint64), missing_pos=np.zeros(max(m, 1), np.uint32))
        order = ("local_reference", "key_position", "key_chunk", "key_symbol", "typecode", "cardinality",
                 "payload_offset", "payload", "is_null", "cov_start", "cov_end", "missing_ptr", "missing_pos")
        self._check(lib().or_export_sequence_column(self._h, c, ctypes.byref(nc), ctypes.byref(nb), ctypes.byref(nm),
                                                    *[_p(out[key]) for key in order]))
        out["payload"] = out["payload"][:b]
        out["missing_pos"] = out["missing_pos"][:m]
        return FinalizedSeqColumnData(name=column, alphabet=self.alphabets[column],
                                      reference=self.references[column].copy(), **out)

    def index_stats(self, column: str) -> tuple[int, int]:
        n = ctypes.c_uint64()
        b = ctypes.c_uint64()
        self._check(lib().or_index_stats(self._h, column.encode(), ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def filter(self, expr: str) -> np.ndarray:
        card = ctypes.c_uint64()
        self._check(lib().or_filter(self._h, expr.encode(), None, 0, ctypes.byref(card)))
        out = np.zeros(max(card.value, 1), dtype=np.uint32)
        self._check(lib().or_filter(self._h, expr.encode(), _p(out), card.value, ctypes.byref(card)))
        return out[: card.value]

    OP_TYPES = ("Empty", "Full", "IndexScan", "Intersection", "Complement", "RangeSelection", "Selection",
                "Threshold", "Union")

    def root_operator(self, expr: str) -> str:
        """Name of the operator the filter compiles to (its root)."""
        k = ctypes.c_int()
        self._check(lib().or_filter_root_type(self._h, expr.encode(), ctypes.byref(k)))
        return self.OP_TYPES[k.value]

    def count(self, expr: str) -> int:
        card = ctypes.c_uint64()
        self._check(lib().or_filter(self._h, expr.encode(), None, 0, ctypes.byref(card)))
        return card.value

    def mutation_counts(self, expr: str, column: str) -> np.ndarray:
        n_sym = 16 if self.alphabets[column] == 0 else 28
        out = np.zeros((n_sym, self.seq_lengths[column]), dtype=np.uint32)
        self._check(lib().or_mutation_counts(self._h, expr.encode(), column.encode(), _p(out)))
        return out

    def mutations(self, expr: str, column: str, min_proportion: float) -> list:
        cap = 32 * self.seq_lengths[column] + 16
        pos = np.zeros(cap, np.uint32)
        fr = np.zeros(cap, np.uint8)
        to = np.zeros(cap, np.uint8)
        cnt = np.zeros(cap, np.uint32)
        cov = np.zeros(cap, np.uint32)
        prop = np.zeros(cap, np.float64)
        n = ctypes.c_uint64()
        self._check(lib().or_mutations(self._h, expr.encode(), column.encode(), float(min_proportion), _p(pos),
                                       _p(fr), _p(to), _p(cnt), _p(cov), _p(prop), cap, ctypes.byref(n)))
        k = n.value
        return [(int(pos[i]), int(fr[i]), int(to[i]), int(cnt[i]), int(cov[i]), float(prop[i])) for i in range(k)]

    def aggregate(self, expr: str, dims: list) -> list:
        cap = 1 << 20
        nd = max(len(dims), 1)
        ids = np.zeros(cap * nd, np.uint32)
        counts = np.zeros(cap, np.uint64)
        n = ctypes.c_uint64()
        self._check(lib().or_aggregate(self._h, expr.encode(), ",".join(dims).encode(), _p(ids), _p(counts), cap,
                                       ctypes.byref(n)))
        k = n.value
        out = []
        for i in range(k):
            key = tuple(None if int(x) == 0xFFFFFFFF else int(x) for x in ids[i * len(dims):(i + 1) * len(dims)])
            out.append((key, int(counts[i])))
        return out


def diff_sequence(sequence: str, reference: str, alphabet: int = 0):
    """Oracle restatement of `extractCoverageAndMutationsFromSequence` (offset 0)."""
    n = len(sequence)
    start = ctypes.c_uint32()
    end = ctypes.c_uint32()
    miss = np.zeros(n + 1, np.uint32)
    mpos = np.zeros(n + 1, np.uint32)
    msym = np.zeros(n + 1, np.uint8)
    nm = ctypes.c_uint64()
    nmut = ctypes.c_uint64()
    rc = lib().or_diff_sequence(alphabet, sequence.encode(), n, reference.encode(), ctypes.byref(start),
                                ctypes.byref(end), _p(miss), n + 1, ctypes.byref(nm), _p(mpos), _p(msym), n + 1,
                                ctypes.byref(nmut))
    if rc != 0:
        raise OracleError(lib().or_last_error(None).decode())
    return (start.value, end.value, miss[: nm.value].tolist(),
            list(zip(mpos[: nmut.value].tolist(), msym[: nmut.value].tolist())))
