"""Column data handed to the engine at upload time (the post-diff ingest form).

A `TableSpec` mirrors what the reference's `Table` holds after ingestion and before `finalize()`
(`storage/table.h:16-68`, `table.cpp:75-94`): a row layout of 2^16-row chunks, sequence columns in
reference-diffed form, and int/date/dictionary metadata columns.  Values are dense in row order;
row `i` of chunk `c` has the global id `c << 16 | i`.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .datagen import CHUNK, Dataset, chunk_sizes

NUC_CHARS = "-ACGTRYSWKMBDHVN"
AA_CHARS = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X"
ALPHABET_NUC = 0
ALPHABET_AA = 1


def alphabet_chars(alphabet: int) -> str:
    return NUC_CHARS if alphabet == ALPHABET_NUC else AA_CHARS


def symbol_of(alphabet: int, ch: str) -> int:
    c = ch.upper()
    if alphabet == ALPHABET_NUC and c == "U":
        c = "T"
    return alphabet_chars(alphabet).index(c)


@dataclass
class SeqColumnData:
    name: str
    alphabet: int
    reference: np.ndarray  # u8 symbols
    is_null: np.ndarray
    cov_start: np.ndarray
    cov_end: np.ndarray
    mut_ptr: np.ndarray
    mut_pos: np.ndarray
    mut_sym: np.ndarray
    nrun_ptr: np.ndarray
    nrun_start: np.ndarray
    nrun_end: np.ndarray


@dataclass
class FinalizedSeqColumnData:
    """A sequence column as the reference's SequenceColumn holds it after finalize
    (`storage/column/sequence_column.h:86-97`): the local reference, the vertical index's containers
    keyed (position, chunk, symbol) with cardinality, CRoaring typecode (1 bitset, 2 array, 3 run)
    and `container_write` bytes, and per row the null flag, the covered range and the missing
    positions inside it (CSR)."""
    name: str
    alphabet: int
    reference: np.ndarray        # u8 symbols (the column's reference)
    local_reference: np.ndarray  # u8 symbols
    key_position: np.ndarray     # u32
    key_chunk: np.ndarray        # u16
    key_symbol: np.ndarray       # u8
    typecode: np.ndarray         # u8
    cardinality: np.ndarray      # u32
    payload_offset: np.ndarray   # u64 byte offsets into payload
    payload: np.ndarray          # u8
    is_null: np.ndarray
    cov_start: np.ndarray
    cov_end: np.ndarray
    missing_ptr: np.ndarray | None  # u64 [n_rows + 1]
    missing_pos: np.ndarray | None  # u32
    # or the missing positions as runs [start, end) per row, as a saved state holds them
    missing_run_ptr: np.ndarray | None = None    # u64 [n_rows + 1]
    missing_run_start: np.ndarray | None = None  # u32
    missing_run_end: np.ndarray | None = None    # u32

    def missing_runs(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(ptr, start, end) of every row's missing runs, from whichever form the column holds."""
        if self.missing_run_ptr is not None:
            return (np.asarray(self.missing_run_ptr, np.uint64), np.asarray(self.missing_run_start, np.uint32),
                    np.asarray(self.missing_run_end, np.uint32))
        ptr = np.asarray(self.missing_ptr, np.int64)
        pos = np.asarray(self.missing_pos, np.int64)[: int(ptr[-1])]
        row = np.repeat(np.arange(len(ptr) - 1, dtype=np.int64), np.diff(ptr))
        first = np.ones(len(pos), bool)
        first[1:] = (row[1:] != row[:-1]) | (pos[1:] != pos[:-1] + 1)
        starts = np.nonzero(first)[0]
        ends = np.concatenate([starts[1:], [len(pos)]])
        run_ptr = np.zeros(len(ptr), np.uint64)
        np.cumsum(np.bincount(row[starts], minlength=len(ptr) - 1), out=run_ptr[1:])
        return run_ptr, pos[starts].astype(np.uint32), (pos[ends - 1] + 1).astype(np.uint32)

    def missing_positions(self) -> tuple[np.ndarray, np.ndarray]:
        """(ptr, positions) of every row's missing positions, from whichever form the column holds."""
        if self.missing_run_ptr is None:
            return np.asarray(self.missing_ptr, np.uint64), np.asarray(self.missing_pos, np.uint32)
        rp = np.asarray(self.missing_run_ptr, np.int64)
        s = np.asarray(self.missing_run_start, np.int64)
        e = np.asarray(self.missing_run_end, np.int64)
        length = e - s
        row = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp))
        ptr = np.zeros(len(rp), np.uint64)
        np.cumsum(np.bincount(row, weights=length, minlength=len(rp) - 1).astype(np.int64), out=ptr[1:])
        total = int(length.sum())
        base = np.repeat(s - np.concatenate([[0], np.cumsum(length)[:-1]]), length)
        return ptr, (base + np.arange(total, dtype=np.int64)).astype(np.uint32)


@dataclass
class AlignedSeqColumnData:
    """A sequence column given as aligned sequences, diffed and indexed on the GPU by the column
    builder (the reference's SequenceColumnBuilder insert / insertNull, storage/column/
    sequence_column.h:178-233).  `batches` yields, in row order, tuples (chars u8, row_ptr u64
    [n + 1], offsets u32 or None, is_null u8 or None): row i of a batch is chars[row_ptr[i] ..
    row_ptr[i + 1]) placed from position offsets[i]."""
    name: str
    alphabet: int
    reference: np.ndarray  # u8 symbols
    batches: object        # an iterable of batches, or a callable returning one


def render_rows(col: SeqColumnData, row0: int, n: int) -> np.ndarray:
    """Rows [row0, row0 + n) of a column in diff form as aligned characters, shape (n, L) u8: the
    reference's characters with the mutations, and the missing symbol's character outside the
    covered range, inside missing runs and on null rows."""
    from . import datagen

    chars = alphabet_chars(col.alphabet)
    missing = "N" if col.alphabet == ALPHABET_NUC else "X"
    L = len(col.reference)
    out = np.empty((n, L), np.uint8)
    arrs = [np.ascontiguousarray(np.asarray(a, dt)) for a, dt in (
        (col.reference, np.uint8), (col.is_null, np.uint8), (col.cov_start, np.uint32), (col.cov_end, np.uint32),
        (col.mut_ptr, np.uint64), (col.mut_pos, np.uint32), (col.mut_sym, np.uint8), (col.nrun_ptr, np.uint64),
        (col.nrun_start, np.uint32), (col.nrun_end, np.uint32))]
    if n:
        datagen._lib().dg_render(L, arrs[0].ctypes.data, chars.encode(), missing.encode(), row0, n,
                                 *[a.ctypes.data for a in arrs[1:]], out.ctypes.data)
    return out


def aligned_batches(col: SeqColumnData, batch_rows: int = 65536):
    """The rows of a column in diff form as AlignedSeqColumnData batches of full-length rows."""
    L = len(col.reference)
    n = len(col.is_null)
    for a in range(0, n, batch_rows):
        m = min(batch_rows, n - a)
        chars = render_rows(col, a, m)
        row_ptr = np.arange(m + 1, dtype=np.uint64) * np.uint64(L)
        yield chars.reshape(-1), row_ptr, None, np.ascontiguousarray(col.is_null[a:a + m], np.uint8)


@dataclass
class IntColumnData:
    name: str
    is_date: bool
    values: np.ndarray  # i32
    is_null: np.ndarray  # u8


@dataclass
class DictColumnData:
    name: str
    ids: np.ndarray  # u32
    is_null: np.ndarray  # u8
    names: list
    lineage_parent: np.ndarray | None = None  # i32 per value, -1 = root
    lineage_alias: np.ndarray | None = None   # i32 per value: the value its alias resolves to, -1 = none


@dataclass
class TableSpec:
    chunk_sizes: np.ndarray
    seq_columns: list = field(default_factory=list)
    int_columns: list = field(default_factory=list)
    dict_columns: list = field(default_factory=list)
    bitmaps: list = field(default_factory=list)  # literal bitmaps: arrays of global row ids

    @property
    def n_rows(self) -> int:
        return int(np.sum(self.chunk_sizes, dtype=np.uint64))

    def global_ids(self) -> np.ndarray:
        out = []
        for c, size in enumerate(self.chunk_sizes.tolist()):
            out.append((np.uint32(c) << np.uint32(16)) + np.arange(size, dtype=np.uint32))
        return np.concatenate(out) if out else np.zeros(0, dtype=np.uint32)


def spec_from_dataset(ds: Dataset, chunk: int = CHUNK, seq_name: str = "main", aa_columns: list | None = None) -> TableSpec:
    """The dataset as ingest columns; `aa_columns` (datagen.translate) adds amino-acid columns."""
    spec = TableSpec(chunk_sizes=chunk_sizes(ds.n_rows, chunk))
    spec.seq_columns.append(SeqColumnData(
        seq_name, ALPHABET_NUC, ds.reference, ds.is_null, ds.cov_start, ds.cov_end, ds.mut_ptr, ds.mut_pos,
        ds.mut_sym, ds.nrun_ptr, ds.nrun_start, ds.nrun_end))
    for (name, ref, cs, ce, mptr, mpos, msym, nptr, ns, ne) in aa_columns or []:
        spec.seq_columns.append(SeqColumnData(name, ALPHABET_AA, ref, ds.is_null, cs, ce, mptr, mpos, msym, nptr, ns, ne))
    spec.int_columns.append(IntColumnData("date", True, ds.date, ds.date_null))
    spec.dict_columns.append(DictColumnData("country", ds.country, ds.country_null, list(ds.country_names)))
    spec.dict_columns.append(DictColumnData("lineage", ds.lineage, ds.lineage_null, list(ds.lineage_names),
                                            ds.lineage_parent))
    return spec


def encode_names(names: list) -> tuple[bytes, np.ndarray]:
    blobs = [n.encode("utf-8") for n in names]
    offsets = np.zeros(len(blobs) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(b) for b in blobs], dtype=np.uint64).astype(np.uint32)
    return b"".join(blobs), offsets


def seq_column_from_strings(name: str, reference: str, rows: list, alphabet: int = ALPHABET_NUC) -> SeqColumnData:
    """Diffs aligned sequence strings (None = null row) with the engine's restatement of
    `extractCoverageAndMutationsFromSequence` and returns the ingest arrays."""
    from .engine import diff_sequence

    n = len(rows)
    is_null = np.zeros(n, np.uint8)
    cs = np.zeros(n, np.uint32)
    ce = np.zeros(n, np.uint32)
    mptr = np.zeros(n + 1, np.uint64)
    nptr = np.zeros(n + 1, np.uint64)
    mpos, msym, ns, ne = [], [], [], []
    for i, s in enumerate(rows):
        if s is None:
            is_null[i] = 1
        else:
            if len(s) != len(reference):
                raise ValueError("aligned sequences must have the reference's length")
            start, end, missing, muts = diff_sequence(s, reference, alphabet)
            cs[i], ce[i] = start, end
            for p, sym in muts:
                mpos.append(p)
                msym.append(sym)
            k = 0
            while k < len(missing):
                j = k
                while j + 1 < len(missing) and missing[j + 1] == missing[j] + 1:
                    j += 1
                ns.append(missing[k])
                ne.append(missing[j] + 1)
                k = j + 1
        mptr[i + 1] = len(mpos)
        nptr[i + 1] = len(ns)
    ref = np.array([symbol_of(alphabet, c) for c in reference], np.uint8)
    return SeqColumnData(name, alphabet, ref, is_null, cs, ce, mptr, np.array(mpos, np.uint32), np.array(msym, np.uint8),
                         nptr, np.array(ns, np.uint32), np.array(ne, np.uint32))
