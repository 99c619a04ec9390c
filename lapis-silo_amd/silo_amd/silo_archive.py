"""Reader of the reference's saved database state (`.silo` files, Boost binary archives).

A saved state is a directory holding `database_schema.silo` (the table schemas with each column's
metadata), one `<table>.silo` per table (its columns, row count and row layout) and
`data_version.silo` (`database.cpp:426-455`, `table.cpp:220-233`, `database_schema.cpp:121-132`).
This module reads those archives and turns a table into the engine's upload form: sequence columns
as `FinalizedSeqColumnData` (the vertical index's containers as written, per-row coverage and
missing positions), int32 and date32 columns as `IntColumnData`, dictionary-encoded columns as
`DictColumnData` (with the lineage tree's parents when the column has one).  Column types off the
query path (plain and zstd-compressed strings, bool, float, int64) are read so the archive stays
in step and are returned as host values.

Archive layout (Boost.Serialization binary archive, library version 20, little endian):
- header: u64 length + "serialization::archive", u16 library version, the sizes of int, long, float
  and double as one byte each, int 1;
- primitives are written raw: integers and floats at their width, bool as one byte, enums as int32,
  std::string as u64 length + bytes;
- the first object of each class type is preceded by its class information: a tracking byte and a
  u32 class version (tracked classes then carry a u32 object id per new object);
- std::vector of arithmetic types: u64 count + the raw elements; of other types: u64 count, u32 item
  version, the elements;  std::map / std::deque: u64 count, u32 item version, the elements
  (std::pair = first, second);  std::unordered_map: u64 count, u64 bucket count, u32 item version,
  the elements;  std::optional: a bool, then the value when set;
- roaring::Roaring: u64 size + CRoaring's portable serialization (`roaring_serialize.h:16-48`).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

from .table import (ALPHABET_AA, ALPHABET_NUC, DictColumnData, FinalizedSeqColumnData, IntColumnData,
                    TableSpec)

# schema::ColumnType (`schema/database_schema.h:20-31`)
STRING, DICTIONARY_ENCODED, DATE32, BOOL, INT32, INT64, FLOAT, AA_SEQUENCE, NUC_SEQUENCE, ZSTD_STRING = range(10)
COLUMN_TYPE_NAMES = ["string", "dictionary_encoded", "date32", "bool", "int32", "int64", "float",
                     "amino_acid_sequence", "nucleotide_sequence", "zstd_compressed_string"]
# storage order of the column maps in ColumnGroup::serialize (`storage/column_group.h:28-60`)
TABLE_COLUMN_ORDER = [STRING, DICTIONARY_ENCODED, BOOL, INT32, FLOAT, DATE32, INT64, NUC_SEQUENCE, AA_SEQUENCE,
                      ZSTD_STRING]

PAGE_SIZE = 16384  # storage/buffer/page.h:10
CHUNK_ROWS = 1 << 16
SCHEMA_FILENAME = "database_schema.silo"
DATA_VERSION_FILENAME = "data_version.silo"


class SiloFormatError(ValueError):
    """The file is not a saved state this reader understands."""


class ArchiveReader:
    """A cursor over one Boost binary archive with the per-class information it has seen."""

    def __init__(self, data: bytes, name: str):
        self.data = memoryview(data)
        self.name = name
        self.pos = 0
        self.classes: dict[str, tuple[int, int]] = {}
        self._header()

    def fail(self, what: str):
        raise SiloFormatError(f"{self.name}: {what} at byte {self.pos}")

    def take(self, n: int) -> memoryview:
        if n < 0 or self.pos + n > len(self.data):
            self.fail(f"truncated archive (wanted {n} bytes)")
        out = self.data[self.pos:self.pos + n]
        self.pos += n
        return out

    def _unpack(self, fmt: str):
        size = struct.calcsize(fmt)
        return struct.unpack_from(fmt, self.take(size))[0]

    def u8(self) -> int:
        return self._unpack("<B")

    def boolean(self) -> bool:
        v = self.u8()
        if v > 1:
            self.fail(f"bool byte {v}")
        return bool(v)

    def u16(self) -> int:
        return self._unpack("<H")

    def u32(self) -> int:
        return self._unpack("<I")

    def i32(self) -> int:
        return self._unpack("<i")

    def u64(self) -> int:
        return self._unpack("<Q")

    def f32(self) -> float:
        return self._unpack("<f")

    def count(self, limit: int | None = None) -> int:
        n = self.u64()
        if n > (limit if limit is not None else len(self.data)):
            self.fail(f"implausible count {n}")
        return n

    def string_bytes(self) -> bytes:
        return bytes(self.take(self.count()))

    def string(self) -> str:
        b = self.string_bytes()
        try:
            return b.decode("utf-8")
        except UnicodeDecodeError:
            self.fail("a string that is not UTF-8")

    def _header(self):
        magic = self.string_bytes()
        if magic != b"serialization::archive":
            self.fail("not a Boost serialization archive")
        self.library_version = self.u16()
        sizes = bytes(self.take(4))
        if sizes != b"\x04\x08\x04\x08":
            self.fail(f"unsupported primitive sizes {sizes.hex()}")
        if self.i32() != 1:
            self.fail("unsupported byte order")
        if self.library_version < 10:
            self.fail(f"unsupported archive library version {self.library_version}")

    def cls(self, key: str) -> int:
        """Class information of `key`: read on its first object; per object the id of a tracked class."""
        info = self.classes.get(key)
        if info is None:
            tracking = self.boolean()
            version = self.u32()
            info = (tracking, version)
            self.classes[key] = info
        if info[0]:
            self.u32()  # object id
        return info[1]

    def items(self, key: str) -> int:
        """An object collection's preamble: class information, count, item version."""
        self.cls(key)
        n = self.count()
        self.u32()
        return n

    def unordered_items(self, key: str) -> int:
        self.cls(key)
        n = self.count()
        self.u64()  # bucket count
        self.u32()
        return n

    def array(self, dtype, key: str | None = None) -> np.ndarray:
        """std::vector of an arithmetic type (raw elements; class information only for `key`)."""
        if key is not None:
            self.cls(key)
        dt = np.dtype(dtype)
        n = self.count(len(self.data) // max(dt.itemsize, 1))
        return np.frombuffer(self.take(n * dt.itemsize), dtype=dt).copy()

    def enum_vector(self, key: str) -> np.ndarray:
        """std::vector of an enum: object collection of int32 values."""
        n = self.items(key)
        return np.frombuffer(self.take(4 * n), dtype="<i4").copy()

    def optional(self, key: str, read):
        self.cls(key)
        if not self.boolean():
            return None
        return read()

    def roaring(self) -> np.ndarray:
        """A roaring::Roaring as a sorted u32 array."""
        self.cls("roaring::Roaring")
        n = self.count()
        return roaring_values(bytes(self.take(n)), self.name)

    def at_end(self) -> bool:
        return self.pos == len(self.data)


# ------------------------------------------------------------------------------------------------
# CRoaring's portable serialization (`roaring_bitmap_portable_serialize`) and container_write forms
# ------------------------------------------------------------------------------------------------
_ENTRY = struct.Struct("<IHiIBQ")
_ROW_BITMAP = struct.Struct("<IQ")
_SERIAL_COOKIE_NO_RUN = 12346
_SERIAL_COOKIE = 12347
_NO_OFFSET_THRESHOLD = 4


def _expand_runs(runs: np.ndarray) -> np.ndarray:
    """(start, length - 1) u16 pairs to the u32 values they cover."""
    start = runs[:, 0].astype(np.int64)
    length = runs[:, 1].astype(np.int64) + 1
    if not len(start):
        return np.zeros(0, np.uint32)
    base = np.repeat(start - np.concatenate([[0], np.cumsum(length)[:-1]]), length)
    return (base + np.arange(int(length.sum()), dtype=np.int64)).astype(np.uint32)


def container_values(typecode: int, cardinality: int, payload: bytes) -> np.ndarray:
    """The u16 values of one container in its `container_write` form (1 bitset, 2 array, 3 run)."""
    if typecode == 1:
        if len(payload) < 8192:
            raise SiloFormatError("truncated bitset container")
        words = np.frombuffer(payload[:8192], dtype="<u8")
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return np.nonzero(bits)[0].astype(np.uint16)
    if typecode == 2:
        if len(payload) < 2 * cardinality:
            raise SiloFormatError("truncated array container")
        return np.frombuffer(payload[:2 * cardinality], dtype="<u2").copy()
    if typecode == 3:
        if len(payload) < 2:
            raise SiloFormatError("truncated run container")
        n_runs = struct.unpack_from("<H", payload)[0]
        if len(payload) < 2 + 4 * n_runs:
            raise SiloFormatError("truncated run container")
        runs = np.frombuffer(payload[2:2 + 4 * n_runs], dtype="<u2").reshape(-1, 2)
        if n_runs == 0:
            return np.zeros(0, np.uint16)
        if np.any(runs[:, 0].astype(np.int64) + runs[:, 1] > 0xFFFF):
            raise SiloFormatError("run past the end of its block")
        return _expand_runs(runs).astype(np.uint16)
    raise SiloFormatError(f"unknown container typecode {typecode}")


def _gather_u16(data: np.ndarray, at: np.ndarray) -> np.ndarray:
    return data[at].astype(np.uint32) | (data[at + 1].astype(np.uint32) << np.uint32(8))


def _ranges(starts: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Concatenated arange(starts[i], starts[i] + counts[i])."""
    counts = counts.astype(np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    base = np.repeat(starts.astype(np.int64) - np.concatenate([[0], np.cumsum(counts)[:-1]]), counts)
    return base + np.arange(total, dtype=np.int64)


def decode_bitmaps(data: np.ndarray, starts: np.ndarray, lens: np.ndarray, name: str = "roaring",
                   block: int = 2048):
    """Many portable roaring bitmaps in one buffer (bitmap i at data[starts[i]:starts[i] + lens[i]])
    to a CSR (ptr u64 [n + 1], values u32), `block` bitmaps at a time (the temporaries stay in
    cache)."""
    n = len(starts)
    ptrs, vals = [np.zeros(1, np.uint64)], []
    total = 0
    for b in range(0, n, block):
        p, v = _decode_block(data, starts[b:b + block], lens[b:b + block], name)
        ptrs.append(p[1:] + np.uint64(total))
        vals.append(v)
        total += len(v)
    return np.concatenate(ptrs), (np.concatenate(vals) if vals else np.zeros(0, np.uint32))


def decode_bitmap_runs(data: np.ndarray, starts: np.ndarray, lens: np.ndarray, name: str = "roaring"):
    """Many portable roaring bitmaps (bitmap i at data[starts[i]:starts[i] + lens[i]]) to their
    maximal runs: CSR ptr u64 [n + 1] and the runs [start, end) u32, decoded natively
    (rdb_archive_bitmap_runs, two passes: counts, then runs).  Malformed bitmaps raise
    SiloFormatError."""
    import ctypes

    from .engine import _p, lib

    data = np.ascontiguousarray(data, np.uint8)
    starts = np.ascontiguousarray(starts, np.uint64)
    lens = np.ascontiguousarray(lens, np.uint64)
    n = len(starts)
    ptr = np.zeros(n + 1, np.uint64)
    base = ctypes.c_void_p(data.ctypes.data)
    rc = lib().rdb_archive_bitmap_runs(base, len(data), _p(starts), _p(lens), n, _p(ptr), None, None, 0)
    if rc == 0:
        total = int(ptr[-1])
        rs, re_ = np.empty(max(total, 1), np.uint32), np.empty(max(total, 1), np.uint32)
        rc = lib().rdb_archive_bitmap_runs(base, len(data), _p(starts), _p(lens), n, _p(ptr), _p(rs), _p(re_), total)
    if rc != 0:
        raise SiloFormatError(f"{name}: {lib().rdb_last_error(None).decode()}")
    return ptr, rs[:total], re_[:total]


def _decode_block(data: np.ndarray, starts: np.ndarray, lens: np.ndarray, name: str):
    """decode_bitmaps for one block: bitmaps of one array or one run container (key and
    cardinality, then the values or the runs) are decoded together; the rest one by one."""
    n = len(starts)
    ptr = np.zeros(n + 1, np.uint64)
    if n == 0:
        return ptr, np.zeros(0, np.uint32)
    starts = starts.astype(np.int64)
    lens = lens.astype(np.int64)
    if np.any(lens < 8):
        raise SiloFormatError(f"{name}: roaring bitmap too short")
    u32 = lambda at: _gather_u16(data, at) | (_gather_u16(data, at + 2) << np.uint32(16))
    cookie = u32(starts)
    # one array container without runs: cookie, count 1, key, cardinality - 1, offset 16, values
    plain = (cookie == _SERIAL_COOKIE_NO_RUN) & (lens >= 18)
    plain &= np.where(plain, u32(np.where(plain, starts + 4, 0)), 0) == 1
    p_card = np.where(plain, _gather_u16(data, np.where(plain, starts + 10, 0)).astype(np.int64) + 1, 0)
    plain &= (p_card <= 4096) & (lens == 16 + 2 * p_card)
    # one container with the run cookie: cookie (size 1), flag byte, key, cardinality - 1, then
    # runs (count, (start, length - 1) pairs) or array values
    single = (cookie == _SERIAL_COOKIE) & (lens >= 9)
    flag = np.where(single, data[np.where(single, starts + 4, 0)], 0)
    s_card = np.where(single, _gather_u16(data, np.where(single, starts + 7, 0)).astype(np.int64) + 1, 0)
    run = single & (flag & 1 == 1) & (lens >= 11)
    n_runs = np.where(run, _gather_u16(data, np.where(run, starts + 9, 0)).astype(np.int64), 0)
    run &= lens == 11 + 4 * n_runs
    arr1 = single & (flag & 1 == 0) & (s_card <= 4096) & (lens == 9 + 2 * s_card)
    key = np.zeros(n, np.uint32)
    key[plain] = _gather_u16(data, starts[plain] + 8)
    key[run | arr1] = _gather_u16(data, starts[run | arr1] + 5)
    counts = np.zeros(n, np.int64)
    counts[plain] = p_card[plain]
    counts[arr1] = s_card[arr1]
    # runs: lengths from the pairs
    r_idx = np.nonzero(run)[0]
    pair_at = _ranges(starts[r_idx] + 11, n_runs[r_idx] * 4)[::4]
    r_start = _gather_u16(data, pair_at).astype(np.int64)
    r_len = _gather_u16(data, pair_at + 2).astype(np.int64) + 1
    run_owner = np.repeat(r_idx, n_runs[r_idx])
    np.add.at(counts, run_owner, r_len)
    other = np.nonzero(~(plain | run | arr1))[0]
    decoded = {}
    for i in other.tolist():
        v = roaring_values(data[starts[i]:starts[i] + lens[i]].tobytes(), name)
        decoded[i] = v
        counts[i] = len(v)
    np.cumsum(counts, out=ptr[1:])
    values = np.zeros(int(ptr[-1]), np.uint32)
    for sel, off in ((plain, 16), (arr1, 9)):
        idx = np.nonzero(sel)[0]
        at = _ranges(starts[idx] + off, counts[idx] * 2)[::2]
        dest = _ranges(ptr[:-1][idx].astype(np.int64), counts[idx])
        values[dest] = _gather_u16(data, at) | (np.repeat(key[idx], counts[idx]) << np.uint32(16))
    if len(r_idx):
        run_first = np.zeros(len(r_idx), np.int64)
        run_first[1:] = np.cumsum(n_runs[r_idx])[:-1]
        # each run's place in the output: its bitmap's start plus the lengths of the runs before it
        within = np.cumsum(r_len) - r_len - np.repeat(np.cumsum(r_len)[run_first] - r_len[run_first], n_runs[r_idx])
        dest = _ranges(ptr[:-1][run_owner].astype(np.int64) + within, r_len)
        values[dest] = (_ranges(r_start, r_len) | (np.repeat(key[run_owner], r_len).astype(np.int64) << 16)).astype(
            np.uint32)
    for i, v in decoded.items():
        values[int(ptr[i]):int(ptr[i + 1])] = v
    return ptr, values


def roaring_values(buf: bytes, name: str = "roaring") -> np.ndarray:
    """Decodes CRoaring's portable format (the format `Roaring::write` produces) to sorted u32.
    Malformed input (a short buffer, keys out of order, a container whose values do not match its
    cardinality) raises SiloFormatError."""
    def bad(what: str):
        raise SiloFormatError(f"{name}: {what}")

    n_buf = len(buf)
    if n_buf < 4:
        bad("roaring bitmap too short")
    cookie = struct.unpack_from("<I", buf)[0]
    # one array container, the common case of small bitmaps: cookie, count 1, key, cardinality - 1,
    # offset, values
    if cookie == _SERIAL_COOKIE_NO_RUN and n_buf >= 16 and struct.unpack_from("<I", buf, 4)[0] == 1:
        key, c1 = struct.unpack_from("<HH", buf, 8)
        if c1 < 4096 and n_buf == 16 + 2 * (c1 + 1):
            vals = np.frombuffer(buf, "<u2", c1 + 1, 16)
            if c1 and np.any(np.diff(vals.astype(np.int32)) <= 0):
                bad("array container values out of order")
            return vals.astype(np.uint32) | np.uint32(key << 16)
    pos = 4
    run_flags = None
    if cookie & 0xFFFF == _SERIAL_COOKIE:
        size = (cookie >> 16) + 1
        nbytes = (size + 7) // 8
        if pos + nbytes > n_buf:
            bad("truncated roaring bitmap")
        run_flags = np.unpackbits(np.frombuffer(buf[pos:pos + nbytes], np.uint8), bitorder="little")[:size]
        pos += nbytes
    elif cookie == _SERIAL_COOKIE_NO_RUN:
        if n_buf < 8:
            bad("truncated roaring bitmap")
        size = struct.unpack_from("<I", buf, pos)[0]
        pos += 4
    else:
        bad(f"bad roaring cookie {cookie:#x}")
    if size > 65536 or pos + 4 * size > n_buf:
        bad("truncated roaring bitmap")
    hdr = np.frombuffer(buf[pos:pos + 4 * size], dtype="<u2").reshape(-1, 2) if size else np.zeros((0, 2), np.uint16)
    if size > 1 and np.any(np.diff(hdr[:, 0].astype(np.int32)) <= 0):
        bad("container keys out of order")
    pos += 4 * size
    if run_flags is None or size >= _NO_OFFSET_THRESHOLD:
        pos += 4 * size  # container offsets
    out = []
    for i in range(size):
        key = int(hdr[i, 0])
        card = int(hdr[i, 1]) + 1
        if run_flags is not None and run_flags[i]:
            if pos + 2 > n_buf:
                bad("truncated run container")
            n_runs = struct.unpack_from("<H", buf, pos)[0]
            if pos + 2 + 4 * n_runs > n_buf:
                bad("truncated run container")
            runs = np.frombuffer(buf[pos + 2:pos + 2 + 4 * n_runs], dtype="<u2").reshape(-1, 2)
            pos += 2 + 4 * n_runs
            if np.any(runs[:, 0].astype(np.int64) + runs[:, 1] > 0xFFFF):
                bad("run past the end of its block")
            vals = _expand_runs(runs)
        elif card > 4096:
            if pos + 8192 > n_buf:
                bad("truncated bitset container")
            words = np.frombuffer(buf[pos:pos + 8192], dtype="<u8")
            pos += 8192
            vals = np.nonzero(np.unpackbits(words.view(np.uint8), bitorder="little"))[0].astype(np.uint32)
        else:
            if pos + 2 * card > n_buf:
                bad("truncated array container")
            vals = np.frombuffer(buf[pos:pos + 2 * card], dtype="<u2").astype(np.uint32)
            pos += 2 * card
        if len(vals) != card or (len(vals) > 1 and np.any(np.diff(vals.astype(np.int64)) <= 0)):
            bad("container values do not match its cardinality")
        out.append((np.uint32(key) << np.uint32(16)) | vals.astype(np.uint32))
    return np.concatenate(out).astype(np.uint32) if out else np.zeros(0, np.uint32)


def _optional(r: ArchiveReader, key: str, read):
    """std::optional<T>: class information, a bool, then (class version 0 only) an item version
    and the value."""
    version = r.cls(key)
    if not r.boolean():
        return None
    if version == 0:
        r.u32()
    return read()


# ------------------------------------------------------------------------------------------------
# database_schema.silo (`schema/database_schema.cpp:72-132`, metadata loaders of each column type)
# ------------------------------------------------------------------------------------------------
@dataclass
class LineageTree:
    """`common/lineage_tree.h:40-140`: parents per lineage id, recombinants' clade ancestors, aliases
    (ids in `lineage_id_lookup_map`), and the definition file's text."""
    child_to_parent: list
    recombinant_clade_ancestors: dict
    alias_mapping: dict
    names: list
    file: str


@dataclass
class ColumnSchema:
    name: str
    type: int
    dictionary: list | None = None          # string and dictionary-encoded columns
    lineage_tree: LineageTree | None = None  # dictionary-encoded columns with a lineage definition
    treat_unknown_lineages_as_null: bool = False
    reference: np.ndarray | None = None     # sequence columns: reference symbols (u8)
    phylo_tree: dict | None = None           # string columns with a phylogenetic tree: node -> parent
    zstd_dictionary: bytes | None = None

    @property
    def type_name(self) -> str:
        return COLUMN_TYPE_NAMES[self.type]


@dataclass
class TableSchema:
    name: str
    primary_key: str
    columns: list  # ColumnSchema in the schema's order (by name, then type)

    def column(self, name: str) -> ColumnSchema:
        for c in self.columns:
            if c.name == name:
                return c
        raise KeyError(name)


def _string_map(r: ArchiveReader) -> list:
    """BidirectionalStringMap (`common/bidirectional_string_map.h:83-94`): id -> value."""
    r.cls("BidirectionalStringMap")
    n = r.items("vector<string>")
    return [r.string() for _ in range(n)]


def _lineage_tree(r: ArchiveReader) -> LineageTree:
    r.cls("LineageTreeAndIdMap")
    r.cls("LineageTree")
    n = r.items("vector<vector<u32>>")
    parents = [r.array(np.uint32).tolist() for _ in range(n)]
    recomb = {}
    for _ in range(r.unordered_items("unordered_map<u32,optional<u32>>")):
        r.cls("pair<const u32,optional<u32>>")
        k = r.u32()
        recomb[k] = _optional(r, "optional<u32>", r.u32)
    alias = {}
    for _ in range(r.unordered_items("unordered_map<u32,u32>")):
        r.cls("pair<const u32,u32>")
        k = r.u32()
        alias[k] = r.u32()
    names = _string_map(r)
    file = r.string()
    return LineageTree(parents, recomb, alias, names, file)


def _tree_node_id(r: ArchiveReader) -> str:
    r.cls("TreeNodeId")
    return r.string()


def _phylo_tree(r: ArchiveReader) -> dict:
    """PhyloTree (`common/phylo_tree.cpp:33-58`) as {node: (parent or None, row index or None)}."""
    r.cls("PhyloTree")
    ids = [_tree_node_id(r) for _ in range(r.items("vector<TreeNodeId>"))]
    nodes = {}
    for node in ids:
        r.cls("TreeNode")
        node_id = _tree_node_id(r)
        for _ in range(r.items("vector<TreeNodeId>")):
            _tree_node_id(r)
        parent = _optional(r, "optional<TreeNodeId>", lambda: _tree_node_id(r))
        r.i32()  # depth
        row = _optional(r, "optional<u64>", r.u64)
        _optional(r, "optional<float>", r.f32)
        if node_id != node:
            r.fail(f"phylogenetic tree node {node_id!r} out of order")
        nodes[node_id] = (parent, row)
    return nodes


def _column_metadata(r: ArchiveReader, col: ColumnSchema):
    t = col.type
    if t in (DATE32, BOOL, INT32, INT64, FLOAT):
        r.cls("ColumnMetadata")
        name = r.string()
    elif t == STRING:
        r.cls("StringColumnMetadata")
        name = r.string()
        col.dictionary = _string_map(r)
        col.phylo_tree = _optional(r, "optional<PhyloTree>", lambda: _phylo_tree(r))
    elif t == DICTIONARY_ENCODED:
        r.cls("DictionaryEncodedColumnMetadata")
        name = r.string()
        col.dictionary = _string_map(r)
        col.lineage_tree = _optional(r, "optional<LineageTreeAndIdMap>", lambda: _lineage_tree(r))
        col.treat_unknown_lineages_as_null = r.boolean()
    elif t in (NUC_SEQUENCE, AA_SEQUENCE):
        tag = "Nucleotide" if t == NUC_SEQUENCE else "AminoAcid"
        r.cls(f"SequenceColumnMetadata<{tag}>")
        name = r.string()
        ref = r.enum_vector(f"vector<{tag}::Symbol>")
        limit = 16 if t == NUC_SEQUENCE else 28
        if ref.size and (ref.min() < 0 or ref.max() >= limit):
            r.fail(f"reference symbol out of range in column {name!r}")
        col.reference = ref.astype(np.uint8)
    elif t == ZSTD_STRING:
        r.cls("ZstdCompressedStringColumnMetadata")
        name = r.string()
        col.zstd_dictionary = r.string_bytes()
    else:
        r.fail(f"unknown column type {t}")
    # a zstd-compressed column's metadata carries the name of the sequence whose reference is its
    # compression dictionary (`unaligned_main` -> `main`)
    if name != col.name and t != ZSTD_STRING:
        r.fail(f"metadata of column {name!r} where {col.name!r} was expected")


def _column_identifier(r: ArchiveReader) -> tuple[str, int]:
    r.cls("ColumnIdentifier")
    name = r.string()
    t = r.i32()
    if not 0 <= t < len(COLUMN_TYPE_NAMES):
        r.fail(f"unknown column type {t}")
    return name, t


def read_schema(data: bytes, name: str = SCHEMA_FILENAME) -> dict:
    """database_schema.silo: {table name: TableSchema}."""
    r = ArchiveReader(data, name)
    tables = {}
    for _ in range(r.items("map<TableName,shared_ptr<TableSchema>>")):
        r.cls("pair<const TableName,shared_ptr<TableSchema>>")
        r.cls("TableName")
        table = r.string()
        r.cls("shared_ptr<TableSchema>")
        r.cls("TableSchema")
        pk, _ = _column_identifier(r)
        idents = [_column_identifier(r) for _ in range(r.items("vector<ColumnIdentifier>"))]
        cols = []
        for cname, t in idents:
            col = ColumnSchema(cname, t)
            _column_metadata(r, col)
            cols.append(col)
        tables[table] = TableSchema(table, pk, cols)
    if not r.at_end():
        r.fail("trailing bytes after the schema")
    return tables


# ------------------------------------------------------------------------------------------------
# <table>.silo (`storage/table.h:35-42`, `storage/column_group.h:28-60`, each column's serialize)
# ------------------------------------------------------------------------------------------------
@dataclass
class SequenceColumnState:
    """A saved SequenceColumn (`storage/column/sequence_column.h:86-97`)."""
    name: str
    alphabet: int
    reference: np.ndarray
    local_reference: np.ndarray     # u8 symbols
    key_position: np.ndarray        # vertical index keys, in the map's order
    key_chunk: np.ndarray
    key_symbol: np.ndarray
    typecode: np.ndarray
    cardinality: np.ndarray
    payload_offset: np.ndarray
    payload: np.ndarray
    missing_rows: np.ndarray        # u32 global ids of the rows with internal missing symbols
    missing_ptr: np.ndarray         # u64 [len(missing_rows) + 1]: their runs in missing_start/end
    missing_start: np.ndarray       # u32 runs [start, end) of missing positions, ascending per row
    missing_end: np.ndarray
    start_end: np.ndarray           # (rows, 2) u32 covered ranges, in row order
    null_rows: np.ndarray           # u32 row ids
    sequence_count: int
    num_chunks: int
    info: tuple                     # SequenceColumnInfo: (sequence_count, vertical, horizontal sizes)
    insertion_count: int
    span: tuple = (0, 0)            # its bytes in the table archive
    classes_before: frozenset = frozenset()  # class information written before it


@dataclass
class TableState:
    """A loaded table: its schema, row layout and columns by name."""
    schema: TableSchema
    sequence_count: int
    chunk_sizes: np.ndarray
    columns: dict = field(default_factory=dict)


def _symbol_chars(alphabet: int) -> str:
    from .table import alphabet_chars
    return alphabet_chars(alphabet)


def _chunked(r: ArchiveReader, tag: str, dtype) -> list:
    """ChunkedValueBuffer<T> (`storage/column/chunked_value_buffer.h:48-54`): its chunks."""
    r.cls(f"ChunkedValueBuffer<{tag}>")
    n = r.items(f"vector<vector<{tag}>>")
    return [r.array(dtype) for _ in range(n)]


def _page(r: ArchiveReader) -> bytes:
    r.cls("Page")
    return bytes(r.take(PAGE_SIZE))


def _string_column(r: ArchiveReader) -> dict:
    """StringColumn (`storage/column/string_column.h:93-98,147-152`): pages as written."""
    r.cls("StringColumn")
    nulls = r.roaring()
    chunks = []
    for _ in range(r.items("deque<StringColumnChunk>")):
        r.cls("StringColumnChunk")
        r.cls("GermanStringRegistry")
        fixed = []
        for _ in range(r.items("deque<GermanStringPage>")):
            r.cls("GermanStringPage")
            fixed.append(_page(r))
        r.cls("VariableDataRegistry")
        variable = [_page(r) for _ in range(r.items("deque<Page>"))]
        offset = r.u16()
        chunks.append((fixed, variable, offset))
    return {"null_rows": nulls, "chunks": chunks}


class SavedBitmap:
    """A roaring bitmap of the archive kept in its portable form; `values()` decodes it."""
    __slots__ = ("raw",)

    def __init__(self, raw: bytes):
        self.raw = raw

    def values(self) -> np.ndarray:
        return roaring_values(self.raw)


def _unordered_bitmaps(r: ArchiveReader) -> dict:
    """unordered_map<u32, Roaring>: value id -> SavedBitmap (the engine rebuilds these row sets from
    the value ids, so they are not decoded here)."""
    out = {}
    for _ in range(r.unordered_items("unordered_map<u32,Roaring>")):
        r.cls("pair<const u32,Roaring>")
        k = r.u32()
        r.cls("roaring::Roaring")
        out[k] = SavedBitmap(bytes(r.take(r.count())))
    return out


def _dictionary_column(r: ArchiveReader, col: ColumnSchema) -> dict:
    """DictionaryEncodedColumn (`storage/column/dictionary_encoded_column.h:129-139`)."""
    r.cls("DictionaryEncodedColumn")
    ids = _chunked(r, "u32", np.uint32)
    indexed = _unordered_bitmaps(r)
    nulls = r.roaring()
    lineage = None
    if col.lineage_tree is not None:
        r.cls("LineageIndex")
        including = {}
        for _ in range(r.unordered_items("unordered_map<RecombinantEdgeFollowingMode,unordered_map<u32,Roaring>>")):
            r.cls("pair<const RecombinantEdgeFollowingMode,unordered_map<u32,Roaring>>")
            mode = r.i32()
            including[mode] = _unordered_bitmaps(r)
        excluding = _unordered_bitmaps(r)
        lineage = {"including_sublineages": including, "excluding_sublineages": excluding}
    return {"ids": ids, "indexed_values": indexed, "null_rows": nulls, "lineage_index": lineage}


def _pair_array(r: ArchiveReader) -> np.ndarray:
    r.cls("vector<pair<u32,u32>>")
    n = r.count()
    return np.frombuffer(r.take(8 * n), dtype="<u4").reshape(n, 2).copy()


def _insertion_index(r: ArchiveReader, tag: str) -> int:
    """InsertionIndex (`storage/column/insertion_index.h:27-95`); read and counted, not kept."""
    r.cls(f"InsertionIndex<{tag}>")
    count = 0
    for _ in range(r.unordered_items(f"unordered_map<u32,InsertionPosition<{tag}>>")):
        r.cls(f"pair<const u32,InsertionPosition<{tag}>>")
        r.u32()
        r.cls(f"InsertionPosition<{tag}>")
        for _ in range(r.items("vector<Insertion>")):
            r.cls("Insertion")
            r.string_bytes()
            r.roaring()
            count += 1
        for _ in range(r.unordered_items(f"unordered_map<array<{tag}::Symbol,3>,vector<u32>>")):
            r.cls(f"pair<const array<{tag}::Symbol,3>,vector<u32>>")
            r.cls(f"array<{tag}::Symbol,3>")
            if r.count() != 3:
                r.fail("three-mer of the wrong length")
            r.take(12)
            r.array(np.uint32)
    for _ in range(r.unordered_items("unordered_map<u32,unordered_map<string,Roaring>>")):
        r.cls("pair<const u32,unordered_map<string,Roaring>>")
        r.u32()
        for _ in range(r.unordered_items("unordered_map<string,Roaring>")):
            r.cls("pair<const string,Roaring>")
            r.string_bytes()
            r.roaring()
    return count


def _buffer(r: ArchiveReader):
    import ctypes

    arr = np.frombuffer(r.data, np.uint8)
    return arr, ctypes.c_void_p(arr.ctypes.data)


def _scan_containers(r: ArchiveReader, n: int, name: str):
    """n vertical-index records from r.pos (rdb_archive_scan_containers): position, chunk, symbol,
    cardinality, typecode, payload offsets and the packed payload."""
    import ctypes

    from .engine import _p, lib

    arr, base = _buffer(r)
    out = [np.empty(n, dt) for dt in (np.uint32, np.uint16, np.int32, np.uint32, np.uint8, np.uint64)]
    total, end = ctypes.c_uint64(), ctypes.c_uint64()
    for payload in (None, True):
        buf = None if payload is None else np.empty(max(total.value, 1), np.uint8)
        rc = lib().rdb_archive_scan_containers(base, len(arr), r.pos, n, *[_p(a) for a in out],
                                               None if buf is None else _p(buf), 0 if buf is None else len(buf),
                                               ctypes.byref(total), ctypes.byref(end))
        if rc != 0:
            r.fail(f"column {name!r}: {lib().rdb_last_error(None).decode()}")
    r.pos = end.value
    return out + [buf[:total.value]]


def _scan_row_bitmaps(r: ArchiveReader, n: int, name: str):
    """n horizontal-index records from r.pos (rdb_archive_scan_row_bitmaps): row ids and where each
    bitmap lies."""
    import ctypes

    from .engine import _p, lib

    arr, base = _buffer(r)
    rows, starts, lens = np.empty(n, np.uint32), np.empty(n, np.uint64), np.empty(n, np.uint64)
    end = ctypes.c_uint64()
    rc = lib().rdb_archive_scan_row_bitmaps(base, len(arr), r.pos, n, _p(rows), _p(starts), _p(lens),
                                            ctypes.byref(end))
    if rc != 0:
        r.fail(f"column {name!r}: {lib().rdb_last_error(None).decode()}")
    r.pos = end.value
    return rows, starts.astype(np.int64), lens.astype(np.int64)


def _sequence_column(r: ArchiveReader, col: ColumnSchema) -> SequenceColumnState:
    alphabet = ALPHABET_NUC if col.type == NUC_SEQUENCE else ALPHABET_AA
    tag = "Nucleotide" if alphabet == ALPHABET_NUC else "AminoAcid"
    n_symbols = 16 if alphabet == ALPHABET_NUC else 28
    r.cls(f"SequenceColumn<{tag}>")
    chars = _symbol_chars(alphabet)
    local = r.string()
    try:
        local_ref = np.array([chars.index(c) for c in local], np.uint8)
    except ValueError:
        r.fail(f"column {col.name!r}: local reference holds a character outside the alphabet")
    r.cls(f"VerticalSequenceIndex<{tag}>")
    n = r.items(f"map<SequenceDiffKey<{tag}>,RoaringContainer>")
    keys = np.empty((n, 5), np.int64)  # position, chunk, symbol, cardinality, typecode
    offs = np.zeros(n, np.uint64)
    payload = []
    total = 0
    classes = (f"pair<const SequenceDiffKey<{tag}>,RoaringContainer>", f"SequenceDiffKey<{tag}>", "RoaringContainer")
    i = 0
    while i < n:
        if all(k in r.classes and not r.classes[k][0] for k in classes):
            # every class of an entry is known and untracked: the rest are fixed 23-byte headers
            # (u32 position, u16 chunk, i32 symbol, u32 cardinality, u8 typecode, u64 length) and
            # the container bytes, scanned natively
            rest = _scan_containers(r, n - i, col.name)
            keys[i:, 0], keys[i:, 1], keys[i:, 2], keys[i:, 3], keys[i:, 4] = rest[:5]
            offs[i:] = rest[5] + np.uint64(total)
            payload.append(rest[6])
            total += len(rest[6])
            break
        for k in classes[:2]:
            r.cls(k)
        keys[i, 0] = r.u32()
        keys[i, 1] = r.u16()
        keys[i, 2] = r.i32()
        r.cls(classes[2])
        keys[i, 3] = r.u32()
        keys[i, 4] = r.u8()
        buf = r.string_bytes()
        offs[i] = total
        payload.append(buf)
        total += len(buf)
        i += 1
    if n and (keys[:, 2].min() < 0 or keys[:, 2].max() >= n_symbols):
        r.fail(f"column {col.name!r}: symbol out of range")
    pos = keys[:, 0].astype(np.uint32)
    chunk = keys[:, 1].astype(np.uint16)
    sym = keys[:, 2].astype(np.uint8)
    card = keys[:, 3].astype(np.uint32)
    tc = keys[:, 4].astype(np.uint8)
    r.cls("HorizontalCoverageIndex")
    n_h = r.items("map<u32,Roaring>")
    h_rows = np.empty(n_h, np.uint32)
    h_start = np.empty(n_h, np.int64)
    h_len = np.empty(n_h, np.int64)
    i = 0
    entry_classes = ("pair<const u32,Roaring>", "roaring::Roaring")
    while i < n_h:
        if all(k in r.classes and not r.classes[k][0] for k in entry_classes):
            # the rest are u32 row, u64 length, the bitmap: scanned natively
            h_rows[i:], h_start[i:], h_len[i:] = _scan_row_bitmaps(r, n_h - i, col.name)
            break
        r.cls(entry_classes[0])
        h_rows[i] = r.u32()
        r.cls(entry_classes[1])
        h_len[i] = r.count()
        h_start[i] = r.pos
        r.take(int(h_len[i]))
        i += 1
    if n_h > 1 and np.any(np.diff(h_rows.astype(np.int64)) <= 0):
        r.fail(f"column {col.name!r}: coverage index rows out of order")
    h_ptr, h_run_start, h_run_end = decode_bitmap_runs(np.frombuffer(r.data, np.uint8), h_start, h_len, r.name)
    chunks = [_pair_array(r) for _ in range(r.items("vector<vector<pair<u32,u32>>>"))]
    _pair_array(r)  # batch_start_ends
    insertions = _insertion_index(r, tag)
    r.cls("SequenceColumnInfo")
    info = (r.u32(), r.u64(), r.u64())
    seq_count = r.u32()
    nulls = r.roaring()
    num_chunks = r.u16()
    start_end = np.concatenate(chunks) if chunks else np.zeros((0, 2), np.uint32)
    # the packed payload: the native scan's own array when it read every record (no copy), else the
    # records' bytes concatenated
    parts = [p if isinstance(p, np.ndarray) else np.frombuffer(p, np.uint8) for p in payload]
    if len(parts) == 1 and isinstance(payload[0], np.ndarray):
        packed = parts[0]
    else:
        packed = np.concatenate(parts) if parts else np.zeros(0, np.uint8)
    return SequenceColumnState(
        col.name, alphabet, col.reference, local_ref, pos, chunk, sym, tc, card, offs,
        packed, h_rows, h_ptr, h_run_start, h_run_end, start_end, nulls,
        seq_count,
        num_chunks, info,
        insertions)


def read_table(data: bytes, schema: TableSchema, name: str | None = None) -> TableState:
    """<table>.silo for `schema`: every column, the row count and the row layout."""
    r = ArchiveReader(data, name or f"{schema.name}.silo")
    r.cls("ColumnGroup")
    by_type = {t: sorted((c for c in schema.columns if c.type == t), key=lambda c: c.name) for t in TABLE_COLUMN_ORDER}
    cols = {}
    for t in TABLE_COLUMN_ORDER:
        for col in by_type[t]:
            if t == STRING:
                cols[col.name] = _string_column(r)
            elif t == DICTIONARY_ENCODED:
                cols[col.name] = _dictionary_column(r, col)
            elif t == BOOL:
                r.cls("BoolColumn")
                cols[col.name] = {"true_rows": r.roaring(), "false_rows": r.roaring(), "null_rows": r.roaring(),
                                  "num_chunks": r.u16()}
            elif t in (INT32, INT64):
                tag, dt = ("i32", np.int32) if t == INT32 else ("i64", np.int64)
                r.cls(f"IntColumn<{tag}>")
                cols[col.name] = {"chunks": _chunked(r, tag, dt), "null_rows": r.roaring()}
            elif t == FLOAT:
                r.cls("FloatColumn")
                cols[col.name] = {"chunks": _chunked(r, "f64", np.float64), "null_rows": r.roaring()}
            elif t == DATE32:
                r.cls("Date32Column")
                nulls = r.roaring()
                cols[col.name] = {"null_rows": nulls, "chunks": _chunked(r, "i32", np.int32),
                                  "is_sorted": r.boolean()}
            elif t in (NUC_SEQUENCE, AA_SEQUENCE):
                start, before = r.pos, frozenset(r.classes)
                cols[col.name] = _sequence_column(r, col)
                cols[col.name].span = (start, r.pos)
                cols[col.name].classes_before = before
            elif t == ZSTD_STRING:
                r.cls("ZstdCompressedStringColumn")
                r.cls("ChunkedValueBuffer<string>")
                chunks = []
                for _ in range(r.items("vector<vector<string>>")):
                    chunks.append([r.string_bytes() for _ in range(r.items("vector<string>"))])
                cols[col.name] = {"compressed_chunks": chunks, "null_rows": r.roaring()}
    seq_count = r.u32()
    r.cls("RowLayout")
    chunk_sizes = r.array(np.uint32)
    if not r.at_end():
        r.fail("trailing bytes after the table")
    return TableState(schema, seq_count, chunk_sizes, cols)


# ------------------------------------------------------------------------------------------------
# The engine's upload form
# ------------------------------------------------------------------------------------------------
def _dense_rows(global_ids: np.ndarray, chunk_sizes: np.ndarray, what: str) -> np.ndarray:
    """Global row ids (chunk << 16 | row in chunk, `storage/column/row_id.h:16-37`) to row-layout
    indices."""
    g = np.asarray(global_ids, np.uint64)
    chunk = (g >> np.uint64(16)).astype(np.int64)
    row = (g & np.uint64(0xFFFF)).astype(np.int64)
    if g.size and (chunk.max() >= len(chunk_sizes) or np.any(row >= chunk_sizes[chunk])):
        raise SiloFormatError(f"{what}: row id outside the row layout")
    offsets = np.zeros(len(chunk_sizes) + 1, np.int64)
    np.cumsum(chunk_sizes, out=offsets[1:])
    return offsets[chunk] + row


def _flags(rows: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint8)
    out[rows] = 1
    return out


def _flat(chunks: list, dtype) -> np.ndarray:
    return np.concatenate(chunks).astype(dtype) if chunks else np.zeros(0, dtype)


def finalized_column(col: SequenceColumnState, chunk_sizes: np.ndarray) -> FinalizedSeqColumnData:
    n = int(np.sum(chunk_sizes, dtype=np.int64))
    if col.start_end.shape[0] != n:
        raise SiloFormatError(f"column {col.name!r}: {col.start_end.shape[0]} covered ranges for {n} rows")
    if col.reference is None or len(col.local_reference) != len(col.reference):
        raise SiloFormatError(f"column {col.name!r}: local reference and reference differ in length")
    # rows come in ascending global id, which is row-layout order
    dense = _dense_rows(col.missing_rows, chunk_sizes, col.name)
    counts = np.zeros(n, np.uint64)
    counts[dense] = np.diff(col.missing_ptr.astype(np.int64))
    ptr = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=ptr[1:])
    is_null = _flags(_dense_rows(col.null_rows, chunk_sizes, col.name), n)
    return FinalizedSeqColumnData(
        col.name, col.alphabet, col.reference, col.local_reference, col.key_position, col.key_chunk, col.key_symbol,
        col.typecode, col.cardinality, col.payload_offset, col.payload, is_null,
        np.ascontiguousarray(col.start_end[:, 0]), np.ascontiguousarray(col.start_end[:, 1]), None, None,
        ptr, col.missing_start, col.missing_end)


def lineage_arrays(tree: LineageTree, n_values: int) -> tuple[np.ndarray, np.ndarray]:
    """Per dictionary value its lineage parent (-1 for roots, recombinants of several parents and
    aliases) and the value its alias resolves to (-1 = none).  A recombinant's sublineage rows stay
    apart from its parents', as the reference's default `DO_NOT_FOLLOW` mode keeps them
    (`saneql/ast_to_query.cpp:410-430`, `common/lineage_tree.h:30-45`)."""
    parent = np.full(n_values, -1, np.int32)
    for v, ps in enumerate(tree.child_to_parent[:n_values]):
        if len(ps) == 1:
            parent[v] = ps[0]
    alias = np.full(n_values, -1, np.int32)
    for a, c in tree.alias_mapping.items():
        if a < n_values and c < n_values:
            alias[a] = c
    return parent, alias


def table_spec(state: TableState) -> TableSpec:
    """The loaded table in the engine's upload form (query-path columns only)."""
    chunk_sizes = np.asarray(state.chunk_sizes, np.uint32)
    n = int(np.sum(chunk_sizes, dtype=np.int64))
    if n != state.sequence_count:
        raise SiloFormatError(f"table {state.schema.name!r}: row layout holds {n} rows, the table {state.sequence_count}")
    spec = TableSpec(chunk_sizes=chunk_sizes)
    for col in state.schema.columns:
        data = state.columns[col.name]
        if col.type in (NUC_SEQUENCE, AA_SEQUENCE):
            spec.seq_columns.append(finalized_column(data, chunk_sizes))
        elif col.type in (INT32, DATE32):
            values = _flat(data["chunks"], np.int32)
            if len(values) != n:
                raise SiloFormatError(f"column {col.name!r}: {len(values)} values for {n} rows")
            nulls = _flags(_dense_rows(data["null_rows"], chunk_sizes, col.name), n)
            spec.int_columns.append(IntColumnData(col.name, col.type == DATE32, values, nulls))
        elif col.type == DICTIONARY_ENCODED:
            ids = _flat(data["ids"], np.uint32)
            if len(ids) != n:
                raise SiloFormatError(f"column {col.name!r}: {len(ids)} values for {n} rows")
            nulls = _flags(_dense_rows(data["null_rows"], chunk_sizes, col.name), n)
            ids = np.where(nulls.astype(bool), 0, ids).astype(np.uint32)
            names = list(col.dictionary or [])
            if ids.size and int(ids.max()) >= len(names):
                raise SiloFormatError(f"column {col.name!r}: value id outside the dictionary")
            parent = alias = None
            if col.lineage_tree is not None:
                parent, alias = lineage_arrays(col.lineage_tree, len(names))
            spec.dict_columns.append(DictColumnData(col.name, ids, nulls, names, parent, alias))
    return spec


# the serialization version this reader follows (`src/rhydb/common/serialization_version.txt`); a
# saved state of another version is refused, as `DataVersion::isCompatibleVersion` refuses it
# (`common/data_version.cpp:105-108`)
SERIALIZATION_VERSION = 1785915114


@dataclass
class SavedDatabase:
    """A saved database directory: the schema, each table's state and the data version."""
    path: str
    tables: dict
    timestamp: int | None
    serialization_version: int | None

    def table_spec(self, name: str = "default") -> TableSpec:
        return table_spec(self.tables[name])


def read_data_version(text: str) -> tuple[int | None, int | None]:
    """data_version.silo (YAML, `common/data_version.cpp:110-152`): (timestamp, serialization
    version)."""
    fields = {}
    for line in text.splitlines():
        key, sep, value = line.partition(":")
        if sep:
            fields[key.strip()] = value.strip().strip("'\"")
    def number(k):
        v = fields.get(k)
        if v is None:
            return None
        if not v.isdigit():
            raise SiloFormatError(f"{DATA_VERSION_FILENAME}: {k} {v!r} is not a number")
        return int(v)
    return number("timestamp"), number("serializationVersion")


def load_database(path: str) -> SavedDatabase:
    """Reads a saved database directory (`Database::loadDatabaseState`, `database.cpp:426-455`):
    its data version, the schema, then `<table>.silo` for each table of it."""
    with open(os.path.join(path, DATA_VERSION_FILENAME), "r", encoding="utf-8") as f:
        timestamp, version = read_data_version(f.read())
    if version != SERIALIZATION_VERSION:
        raise SiloFormatError(f"{path}: serialization version {version} is not the supported {SERIALIZATION_VERSION}")
    with open(os.path.join(path, SCHEMA_FILENAME), "rb") as f:
        schema = read_schema(f.read(), SCHEMA_FILENAME)
    tables = {}
    for name, ts in schema.items():
        fname = name + ".silo"
        with open(os.path.join(path, fname), "rb") as f:
            tables[name] = read_table(f.read(), ts, fname)
    return SavedDatabase(path, tables, timestamp, version)


def most_recent_data_directory(root: str) -> str | None:
    """The newest saved state under `root` whose data version this reader supports
    (`RhyDBDirectory::getMostRecentDataDirectory`, `common/silo_directory.cpp:40-92`): subdirectories
    named by their timestamp, each holding a data version file."""
    found = []
    for entry in os.listdir(root):
        d = os.path.join(root, entry)
        vpath = os.path.join(d, DATA_VERSION_FILENAME)
        if not entry.isdigit() or not os.path.isfile(vpath):
            continue
        with open(vpath, "r", encoding="utf-8") as f:
            try:
                timestamp, version = read_data_version(f.read())
            except SiloFormatError:
                continue
        # the timestamp must name the directory (`common/silo_directory.cpp:40-45`)
        if version == SERIALIZATION_VERSION and timestamp == int(entry):
            found.append((timestamp, d))
    return max(found)[1] if found else None


# ------------------------------------------------------------------------------------------------
# Writing a saved state (`Database::saveDatabaseState`, `database.cpp:380-415`) for the column types
# the engine holds: sequence, int32, date32 and dictionary-encoded columns
# ------------------------------------------------------------------------------------------------
class ArchiveWriter:
    """A Boost binary archive being written, with the class information already written."""

    def __init__(self):
        self.buf = bytearray()
        self.classes: set = set()
        self.string(b"serialization::archive")
        self.buf += struct.pack("<H", 20) + b"\x04\x08\x04\x08" + struct.pack("<i", 1)

    def cls(self, key: str, version: int = 0):
        if key not in self.classes:
            self.classes.add(key)
            self.buf += struct.pack("<BI", 0, version)

    def u8(self, v: int):
        self.buf += struct.pack("<B", v)

    def u16(self, v: int):
        self.buf += struct.pack("<H", v)

    def u32(self, v: int):
        self.buf += struct.pack("<I", v)

    def i32(self, v: int):
        self.buf += struct.pack("<i", v)

    def u64(self, v: int):
        self.buf += struct.pack("<Q", v)

    def string(self, s):
        b = s.encode("utf-8") if isinstance(s, str) else bytes(s)
        self.u64(len(b))
        self.buf += b

    def items(self, key: str, n: int):
        self.cls(key)
        self.u64(n)
        self.u32(0)

    def unordered_items(self, key: str, n: int, buckets: int | None = None):
        self.cls(key)
        self.u64(n)
        self.u64(_bucket_count(n) if buckets is None else buckets)
        self.u32(0)

    def array(self, values: np.ndarray, dtype, key: str | None = None):
        if key is not None:
            self.cls(key)
        a = np.ascontiguousarray(values, dtype=np.dtype(dtype).newbyteorder("<"))
        self.u64(a.size)
        self.buf += a.tobytes()

    def roaring(self, values: np.ndarray):
        self.cls("roaring::Roaring")
        b = portable_bytes(values)
        self.u64(len(b))
        self.buf += b


def _bucket_count(n: int) -> int:
    """The bucket count written for an unordered_map of n elements (libstdc++'s prime list, 1 for a
    default-constructed empty map); readers ignore it."""
    if n == 0:
        return 1
    for p in (13, 29, 59, 127, 257, 541, 1109, 2357, 5087, 10273, 20753, 42043, 85229, 172933, 351061, 712697,
              1447153, 2938679, 5967347, 12117689, 24607243, 49969847, 101473717):
        if p >= n:
            return p
    return n | 1


def _runs_of(low: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    breaks = np.nonzero(np.diff(low.astype(np.int64)) != 1)[0]
    starts = np.concatenate([low[:1], low[breaks + 1]]).astype(np.int64)
    ends = np.concatenate([low[breaks], low[-1:]]).astype(np.int64)
    return starts, ends


def _containers(values: np.ndarray, run_optimize: bool) -> list:
    """(key, cardinality, is_run, container bytes) per 2^16 block: arrays up to 4,096 values, else
    bitsets; with run_optimize a container becomes runs where that is smaller (CRoaring's
    `runOptimize`)."""
    v = np.unique(np.asarray(values, np.uint32))
    out = []
    if not v.size:
        return out
    hi = v >> np.uint32(16)
    cuts = np.nonzero(np.diff(hi))[0] + 1
    for part in np.split(v, cuts):
        key = int(part[0] >> 16)
        low = (part & np.uint32(0xFFFF)).astype(np.uint16)
        card = len(low)
        size = 2 * card if card <= 4096 else 8192
        starts, ends = _runs_of(low)
        if run_optimize and 2 + 4 * len(starts) < size:
            body = struct.pack("<H", len(starts)) + np.stack([starts, ends - starts], 1).astype("<u2").tobytes()
            out.append((key, card, True, body))
        elif card <= 4096:
            out.append((key, card, False, low.astype("<u2").tobytes()))
        else:
            bits = np.zeros(65536, np.uint8)
            bits[low] = 1
            out.append((key, card, False, np.packbits(bits, bitorder="little").tobytes()))
    return out


def portable_bytes(values: np.ndarray, run_optimize: bool = True) -> bytes:
    """CRoaring's portable serialization of a set of u32 (RoaringFormatSpec): cookie (with run flags
    when a container is a run), keys and cardinalities, offsets (without run containers, or from 4
    containers on), the containers."""
    conts = _containers(values, run_optimize)
    n = len(conts)
    has_run = any(c[2] for c in conts)
    if has_run:
        flags = bytearray((n + 7) // 8)
        for i, c in enumerate(conts):
            if c[2]:
                flags[i // 8] |= 1 << (i % 8)
        head = bytearray(struct.pack("<I", _SERIAL_COOKIE | ((n - 1) << 16)) + bytes(flags))
    else:
        head = bytearray(struct.pack("<II", _SERIAL_COOKIE_NO_RUN, n))
    for key, card, _, _ in conts:
        head += struct.pack("<HH", key, card - 1)
    if not has_run or n >= _NO_OFFSET_THRESHOLD:
        off = len(head) + 4 * n
        for c in conts:
            head += struct.pack("<I", off)
            off += len(c[3])
    return bytes(head) + b"".join(c[3] for c in conts)


def _one_container_bitmap(starts: list, ends: list) -> bytes:
    """The run-optimized portable form of a set of runs inside one 2^16 block (a row's missing
    positions: the genome is shorter than 2^16)."""
    key = starts[0] >> 16
    card = sum(e - s for s, e in zip(starts, ends))
    size = 2 * card if card <= 4096 else 8192
    if 2 + 4 * len(starts) < size:
        body = b"".join(struct.pack("<HH", s & 0xFFFF, e - s - 1) for s, e in zip(starts, ends))
        return struct.pack("<IBHHH", _SERIAL_COOKIE, 1, key, card - 1, len(starts)) + body
    values = np.concatenate([np.arange(s, e, dtype=np.uint32) for s, e in zip(starts, ends)])
    return portable_bytes(values, run_optimize=False)


def _frozen_size(values: np.ndarray) -> int:
    """`Roaring::getSizeInBytes(false)`: the smaller of the portable form + 1 and a plain u32 list."""
    card = int(np.unique(np.asarray(values, np.uint32)).size)
    return min(len(portable_bytes(values)) + 1, 1 + 4 + 4 * card)


def _global_ids(rows: np.ndarray, chunk_sizes: np.ndarray) -> np.ndarray:
    offsets = np.zeros(len(chunk_sizes) + 1, np.int64)
    np.cumsum(chunk_sizes, out=offsets[1:])
    rows = np.asarray(rows, np.int64)
    chunk = np.searchsorted(offsets, rows, side="right") - 1
    return ((chunk << 16) | (rows - offsets[chunk])).astype(np.uint32)


_CONTAINER_HEADER = np.dtype([("position", "<u4"), ("chunk", "<u2"), ("symbol", "<i4"), ("cardinality", "<u4"),
                              ("typecode", "u1"), ("length", "<u8")])
_RUN_BITMAP_HEADER = np.dtype([("row", "<u4"), ("length", "<u8"), ("cookie", "<u4"), ("flags", "u1"), ("key", "<u2"),
                               ("card", "<u2"), ("runs", "<u2")])


def _interleave(headers: np.ndarray, body_src: np.ndarray, body_at: np.ndarray, body_len: np.ndarray) -> np.ndarray:
    """Records of a fixed header each followed by body_len[i] bytes of body_src from body_at[i]."""
    h = headers.dtype.itemsize
    size = h + body_len
    start = np.zeros(len(size), np.int64)
    start[1:] = np.cumsum(size)[:-1]
    out = np.empty(int(size.sum()), np.uint8)
    out[(start[:, None] + np.arange(h)).ravel()] = headers.view(np.uint8).reshape(-1, h).ravel()
    out[_ranges(start + h, body_len)] = body_src[_ranges(body_at, body_len)]
    return out


def _records(w: ArchiveWriter, order: np.ndarray, c: FinalizedSeqColumnData, payload: np.ndarray, offs: np.ndarray,
             sizes: np.ndarray, block: int = 1 << 16):
    """Vertical-index records after the first: header + container bytes, block by block."""
    for b in range(0, len(order), block):
        sel = order[b:b + block]
        hdr = np.empty(len(sel), _CONTAINER_HEADER)
        hdr["position"] = np.asarray(c.key_position)[sel]
        hdr["chunk"] = np.asarray(c.key_chunk)[sel]
        hdr["symbol"] = np.asarray(c.key_symbol)[sel]
        hdr["cardinality"] = np.asarray(c.cardinality)[sel]
        hdr["typecode"] = np.asarray(c.typecode)[sel]
        hdr["length"] = sizes[b:b + block]
        w.buf += _interleave(hdr, payload, offs[b:b + block], sizes[b:b + block]).tobytes()


def _row_bitmap_records(w: ArchiveWriter, rows: np.ndarray, gids: np.ndarray, rptr: np.ndarray, rstart: np.ndarray,
                        rend: np.ndarray, block: int = 1 << 15) -> int:
    """Horizontal-index records after the first: row id, then the row's missing positions as a
    run-optimized portable bitmap of one container (the genome is shorter than 2^16).  Rows whose
    runs would not be the smaller form are written one by one.  Returns their frozen sizes' sum."""
    total = 0
    rstart = rstart.astype(np.int64)
    rend = rend.astype(np.int64)
    for b in range(0, len(rows), block):
        sel = rows[b:b + block].astype(np.int64)
        g = gids[b:b + block]
        nr = (rptr[sel + 1] - rptr[sel]).astype(np.int64)
        idx = _ranges(rptr[sel].astype(np.int64), nr)
        lens = rend[idx] - rstart[idx]
        owner = np.repeat(np.arange(len(sel)), nr)
        card = np.bincount(owner, weights=lens, minlength=len(sel)).astype(np.int64)
        as_runs = 2 + 4 * nr < np.where(card <= 4096, 2 * card, 8192)
        if as_runs.all():
            hdr = np.empty(len(sel), _RUN_BITMAP_HEADER)
            hdr["row"] = g
            hdr["length"] = 11 + 4 * nr
            hdr["cookie"] = _SERIAL_COOKIE
            hdr["flags"] = 1
            hdr["key"] = 0
            hdr["card"] = card - 1
            hdr["runs"] = nr
            pairs = np.stack([rstart[idx], lens - 1], 1).astype("<u2").view(np.uint8).ravel()
            body_at = np.zeros(len(sel), np.int64)
            body_at[1:] = np.cumsum(4 * nr)[:-1]
            w.buf += _interleave(hdr, pairs, body_at, 4 * nr).tobytes()
            total += int(np.minimum(11 + 4 * nr + 1, 5 + 4 * card).sum())
            continue
        for j in range(len(sel)):
            a, e = int(rptr[sel[j]]), int(rptr[sel[j] + 1])
            bm = _one_container_bitmap(rstart[a:e].tolist(), rend[a:e].tolist())
            w.u32(int(g[j]))
            w.u64(len(bm))
            w.buf += bm
            total += min(len(bm) + 1, 5 + 4 * int(card[j]))
    return total


def _write_sequence_column(w: ArchiveWriter, c: FinalizedSeqColumnData, chunk_sizes: np.ndarray):
    tag = "Nucleotide" if c.alphabet == ALPHABET_NUC else "AminoAcid"
    chars = _symbol_chars(c.alphabet)
    w.cls(f"SequenceColumn<{tag}>")
    w.string("".join(chars[s] for s in np.asarray(c.local_reference).tolist()))
    w.cls(f"VerticalSequenceIndex<{tag}>")
    k = len(c.key_position)
    order = np.lexsort((np.asarray(c.key_symbol), np.asarray(c.key_chunk), np.asarray(c.key_position)))
    w.items(f"map<SequenceDiffKey<{tag}>,RoaringContainer>", k)
    payload = np.asarray(c.payload, np.uint8)
    tcs = np.asarray(c.typecode, np.int64)[order]
    cards = np.asarray(c.cardinality, np.int64)[order]
    offs = np.asarray(c.payload_offset, np.int64)[order]
    n_runs = np.zeros(k, np.int64)
    runs = np.nonzero(tcs == 3)[0]
    n_runs[runs] = payload[offs[runs]].astype(np.int64) | (payload[offs[runs] + 1].astype(np.int64) << 8)
    sizes = np.where(tcs == 1, 8192, np.where(tcs == 2, 2 * cards, 2 + 4 * n_runs))
    vertical_size = int(sizes.sum())
    if k:
        # the first entry carries the class information; the rest are fixed headers + bytes
        i = int(order[0])
        w.cls(f"pair<const SequenceDiffKey<{tag}>,RoaringContainer>")
        w.cls(f"SequenceDiffKey<{tag}>")
        w.u32(int(c.key_position[i]))
        w.u16(int(c.key_chunk[i]))
        w.i32(int(c.key_symbol[i]))
        w.cls("RoaringContainer")
        w.u32(int(c.cardinality[i]))
        w.u8(int(c.typecode[i]))
        w.string(payload[offs[0]:offs[0] + sizes[0]].tobytes())
        _records(w, order[1:], c, payload, offs[1:], sizes[1:])
    w.cls("HorizontalCoverageIndex")
    n = len(c.cov_start)
    rptr, rstart, rend = c.missing_runs()
    rptr = rptr.astype(np.int64)
    with_missing = np.nonzero(np.diff(rptr) > 0)[0]
    gids = _global_ids(with_missing, chunk_sizes)
    w.items("map<u32,Roaring>", len(with_missing))
    horizontal_size = 0
    if len(with_missing):
        row = int(with_missing[0])
        w.cls("pair<const u32,Roaring>")
        w.u32(int(gids[0]))
        w.cls("roaring::Roaring")
        a, b = int(rptr[row]), int(rptr[row + 1])
        bm = _one_container_bitmap(rstart[a:b].tolist(), rend[a:b].tolist())
        w.u64(len(bm))
        w.buf += bm
        card = int((rend[a:b].astype(np.int64) - rstart[a:b]).sum())
        horizontal_size = min(len(bm) + 1, 5 + 4 * card)
        horizontal_size += _row_bitmap_records(w, with_missing[1:], gids[1:], rptr, rstart, rend)
    start_end = np.stack([np.asarray(c.cov_start, np.uint32), np.asarray(c.cov_end, np.uint32)], 1)
    offsets = np.concatenate([[0], np.cumsum(chunk_sizes)]).astype(np.int64)
    w.items("vector<vector<pair<u32,u32>>>", len(chunk_sizes))
    batch = []
    for ci in range(len(chunk_sizes)):
        part = start_end[offsets[ci]:offsets[ci + 1]]
        w.cls("vector<pair<u32,u32>>")
        w.u64(len(part))
        w.buf += part.astype("<u4").tobytes()
        batch.append((int(part[:, 0].min()) if len(part) else 0xFFFFFFFF, int(part[:, 1].max()) if len(part) else 0))
    w.cls("vector<pair<u32,u32>>")
    w.u64(len(batch))
    w.buf += np.asarray(batch, "<u4").reshape(-1, 2).tobytes()
    w.cls(f"InsertionIndex<{tag}>")
    # an empty insertion index as the reference saves it (its position map shows 2 buckets)
    w.unordered_items(f"unordered_map<u32,InsertionPosition<{tag}>>", 0, 2)
    w.unordered_items("unordered_map<u32,unordered_map<string,Roaring>>", 0)
    w.cls("SequenceColumnInfo")
    w.u32(n)
    w.u64(vertical_size)
    w.u64(horizontal_size)
    w.u32(n)
    w.roaring(_global_ids(np.nonzero(np.asarray(c.is_null))[0], chunk_sizes))
    w.u16(len(chunk_sizes))


def _chunks(values: np.ndarray, chunk_sizes: np.ndarray) -> list:
    offsets = np.concatenate([[0], np.cumsum(chunk_sizes)]).astype(np.int64)
    return [values[offsets[i]:offsets[i + 1]] for i in range(len(chunk_sizes))]


def _write_chunked(w: ArchiveWriter, tag: str, dtype, values: np.ndarray, chunk_sizes: np.ndarray):
    w.cls(f"ChunkedValueBuffer<{tag}>")
    parts = _chunks(np.asarray(values), chunk_sizes)
    w.items(f"vector<vector<{tag}>>", len(parts))
    for part in parts:
        w.array(part, dtype)


def _resolved(c: DictColumnData) -> np.ndarray:
    n = len(c.names)
    alias = np.full(n, -1, np.int64) if c.lineage_alias is None else np.asarray(c.lineage_alias, np.int64)
    return np.where(alias >= 0, alias, np.arange(n))


def _write_bitmap_map(w: ArchiveWriter, groups: dict, encoded: dict | None = None):
    """unordered_map<u32, Roaring>; `encoded` caches the portable bytes of maps written again."""
    w.unordered_items("unordered_map<u32,Roaring>", len(groups))
    for k, rows in groups.items():
        w.cls("pair<const u32,Roaring>")
        w.u32(k)
        w.cls("roaring::Roaring")
        b = encoded.get(k) if encoded is not None else None
        if b is None:
            b = portable_bytes(rows)
            if encoded is not None:
                encoded[k] = b
        w.u64(len(b))
        w.buf += b


def _write_dictionary_column(w: ArchiveWriter, c: DictColumnData, chunk_sizes: np.ndarray):
    w.cls("DictionaryEncodedColumn")
    ids = np.asarray(c.ids, np.uint32)
    nulls = np.asarray(c.is_null).astype(bool)
    _write_chunked(w, "u32", np.uint32, ids, chunk_sizes)
    gids = _global_ids(np.arange(len(ids)), chunk_sizes)
    valid = np.nonzero(~nulls)[0]
    by_value = {}
    for v in np.unique(ids[valid]).tolist():
        by_value[v] = gids[valid[ids[valid] == v]]
    _write_bitmap_map(w, by_value)
    w.roaring(gids[nulls])
    if c.lineage_parent is not None:
        # lineage_index.cpp:17-25: rows under their value's resolved lineage, and under every
        # ancestor of it for each recombinant-edge mode (one parent per lineage: the modes agree)
        res = _resolved(c)
        parent = np.asarray(c.lineage_parent, np.int64)
        own = {}
        for v, rows in by_value.items():
            r_ = int(res[v])
            own[r_] = np.concatenate([own[r_], rows]) if r_ in own else rows
        own = {k: np.sort(v) for k, v in own.items()}
        including = {}
        for v, rows in own.items():
            a, seen = v, 0
            while a >= 0 and seen <= len(parent):
                including.setdefault(a, []).append(rows)
                a = int(parent[a])
                seen += 1
        including = {k: np.sort(np.concatenate(v)) for k, v in including.items()}
        w.cls("LineageIndex")
        w.unordered_items("unordered_map<RecombinantEdgeFollowingMode,unordered_map<u32,Roaring>>", 3)
        encoded = {}
        for mode in range(3):
            w.cls("pair<const RecombinantEdgeFollowingMode,unordered_map<u32,Roaring>>")
            w.i32(mode)
            _write_bitmap_map(w, including, encoded)
        _write_bitmap_map(w, own)


def _lineage_yaml(c: DictColumnData) -> str:
    """A lineage definition file (`preprocessing/lineage_definition_file.cpp` format) of the
    column's lineages: per lineage its parent and aliases."""
    res = _resolved(c)
    parent = np.asarray(c.lineage_parent, np.int64)
    aliases = {}
    for v in range(len(c.names)):
        if res[v] != v:
            aliases.setdefault(int(res[v]), []).append(c.names[v])
    lines = []
    for v, name in enumerate(c.names):
        if res[v] != v:
            continue
        body = []
        if v in aliases:
            body += ["  aliases:"] + [f"  - {json_quote(a)}" for a in aliases[v]]
        if parent[v] >= 0:
            body += ["  parents:", f"  - {json_quote(c.names[parent[v]])}"]
        lines.append(f"{json_quote(name)}:" + ("" if body else " {}"))
        lines += body
    return "\n".join(lines) + "\n"


def json_quote(s: str) -> str:
    import json
    return json.dumps(s)


def _write_column_metadata(w: ArchiveWriter, kind: int, name: str, col):
    if kind in (INT32, DATE32):
        w.cls("ColumnMetadata")
        w.string(name)
    elif kind == DICTIONARY_ENCODED:
        w.cls("DictionaryEncodedColumnMetadata")
        w.string(name)
        w.cls("BidirectionalStringMap")
        w.items("vector<string>", len(col.names))
        for s in col.names:
            w.string(s)
        w.cls("optional<LineageTreeAndIdMap>", 1)
        w.u8(1 if col.lineage_parent is not None else 0)
        if col.lineage_parent is not None:
            res = _resolved(col)
            parent = np.asarray(col.lineage_parent, np.int64)
            w.cls("LineageTreeAndIdMap")
            w.cls("LineageTree")
            w.items("vector<vector<u32>>", len(col.names))
            for v in range(len(col.names)):
                w.array(np.asarray([parent[v]] if parent[v] >= 0 and res[v] == v else [], np.uint32), np.uint32)
            w.unordered_items("unordered_map<u32,optional<u32>>", 0)
            alias = [(v, int(res[v])) for v in range(len(col.names)) if res[v] != v]
            w.unordered_items("unordered_map<u32,u32>", len(alias))
            for a, t in alias:
                w.cls("pair<const u32,u32>")
                w.u32(a)
                w.u32(t)
            w.cls("BidirectionalStringMap")
            w.items("vector<string>", len(col.names))
            for s in col.names:
                w.string(s)
            w.string(_lineage_yaml(col))
        w.u8(0)  # treat_unknown_lineages_as_null
    else:
        tag = "Nucleotide" if kind == NUC_SEQUENCE else "AminoAcid"
        w.cls(f"SequenceColumnMetadata<{tag}>")
        w.string(name)
        ref = np.asarray(col.reference, np.int64)
        w.items(f"vector<{tag}::Symbol>", len(ref))
        w.buf += ref.astype("<i4").tobytes()


def _spec_columns(spec: TableSpec) -> list:
    """(name, type, data) of every column of the spec, in the schema's (name, type) order."""
    cols = []
    for c in spec.seq_columns:
        if not isinstance(c, FinalizedSeqColumnData):
            raise TypeError(f"column {c.name!r}: only finalized sequence columns are written")
        cols.append((c.name, NUC_SEQUENCE if c.alphabet == ALPHABET_NUC else AA_SEQUENCE, c))
    cols += [(c.name, DATE32 if c.is_date else INT32, c) for c in spec.int_columns]
    cols += [(c.name, DICTIONARY_ENCODED, c) for c in spec.dict_columns]
    names = [n for n, _, _ in cols]
    if len(set(names)) != len(names):
        raise ValueError("duplicate column names")
    return sorted(cols, key=lambda x: (x[0].encode("utf-8"), x[1]))


def write_schema(tables: dict) -> bytes:
    """database_schema.silo for {table name: TableSpec}.  The primary key names the table's first
    column in schema order (the reference's string primary-key column is not part of the engine's
    tables)."""
    w = ArchiveWriter()
    w.items("map<TableName,shared_ptr<TableSchema>>", len(tables))
    for tname in sorted(tables, key=lambda s: s.encode("utf-8")):
        cols = _spec_columns(tables[tname])
        if not cols:
            raise ValueError(f"table {tname!r} has no columns")
        w.cls("pair<const TableName,shared_ptr<TableSchema>>")
        w.cls("TableName")
        w.string(tname)
        w.cls("shared_ptr<TableSchema>")
        w.cls("TableSchema")
        w.cls("ColumnIdentifier")
        w.string(cols[0][0])
        w.i32(cols[0][1])
        w.items("vector<ColumnIdentifier>", len(cols))
        for name, kind, _ in cols:
            w.cls("ColumnIdentifier")
            w.string(name)
            w.i32(kind)
        for name, kind, col in cols:
            _write_column_metadata(w, kind, name, col)
    return bytes(w.buf)


def write_table(spec: TableSpec) -> bytes:
    """<table>.silo for a TableSpec of finalized sequence, int32, date32 and dictionary columns."""
    chunk_sizes = np.asarray(spec.chunk_sizes, np.uint32)
    cols = _spec_columns(spec)
    w = ArchiveWriter()
    w.cls("ColumnGroup")
    for t in TABLE_COLUMN_ORDER:
        for name, kind, col in sorted((c for c in cols if c[1] == t), key=lambda c: c[0].encode("utf-8")):
            if kind == DICTIONARY_ENCODED:
                _write_dictionary_column(w, col, chunk_sizes)
            elif kind == INT32:
                w.cls("IntColumn<i32>")
                _write_chunked(w, "i32", np.int32, col.values, chunk_sizes)
                w.roaring(_global_ids(np.nonzero(np.asarray(col.is_null))[0], chunk_sizes))
            elif kind == DATE32:
                w.cls("Date32Column")
                w.roaring(_global_ids(np.nonzero(np.asarray(col.is_null))[0], chunk_sizes))
                _write_chunked(w, "i32", np.int32, col.values, chunk_sizes)
                w.u8(0)  # is_sorted
            else:
                _write_sequence_column(w, col, chunk_sizes)
    w.u32(int(np.sum(chunk_sizes, dtype=np.int64)))
    w.cls("RowLayout")
    w.array(chunk_sizes, np.uint32)
    return bytes(w.buf)


def save_database(root: str, tables: dict, timestamp: int) -> str:
    """Writes {table name: TableSpec} as a saved state under root/<timestamp>/ (the directory layout
    `Database::saveDatabaseState` writes, `database.cpp:380-415`) and returns that directory."""
    d = os.path.join(root, str(int(timestamp)))
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, SCHEMA_FILENAME), "wb") as f:
        f.write(write_schema(tables))
    for name, spec in tables.items():
        with open(os.path.join(d, name + ".silo"), "wb") as f:
            f.write(write_table(spec))
    with open(os.path.join(d, DATA_VERSION_FILENAME), "w", encoding="utf-8") as f:
        f.write(f"timestamp: {int(timestamp)}\nserializationVersion: {SERIALIZATION_VERSION}")
    return d
