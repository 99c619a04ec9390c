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
        return self.string_bytes().decode("utf-8")

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
_SERIAL_COOKIE_NO_RUN = 12346
_SERIAL_COOKIE = 12347
_NO_OFFSET_THRESHOLD = 4


def container_values(typecode: int, cardinality: int, payload: bytes) -> np.ndarray:
    """The u16 values of one container in its `container_write` form (1 bitset, 2 array, 3 run)."""
    if typecode == 1:
        words = np.frombuffer(payload[:8192], dtype="<u8")
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return np.nonzero(bits)[0].astype(np.uint16)
    if typecode == 2:
        return np.frombuffer(payload[:2 * cardinality], dtype="<u2").copy()
    if typecode == 3:
        n_runs = struct.unpack_from("<H", payload)[0]
        runs = np.frombuffer(payload[2:2 + 4 * n_runs], dtype="<u2").reshape(-1, 2)
        if n_runs == 0:
            return np.zeros(0, np.uint16)
        return np.concatenate([np.arange(s, int(s) + int(l) + 1, dtype=np.uint32) for s, l in runs]).astype(np.uint16)
    raise SiloFormatError(f"unknown container typecode {typecode}")


def roaring_values(buf: bytes, name: str = "roaring") -> np.ndarray:
    """Decodes CRoaring's portable format (the format `Roaring::write` produces) to sorted u32."""
    if len(buf) < 4:
        raise SiloFormatError(f"{name}: roaring bitmap too short")
    cookie = struct.unpack_from("<I", buf)[0]
    pos = 4
    run_flags = None
    if cookie & 0xFFFF == _SERIAL_COOKIE:
        size = (cookie >> 16) + 1
        nbytes = (size + 7) // 8
        run_flags = np.unpackbits(np.frombuffer(buf[pos:pos + nbytes], np.uint8), bitorder="little")[:size]
        pos += nbytes
    elif cookie == _SERIAL_COOKIE_NO_RUN:
        size = struct.unpack_from("<I", buf, pos)[0]
        pos += 4
    else:
        raise SiloFormatError(f"{name}: bad roaring cookie {cookie:#x}")
    hdr = np.frombuffer(buf[pos:pos + 4 * size], dtype="<u2").reshape(-1, 2) if size else np.zeros((0, 2), np.uint16)
    pos += 4 * size
    if run_flags is None or size >= _NO_OFFSET_THRESHOLD:
        pos += 4 * size  # container offsets
    out = []
    for i in range(size):
        key = int(hdr[i, 0])
        card = int(hdr[i, 1]) + 1
        if run_flags is not None and run_flags[i]:
            n_runs = struct.unpack_from("<H", buf, pos)[0]
            runs = np.frombuffer(buf[pos + 2:pos + 2 + 4 * n_runs], dtype="<u2").reshape(-1, 2)
            pos += 2 + 4 * n_runs
            vals = np.concatenate([np.arange(s, int(s) + int(l) + 1, dtype=np.uint32) for s, l in runs])
        elif card > 4096:
            words = np.frombuffer(buf[pos:pos + 8192], dtype="<u8")
            pos += 8192
            vals = np.nonzero(np.unpackbits(words.view(np.uint8), bitorder="little"))[0].astype(np.uint32)
        else:
            vals = np.frombuffer(buf[pos:pos + 2 * card], dtype="<u2").astype(np.uint32)
            pos += 2 * card
        out.append((np.uint32(key) << np.uint32(16)) | vals.astype(np.uint32))
    if pos > len(buf):
        raise SiloFormatError(f"{name}: truncated roaring bitmap")
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
    missing: dict                   # row -> sorted u32 positions of its internal missing symbols
    start_end: np.ndarray           # (rows, 2) u32 covered ranges, in row order
    null_rows: np.ndarray           # u32 row ids
    sequence_count: int
    num_chunks: int
    info: tuple                     # SequenceColumnInfo: (sequence_count, vertical, horizontal sizes)
    insertion_count: int


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


def _unordered_bitmaps(r: ArchiveReader) -> dict:
    out = {}
    for _ in range(r.unordered_items("unordered_map<u32,Roaring>")):
        r.cls("pair<const u32,Roaring>")
        k = r.u32()
        out[k] = r.roaring()
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
    pos = np.empty(n, np.uint32)
    chunk = np.empty(n, np.uint16)
    sym = np.empty(n, np.uint8)
    tc = np.empty(n, np.uint8)
    card = np.empty(n, np.uint32)
    offs = np.zeros(n, np.uint64)
    payload = []
    total = 0
    for i in range(n):
        r.cls(f"pair<const SequenceDiffKey<{tag}>,RoaringContainer>")
        r.cls(f"SequenceDiffKey<{tag}>")
        pos[i] = r.u32()
        chunk[i] = r.u16()
        s = r.i32()
        if not 0 <= s < n_symbols:
            r.fail(f"column {col.name!r}: symbol {s} out of range")
        sym[i] = s
        r.cls("RoaringContainer")
        card[i] = r.u32()
        tc[i] = r.u8()
        buf = r.string_bytes()
        offs[i] = total
        payload.append(buf)
        total += len(buf)
    r.cls("HorizontalCoverageIndex")
    missing = {}
    for _ in range(r.items("map<u32,Roaring>")):
        r.cls("pair<const u32,Roaring>")
        row = r.u32()
        missing[row] = r.roaring()
    chunks = [_pair_array(r) for _ in range(r.items("vector<vector<pair<u32,u32>>>"))]
    _pair_array(r)  # batch_start_ends
    insertions = _insertion_index(r, tag)
    r.cls("SequenceColumnInfo")
    info = (r.u32(), r.u64(), r.u64())
    seq_count = r.u32()
    nulls = r.roaring()
    num_chunks = r.u16()
    start_end = np.concatenate(chunks) if chunks else np.zeros((0, 2), np.uint32)
    return SequenceColumnState(
        col.name, alphabet, col.reference, local_ref, pos, chunk, sym, tc, card, offs,
        np.frombuffer(b"".join(payload), np.uint8).copy(), missing, start_end, nulls, seq_count, num_chunks, info,
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
                cols[col.name] = _sequence_column(r, col)
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
    missing_rows = np.array(sorted(col.missing), np.uint32)
    dense = _dense_rows(missing_rows, chunk_sizes, col.name)
    counts = np.zeros(n, np.uint64)
    counts[dense] = [len(col.missing[int(g)]) for g in missing_rows]
    ptr = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=ptr[1:])
    pos = np.zeros(int(ptr[-1]), np.uint32)
    for g, d in zip(missing_rows.tolist(), dense.tolist()):
        pos[int(ptr[d]):int(ptr[d + 1])] = col.missing[g]
    is_null = _flags(_dense_rows(col.null_rows, chunk_sizes, col.name), n)
    return FinalizedSeqColumnData(
        col.name, col.alphabet, col.reference, col.local_reference, col.key_position, col.key_chunk, col.key_symbol,
        col.typecode, col.cardinality, col.payload_offset, col.payload, is_null,
        np.ascontiguousarray(col.start_end[:, 0]), np.ascontiguousarray(col.start_end[:, 1]), ptr, pos)


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
