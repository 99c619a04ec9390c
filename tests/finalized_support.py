"""Reading a finalized sequence column (silo_amd.table.FinalizedSeqColumnData, the state a finalized
reference SequenceColumn holds) back into the diff form, in numpy: the test-side statement of the
rule the engine's upload applies (a stored container's rows carry its symbol unless it is the
column's reference symbol; at a position whose local reference differs from the reference, the
covered rows in no container carry the local reference symbol; missing positions form runs)."""
import numpy as np


def container_rows(f, k: int) -> np.ndarray:
    o = int(f.payload_offset[k])
    t = int(f.typecode[k])
    if t == 1:
        words = np.frombuffer(f.payload[o:o + 8192].tobytes(), dtype="<u8")
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return np.flatnonzero(bits).astype(np.uint32)
    if t == 2:
        return np.frombuffer(f.payload[o:o + 2 * int(f.cardinality[k])].tobytes(), dtype="<u2").astype(np.uint32)
    n = int(np.frombuffer(f.payload[o:o + 2].tobytes(), dtype="<u2")[0])
    pairs = np.frombuffer(f.payload[o + 2:o + 2 + 4 * n].tobytes(), dtype="<u2").reshape(-1, 2).astype(np.uint32)
    return np.concatenate([np.arange(a, a + ln + 1, dtype=np.uint32) for a, ln in pairs]) if n else np.zeros(0, np.uint32)


def to_diff_form(f, chunk_sizes):
    """(mut_ptr, mut_pos, mut_sym, nrun_ptr, nrun_start, nrun_end) of the finalized column."""
    sizes = np.asarray(chunk_sizes, np.uint64)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    n = int(offsets[-1])
    ref = np.asarray(f.reference)
    local = np.asarray(f.local_reference)
    rows, pos, sym = [], [], []
    stored = {}
    for k in range(len(f.key_position)):
        p, c, s = int(f.key_position[k]), int(f.key_chunk[k]), int(f.key_symbol[k])
        dense = container_rows(f, k).astype(np.uint64) + offsets[c]
        if local[p] != ref[p]:
            stored.setdefault(p, []).append(dense)
        if s != ref[p]:
            rows.append(dense)
            pos.append(np.full(len(dense), p, np.uint32))
            sym.append(np.full(len(dense), s, np.uint8))
    miss_rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(f.missing_ptr.astype(np.int64)))
    for p in np.flatnonzero(local != ref).tolist():
        covered = (f.is_null == 0) & (f.cov_start <= p) & (p < f.cov_end)
        covered[miss_rows[f.missing_pos == p].astype(np.int64)] = False
        if p in stored:
            covered[np.concatenate(stored[p]).astype(np.int64)] = False
        dense = np.flatnonzero(covered).astype(np.uint64)
        rows.append(dense)
        pos.append(np.full(len(dense), p, np.uint32))
        sym.append(np.full(len(dense), local[p], np.uint8))
    rows = np.concatenate(rows) if rows else np.zeros(0, np.uint64)
    pos = np.concatenate(pos) if pos else np.zeros(0, np.uint32)
    sym = np.concatenate(sym) if sym else np.zeros(0, np.uint8)
    order = np.lexsort((pos, rows))
    rows, pos, sym = rows[order], pos[order], sym[order]
    mut_ptr = np.zeros(n + 1, np.uint64)
    mut_ptr[1:] = np.cumsum(np.bincount(rows.astype(np.int64), minlength=n))
    # missing positions -> runs
    mp = f.missing_pos.astype(np.int64)
    new_run = np.ones(len(mp), bool)
    if len(mp):
        new_run[1:] = (mp[1:] != mp[:-1] + 1) | (miss_rows[1:] != miss_rows[:-1])
    starts = np.flatnonzero(new_run)
    ends = np.r_[starts[1:], len(mp)] - 1
    nrun_ptr = np.zeros(n + 1, np.uint64)
    nrun_ptr[1:] = np.cumsum(np.bincount(miss_rows[starts].astype(np.int64), minlength=n))
    return mut_ptr, pos, sym, nrun_ptr, mp[starts].astype(np.uint32), (mp[ends] + 1).astype(np.uint32)
